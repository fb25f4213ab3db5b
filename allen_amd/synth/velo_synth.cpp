// Synthetic VELO event generator (see velo_synth.h for the model and the encoding).
#include "velo_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

constexpr unsigned n_modules = 52;
constexpr unsigned n_sensors = 208;
constexpr unsigned n_cols = 768;
constexpr unsigned n_rows = 256;
constexpr float pixel_size = 0.055f;

struct Geometry {
  float module_zs[n_modules];
  float local_x[n_cols];
  float x_pitch[n_cols];
  float ltg[12 * n_sensors];
  double col_edge[n_cols + 1];
};

bool load_geometry(const char* p0, size_t n, Geometry& g)
{
  const char* p = p0;
  auto arr = [&](size_t expected, float* d) {
    uint64_t m;
    std::memcpy(&m, p, 8);
    p += 8;
    if (m != expected) return false;
    std::memcpy(d, p, 4 * m);
    p += 4 * m;
    return true;
  };
  if (!arr(n_modules, g.module_zs) || !arr(n_cols, g.local_x) || !arr(n_cols, g.x_pitch)) return false;
  uint64_t n_ltg, n_trans;
  std::memcpy(&n_ltg, p, 8);
  std::memcpy(&n_trans, p + 8, 8);
  p += 16;
  if (n_ltg != n_sensors || n_trans != 12) return false;
  std::memcpy(g.ltg, p, 4 * n_ltg * n_trans);
  p += 4 * n_ltg * n_trans;
  if (static_cast<size_t>(p - p0) != n) return false;
  // Column c spans [edge[c], edge[c+1]); local_x[c] is its centre and x_pitch[c] the
  // distance to the next centre, so edges sit half-way between centres.
  g.col_edge[0] = g.local_x[0] - 0.5 * pixel_size;
  for (unsigned c = 0; c + 1 < n_cols; ++c) g.col_edge[c + 1] = g.local_x[c] + 0.5 * g.x_pitch[c];
  g.col_edge[n_cols] = g.local_x[n_cols - 1] + 0.5 * pixel_size;
  return true;
}

struct Pixel {
  uint16_t col, row;
};

struct EventOut {
  std::vector<uint32_t> block;             // event block: n_banks, offsets, banks
  std::vector<uint16_t> bank_sizes;
  std::vector<uint8_t> bank_types;
  std::vector<velo_synth_particle> particles;
  std::vector<velo_synth_hit> hits;         // particle index local to the event
};

uint32_t channel_lhcb_id(unsigned sensor, unsigned col, unsigned row)
{
  const unsigned chip = col >> 8;
  const uint32_t cid = (sensor << 18) | (chip << 16) | ((col & 0xFF) << 8) | row;
  return (0x2u << 28) | (cid & 0x0FFFFFFFu);
}

double neg_binomial(std::mt19937_64& rng, double mean, double k)
{
  if (mean <= 0) return 0;
  std::gamma_distribution<double> gam(k, mean / k);
  std::poisson_distribution<int> po(gam(rng));
  return po(rng);
}

// Uniform (0, 1] from a 64-bit key (splitmix64 finaliser), for draws that must not disturb the
// event's main random stream.
double keyed_uniform(uint64_t key)
{
  key += 0x9E3779B97F4A7C15ull;
  key = (key ^ (key >> 30)) * 0xBF58476D1CE4E5B9ull;
  key = (key ^ (key >> 27)) * 0x94D049BB133111EBull;
  key ^= key >> 31;
  return (static_cast<double>(key >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

void generate_event(const velo_synth_config& c, const Geometry& g, unsigned ev, EventOut& out)
{
  // Per-event stream: reproducible whatever the thread partition.
  uint64_t s = c.seed * 0x9E3779B97F4A7C15ull + (ev + 1) * 0xBF58476D1CE4E5B9ull;
  s ^= s >> 31;
  std::mt19937_64 rng(s);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  std::normal_distribution<double> gauss(0.0, 1.0);

  // Dead ASICs of this event (sensor * 3 + chip).
  std::vector<uint8_t> dead(n_sensors * 3, 0);
  if (c.dead_asic_fraction > 0) {
    std::mt19937_64 drng(c.seed ^ 0xDEADA51Cull);
    for (auto& d : dead) d = uni(drng) < c.dead_asic_fraction;
  }

  std::vector<std::vector<Pixel>> sensor_pixels(n_sensors);
  // Retina encoding: one word per generated cluster (centroid in quarter pixels).
  std::vector<std::vector<uint32_t>> sensor_retina(n_sensors);
  auto retina_word = [&](unsigned sensor, unsigned sc, unsigned sr, unsigned n, uint32_t* lhcb_id) {
    const unsigned cx4 = 4 * sc / n, cy4 = 4 * sr / n;
    const unsigned cx = cx4 >> 2, cy = cy4 >> 2, fx = cx4 & 3u, fy = cy4 & 3u;
    const uint32_t cid = ((fx != 0) << 27) | ((fy != 0) << 26) | (sensor << 18) | ((cx >> 8) << 16) |
                         ((cx & 0xFFu) << 8) | cy;
    *lhcb_id = (0x2u << 28) | (cid & 0x0FFFFFFFu);
    const uint32_t w = (cx << 12) | (fx << 10) | (cy << 2) | fy;
    return c.bank_version > 3 ? (w | ((sensor & 1u) << 22)) : w;
  };

  // Collisions and tracks.
  struct Vtx {
    double x, y, z;
    unsigned n;
  };
  std::vector<Vtx> pvs;
  auto draw_pv = [&](unsigned n) {
    pvs.push_back({c.pv_sigma_xy * gauss(rng), c.pv_sigma_xy * gauss(rng), c.pv_sigma_z * gauss(rng), n});
  };
  if (c.pileup) {
    const double nu = (c.high_occupancy_fraction > 0 && uni(rng) < c.high_occupancy_fraction) ? c.high_nu : c.nu;
    std::poisson_distribution<int> pn(nu);
    const int n_coll = pn(rng);
    for (int i = 0; i < n_coll; ++i) draw_pv(static_cast<unsigned>(neg_binomial(rng, c.nb_mean, c.nb_k)));
  }
  else if (c.high_occupancy_fraction > 0 && uni(rng) < c.high_occupancy_fraction) {
    // A high-occupancy event in a single-collision workload: nu = high_nu pile-up collisions.
    std::poisson_distribution<int> pn(c.high_nu);
    const int n_coll = pn(rng);
    for (int i = 0; i < n_coll; ++i) draw_pv(static_cast<unsigned>(neg_binomial(rng, c.nb_mean, c.nb_k)));
  }
  else {
    std::poisson_distribution<int> pt(c.mean_tracks);
    draw_pv(static_cast<unsigned>(pt(rng)));
  }

  unsigned signal_pixels = 0;
  for (unsigned v = 0; v < pvs.size(); ++v) {
    for (unsigned t = 0; t < pvs[v].n; ++t) {
      double eta = c.eta_min + (c.eta_max - c.eta_min) * uni(rng);
      if (c.backward_fraction > 0 && uni(rng) < c.backward_fraction) eta = -eta;
      const double phi = 2.0 * M_PI * uni(rng);
      const double theta = 2.0 * std::atan(std::exp(-eta));
      const double tr = std::tan(theta);
      const double dir = theta < M_PI / 2 ? 1.0 : -1.0;
      const double tx = tr * std::cos(phi), ty = tr * std::sin(phi);
      velo_synth_particle p {};
      p.event = ev;
      p.pv = v;
      p.x0 = pvs[v].x;
      p.y0 = pvs[v].y;
      p.z0 = pvs[v].z;
      p.tx = tx;
      p.ty = ty;
      p.eta = eta;
      p.phi = phi;
      p.first_module = 0xFFFFFFFFu;
      p.last_module = 0;
      {
        const uint64_t key = (c.seed * 0xD1342543DE82EF95ull) ^ (static_cast<uint64_t>(ev) << 32) ^ out.particles.size();
        const double pt = -500.0 * std::log(keyed_uniform(key));
        p.pt = static_cast<float>(pt);
        p.p = static_cast<float>(pt * std::cosh(eta));
      }
      const unsigned pidx = out.particles.size();
      std::vector<uint8_t> module_hit(n_modules, 0);
      for (unsigned sensor = 0; sensor < n_sensors; ++sensor) {
        const float* m = g.ltg + 12 * sensor;
        const double zs = m[11];
        if ((zs - pvs[v].z) * dir <= 0) continue;
        const double gx = pvs[v].x + tx * (zs - pvs[v].z);
        const double gy = pvs[v].y + ty * (zs - pvs[v].z);
        // Solve [r0 r1; r3 r4] (lx, ly) = (gx - t0, gy - t1).
        const double a = m[0], b = m[1], cc = m[3], d = m[4];
        const double det = a * d - b * cc;
        const double rx = gx - m[9], ry = gy - m[10];
        const double lx = (d * rx - b * ry) / det;
        const double ly = (-cc * rx + a * ry) / det;
        if (lx < g.col_edge[0] || lx >= g.col_edge[n_cols] || ly < 0 || ly >= n_rows * pixel_size) continue;
        if (uni(rng) >= c.hit_efficiency) continue;
        const unsigned col = std::upper_bound(g.col_edge, g.col_edge + n_cols + 1, lx) - g.col_edge - 1;
        const unsigned row = static_cast<unsigned>(ly / pixel_size);
        if (col >= n_cols || row >= n_rows) continue;
        // Cluster shape: 1, 2 or 3 pixels growing toward the sub-pixel position.
        const double u = uni(rng);
        const unsigned size = u < c.cluster_size_p[0] ? 1 : (u < c.cluster_size_p[0] + c.cluster_size_p[1] ? 2 : 3);
        const double fxp = (lx - g.col_edge[col]) / (g.col_edge[col + 1] - g.col_edge[col]);
        const double fyp = ly / pixel_size - row;
        const int dc = fxp >= 0.5 ? 1 : -1;
        const int dr = fyp >= 0.5 ? 1 : -1;
        Pixel px[3];
        unsigned np = 0;
        px[np++] = {static_cast<uint16_t>(col), static_cast<uint16_t>(row)};
        auto add = [&](int cc2, int rr2) {
          if (cc2 < 0 || cc2 >= (int) n_cols || rr2 < 0 || rr2 >= (int) n_rows) return;
          px[np++] = {static_cast<uint16_t>(cc2), static_cast<uint16_t>(rr2)};
        };
        if (size >= 2) {
          if (uni(rng) < 0.5) add(col + dc, row);
          else add(col, row + dr);
        }
        if (size >= 3) {
          if (px[np - 1].col != col) add(col, row + dr);
          else add(col + dc, row);
        }
        unsigned kept = 0, sc = 0, sr = 0;
        for (unsigned i = 0; i < np; ++i) {
          if (dead[sensor * 3 + (px[i].col >> 8)]) continue;
          sensor_pixels[sensor].push_back(px[i]);
          sc += px[i].col;
          sr += px[i].row;
          ++kept;
        }
        if (!kept) continue;
        signal_pixels += kept;
        velo_synth_hit h {};
        h.event = ev;
        h.particle = pidx;
        h.lhcb_id = channel_lhcb_id(sensor, sc / kept, sr / kept);
        if (c.retina) sensor_retina[sensor].push_back(retina_word(sensor, sc, sr, kept, &h.lhcb_id));
        h.sensor = sensor;
        h.n_pixels = kept;
        h.x = gx;
        h.y = gy;
        h.z = zs;
        out.hits.push_back(h);
        ++p.n_hits;
        const unsigned module = sensor / 4;
        if (!module_hit[module]) {
          module_hit[module] = 1;
          ++p.n_modules;
          p.first_module = std::min(p.first_module, module);
          p.last_module = std::max(p.last_module, module);
        }
      }
      out.particles.push_back(p);
    }
  }

  // Uniform noise pixels.
  const unsigned n_noise = static_cast<unsigned>(std::lround(c.noise_fraction * signal_pixels));
  for (unsigned i = 0; i < n_noise; ++i) {
    const unsigned sensor = static_cast<unsigned>(uni(rng) * n_sensors) % n_sensors;
    const unsigned col = static_cast<unsigned>(uni(rng) * n_cols) % n_cols;
    const unsigned row = static_cast<unsigned>(uni(rng) * n_rows) % n_rows;
    if (dead[sensor * 3 + (col >> 8)]) continue;
    sensor_pixels[sensor].push_back({static_cast<uint16_t>(col), static_cast<uint16_t>(row)});
    velo_synth_hit h {};
    h.event = ev;
    h.particle = 0xFFFFFFFFu;
    h.lhcb_id = channel_lhcb_id(sensor, col, row);
    if (c.retina) sensor_retina[sensor].push_back(retina_word(sensor, col, row, 1, &h.lhcb_id));
    h.sensor = sensor;
    h.n_pixels = 1;
    out.hits.push_back(h);
  }

  // Super-pixel encoding per sensor.
  std::vector<std::vector<uint32_t>> sensor_words(n_sensors);
  for (unsigned sensor = 0; sensor < n_sensors; ++sensor) {
    std::map<uint32_t, uint8_t> sps; // key = sp_col << 6 | sp_row
    for (const auto& px : sensor_pixels[sensor]) {
      const uint32_t key = ((px.col >> 1) << 6) | (px.row >> 2);
      sps[key] |= static_cast<uint8_t>(1u << (4 * (px.col & 1) + (px.row & 3)));
    }
    for (const auto& kv : sps) {
      const int spc = kv.first >> 6, spr = kv.first & 0x3F;
      bool isolated = true;
      for (int dcol = -1; dcol <= 1 && isolated; ++dcol)
        for (int drow = -1; drow <= 1; ++drow) {
          if (!dcol && !drow) continue;
          const int nc = spc + dcol, nr = spr + drow;
          if (nc < 0 || nr < 0 || nc >= 384 || nr >= 64) continue;
          if (sps.count((nc << 6) | nr)) {
            isolated = false;
            break;
          }
        }
      uint32_t w = (isolated ? 0x80000000u : 0u) | (static_cast<uint32_t>(spc) << 14) |
                   (static_cast<uint32_t>(spr) << 8) | kv.second;
      if (c.bank_version > 3) w |= (sensor & 1u) << 23;
      sensor_words[sensor].push_back(w);
    }
  }

  // Banks.
  std::vector<std::vector<uint32_t>> banks;
  std::vector<uint32_t> source_ids;
  uint8_t bank_type = 63;
  if (c.retina) {
    bank_type = 73;
    // Identical words of a sensor merged, in generation order.
    for (auto& w : sensor_retina) {
      std::vector<uint32_t> u;
      for (const uint32_t x : w)
        if (std::find(u.begin(), u.end(), x) == u.end()) u.push_back(x);
      if (c.shuffle_sp) std::shuffle(u.begin(), u.end(), rng);
      w = std::move(u);
    }
    if (c.bank_version > 3) {
      for (unsigned pair = 0; pair < n_sensors / 2; ++pair) {
        std::vector<uint32_t> w = sensor_retina[2 * pair];
        w.insert(w.end(), sensor_retina[2 * pair + 1].begin(), sensor_retina[2 * pair + 1].end());
        if (w.size() & 1u) w.push_back(0u);
        banks.push_back(std::move(w));
        source_ids.push_back(0x2000u | pair);
      }
    }
    else {
      for (unsigned sensor = 0; sensor < n_sensors; ++sensor) {
        std::vector<uint32_t> w = sensor_retina[sensor];
        w.insert(w.begin(), static_cast<uint32_t>(w.size()));
        banks.push_back(std::move(w));
        source_ids.push_back(sensor);
      }
    }
  }
  else if (c.bank_version > 3) {
    for (unsigned pair = 0; pair < n_sensors / 2; ++pair) {
      std::vector<uint32_t> w = sensor_words[2 * pair];
      w.insert(w.end(), sensor_words[2 * pair + 1].begin(), sensor_words[2 * pair + 1].end());
      if (c.shuffle_sp) std::shuffle(w.begin(), w.end(), rng);
      banks.push_back(std::move(w));
      source_ids.push_back(0x2000u | pair);
    }
  }
  else {
    for (unsigned sensor = 0; sensor < n_sensors; ++sensor) {
      std::vector<uint32_t> w = sensor_words[sensor];
      if (c.shuffle_sp) std::shuffle(w.begin(), w.end(), rng);
      w.insert(w.begin(), static_cast<uint32_t>(w.size()));
      banks.push_back(std::move(w));
      source_ids.push_back(sensor);
    }
  }
  const unsigned nb = banks.size();
  out.block.clear();
  out.block.push_back(nb);
  const size_t off_pos = out.block.size();
  out.block.resize(out.block.size() + nb + 1, 0);
  uint32_t offset = 0;
  for (unsigned b = 0; b < nb; ++b) {
    out.block[off_pos + b] = offset;
    out.block.push_back(source_ids[b]);
    out.block.insert(out.block.end(), banks[b].begin(), banks[b].end());
    offset += 4 * (1 + banks[b].size());
    out.bank_sizes.push_back(static_cast<uint16_t>(4 * banks[b].size()));
    out.bank_types.push_back(bank_type);
  }
  out.block[off_pos + nb] = offset;
}

struct Result {
  std::vector<char> raw;
  std::vector<uint32_t> offsets, sizes, types;
  std::vector<velo_synth_particle> particles;
  std::vector<velo_synth_hit> hits;
  std::vector<uint32_t> particle_offsets, hit_offsets;
};

} // namespace

extern "C" {

void velo_synth_default_config(velo_synth_config* c)
{
  std::memset(c, 0, sizeof(*c));
  c->n_events = 200;
  c->seed = 1;
  c->pileup = 0;
  c->mean_tracks = 120;
  c->nu = 7.6;
  c->nb_mean = 16;
  c->nb_k = 1.5;
  c->eta_min = 2;
  c->eta_max = 5;
  c->pv_sigma_xy = 0.04;
  c->pv_sigma_z = 45;
  c->hit_efficiency = 0.98;
  c->noise_fraction = 0.03;
  c->cluster_size_p[0] = 1.0 / 3;
  c->cluster_size_p[1] = 1.0 / 3;
  c->cluster_size_p[2] = 1.0 / 3;
  c->bank_version = 4;
}

void* velo_synth_generate(const velo_synth_config* c, const char* geometry, size_t geometry_bytes, unsigned n_threads)
{
  Geometry* gp = new Geometry;
  if (!load_geometry(geometry, geometry_bytes, *gp)) {
    delete gp;
    return nullptr;
  }
  if (c->retina ? (c->bank_version != 3 && c->bank_version != 4) : (c->bank_version != 2 && c->bank_version != 4)) {
    delete gp;
    return nullptr;
  }
  const unsigned n = c->n_events;
  std::vector<EventOut> events(n);
  if (n_threads == 0) n_threads = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < n_threads; ++t)
    th.emplace_back([&, t]() {
      for (unsigned e = t; e < n; e += n_threads) generate_event(*c, *gp, e, events[e]);
    });
  for (auto& t : th) t.join();
  delete gp;

  auto* r = new Result;
  size_t total_banks = 0;
  for (auto& e : events) total_banks += e.bank_sizes.size();
  r->offsets.resize(n + 1);
  r->sizes.assign((n + 1) + (total_banks + 1) / 2, 0);
  r->types.assign((n + 1) + (total_banks + 3) / 4, 0);
  size_t raw_bytes = 0;
  for (auto& e : events) raw_bytes += 4 * e.block.size();
  r->raw.resize(raw_bytes);
  uint16_t* sz16 = reinterpret_cast<uint16_t*>(r->sizes.data());
  uint8_t* ty8 = reinterpret_cast<uint8_t*>(r->types.data());
  size_t pos = 0, s16 = 2 * (n + 1), t8 = 4 * (n + 1);
  r->particle_offsets.push_back(0);
  r->hit_offsets.push_back(0);
  for (unsigned e = 0; e < n; ++e) {
    r->offsets[e] = pos;
    std::memcpy(r->raw.data() + pos, events[e].block.data(), 4 * events[e].block.size());
    pos += 4 * events[e].block.size();
    r->sizes[e] = s16;
    r->types[e] = t8;
    for (auto v : events[e].bank_sizes) sz16[s16++] = v;
    for (auto v : events[e].bank_types) ty8[t8++] = v;
    const uint32_t pbase = r->particles.size();
    r->particles.insert(r->particles.end(), events[e].particles.begin(), events[e].particles.end());
    for (auto h : events[e].hits) {
      if (h.particle != 0xFFFFFFFFu) h.particle += pbase;
      r->hits.push_back(h);
    }
    r->particle_offsets.push_back(r->particles.size());
    r->hit_offsets.push_back(r->hits.size());
  }
  r->offsets[n] = pos;
  r->sizes[n] = s16;
  r->types[n] = t8;
  return r;
}

int velo_synth_get(void* h, const char* name, const void** ptr, size_t* nbytes)
{
  auto* r = static_cast<Result*>(h);
  const std::string s(name);
  auto set = [&](const void* p, size_t b) {
    *ptr = p;
    *nbytes = b;
    return 0;
  };
  if (s == "raw") return set(r->raw.data(), r->raw.size());
  if (s == "offsets") return set(r->offsets.data(), 4 * r->offsets.size());
  if (s == "sizes") return set(r->sizes.data(), 4 * r->sizes.size());
  if (s == "types") return set(r->types.data(), 4 * r->types.size());
  if (s == "particles") return set(r->particles.data(), sizeof(velo_synth_particle) * r->particles.size());
  if (s == "hits") return set(r->hits.data(), sizeof(velo_synth_hit) * r->hits.size());
  if (s == "particle_offsets") return set(r->particle_offsets.data(), 4 * r->particle_offsets.size());
  if (s == "hit_offsets") return set(r->hit_offsets.data(), 4 * r->hit_offsets.size());
  return 1;
}

void velo_synth_free(void* h) { delete static_cast<Result*>(h); }

} // extern "C"
