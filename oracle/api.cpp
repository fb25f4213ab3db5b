// TEST INFRASTRUCTURE — C ABI of the CPU oracle (see common.hpp for the rules of use).
// Runs the VELO-SP + Kalman sequence in the order wired by
// configuration/python/AllenConf/velo_reconstruction.py:72-321, with host prefix sums
// between the stages exactly where the reference places them.
#include "oracle.hpp"
#include "velo_oracle.h"
#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace vo {

bool load_geometry(const char* p0, size_t n, Geometry& g)
{
  // ClusteringDefinitions.cuh:164-197: [u64 n][f32 x n] x3, then [u64 n_ltg][u64 n_trans][f32 ...]
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
  if (!arr(n_modules, g.module_zs) || !arr(n_sensor_columns, g.local_x) || !arr(n_sensor_columns, g.x_pitch))
    return false;
  uint64_t n_ltg, n_trans;
  std::memcpy(&n_ltg, p, 8);
  p += 8;
  std::memcpy(&n_trans, p, 8);
  p += 8;
  if (n_ltg != n_sensors || n_trans != 12) return false;
  std::memcpy(g.ltg, p, 4 * n_ltg * n_trans);
  p += 4 * n_ltg * n_trans;
  return static_cast<size_t>(p - p0) == n;
}

struct Sequence {
  std::map<std::string, std::vector<char>> bufs;
  uint32_t status = 0;
  SearchStats stats;

  template<typename T>
  T* make(const std::string& name, size_t n)
  {
    auto& v = bufs[name];
    v.assign(n * sizeof(T) + 8, 0);
    v.resize(n * sizeof(T));
    return reinterpret_cast<T*>(v.data());
  }
  template<typename T>
  T* get(const std::string& name)
  {
    return reinterpret_cast<T*>(bufs[name].data());
  }
};

struct PrebuiltClusters {
  const char* sorted = nullptr;
  const uint32_t* offsets = nullptr;
  const uint32_t* counts = nullptr;
};

struct Params {
  SearchParams search;
  PrebuiltClusters clusters;
  float max_chi2 = 20.f;
  unsigned block_dim_x = 256;
  bool f64 = true;
  bool kalman_fp64_fallback = false;
  bool retina = false;
  bool pv_finder = false;
};

void run_sequence(
  Sequence& s,
  const RawInput& in,
  unsigned n_events,
  unsigned event_start,
  const Geometry& g,
  const SpTables& sp,
  const float* beamline,
  const Params& prm,
  SeedSink* sink,
  const uint32_t* list = nullptr,
  unsigned list_size = 0)
{
  // MASK_INPUT dev_event_list: the events the kernels visit, in list order (the identity when
  // none is given). Buffers span all n_events; what an unlisted event owns stays as the
  // reference's memsets leave it (every buffer here starts zeroed).
  std::vector<uint32_t> event_list(n_events);
  std::iota(event_list.begin(), event_list.end(), 0u);
  const uint32_t* el = list ? list : event_list.data();
  const unsigned nl = list ? list_size : n_events;

  uint32_t total_clusters = 0;
  uint32_t* est_off = nullptr;
  uint32_t* mcnum = nullptr;
  char* sorted = nullptr;
  if (prm.clusters.sorted) {
    // Pre-built sorted clusters: the sequence starts at velo_search_by_triplet.
    mcnum = s.make<uint32_t>("module_cluster_num", n_events * n_module_pairs);
    std::copy(prm.clusters.counts, prm.clusters.counts + (size_t) n_events * n_module_pairs, mcnum);
    est_off = s.make<uint32_t>("offsets_estimated_input_size", n_events * n_module_pairs + 1);
    std::copy(prm.clusters.offsets, prm.clusters.offsets + (size_t) n_events * n_module_pairs + 1, est_off);
    total_clusters = est_off[(size_t) n_events * n_module_pairs];
    sorted = s.make<char>("sorted_velo_cluster_container", (size_t) total_clusters * cluster_element_size);
    std::copy(prm.clusters.sorted, prm.clusters.sorted + (size_t) total_clusters * cluster_element_size, sorted);
  }
  else if (prm.retina) {
    // velo_reconstruction.py:18-71: calculate_number_of_retinaclusters_each_sensor_pair,
    // prefix_sum_offsets_estimated_input_size, decode_retinaclusters
    const unsigned slots = retina_sensor_slots(in.version);
    auto* size = s.make<uint32_t>("each_sensor_pair_size", (size_t) n_events * slots);
    auto* bank_index = s.make<uint32_t>("retina_bank_index", slots);
    calculate_number_of_retinaclusters_each_sensor_pair(in, el, nl, event_start, size, bank_index);
    auto* off = s.make<uint32_t>("offsets_each_sensor_pair_size", (size_t) n_events * slots + 1);
    total_clusters = prefix_sum(size, (size_t) n_events * slots, off);
    mcnum = s.make<uint32_t>("module_cluster_num", n_events * n_module_pairs);
    est_off = s.make<uint32_t>("offsets_estimated_input_size", n_events * n_module_pairs + 1);
    sorted = s.make<char>("sorted_velo_cluster_container", (size_t) total_clusters * cluster_element_size);
    auto* perm = s.make<uint32_t>("hit_permutations", total_clusters);
    auto* keys = s.make<int64_t>("hit_sorting_key", total_clusters);
    decode_retinaclusters(
      in, el, nl, n_events, event_start, g, off, bank_index, mcnum, est_off, sorted, perm, keys);
  }
  else {
    // velo_calculate_number_of_candidates + prefix_sum_offsets_velo_candidates
    auto* ncand = s.make<uint32_t>("number_of_candidates", n_events);
    auto* bank_index = s.make<uint32_t>("velo_bank_index", n_sensors);
    calculate_number_of_candidates(in, el, nl, event_start, prm.block_dim_x, ncand, bank_index);
    auto* cand_off = s.make<uint32_t>("candidates_offsets", n_events + 1);
    const uint32_t total_candidates = prefix_sum(ncand, n_events, cand_off);

    // velo_estimate_input_size + prefix_sum_offsets_estimated_input_size
    auto* est = s.make<uint32_t>("estimated_input_size", n_events * n_module_pairs);
    auto* mcn = s.make<uint32_t>("module_candidate_num", n_events);
    auto* cands = s.make<uint32_t>("cluster_candidates", total_candidates);
    estimate_input_size(in, el, nl, event_start, cand_off, est, mcn, cands);
    est_off = s.make<uint32_t>("offsets_estimated_input_size", n_events * n_module_pairs + 1);
    total_clusters = prefix_sum(est, n_events * n_module_pairs, est_off);

    // velo_masked_clustering
    mcnum = s.make<uint32_t>("module_cluster_num", n_events * n_module_pairs);
    auto* container = s.make<char>("velo_cluster_container", (size_t) total_clusters * cluster_element_size);
    masked_clustering(
      in, el, nl, n_events, event_start, g, sp, est_off, mcn, cands, cand_off, bank_index, mcnum, container);

    // velo_sort_by_phi
    auto* perm = s.make<uint32_t>("hit_permutation", total_clusters);
    sorted = s.make<char>("sorted_velo_cluster_container", (size_t) total_clusters * cluster_element_size);
    sort_by_phi(el, nl, n_events, est_off, mcnum, container, sorted, perm);
  }

  // velo_search_by_triplet
  const unsigned tcs = track_container_size(total_clusters, n_events);
  auto* tracks = s.make<TrackHits>("tracks", tcs);
  auto* tracklets = s.make<TrackletHits>("tracklets", (size_t) n_events * max_tracks_to_follow);
  auto* ttf = s.make<uint32_t>("tracks_to_follow", (size_t) n_events * max_tracks_to_follow);
  auto* three = s.make<TrackletHits>("three_hit_tracks", tcs);
  auto* used = s.make<bool>("hit_used", total_clusters);
  auto* atomics = s.make<uint32_t>("atomics_velo", n_events * num_atomics);
  auto* rel = s.make<uint16_t>("rel_indices", total_clusters);
  auto* nvt = s.make<uint32_t>("number_of_velo_tracks", n_events);
  search_by_triplet(
    el, nl, n_events, sorted, est_off, mcnum, g, tracks, tracklets, ttf, three, used, atomics, rel, nvt, prm.search,
    s.stats, s.status, sink);

  // prefix_sum_offsets_velo_tracks
  auto* off_tracks = s.make<uint32_t>("offsets_velo_tracks", n_events + 1);
  const uint32_t n4 = prefix_sum(nvt, n_events, off_tracks);

  // velo_three_hit_tracks_filter + prefix sum
  auto* three_out = s.make<TrackletHits>("three_hit_tracks_output", tcs);
  auto* n_three_out = s.make<uint32_t>("number_of_three_hit_tracks_output", n_events);
  three_hit_tracks_filter(el, nl, n_events, sorted, est_off, three, atomics, used, prm.max_chi2, three_out, n_three_out);
  auto* off_three = s.make<uint32_t>("offsets_three_hit_tracks_filtered", n_events + 1);
  const uint32_t n3 = prefix_sum(n_three_out, n_events, off_three);

  // velo_copy_track_hit_number + prefix_sum_offsets_velo_track_hit_number
  auto* thn = s.make<uint32_t>("velo_track_hit_number", n4 + n3);
  auto* off_all = s.make<uint32_t>("offsets_all_velo_tracks", n_events + 1);
  copy_track_hit_number(n_events, est_off, tracks, off_tracks, off_three, thn, off_all);
  auto* off_hits = s.make<uint32_t>("offsets_velo_track_hit_number", n4 + n3 + 1);
  const uint32_t total_hits = prefix_sum(thn, n4 + n3, off_hits);

  // velo_consolidate_tracks
  auto* track_hits = s.make<char>("velo_track_hits", (size_t) total_hits * cluster_element_size);
  consolidate_tracks(el, nl, n_events, est_off, tracks, three_out, off_three, off_all, off_hits, sorted, track_hits);

  // velo_kalman_filter
  const unsigned nt = n4 + n3;
  auto* bwd = s.make<bool>("is_backward", nt);
  auto* bl = s.make<float>("velo_kalman_beamline_states", (size_t) nt * 14);
  auto* ev = s.make<float>("velo_kalman_endvelo_states", (size_t) nt * 14);
  double* bl64 = prm.f64 ? s.make<double>("velo_kalman_beamline_states_f64", (size_t) nt * 14) : nullptr;
  double* ev64 = prm.f64 ? s.make<double>("velo_kalman_endvelo_states_f64", (size_t) nt * 14) : nullptr;
  kalman_filter(el, nl, n_events, off_all, off_hits, track_hits, beamline, bwd, bl, ev, bl64, ev64, prm.kalman_fp64_fallback);

  if (prm.pv_finder) {
    // primary_vertex_reconstruction.py: pv_beamline_extrapolate ... pv_beamline_cleanup
    auto* pvtracks = s.make<PVTrack>("pvtracks", nt);
    auto* zhisto = s.make<float>("zhisto", (size_t) pv_nbins * n_events);
    auto* zpeaks = s.make<float>("zpeaks", (size_t) pv_max_vertices * n_events);
    auto* nzp = s.make<uint32_t>("number_of_zpeaks", n_events);
    auto* denom = s.make<float>("pvtracks_denom", nt);
    auto* fit = s.make<PVVertex>("multi_fit_vertices", (size_t) pv_max_vertices * n_events);
    auto* nfit = s.make<uint32_t>("number_of_multi_fit_vertices", n_events);
    auto* fin = s.make<PVVertex>("multi_final_vertices", (size_t) pv_max_vertices * n_events);
    auto* nfin = s.make<uint32_t>("number_of_multi_final_vertices", n_events);
    pv_beamline(n_events, off_all, bl, beamline, pvtracks, zhisto, zpeaks, nzp, denom, fit, nfit, fin, nfin, s.status);
  }

  auto* st = s.make<uint64_t>("search_stats", 4);
  st[0] = s.stats.seed_evaluations;
  st[1] = s.stats.forward_evaluations;
  st[2] = s.stats.extrapolations;
  st[3] = s.stats.search_steps;
  *s.make<uint32_t>("status", 1) = s.status;
}

} // namespace vo

namespace {
struct Handle {
  vo::Sequence seq;
};

// Every seed written by seeding: (event, module pair of h1, h0, h1, h2) and its features.
struct SeedLog : vo::SeedSink {
  unsigned event = 0;
  std::vector<uint32_t> records;
  std::vector<float> features;
  void begin_event(unsigned e) override { event = e; }
  void seed(unsigned module_pair, unsigned h0, unsigned h1, unsigned h2, const float* f) override
  {
    records.insert(records.end(), {event, module_pair, h0, h1, h2});
    features.insert(features.end(), f, f + vo::n_seed_features);
  }
};

bool make_inputs(const vo_inputs* a, vo::RawInput& in, vo::Geometry& g, vo::SpTables& sp, vo::Params& p)
{
  in.raw = a->raw;
  in.offsets = a->offsets;
  in.sizes = a->sizes;
  in.types = a->types;
  in.mep = a->mep_layout != 0;
  in.version = a->bank_version;
  if (!vo::load_geometry(a->geometry, a->geometry_bytes, g)) return false;
  vo::cache_sp_patterns(sp);
  p.search.phi_tolerance = a->phi_tolerance;
  p.search.max_scatter = a->max_scatter;
  p.search.max_skipped_modules = a->max_skipped_modules;
  p.max_chi2 = a->max_chi2;
  p.block_dim_x = a->block_dim_x ? a->block_dim_x : 256;
  p.f64 = a->compute_f64 != 0;
  p.kalman_fp64_fallback = a->kalman_fp64_fallback != 0;
  p.retina = a->retina != 0;
  p.pv_finder = a->pv_finder != 0;
  p.search.vectorized_seeding = a->vectorized_seeding != 0;
  p.search.seed_logits = a->seed_logits;
  p.search.seed_threshold = a->seed_threshold;
  p.search.seed_block = a->seed_block ? a->seed_block : 256;
  if (a->seed_logits && (a->vectorized_seeding || p.search.seed_block % 8 != 0)) return false;
  p.clusters.sorted = a->sorted_clusters;
  p.clusters.offsets = a->clusters_offsets;
  p.clusters.counts = a->clusters_counts;
  if (p.clusters.sorted && (!p.clusters.offsets || !p.clusters.counts)) return false;
  return true;
}
} // namespace

extern "C" {

void vo_default_inputs(vo_inputs* a)
{
  std::memset(a, 0, sizeof(*a));
  a->bank_version = 4;
  a->phi_tolerance = 0.045f;
  a->max_scatter = 0.08f;
  a->max_skipped_modules = 1;
  a->max_chi2 = 20.f;
  a->block_dim_x = 256;
  a->compute_f64 = 1;
}

void* vo_run(const vo_inputs* a)
{
  if (!a->sorted_clusters && a->bank_version != 2 && a->bank_version != 3 && a->bank_version != 4) return nullptr;
  vo::RawInput in;
  vo::Geometry g;
  vo::SpTables sp;
  vo::Params p;
  if (!make_inputs(a, in, g, sp, p)) return nullptr;
  auto* h = new Handle;
  SeedLog log;
  if (a->event_list)
    for (unsigned i = 0; i < a->event_list_size; ++i)
      if (a->event_list[i] >= a->n_events) {
        delete h;
        return nullptr;
      }
  vo::run_sequence(
    h->seq, in, a->n_events, a->event_start, g, sp, a->beamline, p, a->collect_seeds ? &log : nullptr,
    a->event_list, a->event_list_size);
  if (a->collect_seeds) {
    auto* rec = h->seq.make<uint32_t>("seed_log", log.records.size());
    std::copy(log.records.begin(), log.records.end(), rec);
    auto* feat = h->seq.make<float>("seed_features", log.features.size());
    std::copy(log.features.begin(), log.features.end(), feat);
  }
  return h;
}

int vo_get(void* handle, const char* name, const void** ptr, size_t* nbytes)
{
  auto* h = static_cast<Handle*>(handle);
  auto it = h->seq.bufs.find(name);
  if (it == h->seq.bufs.end()) return 1;
  *ptr = it->second.data();
  *nbytes = it->second.size();
  return 0;
}

void vo_free(void* handle) { delete static_cast<Handle*>(handle); }

double vo_throughput(const vo_inputs* a, unsigned n_threads, unsigned repetitions, double* seconds)
{
  vo::RawInput in;
  vo::Geometry g;
  vo::SpTables sp;
  vo::Params p;
  if (!make_inputs(a, in, g, sp, p)) return -1.0;
  p.f64 = false;
  if (n_threads == 0) n_threads = 1;
  std::atomic<unsigned> next {0};
  const unsigned total = a->n_events * (repetitions ? repetitions : 1);
  auto worker = [&]() {
    vo::Sequence s;
    for (;;) {
      const unsigned k = next.fetch_add(1);
      if (k >= total) break;
      vo::run_sequence(s, in, 1, a->event_start + (k % a->n_events), g, sp, a->beamline, p, nullptr);
    }
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> threads;
  for (unsigned t = 0; t < n_threads; ++t) threads.emplace_back(worker);
  for (auto& t : threads) t.join();
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (seconds) *seconds = dt;
  return total / dt;
}

void vo_cache_sp_patterns(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy)
{
  vo::SpTables t;
  vo::cache_sp_patterns(t);
  std::memcpy(patterns, t.patterns, 256);
  std::memcpy(sizes, t.sizes, 256);
  std::memcpy(fx, t.fx, sizeof(t.fx));
  std::memcpy(fy, t.fy, sizeof(t.fy));
}

float vo_fast_atan2f(float y, float x) { return vo::fast_atan2f(y, x); }
int16_t vo_hit_phi_16(float x, float y) { return vo::hit_phi_16(x, y); }
int16_t vo_hit_phi_float_to_16(float phi) { return vo::hit_phi_float_to_16(phi); }

uint32_t vo_prefix_sum(const uint32_t* input, size_t n, uint32_t* output) { return vo::prefix_sum(input, n, output); }

int vo_binary_search_leftmost_i16(const int16_t* array, unsigned n, int16_t value)
{
  return vo::binary_search_leftmost(array, n, value);
}

uint64_t vo_make_cluster(uint64_t pixel_map, uint64_t start, unsigned col_lower, unsigned row_lower, uint32_t* xyn)
{
  const vo::ClusterGrow c = vo::make_cluster(pixel_map, start, col_lower, row_lower);
  xyn[0] = c.x;
  xyn[1] = c.y;
  xyn[2] = c.n;
  return c.pixels;
}

float vo_three_hit_chi2(const float* xyz9)
{
  char buf[18 * 3];
  vo::Clusters c {buf, 3, 0};
  for (unsigned i = 0; i < 3; ++i) c.set(i, 0, xyz9[3 * i], xyz9[3 * i + 1], xyz9[3 * i + 2], 0);
  vo::TrackletHits t {{0, 1, 2}};
  return vo::means_square_fit_chi2(c, t);
}

int vo_load_geometry(const char* bytes, size_t n, float* module_zs, float* local_x, float* x_pitch, float* ltg)
{
  vo::Geometry g;
  if (!vo::load_geometry(bytes, n, g)) return 1;
  std::memcpy(module_zs, g.module_zs, sizeof(g.module_zs));
  std::memcpy(local_x, g.local_x, sizeof(g.local_x));
  std::memcpy(x_pitch, g.x_pitch, sizeof(g.x_pitch));
  std::memcpy(ltg, g.ltg, sizeof(g.ltg));
  return 0;
}

} // extern "C"
