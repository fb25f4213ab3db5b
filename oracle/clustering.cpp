// TEST INFRASTRUCTURE — CPU oracle (see common.hpp for the rules of use).
// Restates, with CPU-target semantics:
//   cache_sp_patterns        host/velo/clustering/src/ClusteringCommon.cpp:6-104
//   K1 number of candidates  device/velo/mask_clustering/src/VeloCalculateNumberOfCandidates.cu:10-48
//   K2 estimate input size   device/velo/mask_clustering/src/EstimateInputSize.cu:10-239
//   K3 masked clustering     device/velo/mask_clustering/src/MaskedVeloClustering.cu:21-576
#include "oracle.hpp"

namespace vo {

// ClusteringCommon.cpp:6-104 — 8-connected flood fill inside one 2x4 super pixel.
void cache_sp_patterns(SpTables& t)
{
  std::memset(&t, 0, sizeof(t));
  const int dx[] = {-1, 0, 1, -1, 0, 1, -1, 1};
  const int dy[] = {-1, -1, -1, 1, 1, 1, 0, 0};
  for (unsigned sp = 0; sp < 256; ++sp) {
    unsigned char buffer[8];
    unsigned char idxs[8];
    unsigned n_idx = 0;
    for (unsigned s = 0; s < 8; ++s) {
      buffer[s] = sp & (1u << s);
      if (buffer[s]) idxs[n_idx++] = s;
    }
    unsigned clu = 0;
    for (unsigned ip = 0; ip < n_idx; ++ip) {
      unsigned char idx = idxs[ip];
      if (!buffer[idx]) continue;
      unsigned char stack[8];
      unsigned sptr = 0;
      stack[sptr++] = idx;
      buffer[idx] = 0;
      unsigned char x = 0, y = 0, n = 0;
      while (sptr) {
        idx = stack[--sptr];
        const int row = idx % 4, col = idx / 4;
        x += col;
        y += row;
        ++n;
        for (int k = 0; k < 8; ++k) {
          const int nc = col + dx[k];
          if (nc < 0 || nc > 1) continue;
          const int nr = row + dy[k];
          if (nr < 0 || nr > 3) continue;
          const unsigned nidx = (nc << 2) | nr;
          if (!buffer[nidx]) continue;
          stack[sptr++] = nidx;
          buffer[nidx] = 0;
        }
      }
      const uint32_t cx = x / n;
      const uint32_t cy = y / n;
      const float fx = x / static_cast<float>(n) - cx;
      const float fy = y / static_cast<float>(n) - cy;
      t.sizes[sp] |= n << (4 * clu);
      t.patterns[sp] |= ((cx << 2) | cy) << (4 * clu);
      t.patterns[sp] |= clu << 3;
      t.fx[2 * sp + clu] = fx;
      t.fy[2 * sp + clu] = fy;
      ++clu;
    }
  }
}

// VeloCalculateNumberOfCandidates.cu:10-48; CPU target: grid = ceil(n/block_dim_x) blocks of one thread.
void calculate_number_of_candidates(
  const RawInput& in,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  unsigned block_dim_x,
  uint32_t* number_of_candidates,
  uint32_t* bank_index)
{
  const unsigned grid = (n_list + block_dim_x - 1) / block_dim_x;
  for (unsigned block = 0; block < grid; ++block) {
    for (unsigned ev_idx = block; ev_idx < n_list; ev_idx += grid) {
      const unsigned event_number = event_list[ev_idx];
      RawEvent ev(in, event_number + event_start);
      unsigned n = 0;
      for (unsigned b = 0; b < ev.n_banks; ++b) {
        const RawBank bank = ev.bank(b);
        if (is_velo_bank(bank.type)) n += bank.n_sp(in.version);
        if (block == 0) {
          if (in.version > 3) {
            bank_index[bank.sensor_index0()] = b;
            bank_index[bank.sensor_index1()] = b;
          }
          else {
            bank_index[bank.sensor_pair()] = b;
          }
        }
      }
      number_of_candidates[event_number] = 2 * n;
    }
  }
}

// EstimateInputSize.cu:10-212 (estimate_raw_bank_size), one bank, sequential over SPs.
static void estimate_raw_bank_size(
  uint32_t* estimated_input_size,
  uint32_t* cluster_candidates,
  uint32_t* event_candidate_num,
  const RawBank& bank,
  int version)
{
  uint32_t* module_pair_size;
  uint32_t n_sp;
  if (version == 2 || version == 3) {
    module_pair_size = estimated_input_size + (bank.sensor_pair() / 8);
    n_sp = bank.count;
  }
  else {
    module_pair_size = estimated_input_size + (bank.sensor_index0() / 8);
    n_sp = bank.size / 4;
  }
  unsigned found = 0;
  for (unsigned sp_index = 0; sp_index < n_sp; ++sp_index) {
    const uint32_t sp_word = bank.word[sp_index];
    const uint32_t no_sp_neighbours = sp_word & 0x80000000u;
    const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
    const uint8_t sp = sp_word & 0xFFu;
    const uint32_t sensor_number =
      (version == 2 || version == 3) ? bank.sensor_pair() : (bank.sensor_index0() | ((sp_word >> 23) & 0x1u));
    if (no_sp_neighbours) {
      const bool pattern_0 = (sp & 0x88) && !(sp & 0x44) && (sp & 0x33);
      const bool pattern_1 = (sp & 0xCC) && !(sp & 0x22) && (sp & 0x11);
      *module_pair_size += (pattern_0 | pattern_1) ? 2 : 1;
    }
    else {
      uint32_t pixels = (sp & 0x0F) | ((sp & 0xF0) << 2);
      const uint32_t sp_row = sp_addr & 0x3Fu;
      const uint32_t sp_col = sp_addr >> 6;
      for (unsigned k = 0; k < n_sp; ++k) {
        const uint32_t other = bank.word[k];
        // EstimateInputSize.cu:113 tests the *current* word's isolation flag, which is
        // always clear on this branch, so every other SP of the bank (either sensor) is scanned.
#if VO_MUTANT == 1
        const uint32_t other_no_sp_neighbours = other & 0x80000000u;
#else
        const uint32_t other_no_sp_neighbours = sp_word & 0x80000000u;
#endif
#if VO_MUTANT == 2
        if (version > 3 && ((other >> 23) & 1u) != ((sp_word >> 23) & 1u)) continue;
#endif
        if (!other_no_sp_neighbours) {
          const uint32_t other_addr = (other & 0x007FFF00u) >> 8;
          const uint32_t orow = other_addr & 0x3Fu;
          const uint32_t ocol = other_addr >> 6;
          const uint8_t osp = other & 0xFFu;
          const bool is_top = orow == (sp_row + 1) && ocol == sp_col;
          const bool is_top_right = orow == (sp_row + 1) && ocol == (sp_col + 1);
          const bool is_right = orow == sp_row && ocol == (sp_col + 1);
          const bool is_right_bottom = orow == (sp_row - 1) && ocol == (sp_col + 1);
          const bool is_bottom = orow == (sp_row - 1) && ocol == sp_col;
          if (is_top || is_top_right || is_right || is_right_bottom || is_bottom) {
            pixels |= is_top * (((osp & 0x01) | ((osp & 0x10) << 2)) << 4);
            pixels |= is_top_right * ((osp & 0x01) << 16);
            pixels |= is_right * ((osp & 0x0F) << 12);
            pixels |= is_right_bottom * ((osp & 0x08) << 8);
            pixels |= is_bottom * ((osp & 0x80) >> 2);
          }
        }
      }
      const uint32_t inside = pixels & 0x3CF;
      const uint32_t mask = (inside << 1) | (inside << 5) | (inside << 6) | (inside << 7);
      const uint32_t working = mask & (~pixels);
      const uint32_t cand_tmp = (working >> 1) & (working >> 5) & (working >> 6) & (working >> 7);
      const uint32_t cand = cand_tmp & pixels;
      const uint32_t consolidated = (cand & 0x0F) | ((cand >> 2) & 0xF0);
      const uint32_t first = consolidated & 0x33;
      const uint32_t second = consolidated & 0xCC;
      if (first) {
#if VO_MUTANT == 3
        const uint32_t pixel = 31 - __builtin_clz(first);
#else
        const uint32_t pixel = __builtin_clz(first) - 24;
#endif
        const uint32_t c = (*event_candidate_num)++;
        cluster_candidates[c] = (sp_index << 11) | (sensor_number << 3) | pixel;
        ++found;
      }
      if (second) {
        const uint32_t pixel = __builtin_clz(second) - 24;
        const uint32_t c = (*event_candidate_num)++;
        cluster_candidates[c] = (sp_index << 11) | (sensor_number << 3) | pixel;
        ++found;
      }
    }
  }
  if (found > 0) *module_pair_size += found;
}

// EstimateInputSize.cu:215-239; blockDim == 1 so banks and SPs are walked in order.
void estimate_input_size(
  const RawInput& in,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  const uint32_t* candidates_offsets,
  uint32_t* estimated_input_size,
  uint32_t* module_candidate_num,
  uint32_t* cluster_candidates)
{
  for (unsigned block = 0; block < n_list; ++block) {
    const unsigned event_number = event_list[block];
    uint32_t* est = estimated_input_size + event_number * n_module_pairs;
    uint32_t* cnum = module_candidate_num + event_number;
    uint32_t* cands = cluster_candidates + candidates_offsets[event_number];
    RawEvent ev(in, event_number + event_start);
    for (unsigned b = 0; b < ev.n_banks; ++b) {
      const RawBank bank = ev.bank(b);
      if (!is_velo_bank(bank.type)) continue;
      estimate_raw_bank_size(est, cands, cnum, bank, in.version);
    }
  }
}

// MaskedVeloClustering.cu:21-26
static inline uint64_t make_8con_mask(uint64_t c)
{
#if VO_MUTANT == 6
  return c | ((c & 0x7FFF7FFF7FFF7FFFull) << 1) | ((c & 0xFFFEFFFEFFFEFFFEull) << 15) | (c << 16) |
         ((c & 0x7FFF7FFF7FFF7FFFull) << 17) | ((c & 0xFFFEFFFEFFFEFFFEull) >> 1) |
         ((c & 0x7FFF7FFF7FFF7FFFull) >> 15) | (c >> 16) | ((c & 0xFFFEFFFEFFFEFFFEull) >> 17);
#else
  return c | ((c & 0x7FFF7FFF7FFF7FFFull) << 1) | ((c & 0xFFFEFFFEFFFEFFFEull) << 15) | (c << 16) |
         ((c & 0xFFFCFFFCFFFCFFFCull) << 17) | ((c & 0xFFFEFFFEFFFEFFFEull) >> 1) |
         ((c & 0x7FFF7FFF7FFF7FFFull) >> 15) | (c >> 16) | ((c & 0x3FFF3FFF3FFF3FFFull) >> 17);
#endif
}
// MaskedVeloClustering.cu:31-44
static inline uint64_t mask_east(uint64_t c)
{
  const uint64_t m = c >> 48;
  return m | ((m & 0x7FFF7FFF7FFF7FFFull) << 1) | (m >> 1);
}
static inline uint64_t mask_west(uint64_t c)
{
  const uint64_t m = c << 48;
  return m | (m << 1) | ((m & 0xFFFEFFFEFFFEFFFEull) >> 1);
}
static inline unsigned popc64(uint64_t v) { return __builtin_popcountll(v); }

// MaskedVeloClustering.cu:49-84
ClusterGrow make_cluster(uint64_t pixel_map, uint64_t start, unsigned col_lower, unsigned row_lower)
{
  uint64_t cur = 0, next = start;
  while (cur != next) {
    cur = next;
    next = pixel_map & make_8con_mask(cur);
  }
  const unsigned n = popc64(cur);
  const unsigned x = col_lower * n + popc64(cur & 0x00000000FFFF0000ull) + popc64(cur & 0x0000FFFF00000000ull) * 2 +
                     popc64(cur & 0xFFFF000000000000ull) * 3;
  unsigned y = row_lower * n;
  for (unsigned r = 1; r < 15; ++r) {
    y += popc64(cur & (0x0001000100010001ull << r)) * r;
  }
  return {cur, x, y, n};
}

static inline void write_cluster(
  const Clusters& c,
  unsigned index,
  uint32_t id,
  const float* ltg,
  float local_x,
  float local_y)
{
  const float gx = ltg[0] * local_x + ltg[1] * local_y + ltg[9];
  const float gy = ltg[3] * local_x + ltg[4] * local_y + ltg[10];
  const float gz = ltg[6] * local_x + ltg[7] * local_y + ltg[11];
  c.set(index, id, gx, gy, gz, hit_phi_16(gx, gy));
}

// MaskedVeloClustering.cu:131-288 (no_neighbour_sp), one bank.
static void no_neighbour_sp(
  const SpTables& sp_tables,
  const Clusters& clusters,
  uint32_t* module_pair_cluster_num,
  const Geometry& g,
  unsigned module_pair_number,
  unsigned cluster_start,
  const RawBank& bank,
  int version)
{
  const bool v23 = (version == 2 || version == 3);
  const uint32_t n_sp = bank.n_sp(version);
  for (unsigned sp_index = 0; sp_index < n_sp; ++sp_index) {
    const uint32_t sp_word = bank.word[sp_index];
    const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
    if (!(sp_word & 0x80000000u)) continue;
    const bool sensor_bit = (sp_word & 0x800000u) != 0;
    const uint32_t sensor_index = v23 ? bank.sourceID : (sensor_bit ? bank.sensor_index1() : bank.sensor_index0());
    const float* ltg = g.ltg + 12 * (v23 ? bank.sensor_pair() : (sensor_bit ? bank.sensor_index1() : bank.sensor_index0()));
    const int32_t sp_row = sp_addr & 0x3Fu;
    const int32_t sp_col = sp_addr >> 6;
    const uint8_t sp = sp_word & 0xFFu;
    const uint32_t idx = sp_tables.patterns[sp];
    const uint32_t chip = sp_col >> 7;
    for (unsigned k = 0; k < 2; ++k) {
      if (k == 1 && !(idx & 8)) break;
      const uint32_t row = k == 0 ? (idx & 0x03u) : ((idx >> 4) & 3u);
      const uint32_t col = k == 0 ? ((idx >> 2) & 1u) : ((idx >> 6) & 1u);
      const uint32_t cx = sp_col * 2 + col;
      const uint32_t cy = sp_row * 4 + row;
      const uint32_t cid = get_channel_id(sensor_index, chip, cx & 0xFFu, cy);
      const float fx = sp_tables.fx[sp * 2 + k];
      const float fy = sp_tables.fy[sp * 2 + k];
      const float local_x = g.local_x[cx] + fx * g.x_pitch[cx];
      const float local_y = (cy + 0.5f + fy) * pixel_size;
      const unsigned cluster_num = module_pair_cluster_num[module_pair_number]++;
      write_cluster(clusters, cluster_start + cluster_num, get_lhcb_id(cid), ltg, local_x, local_y);
    }
  }
}

// MaskedVeloClustering.cu:290-496 (rest_of_clusters), one candidate.
static void rest_of_clusters(
  const uint32_t* module_pair_cluster_start,
  const Clusters& clusters,
  uint32_t* module_pair_cluster_num,
  const Geometry& g,
  uint32_t candidate,
  const RawBank& bank,
  int version)
{
  const uint32_t sp_index = candidate >> 11;
  const uint32_t sensor_number = (candidate >> 3) & 0xFFu;
  const uint32_t module_pair_number = sensor_number / 8;
  const uint32_t starting_pixel_location = candidate & 0x7u;
  const float* ltg = g.ltg + 12 * sensor_number;
  const uint32_t n_sp = bank.n_sp(version);
  const uint32_t sp_word = bank.word[sp_index];
  const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
  const int32_t sp_row = sp_addr & 0x3Fu;
  const int32_t sp_col = sp_addr >> 6;
  uint32_t pixel_array[5] = {0, 0, 0, 0, 0};
  const int32_t sp_row_lower = sp_row - 2, sp_row_upper = sp_row + 1;
  const int32_t sp_col_lower = sp_col - 3, sp_col_upper = sp_col + 1;
  const int32_t row_lower_limit = sp_row_lower * 4;
  const int32_t col_lower_limit = sp_col_lower * 2;
  for (unsigned k = 0; k < n_sp; ++k) {
    const uint32_t other = bank.word[k];
    if (version > 3) {
      const uint32_t other_sensor_bit = (other & 0x800000u) ? 1u : 0u;
      if (other_sensor_bit != (sensor_number & 0x1u)) continue;
    }
#if VO_MUTANT != 8
    if (other & 0x80000000u) continue;
#endif
    const uint32_t oaddr = (other & 0x007FFF00u) >> 8;
    const int32_t orow = oaddr & 0x3Fu;
    const int32_t ocol = oaddr >> 6;
    const uint8_t osp = other & 0xFFu;
    if (orow >= sp_row_lower && orow <= sp_row_upper && ocol >= sp_col_lower && ocol <= sp_col_upper) {
      const int rr = orow - sp_row_lower;
      const int rc = ocol - sp_col_lower;
      pixel_array[rc] |= ((osp & 0x0Fu) << (4 * rr)) | ((osp & 0xF0u) << (12 + 4 * rr));
    }
  }
  const uint32_t sp_relative_row = starting_pixel_location & 0x3u;
  const uint32_t sp_relative_col = starting_pixel_location < 4;
  const uint32_t col = sp_col * 2 + sp_relative_col;
  const uint64_t starting_pixel = ((uint64_t)(0x01u << (11 - sp_relative_row)) << (16 * (col & 0x01))) << 32;
  const uint64_t pixel_map = (((uint64_t) pixel_array[3]) << 32) | pixel_array[2];
  const ClusterGrow cl = make_cluster(pixel_map, starting_pixel, col_lower_limit + 4, row_lower_limit);
  const uint64_t precedence =
#if VO_MUTANT == 7
    (cl.pixels & (starting_pixel ^ (0 - starting_pixel)));
#else
    (mask_east(cl.pixels) & pixel_array[4]) | (cl.pixels & (starting_pixel ^ (0 - starting_pixel)));
#endif
#if VO_MUTANT == 4
  bool keep = precedence == 0 && !popc64(cl.pixels & 0x8000800080008000ull);
#else
  bool keep = precedence == 0 && !popc64(cl.pixels & 0x8000800080008000ull) &&
              !popc64(cl.pixels & 0x0001000100010001ull);
#endif
  if (!keep) return;
  unsigned x = cl.x, y = cl.y, n = cl.n;
  if (popc64(cl.pixels & 0xFFFFull)) {
    const uint64_t west_map = (((uint64_t) pixel_array[1]) << 32) | pixel_array[0];
    const uint64_t west_start = mask_west(cl.pixels) & west_map;
    const ClusterGrow ext = make_cluster(west_map, west_start, col_lower_limit, row_lower_limit);
    x += ext.x;
    y += ext.y;
    n += ext.n;
#if VO_MUTANT == 5
    keep &= !popc64(ext.pixels & 0x0001000100010001ull);
#endif
    keep &= !popc64(ext.pixels & 0x8000800080008000ull) && !popc64(ext.pixels & 0xFFFFull) &&
            !popc64(cl.pixels & 0x0001000100010001ull);
  }
  if (!keep) return;
  const unsigned cx = x / n;
  const unsigned cy = y / n;
  const float fx = x / static_cast<float>(n) - cx;
  const float fy = y / static_cast<float>(n) - cy;
  const uint32_t chip = cx >> 8;
  const uint32_t cid = get_channel_id(sensor_number, chip, cx & 0xFFu, cy);
  const float local_x = g.local_x[cx] + fx * g.x_pitch[cx];
  const float local_y = (cy + 0.5f + fy) * pixel_size;
  const unsigned cluster_num = module_pair_cluster_num[module_pair_number]++;
  const unsigned cluster_start = module_pair_cluster_start[module_pair_number];
  write_cluster(clusters, cluster_start + cluster_num, get_lhcb_id(cid), ltg, local_x, local_y);
}

// MaskedVeloClustering.cu:498-576
void masked_clustering(
  const RawInput& in,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  unsigned event_start,
  const Geometry& g,
  const SpTables& sp_tables,
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* module_candidate_num,
  const uint32_t* cluster_candidates,
  const uint32_t* candidates_offsets,
  const uint32_t* bank_index,
  uint32_t* module_pair_cluster_num_all,
  char* cluster_container)
{
  const unsigned total = offsets_estimated_input_size[n_module_pairs * n_events];
  for (unsigned block = 0; block < n_list; ++block) {
    const unsigned event_number = event_list[block];
    const uint32_t* mp_start = offsets_estimated_input_size + event_number * n_module_pairs;
    uint32_t* mp_num = module_pair_cluster_num_all + event_number * n_module_pairs;
    const unsigned n_cand = module_candidate_num[event_number];
    const uint32_t* cands = cluster_candidates + candidates_offsets[event_number];
    const Clusters clusters {cluster_container, total, 0};
    RawEvent ev(in, event_number + event_start);
    for (unsigned b = 0; b < ev.n_banks; ++b) {
      const RawBank bank = ev.bank(b);
      if (!is_velo_bank(bank.type)) continue;
      const unsigned mp = (in.version == 2 || in.version == 3) ? bank.sensor_pair() / 8 : bank.sensor_index0() / 8;
      no_neighbour_sp(sp_tables, clusters, mp_num, g, mp, mp_start[mp], bank, in.version);
    }
    for (unsigned c = 0; c < n_cand; ++c) {
      const uint32_t candidate = cands[c];
      const uint32_t sensor_number = (candidate >> 3) & 0xFFu;
      const RawBank bank = ev.bank(bank_index[sensor_number]);
      rest_of_clusters(mp_start, clusters, mp_num, g, candidate, bank, in.version);
    }
  }
}

} // namespace vo
