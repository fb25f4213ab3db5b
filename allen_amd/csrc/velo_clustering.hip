// Super-pixel decoding, masked clustering and phi sort for gfx950.
//
//   k_number_of_candidates   VeloCalculateNumberOfCandidates.cu:10-48   one wave per event
//   k_bank_index             VeloCalculateNumberOfCandidates.cu:35-43   last-writer-wins in CPU order
//   k_estimate_input_size    EstimateInputSize.cu:10-239                one workgroup per event
//   k_masked_clustering      MaskedVeloClustering.cu:131-576            one workgroup per event
//   k_sort_by_phi            SortByPhi.cu:49-153                        one workgroup per module pair
//   k_cluster_sorted         K2+K3+K4 fused: one workgroup per (event, module pair); the
//                            module pair's SP words are staged in LDS, clusters are built and
//                            rank-sorted by (phi, id) in LDS and written once, already sorted.
//
// The cluster arithmetic (pattern lookup, 64-bit mask growth, centroid, local->global
// transform, hit_phi_16) is evaluated exactly as in the reference with fp contraction off.
#include "velo_common.cuh"
#include "velo_kernels.h"

namespace velo {

constexpr unsigned max_banks_per_mp = 8; // v2/v3: 8 sensors per module pair; v4: 4 sensor pairs

// ---------------------------------------------------------------------------------------
// Bit tricks of MaskedVeloClustering.cu:21-84.
__device__ __forceinline__ uint64_t make_8con_mask(uint64_t c)
{
  return c | ((c & 0x7FFF7FFF7FFF7FFFull) << 1) | ((c & 0xFFFEFFFEFFFEFFFEull) << 15) | (c << 16) |
         ((c & 0xFFFCFFFCFFFCFFFCull) << 17) | ((c & 0xFFFEFFFEFFFEFFFEull) >> 1) |
         ((c & 0x7FFF7FFF7FFF7FFFull) >> 15) | (c >> 16) | ((c & 0x3FFF3FFF3FFF3FFFull) >> 17);
}
__device__ __forceinline__ uint64_t mask_east(uint64_t c)
{
  const uint64_t m = c >> 48;
  return m | ((m & 0x7FFF7FFF7FFF7FFFull) << 1) | (m >> 1);
}
__device__ __forceinline__ uint64_t mask_west(uint64_t c)
{
  const uint64_t m = c << 48;
  return m | (m << 1) | ((m & 0xFFFEFFFEFFFEFFFEull) >> 1);
}

struct Grown {
  uint64_t pixels;
  uint32_t x, y, n;
};

__device__ __forceinline__ Grown grow_cluster(uint64_t pixel_map, uint64_t start, unsigned col_lower, unsigned row_lower)
{
  uint64_t cur = 0, next = start;
  while (cur != next) {
    cur = next;
    next = pixel_map & make_8con_mask(cur);
  }
  const unsigned n = __popcll(cur);
  const unsigned x = col_lower * n + __popcll(cur & 0x00000000FFFF0000ull) + __popcll(cur & 0x0000FFFF00000000ull) * 2 +
                     __popcll(cur & 0xFFFF000000000000ull) * 3;
  // Sum of the row offsets r (bit r of each 16-bit column) by bit planes of r. The reference sums
  // r = 1..14; a pixel at r = 15 makes every caller drop the cluster, so the sums agree wherever
  // they are used.
  const unsigned y = row_lower * n + __popcll(cur & 0xAAAAAAAAAAAAAAAAull) + 2 * __popcll(cur & 0xCCCCCCCCCCCCCCCCull) +
                     4 * __popcll(cur & 0xF0F0F0F0F0F0F0F0ull) + 8 * __popcll(cur & 0xFF00FF00FF00FF00ull);
  return {cur, x, y, n};
}

// Global position and LHCbID of a cluster (MaskedVeloClustering.cu:192-225,459-494).
struct ClusterOut {
  uint32_t id;
  float x, y, z;
  int16_t phi;
};

__device__ __forceinline__ ClusterOut
place_cluster(const Geometry& g, const float* ltg, uint32_t lhcb_id, unsigned cx, unsigned cy, float fx, float fy)
{
  // Every geometry load is issued before the first is used (the scheduler would otherwise wait
  // out several memory round trips one after another).
  float m[12];
#pragma unroll
  for (unsigned k = 0; k < 12; ++k) m[k] = ltg[k];
  const float lx = g.local_x[cx], pitch = g.x_pitch[cx];
  __builtin_amdgcn_sched_barrier(0);
  const float local_x = lx + fx * pitch;
  const float local_y = (cy + 0.5f + fy) * pixel_size;
  ClusterOut c;
  c.id = lhcb_id;
  c.x = m[0] * local_x + m[1] * local_y + m[9];
  c.y = m[3] * local_x + m[4] * local_y + m[10];
  c.z = m[6] * local_x + m[7] * local_y + m[11];
  c.phi = hit_phi_16(c.x, c.y);
  return c;
}

// One or two clusters of an isolated SP from the pattern table (MaskedVeloClustering.cu:165-286).
// Returns the number written into out[].
template<int version>
__device__ __forceinline__ unsigned isolated_sp_clusters(
  const Geometry& g,
  const uint8_t* sp_patterns,
  const float* sp_fx,
  const float* sp_fy,
  const RawBank& bank,
  uint32_t sp_word,
  ClusterOut& out0,
  ClusterOut& out1)
{
  constexpr bool v23 = (version == 2 || version == 3);
  const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
  const bool sensor_bit = (sp_word & 0x800000u) != 0;
  const uint32_t pair = bank.sourceID & 0x1FFu;
  const uint32_t sensor_index = v23 ? bank.sourceID : ((pair << 1) | (sensor_bit ? 1u : 0u));
  const float* ltg = g.ltg + 12 * (v23 ? pair : ((pair << 1) | (sensor_bit ? 1u : 0u)));
  const uint32_t sp_row = sp_addr & 0x3Fu;
  const uint32_t sp_col = sp_addr >> 6;
  const uint32_t sp = sp_word & 0xFFu;
  const uint32_t idx = sp_patterns[sp];
  const uint32_t chip = sp_col >> 7;
  unsigned n = 0;
#pragma unroll
  for (unsigned k = 0; k < 2; ++k) {
    if (k == 1 && !(idx & 8)) break;
    const uint32_t row = k == 0 ? (idx & 0x03u) : ((idx >> 4) & 3u);
    const uint32_t col = k == 0 ? ((idx >> 2) & 1u) : ((idx >> 6) & 1u);
    const uint32_t cx = sp_col * 2 + col;
    const uint32_t cy = sp_row * 4 + row;
    const ClusterOut c =
      place_cluster(g, ltg, lhcb_id_of(sensor_index, chip, cx & 0xFFu, cy), cx, cy, sp_fx[sp * 2 + k], sp_fy[sp * 2 + k]);
    if (k == 0) out0 = c;
    else out1 = c;
    ++n;
  }
  return n;
}

// Candidate pixels of a non-isolated SP (EstimateInputSize.cu:104-204), in two steps:
// neighbour_acc() ORs into the 18-bit neighbourhood `pixels` what one other SP word of the bank
// contributes, candidates_final() applies the candidate mask. The neighbour scan keeps the
// reference's behaviour: every word of the bank, of either sensor and isolated or not,
// contributes (EstimateInputSize.cu:111-141).
__device__ __forceinline__ uint32_t neighbourhood_init(uint32_t sp_word)
{
  const uint32_t sp = sp_word & 0xFFu;
  return (sp & 0x0F) | ((sp & 0xF0) << 2);
}

__device__ __forceinline__ void neighbour_acc(uint32_t sp_row, uint32_t sp_col, uint32_t other, uint32_t& pixels)
{
  const uint32_t oaddr = (other & 0x007FFF00u) >> 8;
  const uint32_t orow = oaddr & 0x3Fu;
  const uint32_t ocol = oaddr >> 6;
  const uint32_t osp = other & 0xFFu;
  const bool top = orow == sp_row + 1, mid = orow == sp_row, bot = orow == sp_row - 1;
  const bool same_col = ocol == sp_col, right_col = ocol == sp_col + 1;
  if (top && same_col) pixels |= ((osp & 0x01) | ((osp & 0x10) << 2)) << 4;
  if (top && right_col) pixels |= (osp & 0x01) << 16;
  if (mid && right_col) pixels |= (osp & 0x0F) << 12;
  if (bot && right_col) pixels |= (osp & 0x08) << 8;
  if (bot && same_col) pixels |= (osp & 0x80) >> 2;
}

// Returns a 2-bit mask: bit0 = first candidate (pixels 0,1,4,5), bit1 = second; the pixel
// numbers are returned in *p0, *p1.
__device__ __forceinline__ unsigned candidates_final(uint32_t pixels, unsigned* p0, unsigned* p1)
{
  const uint32_t inside = pixels & 0x3CF;
  const uint32_t mask = (inside << 1) | (inside << 5) | (inside << 6) | (inside << 7);
  const uint32_t working = mask & (~pixels);
  const uint32_t cand = ((working >> 1) & (working >> 5) & (working >> 6) & (working >> 7)) & pixels;
  const uint32_t consolidated = (cand & 0x0F) | ((cand >> 2) & 0xF0);
  const uint32_t first = consolidated & 0x33;
  const uint32_t second = consolidated & 0xCC;
  unsigned m = 0;
  if (first) {
    *p0 = __clz(first) - 24;
    m |= 1;
  }
  if (second) {
    *p1 = __clz(second) - 24;
    m |= 2;
  }
  return m;
}

__device__ __forceinline__ unsigned
sp_candidates(uint32_t sp_word, const uint32_t* words, unsigned n_sp, unsigned* p0, unsigned* p1)
{
  const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
  const uint32_t sp_row = sp_addr & 0x3Fu, sp_col = sp_addr >> 6;
  uint32_t pixels = neighbourhood_init(sp_word);
  for (unsigned k = 0; k < n_sp; ++k) neighbour_acc(sp_row, sp_col, words[k], pixels);
  return candidates_final(pixels, p0, p1);
}

__device__ __forceinline__ unsigned isolated_cluster_count(uint32_t sp)
{
  const bool pattern_0 = (sp & 0x88) && !(sp & 0x44) && (sp & 0x33);
  const bool pattern_1 = (sp & 0xCC) && !(sp & 0x22) && (sp & 0x11);
  return (pattern_0 | pattern_1) ? 2 : 1;
}

// rest_of_clusters (MaskedVeloClustering.cu:290-496) for one candidate pixel, in two steps:
// window_acc() ORs one other SP word of the bank into the 5-column x 4-row SP window around the
// candidate's SP (pa[c]: one SP column, rows bottom to top, 16 bits per pixel column), and
// masked_cluster_final() grows the cluster, applies the precedence and window-edge rules and
// places it.
struct Window {
  uint32_t pa0, pa1, pa2, pa3, pa4;
};

template<int version>
__device__ __forceinline__ void
window_acc(int32_t sp_row_lower, int32_t sp_col_lower, uint32_t sensor_number, uint32_t other, Window& w)
{
  if constexpr (version > 3) {
    if (((other >> 23) & 1u) != (sensor_number & 1u)) return;
  }
  if (other & 0x80000000u) return;
  const uint32_t oaddr = (other & 0x007FFF00u) >> 8;
  const int32_t rr = static_cast<int32_t>(oaddr & 0x3Fu) - sp_row_lower;
  const int32_t rc = static_cast<int32_t>(oaddr >> 6) - sp_col_lower;
  if (rr >= 0 && rr <= 3 && rc >= 0 && rc <= 4) {
    const uint32_t osp = other & 0xFFu;
    const uint32_t bits = ((osp & 0x0Fu) << (4 * rr)) | ((osp & 0xF0u) << (12 + 4 * rr));
    w.pa0 |= rc == 0 ? bits : 0u;
    w.pa1 |= rc == 1 ? bits : 0u;
    w.pa2 |= rc == 2 ? bits : 0u;
    w.pa3 |= rc == 3 ? bits : 0u;
    w.pa4 |= rc == 4 ? bits : 0u;
  }
}

// window_acc() for a word already known to be non-isolated and on the candidate's sensor (the
// sorted lists of k_cluster_mp hold only such words in a candidate's range). Bits 24-31 of
// `other` are not read.
__device__ __forceinline__ void window_acc_same_sensor(int32_t sp_row_lower, int32_t sp_col_lower, uint32_t other, Window& w)
{
  const uint32_t oaddr = (other & 0x007FFF00u) >> 8;
  const int32_t rr = static_cast<int32_t>(oaddr & 0x3Fu) - sp_row_lower;
  const int32_t rc = static_cast<int32_t>(oaddr >> 6) - sp_col_lower;
  if (rr >= 0 && rr <= 3 && rc >= 0 && rc <= 4) {
    const uint32_t osp = other & 0xFFu;
    const uint32_t bits = ((osp & 0x0Fu) << (4 * rr)) | ((osp & 0xF0u) << (12 + 4 * rr));
    w.pa0 |= rc == 0 ? bits : 0u;
    w.pa1 |= rc == 1 ? bits : 0u;
    w.pa2 |= rc == 2 ? bits : 0u;
    w.pa3 |= rc == 3 ? bits : 0u;
    w.pa4 |= rc == 4 ? bits : 0u;
  }
}

__device__ __forceinline__ bool masked_cluster_final(
  const Geometry& g,
  uint32_t sp_word,
  uint32_t sensor_number,
  uint32_t starting_pixel_location,
  const Window& pa,
  ClusterOut* out)
{
  const float* ltg = g.ltg + 12 * sensor_number;
  const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
  const int32_t sp_row = sp_addr & 0x3Fu;
  const int32_t sp_col = sp_addr >> 6;
  const int32_t row_lower_limit = (sp_row - 2) * 4;
  const int32_t col_lower_limit = (sp_col - 3) * 2;
  const uint32_t rel_row = starting_pixel_location & 0x3u;
  const uint32_t rel_col = starting_pixel_location < 4;
  const uint32_t col = sp_col * 2 + rel_col;
  const uint64_t starting_pixel = ((uint64_t)(0x01u << (11 - rel_row)) << (16 * (col & 0x01))) << 32;
  const uint64_t pixel_map = (((uint64_t) pa.pa3) << 32) | pa.pa2;
  const Grown cl = grow_cluster(pixel_map, starting_pixel, col_lower_limit + 4, row_lower_limit);
  const uint64_t precedence = (mask_east(cl.pixels) & pa.pa4) | (cl.pixels & (starting_pixel ^ (0ull - starting_pixel)));
  bool keep = precedence == 0 && !(cl.pixels & 0x8000800080008000ull) && !(cl.pixels & 0x0001000100010001ull);
  if (!keep) return false;
  unsigned x = cl.x, y = cl.y, n = cl.n;
  if (cl.pixels & 0xFFFFull) {
    const uint64_t west_map = (((uint64_t) pa.pa1) << 32) | pa.pa0;
    const uint64_t west_start = mask_west(cl.pixels) & west_map;
    const Grown ext = grow_cluster(west_map, west_start, col_lower_limit, row_lower_limit);
    x += ext.x;
    y += ext.y;
    n += ext.n;
    keep = !(ext.pixels & 0x8000800080008000ull) && !(ext.pixels & 0xFFFFull);
  }
  if (!keep) return false;
  const unsigned cx = x / n;
  const unsigned cy = y / n;
  const float fx = x / static_cast<float>(n) - cx;
  const float fy = y / static_cast<float>(n) - cy;
  *out = place_cluster(g, ltg, lhcb_id_of(sensor_number, cx >> 8, cx & 0xFFu, cy), cx, cy, fx, fy);
  return true;
}

// The whole of rest_of_clusters for the SP at `sp_index` of a bank whose words are `words`.
template<int version>
__device__ __forceinline__ bool masked_cluster(
  const Geometry& g,
  const uint32_t* words,
  unsigned n_sp,
  unsigned sp_index,
  uint32_t sensor_number,
  uint32_t starting_pixel_location,
  ClusterOut* out)
{
  const uint32_t sp_word = words[sp_index];
  const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
  const int32_t sp_row_lower = static_cast<int32_t>(sp_addr & 0x3Fu) - 2;
  const int32_t sp_col_lower = static_cast<int32_t>(sp_addr >> 6) - 3;
  Window w {0, 0, 0, 0, 0};
  for (unsigned k = 0; k < n_sp; ++k) window_acc<version>(sp_row_lower, sp_col_lower, sensor_number, words[k], w);
  return masked_cluster_final(g, sp_word, sensor_number, starting_pixel_location, w, out);
}

__device__ __forceinline__ void store_cluster(const Clusters& c, unsigned i, const ClusterOut& o)
{
  c.ids()[i] = o.id;
  c.xyz()[3 * (size_t) i] = o.x;
  c.xyz()[3 * (size_t) i + 1] = o.y;
  c.xyz()[3 * (size_t) i + 2] = o.z;
  c.phis()[i] = o.phi;
}

// ---------------------------------------------------------------------------------------
// K1: one wave per event sums the SP words of its VELO banks.
template<int version>
__global__ void k_number_of_candidates(
  const char* raw,
  const uint32_t* offsets,
  const uint32_t* sizes,
  const uint32_t* types,
  unsigned mep,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  uint32_t* number_of_candidates)
{
  const unsigned waves = blockDim.x / wave_size;
  const unsigned ev_idx = blockIdx.x * waves + threadIdx.x / wave_size;
  if (ev_idx >= n_list) return;
  const unsigned event_number = event_list ? event_list[ev_idx] : ev_idx;
  const RawEvent ev(raw, offsets, sizes, types, mep, event_number + event_start);
  uint32_t n = 0;
  for (unsigned b = lane_id(); b < ev.n_banks; b += wave_size) {
    const RawBank bank = raw_bank<version>(ev, b);
    if (is_velo_bank(bank.type)) n += bank_n_sp<version>(bank);
  }
#pragma unroll
  for (unsigned d = 32; d > 0; d >>= 1) n += __shfl_xor(n, d, wave_size);
  if (lane_id() == 0) number_of_candidates[event_number] = 2 * n;
}

// Bank directory of the whole-sequence path: one wave per event reads the bank headers once
// and writes, per (event, module pair), the module pair's banks in bank order (at most 8) as
// {first SP word (index into the raw buffer in 32-bit words), SP words, sourceID}, plus their
// number. Also writes dev_number_of_candidates (2 x SP words of the VELO banks).
template<int version>
__global__ void k_bank_directory(
  const char* raw,
  const uint32_t* offsets,
  const uint32_t* sizes,
  const uint32_t* types,
  unsigned mep,
  unsigned n_events,
  uint32_t* number_of_candidates,
  BankEntry* directory,
  uint32_t* directory_count)
{
  constexpr bool v23 = (version == 2 || version == 3);
  const unsigned waves = blockDim.x / wave_size;
  const unsigned e = blockIdx.x * waves + threadIdx.x / wave_size;
  if (e >= n_events) return;
  const RawEvent ev(raw, offsets, sizes, types, mep, e);
  uint32_t total = 0;
  unsigned my_count = 0; // lane m: banks of module pair m so far
  for (unsigned base = 0; base < ev.n_banks; base += wave_size) {
    const unsigned b = base + lane_id();
    unsigned mp = n_module_pairs;
    RawBank bank;
    if (b < ev.n_banks) {
      bank = raw_bank<version>(ev, b);
      if (is_velo_bank(bank.type)) {
        total += bank_n_sp<version>(bank);
        const uint32_t pair = bank.sourceID & 0x1FFu;
        mp = v23 ? pair / 8 : (pair << 1) / 8;
      }
    }
    // One round per module pair present in this chunk: its banks take the next ranks in bank
    // order (lane order within the chunk).
    for (uint64_t todo = __ballot(mp < n_module_pairs); todo;) {
      const unsigned m = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(mp), __ffsll(static_cast<unsigned long long>(todo)) - 1));
      const uint64_t bal = __ballot(mp == m);
      const unsigned before = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(my_count), static_cast<int>(m)));
      if (mp == m) {
        const unsigned rank = before + __popcll(bal & lanes_below_mask());
        if (rank < max_banks_per_mp) {
          BankEntry en;
          en.word_offset = static_cast<uint32_t>((reinterpret_cast<const char*>(bank.word) - raw) / 4);
          en.n_sp = bank_n_sp<version>(bank);
          en.source = bank.sourceID;
          en.pad = 0;
          directory[(static_cast<size_t>(e) * n_module_pairs + m) * max_banks_per_mp + rank] = en;
        }
      }
      if (lane_id() == m) my_count += __popcll(bal);
      todo &= ~bal;
    }
  }
#pragma unroll
  for (unsigned d = 32; d > 0; d >>= 1) total += __shfl_xor(total, d, wave_size);
  if (lane_id() < n_module_pairs) directory_count[e * n_module_pairs + lane_id()] = min(my_count, max_banks_per_mp);
  if (lane_id() == 0) number_of_candidates[e] = 2 * total;
}

// dev_velo_bank_index as the CPU target leaves it: events of CPU block 0 (ev_idx = 0, G, 2G, ...)
// write in order and the last write of each entry wins. Walk those events backwards and keep the
// first value seen per entry; stop once every sensor has one.
template<int version>
__global__ void k_bank_index(
  const char* raw,
  const uint32_t* offsets,
  const uint32_t* sizes,
  const uint32_t* types,
  unsigned mep,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  unsigned cpu_grid,
  uint32_t* bank_index)
{
  __shared__ int best[n_sensors];
  __shared__ uint8_t written[n_sensors];
  __shared__ unsigned n_written;
  for (unsigned i = threadIdx.x; i < n_sensors; i += blockDim.x) written[i] = 0;
  if (threadIdx.x == 0) n_written = 0;
  __syncthreads();
  if (n_list == 0) return;
  const unsigned last = ((n_list - 1) / cpu_grid) * cpu_grid;
  for (int ev_idx = last; ev_idx >= 0; ev_idx -= cpu_grid) {
    for (unsigned i = threadIdx.x; i < n_sensors; i += blockDim.x) best[i] = -1;
    __syncthreads();
    const unsigned event_number = event_list ? event_list[ev_idx] : ev_idx;
    const RawEvent ev(raw, offsets, sizes, types, mep, event_number + event_start);
    for (unsigned b = threadIdx.x; b < ev.n_banks; b += blockDim.x) {
      const RawBank bank = raw_bank<version>(ev, b);
      const uint32_t pair = bank.sourceID & 0x1FFu;
      // Entries beyond the 208 sensors would be written out of bounds by the reference;
      // they are skipped here.
      if constexpr (version > 3) {
        if ((pair << 1) < n_sensors) atomicMax(&best[pair << 1], (int) b);
        if (((pair << 1) | 1u) < n_sensors) atomicMax(&best[(pair << 1) | 1u], (int) b);
      }
      else {
        if (pair < n_sensors) atomicMax(&best[pair], (int) b);
      }
    }
    __syncthreads();
    for (unsigned i = threadIdx.x; i < n_sensors; i += blockDim.x) {
      if (best[i] >= 0 && !written[i]) {
        bank_index[i] = best[i];
        written[i] = 1;
        atomicAdd(&n_written, 1u);
      }
    }
    __syncthreads();
    if (n_written >= n_sensors) break;
  }
}

// K2: one workgroup per event; wave w takes banks w, w+W, ...; lanes take SP words.
template<int version>
__global__ void k_estimate_input_size(
  const char* raw,
  const uint32_t* offsets,
  const uint32_t* sizes,
  const uint32_t* types,
  unsigned mep,
  const uint32_t* event_list,
  unsigned event_start,
  const uint32_t* candidates_offsets,
  uint32_t* estimated_input_size,
  uint32_t* module_candidate_num,
  uint32_t* cluster_candidates)
{
  __shared__ uint32_t est[n_module_pairs];
  __shared__ uint32_t n_cand;
  const unsigned event_number = event_list ? event_list[blockIdx.x] : blockIdx.x;
  for (unsigned i = threadIdx.x; i < n_module_pairs; i += blockDim.x) est[i] = 0;
  if (threadIdx.x == 0) n_cand = 0;
  __syncthreads();
  uint32_t* cands = cluster_candidates + candidates_offsets[event_number];
  const RawEvent ev(raw, offsets, sizes, types, mep, event_number + event_start);
  const unsigned waves = blockDim.x / wave_size;
  for (unsigned b = threadIdx.x / wave_size; b < ev.n_banks; b += waves) {
    const RawBank bank = raw_bank<version>(ev, b);
    if (!is_velo_bank(bank.type)) continue;
    const unsigned n_sp = bank_n_sp<version>(bank);
    const uint32_t pair = bank.sourceID & 0x1FFu;
    const unsigned mp = (version == 2 || version == 3) ? pair / 8 : (pair << 1) / 8;
    if (mp >= n_module_pairs) continue; // sourceID outside the 208 sensors
    uint32_t count = 0;
    for (unsigned s = lane_id(); s < n_sp; s += wave_size) {
      const uint32_t w = bank.word[s];
      if (w & 0x80000000u) {
        count += isolated_cluster_count(w & 0xFFu);
      }
      else {
        unsigned p0 = 0, p1 = 0;
        const unsigned m = sp_candidates(w, bank.word, n_sp, &p0, &p1);
        const uint32_t sensor_number =
          (version == 2 || version == 3) ? pair : ((pair << 1) | ((w >> 23) & 0x1u));
        if (m & 1) cands[atomicAdd(&n_cand, 1u)] = (s << 11) | (sensor_number << 3) | p0;
        if (m & 2) cands[atomicAdd(&n_cand, 1u)] = (s << 11) | (sensor_number << 3) | p1;
        count += __popc(m);
      }
    }
    if (count) atomicAdd(&est[mp], count);
  }
  __syncthreads();
  for (unsigned i = threadIdx.x; i < n_module_pairs; i += blockDim.x)
    estimated_input_size[event_number * n_module_pairs + i] = est[i];
  if (threadIdx.x == 0) module_candidate_num[event_number] = n_cand;
}

// K3: one workgroup per event. Phase 1 takes isolated SPs of all banks, phase 2 the candidates.
template<int version>
__global__ void k_masked_clustering(
  const char* raw,
  const uint32_t* offsets,
  const uint32_t* sizes,
  const uint32_t* types,
  unsigned mep,
  const uint32_t* event_list,
  unsigned event_start,
  const uint32_t* number_of_events,
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* module_candidate_num,
  const uint32_t* cluster_candidates,
  const uint32_t* candidates_offsets,
  const uint32_t* bank_index,
  const Geometry* geometry,
  const uint8_t* sp_patterns,
  const float* sp_fx,
  const float* sp_fy,
  uint32_t* module_cluster_num,
  char* cluster_container)
{
  __shared__ uint32_t count[n_module_pairs];
  const unsigned event_number = event_list ? event_list[blockIdx.x] : blockIdx.x;
  const unsigned n_events = number_of_events[0];
  const Clusters clusters {cluster_container, offsets_estimated_input_size[n_module_pairs * n_events]};
  const uint32_t* mp_start = offsets_estimated_input_size + event_number * n_module_pairs;
  for (unsigned i = threadIdx.x; i < n_module_pairs; i += blockDim.x) count[i] = 0;
  __syncthreads();
  const Geometry& g = *geometry;
  const RawEvent ev(raw, offsets, sizes, types, mep, event_number + event_start);
  const unsigned waves = blockDim.x / wave_size;
  for (unsigned b = threadIdx.x / wave_size; b < ev.n_banks; b += waves) {
    const RawBank bank = raw_bank<version>(ev, b);
    if (!is_velo_bank(bank.type)) continue;
    const unsigned n_sp = bank_n_sp<version>(bank);
    const uint32_t pair = bank.sourceID & 0x1FFu;
    const unsigned mp = (version == 2 || version == 3) ? pair / 8 : (pair << 1) / 8;
    if (mp >= n_module_pairs) continue;
    for (unsigned s = lane_id(); s < n_sp; s += wave_size) {
      const uint32_t w = bank.word[s];
      if (!(w & 0x80000000u)) continue;
      ClusterOut c0, c1;
      const unsigned n = isolated_sp_clusters<version>(g, sp_patterns, sp_fx, sp_fy, bank, w, c0, c1);
      const unsigned slot = atomicAdd(&count[mp], n);
      const unsigned cap = mp_start[mp + 1] - mp_start[mp];
      if (slot < cap) store_cluster(clusters, mp_start[mp] + slot, c0);
      if (n > 1 && slot + 1 < cap) store_cluster(clusters, mp_start[mp] + slot + 1, c1);
    }
  }
  __syncthreads();
  const unsigned n_cand = module_candidate_num[event_number];
  const uint32_t* cands = cluster_candidates + candidates_offsets[event_number];
  for (unsigned i = threadIdx.x; i < n_cand; i += blockDim.x) {
    const uint32_t cand = cands[i];
    const uint32_t sensor_number = (cand >> 3) & 0xFFu;
    if (sensor_number >= n_sensors) continue;
    const RawBank bank = raw_bank<version>(ev, bank_index[sensor_number]);
    ClusterOut c;
    if (masked_cluster<version>(g, bank.word, bank_n_sp<version>(bank), cand >> 11, sensor_number, cand & 7u, &c)) {
      const unsigned mp = sensor_number / 8;
      const unsigned slot = atomicAdd(&count[mp], 1u);
      if (mp < n_module_pairs && slot < mp_start[mp + 1] - mp_start[mp]) store_cluster(clusters, mp_start[mp] + slot, c);
    }
  }
  __syncthreads();
  for (unsigned i = threadIdx.x; i < n_module_pairs; i += blockDim.x)
    module_cluster_num[event_number * n_module_pairs + i] = min(count[i], mp_start[i + 1] - mp_start[i]);
}

// (phi, id) rank of element i among n (SortByPhi.cu:90-125, including its tie rule).
template<typename PhiT, typename IdT>
__device__ __forceinline__ unsigned phi_rank(const PhiT* phi, const IdT* id, unsigned n, unsigned i)
{
  const int16_t pi = phi[i];
  const uint32_t ii = id[i];
  // Eight elements read at a time, so that their loads are in flight together.
  unsigned position = 0, j = 0;
  for (; j + 8 <= n; j += 8) {
    int16_t p[8];
    uint32_t d[8];
#pragma unroll
    for (unsigned u = 0; u < 8; ++u) {
      p[u] = phi[j + u];
      d[u] = static_cast<uint32_t>(id[j + u]);
    }
#pragma unroll
    for (unsigned u = 0; u < 8; ++u) position += pi > p[u] || (i != j + u && pi == p[u] && ii > d[u]);
  }
  for (; j < n; ++j) {
    const int16_t pj = phi[j];
    position += pi > pj || (i != j && pi == pj && ii > (uint32_t) id[j]);
  }
  return position;
}

// K4: one workgroup per (module pair, event). Keys are staged in LDS when they fit.
constexpr unsigned sort_lds_cap = 2048;
__global__ void k_sort_by_phi(
  const uint32_t* event_list,
  const uint32_t* number_of_events,
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* module_cluster_num,
  const char* cluster_container,
  char* sorted_container,
  uint32_t* hit_permutation)
{
  __shared__ int16_t s_phi[sort_lds_cap];
  __shared__ uint32_t s_id[sort_lds_cap];
  const unsigned mp = blockIdx.x;
  const unsigned event_number = event_list ? event_list[blockIdx.y] : blockIdx.y;
  const unsigned total = offsets_estimated_input_size[n_module_pairs * number_of_events[0]];
  const Clusters in {const_cast<char*>(cluster_container), total};
  const Clusters out {sorted_container, total};
  const unsigned start = offsets_estimated_input_size[event_number * n_module_pairs + mp];
  const unsigned n = module_cluster_num[event_number * n_module_pairs + mp];
  const bool lds = n <= sort_lds_cap;
  if (lds) {
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
      s_phi[i] = in.phis()[start + i];
      s_id[i] = in.ids()[start + i];
    }
    __syncthreads();
  }
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned r = lds ? phi_rank(s_phi, s_id, n, i) : phi_rank(in.phis() + start, in.ids() + start, n, i);
    hit_permutation[start + r] = start + i;
  }
  __threadfence_block();
  __syncthreads();
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned src = hit_permutation[start + i];
    out.ids()[start + i] = in.ids()[src];
    out.phis()[start + i] = in.phis()[src];
    out.xyz()[3 * (size_t)(start + i)] = in.xyz()[3 * (size_t) src];
    out.xyz()[3 * (size_t)(start + i) + 1] = in.xyz()[3 * (size_t) src + 1];
    out.xyz()[3 * (size_t)(start + i) + 2] = in.xyz()[3 * (size_t) src + 2];
  }
}

// ---------------------------------------------------------------------------------------
// Fused K2+K3+K4 per (module pair, event), used by the whole-sequence path after the
// estimate scan. The module pair's SP words are staged in LDS; clusters are built into LDS,
// ranked by (phi, id) and written to their sorted slot. A module pair that does not fit
// the LDS budget is processed the same way out of global memory.
constexpr unsigned fused_word_cap = 1024;
constexpr unsigned fused_cluster_cap = 1024;

template<int version>
__global__ void __launch_bounds__(256) k_cluster_sorted(
  const char* raw,
  const uint32_t* offsets,
  const uint32_t* sizes,
  const uint32_t* types,
  unsigned mep,
  unsigned n_events,
  const uint32_t* offsets_estimated_input_size,
  const Geometry* geometry,
  const uint8_t* sp_patterns,
  const float* sp_fx,
  const float* sp_fy,
  uint32_t* module_cluster_num,
  char* sorted_container,
  char* unsorted_scratch,
  uint32_t* status)
{
  __shared__ uint32_t s_words[fused_word_cap];
  __shared__ int16_t s_phi[fused_cluster_cap];
  __shared__ uint32_t s_id[fused_cluster_cap];
  __shared__ float s_xyz[3 * fused_cluster_cap];
  __shared__ uint32_t s_bank_start[max_banks_per_mp + 1];
  __shared__ uint32_t s_bank_num[max_banks_per_mp];
  __shared__ uint32_t s_bank_src[max_banks_per_mp];
  __shared__ unsigned s_nbanks, s_count;

  constexpr bool v23 = (version == 2 || version == 3);
  const unsigned mp = blockIdx.x;
  const unsigned event_number = blockIdx.y;
  const unsigned total = offsets_estimated_input_size[n_module_pairs * n_events];
  const unsigned start = offsets_estimated_input_size[event_number * n_module_pairs + mp];
  const unsigned estimate = offsets_estimated_input_size[event_number * n_module_pairs + mp + 1] - start;
  const Geometry& g = *geometry;
  const RawEvent ev(raw, offsets, sizes, types, mep, event_number);

  // Banks of this module pair, in bank order.
  if (threadIdx.x == 0) s_nbanks = 0;
  __syncthreads();
  for (unsigned b = threadIdx.x; b < ev.n_banks; b += blockDim.x) {
    const RawBank bank = raw_bank<version>(ev, b);
    if (!is_velo_bank(bank.type)) continue;
    const uint32_t pair = bank.sourceID & 0x1FFu;
    const unsigned bmp = v23 ? pair / 8 : (pair << 1) / 8;
    if (bmp == mp) {
      const unsigned k = atomicAdd(&s_nbanks, 1u);
      if (k < max_banks_per_mp) s_bank_num[k] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nb = min(s_nbanks, max_banks_per_mp);
    // insertion sort: a handful of banks
    for (unsigned i = 1; i < nb; ++i) {
      const uint32_t v = s_bank_num[i];
      int j = i - 1;
      while (j >= 0 && s_bank_num[j] > v) {
        s_bank_num[j + 1] = s_bank_num[j];
        --j;
      }
      s_bank_num[j + 1] = v;
    }
    uint32_t acc = 0;
    for (unsigned i = 0; i < nb; ++i) {
      const RawBank bank = raw_bank<version>(ev, s_bank_num[i]);
      s_bank_start[i] = acc;
      s_bank_src[i] = bank.sourceID;
      acc += bank_n_sp<version>(bank);
    }
    s_bank_start[nb] = acc;
    s_nbanks = nb;
    s_count = 0;
  }
  __syncthreads();
  const unsigned nb = s_nbanks;
  const unsigned n_words = s_bank_start[nb];
  const bool words_in_lds = n_words <= fused_word_cap;
  const bool clusters_in_lds = estimate <= fused_cluster_cap;
  if (!words_in_lds || !clusters_in_lds) {
    if (threadIdx.x == 0) atomicOr(status, (uint32_t) status_lds_fallback);
  }
  if (words_in_lds) {
    for (unsigned bi = 0; bi < nb; ++bi) {
      const RawBank bank = raw_bank<version>(ev, s_bank_num[bi]);
      const unsigned n = s_bank_start[bi + 1] - s_bank_start[bi];
      for (unsigned s = threadIdx.x; s < n; s += blockDim.x) s_words[s_bank_start[bi] + s] = bank.word[s];
    }
  }
  __syncthreads();

  // Clusters go to LDS, or to the unsorted scratch container (same gapped layout) if too many.
  const Clusters scratch {unsorted_scratch, total};
  auto emit = [&](const ClusterOut& c) {
    const unsigned slot = atomicAdd(&s_count, 1u);
    if (slot >= estimate) return; // cannot happen for well-formed banks: estimate bounds the count
    if (clusters_in_lds) {
      s_phi[slot] = c.phi;
      s_id[slot] = c.id;
      s_xyz[3 * slot] = c.x;
      s_xyz[3 * slot + 1] = c.y;
      s_xyz[3 * slot + 2] = c.z;
    }
    else {
      store_cluster(scratch, start + slot, c);
    }
  };

  for (unsigned w = threadIdx.x; w < n_words; w += blockDim.x) {
    unsigned bi = 0;
    while (bi + 1 < nb && s_bank_start[bi + 1] <= w) ++bi;
    const unsigned s = w - s_bank_start[bi];
    RawBank bank;
    bank.sourceID = s_bank_src[bi];
    const unsigned n_sp = s_bank_start[bi + 1] - s_bank_start[bi];
    const uint32_t* words;
    if (words_in_lds) words = s_words + s_bank_start[bi];
    else words = raw_bank<version>(ev, s_bank_num[bi]).word;
    const uint32_t sp_word = words[s];
    if (sp_word & 0x80000000u) {
      ClusterOut c0, c1;
      const unsigned n = isolated_sp_clusters<version>(g, sp_patterns, sp_fx, sp_fy, bank, sp_word, c0, c1);
      emit(c0);
      if (n > 1) emit(c1);
    }
    else {
      unsigned p0 = 0, p1 = 0;
      const unsigned m = sp_candidates(sp_word, words, n_sp, &p0, &p1);
      const uint32_t pair = bank.sourceID & 0x1FFu;
      const uint32_t sensor_number = v23 ? pair : ((pair << 1) | ((sp_word >> 23) & 0x1u));
      ClusterOut c;
      if ((m & 1) && masked_cluster<version>(g, words, n_sp, s, sensor_number, p0, &c)) emit(c);
      if ((m & 2) && masked_cluster<version>(g, words, n_sp, s, sensor_number, p1, &c)) emit(c);
    }
  }
  __threadfence_block();
  __syncthreads();
  const unsigned n = min(s_count, estimate);
  if (threadIdx.x == 0) module_cluster_num[event_number * n_module_pairs + mp] = n;
  const Clusters out {sorted_container, total};
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    unsigned r;
    ClusterOut c;
    if (clusters_in_lds) {
      r = phi_rank(s_phi, s_id, n, i);
      c.id = s_id[i];
      c.phi = s_phi[i];
      c.x = s_xyz[3 * i];
      c.y = s_xyz[3 * i + 1];
      c.z = s_xyz[3 * i + 2];
    }
    else {
      r = phi_rank(scratch.phis() + start, scratch.ids() + start, n, i);
      c.id = scratch.ids()[start + i];
      c.phi = scratch.phis()[start + i];
      c.x = scratch.xyz()[3 * (size_t)(start + i)];
      c.y = scratch.xyz()[3 * (size_t)(start + i) + 1];
      c.z = scratch.xyz()[3 * (size_t)(start + i) + 2];
    }
    store_cluster(out, start + r, c);
  }
}

// ---------------------------------------------------------------------------------------
// Host launchers.
template<int version>
static hipError_t launch_candidates_v(const CandidatesLaunch& a, hipStream_t s)
{
  const unsigned waves = 4;
  const unsigned grid = (a.n_list + waves - 1) / waves;
  if (grid && a.directory) {
    if (a.event_list || a.event_start) return hipErrorInvalidValue;
    hipLaunchKernelGGL(
      k_bank_directory<version>, dim3(grid), dim3(waves * wave_size), 0, s, a.raw, a.offsets, a.sizes, a.types, unsigned(a.mep),
      a.n_list, a.number_of_candidates, static_cast<BankEntry*>(a.directory), a.directory_count);
  }
  else if (grid) {
    hipLaunchKernelGGL(
      k_number_of_candidates<version>, dim3(grid), dim3(waves * wave_size), 0, s, a.raw, a.offsets, a.sizes, a.types, unsigned(a.mep),
      a.event_list, a.n_list, a.event_start, a.number_of_candidates);
  }
  const unsigned cpu_grid = (a.n_list + a.block_dim_x - 1) / a.block_dim_x;
  if (a.n_list) {
    hipLaunchKernelGGL(
      k_bank_index<version>, dim3(1), dim3(256), 0, s, a.raw, a.offsets, a.sizes, a.types, unsigned(a.mep), a.event_list, a.n_list,
      a.event_start, cpu_grid, a.bank_index);
  }
  return hipGetLastError();
}

hipError_t launch_number_of_candidates(const CandidatesLaunch& a, hipStream_t s)
{
  switch (a.version) {
  case 2: return launch_candidates_v<2>(a, s);
  case 3: return launch_candidates_v<3>(a, s);
  default: return launch_candidates_v<4>(a, s);
  }
}

hipError_t launch_estimate_input_size(const EstimateLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
#define VELO_LAUNCH_EST(V)                                                                                          \
  hipLaunchKernelGGL(                                                                                               \
    k_estimate_input_size<V>, dim3(a.n_list), dim3(256), 0, s, a.raw, a.offsets, a.sizes, a.types, unsigned(a.mep), a.event_list,    \
    a.event_start, a.candidates_offsets, a.estimated_input_size, a.module_candidate_num, a.cluster_candidates)
  switch (a.version) {
  case 2: VELO_LAUNCH_EST(2); break;
  case 3: VELO_LAUNCH_EST(3); break;
  default: VELO_LAUNCH_EST(4); break;
  }
#undef VELO_LAUNCH_EST
  return hipGetLastError();
}

hipError_t launch_masked_clustering(const MaskedClusteringLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
#define VELO_LAUNCH_MC(V)                                                                                           \
  hipLaunchKernelGGL(                                                                                               \
    k_masked_clustering<V>, dim3(a.n_list), dim3(256), 0, s, a.raw, a.offsets, a.sizes, a.types, unsigned(a.mep), a.event_list,      \
    a.event_start, a.number_of_events, a.offsets_estimated_input_size, a.module_candidate_num, a.cluster_candidates, \
    a.candidates_offsets, a.bank_index, reinterpret_cast<const Geometry*>(a.geometry), a.sp_patterns, a.sp_fx,       \
    a.sp_fy, a.module_cluster_num, a.cluster_container)
  switch (a.version) {
  case 2: VELO_LAUNCH_MC(2); break;
  case 3: VELO_LAUNCH_MC(3); break;
  default: VELO_LAUNCH_MC(4); break;
  }
#undef VELO_LAUNCH_MC
  return hipGetLastError();
}

hipError_t launch_sort_by_phi(const SortLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
  hipLaunchKernelGGL(
    k_sort_by_phi, dim3(n_module_pairs, a.n_list), dim3(256), 0, s, a.event_list, a.number_of_events,
    a.offsets_estimated_input_size, a.module_cluster_num, a.cluster_container, a.sorted_container, a.hit_permutation);
  return hipGetLastError();
}

hipError_t launch_cluster_sorted(const FusedClusterLaunch& a, hipStream_t s)
{
  if (!a.n_events) return hipSuccess;
#define VELO_LAUNCH_FUSED(V)                                                                                        \
  hipLaunchKernelGGL(                                                                                               \
    k_cluster_sorted<V>, dim3(n_module_pairs, a.n_events), dim3(256), 0, s, a.raw, a.offsets, a.sizes, a.types, unsigned(a.mep),     \
    a.n_events, a.offsets_estimated_input_size, reinterpret_cast<const Geometry*>(a.geometry), a.sp_patterns,        \
    a.sp_fx, a.sp_fy, a.module_cluster_num, a.sorted_container, a.unsorted_scratch, a.status)
  switch (a.version) {
  case 2: VELO_LAUNCH_FUSED(2); break;
  case 3: VELO_LAUNCH_FUSED(3); break;
  default: VELO_LAUNCH_FUSED(4); break;
  }
#undef VELO_LAUNCH_FUSED
  return hipGetLastError();
}


// ---------------------------------------------------------------------------------------
// Whole-sequence clustering, one wave per (event, module pair), no block barriers.
//
//   k_estimate_mp   the K2 estimate of one module pair; also records, for every non-isolated
//                   SP word, its candidate pixels in one byte indexed by the word's position in
//                   the raw buffer (bit0/bit1: first/second candidate, bits 2-4 and 5-7: pixels)
//   k_cluster_mp    K3 + K4 of one module pair: isolated SPs through the pattern table, the
//                   recorded candidates through the masked clustering window (each bank's
//                   non-isolated words sorted by sensor, column and row, so a window reads one
//                   run of them), clusters appended with wave ballots, then ranked by (phi, id)
//                   and written sorted
//
// Each wave stages the SP words of its module pair's banks in its own LDS slice; a module pair
// larger than the slice runs the same code on global memory.
namespace mp {
constexpr unsigned waves_per_block = 4;
constexpr unsigned word_cap = 384;
// Clusters of a module pair kept in LDS (more go through global memory): 192 keeps k_cluster_mp's
// workgroup at 22 880 B of LDS, so seven fit a CU; about 0.1 % of pileup_1000's module pairs and
// 1.4 % of batched_20k's have a larger estimate.
constexpr unsigned cluster_cap = 192;
constexpr unsigned hash_bits = 10;  // estimate: (bank, SP address) table of a resident module pair,
constexpr unsigned hash_slots = 1u << hash_bits; // more slots than words, so an empty slot ends every probe
constexpr uint32_t hash_empty = 0xFFFFFFFFu; // an empty slot; keys have 18 bits, so no slot of a key is all ones
constexpr uint32_t hash_no_key = 0x01000000u; // a lookup key that matches no slot (nor an empty one)
constexpr unsigned hash_min_bank = 16; // largest bank (words) above which the table beats the scan (measured)
static_assert(hash_slots > word_cap, "hash table load");

// Banks of one module pair (at most 8), kept in the wave's LDS slice. Bank k's words start at
// base + word_offset[k] (32-bit offsets in place of pointers: 112 B per wave, so that seven
// workgroups of k_estimate_mp fit a CU's LDS).
struct Banks {
  const uint32_t* base;
  unsigned n;
  uint32_t start[max_banks_per_mp + 1];
  uint32_t source[max_banks_per_mp];
  uint32_t word_offset[max_banks_per_mp];
  __device__ __forceinline__ const uint32_t* word(unsigned k) const { return base + word_offset[k]; }
};

// k_cluster_mp's sorted banks: each bank's non-isolated words first, then its isolated ones.
struct SortedBanks {
  uint32_t sorted[max_banks_per_mp + 1];          // non-isolated words of the banks before each
  uint32_t first_isolated[max_banks_per_mp];      // a bank's first isolated position
  uint32_t isolated_before[max_banks_per_mp + 1]; // isolated words of the banks before each
};

__device__ __forceinline__ void wave_sync_lds()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The bank directory record of one (event, module pair) as this lane holds it: lanes 0-7 load
// one entry each, every lane the number of entries; both loads are issued together, without
// waiting, so a wave can fetch the next item's record while it works on the current one.
struct DirFetch {
  BankEntry en;
  unsigned nb;
};

__device__ __forceinline__ DirFetch fetch_dir(const BankEntry* directory, const uint32_t* directory_count, unsigned idx)
{
  DirFetch f;
  f.en = BankEntry {0, 0, 0, 0};
  const unsigned l = lane_id();
  if (l < max_banks_per_mp) f.en = directory[static_cast<size_t>(idx) * max_banks_per_mp + l];
  f.nb = min(directory_count[idx], max_banks_per_mp);
  return f;
}

// The banks of one module pair from its directory record.
__device__ __forceinline__ void load_banks(const char* raw, const DirFetch& f, Banks& B)
{
  const unsigned nb = f.nb;
  const unsigned l = lane_id();
  const BankEntry en = l < nb ? f.en : BankEntry {0, 0, 0, 0};
  const uint32_t inc = wave_inclusive_scan(l < nb ? en.n_sp : 0u);
  if (l < nb) {
    B.start[l] = inc - en.n_sp;
    B.source[l] = en.source;
    B.word_offset[l] = en.word_offset;
  }
  if (l == 0) {
    B.n = nb;
    B.base = reinterpret_cast<const uint32_t*>(raw);
  }
  if (l == nb) B.start[nb] = inc - (l < nb ? en.n_sp : 0u);
  wave_sync_lds();
}

// The bank holding position w of a list laid out bank after bank (starts[k]: bank k's first
// position): the starts are non-decreasing, so the last bank starting at or before w is the
// number of banks after the first that do (all starts read at once, one LDS round trip).
__device__ __forceinline__ unsigned bank_at(const uint32_t* starts, unsigned n, unsigned w)
{
  unsigned bi = 0;
#pragma unroll
  for (unsigned k = 1; k < max_banks_per_mp; ++k) bi += (k < n && starts[k] <= w) ? 1u : 0u;
  return bi;
}
__device__ __forceinline__ unsigned bank_of(const Banks& B, unsigned w) { return bank_at(B.start, B.n, w); }
} // namespace mp

namespace mp {
// Word `w` of the module pair (flattened over its banks), from global memory.
__device__ __forceinline__ uint32_t word_at(const Banks& B, unsigned w)
{
  const unsigned bi = bank_of(B, w);
  return B.word(bi)[w - B.start[bi]];
}

// Copies words [t0, t0 + n) of the module pair into the wave's LDS slice.
__device__ __forceinline__ void load_words(const Banks& B, unsigned t0, unsigned n, uint32_t* lw)
{
  wave_sync_lds();
  for (unsigned i = lane_id(); i < n; i += wave_size) lw[i] = word_at(B, t0 + i);
  wave_sync_lds();
}

// Calls f(word) for every word of the flattened range [lo, hi): straight from the resident
// slice, or tile by tile through it (wave-uniform tile loop, per-lane ranges).
template<typename F>
__device__ __forceinline__ void
for_words(const Banks& B, unsigned n_words, bool resident, uint32_t* lw, bool active, unsigned lo, unsigned hi, F&& f)
{
  // Words [a, b) of the slice in order, eight LDS reads in flight at a time.
  auto scan = [&](const uint32_t* w, unsigned a, unsigned b) {
    unsigned k = a;
    for (; k + 8 <= b; k += 8) {
      uint32_t v[8];
#pragma unroll
      for (unsigned u = 0; u < 8; ++u) v[u] = w[k + u];
#pragma unroll
      for (unsigned u = 0; u < 8; ++u) f(v[u]);
    }
    for (; k < b; ++k) f(w[k]);
  };
  if (resident) {
    if (active) scan(lw, lo, hi);
    return;
  }
  for (unsigned t0 = 0; t0 < n_words; t0 += word_cap) {
    const unsigned tn = min(word_cap, n_words - t0);
    load_words(B, t0, tn, lw);
    if (active) {
      const unsigned a = max(lo, t0), b = min(hi, t0 + tn);
      if (a < b) scan(lw, a - t0, b - t0);
    }
  }
}
} // namespace mp

namespace mp {
// Slot of a (bank, SP address) key in a table of 2^bits slots: multiplicative hashing.
__device__ __forceinline__ unsigned hash_slot(uint32_t key, unsigned bits) { return (key * 2654435761u) >> (32 - bits); }
} // namespace mp

// K2 of one module pair (idx = event * 26 + module pair) by one wave. The neighbourhood of a
// non-isolated SP is what the reference collects by scanning every word of its bank: the SP
// bytes of the words at five neighbouring addresses, OR-ed (EstimateInputSize.cu:111-141). For a
// module pair resident in LDS the words are first entered in a hash table keyed by (bank, SP
// address) whose value is the OR of their SP bytes (one word per slot: key << 8 | value), and
// each SP looks its five neighbours up; OR-ing is order-free, so the result is the scan's.
// Larger module pairs, and those whose banks are all small, scan.
__device__ __forceinline__ void estimate_item(
  const char* raw,
  const mp::DirFetch& dir,
  unsigned idx,
  mp::Banks& B,
  uint32_t* lw,
  uint32_t* table_slots,
  uint32_t* estimated_input_size,
  uint8_t* sp_flags)
{
  mp::load_banks(raw, dir, B);
  const unsigned n_words = B.start[B.n];
  const bool resident = n_words <= mp::word_cap;
  // A scan costs a wave about as many steps as its largest bank has words; the table, a fixed
  // number of LDS round trips. The table serves resident module pairs whose largest bank has
  // more than hash_min_bank words.
  unsigned largest_bank = 0;
  for (unsigned k = 0; k < B.n; ++k) largest_bank = max(largest_bank, B.start[k + 1] - B.start[k]);
  const bool table = resident && largest_bank > mp::hash_min_bank;
  // Table of four times as many slots as words, 64 to 1024 (load at most 3/8: the wave waits for
  // its longest probe, so the probes must stay short everywhere).
  const unsigned bits = min(mp::hash_bits, max(6u, 32u - __clz(4 * n_words - 1)));
  const unsigned hmask = (1u << bits) - 1;
  if (resident) mp::load_words(B, 0, n_words, lw);
  if (table) {
    for (unsigned h = lane_id(); h <= hmask; h += wave_size) table_slots[h] = mp::hash_empty;
    mp::wave_sync_lds();
    for (unsigned w = lane_id(); w < n_words; w += wave_size) {
      const uint32_t word = lw[w];
      const uint32_t key = (mp::bank_of(B, w) << 15) | ((word & 0x007FFF00u) >> 8);
      for (unsigned h = mp::hash_slot(key, bits);; h = (h + 1) & hmask) {
        const uint32_t old = atomicCAS(&table_slots[h], mp::hash_empty, (key << 8) | (word & 0xFFu));
        if (old == mp::hash_empty) break;
        if ((old >> 8) == key) {
          atomicOr(&table_slots[h], word & 0xFFu);
          break;
        }
      }
    }
    mp::wave_sync_lds();
  }
  uint32_t count = 0;
  for (unsigned base = 0; base < n_words; base += wave_size) {
    const unsigned w = base + lane_id();
    const bool active = w < n_words;
    const unsigned bi = active ? mp::bank_of(B, w) : 0;
    const uint32_t sp_word = active ? (resident ? lw[w] : B.word(bi)[w - B.start[bi]]) : 0u;
    const bool isolated = (sp_word & 0x80000000u) != 0;
    const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
    const uint32_t sp_row = sp_addr & 0x3Fu, sp_col = sp_addr >> 6;
    uint32_t pixels = neighbourhood_init(sp_word);
    if (table) {
      if (active && !isolated) {
        // neighbour_acc's five cases, each from the OR of the words at that address (rows and
        // columns outside the SP grid match no word).
        // The five first probes are issued together, then the rare longer probes.
        const uint32_t kb = bi << 15;
        const uint32_t cols[5] = {sp_col, sp_col + 1, sp_col + 1, sp_col + 1, sp_col};
        const uint32_t rows[5] = {sp_row + 1, sp_row + 1, sp_row, sp_row - 1, sp_row - 1};
        uint32_t key[5], got[5], v[5];
        unsigned h[5];
#pragma unroll
        for (unsigned k = 0; k < 5; ++k) {
          const bool valid = rows[k] < 64u && cols[k] < 512u;
          key[k] = valid ? (kb | (cols[k] << 6) | rows[k]) : mp::hash_no_key;
          h[k] = mp::hash_slot(key[k], bits);
          got[k] = valid ? table_slots[h[k]] : mp::hash_empty;
        }
#pragma unroll
        for (unsigned k = 0; k < 5; ++k)
          while ((got[k] >> 8) != key[k] && got[k] != mp::hash_empty) {
            h[k] = (h[k] + 1) & hmask;
            got[k] = table_slots[h[k]];
          }
#pragma unroll
        for (unsigned k = 0; k < 5; ++k) v[k] = (got[k] >> 8) == key[k] ? (got[k] & 0xFFu) : 0u;
        const uint32_t top = v[0], top_right = v[1], right = v[2], bot_right = v[3], bot = v[4];
        pixels |= ((top & 0x01) | ((top & 0x10) << 2)) << 4;
        pixels |= (top_right & 0x01) << 16;
        pixels |= (right & 0x0F) << 12;
        pixels |= (bot_right & 0x08) << 8;
        pixels |= (bot & 0x80) >> 2;
      }
    }
    else {
      mp::for_words(B, n_words, resident, lw, active && !isolated, B.start[bi], B.start[bi + 1],
                    [&](uint32_t other) { neighbour_acc(sp_row, sp_col, other, pixels); });
    }
    if (active) {
      if (isolated) {
        count += isolated_cluster_count(sp_word & 0xFFu);
      }
      else {
        unsigned p0 = 0, p1 = 0;
        const unsigned mk = candidates_final(pixels, &p0, &p1);
        count += __popc(mk);
        const uint32_t* wp = B.word(bi) + (w - B.start[bi]);
        sp_flags[static_cast<size_t>(reinterpret_cast<const char*>(wp) - raw) / 4] =
          static_cast<uint8_t>(mk | (p0 << 2) | (p1 << 5));
      }
    }
  }
#pragma unroll
  for (unsigned d = 32; d > 0; d >>= 1) count += __shfl_xor(count, d, wave_size);
  if (lane_id() == 0) estimated_input_size[idx] = count;
}

// Resident grid: each wave takes items idx, idx + stride, ... and fetches the next item's
// bank directory record before working on the current one.
template<int version>
__global__ void __launch_bounds__(mp::waves_per_block* wave_size) k_estimate_mp(
  const char* raw,
  const BankEntry* directory,
  const uint32_t* directory_count,
  unsigned n_events,
  uint32_t* estimated_input_size,
  uint8_t* sp_flags)
{
  __shared__ uint32_t s_words[mp::waves_per_block][mp::word_cap];
  __shared__ uint32_t s_table[mp::waves_per_block][mp::hash_slots];
  __shared__ mp::Banks s_banks[mp::waves_per_block];
  const unsigned wave = threadIdx.x / wave_size;
  const unsigned n_items = n_events * n_module_pairs;
  const unsigned stride = gridDim.x * mp::waves_per_block;
  unsigned idx = blockIdx.x * mp::waves_per_block + wave;
  if (idx >= n_items) return;
  mp::DirFetch cur = mp::fetch_dir(directory, directory_count, idx);
  for (; idx < n_items; idx += stride) {
    mp::DirFetch next = cur;
    if (idx + stride < n_items) next = mp::fetch_dir(directory, directory_count, idx + stride);
    estimate_item(raw, cur, idx, s_banks[wave], s_words[wave], s_table[wave], estimated_input_size,
                  sp_flags);
    cur = next;
  }
}

// The wave-private LDS slices of k_cluster_mp.
struct ClusterSlices {
  uint32_t* lw;
  uint64_t* lkey;
  float* lxyz;
  uint8_t* lcand; // 2 x wave_size: the candidate pixels of one word round, (lane | pixel << 7)
};

// (phi, id) as one unsigned 64-bit key: phi (signed) in the high word, offset to unsigned. Key
// order is the sort's (phi, id) order (SortByPhi.cu:90-125), and distinct clusters of a module
// pair have distinct ids, so the rank of a cluster is the number of keys below its own.
__device__ __forceinline__ uint64_t phi_id_key(int16_t phi, uint32_t id)
{
  return (static_cast<uint64_t>(static_cast<uint16_t>(phi) ^ 0x8000u) << 32) | id;
}

// The number of keys[0 .. n) below `key` (n is wave-uniform: every lane reads the same key,
// eight reads in flight at a time).
__device__ __forceinline__ unsigned keys_below(const uint64_t* keys, unsigned n, uint64_t key)
{
  unsigned below = 0, j = 0;
  for (; j + 8 <= n; j += 8) {
    uint64_t v[8];
#pragma unroll
    for (unsigned u = 0; u < 8; ++u) v[u] = keys[j + u];
#pragma unroll
    for (unsigned u = 0; u < 8; ++u) below += v[u] < key;
  }
  for (; j < n; ++j) below += keys[j] < key;
  return below;
}

// K3 + K4 of one module pair (idx = event * 26 + module pair) by one wave.
template<int version>
__device__ __forceinline__ void cluster_item(
  const char* raw,
  const mp::DirFetch& dir,
  unsigned idx,
  unsigned total,
  const uint32_t* offsets_estimated_input_size,
  const uint8_t* sp_flags,
  const Geometry& g,
  const uint8_t* sp_patterns,
  const float* sp_fx,
  const float* sp_fy,
  uint32_t* module_cluster_num,
  char* sorted_container,
  char* unsorted_scratch,
  mp::Banks& B,
  mp::SortedBanks& S,
  const ClusterSlices& sl)
{
  constexpr bool v23 = (version == 2 || version == 3);
  const unsigned start = offsets_estimated_input_size[idx];
  const unsigned estimate = offsets_estimated_input_size[idx + 1] - start;
  mp::load_banks(raw, dir, B);
  const unsigned n_words = B.start[B.n];
  const bool resident = n_words <= mp::word_cap;
  const bool clusters_in_lds = estimate <= mp::cluster_cap;
  uint32_t* lw = sl.lw;
  if (resident) mp::load_words(B, 0, n_words, lw);
  const Clusters scratch {unsorted_scratch, total};
  uint64_t* lkey = sl.lkey;
  float* lxyz = sl.lxyz;
  unsigned n = 0; // clusters so far (wave-uniform)
  // Appends up to two clusters per lane in lane order (wave scan of the per-lane counts).
  auto emit = [&](unsigned k, const ClusterOut& c0, const ClusterOut& c1) {
    const unsigned inc = wave_inclusive_scan(k);
    const unsigned slot0 = n + inc - k;
    for (unsigned j = 0; j < k; ++j) {
      const unsigned slot = slot0 + j;
      const ClusterOut& c = j == 0 ? c0 : c1;
      if (slot < estimate) {
        if (clusters_in_lds) {
          lkey[slot] = phi_id_key(c.phi, c.id);
          lxyz[3 * slot] = c.x;
          lxyz[3 * slot + 1] = c.y;
          lxyz[3 * slot + 2] = c.z;
        }
        else {
          store_cluster(scratch, start + slot, c);
        }
      }
    }
    n += __shfl(inc, wave_size - 1, wave_size);
  };
  // The order in which clusters are appended does not reach the output: the module pair's
  // clusters are ranked by their distinct (phi, id) keys below.
  constexpr uint32_t field = v23 ? 0x7FFFu : 0xFFFFu;
  if (resident) {
    // Every bank's words sorted in the slice by (isolated, sensor, SP column, SP row, position)
    // (v2/v3: the bank is the sensor, bit 23 cleared), each with its candidate byte in bits 24-31:
    // a bank's non-isolated words first, then its isolated ones. A candidate's window then takes
    // only the words of its bank and sensor in SP columns col - 3 .. col + 1, a contiguous run
    // found by binary search, instead of every word of the bank; OR-ing the window is order-free,
    // so the windows are the scan's. The cluster slices are still empty here and serve as the
    // sort's scratch: keys in the xyz slice, sorted entries in the key slice, copied back over
    // the words at the end.
    static_assert(3 * mp::cluster_cap >= mp::word_cap && 2 * mp::cluster_cap >= mp::word_cap, "sort scratch");
    uint32_t* keys = reinterpret_cast<uint32_t*>(lxyz);
    uint32_t* sorted = reinterpret_cast<uint32_t*>(lkey);
    // Distinct for distinct positions k; isolated words above the non-isolated ones.
    auto sort_key = [](uint32_t word, unsigned k) { return (word & 0x80000000u) | (((word >> 8) & field) << 9) | k; };
    uint32_t bank_nonisolated = 0; // lane k counts bank k's non-isolated words
    for (unsigned base = 0; base < n_words; base += wave_size) {
      const unsigned w = base + lane_id();
      uint32_t word = 0x80000000u;
      if (w < n_words) {
        word = lw[w];
        keys[w] = sort_key(word, w);
      }
      const uint64_t m = __ballot(!(word & 0x80000000u));
      const unsigned l = lane_id();
      if (l < B.n) {
        auto below = [](unsigned x) { return x >= wave_size ? ~0ull : ((1ull << x) - 1); };
        const unsigned lo = B.start[l] <= base ? 0u : min(B.start[l] - base, wave_size);
        const unsigned hi = B.start[l + 1] <= base ? 0u : min(B.start[l + 1] - base, wave_size);
        bank_nonisolated += __popcll(m & below(hi)) - __popcll(m & below(lo));
      }
    }
    {
      const unsigned l = lane_id();
      const uint32_t cnt = l < B.n ? bank_nonisolated : 0u;
      const uint32_t inc = wave_inclusive_scan(cnt);
      if (l <= B.n) {
        S.sorted[l] = inc - cnt;
        S.isolated_before[l] = B.start[l] - (inc - cnt);
      }
      if (l < B.n) S.first_isolated[l] = B.start[l] + cnt;
    }
    mp::wave_sync_lds(); // keys and counts written
    for (unsigned base = 0; base < n_words; base += wave_size) {
      const unsigned w = base + lane_id();
      if (w < n_words) {
        const uint32_t key = keys[w];
        const unsigned bi = mp::bank_of(B, w);
        const unsigned a = B.start[bi], b = B.start[bi + 1];
        unsigned rank = 0, k = a;
        for (; k + 4 <= b; k += 4) {
          uint32_t v[4];
#pragma unroll
          for (unsigned u = 0; u < 4; ++u) v[u] = keys[k + u];
#pragma unroll
          for (unsigned u = 0; u < 4; ++u) rank += v[u] < key ? 1u : 0u;
        }
        for (; k < b; ++k) rank += keys[k] < key ? 1u : 0u;
        const uint32_t word = lw[w];
        uint32_t f = 0;
        if (!(word & 0x80000000u)) {
          const uint32_t* wp = B.word(bi) + (w - a);
          f = sp_flags[static_cast<size_t>(reinterpret_cast<const char*>(wp) - raw) / 4];
        }
        sorted[a + rank] = (f << 24) | (word & (v23 ? 0x007FFFFFu : 0x00FFFFFFu));
      }
    }
    mp::wave_sync_lds(); // every entry written
    for (unsigned w = lane_id(); w < n_words; w += wave_size) lw[w] = sorted[w];
    mp::wave_sync_lds(); // the cluster slices are free again
  }
  // Isolated SPs. Resident: the isolated tails of the sorted banks, one after another, so every
  // lane of a round takes an isolated word (positions in the same order as a walk over all words
  // would visit them). Otherwise every word, in place, by its isolated bit.
  if (resident) {
    const unsigned n_isolated = S.isolated_before[B.n];
    for (unsigned base = 0; base < n_isolated; base += wave_size) {
      const unsigned t = base + lane_id();
      ClusterOut c0, c1;
      unsigned k = 0;
      if (t < n_isolated) {
        const unsigned bi = mp::bank_at(S.isolated_before, B.n, t);
        const uint32_t sp_word = lw[S.first_isolated[bi] + (t - S.isolated_before[bi])];
        RawBank bank;
        bank.sourceID = B.source[bi];
        k = isolated_sp_clusters<version>(g, sp_patterns, sp_fx, sp_fy, bank, sp_word, c0, c1);
      }
      emit(k, c0, c1);
    }
  }
  else {
    for (unsigned base = 0; base < n_words; base += wave_size) {
      const unsigned w = base + lane_id();
      ClusterOut c0, c1;
      unsigned k = 0;
      if (w < n_words) {
        const unsigned bi = mp::bank_of(B, w);
        const uint32_t sp_word = B.word(bi)[w - B.start[bi]];
        if (sp_word & 0x80000000u) {
          RawBank bank;
          bank.sourceID = B.source[bi];
          k = isolated_sp_clusters<version>(g, sp_patterns, sp_fx, sp_fy, bank, sp_word, c0, c1);
        }
      }
      emit(k, c0, c1);
    }
  }
  // The candidates of one round of (at most 64) SPs: their pixels as one list, an SP's first
  // pixel before its second; each lane grows one of them with its SP's window and word fetched
  // from the SP's lane, and the kept clusters are appended in list order. A round has at most 128
  // candidates.
  auto candidates_round = [&](uint32_t sp_word, uint32_t sensor_number, uint32_t f, const Window& win) {
    const unsigned nc = __popc(f & 3u);
    const unsigned nc_inc = wave_inclusive_scan(nc);
    const unsigned n_cand = __shfl(nc_inc, wave_size - 1, wave_size);
    const unsigned first = nc_inc - nc;
    mp::wave_sync_lds(); // the previous round's list has been read
    if (f & 1u) sl.lcand[first] = static_cast<uint8_t>(lane_id());
    if (f & 2u) sl.lcand[first + (f & 1u)] = static_cast<uint8_t>(lane_id() | 0x80u);
    mp::wave_sync_lds();
    for (unsigned c0 = 0; c0 < n_cand; c0 += wave_size) {
      const unsigned j = c0 + lane_id();
      const unsigned e = j < n_cand ? sl.lcand[j] : 0u;
      const int src = static_cast<int>(e & 0x3Fu);
      Window ws;
      ws.pa0 = __shfl(win.pa0, src, wave_size);
      ws.pa1 = __shfl(win.pa1, src, wave_size);
      ws.pa2 = __shfl(win.pa2, src, wave_size);
      ws.pa3 = __shfl(win.pa3, src, wave_size);
      ws.pa4 = __shfl(win.pa4, src, wave_size);
      const uint32_t word_s = __shfl(sp_word, src, wave_size);
      const uint32_t sensor_s = __shfl(sensor_number, src, wave_size);
      const uint32_t f_s = __shfl(f, src, wave_size);
      const uint32_t location = (e & 0x80u) ? (f_s >> 5) & 7u : (f_s >> 2) & 7u;
      ClusterOut c, unused;
      const bool keep = j < n_cand && masked_cluster_final(g, word_s, sensor_s, location, ws, &c);
      emit(keep ? 1u : 0u, c, unused);
    }
  };
  if (resident) {
    // The non-isolated words, bank after bank: word t of the list is position
    // start[bank] + t - sorted[bank] of the slice.
    const unsigned n_sorted = S.sorted[B.n];
    for (unsigned base = 0; base < n_sorted; base += wave_size) {
      const unsigned t = base + lane_id();
      const bool active = t < n_sorted;
      const unsigned bi = active ? mp::bank_at(S.sorted, B.n, t) : 0u;
      const unsigned j = B.start[bi] + (t - S.sorted[bi]);
      const uint32_t e = active ? lw[j] : 0u;
      const uint32_t f = e >> 24;
      const uint32_t sp_word = e & 0x00FFFFFFu;
      const uint32_t pair = B.source[bi] & 0x1FFu;
      const uint32_t sensor_number = v23 ? pair : ((pair << 1) | ((sp_word >> 23) & 0x1u));
      Window win {0, 0, 0, 0, 0};
      if (f & 3u) {
        const uint32_t key = (e >> 8) & 0xFFFFu;
        const uint32_t sensor = key & 0x8000u;
        const uint32_t col = (key >> 6) & 0x1FFu;
        const uint32_t lower = sensor | ((max(col, 3u) - 3u) << 6);
        const uint32_t upper = sensor + min((col + 2u) << 6, 0x8000u);
        unsigned lo = B.start[bi], hi = S.first_isolated[bi];
        const unsigned end = hi;
        while (lo < hi) {
          const unsigned mid = (lo + hi) >> 1;
          if (((lw[mid] >> 8) & 0xFFFFu) < lower) lo = mid + 1;
          else hi = mid;
        }
        const int32_t row_lower = static_cast<int32_t>(key & 0x3Fu) - 2;
        const int32_t col_lower = static_cast<int32_t>(col) - 3;
        for (unsigned k = lo; k < end; ++k) {
          const uint32_t o = lw[k];
          if (((o >> 8) & 0xFFFFu) >= upper) break;
          window_acc_same_sensor(row_lower, col_lower, o, win);
        }
      }
      candidates_round(sp_word, sensor_number, f, win);
    }
  }
  else {
    // Larger module pairs: every candidate scans its bank, tile by tile through the slice.
    for (unsigned base = 0; base < n_words; base += wave_size) {
      const unsigned w = base + lane_id();
      const bool active = w < n_words;
      const unsigned bi = active ? mp::bank_of(B, w) : 0;
      const uint32_t sp_word = active ? B.word(bi)[w - B.start[bi]] : 0x80000000u;
      uint32_t f = 0;
      if (active && !(sp_word & 0x80000000u)) {
        const uint32_t* wp = B.word(bi) + (w - B.start[bi]);
        f = sp_flags[static_cast<size_t>(reinterpret_cast<const char*>(wp) - raw) / 4];
      }
      const uint32_t pair = B.source[bi] & 0x1FFu;
      const uint32_t sensor_number = v23 ? pair : ((pair << 1) | ((sp_word >> 23) & 0x1u));
      const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
      const int32_t sp_row_lower = static_cast<int32_t>(sp_addr & 0x3Fu) - 2;
      const int32_t sp_col_lower = static_cast<int32_t>(sp_addr >> 6) - 3;
      Window win {0, 0, 0, 0, 0};
      // Both candidates of an SP share its window.
      mp::for_words(B, n_words, resident, lw, (f & 3) != 0, B.start[bi], B.start[bi + 1], [&](uint32_t other) {
        window_acc<version>(sp_row_lower, sp_col_lower, sensor_number, other, win);
      });
      candidates_round(sp_word, sensor_number, f, win);
    }
  }
  n = min(n, estimate);
  mp::wave_sync_lds();
  if (lane_id() == 0) module_cluster_num[idx] = n;
  const Clusters out {sorted_container, total};
  if (clusters_in_lds) {
    for (unsigned i = lane_id(); i < n; i += wave_size) {
      const uint64_t key = lkey[i];
      ClusterOut c;
      c.id = static_cast<uint32_t>(key);
      c.phi = static_cast<int16_t>(static_cast<uint16_t>(key >> 32) ^ 0x8000u);
      c.x = lxyz[3 * i];
      c.y = lxyz[3 * i + 1];
      c.z = lxyz[3 * i + 2];
      const unsigned rank = keys_below(lkey, n, key);
      store_cluster(out, start + rank, c);
    }
    return;
  }
  // Too many clusters for the slice: they sit in the scratch container; rank them against
  // tiles of (phi, id) staged through the slice.
  __threadfence_block();
  for (unsigned i0 = 0; i0 < n; i0 += wave_size) {
    const unsigned i = i0 + lane_id();
    const bool mine = i < n;
    ClusterOut c;
    if (mine) {
      c.id = scratch.ids()[start + i];
      c.phi = scratch.phis()[start + i];
      c.x = scratch.xyz()[3 * (size_t)(start + i)];
      c.y = scratch.xyz()[3 * (size_t)(start + i) + 1];
      c.z = scratch.xyz()[3 * (size_t)(start + i) + 2];
    }
    unsigned rank = 0;
    for (unsigned t0 = 0; t0 < n; t0 += mp::cluster_cap) {
      const unsigned tn = min(mp::cluster_cap, n - t0);
      mp::wave_sync_lds();
      for (unsigned j = lane_id(); j < tn; j += wave_size)
        lkey[j] = phi_id_key(scratch.phis()[start + t0 + j], scratch.ids()[start + t0 + j]);
      mp::wave_sync_lds();
      if (mine) {
        const uint64_t key = phi_id_key(c.phi, c.id);
        rank += keys_below(lkey, tn, key);
      }
    }
    if (mine) store_cluster(out, start + rank, c);
  }
}

// One wave per item (loop = false, the default), or a resident grid (loop = true) whose waves
// take items idx, idx + stride, ... and fetch the next item's bank directory record before
// working on the current one. Six workgroups per CU (80 VGPRs, no spills for one wave per item:
// the sort stages its keys and entries in the cluster slices rather than in registers).
template<int version, bool loop>
__global__ void __launch_bounds__(mp::waves_per_block* wave_size, 6) k_cluster_mp(
  const char* raw,
  const BankEntry* directory,
  const uint32_t* directory_count,
  unsigned n_events,
  const uint32_t* offsets_estimated_input_size,
  const uint8_t* sp_flags,
  const Geometry* geometry,
  const uint8_t* sp_patterns,
  const float* sp_fx,
  const float* sp_fy,
  uint32_t* module_cluster_num,
  char* sorted_container,
  char* unsorted_scratch)
{
  __shared__ uint32_t s_words[mp::waves_per_block][mp::word_cap];
  __shared__ uint64_t s_key[mp::waves_per_block][mp::cluster_cap];
  __shared__ float s_xyz[mp::waves_per_block][3 * mp::cluster_cap];
  __shared__ uint8_t s_cand[mp::waves_per_block][2 * wave_size];
  __shared__ mp::Banks s_banks[mp::waves_per_block];
  __shared__ mp::SortedBanks s_sorted[mp::waves_per_block];
  const unsigned wave = threadIdx.x / wave_size;
  const unsigned n_items = n_events * n_module_pairs;
  const unsigned stride = gridDim.x * mp::waves_per_block;
  unsigned idx = blockIdx.x * mp::waves_per_block + wave;
  if (idx >= n_items) return;
  const unsigned total = offsets_estimated_input_size[n_items];
  const ClusterSlices sl {s_words[wave], s_key[wave], s_xyz[wave], s_cand[wave]};
  if constexpr (!loop) {
    cluster_item<version>(
      raw, mp::fetch_dir(directory, directory_count, idx), idx, total, offsets_estimated_input_size, sp_flags,
      *geometry, sp_patterns, sp_fx, sp_fy, module_cluster_num, sorted_container, unsorted_scratch, s_banks[wave], s_sorted[wave], sl);
    return;
  }
  mp::DirFetch cur = mp::fetch_dir(directory, directory_count, idx);
  for (; idx < n_items; idx += stride) {
    mp::DirFetch next = cur;
    if (idx + stride < n_items) next = mp::fetch_dir(directory, directory_count, idx + stride);
    cluster_item<version>(
      raw, cur, idx, total, offsets_estimated_input_size, sp_flags, *geometry, sp_patterns, sp_fx, sp_fy,
      module_cluster_num, sorted_container, unsorted_scratch, s_banks[wave], s_sorted[wave], sl);
    cur = next;
  }
}

// One wave per item, or (resident) at most the blocks that stay resident, whose waves loop
// over the items. One wave per item balances uneven module pairs better (the dispatcher starts
// a new wave wherever one ends), and is the default.
static unsigned mp_grid(const void* kernel, unsigned n_items, bool resident)
{
  const unsigned need = (n_items + mp::waves_per_block - 1) / mp::waves_per_block;
  if (!resident) return need;
  const unsigned blocks = device_resident_blocks(kernel, mp::waves_per_block * wave_size, 0);
  return blocks ? min(need, blocks) : need;
}

hipError_t launch_estimate_mp(const EstimateMpLaunch& a, hipStream_t s)
{
  if (!a.n_events) return hipSuccess;
  const unsigned v = a.version == 2 || a.version == 3 ? a.version : 4;
  const void* kernel = v == 2 ? reinterpret_cast<const void*>(&k_estimate_mp<2>)
                     : v == 3 ? reinterpret_cast<const void*>(&k_estimate_mp<3>)
                              : reinterpret_cast<const void*>(&k_estimate_mp<4>);
  const unsigned grid = mp_grid(kernel, a.n_events * n_module_pairs, a.resident_grid);
#define VELO_LAUNCH_EMP(V)                                                                                          \
  hipLaunchKernelGGL(                                                                                               \
    k_estimate_mp<V>, dim3(grid), dim3(mp::waves_per_block * wave_size), 0, s, a.raw,                               \
    static_cast<const BankEntry*>(a.directory), a.directory_count, a.n_events, a.estimated_input_size, a.sp_flags)
  switch (a.version) {
  case 2: VELO_LAUNCH_EMP(2); break;
  case 3: VELO_LAUNCH_EMP(3); break;
  default: VELO_LAUNCH_EMP(4); break;
  }
#undef VELO_LAUNCH_EMP
  return hipGetLastError();
}

hipError_t launch_cluster_mp(const ClusterMpLaunch& a, hipStream_t s)
{
  if (!a.n_events) return hipSuccess;
  const unsigned v = a.version == 2 || a.version == 3 ? a.version : 4;
  const void* kernel = v == 2 ? reinterpret_cast<const void*>(&k_cluster_mp<2, true>)
                     : v == 3 ? reinterpret_cast<const void*>(&k_cluster_mp<3, true>)
                              : reinterpret_cast<const void*>(&k_cluster_mp<4, true>);
  const unsigned grid = mp_grid(kernel, a.n_events * n_module_pairs, a.resident_grid);
#define VELO_LAUNCH_CMP(V, L)                                                                                       \
  hipLaunchKernelGGL(                                                                                               \
    (k_cluster_mp<V, L>), dim3(grid), dim3(mp::waves_per_block * wave_size), 0, s, a.raw,                            \
    static_cast<const BankEntry*>(a.directory), a.directory_count, a.n_events, a.offsets_estimated_input_size, a.sp_flags, reinterpret_cast<const Geometry*>(a.geometry),           \
    a.sp_patterns, a.sp_fx, a.sp_fy, a.module_cluster_num, a.sorted_container, a.unsorted_scratch)
  switch (a.version * 2 + (a.resident_grid ? 1 : 0)) {
  case 4: VELO_LAUNCH_CMP(2, false); break;
  case 5: VELO_LAUNCH_CMP(2, true); break;
  case 6: VELO_LAUNCH_CMP(3, false); break;
  case 7: VELO_LAUNCH_CMP(3, true); break;
  default:
    if (a.resident_grid) VELO_LAUNCH_CMP(4, true);
    else VELO_LAUNCH_CMP(4, false);
    break;
  }
#undef VELO_LAUNCH_CMP
  return hipGetLastError();
}

} // namespace velo
