// search_by_triplet and the three-hit filter for gfx950.
//
//   k_search_by_triplet     SearchByTriplet.cu:110-609   one workgroup per event
//   k_three_hit_filter      ThreeHitTracksFilter.cu:48-169
//
// The reference appends seeds, tracks to follow, promoted tracks and three-hit tracks with
// atomicAdd, so its GPU output order depends on scheduling. Here every append is an ordered
// compaction (wave ballot + popcount, then a prefix over the waves of the block) over the
// items in the order the CPU target visits them (h1 index order in seeding, ring order in
// forwarding). Ring slots, track numbers and three-hit slots therefore land exactly where the
// CPU target puts them, and everything downstream is bit-identical.
//
// The event's hits (x, y, z, phi, used) live in LDS when they fit the per-launch budget;
// a larger event runs the same code on the global containers.
#include "velo_common.cuh"
#include "velo_kernels.h"
#include "velo_mfma.cuh"

namespace velo {

// LDS bytes per hit of the full layout: x, y, z (f32), phi (i16), flags (u8).
constexpr unsigned search_bytes_per_hit = 15;
// Default capacity of the list of unused hits of the previous module pair used by seeding step 1
// (a launch parameter, at most 1024); a larger module pair takes the pendulum walk.
constexpr unsigned unused_list_cap = 1024;
// Default number of tracks to follow kept in LDS (a launch parameter: a power of two from 64 to
// 512): the last ring_window positions of the ring (written through to
// the ring in HBM, which holds all max_tracks_to_follow).
constexpr unsigned ring_window = 512;
constexpr unsigned ring_window_bytes = 4 + 4 + 2 + 1; // entry, last two hits, first hit, number of hits
static_assert(max_tracks_to_follow % ring_window == 0, "window positions follow ring positions");
static_assert(ring_window % 16 == 0, "the arrays after the window stay aligned");

// Three 0/1 flags scanned at once over the block. Exclusive ranks in r[], totals in t[].
// One barrier: consecutive scans must alternate between two scratch buffers (a scan's reads of
// its buffer are ordered before the next scan's barrier, hence before the reuse two scans on).
struct Scan3 {
  uint32_t r[3];
  uint32_t t[3];
};
__device__ __forceinline__ Scan3 block_flag_scan3(bool f0, bool f1, bool f2, uint32_t* scratch)
{
  if (blockDim.x == wave_size) {
    // One wave: ranks straight from the ballots.
    const uint64_t below = lanes_below_mask();
    const uint64_t b0 = __ballot(f0), b1 = __ballot(f1), b2 = __ballot(f2);
    Scan3 s;
    s.r[0] = __popcll(b0 & below);
    s.r[1] = __popcll(b1 & below);
    s.r[2] = __popcll(b2 & below);
    s.t[0] = __popcll(b0);
    s.t[1] = __popcll(b1);
    s.t[2] = __popcll(b2);
    return s;
  }
  const unsigned wave = threadIdx.x / wave_size;
  const unsigned n_waves = blockDim.x / wave_size;
  const uint64_t below = lanes_below_mask();
  const uint64_t b0 = __ballot(f0), b1 = __ballot(f1), b2 = __ballot(f2);
  if (lane_id() == 0) {
    scratch[3 * wave] = __popcll(b0);
    scratch[3 * wave + 1] = __popcll(b1);
    scratch[3 * wave + 2] = __popcll(b2);
  }
  __syncthreads();
  Scan3 s;
  s.r[0] = __popcll(b0 & below);
  s.r[1] = __popcll(b1 & below);
  s.r[2] = __popcll(b2 & below);
  s.t[0] = s.t[1] = s.t[2] = 0;
  for (unsigned w = 0; w < n_waves; ++w) {
    const uint32_t c0 = scratch[3 * w], c1 = scratch[3 * w + 1], c2 = scratch[3 * w + 2];
    if (w < wave) {
      s.r[0] += c0;
      s.r[1] += c1;
      s.r[2] += c2;
    }
    s.t[0] += c0;
    s.t[1] += c1;
    s.t[2] += c2;
  }
  return s;
}

struct ModulePairData {
  unsigned hit_start;
  unsigned hit_num;
  float z[2];
  unsigned table; // offset of the module pair's phi bucket table, or no_phi_table
};

// Phi bucket tables of an event's module pairs (LDS, u8 entries, module pairs of at most 255
// hits): entry b < phi_buckets is the first position whose phi lies in bucket b or above (bucket
// of an int16 phi: (phi + 32768) >> 10), entry phi_buckets the number of hits.
constexpr unsigned phi_buckets = 64;
constexpr unsigned phi_table_stride = phi_buckets + 1;
constexpr unsigned phi_tables_bytes = (n_module_pairs * phi_table_stride + 15) & ~15u;
constexpr unsigned no_phi_table = 0xFFFFFFFFu;

// Hits of one event, structure-of-arrays in LDS.
// Per-hit flags byte of the LDS layouts: bit 0 = used (dev_hit_used), bit 1 = module parity of
// the hit's sensor, ((id >> 18) & 0xFF) / 4 % 2. Marking a hit used ORs bit 0 in; lanes that mark
// the same hit concurrently all store the same byte, so the read-modify-write cannot lose a bit.
constexpr uint8_t flag_used = 1, flag_odd = 2;

struct LdsHits {
  const float* xs;
  const float* ys;
  const float* zs;
  const int16_t* ph;
  uint8_t* flags;
  const uint8_t* tables;
  __device__ __forceinline__ float x(unsigned i) const { return xs[i]; }
  __device__ __forceinline__ float y(unsigned i) const { return ys[i]; }
  __device__ __forceinline__ float z(unsigned i) const { return zs[i]; }
  __device__ __forceinline__ int16_t phi(unsigned i) const { return ph[i]; }
  __device__ __forceinline__ const int16_t* phi_begin() const { return ph; }
  __device__ __forceinline__ bool used(unsigned i) const { return (flags[i] & flag_used) != 0; }
  __device__ __forceinline__ void mark_used(unsigned i) const { flags[i] |= flag_used; }
  // Four hits at once: all four flags are read before any is written, so the four LDS round
  // trips overlap (equal indices write equal values).
  __device__ __forceinline__ void mark_used4(unsigned a, unsigned b, unsigned c, unsigned d) const
  {
    const uint8_t fa = flags[a], fb = flags[b], fc = flags[c], fd = flags[d];
    flags[a] = fa | flag_used;
    flags[b] = fb | flag_used;
    flags[c] = fc | flag_used;
    flags[d] = fd | flag_used;
  }
  __device__ __forceinline__ bool module_odd(unsigned i) const { return (flags[i] & flag_odd) != 0; }
};

// Module parity of a hit from its LHCbID (SearchByTriplet.cu:498-499).
__device__ __forceinline__ uint8_t module_parity(uint32_t id) { return static_cast<uint8_t>(((id & 0x3fc0000u) >> 18) / 4 % 2); }

// Hits of one event read from the sorted cluster container in global memory.
struct GlobalHits {
  const float* xyz;
  const int16_t* ph;
  uint8_t* used_flags; // dev_hit_used
  const uint32_t* ids;
  const uint8_t* tables;
  __device__ __forceinline__ float x(unsigned i) const { return xyz[3 * i]; }
  __device__ __forceinline__ float y(unsigned i) const { return xyz[3 * i + 1]; }
  __device__ __forceinline__ float z(unsigned i) const { return xyz[3 * i + 2]; }
  __device__ __forceinline__ int16_t phi(unsigned i) const { return ph[i]; }
  __device__ __forceinline__ const int16_t* phi_begin() const { return ph; }
  __device__ __forceinline__ bool used(unsigned i) const { return used_flags[i] != 0; }
  __device__ __forceinline__ void mark_used(unsigned i) const { used_flags[i] = 1; }
  __device__ __forceinline__ void mark_used4(unsigned a, unsigned b, unsigned c, unsigned d) const
  {
    used_flags[a] = 1;
    used_flags[b] = 1;
    used_flags[c] = 1;
    used_flags[d] = 1;
  }
  __device__ __forceinline__ bool module_odd(unsigned i) const { return module_parity(ids[i]) != 0; }
};

// Work counters of one thread (seeding evaluations, forwarding evaluations, extrapolations,
// search probes); only kept by the counting instantiation of the kernel.
struct WorkCount {
  uint32_t seed_eval = 0, fwd_eval = 0, extrap = 0, steps = 0;
  // Wall-clock ticks spent by the workgroup in each Phase, kept by thread 0 (which takes part
  // in every phase, so its marks follow the slowest wave only at barriers).
  unsigned long long phase[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long last = 0;
};

// forwarding; h1 compaction; seeding step 1 split into the h0 search (h1 and phi reads, then
// the search in the previous module pair), the pendulum walk, and the wait at the barrier that
// ends the step; seeding steps 2 and 3; the rest.
enum Phase {
  phase_forwarding,
  phase_compaction,
  phase_seed1_search,
  phase_seed1_pendulum,
  phase_seed1_wait,
  phase_seed2,
  phase_seed3,
  phase_other,
  n_phases
};

template<bool count>
__device__ __forceinline__ void phase_mark(WorkCount& wc, Phase p)
{
  if constexpr (count) {
    if (threadIdx.x == 0) {
      const unsigned long long now = wall_clock64();
      wc.phase[p] += now - wc.last;
      wc.last = now;
    }
  }
}

// First index i of the sorted array with array[i] >= value (the leftmost binary search of
// BinarySearch.cuh, same result on every input). Eight-way: each round probes seven split
// points at once (independent LDS reads) and keeps the eighth of the range that holds the
// answer, so a module pair of n hits takes about log8(n) dependent round trips instead of
// log2(n). `steps` counts the probes.
template<bool count, typename T>
__device__ __forceinline__ int counted_search(const T* array, unsigned size, T value, WorkCount& wc)
{
  unsigned lo = 0, hi = size; // answer in [lo, hi]; array[< lo] < value <= array[>= hi]
  while (lo < hi) {
    const unsigned span = hi - lo;
    unsigned p[7];
    bool below[7];
    T v[7];
#pragma unroll
    for (unsigned j = 0; j < 7; ++j) {
      p[j] = lo + ((j + 1) * span) / 8; // lo <= p[j] < hi, non-decreasing in j
      v[j] = array[p[j]];
    }
    // All seven reads are issued before the first compare waits for one (the scheduler would
    // otherwise interleave read, wait and compare, one LDS round trip per probe).
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (unsigned j = 0; j < 7; ++j) below[j] = v[j] < value; // true for a prefix of j (sorted array)
    if constexpr (count) wc.steps += 7;
    unsigned nlo = lo, nhi = hi;
#pragma unroll
    for (unsigned j = 0; j < 7; ++j)
      if (below[j]) nlo = p[j] + 1;
#pragma unroll
    for (int j = 6; j >= 0; --j)
      if (!below[j]) nhi = p[j];
    lo = nlo;
    hi = nhi;
  }
  return static_cast<int>(lo);
}

// The same lower bound in a module pair's phi array through its bucket table: positions before
// the value's bucket hold smaller phis and positions after it larger ones, so the answer is the
// bucket's start plus the bucket's elements below the value (read directly for up to four, by
// the 8-way search for more). Module pairs without a table take the 8-way search.
template<bool count, typename H>
__device__ __forceinline__ int mp_search(const H& hits, const ModulePairData& mp, int16_t value, WorkCount& wc)
{
  const int16_t* a = hits.phi_begin() + mp.hit_start;
  if (mp.table == no_phi_table) return counted_search<count>(a, mp.hit_num, value, wc);
  const unsigned b = static_cast<unsigned>(static_cast<uint16_t>(value) ^ 0x8000u) >> 10;
  const uint8_t* t = hits.tables + mp.table;
  const unsigned lo = t[b], hi = t[b + 1];
  if (hi - lo > 4) return static_cast<int>(lo) + counted_search<count>(a + lo, hi - lo, value, wc);
  unsigned c = lo;
#pragma unroll
  for (unsigned k = 0; k < 4; ++k)
    if (lo + k < hi && a[lo + k] < value) ++c;
  return static_cast<int>(c);
}

// Entry b of module pair mp's table: the lower bound of bucket b's first phi (b < phi_buckets),
// or the number of hits.
template<typename H>
__device__ __forceinline__ uint8_t phi_table_entry(const H& hits, const ModulePairData& mp, unsigned b)
{
  WorkCount wc;
  return static_cast<uint8_t>(
    b == phi_buckets ? mp.hit_num
                     : counted_search<false>(hits.phi_begin() + mp.hit_start, mp.hit_num,
                                             static_cast<int16_t>(static_cast<int>(b << 10) - 32768), wc));
}

// Builds the tables of module pairs [m0, n_module_pairs) that have one (whole block; the
// module-pair descriptors are in LDS and visible).
template<typename H>
__device__ __forceinline__ void build_phi_tables(const H& hits, const ModulePairData* mp, uint8_t* tables, unsigned m0)
{
  for (unsigned t = m0 * phi_table_stride + threadIdx.x; t < n_module_pairs * phi_table_stride; t += blockDim.x) {
    const unsigned m = t / phi_table_stride;
    if (mp[m].table != no_phi_table) tables[t] = phi_table_entry(hits, mp[m], t % phi_table_stride);
  }
}

// Module pairs whose tables are built at the start of an event; every other table is built
// during an earlier seeding (see track_seeding's `build`), before its first use.
constexpr unsigned phi_tables_at_start = 4;

// Hits of one event: phi and used flags in LDS, coordinates in global memory (events too large
// for the full LDS layout).
struct MixedHits {
  const float* xyz;
  const int16_t* ph;
  uint8_t* flags;
  const uint8_t* tables;
  __device__ __forceinline__ float x(unsigned i) const { return xyz[3 * i]; }
  __device__ __forceinline__ float y(unsigned i) const { return xyz[3 * i + 1]; }
  __device__ __forceinline__ float z(unsigned i) const { return xyz[3 * i + 2]; }
  __device__ __forceinline__ int16_t phi(unsigned i) const { return ph[i]; }
  __device__ __forceinline__ const int16_t* phi_begin() const { return ph; }
  __device__ __forceinline__ bool used(unsigned i) const { return (flags[i] & flag_used) != 0; }
  __device__ __forceinline__ void mark_used(unsigned i) const { flags[i] |= flag_used; }
  // Four hits at once: all four flags are read before any is written, so the four LDS round
  // trips overlap (equal indices write equal values).
  __device__ __forceinline__ void mark_used4(unsigned a, unsigned b, unsigned c, unsigned d) const
  {
    const uint8_t fa = flags[a], fb = flags[b], fc = flags[c], fd = flags[d];
    flags[a] = fa | flag_used;
    flags[b] = fb | flag_used;
    flags[c] = fc | flag_used;
    flags[d] = fd | flag_used;
  }
  __device__ __forceinline__ bool module_odd(unsigned i) const { return (flags[i] & flag_odd) != 0; }
};

struct SearchShared {
  ModulePairData mp[n_module_pairs]; // hit range (event-relative) and module z of every module pair
  uint32_t scan[2][3 * 16 + 4];      // block_flag_scan3 scratch, used alternately
};

// The event's counters, identical in every lane of the workgroup (updated from block-wide
// scan totals), so no lane has to publish them through LDS.
struct Counters {
  uint32_t ttf = 0;        // atomics::tracks_to_follow
  uint32_t three = 0;      // atomics::number_of_three_hit_tracks
  uint32_t local_hits = 0; // atomics::local_number_of_hits
  uint32_t n_tracks = 0;   // dev_number_of_velo_tracks[event]
  uint32_t status = 0;
  unsigned scan_buffer = 0;
};

__device__ __forceinline__ Scan3 scan3(SearchShared& sh, Counters& c, bool f0, bool f1, bool f2)
{
  const Scan3 s = block_flag_scan3(f0, f1, f2, sh.scan[c.scan_buffer]);
  c.scan_buffer ^= 1u;
  return s;
}

// Seeding scratch for one chunk of h1 hits, blockDim entries each, in dynamic LDS.
// Seeding scratch of a chunk of blockDim h1 hits: per h1, its h0 candidates (up to
// max_h0_candidates), their number and the chosen seed (h0 | h2 << 16, or no_seed).
static_assert(max_h0_candidates >= initial_seeding_h0_candidates, "room for every h0 candidate");
static_assert(initial_seeding_h0_candidates == 8 && seeding_h0_candidates == 4, "track_seeding splits lanes by K = 4 or 8");
constexpr uint32_t no_seed = 0xFFFFFFFFu;
constexpr unsigned seed_scratch_bytes_per_lane = 2 * max_h0_candidates + 4 + 2 + 1;
struct SeedScratch {
  uint16_t* h0;
  uint32_t* seed;
  uint16_t* h1_chunk; // unused hits of the current module pair, one compaction chunk
  uint8_t* n_h0;
  uint16_t* unused;   // unused hits of the previous module pair (relative indices, ascending)
  uint16_t* before;   // per position of the previous module pair: the unused hits before it
  unsigned list_cap;  // entries of `unused` and `before`
};

struct ScorerLds;

struct EventOutputs {
  TrackHits* tracks;
  TrackletHits* three_hit_tracks;
  TrackletHits* tracklets;
  uint32_t* ttf;        // the ring of tracks to follow in HBM (max_tracks_to_follow entries)
  uint32_t* ttf_window; // its last `window` positions, in LDS (position p at p & window_mask)
  // Beside each window entry, what forwarding reads of its track: the last two hits (low and
  // high 16 bits), the first hit (read for a three-hit seed) and the number of hits.
  uint32_t* ttf_last2;
  uint16_t* ttf_first;
  uint8_t* ttf_n;
  uint16_t* h1_indices;
  unsigned track_capacity;
  unsigned event_number;
  SeedLog seed_log;
  const ScorerLds* scorer;
  float scorer_threshold;
  unsigned window;      // LDS window positions (a power of two dividing max_tracks_to_follow)
  unsigned window_mask; // window - 1
};

// Four consecutive positions of an h2 window, read at once: position j of the window is hit
// (c + j) % n of the module pair (SearchByTriplet.cu:393-395, 520-522); positions past the
// module pair read the window's first hit and are not used.
struct Window4 {
  unsigned index[4];
  int16_t phi[4];
  bool used[4];
  float x[4], y[4], z[4];
  template<typename H>
  __device__ __forceinline__ void
  load(const H& hits, unsigned start, unsigned n, unsigned c, unsigned j0, bool with_used)
  {
    // c <= n (a lower bound) and j < n, so c + j < 2 n and one conditional subtraction wraps it.
#pragma unroll
    for (unsigned k = 0; k < 4; ++k) {
      const unsigned j = j0 + k < n ? j0 + k : j0;
      const unsigned w = c + j;
      index[k] = start + (w >= n ? w - n : w);
    }
#pragma unroll
    for (unsigned k = 0; k < 4; ++k) {
      phi[k] = hits.phi(index[k]);
      used[k] = with_used ? hits.used(index[k]) : false;
      x[k] = hits.x(index[k]);
      y[k] = hits.y(index[k]);
      z[k] = hits.z(index[k]);
    }
  }
};

// SearchByTriplet.cu:272-292
template<bool count, typename H>
__device__ __forceinline__ void find_forward_candidate(
  WorkCount& wc,
  const H& hits,
  const ModulePairData& mp,
  float h0x,
  float h0y,
  float tx,
  float ty,
  float dz,
  int16_t phi_tolerance,
  int16_t& candidate,
  int16_t& extrapolated_phi)
{
  const float predx = tx * dz;
  const float predy = ty * dz;
  const float xp = h0x + predx;
  const float yp = h0y + predy;
  const int16_t phi = hit_phi_16(xp, yp);
  if constexpr (count) ++wc.extrap;
  candidate = static_cast<int16_t>(mp_search<count>(hits, mp, static_cast<int16_t>(phi - phi_tolerance), wc));
  extrapolated_phi = phi;
}

// The learned scorer's weights in LDS (scored mode): W1 (K padded to 16) and W2 as bf16, the
// biases and w3 (bf16-rounded) as FP32.
// The scorer's weights in LDS, transposed for the products below: w1t[j][k] = bf16(W1[k][j])
// (k < 16, zero from 10), w2t[j2][j1] = bf16(W2[j1][j2]); biases and bf16(w3) in FP32.
struct ScorerLds {
  uint16_t w1t[32 * 16];
  uint16_t w2t[32 * 32];
  float b1[32], b2[32], w3[32];
  float b3;
};
constexpr unsigned scorer_lds_bytes = (sizeof(ScorerLds) + 15) & ~15u;

__device__ void load_scorer(const float* p, ScorerLds& w)
{
  using namespace mlp;
  for (unsigned t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
    const unsigned j = t / 32, k = t % 32;
    w.w2t[t] = to_bf16(p[param_w2 + k * 32 + j]);
    if (t < 32 * 16) {
      const unsigned jj = t / 16, kk = t % 16;
      w.w1t[t] = kk < n_inputs ? to_bf16(p[param_w1 + kk * 32 + jj]) : 0;
    }
  }
  for (unsigned t = threadIdx.x; t < 32; t += blockDim.x) {
    w.b1[t] = p[param_b1 + t];
    w.b2[t] = p[param_b2 + t];
    w.w3[t] = from_bf16(to_bf16(p[param_w3 + t]));
  }
  if (threadIdx.x == 0) w.b3 = p[param_b3];
}

__device__ __forceinline__ mlp::bf16x4 bf16x4_from(uint2 v)
{
  mlp::bf16x4 r;
  r[0] = static_cast<short>(v.x & 0xFFFFu);
  r[1] = static_cast<short>(v.x >> 16);
  r[2] = static_cast<short>(v.y & 0xFFFFu);
  r[3] = static_cast<short>(v.y >> 16);
  return r;
}

// Logit of the seed of lane l (features fp[0..9]) for the lanes set in `live` (a ballot, so
// wave-uniform), with the products of k_mlp_score computed transposed: Z1^T = W1b^T xb^T and
// Z2^T = W2b^T A1b^T on v_mfma_f32_32x32x8_bf16. Seeds 32 rb + c form row block rb (computed
// only when one of its seeds is live), with seed c in column c. In that layout the hidden layer
// leaves the first product exactly where the second takes its operand (element 4 s + i of
// lane l is hidden unit 8 s + 4 (l / 32) + i of seed l % 32), and the output layer is a sum
// over the lane's 16 hidden units plus one exchange with lane l ^ 32 (the other 16), so the
// wave needs no LDS beyond the weights. The output sum runs in another order than
// k_mlp_score's. The whole wave must be active.
__device__ __forceinline__ float scorer_logit(const float* fp, uint64_t live, const ScorerLds& w)
{
  using namespace mlp;
  const unsigned l = lane_id(), h = l / 32, c = l % 32;
  // Features of the seed in lane l ^ 32: lanes l < 32 send features 4..7 (the operand of the
  // upper half in row block 0), lanes l >= 32 send 0..3, 8 and 9 (that of the lower half in row
  // block 1).
  float other[6];
#pragma unroll
  for (unsigned i = 0; i < 4; ++i) other[i] = __shfl_xor(h ? fp[i] : fp[4 + i], 32, wave_size);
  other[4] = __shfl_xor(fp[8], 32, wave_size);
  other[5] = __shfl_xor(fp[9], 32, wave_size);
  float logit = 0.f;
#pragma unroll
  for (unsigned rb = 0; rb < 2; ++rb) {
    if (!static_cast<uint32_t>(live >> (32 * rb))) continue;
    const bool own = h == rb; // this lane's seed is column c of the row block
    // B operand of the first product: features 4 h + i (k-step 0) and 8 + 4 h + i (k-step 1).
    bf16x4 x0, x1;
#pragma unroll
    for (unsigned i = 0; i < 4; ++i) {
      const float f0 = own ? fp[4 * h + i] : other[i];
      x0[i] = static_cast<short>(to_bf16(f0));
      const float f1 = h ? 0.f : (i < 2 ? (own ? fp[8 + i] : other[4 + i]) : 0.f);
      x1[i] = static_cast<short>(to_bf16(f1));
    }
    f32x16 acc = zero16();
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(
      bf16x4_from(*reinterpret_cast<const uint2*>(w.w1t + 16 * c + 4 * h)), x0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(
      bf16x4_from(*reinterpret_cast<const uint2*>(w.w1t + 16 * c + 8 + 4 * h)), x1, acc, 0, 0, 0);
    f32x16 acc2 = zero16();
#pragma unroll
    for (unsigned st = 0; st < 4; ++st) {
      const float4 b1 = *reinterpret_cast<const float4*>(w.b1 + 8 * st + 4 * h);
      const float z[4] = {acc[4 * st] + b1.x, acc[4 * st + 1] + b1.y, acc[4 * st + 2] + b1.z,
                          acc[4 * st + 3] + b1.w};
      bf16x4 a1;
#pragma unroll
      for (unsigned i = 0; i < 4; ++i) a1[i] = static_cast<short>(to_bf16(z[i] > 0.f ? z[i] : 0.f));
      acc2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(
        bf16x4_from(*reinterpret_cast<const uint2*>(w.w2t + 32 * c + 8 * st + 4 * h)), a1, acc2, 0, 0, 0);
    }
    float part = 0.f;
#pragma unroll
    for (unsigned st = 0; st < 4; ++st) {
      const float4 b2 = *reinterpret_cast<const float4*>(w.b2 + 8 * st + 4 * h);
      const float4 w3 = *reinterpret_cast<const float4*>(w.w3 + 8 * st + 4 * h);
      const float z[4] = {acc2[4 * st] + b2.x, acc2[4 * st + 1] + b2.y, acc2[4 * st + 2] + b2.z,
                          acc2[4 * st + 3] + b2.w};
      const float ww[4] = {w3.x, w3.y, w3.z, w3.w};
#pragma unroll
      for (unsigned i = 0; i < 4; ++i) part += (z[i] > 0.f ? z[i] : 0.f) * ww[i];
    }
    const float sum = part + __shfl_xor(part, 32, wave_size);
    if (own) logit = sum + w.b3;
  }
  return logit;
}

// Scorer features of a seed (h0, h1, h2) whose h1 lies in module pair `module_pair` (no
// reference counterpart): h2's residual to the (h0, h1) line and h1's residual to the (h0, h2)
// line in x and y, the phi steps h0 -> h1 and h1 -> h2, the (h0, h1) slopes, the scatter and the
// module pair, each scaled to order one.
template<typename H>
__device__ __forceinline__ void
seed_features(const H& hits, unsigned h0, unsigned h1, unsigned h2, unsigned module_pair, float* f)
{
  const float x0 = hits.x(h0), y0 = hits.y(h0), z0 = hits.z(h0);
  const float x1 = hits.x(h1), y1 = hits.y(h1), z1 = hits.z(h1);
  const float x2 = hits.x(h2), y2 = hits.y(h2), z2 = hits.z(h2);
  const float td = 1.0f / (z1 - z0);
  const float tx = (x1 - x0) * td;
  const float ty = (y1 - y0) * td;
  const float dz2 = z2 - z0;
  const float dx2 = (x0 + tx * dz2) - x2;
  const float dy2 = (y0 + ty * dz2) - y2;
  const float td02 = 1.0f / (z2 - z0);
  const float dz1 = z1 - z0;
  const float dx1 = (x0 + ((x2 - x0) * td02) * dz1) - x1;
  const float dy1 = (y0 + ((y2 - y0) * td02) * dz1) - y1;
  const int16_t dphi01 = static_cast<int16_t>(hits.phi(h1) - hits.phi(h0));
  const int16_t dphi12 = static_cast<int16_t>(hits.phi(h2) - hits.phi(h1));
  f[0] = dx2 * 4.0f;
  f[1] = dy2 * 4.0f;
  f[2] = dx1 * 4.0f;
  f[3] = dy1 * 4.0f;
  f[4] = static_cast<float>(dphi01) * (1.0f / 512.0f);
  f[5] = static_cast<float>(dphi12) * (1.0f / 512.0f);
  f[6] = tx * 4.0f;
  f[7] = ty * 4.0f;
  f[8] = (dx2 * dx2 + dy2 * dy2) * 12.5f;
  f[9] = static_cast<float>(module_pair) * (1.0f / 32.0f);
}

// The first K unused hits of the previous module pair in the pendulum order of track_seeding
// (SearchByTriplet.cu:348-370), from the ascending list `unused` (u entries) of the module
// pair's unused hits (relative indices) instead of a walk over every position.
// The pendulum visits position start + off, off = 0, +1, -1, +2, -2, ..., each wrapped once
// into [0, n): position x gets rank 2d - 1 if d = (x - start) mod n lies in [1, n/2] (right of
// start), else 2 ((start - x) mod n) (left, start itself included); ranks are a permutation of
// 0..n-1, and the walk keeps the unused positions of the K smallest ranks, in rank order. Those
// are among the K nearest unused entries on either side of start in the circular list, so the
// 2K entries around its lower bound (all u if u <= 2K) are ranked, sorted by a fixed Batcher
// odd-even merge network on (rank, position) keys, and the first K taken. Writes hit_start +
// position to out[0..) and returns how many (<= K).
template<unsigned K>
__device__ __forceinline__ unsigned h0_candidates_from_list(
  const uint16_t* unused,
  const uint16_t* before,
  unsigned u,
  unsigned n,
  unsigned hit_start,
  int start,
  uint16_t* out)
{
  if (u == 0) return 0;
  const unsigned s0 = static_cast<unsigned>(start) >= n ? static_cast<unsigned>(start) - n : static_cast<unsigned>(start);
  // The lower bound of s0 in the list: the number of unused positions below s0.
  const unsigned p = before[s0];
  // key = rank << 16 | position (n <= 1024 on this path: rank < 2048, position < 1024).
  constexpr unsigned M = 2 * K;
  constexpr unsigned none = 0xFFFFFFFFu;
  const bool all = u <= M;
  unsigned key[M];
#pragma unroll
  for (unsigned j = 0; j < M; ++j) {
    unsigned ix;
    if (all) {
      ix = j < u ? j : 0u;
    }
    else {
      const int q = static_cast<int>(p) + static_cast<int>(j) - static_cast<int>(K);
      ix = static_cast<unsigned>(q < 0 ? q + static_cast<int>(u) : (q >= static_cast<int>(u) ? q - static_cast<int>(u) : q));
    }
    const unsigned x = unused[ix];
    const unsigned d = x >= s0 ? x - s0 : x + n - s0;
    const unsigned r = d == 0 ? 0u : (d <= n / 2 ? 2 * d - 1 : 2 * (n - d));
    key[j] = (!all || j < u) ? (r << 16 | x) : none;
  }
  // Batcher's odd-even merge sort of the M keys (compile-time indices: a fixed network of
  // min/max pairs on registers).
#pragma unroll
  for (unsigned pp = 1; pp < M; pp <<= 1) {
#pragma unroll
    for (unsigned k = pp; k >= 1; k >>= 1) {
#pragma unroll
      for (unsigned j = k % pp; j + k < M; j += 2 * k) {
#pragma unroll
        for (unsigned i = 0; i < k && i + j + k < M; ++i) {
          if ((i + j) / (2 * pp) == (i + j + k) / (2 * pp)) {
            const unsigned a = key[i + j], b = key[i + j + k];
            key[i + j] = min(a, b);
            key[i + j + k] = max(a, b);
          }
        }
      }
    }
  }
  unsigned found = 0;
#pragma unroll
  for (unsigned k = 0; k < K; ++k) {
    if (key[k] != none) {
      out[k] = static_cast<uint16_t>(hit_start + (key[k] & 0xFFFFu));
      found = k + 1;
    }
  }
  return found;
}

// SearchByTriplet.cu:298-446, ordered, with the (h1, h0) pairs spread over lanes.
//
// The reference walks, per h1, its h0 candidates in pendulum order and, per h0, the h2 window
// in phi order, keeping the first strictly smaller scatter. Here a chunk of blockDim unused h1
// hits is processed at once (K = h0 candidates considered): every lane collects the h0
// candidates of one h1, then in rounds of C = blockDim / K h1 hits every lane scores one
// (h1, h0) pair over its h2 window (first strict minimum) and the K lanes of an h1 keep the
// first strict minimum over its h0 candidates in pendulum order by cross-lane exchanges: the
// same pair the sequential double loop selects. The chunk's seeds are then appended in h1 order.
template<bool count, int seed_mode, typename H>
__device__ __forceinline__ void track_seeding(
  WorkCount& wc,
  const H& hits,
  const SeedScratch& ss,
  SearchShared& sh,
  Counters& cn,
  const EventOutputs& o,
  const ModulePairData prev,
  const ModulePairData cur,
  const ModulePairData next,
  unsigned module_pair,
  float max_scatter,
  int16_t phi_tolerance,
  unsigned h0_candidates_to_consider,
  const ModulePairData* build,
  uint8_t* tables)
{
  const unsigned K = h0_candidates_to_consider;
  const unsigned log2K = K == 8 ? 3u : 2u; // K is 4 or 8
  const unsigned C = blockDim.x >> log2K;
  const float dz_next = next.z[0] - prev.z[0];

  // The list of unused hits of the previous module pair for step 1 (its writes are ordered
  // before step 1 by the barrier of the compaction scan below).
  // The first blockDim hits of both lists are compacted by one scan (seeding marks no hit used,
  // so both read the flags the last forwarding left).
  const bool use_list = prev.hit_num <= ss.list_cap;
  const bool free0 = use_list && threadIdx.x < prev.hit_num && !hits.used(prev.hit_start + threadIdx.x);
  const bool act0 = threadIdx.x < cur.hit_num && !hits.used(cur.hit_start + threadIdx.x);
  const Scan3 first = scan3(sh, cn, free0, act0, false);
  if (free0) ss.unused[first.r[0]] = static_cast<uint16_t>(threadIdx.x);
  if (use_list && threadIdx.x < prev.hit_num) ss.before[threadIdx.x] = static_cast<uint16_t>(first.r[0]);
  unsigned n_unused = first.t[0];
  if (use_list) {
    for (unsigned c0 = blockDim.x; c0 < prev.hit_num; c0 += blockDim.x) {
      const unsigned r = c0 + threadIdx.x;
      const bool free_hit = r < prev.hit_num && !hits.used(prev.hit_start + r);
      const Scan3 sc = scan3(sh, cn, free_hit, false, false);
      if (free_hit) ss.unused[n_unused + sc.r[0]] = static_cast<uint16_t>(r);
      if (r < prev.hit_num) ss.before[r] = static_cast<uint16_t>(n_unused + sc.r[0]);
      n_unused += sc.t[0];
    }
  }

  // The phi table of module pair `build` (first searched in the forwarding after the next
  // seeding, so any later barrier orders it), by the block's last wave: during step 1 of the
  // first chunk, where that wave is idle when the chunk has no more h1 hits than the other waves
  // have lanes, or here when the module pair has no hit to seed from.
  const auto build_table = [&]() {
    if (!build || build->table == no_phi_table) return;
    const unsigned builders = blockDim.x > wave_size ? blockDim.x - wave_size : 0u;
    if (threadIdx.x >= builders) {
      const unsigned b = threadIdx.x - builders;
      tables[build->table + b] = phi_table_entry(hits, *build, b);
      if (b == 0) tables[build->table + phi_buckets] = static_cast<uint8_t>(build->hit_num);
    }
  };
  if (cur.hit_num == 0) build_table();

  // Unused hits of the current module pair, compacted in index order one chunk of blockDim
  // hits at a time (dev_rel_indices in global memory, the chunk's list in LDS); the seeds of a
  // chunk are made before the next chunk, which keeps the h1 order of the reference.
  uint32_t base = 0;
  for (unsigned c0 = 0; c0 < cur.hit_num; c0 += blockDim.x) {
    const unsigned r = c0 + threadIdx.x;
    bool act = act0;
    Scan3 sc;
    if (c0 == 0) {
      sc.r[0] = first.r[1];
      sc.t[0] = first.t[1];
    }
    else {
      act = r < cur.hit_num && !hits.used(cur.hit_start + r);
      sc = scan3(sh, cn, act, false, false);
    }
    if (act) {
      o.h1_indices[base + sc.r[0]] = cur.hit_start + r;
      ss.h1_chunk[sc.r[0]] = cur.hit_start + r;
    }
    base += sc.t[0];
    const unsigned n_h1 = sc.t[0];
    __syncthreads();
    phase_mark<count>(wc, phase_compaction);

    if (c0 == 0) build_table();
    // 1. h0 candidates of every h1 of the chunk (lane t: the chunk's h1 number t), pendulum order.
    {
      const unsigned rel = threadIdx.x;
      unsigned n_found = 0;
      if (rel < n_h1) {
        const unsigned h1_index = ss.h1_chunk[rel];
        const int16_t h1_phi = hits.phi(h1_index);
        const int start = mp_search<count>(hits, prev, h1_phi, wc);
        phase_mark<count>(wc, phase_seed1_search);
        uint16_t* h0_out = ss.h0 + rel * max_h0_candidates;
        if (use_list) {
          n_found = K == initial_seeding_h0_candidates
                      ? h0_candidates_from_list<initial_seeding_h0_candidates>(
                          ss.unused, ss.before, n_unused, prev.hit_num, prev.hit_start, start, h0_out)
                      : h0_candidates_from_list<seeding_h0_candidates>(
                          ss.unused, ss.before, n_unused, prev.hit_num, prev.hit_start, start, h0_out);
        }
        else {
          // Module pairs beyond the list's capacity: the pendulum walk itself, start, +1, -1,
          // +2, -2, ... (each wrapped once into the module pair).
          const int n = static_cast<int>(prev.hit_num);
          int phi_index = start;
          for (unsigned i = 0; i < prev.hit_num && n_found < K; ++i) {
            phi_index += (i & 1u) ? static_cast<int>(i) : -static_cast<int>(i);
            const int in_bounds = phi_index < 0 ? phi_index + n : (phi_index >= n ? phi_index - n : phi_index);
            const unsigned h0_index = prev.hit_start + static_cast<unsigned>(in_bounds);
            if (!hits.used(h0_index)) h0_out[n_found++] = h0_index;
          }
        }
        phase_mark<count>(wc, phase_seed1_pendulum);
      }
      ss.n_h0[rel] = n_found;
    }
    __syncthreads();
    phase_mark<count>(wc, phase_seed1_wait);

    // 2. Rounds of C h1 hits: lane t K + k scores the pair (h1 number r C + t, its h0 number k)
    // over the h2 window (first strict minimum), and the K lanes of an h1 (consecutive lanes of
    // one wave) keep the least (scatter, k) by exchanging keys: the first strict minimum over
    // the h0 candidates in pendulum order, the pair the sequential double loop selects. A
    // scatter below max_scatter is a non-negative float, ordered as its bit pattern.
    for (unsigned r0 = 0; r0 < n_h1; r0 += C) {
      const unsigned t = threadIdx.x >> log2K;
      const unsigned k = threadIdx.x & (K - 1);
      const unsigned q = r0 + t;
      float best = max_scatter;
      uint16_t best_h2 = 0;
      // The pair's h1 and h0 are read together with the candidate count (all within the
      // chunk's scratch for q < n_h1), one LDS round trip instead of two.
      const bool in_chunk = t < C && q < n_h1;
      const unsigned qq = in_chunk ? q : 0u;
      const unsigned n_h0_q = ss.n_h0[qq];
      const unsigned h1_index = ss.h1_chunk[qq];
      const unsigned h0_index = ss.h0[qq * max_h0_candidates + k];
      if (in_chunk && k < n_h0_q) {
        const float h1x = hits.x(h1_index), h1y = hits.y(h1_index), h1z = hits.z(h1_index);
        const float h0x = hits.x(h0_index), h0y = hits.y(h0_index), h0z = hits.z(h0_index);
        const float td = 1.0f / (h1z - h0z);
        const float txn = h1x - h0x;
        const float tyn = h1y - h0y;
        const float tx = txn * td;
        const float ty = tyn * td;
        int16_t cand, ephi;
        find_forward_candidate<count>(wc, hits, next, h0x, h0y, tx, ty, dz_next, phi_tolerance, cand, ephi);
        // The h2 window in phi order from the candidate, four positions read per round.
        const unsigned n2 = next.hit_num;
        const unsigned c2 = static_cast<unsigned>(static_cast<int>(cand));
        for (unsigned j0 = 0; j0 < n2; j0 += 4) {
          Window4 w;
          w.load(hits, next.hit_start, n2, c2, j0, true);
          bool stop = false;
#pragma unroll
          for (unsigned kk = 0; kk < 4; ++kk) {
            if (stop || j0 + kk >= n2) continue;
            const int16_t phi_diff = w.phi[kk] - ephi;
            const int16_t abs_phi_diff = phi_diff < 0 ? -phi_diff : phi_diff;
            if (abs_phi_diff > phi_tolerance) {
              stop = true;
              continue;
            }
            if (!w.used[kk]) {
              if constexpr (count) ++wc.seed_eval;
              const float dz = w.z[kk] - h0z;
              const float predx = h0x + tx * dz;
              const float predy = h0y + ty * dz;
              const float dx = predx - w.x[kk];
              const float dy = predy - w.y[kk];
              const float scatter = dx * dx + dy * dy;
              if (scatter < best) {
                best = scatter;
                best_h2 = w.index[kk];
              }
            }
          }
          if (stop) break;
        }
      }
      // key = scatter bits, k, h2; the whole wave takes part in the exchange.
      unsigned long long key = (static_cast<unsigned long long>(__float_as_uint(best)) << 32) | (k << 16) | best_h2;
      for (unsigned d = K / 2; d > 0; d >>= 1) {
        const unsigned long long other = __shfl_xor(key, d, wave_size);
        key = other < key ? other : key;
      }
      if (k == 0 && t < C && q < n_h1) {
        const bool found = __uint_as_float(static_cast<uint32_t>(key >> 32)) < max_scatter;
        const unsigned kb = static_cast<unsigned>(key >> 16) & 0xFFu;
        ss.seed[q] = found ? (static_cast<uint32_t>(ss.h0[q * max_h0_candidates + kb]) |
                              (static_cast<uint32_t>(key & 0xFFFFu) << 16))
                           : no_seed;
      }
    }
    __syncthreads();
    phase_mark<count>(wc, phase_seed2);

    // 3. The chunk's seeds in h1 order (lane t: h1 number t).
    uint32_t seed = no_seed;
    uint16_t h1_index = 0;
    if (threadIdx.x < n_h1) {
      seed = ss.seed[threadIdx.x];
      h1_index = ss.h1_chunk[threadIdx.x];
    }
    bool found = seed != no_seed;
    const uint16_t best_h0 = static_cast<uint16_t>(seed & 0xFFFFu);
    const uint16_t best_h2 = static_cast<uint16_t>(seed >> 16);
    if constexpr (seed_mode == seed_mode_log) {
      if (found) {
        float f[n_seed_features];
        seed_features(hits, best_h0, h1_index, best_h2, module_pair, f);
        const unsigned slot = atomicAdd(o.seed_log.counter, 1u);
        if (slot < o.seed_log.capacity) {
          uint32_t* rec = o.seed_log.records + 5 * (size_t) slot;
          rec[0] = o.event_number;
          rec[1] = module_pair;
          rec[2] = best_h0;
          rec[3] = h1_index;
          rec[4] = best_h2;
          float* fo = o.seed_log.features + n_seed_features * (size_t) slot;
#pragma unroll
          for (unsigned kk = 0; kk < n_seed_features; ++kk) fo[kk] = f[kk];
        }
      }
    }
    if constexpr (seed_mode == seed_mode_scored) {
      // Seeds are decided in groups of C consecutive h1 hits (a wave, or half of one when
      // C = 32). A seed is kept when its logit reaches the threshold; a group whose seeds would
      // all go keeps them (plain search as fallback).
      float fp[n_seed_features];
#pragma unroll
      for (unsigned kk = 0; kk < n_seed_features; ++kk) fp[kk] = 0.f;
      if (found) seed_features(hits, best_h0, h1_index, best_h2, module_pair, fp);
      const uint64_t live = __ballot(found);
      const float logit = live ? scorer_logit(fp, live, *o.scorer) : 0.f;
      const bool keep = found && logit >= o.scorer_threshold;
      if (found && o.seed_log.counter) {
        const unsigned slot = atomicAdd(o.seed_log.counter, 1u);
        if (slot < o.seed_log.capacity) {
          uint32_t* rec = o.seed_log.records + 5 * (size_t) slot;
          rec[0] = o.event_number;
          rec[1] = module_pair;
          rec[2] = best_h0;
          rec[3] = h1_index;
          rec[4] = best_h2;
          float* fo = o.seed_log.features + n_seed_features * (size_t) slot;
#pragma unroll
          for (unsigned kk = 0; kk < n_seed_features; ++kk) fo[kk] = fp[kk];
          if (o.seed_log.logits) o.seed_log.logits[slot] = logit;
        }
      }
      const uint64_t group = C >= wave_size ? ~0ull : (((1ull << C) - 1) << (lane_id() / C * C));
      if (__ballot(keep) & group) found = keep;
    }
    const Scan3 s = scan3(sh, cn, found, false, false);
    if (found) {
      const unsigned tn = (cn.ttf + s.r[0]) % max_tracks_to_follow;
      TrackletHits tl;
      tl.hits[0] = best_h0;
      tl.hits[1] = h1_index;
      tl.hits[2] = best_h2;
      o.tracklets[tn] = tl;
      o.ttf[tn] = bit_seed | tn;
      const unsigned wslot = tn & o.window_mask;
      o.ttf_window[wslot] = bit_seed | tn;
      o.ttf_last2[wslot] = static_cast<uint32_t>(h1_index) | (static_cast<uint32_t>(best_h2) << 16);
      o.ttf_first[wslot] = best_h0;
      o.ttf_n[wslot] = 3;
    }
    cn.ttf += s.t[0];
    phase_mark<count>(wc, phase_seed3);
  }
  cn.local_hits = base;
}

// SearchByTriplet.cu:451-609, ordered.
template<bool count, typename H>
__device__ __forceinline__ void track_forwarding(
  WorkCount& wc,
  const H& hits,
  SearchShared& sh,
  Counters& cn,
  const EventOutputs& o,
  const ModulePairData next,
  unsigned prev_ttf,
  unsigned diff_ttf,
  bool window,
  int16_t phi_tolerance,
  float max_scatter,
  unsigned max_skipped_modules)
{
  for (unsigned c = 0; c < diff_ttf; c += blockDim.x) {
    const unsigned e = c + threadIdx.x;
    bool promote = false, reinsert = false, three = false;
    uint32_t full = 0;
    unsigned number_of_hits = 0;
    int best_h2 = -1;
    uint16_t th[3] = {0, 0, 0};
    unsigned track_number = 0;
    bool valid = false;
    unsigned h0_num = 0, h1_num = 0;
    if (e < diff_ttf) {
      const unsigned q = (prev_ttf + e) % max_tracks_to_follow;
      if (window) {
        // The entry and its track's last hits from LDS (written with the entry).
        const unsigned wslot = q & o.window_mask;
        full = o.ttf_window[wslot];
        const uint32_t last2 = o.ttf_last2[wslot];
        h0_num = last2 & 0xFFFFu;
        h1_num = last2 >> 16;
        number_of_hits = o.ttf_n[wslot];
        th[0] = o.ttf_first[wslot];
        th[1] = static_cast<uint16_t>(h0_num);
        th[2] = static_cast<uint16_t>(h1_num);
        track_number = full & bit_track_number;
        if ((full & bit_seed) == bit_seed) valid = track_number < max_tracks_to_follow;
        // Guards that only matter once the slots of an event have overflowed (status raised).
        else valid = track_number < o.track_capacity && number_of_hits >= 3 && number_of_hits < max_track_size;
      }
      else {
        full = o.ttf[q];
        track_number = full & bit_track_number;
        const uint16_t* t_hits = nullptr;
        if ((full & bit_seed) == bit_seed) {
          valid = track_number < max_tracks_to_follow;
          if (valid) t_hits = o.tracklets[track_number].hits;
          number_of_hits = 3;
        }
        else if (track_number < o.track_capacity) {
          // Guards that only matter once the slots of an event have overflowed (status raised).
          t_hits = o.tracks[track_number].hits;
          number_of_hits = o.tracks[track_number].hitsNum;
          valid = number_of_hits >= 3 && number_of_hits < max_track_size;
        }
        if (valid) {
          h0_num = t_hits[number_of_hits - 2];
          h1_num = t_hits[number_of_hits - 1];
          th[0] = t_hits[0];
          th[1] = t_hits[1];
          th[2] = t_hits[2];
        }
      }
    }
    if (valid) {
      const unsigned skipped = (full & bit_skipped) >> skipped_pos;
      const float h0x = hits.x(h0_num), h0y = hits.y(h0_num), h0z = hits.z(h0_num);
      const float h1x = hits.x(h1_num), h1y = hits.y(h1_num), h1z = hits.z(h1_num);
      const float td = 1.0f / (h1z - h0z);
      const float txn = h1x - h0x;
      const float tyn = h1y - h0y;
      const float tx = txn * td;
      const float ty = tyn * td;
      float best_fit = max_scatter;
      int16_t cand, ephi;
      find_forward_candidate<count>(
        wc, hits, next, h0x, h0y, tx, ty, (hits.module_odd(h0_num) ? next.z[1] : next.z[0]) - h0z, phi_tolerance, cand, ephi);
      const unsigned n2 = next.hit_num;
      const unsigned c2 = static_cast<unsigned>(static_cast<int>(cand));
      for (unsigned j0 = 0; j0 < n2; j0 += 4) {
        Window4 w;
        w.load(hits, next.hit_start, n2, c2, j0, false);
        bool stop = false;
#pragma unroll
        for (unsigned k = 0; k < 4; ++k) {
          if (stop || j0 + k >= n2) continue;
          const int16_t phi_diff = w.phi[k] - ephi;
          const int16_t abs_phi_diff = phi_diff < 0 ? -phi_diff : phi_diff;
          if (abs_phi_diff > phi_tolerance) {
            stop = true;
            continue;
          }
          if constexpr (count) ++wc.fwd_eval;
          const float dz = w.z[k] - h0z;
          const float predx = h0x + tx * dz;
          const float predy = h0y + ty * dz;
          const float dx = predx - w.x[k];
          const float dy = predy - w.y[k];
          const float scatter = dx * dx + dy * dy;
          if (scatter < best_fit) {
            best_fit = scatter;
            best_h2 = w.index[k];
          }
        }
        if (stop) break;
      }
      if (best_h2 != -1) {
        promote = number_of_hits == 3;
        reinsert = number_of_hits + 1 < max_track_size;
      }
      else if (skipped < max_skipped_modules) {
        reinsert = true;
        track_number = ((skipped + 1) << skipped_pos) | (full & (bit_seed | bit_track_number));
      }
      else if (number_of_hits == 3) {
        three = true;
      }
    }
    const Scan3 s = scan3(sh, cn, promote, reinsert, three);
    if (valid) {
      if (best_h2 != -1) {
        if (promote) {
          hits.mark_used4(best_h2, th[0], th[1], th[2]);
          track_number = cn.n_tracks + s.r[0];
          if (track_number < o.track_capacity) {
            TrackHits& t = o.tracks[track_number];
            t.hits[0] = th[0];
            t.hits[1] = th[1];
            t.hits[2] = th[2];
            t.hits[3] = best_h2;
            t.hitsNum = 4;
          }
        }
        else {
          hits.mark_used(best_h2);
          TrackHits& t = o.tracks[track_number];
          t.hits[number_of_hits] = best_h2;
          t.hitsNum = number_of_hits + 1;
        }
      }
      if (reinsert) {
        const unsigned p = (cn.ttf + s.r[1]) % max_tracks_to_follow;
        // A promoted track beyond the slots of its event is skipped when read back.
        o.ttf[p] = track_number;
        const unsigned wslot = p & o.window_mask;
        o.ttf_window[wslot] = track_number;
        if (best_h2 != -1) {
          o.ttf_last2[wslot] = h1_num | (static_cast<uint32_t>(best_h2) << 16);
          o.ttf_n[wslot] = static_cast<uint8_t>(number_of_hits + 1);
        }
        else {
          o.ttf_last2[wslot] = h0_num | (h1_num << 16);
          o.ttf_n[wslot] = static_cast<uint8_t>(number_of_hits);
        }
        o.ttf_first[wslot] = th[0];
      }
      if (three) {
        const unsigned p = cn.three + s.r[2];
        if (p < o.track_capacity) {
          TrackletHits t;
          t.hits[0] = th[0];
          t.hits[1] = th[1];
          t.hits[2] = th[2];
          o.three_hit_tracks[p] = t;
        }
      }
    }
    // Chunks need no barrier between them: a chunk writes ring slots at and beyond the
    // counter, reads only the slots before it (distinct unless the ring overflows, which is
    // reported), and the used flags it sets are read by seeding after the caller's barrier.
    cn.n_tracks += s.t[0];
    cn.ttf += s.t[1];
    cn.three += s.t[2];
    if (cn.n_tracks > o.track_capacity) cn.status |= status_track_slots_overflow;
    if (cn.three > o.track_capacity) cn.status |= status_three_hit_slots_overflow;
  }
}

template<bool count, int seed_mode, typename H>
__device__ __forceinline__ void search_event(
  WorkCount& wc,
  const H& hits,
  const SeedScratch& ss,
  SearchShared& sh,
  Counters& cn,
  const EventOutputs& o,
  float phi_tolerance_f,
  float max_scatter,
  unsigned max_skipped_modules,
  uint8_t* tables)
{
  // Module pairs are visited from the last one upstream: seeding on (first, first - 1,
  // first - 2), forwarding into first - 2 (SearchByTriplet.cu:160-236).
  phase_mark<count>(wc, phase_other);
  const int16_t phi_tolerance = f32_to_i16(phi_tolerance_f * convert_factor_i16);
  unsigned first_module_pair = n_module_pairs - 1;
  unsigned last_ttf = 0;
  // Seeding with first = F searches the tables of F and F - 2, the forwarding after it that of
  // F - 3; seeding F builds the table of F - 4, so the first four are built at the event's start.
  static_assert(phi_tables_at_start == 4 && phi_buckets == wave_size, "table schedule");
  const auto seed = [&](unsigned h0_candidates) {
    track_seeding<count, seed_mode>(
      wc, hits, ss, sh, cn, o, sh.mp[first_module_pair], sh.mp[first_module_pair - 1], sh.mp[first_module_pair - 2],
      first_module_pair - 1, max_scatter, phi_tolerance, h0_candidates,
      first_module_pair >= phi_tables_at_start ? &sh.mp[first_module_pair - phi_tables_at_start] : nullptr,
      tables);
    --first_module_pair;
  };
  const auto forward = [&]() {
    const unsigned prev_ttf = last_ttf;
    last_ttf = cn.ttf;
    const unsigned diff_ttf = last_ttf - prev_ttf;
    if (diff_ttf >= max_tracks_to_follow) cn.status |= status_ttf_ring_overflow;
    // The positions read here, [prev_ttf, last_ttf), and every position written from their
    // writing to their reading (this forwarding reinserts at most diff_ttf entries) span at most
    // 2 diff_ttf positions: within the LDS window they are all distinct. Beyond it the ring is
    // read back from HBM, after a fence that makes this workgroup's ring writes visible to it.
    const bool window = 2 * diff_ttf <= o.window;
    if (!window) __threadfence_block();
    // Orders the ring slots and tracklets written by the last seeding chunk before they are
    // read (seeding chunks end right after their writes, without a barrier).
    __syncthreads();
    track_forwarding<count>(
      wc, hits, sh, cn, o, sh.mp[first_module_pair - 2], prev_ttf, diff_ttf, window, phi_tolerance, max_scatter,
      max_skipped_modules);
    __syncthreads();
    phase_mark<count>(wc, phase_forwarding);
  };
  if constexpr (seed_mode == seed_mode_scored) {
    // One inlined copy of track_seeding (the largest body, here with the scorer) keeps the
    // kernel's code small.
    for (;;) {
      seed(first_module_pair == n_module_pairs - 1 ? initial_seeding_h0_candidates : seeding_h0_candidates);
      if (first_module_pair <= 1) break;
      forward();
    }
  }
  else {
    // The first call peeled: both copies see a constant h0 window.
    seed(initial_seeding_h0_candidates);
    while (first_module_pair > 1) {
      forward();
      seed(seeding_h0_candidates);
    }
  }
  __syncthreads();
  const unsigned prev_ttf = last_ttf;
  last_ttf = cn.ttf;
  const unsigned diff_ttf = last_ttf - prev_ttf;
  if (diff_ttf >= max_tracks_to_follow) cn.status |= status_ttf_ring_overflow;
  const bool window = diff_ttf <= o.window; // nothing is written while these are read
  if (!window) {
    __threadfence_block();
    __syncthreads();
  }
  for (unsigned c = 0; c < diff_ttf; c += blockDim.x) {
    const unsigned e = c + threadIdx.x;
    bool seed = false;
    uint32_t full = 0;
    if (e < diff_ttf) {
      const unsigned q = (prev_ttf + e) % max_tracks_to_follow;
      full = window ? o.ttf_window[q & o.window_mask] : o.ttf[q];
      seed = (full & bit_seed) == bit_seed;
    }
    const Scan3 s = scan3(sh, cn, seed, false, false);
    if (seed) {
      const unsigned p = cn.three + s.r[0];
      if (p < o.track_capacity) o.three_hit_tracks[p] = o.tracklets[full & bit_track_number];
    }
    cn.three += s.t[0];
    if (cn.three > o.track_capacity) cn.status |= status_three_hit_slots_overflow;
  }
}

// One workgroup per event, or (persistent, launched with a work counter) a resident grid whose
// workgroups take the next event slot from the counter until the slots run out; the persistent
// grid is an instantiation of its own, so the one-event kernel keeps no loop state live. Dynamic LDS: the seeding scratch
// (seed_scratch_bytes_per_lane per lane), the LDS window of the ring of tracks to follow
// (ring_window entries; the ring itself, 2048 x u32, is written through to HBM), the list of
// unused hits of the previous module pair and, per position, the number of unused hits before
// it (list_cap x u16 each), the module pairs' phi bucket tables, then `lds_hits` hits as x[],
// y[], z[] (f32), phi[] (i16)
// and a flags byte (used, module parity). Events with up to 5 x lds_hits hits keep phi and the
// flags in LDS and read coordinates from global memory; larger events run on global memory.
// The `count` instantiation also accumulates the work counters into work_stats[4] (one
// atomic per wave). With event_clock, each workgroup stores at event_clock[10 * event_number]
// its event's start and end time (wall clock ticks) and, in the counting instantiation, the
// ticks spent in each Phase.
template<bool count, int seed_mode, bool persistent>
__global__ void __launch_bounds__(seed_mode == seed_mode_scored ? 256 : 1024, seed_mode == seed_mode_scored ? 3 : 1)
  k_search_by_triplet(
  const uint32_t* event_list,
  const uint32_t* number_of_events,
  const char* sorted_container,
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* module_cluster_num,
  const Geometry* geometry,
  TrackHits* tracks_all,
  TrackletHits* tracklets_all,
  uint32_t* ttf_all,
  TrackletHits* three_hit_all,
  uint8_t* hit_used_all,
  uint32_t* atomics_all,
  uint16_t* rel_indices_all,
  uint32_t* number_of_velo_tracks,
  uint32_t* status,
  float phi_tolerance,
  float max_scatter,
  unsigned max_skipped_modules,
  unsigned lds_hits,
  unsigned long long* work_stats,
  unsigned long long* event_clock,
  unsigned n_list,
  uint32_t* work_counter,
  SeedLog seed_log,
  const float* scorer_params,
  float scorer_threshold,
  unsigned window,
  unsigned list_cap)
{
  extern __shared__ __align__(16) char dyn_all[];
  __shared__ SearchShared sh;
  __shared__ unsigned s_slot;
  WorkCount wc;
  SeedScratch ss;
  ss.seed = reinterpret_cast<uint32_t*>(dyn_all);
  ss.h0 = reinterpret_cast<uint16_t*>(ss.seed + blockDim.x);
  ss.h1_chunk = ss.h0 + max_h0_candidates * blockDim.x;
  ss.n_h0 = reinterpret_cast<uint8_t*>(ss.h1_chunk + blockDim.x);
  uint32_t* ring = reinterpret_cast<uint32_t*>(dyn_all + ((seed_scratch_bytes_per_lane * blockDim.x + 15) & ~15u));
  uint32_t* ring_last2 = ring + window;
  uint16_t* ring_first = reinterpret_cast<uint16_t*>(ring_last2 + window);
  uint8_t* ring_n = reinterpret_cast<uint8_t*>(ring_first + window);
  ss.unused = reinterpret_cast<uint16_t*>(ring_n + window);
  ss.before = ss.unused + list_cap;
  ss.list_cap = list_cap;
  uint8_t* phi_tables = reinterpret_cast<uint8_t*>(ss.before + list_cap);
  char* dyn = reinterpret_cast<char*>(phi_tables + phi_tables_bytes);
  ScorerLds* scorer = nullptr;
  if constexpr (seed_mode == seed_mode_scored) {
    scorer = reinterpret_cast<ScorerLds*>(dyn + (((size_t) lds_hits * search_bytes_per_hit + 15) & ~size_t(15)));
    load_scorer(scorer_params, *scorer);
    __syncthreads();
  }
  // One event per workgroup (work_counter == NULL), or a resident grid whose workgroups take
  // event slots from the counter until they run out (the counter starts at 0).
  // One event: slot `slot` of the event list (the whole workgroup, uniform).
  const auto process_event = [&](unsigned slot) {
    const unsigned long long t_start = event_clock ? wall_clock64() : 0ull;
    wc.last = t_start;
    const unsigned event_number = event_list ? event_list[slot] : slot;
    const unsigned n_events = number_of_events[0];
    const unsigned total = offsets_estimated_input_size[n_module_pairs * n_events];
    const uint32_t* mhs = offsets_estimated_input_size + event_number * n_module_pairs;
    const uint32_t* mhn = module_cluster_num + event_number * n_module_pairs;
    const unsigned hit_offset = mhs[0];
    const unsigned n_hits = mhs[n_module_pairs] - hit_offset;
    const Clusters cc {const_cast<char*>(sorted_container), total};

    const unsigned toff = track_offset(offsets_estimated_input_size, event_number);
    const unsigned container = track_container_size(total, n_events);
    const unsigned next_off =
      event_number + 1 < n_events ? track_offset(offsets_estimated_input_size, event_number + 1) : container;
    EventOutputs o;
    o.tracks = tracks_all + toff;
    o.three_hit_tracks = three_hit_all + toff;
    o.tracklets = tracklets_all + (size_t) event_number * max_tracks_to_follow;
    o.ttf = ttf_all + (size_t) event_number * max_tracks_to_follow;
    o.ttf_window = ring;
    o.ttf_last2 = ring_last2;
    o.ttf_first = ring_first;
    o.ttf_n = ring_n;
    o.h1_indices = rel_indices_all + hit_offset;
    o.event_number = event_number;
    o.seed_log = seed_log;
    o.scorer = scorer;
    o.scorer_threshold = scorer_threshold;
    o.window = window;
    o.window_mask = window - 1;
    // Slots of this event; the reference lets a busy event run into its neighbour's slots, here
    // nothing is written past the container and the overflow is reported.
    o.track_capacity = min(next_off > toff ? next_off - toff : 0u, container - min(toff, container));

    Counters cn;
    cn.status = n_hits > 65536u ? status_hit_index_overflow : 0u;
    if (threadIdx.x < n_module_pairs) {
      const unsigned m = threadIdx.x;
      sh.mp[m].hit_start = mhs[m] - hit_offset;
      sh.mp[m].hit_num = mhn[m];
      sh.mp[m].z[0] = geometry->module_zs[2 * m];
      sh.mp[m].z[1] = geometry->module_zs[2 * m + 1];
      sh.mp[m].table = mhn[m] <= 255u ? m * phi_table_stride : no_phi_table;
    }
    const uint32_t* gids = cc.ids() + hit_offset;
    const int16_t* gph = cc.phis() + hit_offset;
    const float* gxyz = cc.xyz() + 3 * (size_t) hit_offset;
    uint8_t* gused = hit_used_all + hit_offset;
    if (n_hits <= lds_hits) {
      float* xs = reinterpret_cast<float*>(dyn);
      float* ys = xs + lds_hits;
      float* zs = ys + lds_hits;
      int16_t* ph = reinterpret_cast<int16_t*>(zs + lds_hits);
      uint8_t* flags = reinterpret_cast<uint8_t*>(ph + lds_hits);
      // Two hits per lane and round, every load issued before the first store waits (the
      // scheduler would otherwise wait out one memory round trip per field).
      for (unsigned i0 = threadIdx.x; i0 < n_hits; i0 += 2 * blockDim.x) {
        const unsigned i1 = i0 + blockDim.x < n_hits ? i0 + blockDim.x : i0;
        const float x0 = gxyz[3 * i0], y0 = gxyz[3 * i0 + 1], z0 = gxyz[3 * i0 + 2];
        const float x1 = gxyz[3 * i1], y1 = gxyz[3 * i1 + 1], z1 = gxyz[3 * i1 + 2];
        const int16_t p0 = gph[i0], p1 = gph[i1];
        const uint32_t id0 = gids[i0], id1 = gids[i1];
        __builtin_amdgcn_sched_barrier(0);
        xs[i0] = x0;
        ys[i0] = y0;
        zs[i0] = z0;
        ph[i0] = p0;
        flags[i0] = module_parity(id0) ? flag_odd : 0;
        xs[i1] = x1;
        ys[i1] = y1;
        zs[i1] = z1;
        ph[i1] = p1;
        flags[i1] = module_parity(id1) ? flag_odd : 0;
      }
      __syncthreads();
      const LdsHits hits {xs, ys, zs, ph, flags, phi_tables};
      build_phi_tables(hits, sh.mp, phi_tables, n_module_pairs - phi_tables_at_start);
      __syncthreads();
      search_event<count, seed_mode>(wc, hits, ss, sh, cn, o, phi_tolerance, max_scatter, max_skipped_modules, phi_tables);
      __syncthreads();
      for (unsigned i = threadIdx.x; i < n_hits; i += blockDim.x) gused[i] = flags[i] & flag_used;
    }
    else if (3 * n_hits <= search_bytes_per_hit * lds_hits) {
      int16_t* ph = reinterpret_cast<int16_t*>(dyn);
      uint8_t* flags = reinterpret_cast<uint8_t*>(ph + n_hits);
      for (unsigned i0 = threadIdx.x; i0 < n_hits; i0 += 2 * blockDim.x) {
        const unsigned i1 = i0 + blockDim.x < n_hits ? i0 + blockDim.x : i0;
        const int16_t p0 = gph[i0], p1 = gph[i1];
        const uint32_t id0 = gids[i0], id1 = gids[i1];
        __builtin_amdgcn_sched_barrier(0);
        ph[i0] = p0;
        flags[i0] = module_parity(id0) ? flag_odd : 0;
        ph[i1] = p1;
        flags[i1] = module_parity(id1) ? flag_odd : 0;
      }
      cn.status |= status_lds_fallback;
      __syncthreads();
      const MixedHits hits {gxyz, ph, flags, phi_tables};
      build_phi_tables(hits, sh.mp, phi_tables, n_module_pairs - phi_tables_at_start);
      __syncthreads();
      search_event<count, seed_mode>(wc, hits, ss, sh, cn, o, phi_tolerance, max_scatter, max_skipped_modules, phi_tables);
      __syncthreads();
      for (unsigned i = threadIdx.x; i < n_hits; i += blockDim.x) gused[i] = flags[i] & flag_used;
    }
    else {
      for (unsigned i = threadIdx.x; i < n_hits; i += blockDim.x) gused[i] = 0;
      cn.status |= status_lds_fallback;
      __threadfence_block();
      __syncthreads();
      const GlobalHits hits {gxyz, gph, gused, gids, phi_tables};
      build_phi_tables(hits, sh.mp, phi_tables, n_module_pairs - phi_tables_at_start);
      __syncthreads();
      search_event<count, seed_mode>(wc, hits, ss, sh, cn, o, phi_tolerance, max_scatter, max_skipped_modules, phi_tables);
      __syncthreads();
    }
    if constexpr (count) {
      // Wave reduction, then one atomic per wave.
      unsigned long long v[4] = {wc.seed_eval, wc.fwd_eval, wc.extrap, wc.steps};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (unsigned d = 32; d > 0; d >>= 1) v[k] += __shfl_xor(v[k], d, wave_size);
        if (lane_id() == 0 && v[k]) atomicAdd(work_stats + k, v[k]);
      }
      wc.seed_eval = wc.fwd_eval = wc.extrap = wc.steps = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t* at = atomics_all + event_number * num_atomics;
      at[atomic_three_hit] = cn.three;
      at[atomic_ttf] = cn.ttf;
      at[atomic_local_hits] = cn.local_hits;
      number_of_velo_tracks[event_number] = min(cn.n_tracks, o.track_capacity);
      if (cn.status && status) atomicOr(status, cn.status);
      if (event_clock) {
        unsigned long long* ec = event_clock + event_clock_words * event_number;
        const unsigned long long t_end = wall_clock64();
        ec[0] = t_start;
        ec[1] = t_end;
        if constexpr (count) {
          phase_mark<count>(wc, phase_other);
#pragma unroll
          for (int k = 0; k < n_phases; ++k) {
            ec[2 + k] = wc.phase[k];
            wc.phase[k] = 0;
          }
        }
      }
    }
    __syncthreads(); // the next event reuses the LDS
  };
  if constexpr (persistent) {
    for (;;) {
      if (threadIdx.x == 0) s_slot = atomicAdd(work_counter, 1u);
      __syncthreads();
      const unsigned slot = s_slot;
      __syncthreads();
      if (slot >= n_list) break;
      process_event(slot);
    }
  }
  else {
    if (blockIdx.x < n_list) process_event(blockIdx.x);
  }
}

// Events in descending order of their hit count (buckets of 32 hits, order within a bucket
// unspecified), for launching search_by_triplet largest events first: with every event in
// flight at once, the largest event sets the duration of the launch. Results do not depend on
// the order. One workgroup of 1024 threads.
constexpr unsigned order_buckets = 1024;
__global__ void __launch_bounds__(1024) k_order_events_by_hits(
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* number_of_events,
  uint32_t* order)
{
  __shared__ uint32_t bucket[order_buckets];
  __shared__ uint32_t scan_scratch[1024 / wave_size + 1];
  const unsigned n = number_of_events[0];
  for (unsigned b = threadIdx.x; b < order_buckets; b += blockDim.x) bucket[b] = 0;
  __syncthreads();
  auto key = [&](unsigned e) {
    const unsigned hits =
      offsets_estimated_input_size[(e + 1) * n_module_pairs] - offsets_estimated_input_size[e * n_module_pairs];
    return order_buckets - 1 - min(hits >> 5, order_buckets - 1); // descending
  };
  for (unsigned e = threadIdx.x; e < n; e += blockDim.x) atomicAdd(&bucket[key(e)], 1u);
  __syncthreads();
  for (unsigned b0 = 0, carry = 0; b0 < order_buckets; b0 += blockDim.x) {
    const unsigned b = b0 + threadIdx.x;
    const uint32_t v = b < order_buckets ? bucket[b] : 0u;
    uint32_t tot = 0;
    const uint32_t ex = block_exclusive_scan(v, scan_scratch, &tot);
    if (b < order_buckets) bucket[b] = carry + ex;
    carry += tot;
    __syncthreads();
  }
  for (unsigned e = threadIdx.x; e < n; e += blockDim.x) order[atomicAdd(&bucket[key(e)], 1u)] = e;
}

hipError_t launch_order_events_by_hits(
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* number_of_events,
  uint32_t* order,
  hipStream_t s)
{
  hipLaunchKernelGGL(
    k_order_events_by_hits, dim3(1), dim3(1024), 0, s, offsets_estimated_input_size, number_of_events, order);
  return hipGetLastError();
}

// ThreeHitTracksFilter.cu:48-107
__device__ __forceinline__ float means_square_fit_chi2(const Clusters& cc, unsigned hit_offset, const TrackletHits& t)
{
  float s0, sx, sz, sxz, sz2, u0, uy, uz, uyz, uz2;
  s0 = sx = sz = sxz = sz2 = 0.0f;
  u0 = uy = uz = uyz = uz2 = 0.0f;
  float xs[3], ys[3], zs[3];
#pragma unroll
  for (unsigned h = 0; h < 3; ++h) {
    const size_t i = (size_t) hit_offset + t.hits[h];
    xs[h] = cc.xyz()[3 * i];
    ys[h] = cc.xyz()[3 * i + 1];
    zs[h] = cc.xyz()[3 * i + 2];
  }
#pragma unroll
  for (unsigned h = 0; h < 3; ++h) {
    const float x = xs[h], y = ys[h], z = zs[h];
    const float wx = param_w;
    const float wx_t_x = wx * x;
    const float wx_t_z = wx * z;
    s0 += wx;
    sx += wx_t_x;
    sz += wx_t_z;
    sxz += wx_t_x * z;
    sz2 += wx_t_z * z;
    const float wy = param_w;
    const float wy_t_y = wy * y;
    const float wy_t_z = wy * z;
    u0 += wy;
    uy += wy_t_y;
    uz += wy_t_z;
    uyz += wy_t_y * z;
    uz2 += wy_t_z * z;
  }
  const float dens = 1.0f / (sz2 * s0 - sz * sz);
  const float state_tx = (sxz * s0 - sx * sz) * dens;
  const float state_x = (sx * sz2 - sxz * sz) * dens;
  const float denu = 1.0f / (uz2 * u0 - uz * uz);
  const float state_ty = (uyz * u0 - uy * uz) * denu;
  const float state_y = (uy * uz2 - uyz * uz) * denu;
  float chi2 = 0.0f;
#pragma unroll
  for (unsigned h = 0; h < 3; ++h) {
    const float z = zs[h];
    const float x = state_x + state_tx * z;
    const float y = state_y + state_ty * z;
    const float dx = x - xs[h];
    const float dy = y - ys[h];
    chi2 += dx * dx * param_w + dy * dy * param_w;
  }
  return chi2 * 0.5f;
}

// ThreeHitTracksFilter.cu:109-169, ordered output.
__global__ void k_three_hit_filter(
  const uint32_t* event_list,
  const uint32_t* number_of_events,
  const char* sorted_container,
  const uint32_t* offsets_estimated_input_size,
  const TrackletHits* input_all,
  const uint32_t* atomics_all,
  const uint8_t* hit_used_all,
  TrackletHits* output_all,
  uint32_t* number_of_output,
  float max_chi2)
{
  __shared__ uint32_t scan[3 * 16 + 4];
  const unsigned event_number = event_list ? event_list[blockIdx.x] : blockIdx.x;
  const unsigned n_events = number_of_events[0];
  const unsigned total = offsets_estimated_input_size[n_module_pairs * n_events];
  const unsigned hit_offset = offsets_estimated_input_size[event_number * n_module_pairs];
  const uint8_t* used = hit_used_all + hit_offset;
  const Clusters cc {const_cast<char*>(sorted_container), total};
  const unsigned toff = track_offset(offsets_estimated_input_size, event_number);
  const TrackletHits* in = input_all + toff;
  const unsigned container = track_container_size(total, n_events);
  const unsigned next_off =
    event_number + 1 < n_events ? track_offset(offsets_estimated_input_size, event_number + 1) : container;
  const unsigned capacity = min(next_off > toff ? next_off - toff : 0u, container - min(toff, container));
  // The count can only exceed the event's slots once they overflowed (a status bit is set then).
  const unsigned n_in = min(atomics_all[event_number * num_atomics + atomic_three_hit], capacity);
  TrackletHits* out = output_all + toff;
  unsigned base = 0;
  for (unsigned c = 0; c < n_in; c += blockDim.x) {
    const unsigned i = c + threadIdx.x;
    bool keep = false;
    TrackletHits t;
    if (i < n_in) {
      t = in[i];
      const bool any_used = used[t.hits[0]] || used[t.hits[1]] || used[t.hits[2]];
      const float chi2 = means_square_fit_chi2(cc, hit_offset, t);
      keep = !any_used && chi2 < max_chi2;
    }
    const Scan3 s = block_flag_scan3(keep, false, false, scan);
    if (keep) out[base + s.r[0]] = t;
    base += s.t[0];
  }
  if (threadIdx.x == 0) number_of_output[event_number] = base;
}

// Launch geometry: the seeding scratch, the ring window, the unused-hit list, then the hits.
size_t search_dynamic_lds(unsigned block, unsigned lds_hits, unsigned window, unsigned list_cap)
{
  return ((seed_scratch_bytes_per_lane * (size_t) block + 15) & ~size_t(15)) + ring_window_bytes * (size_t) window +
         2 * 2 * (size_t) list_cap + phi_tables_bytes + (size_t) lds_hits * search_bytes_per_hit;
}

hipError_t launch_search_by_triplet(const SearchLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
  const unsigned block = a.block_dim ? a.block_dim : 256;
  // Dynamic LDS stays within the device's per-workgroup limit (less the static part).
  const int max_lds = device_max_lds_per_block();
  if (!max_lds) return hipErrorInvalidDevice;
  // Scored mode: seeds are decided in groups of block / K unused h1 hits of each chunk of block
  // positions (the pruning rule with seed_block = block), and a group's seeds must lie in one
  // wave for the in-wave scorer and decision: workgroups of 64, 128 or 256 lanes. The scorer's
  // weights take LDS from the hits (same residency).
  const bool scored = a.scorer_params != nullptr;
  if (scored && block != 64 && block != 128 && block != 256) return hipErrorInvalidConfiguration;
  // The LDS window of the ring: a power of two from 64 to 512 (it divides the ring); the list
  // of unused hits: a multiple of 16 up to 1024 (relative positions fit 16 bits either way).
  const unsigned window = a.ring_window ? a.ring_window : ring_window;
  const unsigned list_cap = a.list_cap ? a.list_cap : unused_list_cap;
  if (window < 64 || window > ring_window || (window & (window - 1)) || list_cap % 16 || list_cap > unused_list_cap)
    return hipErrorInvalidValue;
  const size_t extra = scored ? scorer_lds_bytes + 16 : 0;
  const size_t fixed = search_dynamic_lds(block, 0, window, list_cap) + extra + 1024;
  if ((size_t) max_lds <= fixed) return hipErrorInvalidConfiguration;
  const unsigned max_hits = static_cast<unsigned>(((size_t) max_lds - fixed) / search_bytes_per_hit);
  const unsigned default_hits =
    scored ? search_default_lds_hits - static_cast<unsigned>((extra + search_bytes_per_hit - 1) / search_bytes_per_hit)
           : search_default_lds_hits;
  const unsigned lds_hits = min(a.lds_hits ? a.lds_hits : default_hits, max_hits);
  const size_t dyn = search_dynamic_lds(block, lds_hits, window, list_cap) + extra;
  // Persistent mode: as many workgroups as stay resident (they take events from the counter).
  unsigned grid = a.n_list;
  if (a.work_counter) {
    const void* kernel = a.work_stats ? reinterpret_cast<const void*>(&k_search_by_triplet<true, seed_mode_plain, true>)
                         : a.scorer_params
                           ? reinterpret_cast<const void*>(&k_search_by_triplet<false, seed_mode_scored, true>)
                         : a.seed_log.counter
                           ? reinterpret_cast<const void*>(&k_search_by_triplet<false, seed_mode_log, true>)
                           : reinterpret_cast<const void*>(&k_search_by_triplet<false, seed_mode_plain, true>);
    const unsigned resident = device_resident_blocks(kernel, block, dyn);
    if (resident) grid = min(grid, resident);
  }
#define VELO_LAUNCH_SEARCH_P(C, L, P)                                                                               \
  hipLaunchKernelGGL(                                                                                               \
    (k_search_by_triplet<C, L, P>), dim3(grid), dim3(block), dyn, s, a.event_list, a.number_of_events,              \
    a.sorted_container, a.offsets_estimated_input_size, a.module_cluster_num,                                       \
    reinterpret_cast<const Geometry*>(a.geometry), reinterpret_cast<TrackHits*>(a.tracks),                          \
    reinterpret_cast<TrackletHits*>(a.tracklets), a.tracks_to_follow,                                               \
    reinterpret_cast<TrackletHits*>(a.three_hit_tracks), a.hit_used, a.atomics_velo, a.rel_indices,                \
    a.number_of_velo_tracks, a.status, a.phi_tolerance, a.max_scatter, a.max_skipped_modules, lds_hits,             \
    a.work_stats, a.event_clock, a.n_list, a.work_counter, a.seed_log, a.scorer_params, a.scorer_threshold, window,   \
    list_cap)
#define VELO_LAUNCH_SEARCH(C, L)                                                                                    \
  do {                                                                                                              \
    if (a.work_counter) VELO_LAUNCH_SEARCH_P(C, L, true);                                                           \
    else VELO_LAUNCH_SEARCH_P(C, L, false);                                                                         \
  } while (0)
  if (a.work_stats) VELO_LAUNCH_SEARCH(true, seed_mode_plain);
  else if (a.scorer_params) VELO_LAUNCH_SEARCH(false, seed_mode_scored);
  else if (a.seed_log.counter) VELO_LAUNCH_SEARCH(false, seed_mode_log);
  else VELO_LAUNCH_SEARCH(false, seed_mode_plain);
#undef VELO_LAUNCH_SEARCH
#undef VELO_LAUNCH_SEARCH_P
  return hipGetLastError();
}

// The first n_big events of the list (the largest, when the list is ordered by descending hit
// count) in workgroups of big_block lanes with big_lds_hits hits of LDS on `side`, the others
// with the launch's own knobs on `s`. `side` waits for the work on `s` before it (fork) and `s`
// waits for `side` after both launches (join), so later work on `s` sees every event's outputs.
hipError_t launch_search_split(
  const SearchLaunch& a,
  unsigned n_big,
  unsigned big_block,
  unsigned big_lds_hits,
  hipStream_t s,
  hipStream_t side,
  hipEvent_t fork,
  hipEvent_t join)
{
  if (!a.event_list || !side || !fork || !join || n_big == 0 || n_big >= a.n_list || a.work_counter)
    return launch_search_by_triplet(a, s);
  hipError_t e = hipEventRecord(fork, s);
  if (e != hipSuccess) return e;
  e = hipStreamWaitEvent(side, fork, 0);
  if (e != hipSuccess) return e;
  SearchLaunch big = a;
  big.n_list = n_big;
  big.block_dim = big_block;
  big.lds_hits = big_lds_hits;
  // Launched first, so its large workgroups are placed before the many small ones.
  e = launch_search_by_triplet(big, side);
  SearchLaunch rest = a;
  rest.event_list = a.event_list + n_big;
  rest.n_list = a.n_list - n_big;
  const hipError_t e2 = launch_search_by_triplet(rest, s);
  // The join is recorded whatever happened above, so `side` never stays outside the run.
  const hipError_t e3 = hipEventRecord(join, side);
  const hipError_t e4 = e3 == hipSuccess ? hipStreamWaitEvent(s, join, 0) : e3;
  return e != hipSuccess ? e : (e2 != hipSuccess ? e2 : e4);
}

hipError_t launch_three_hit_tracks_filter(const ThreeHitFilterLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
  hipLaunchKernelGGL(
    k_three_hit_filter, dim3(a.n_list), dim3(256), 0, s, a.event_list, a.number_of_events, a.sorted_container,
    a.offsets_estimated_input_size, reinterpret_cast<const TrackletHits*>(a.three_hit_tracks_input), a.atomics_velo,
    a.hit_used, reinterpret_cast<TrackletHits*>(a.three_hit_tracks_output), a.number_of_three_hit_tracks_output,
    a.max_chi2);
  return hipGetLastError();
}

} // namespace velo
