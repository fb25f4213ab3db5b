// Track consolidation, the simplified VELO Kalman fit, the device prefix sum and fills.
//
//   k_copy_track_hit_number   VeloCopyTrackHitNumber.cu:49-78
//   k_consolidate_tracks      VeloConsolidateTracks.cu:120-192 (one thread per track hit)
//   k_kalman_filter           VeloKalmanFilter.cu:335-418, VeloKalmanFilter.cuh:25-158 (one thread per track)
//   k_create_velo_views       VeloConsolidateTracks.cu:11-46 (Allen's view objects, byte for byte)
//   k_clusters_views          MaskedVeloClustering.cu:519-522, DecodeRetinaClusters.cu:331-346
//   k_scan_*                  replaces host_prefix_sum_t (HostPrefixSum.cpp:76-107) by a device scan
#include "velo_amd.h"
#include "velo_common.cuh"
#include "velo_kernels.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

namespace velo {

// ---------------------------------------------------------------------------------------
// Prefix sum: out[0] = 0, out[i+1] = in[0] + ... + in[i] (u32 wrap), in three passes:
// per-tile sums, a scan of the tile sums, then per-tile scans seeded with their offset.
// Sizes up to one_pass_items run as a single workgroup (a launch instead of three).
constexpr unsigned scan_block = 1024;
constexpr unsigned scan_items = 4;
constexpr unsigned scan_tile = scan_block * scan_items;
constexpr unsigned one_pass_items = 32;
constexpr unsigned one_pass_max = scan_block * one_pass_items;

__device__ __forceinline__ uint32_t block_scan_1024(uint32_t v, uint32_t* scratch, uint32_t* total)
{
  return block_exclusive_scan(v, scratch, total);
}

__device__ __forceinline__ unsigned scan_length(const uint32_t* n_dev, unsigned n_cap)
{
  return n_dev ? min(n_dev[0], n_cap) : n_cap;
}

__global__ void __launch_bounds__(scan_block)
  k_scan_tile_sums(const uint32_t* in, const uint32_t* n_dev, unsigned n_cap, uint32_t* tile_sums)
{
  __shared__ uint32_t scratch[scan_block / wave_size + 1];
  const unsigned n = scan_length(n_dev, n_cap);
  const size_t base = (size_t) blockIdx.x * scan_tile;
  uint32_t s = 0;
#pragma unroll
  for (unsigned k = 0; k < scan_items; ++k) {
    const size_t i = base + (size_t) k * scan_block + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t total;
  block_scan_1024(s, scratch, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Exclusive scan of `n` values in place by one workgroup (n_tiles is small).
__global__ void __launch_bounds__(scan_block) k_scan_small_inplace(uint32_t* v, unsigned n)
{
  __shared__ uint32_t scratch[scan_block / wave_size + 1];
  uint32_t carry = 0;
  for (unsigned c = 0; c < n; c += scan_block) {
    const unsigned i = c + threadIdx.x;
    const uint32_t x = i < n ? v[i] : 0u;
    uint32_t total;
    const uint32_t e = block_scan_1024(x, scratch, &total);
    if (i < n) v[i] = carry + e;
    carry += total;
  }
}

// Per-tile scan: thread t owns items [t*items, (t+1)*items) of its tile (contiguous, so the
// prefix within a thread is a short serial sum).
template<unsigned items>
__global__ void __launch_bounds__(scan_block)
  k_scan_tiles(const uint32_t* in, const uint32_t* n_dev, unsigned n_cap, uint32_t* out, const uint32_t* tile_offsets)
{
  __shared__ uint32_t scratch[scan_block / wave_size + 1];
  const unsigned n = scan_length(n_dev, n_cap);
  const size_t base = (size_t) blockIdx.x * scan_block * items + (size_t) threadIdx.x * items;
  uint32_t v[items];
  uint32_t s = 0;
#pragma unroll
  for (unsigned k = 0; k < items; ++k) {
    const size_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t total;
  uint32_t run = block_scan_1024(s, scratch, &total) + (tile_offsets ? tile_offsets[blockIdx.x] : 0u);
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
#pragma unroll
  for (unsigned k = 0; k < items; ++k) {
    const size_t i = base + k;
    run += v[k];
    if (i < n) out[i + 1] = run;
  }
}

size_t prefix_sum_scratch_words(unsigned n) { return (n + scan_tile - 1) / scan_tile; }

hipError_t launch_prefix_sum_dyn(
  const uint32_t* in,
  const uint32_t* n_dev,
  unsigned n_cap,
  uint32_t* out,
  uint32_t* scratch,
  hipStream_t s)
{
  const unsigned tiles = (n_cap + scan_tile - 1) / scan_tile;
  if (tiles <= 1) {
    hipLaunchKernelGGL(k_scan_tiles<scan_items>, dim3(1), dim3(scan_block), 0, s, in, n_dev, n_cap, out, nullptr);
    return hipGetLastError();
  }
  if (n_cap <= one_pass_max) {
    hipLaunchKernelGGL(k_scan_tiles<one_pass_items>, dim3(1), dim3(scan_block), 0, s, in, n_dev, n_cap, out, nullptr);
    return hipGetLastError();
  }
  if (!scratch) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(tiles), dim3(scan_block), 0, s, in, n_dev, n_cap, scratch);
  hipLaunchKernelGGL(k_scan_small_inplace, dim3(1), dim3(scan_block), 0, s, scratch, tiles);
  hipLaunchKernelGGL(k_scan_tiles<scan_items>, dim3(tiles), dim3(scan_block), 0, s, in, n_dev, n_cap, out, scratch);
  return hipGetLastError();
}

hipError_t launch_prefix_sum(const uint32_t* in, unsigned n, uint32_t* out, uint32_t* scratch, hipStream_t s)
{
  return launch_prefix_sum_dyn(in, nullptr, n, out, scratch, s);
}

// Tile sums of the allocation-free prefix sum: one u32 per scan tile, room for
// prefix_sum_static_max_items items (256 KiB per device, held until the process ends). Every
// launch writes all the tile sums it reads, so nothing needs clearing between calls; calls that
// overlap on one device share the array and race.
constexpr unsigned prefix_sum_static_max_tiles = prefix_sum_static_max_items / scan_tile;
static_assert(prefix_sum_static_max_items % scan_tile == 0, "whole tiles");
static __device__ __align__(256) uint32_t scratch[prefix_sum_static_max_tiles];

namespace {
std::mutex g_scratch_mutex;
std::map<int, uint32_t*> g_scratch_ptr; // device -> address of its copy of `scratch`
} // namespace

// The device address of the current device's `scratch`, looked up once per device (the array's
// name is a host-side stand-in and is never passed to a kernel).
static hipError_t static_scratch(uint32_t** ptr)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) {
    *ptr = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch_ptr.emplace(dev, static_cast<uint32_t*>(p));
  *ptr = static_cast<uint32_t*>(p);
  return hipSuccess;
}

hipError_t prefix_sum_static_init()
{
  uint32_t* p = nullptr;
  return static_scratch(&p);
}

hipError_t launch_prefix_sum_static(const uint32_t* in, unsigned n, uint32_t* out, hipStream_t s)
{
  if (n > prefix_sum_static_max_items) return hipErrorInvalidValue;
  uint32_t* p = nullptr;
  const hipError_t e = static_scratch(&p);
  if (e != hipSuccess) return e;
  return launch_prefix_sum_dyn(in, nullptr, n, out, p, s);
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t value)
{
  for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
    p[i] = value;
}
__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t value)
{
  for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x)
    p[i] = value;
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t value, hipStream_t s)
{
  if (!n) return hipSuccess;
  const unsigned grid = (unsigned) std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, s, p, n, value);
  return hipGetLastError();
}
hipError_t launch_fill_u8(uint8_t* p, size_t n, uint8_t value, hipStream_t s)
{
  if (!n) return hipSuccess;
  const unsigned grid = (unsigned) std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u8, dim3(grid), dim3(256), 0, s, p, n, value);
  return hipGetLastError();
}

namespace {
std::mutex g_query_mutex;
std::map<std::tuple<int, const void*, unsigned, size_t>, unsigned> g_resident; // (device, kernel, block, lds)
std::map<int, int> g_max_lds;                                                   // device -> bytes
} // namespace

int device_max_lds_per_block()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(g_query_mutex);
  auto it = g_max_lds.find(dev);
  if (it != g_max_lds.end()) return it->second;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
  g_max_lds.emplace(dev, v);
  return v;
}

unsigned device_resident_blocks(const void* kernel, unsigned block, size_t dyn_lds)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  const auto key = std::make_tuple(dev, kernel, block, dyn_lds);
  std::lock_guard<std::mutex> lock(g_query_mutex);
  auto it = g_resident.find(key);
  if (it != g_resident.end()) return it->second;
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, static_cast<int>(block), dyn_lds) != hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return 0;
  const unsigned v = static_cast<unsigned>(cus) * static_cast<unsigned>(per_cu);
  g_resident.emplace(key, v);
  return v;
}

// ---------------------------------------------------------------------------------------
// K7 (grid = number of events).
__global__ void k_copy_track_hit_number(
  const TrackHits* tracks_all,
  const uint32_t* offsets_velo_tracks,
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* offsets_three_hit,
  uint32_t* velo_track_hit_number,
  uint32_t* offsets_all_velo_tracks)
{
  const unsigned e = blockIdx.x;
  const TrackHits* tracks = tracks_all + track_offset(offsets_estimated_input_size, e);
  const unsigned n_tracks = offsets_velo_tracks[e + 1] - offsets_velo_tracks[e];
  const unsigned n_three = offsets_three_hit[e + 1] - offsets_three_hit[e];
  uint32_t* out = velo_track_hit_number + offsets_velo_tracks[e] + offsets_three_hit[e];
  for (unsigned i = threadIdx.x; i < n_tracks; i += blockDim.x) out[i] = tracks[i].hitsNum;
  for (unsigned i = threadIdx.x; i < n_three; i += blockDim.x) out[n_tracks + i] = 3;
  if (threadIdx.x == 0) {
    offsets_all_velo_tracks[e + 1] = offsets_velo_tracks[e + 1] + offsets_three_hit[e + 1];
    if (e == 0) offsets_all_velo_tracks[0] = 0;
  }
}

// K8: one thread per (track, hit); the event's tracks are [main tracks][filtered 3-hit tracks].
constexpr unsigned consolidate_lds_tracks = 2047; // tracks of an event whose offsets are staged in LDS
constexpr unsigned consolidate_lds_hits = 4096;    // track hits of an event mapped to their tracks in LDS

__global__ void k_consolidate_tracks(
  const uint32_t* event_list,
  const uint32_t* number_of_events,
  const uint32_t* offsets_all_velo_tracks,
  const TrackHits* tracks_all,
  const uint32_t* offsets_velo_track_hit_number,
  const char* sorted_container,
  const uint32_t* offsets_estimated_input_size,
  const TrackletHits* three_hit_output_all,
  const uint32_t* offsets_three_hit,
  char* velo_track_hits,
  uint8_t* accepted,
  bool lds_map)
{
  const unsigned e = event_list ? event_list[blockIdx.x] : blockIdx.x;
  const unsigned n_events = number_of_events[0];
  const unsigned total_clusters = offsets_estimated_input_size[n_module_pairs * n_events];
  const unsigned total_tracks = offsets_all_velo_tracks[n_events];
  const unsigned total_hits = offsets_velo_track_hit_number[total_tracks];
  const Clusters in {const_cast<char*>(sorted_container), total_clusters};
  const Clusters out {velo_track_hits, total_hits};
  const unsigned hit_offset = offsets_estimated_input_size[e * n_module_pairs];
  const unsigned toff = track_offset(offsets_estimated_input_size, e);
  const unsigned t0 = offsets_all_velo_tracks[e];
  const unsigned n_total = offsets_all_velo_tracks[e + 1] - t0;
  const unsigned n_main = n_total - (offsets_three_hit[e + 1] - offsets_three_hit[e]);
  const unsigned h_begin = offsets_velo_track_hit_number[t0];
  const unsigned h_end = offsets_velo_track_hit_number[t0 + n_total];
  for (unsigned i = threadIdx.x; i < n_total; i += blockDim.x)
    if (accepted) accepted[t0 + i] = 1;
  // The event's slice of the hit-number offsets, in LDS when it fits. With it, every track
  // writes (track, position in the track) for each of its hits into an LDS map (one thread per
  // track), so a hit reads its track in one LDS access; otherwise each hit finds its track by a
  // binary search over the offsets.
  __shared__ uint32_t s_off[consolidate_lds_tracks + 1];
  __shared__ uint16_t s_map[consolidate_lds_hits];
  const bool in_lds = n_total <= consolidate_lds_tracks;
  const bool mapped = lds_map && in_lds && h_end - h_begin <= consolidate_lds_hits;
  if (in_lds)
    for (unsigned i = threadIdx.x; i <= n_total; i += blockDim.x) s_off[i] = offsets_velo_track_hit_number[t0 + i];
  __syncthreads();
  if (mapped) {
    static_assert(consolidate_lds_tracks < (1u << 11) && max_track_size <= 32, "(track, position) in 16 bits");
    for (unsigned t = threadIdx.x; t < n_total; t += blockDim.x) {
      const unsigned a = s_off[t] - h_begin, b = s_off[t + 1] - h_begin;
      for (unsigned j = a; j < b; ++j) s_map[j] = static_cast<uint16_t>((t << 5) | (j - a));
    }
    __syncthreads();
  }
  const uint32_t* off = in_lds ? s_off : offsets_velo_track_hit_number + t0;
  for (unsigned h = h_begin + threadIdx.x; h < h_end; h += blockDim.x) {
    unsigned lo, k;
    if (mapped) {
      const unsigned m = s_map[h - h_begin];
      lo = m >> 5;
      k = m & 31u;
    }
    else {
      lo = 0;
      unsigned hi = n_total;
      while (hi - lo > 1) {
        const unsigned mid = (lo + hi) / 2;
        if (off[mid] <= h) lo = mid;
        else hi = mid;
      }
      k = h - off[lo];
    }
    const unsigned hit = lo < n_main ? tracks_all[toff + lo].hits[k] : three_hit_output_all[toff + lo - n_main].hits[k];
    const size_t src = (size_t) hit_offset + hit;
    // All four fields are read before any is written (the output may not alias the input, but
    // the compiler cannot know that and would otherwise wait out each read in turn).
    const uint32_t id = in.ids()[src];
    const float x = in.xyz()[3 * src], y = in.xyz()[3 * src + 1], z = in.xyz()[3 * src + 2];
    out.ids()[h] = id;
    out.xyz()[3 * (size_t) h] = x;
    out.xyz()[3 * (size_t) h + 1] = y;
    out.xyz()[3 * (size_t) h + 2] = z;
  }
}

// ---------------------------------------------------------------------------------------
// K9
struct KState {
  float x, y, z, tx, ty;
  float c00, c20, c22, c11, c31, c33;
};

// FP64 covariance fallback (no reference counterpart): the FP32 update (1 - gain_p) * pp_pred
// loses about kappa * 2^-24 of relative accuracy, kappa = 1 + w * pp_pred (about 1e7 on a
// track's first step). Above kappa = 100 the step's covariance prediction and update run in
// FP64 from the stored FP32 covariance, the update as (1 - gain_p) = inv / w, and the three
// terms are rounded to FP32; the state update keeps the reference's FP32 arithmetic.
constexpr float kalman_kappa_limit_minus_one = 99.f;

// One coordinate's filter step to the hit at z_hit with measurement m and weight w: position
// p, slope t and their covariance (pp, pt, tt) are predicted over dz and updated with the gain
// (the reference's expression order, VeloKalmanFilter.cuh:25-60).
template<bool fp64_fallback>
__device__ __forceinline__ void
filter_step(float z, float z_hit, float m, float w, float& p, float& t, float& pp, float& pt, float& tt)
{
  const float dz = z_hit - z;
  const float p_pred = p + dz * t;
  const float dz_tt = dz * tt;
  const float pt_pred = pt + dz_tt;
  const float dz_pt = dz * pt;
  const float pp_pred = pp + 2 * dz_pt + dz * dz_tt;
  const float tt_pred = tt;
  const float inv = 1.0f / ((1.0f / w) + pp_pred);
  const float gain_p = pp_pred * inv;
  const float gain_t = pt_pred * inv;
  const float res = m - p_pred;
  p = p_pred + gain_p * res;
  t = t + gain_t * res;
  if (fp64_fallback && pp_pred * w > kalman_kappa_limit_minus_one) {
    const double dzd = dz, c00 = pp, c20 = pt, c22 = tt;
    const double pt_d = c20 + dzd * c22;
    const double pp_d = c00 + 2.0 * (dzd * c20) + dzd * (dzd * c22);
    const double tt_d = c22;
    const double iw = 1.0 / static_cast<double>(w);
    const double inv_d = 1.0 / (iw + pp_d);
    pp = static_cast<float>(pp_d * iw * inv_d);
    pt = static_cast<float>(pt_d * iw * inv_d);
    tt = static_cast<float>(tt_d - pt_d * pt_d * inv_d);
    return;
  }
  pp = (1 - gain_p) * pp_pred;
  pt = (1 - gain_p) * pt_pred;
  tt = tt_pred - gain_t * pt_pred;
}

// The coordinates of a track's first two and last two hits (0, 1, nh - 2, nh - 1; nh >= 2),
// read together before the fits.
struct TrackEnds {
  float f[3], f1[3], l1[3], l[3];
};

template<bool upstream, bool fp64_fallback>
__device__ __forceinline__ KState
simplified_fit(const float* xyz, unsigned nh, const TrackEnds& te, float at_tx, float at_ty, const float* beamline, bool backward)
{
  const int direction = (backward ? 1 : -1) * (upstream ? 1 : -1);
  const float layer_noise = 1e-8f + 7e-6f * (at_tx * at_tx + at_ty * at_ty);
  int first = 0;
  int last = static_cast<int>(nh) - 1;
  int step = 1;
  const bool reverse = (te.l[2] - te.f[2]) * direction < 0;
  if (reverse) {
    const int t = first;
    first = last;
    last = t;
    step = -1;
  }
  KState s;
  s.x = reverse ? te.l[0] : te.f[0];
  s.y = reverse ? te.l[1] : te.f[1];
  s.z = reverse ? te.l[2] : te.f[2];
  s.tx = at_tx;
  s.ty = at_ty;
  s.c00 = param_w_inverted;
  s.c11 = param_w_inverted;
  s.c20 = 0.f;
  s.c31 = 0.f;
  s.c22 = 1.f;
  s.c33 = 1.f;
  // The next hit's coordinates are read while the current step computes (each step waits for
  // nothing but its own arithmetic).
  float nx = reverse ? te.l1[0] : te.f1[0], ny = reverse ? te.l1[1] : te.f1[1], nz = reverse ? te.l1[2] : te.f1[2];
  for (int i = first + step; i != last + step; i += step) {
    const float hx = nx, hy = ny, hz = nz;
    if (i + step != last + step) {
      nx = xyz[3 * (i + step)];
      ny = xyz[3 * (i + step) + 1];
      nz = xyz[3 * (i + step) + 2];
    }
    s.c22 += layer_noise;
    s.c33 += layer_noise;
    filter_step<fp64_fallback>(s.z, hz, hx, param_w, s.x, s.tx, s.c00, s.c20, s.c22);
    filter_step<fp64_fallback>(s.z, hz, hy, param_w, s.y, s.ty, s.c11, s.c31, s.c33);
    s.z = hz;
  }
  s.c22 += layer_noise;
  s.c33 += layer_noise;
  float delta_z = 0.f;
  if constexpr (upstream) {
    delta_z = (s.tx * (beamline[0] - s.x) + s.ty * (beamline[1] - s.y)) / (s.tx * s.tx + s.ty * s.ty);
  }
  else {
    delta_z = z_end_velo - s.z;
  }
  s.x = s.x + s.tx * delta_z;
  s.y = s.y + s.ty * delta_z;
  s.z = s.z + delta_z;
  const float dz2 = delta_z * delta_z;
  s.c00 += dz2 * s.c22 + 2.f * delta_z * s.c20;
  s.c11 += dz2 * s.c33 + 2.f * delta_z * s.c31;
  s.c20 += s.c22 * delta_z;
  s.c31 += s.c33 * delta_z;
  return s;
}

__device__ __forceinline__ void store_state(float* base, unsigned total_tracks, unsigned i, const KState& s)
{
  float* st = base + 6 * (size_t) i;
  st[0] = s.x;
  st[1] = s.y;
  st[2] = s.z;
  st[3] = s.tx;
  st[4] = s.ty;
  float* cv = base + 6 * (size_t) total_tracks + 8 * (size_t) i;
  cv[0] = s.c00;
  cv[1] = s.c20;
  cv[2] = s.c22;
  cv[3] = s.c11;
  cv[4] = s.c31;
  cv[5] = s.c33;
}

// One track: linear_fit, the upstream fit to the beamline and the downstream fit to the end of
// the VELO (VeloKalmanFilter.cu:384-399); xyz holds its nh hits' x, y, z.
__device__ __forceinline__ void fit_track(
  const float* xyz,
  unsigned nh,
  unsigned t,
  unsigned total_tracks,
  const float* bl,
  uint8_t* is_backward,
  float* beamline_states,
  float* endvelo_states,
  unsigned fp64_fallback)
{
  TrackEnds te;
#pragma unroll
  for (unsigned c = 0; c < 3; ++c) {
    te.f[c] = xyz[c];
    te.f1[c] = xyz[3 + c];
    te.l1[c] = xyz[3 * (nh - 2) + c];
    te.l[c] = xyz[3 * (nh - 1) + c];
  }
  __builtin_amdgcn_sched_barrier(0); // the four hits' loads are in flight together
  // linear_fit (VeloKalmanFilter.cu:335-355)
  const float fx = te.f[0], fy = te.f[1], fz = te.f[2];
  const float lx = te.l[0], ly = te.l[1], lz = te.l[2];
  const float ltx = (lx - fx) / (lz - fz);
  const float lty = (ly - fy) / (lz - fz);
  const float ldz = (ltx * (bl[0] - lx) + lty * (bl[1] - ly)) / (ltx * ltx + lty * lty);
  const float lin_z = lz + ldz;
  const bool backward = lin_z > fz;
  is_backward[t] = backward;
  if (fp64_fallback) {
    const KState b = simplified_fit<true, true>(xyz, nh, te, ltx, lty, bl, backward);
    const KState v = simplified_fit<false, true>(xyz, nh, te, b.tx, b.ty, bl, backward);
    store_state(beamline_states, total_tracks, t, b);
    store_state(endvelo_states, total_tracks, t, v);
  }
  else {
    const KState b = simplified_fit<true, false>(xyz, nh, te, ltx, lty, bl, backward);
    const KState v = simplified_fit<false, false>(xyz, nh, te, b.tx, b.ty, bl, backward);
    store_state(beamline_states, total_tracks, t, b);
    store_state(endvelo_states, total_tracks, t, v);
  }
}

// KalmanStates view of event e (VeloKalmanFilter.cu:370-374).
__device__ __forceinline__ void write_states_view(
  velo_kalman_states_view* views, const float* states, const uint32_t* offsets_all_velo_tracks, unsigned e,
  unsigned n_events)
{
  velo_kalman_states_view v;
  v.base_pointer = reinterpret_cast<const char*>(states);
  v.offset = offsets_all_velo_tracks[e];
  v.size = offsets_all_velo_tracks[e + 1] - offsets_all_velo_tracks[e];
  v.total_number_of_tracks = offsets_all_velo_tracks[n_events];
  v.pad_ = 0;
  views[e] = v;
}

// One workgroup per listed event, one thread per track. views: the tracks and their hits come
// through the event's Tracks view (Track -> Hits -> the consolidated container), as the
// reference reads them; otherwise from the hit-number offsets and the container.
template<bool views>
__global__ void k_kalman_filter(
  const uint32_t* event_list,
  const uint32_t* number_of_events,
  const uint32_t* offsets_all_velo_tracks,
  const uint32_t* offsets_velo_track_hit_number,
  const char* velo_track_hits,
  const velo_tracks_view* tracks_view,
  const float* beamline,
  uint8_t* is_backward,
  float* beamline_states,
  float* endvelo_states,
  velo_kalman_states_view* beamline_states_view,
  velo_kalman_states_view* endvelo_states_view,
  unsigned fp64_fallback)
{
  const unsigned e = event_list ? event_list[blockIdx.x] : blockIdx.x;
  const unsigned n_events = number_of_events[0];
  const unsigned total_tracks = offsets_all_velo_tracks[n_events];
  const float bl[2] = {beamline[0], beamline[1]};
  if (threadIdx.x == 0) {
    if (beamline_states_view)
      write_states_view(beamline_states_view, beamline_states, offsets_all_velo_tracks, e, n_events);
    if (endvelo_states_view)
      write_states_view(endvelo_states_view, endvelo_states, offsets_all_velo_tracks, e, n_events);
  }
  if constexpr (views) {
    const velo_tracks_view tv = tracks_view[e];
    for (unsigned i = threadIdx.x; i < tv.size; i += blockDim.x) {
      const velo_track_view tr = tv.track[i];
      const velo_hits_view h = *tr.hits;
      if (tr.number_of_hits < 2) continue;
      const float* xyz = reinterpret_cast<const float*>(h.base_pointer) + h.total_number_of_hits +
                         3 * (size_t) (h.offset + tr.offset);
      fit_track(xyz, tr.number_of_hits, tv.offset + i, total_tracks, bl, is_backward, beamline_states,
                endvelo_states, fp64_fallback);
    }
  }
  else {
    const unsigned total_hits = offsets_velo_track_hit_number[total_tracks];
    const float* xyz_all = reinterpret_cast<const float*>(velo_track_hits) + total_hits;
    for (unsigned t = offsets_all_velo_tracks[e] + threadIdx.x; t < offsets_all_velo_tracks[e + 1];
         t += blockDim.x) {
      const unsigned h0 = offsets_velo_track_hit_number[t];
      const unsigned nh = offsets_velo_track_hit_number[t + 1] - h0;
      if (nh < 2) continue; // every VELO track has >= 3 hits
      fit_track(xyz_all + 3 * (size_t) h0, nh, t, total_tracks, bl, is_backward, beamline_states, endvelo_states,
                fp64_fallback);
    }
  }
}

// create_velo_views (VeloConsolidateTracks.cu:11-46): event e's Hits and Tracks views and one
// Track view per track; the first workgroup also writes the multi-event view and its address.
__global__ void k_create_velo_views(
  const uint32_t* number_of_events,
  const uint32_t* offsets_all_velo_tracks,
  const uint32_t* offsets_velo_track_hit_number,
  const char* velo_track_hits,
  velo_hits_view* hits_view,
  velo_track_view* track_view,
  velo_tracks_view* tracks_view,
  velo_multi_event_tracks_view* multi_view,
  const velo_multi_event_tracks_view** imec)
{
  const unsigned n_events = number_of_events[0];
  const unsigned e = blockIdx.x;
  const unsigned first_track = offsets_all_velo_tracks[e];
  const unsigned n_tracks = offsets_all_velo_tracks[e + 1] - first_track;
  const uint32_t* hit_offsets = offsets_velo_track_hit_number + first_track;
  const unsigned first_hit = hit_offsets[0];
  for (unsigned i = threadIdx.x; i < n_tracks; i += blockDim.x) {
    velo_track_view t;
    t.hits = hits_view + e;
    t.track_index = i;
    t.track_container_offset = first_track;
    t.offset = hit_offsets[i] - first_hit;
    t.number_of_hits = hit_offsets[i + 1] - hit_offsets[i];
    track_view[first_track + i] = t;
  }
  if (threadIdx.x == 0) {
    velo_hits_view h;
    h.base_pointer = velo_track_hits;
    h.offset = first_hit;
    h.size = offsets_velo_track_hit_number[first_track + n_tracks] - first_hit;
    h.total_number_of_hits = offsets_velo_track_hit_number[offsets_all_velo_tracks[n_events]];
    h.pad_ = 0;
    hits_view[e] = h;
    velo_tracks_view ts;
    ts.track = track_view + first_track;
    ts.size = n_tracks;
    ts.offset = first_track;
    tracks_view[e] = ts;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    velo_multi_event_tracks_view m;
    m.type_id = VELO_TYPE_ID_VELO_TRACKS;
    m.pad_ = 0;
    m.container = tracks_view;
    m.number_of_events = n_events;
    m.pad2_ = 0;
    *multi_view = m;
    imec[0] = multi_view;
  }
}

// The Velo::Clusters view of each listed event (one thread per list entry); with zero_all every
// entry is cleared first by the same threads (the reference's memset of the retina decoding's
// views, DecodeRetinaClusters.cu:442; a single workgroup keeps the clear before the writes).
__global__ void k_clusters_views(
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  unsigned zero_all,
  const uint32_t* number_of_events,
  const uint32_t* totals,
  unsigned totals_stride,
  char* container,
  velo_clusters_view* views)
{
  if (zero_all) {
    for (unsigned i = threadIdx.x; i < n_events; i += blockDim.x) {
      velo_clusters_view z;
      z.base_pointer = nullptr;
      z.total_number_of_hits = 0;
      z.offset = 0;
      views[i] = z;
    }
    __syncthreads();
  }
  const unsigned total = totals[(size_t) totals_stride * number_of_events[0]];
  for (unsigned i = threadIdx.x; i < n_list; i += blockDim.x) {
    velo_clusters_view v;
    v.base_pointer = container;
    v.total_number_of_hits = total;
    v.offset = 0;
    views[event_list ? event_list[i] : i] = v;
  }
}

hipError_t launch_copy_track_hit_number(const CopyHitNumberLaunch& a, hipStream_t s)
{
  if (!a.n_events) return hipSuccess;
  hipLaunchKernelGGL(
    k_copy_track_hit_number, dim3(a.n_events), dim3(256), 0, s, reinterpret_cast<const TrackHits*>(a.tracks),
    a.offsets_velo_tracks, a.offsets_estimated_input_size, a.offsets_three_hit, a.velo_track_hit_number,
    a.offsets_all_velo_tracks);
  return hipGetLastError();
}

hipError_t launch_consolidate_tracks(const ConsolidateLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
  hipLaunchKernelGGL(
    k_consolidate_tracks, dim3(a.n_list), dim3(256), 0, s, a.event_list, a.number_of_events, a.offsets_all_velo_tracks,
    reinterpret_cast<const TrackHits*>(a.tracks), a.offsets_velo_track_hit_number, a.sorted_container,
    a.offsets_estimated_input_size, reinterpret_cast<const TrackletHits*>(a.three_hit_tracks_output),
    a.offsets_three_hit, a.velo_track_hits, a.accepted_velo_tracks, a.lds_map);
  return hipGetLastError();
}

hipError_t launch_kalman_filter(const KalmanLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
  auto* bsv = static_cast<velo_kalman_states_view*>(a.beamline_states_view);
  auto* esv = static_cast<velo_kalman_states_view*>(a.endvelo_states_view);
  const auto* tv = static_cast<const velo_tracks_view*>(a.tracks_view);
  if (tv)
    hipLaunchKernelGGL(
      k_kalman_filter<true>, dim3(a.n_list), dim3(256), 0, s, a.event_list, a.number_of_events,
      a.offsets_all_velo_tracks, a.offsets_velo_track_hit_number, a.velo_track_hits, tv, a.beamline, a.is_backward,
      reinterpret_cast<float*>(a.beamline_states), reinterpret_cast<float*>(a.endvelo_states), bsv, esv,
      a.fp64_fallback);
  else
    hipLaunchKernelGGL(
      k_kalman_filter<false>, dim3(a.n_list), dim3(256), 0, s, a.event_list, a.number_of_events,
      a.offsets_all_velo_tracks, a.offsets_velo_track_hit_number, a.velo_track_hits, tv, a.beamline, a.is_backward,
      reinterpret_cast<float*>(a.beamline_states), reinterpret_cast<float*>(a.endvelo_states), bsv, esv,
      a.fp64_fallback);
  return hipGetLastError();
}

hipError_t launch_create_velo_views(const VeloViewsLaunch& a, hipStream_t s)
{
  if (!a.n_events) return hipSuccess;
  hipLaunchKernelGGL(
    k_create_velo_views, dim3(a.n_events), dim3(256), 0, s, a.number_of_events, a.offsets_all_velo_tracks,
    a.offsets_velo_track_hit_number, a.velo_track_hits, static_cast<velo_hits_view*>(a.hits_view),
    static_cast<velo_track_view*>(a.track_view), static_cast<velo_tracks_view*>(a.tracks_view),
    static_cast<velo_multi_event_tracks_view*>(a.multi_event_tracks_view),
    static_cast<const velo_multi_event_tracks_view**>(a.imec_velo_tracks));
  return hipGetLastError();
}

hipError_t launch_clusters_views(const ClustersViewLaunch& a, hipStream_t s)
{
  if (!a.n_list && !a.zero_all) return hipSuccess;
  hipLaunchKernelGGL(
    k_clusters_views, dim3(1), dim3(1024), 0, s, a.event_list, a.n_list, a.n_events, unsigned(a.zero_all),
    a.number_of_events, a.totals, a.totals_stride, a.container, static_cast<velo_clusters_view*>(a.views));
  return hipGetLastError();
}

} // namespace velo
