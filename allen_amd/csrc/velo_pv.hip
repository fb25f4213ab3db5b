// Beamline primary-vertex finder for gfx950: the consumer of the VELO beamline states
// (device/PV/beamlinePV/src/pv_beamline_{extrapolate,histo,peak,calculate_denom,multi_fitter,
// cleanup}.cu). The reference runs six kernels with the z histogram in HBM and float atomics;
// here one workgroup per event runs them all with the histogram in LDS, and every floating-point
// sum keeps the CPU target's order, so the histogram and the seeds are bit-identical to it:
//
//   1. tracks: PVTrack (z, x, y, tx, ty, 1/c00, 1/c11) of each track
//   2. histogram: lane t owns the bins t, t+256, ...; a track's window (about 30 consecutive bins)
//      spreads over as many lanes, and each lane adds, track by track in track order, the
//      Gaussian-integral increments the reference adds with atomics (no atomics, identical sums)
//   3. peaks: the CPU target's protocluster scan (pv_beamline_peak.cu, the branch without warp
//      intrinsics) in three steps: the empty/non-empty transitions found by all lanes and ranked
//      by a block scan; the integral of each candidate protocluster summed in bin order by one
//      lane each; the extrema split of each kept protocluster by one lane each, its seeds ranked
//      by a block scan
//   4. denominators: one lane per track, seeds in order
//   5. multi-fitter: the seeds iterate in lockstep; per iteration all lanes select the (seed,
//      track) pairs within the chi2 cut, compute the weights and weighted terms of the selected
//      pairs into LDS ranked seed by seed in track order, one lane per (seed, sum) adds them in
//      that order (the CPU target's sums), and one lane per seed solves and updates it; vertices
//      kept in seed order by a ballot
//   6. cleanup: one lane per vertex, unique vertices kept in order by a ballot
//
// Capacities the reference does not check (extrema per protocluster, clusters, vertices) are
// checked: what exceeds them is dropped and status bit 16 is raised.
#include "velo_common.cuh"
#include "velo_kernels.h"

namespace velo {

namespace pv {
// BeamlinePVConstants.cuh
constexpr float zmin = -541.f, zmax = 307.f, dz = 0.25f;
constexpr int nbins = 3392;
constexpr float pp_maxTrackZ0Err = 1.5f, SMOG2_maxTrackZ0Err = 10.f, SMOG2_pp_separation = -341.f;
constexpr float maxTrackBLChi2 = 10.f;
constexpr int order_polynomial = 2;
constexpr float minDensity = 0.f, minDipDensity = 3.f, pp_minTracksInSeed = 2.5f, SMOG2_minTracksInSeed = 1.75f;
constexpr float maxVertexRho2 = 0.3f, maxFitIter = 7.f, maxChi2 = 12.f, minWeight = 0.f;
constexpr float chi2CutExp = 0.135335283f, maxDeltaZConverged = 0.0005f;
constexpr unsigned pp_minNumTracksPerVertex = 4, SMOG2_minNumTracksPerVertex = 3;
constexpr float minChi2Dist = 25.f;
// PV_Definitions.cuh
constexpr unsigned max_vertices = 32, max_clusteredges = 200, max_subclusters = 50;
constexpr unsigned max_protoclusters = max_clusteredges / 2;

constexpr unsigned block = 256, block_waves = block / wave_size;
constexpr unsigned bins_per_lane = (nbins + block - 1) / block; // 14 consecutive bins per lane in the transition scan
constexpr unsigned hist_words = block * bins_per_lane;          // 3584
constexpr unsigned pool_words = 8192;  // LDS shared by the phases (32 KiB)
constexpr unsigned track_chunk = 1024; // histogram windows staged per pass, three words each
constexpr unsigned fit_tracks = 512;   // multi-fitter: tracks resident in LDS per chunk, eight words each
constexpr unsigned fit_terms = 10;     // per selected (seed, track) pair: the ten weighted terms
constexpr unsigned n_sums = fit_terms + 1; // per seed: the ten sums and the number of selected tracks
constexpr uint32_t status_overflow = 1u << 16;
constexpr int dpp_wave_shr1 = 0x138; // DPP control: every lane reads the previous lane (GFX9 wave shift)
static_assert(hist_words + 3 * track_chunk <= pool_words, "histogram staging");
static_assert(8 * fit_tracks + max_vertices * (fit_tracks / 32) + 4 + fit_terms * block <= pool_words, "fit pool");
static_assert(hist_words + (nbins + 1) / 2 + max_clusteredges / 2 <= pool_words, "transition list");
static_assert(block >= max_protoclusters && wave_size >= max_vertices, "one lane per protocluster and per seed");
static_assert((block & (block - 1)) == 0, "bin ownership uses a mask");

struct Extremum {
  unsigned short index;
  float value, integral;
};

__device__ __forceinline__ float gauss_integral(float x)
{
  const float a = sqrtf(float(2 * order_polynomial + 3));
  const float xi = x / a;
  const float eta = 1.f - xi * xi;
  return 0.5f + xi * (0.5f + eta * (0.25f + eta * 0.1875f));
}

// pv_beamline_peak.cu: a bin is empty below dz / (10 maxTrackZ0Err); the scan starts "empty".
__device__ __forceinline__ bool bin_empty(const float* h, unsigned i)
{
  if (i == 0) return true;
  const float zBin = zmin + i * dz;
  const float Z0Err = zBin < SMOG2_pp_separation ? SMOG2_maxTrackZ0Err : pp_maxTrackZ0Err;
  const float inv_maxTrackZ0Err = 1.f / (10.f * Z0Err);
  const float threshold = dz * inv_maxTrackZ0Err;
  return h[i] < threshold;
}

// The seed position of a cluster from its highest bin and the two next to it.
__device__ __forceinline__ float seed_position(const float* h, unsigned short izmax)
{
  const float* b = h + izmax;
  const float d1 = *b - *(b - 1);
  const float d2 = *b - *(b + 1);
  const float idz = d1 + d2 > 0 ? 0.5f * (d1 - d2) / (d1 + d2) : 0.0f;
  return zmin + dz * (izmax + idz + 0.5f);
}
} // namespace pv

using namespace pv;

__global__ void __launch_bounds__(pv::block) k_pv_beamline(
  const uint32_t* event_list,
  const uint32_t* offsets_all_velo_tracks,
  const uint32_t* number_of_events,
  const float* states,
  const float* beamline,
  PVTrackOut* pvtracks,
  float* zhisto_out,
  float* zpeaks,
  uint32_t* number_of_zpeaks,
  float* pvtracks_denom,
  PVVertexOut* multi_fit_vertices,
  uint32_t* number_of_multi_fit_vertices,
  PVVertexOut* multi_final_vertices,
  uint32_t* number_of_multi_final_vertices,
  uint32_t* status)
{
  // pool: the histogram (phases 2-3) followed by the window staging (2) or the transitions and
  // protocluster edges (3); then the track chunks and their (seed, track) terms (5).
  __shared__ __align__(16) float pool[pool_words];
  __shared__ float s_zpeaks[max_vertices];
  __shared__ float4 s_seed[max_vertices]; // seeds not converged: vx, vy, vz, seed z
  __shared__ float s_sum[n_sums * max_vertices];
  __shared__ unsigned s_base[max_vertices], s_count[max_vertices], s_nact_sel;
  __shared__ PVVertexOut s_fit[max_vertices];
  __shared__ uint32_t s_scan[block_waves + 1];
  __shared__ unsigned s_nact, s_nfit;
  float* const h = pool;
  const unsigned t = threadIdx.x, l = lane_id();
  const unsigned event_number = event_list ? event_list[blockIdx.x] : blockIdx.x;
  const unsigned total = offsets_all_velo_tracks[*number_of_events];
  const unsigned t0 = offsets_all_velo_tracks[event_number];
  const unsigned nt = offsets_all_velo_tracks[event_number + 1] - t0;
  const float blx = beamline[0], bly = beamline[1];
  uint32_t st = 0;

  // 1. PVTrack of each track (pv_beamline_extrapolate.cu:31-41).
  for (unsigned i = t; i < nt; i += block) {
    const float* s = states + 6 * (size_t)(t0 + i);
    const float* c = states + 6 * (size_t) total + 8 * (size_t)(t0 + i);
    PVTrackOut p;
    p.z = s[2];
    p.x = s[0];
    p.y = s[1];
    p.tx = s[3];
    p.ty = s[4];
    p.W_00 = 1.f / c[0];
    p.W_11 = 1.f / c[3];
    pvtracks[t0 + i] = p;
  }
  for (unsigned b = t; b < hist_words; b += block) h[b] = 0.f;
  __syncthreads(); // the tracks are read back by other lanes below (global memory, same workgroup)

  // 2. Histogram (pv_beamline_histo.cu:28-86), chunks of tracks in order.
  {
    uint32_t* const s_window = reinterpret_cast<uint32_t*>(pool + hist_words); // minbin | maxbin << 16, 0: none
    float* const s_zerr = pool + hist_words + track_chunk;
    float* const s_z = pool + hist_words + 2 * track_chunk;
    for (unsigned c0 = 0; c0 < nt; c0 += track_chunk) {
      const unsigned cn = min(track_chunk, nt - c0);
      for (unsigned i = t; i < cn; i += block) {
        const PVTrackOut trk = pvtracks[t0 + c0 + i];
        uint32_t window = 0;
        float zerr = 0.f;
        if (zmin < trk.z && trk.z < zmax) {
          const float diffx2 = (trk.x - blx) * (trk.x - blx);
          const float diffy2 = (trk.y - bly) * (trk.y - bly);
          const float blchi2 = diffx2 * trk.W_00 + diffy2 * trk.W_11;
          if (blchi2 < maxTrackBLChi2) {
            const float zbin = (trk.z - zmin) / dz;
            const float zweight = trk.tx * trk.tx * trk.W_00 + trk.ty * trk.ty * trk.W_11;
            zerr = 1.f / sqrtf(zweight);
            const float maxTrackZ0Err = trk.z < SMOG2_pp_separation ? SMOG2_maxTrackZ0Err : pp_maxTrackZ0Err;
            if (zerr < maxTrackZ0Err) {
              const float a = sqrtf(float(2 * order_polynomial + 3));
              const float halfwindow = a * zerr / dz;
              const int minbin = max(int(zbin - halfwindow), 1);
              const int maxbin = min(int(zbin + halfwindow), nbins - 2);
              if (maxbin >= minbin) window = static_cast<uint32_t>(minbin) | (static_cast<uint32_t>(maxbin) << 16);
            }
          }
        }
        s_window[i] = window;
        s_zerr[i] = zerr;
        s_z[i] = trk.z;
      }
      __syncthreads();
      // The lanes of wave w own the bins whose residue mod 256 lies in [64w, 64w + 64). Each
      // wave looks at 64 tracks at a time, keeps those whose window reaches one of its bins (a
      // ballot), and walks them in track order with the window broadcast from the lane that
      // read it.
      const unsigned wave_first = t & ~(wave_size - 1) & (block - 1);
      for (unsigned g = 0; g < cn; g += wave_size) {
        uint32_t window = 0;
        float zerr = 0.f, z = 0.f;
        bool reaches = false;
        if (g + l < cn) {
          window = s_window[g + l];
          zerr = s_zerr[g + l];
          z = s_z[g + l];
          if (window) {
            const unsigned minbin = window & 0xFFFFu, len = (window >> 16) - minbin + 1;
            const unsigned s = (minbin - wave_first) & (block - 1);
            reaches = len >= block || s < wave_size || s + len > block;
          }
        }
        for (uint64_t mask = __ballot(reaches); mask; mask &= mask - 1) {
          const int k = __ffsll(static_cast<unsigned long long>(mask)) - 1;
          const uint32_t win = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(window), k));
          const float ze = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(zerr), k));
          const float zz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(z), k));
          const unsigned minbin = win & 0xFFFFu, maxbin = win >> 16;
          // The increment of bin b: the integral up to its upper edge (1 for the last bin) minus
          // the integral up to its lower edge (0 for the first bin).
          auto edge = [&](unsigned e) { return gauss_integral((zmin + (int) e * dz - zz) / ze); };
          unsigned b = minbin + ((t - minbin) & (block - 1)); // this lane's first bin in the window
          if (maxbin - minbin < block) {
            // At most one bin per lane; the lower edge of bin b is the upper edge of bin b - 1,
            // which the previous lane holds (same expression, same value).
            const bool in = b <= maxbin;
            const float hb = in ? h[b] : 0.f; // read ahead of the integrals
            const float upper = in ? (b < maxbin ? edge(b + 1) : 1.f) : 0.f;
            // The previous lane's value: a DPP shift of the whole wave by one lane.
            const float below =
              __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(upper), dpp_wave_shr1, 0xF, 0xF, false));
            if (in) {
              const float lower = b > minbin ? (l > 0 ? below : edge(b)) : 0.f;
              h[b] = hb + (upper - lower);
            }
          }
          else {
            for (; b <= maxbin; b += block) {
              const float upper = b < maxbin ? edge(b + 1) : 1.f;
              const float lower = b > minbin ? edge(b) : 0.f;
              h[b] += upper - lower;
            }
          }
        }
      }
      __syncthreads();
    }
  }
  if (zhisto_out)
    for (unsigned b = t; b < nbins; b += block) zhisto_out[(size_t) nbins * event_number + b] = h[b];

  // 3. Peaks (pv_beamline_peak.cu:23-236, the CPU target's branch).
  uint16_t* const trans = reinterpret_cast<uint16_t*>(pool + hist_words);
  uint16_t* const edges = reinterpret_cast<uint16_t*>(pool + hist_words + (nbins + 1) / 2);
  // 3a. The bins where the scan flips between empty and non-empty, in bin order. They alternate,
  // the first one opening a protocluster (the scan starts empty).
  uint32_t n_trans;
  {
    const unsigned lo = t * bins_per_lane, hi = min(lo + bins_per_lane, static_cast<unsigned>(nbins));
    uint32_t bits = 0;
    if (lo < hi) {
      bool prev = bin_empty(h, lo == 0 ? 0 : lo - 1);
      for (unsigned i = max(lo, 1u); i < hi; ++i) {
        const bool e = bin_empty(h, i);
        if (e != prev) bits |= 1u << (i - lo);
        prev = e;
      }
    }
    unsigned k = block_exclusive_scan(__popc(bits), s_scan, &n_trans);
    for (; bits; bits &= bits - 1u, ++k) trans[k] = static_cast<uint16_t>(lo + __ffs(bits) - 1);
  }
  __syncthreads();
  // 3b. Candidate m runs from transition 2m to transition 2m+1. The scan keeps it when the
  // histogram summed over it in bin order (the integral the scan carries) exceeds the seed
  // minimum at its end, and stores the edges of the first 100 kept; a protocluster opened after
  // those raises the overflow bit.
  unsigned n_kept = 0;
  for (unsigned m0 = 0; m0 < n_trans / 2; m0 += block) {
    const unsigned m = m0 + t;
    bool keep = false;
    unsigned ibegin = 0, iend = 0;
    if (m < n_trans / 2) {
      ibegin = trans[2 * m];
      iend = trans[2 * m + 1];
      float integral = h[ibegin];
      unsigned i = ibegin + 1;
      for (; i + 8 <= iend + 1; i += 8) { // eight reads in flight, added in order
        float v[8];
#pragma unroll
        for (unsigned u = 0; u < 8; ++u) v[u] = h[i + u];
#pragma unroll
        for (unsigned u = 0; u < 8; ++u) integral += v[u];
      }
      for (; i <= iend; ++i) integral += h[i];
      const float zBin = zmin + iend * dz;
      const float minInSeed = zBin < SMOG2_pp_separation ? SMOG2_minTracksInSeed : pp_minTracksInSeed;
      keep = integral > minInSeed;
    }
    uint32_t kept;
    const unsigned rank = n_kept + block_flag_scan(keep, s_scan, &kept);
    if (keep) {
      if (rank < max_protoclusters) {
        edges[2 * rank] = static_cast<uint16_t>(ibegin);
        edges[2 * rank + 1] = static_cast<uint16_t>(iend);
      }
      if (rank == max_protoclusters - 1 && 2 * (m + 1) < n_trans) st |= status_overflow;
    }
    n_kept += kept;
  }
  __syncthreads();
  const unsigned n_proto = min(n_kept, max_protoclusters);
  // 3c. Lane p splits protocluster p at its extrema; its seeds follow those of the protoclusters
  // before it.
  uint32_t n_clusters;
  {
    const float mindip = minDipDensity * dz;
    const float minpeak = minDensity * dz;
    Extremum extrema[max_vertices + 1];
    int N = 0;
    unsigned n_sub = 0, n_seeds = 0;
    auto subcluster_kept = [&](int i) {
      const float z_extrema = zmin + extrema[2 * i].index * dz;
      const float minInSeed = z_extrema < SMOG2_pp_separation ? SMOG2_minTracksInSeed : pp_minTracksInSeed;
      return extrema[2 * i].integral - extrema[2 * i - 2].integral > minInSeed;
    };
    if (t < n_proto) {
      // One entry more than the capacity, zeroed: the split below reads extrema[N] when N is
      // even (one past the last extremum, unset in the reference); here it reads zeros.
      for (unsigned i = 0; i <= max_vertices; ++i) extrema[i] = Extremum {0, 0.f, 0.f};
      const unsigned short ibegin = edges[2 * t];
      const unsigned short iend = edges[2 * t + 1];
      int number_of_extrema = 0;
      auto push = [&](unsigned short index, float value, float integ) {
        if (number_of_extrema < (int) max_vertices) extrema[number_of_extrema++] = Extremum {index, value, integ};
        else st |= status_overflow;
      };
      bool rising = true;
      float integ = h[ibegin];
      push(ibegin, h[ibegin], integ);
      auto step = [&](unsigned short i, float value, float next) {
        const bool stillrising = next > value;
        if (rising && !stillrising && value >= minpeak) {
          const int n = number_of_extrema;
          if (n >= 2) {
            const float dv1 = extrema[n - 2].value - extrema[n - 1].value;
            const float dv2 = value - extrema[n - 1].value;
            if (dv1 > mindip && dv2 > mindip) {
              push(i, value, integ + 0.5f * value);
            }
            else if (dv1 > dv2) {
              number_of_extrema--;
            }
            else {
              number_of_extrema -= 2;
              push(i, value, integ + 0.5f * value);
            }
          }
          else {
            push(i, value, integ + 0.5f * value);
          }
        }
        else if (rising != stillrising) {
          push(i, value, integ + 0.5f * value);
        }
        rising = stillrising;
        integ += value;
      };
      // Bins ibegin .. iend - 1 in order, the next eight values read ahead.
      float value = h[ibegin];
      for (unsigned i0 = ibegin; i0 < iend; i0 += 8) {
        float next[8];
#pragma unroll
        for (unsigned u = 0; u < 8; ++u) next[u] = h[min(i0 + 1 + u, static_cast<unsigned>(iend))];
#pragma unroll
        for (unsigned u = 0; u < 8; ++u)
          if (i0 + u < iend) {
            step(static_cast<unsigned short>(i0 + u), value, next[u]);
            value = next[u];
          }
      }
      push(iend, h[iend], integ);
      N = number_of_extrema;
      if (N > 3)
        for (int i = 1; i < (N / 2) + 1; ++i)
          if (subcluster_kept(i)) {
            if (n_sub < max_subclusters) ++n_sub;
            else st |= status_overflow;
          }
      n_seeds = n_sub > 0 ? n_sub : (extrema[1].value >= minpeak ? 1u : 0u);
    }
    unsigned k = block_exclusive_scan(n_seeds, s_scan, &n_clusters);
    float* zp = zpeaks + (size_t) max_vertices * event_number;
    auto add_cluster = [&](unsigned short izmax) {
      if (k < max_vertices) {
        const float z = seed_position(h, izmax);
        zp[k] = z;
        s_zpeaks[k] = z;
      }
      else {
        st |= status_overflow;
      }
      ++k;
    };
    if (n_sub > 0) {
      unsigned taken = 0;
      for (int i = 1; i < (N / 2) + 1 && taken < n_sub; ++i)
        if (subcluster_kept(i)) {
          ++taken;
          add_cluster(extrema[2 * i - 1].index);
        }
    }
    else if (n_seeds) {
      add_cluster(extrema[1].index);
    }
  }
  const unsigned n_peaks = min(n_clusters, max_vertices);
  if (t == 0) number_of_zpeaks[event_number] = n_peaks;
  __syncthreads();

  // 4. Denominators (pv_beamline_calculate_denom.cu:24-46).
  for (unsigned i = t; i < nt; i += block) {
    const PVTrackOut trk = pvtracks[t0 + i];
    float d = 0.f;
    for (unsigned j = 0; j < n_peaks; ++j) {
      const float dzs = s_zpeaks[j] - trk.z;
      const float rx = trk.x + trk.tx * dzs, ry = trk.y + trk.ty * dzs;
      const float chi2 = rx * rx * trk.W_00 + ry * ry * trk.W_11;
      d += expf(chi2 * (-0.5f));
    }
    pvtracks_denom[t0 + i] = d;
  }
  __syncthreads();

  // 5. Multi-fitter (pv_beamline_multi_fitter.cu:28-185). Lane s of wave 0 holds seed s; the
  // seeds not yet converged iterate together. Per iteration and chunk of tracks (all of them when
  // there are at most fit_tracks):
  //   a. all lanes test the (seed, track) pairs: a pair is selected when the track is in the z
  //      range and its chi2 to the vertex is below maxChi2 (a bit per pair in LDS);
  //   b. the selected pairs are ranked seed by seed in track order (popcounts and a wave scan);
  //   c. all lanes compute the weights and weighted terms of the selected pairs (two exponentials
  //      and a division each) into LDS at their rank, as many as the pool holds per round;
  //   d. lane 10 j + k adds term k of seed j over its pairs in rank order: the CPU target's sums,
  //      which skip the pairs not selected.
  bool converged = true, accept = false;
  float vtxcov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float vx = blx, vy = bly, vz = 0.f, seedz = 0.f, chi2tot = 0.f, sum_weights = 0.f;
  unsigned nselected = 0, minTracks = 0;
  if (t < n_peaks) {
    converged = false;
    accept = true;
    seedz = s_zpeaks[t];
    vz = seedz;
    minTracks = seedz <= SMOG2_pp_separation ? SMOG2_minNumTracksPerVertex : pp_minNumTracksPerVertex;
  }
  float* const tw = pool; // track i of the chunk: z, x, y, tx, ty, W_00, W_11, denominator at 8 i
  auto stage = [&](unsigned c0, unsigned cn) {
    for (unsigned i = t; i < cn; i += block) {
      const PVTrackOut p = pvtracks[t0 + c0 + i];
      *reinterpret_cast<float4*>(tw + 8 * i) = make_float4(p.z, p.x, p.y, p.tx);
      *reinterpret_cast<float4*>(tw + 8 * i + 4) = make_float4(p.ty, p.W_00, p.W_11, pvtracks_denom[t0 + c0 + i]);
    }
  };
  // The chi2 of track i to vertex j (maxChi2 outside the z range: not selected).
  auto pair_chi2 = [&](unsigned j, unsigned i, float& rx, float& ry) {
    const float4 sv = s_seed[j];
    const float4 ta = *reinterpret_cast<const float4*>(tw + 8 * i);
    const float4 tb = *reinterpret_cast<const float4*>(tw + 8 * i + 4);
    const float z = ta.x, x = ta.y, y = ta.z, tx = ta.w, W_00 = tb.y, W_11 = tb.z;
    const float dzv = sv.z - z;
    rx = sv.x - (x + tx * dzv);
    ry = sv.y - (y + tb.x * dzv);
    const float chi2 = rx * rx * W_00 + ry * ry * W_11;
    return (z < zmin || z >= zmax) ? maxChi2 : chi2;
  };
  // The weighted terms of a selected pair (zeros when its weight is not above minWeight).
  auto pair_terms = [&](unsigned j, unsigned i, float* out) {
    float rx, ry;
    const float chi2 = pair_chi2(j, i, rx, ry);
    const float4 sv = s_seed[j];
    const float4 ta = *reinterpret_cast<const float4*>(tw + 8 * i);
    const float4 tb = *reinterpret_cast<const float4*>(tw + 8 * i + 4);
    const float z = ta.x, x = ta.y, y = ta.z, tx = ta.w, ty = tb.x, W_00 = tb.y, W_11 = tb.z, denom = tb.w;
    const float H_20 = -tx, H_21 = -ty;
    const float HWH_20 = H_20 * W_00, HWH_21 = H_21 * W_11;
    const float HWH_22 = H_20 * H_20 * W_00 + H_21 * H_21 * W_11;
    float q[fit_terms] = {};
    const float dz_seed = sv.w - z;
    const float sx = blx - (x + tx * dz_seed), sy = bly - (y + ty * dz_seed);
    const float chi2_seed = sx * sx * W_00 + sy * sy * W_11;
    const float exp_chi2_0 = expf(chi2_seed * (-0.5f));
    const float nom = expf(chi2 * (-0.5f));
    const float den = chi2CutExp + nom;
    const float w = nom / (den + denom - exp_chi2_0);
    if (w > minWeight) {
      const float hx = rx * W_00, hy = ry * W_11, hz = -tx * rx * W_00 - ty * ry * W_11;
      q[0] = hx * w;
      q[1] = hy * w;
      q[2] = hz * w;
      q[3] = w * W_00;
      q[4] = w * W_11;
      q[5] = w * HWH_20;
      q[6] = w * HWH_21;
      q[7] = w * HWH_22;
      q[8] = w * chi2;
      q[9] = w;
    }
#pragma unroll
    for (unsigned k = 0; k < fit_terms; ++k) out[k] = q[k];
  };
  const bool resident = nt <= fit_tracks;
  if (resident) stage(0, nt); // read after the first barrier below
  for (unsigned iter = 0; iter < maxFitIter; ++iter) {
    unsigned my_j = 0;
    if (t < wave_size) {
      const uint64_t m = __ballot(!converged);
      my_j = __popcll(m & lanes_below_mask());
      if (!converged) s_seed[my_j] = make_float4(vx, vy, vz, seedz);
      if (l == 0) s_nact = __popcll(m);
    }
    __syncthreads();
    const unsigned S = s_nact, rows = fit_terms * S;
    if (S == 0) break;
    // Accumulator thread t < rows adds term k of seed j (t = 10 j + k); t + 256 as well for more
    // than 25 seeds.
    const unsigned aj0 = t / fit_terms, ak0 = t - aj0 * fit_terms;
    const unsigned aj1 = (t + block) / fit_terms, ak1 = t + block - aj1 * fit_terms;
    float acc0 = 0.f, acc1 = 0.f;
    unsigned n_selected = 0; // lane j < S of wave 0: pairs selected for seed j
    for (unsigned c0 = 0; c0 < nt; c0 += fit_tracks) {
      const unsigned cn = min(fit_tracks, nt - c0), mw = (cn + 31) / 32;
      uint32_t* const sel = reinterpret_cast<uint32_t*>(pool + 8 * cn); // seed j: words j mw .. j mw + mw
      const unsigned term_at = (8 * cn + S * mw + 3) & ~3u;
      float* const term = pool + term_at;
      const unsigned cap = (pool_words - term_at) / fit_terms; // selected pairs per round
      if (!resident) stage(c0, cn);
      for (unsigned u = t; u < S * mw; u += block) sel[u] = 0u;
      __syncthreads();
      // a. Selection bits.
      {
        const unsigned step_j = block / cn, step_i = block - step_j * cn;
        unsigned j = t / cn, i = t - j * cn;
        for (unsigned pair = t; pair < S * cn; pair += block) {
          float rx, ry;
          if (pair_chi2(j, i, rx, ry) < maxChi2) atomicOr(&sel[j * mw + (i >> 5)], 1u << (i & 31));
          i += step_i;
          j += step_j;
          if (i >= cn) {
            i -= cn;
            ++j;
          }
        }
      }
      __syncthreads();
      // b. Ranks: seed j's pairs take ranks s_base[j] .. s_base[j] + s_count[j], in track order.
      if (t < wave_size) {
        unsigned n = 0;
        if (l < S)
          for (unsigned w = 0; w < mw; ++w) n += __popc(sel[l * mw + w]);
        const unsigned inc = wave_inclusive_scan(n);
        if (l < S) {
          s_base[l] = inc - n;
          s_count[l] = n;
        }
        if (l == wave_size - 1) s_nact_sel = inc;
        n_selected += n;
      }
      __syncthreads();
      const unsigned total_selected = s_nact_sel;
      for (unsigned r0 = 0; r0 < total_selected; r0 += cap) {
        const unsigned rn = min(cap, total_selected - r0);
        // c. Terms of the pairs of ranks r0 .. r0 + rn.
        for (unsigned x = t; x < rn; x += block) {
          const unsigned rank = r0 + x;
          unsigned j = 0;
          while (j + 1 < S && s_base[j + 1] <= rank) ++j;
          unsigned r = rank - s_base[j], w = 0;
          const uint32_t* sj = sel + j * mw;
          uint32_t bits = sj[0];
          while (static_cast<unsigned>(__popc(bits)) <= r) {
            r -= __popc(bits);
            bits = sj[++w];
          }
          for (; r > 0; --r) bits &= bits - 1u;
          pair_terms(j, 32 * w + __ffs(bits) - 1, term + fit_terms * x);
        }
        __syncthreads();
        // d. Sums in track order, eight reads in flight.
        auto accumulate = [&](float& acc, unsigned aj, unsigned ak) {
          const unsigned lo = max(s_base[aj], r0), hi = min(s_base[aj] + s_count[aj], r0 + rn);
          const float* r = term + ak + fit_terms * (lo - r0);
          const unsigned n = hi > lo ? hi - lo : 0u;
          unsigned i = 0;
          for (; i + 8 <= n; i += 8) {
            float v[8];
#pragma unroll
            for (unsigned u = 0; u < 8; ++u) v[u] = r[fit_terms * (i + u)];
#pragma unroll
            for (unsigned u = 0; u < 8; ++u) acc = acc + v[u];
          }
          for (; i < n; ++i) acc = acc + r[fit_terms * i];
        };
        if (t < rows) accumulate(acc0, aj0, ak0);
        if (t + block < rows) accumulate(acc1, aj1, ak1);
        __syncthreads();
      }
    }
    if (t < rows) s_sum[n_sums * aj0 + ak0] = acc0;
    if (t + block < rows) s_sum[n_sums * aj1 + ak1] = acc1;
    if (t < S) s_sum[n_sums * t + fit_terms] = static_cast<float>(n_selected);
    __syncthreads();
    if (t < wave_size && !converged) {
      const float* sum = s_sum + n_sums * my_j;
      const float g0 = sum[0], g1 = sum[1], g2 = sum[2];
      const float a00 = sum[3], a11 = sum[4], a20 = sum[5], a21 = sum[6], a22 = sum[7];
      chi2tot = 0.f + sum[8];
      sum_weights = 0.f + sum[9];
      nselected = static_cast<unsigned>(sum[10]);
      if (nselected >= minTracks) {
        const float det = a00 * (a22 * a11 - a21 * a21) + a20 * (-a11 * a20);
        const float inv_det = 1.f / det;
        vtxcov[0] = (a22 * a11 - a21 * a21) * inv_det;
        vtxcov[1] = -(-a20 * a21) * inv_det;
        vtxcov[2] = (a22 * a00 - a20 * a20) * inv_det;
        vtxcov[3] = (-a20 * a11) * inv_det;
        vtxcov[4] = -(a21 * a00) * inv_det;
        vtxcov[5] = (a11 * a00) * inv_det;
        const float dx = -1.f * (vtxcov[0] * g0 + vtxcov[1] * g1 + vtxcov[3] * g2);
        const float dy = -1.f * (vtxcov[1] * g0 + vtxcov[2] * g1 + vtxcov[4] * g2);
        const float dzd = -1.f * (vtxcov[3] * g0 + vtxcov[4] * g1 + vtxcov[5] * g2);
        chi2tot += dx * g0 + dy * g1 + dzd * g2;
        vx = vx + dx;
        vy = vy + dy;
        vz = vz + dzd;
        converged = fabsf(dzd) < maxDeltaZConverged;
      }
      else {
        converged = true;
        accept = false;
      }
    }
  }
  if (t < wave_size) {
    bool ok = false;
    PVVertexOut vertex {};
    if (accept) {
      vertex.x = vx;
      vertex.y = vy;
      vertex.z = vz;
      vertex.chi2 = chi2tot;
      vertex.ndof = 0;
      vertex.cov00 = vtxcov[0];
      vertex.cov10 = vtxcov[1];
      vertex.cov11 = vtxcov[2];
      vertex.cov20 = vtxcov[3];
      vertex.cov21 = vtxcov[4];
      vertex.cov22 = vtxcov[5];
      vertex.nTracks = sum_weights;
      const float bdx = vx - blx, bdy = vy - bly;
      const float rho2 = bdx * bdx + bdy * bdy;
      const unsigned minT = vz <= SMOG2_pp_separation ? SMOG2_minNumTracksPerVertex : pp_minNumTracksPerVertex;
      ok = nselected >= minT && rho2 < maxVertexRho2;
    }
    const uint64_t mask = __ballot(ok);
    const unsigned slot = __popcll(mask & lanes_below_mask());
    if (ok) {
      multi_fit_vertices[(size_t) max_vertices * event_number + slot] = vertex;
      s_fit[slot] = vertex;
    }
    if (l == 0) {
      number_of_multi_fit_vertices[event_number] = __popcll(mask);
      s_nfit = __popcll(mask);
    }
  }
  __syncthreads();

  // 6. Cleanup (pv_beamline_cleanup.cu:95-150): lane i of wave 0 keeps vertex i if unique.
  if (t < wave_size) {
    const unsigned n_fit = s_nfit;
    bool unique = false;
    PVVertexOut v1 {};
    if (l < n_fit) {
      unique = true;
      v1 = s_fit[l];
      for (unsigned j = 0; j < n_fit; ++j) {
        if (j == l) continue;
        const PVVertexOut v2 = s_fit[j];
        float chi2_dist = (v1.z - v2.z) * (v1.z - v2.z);
        chi2_dist = chi2_dist / (v1.cov22 + v2.cov22);
        if (chi2_dist < minChi2Dist && v1.nTracks < v2.nTracks) unique = false;
      }
    }
    const uint64_t mask = __ballot(unique);
    if (unique) multi_final_vertices[(size_t) max_vertices * event_number + __popcll(mask & lanes_below_mask())] = v1;
    if (l == 0) number_of_multi_final_vertices[event_number] = __popcll(mask);
  }
  if (st && status) atomicOr(status, st);
}

hipError_t launch_pv_beamline(const PVLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
  hipLaunchKernelGGL(
    k_pv_beamline, dim3(a.n_list), dim3(pv::block), 0, s, a.event_list, a.offsets_all_velo_tracks,
    a.number_of_events, a.beamline_states, a.beamline, a.pvtracks, a.zhisto, a.zpeaks, a.number_of_zpeaks,
    a.pvtracks_denom, a.multi_fit_vertices, a.number_of_multi_fit_vertices, a.multi_final_vertices,
    a.number_of_multi_final_vertices, a.status);
  return hipGetLastError();
}

} // namespace velo
