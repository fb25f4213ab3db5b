// Beamline primary-vertex finder for gfx950: the consumer of the VELO beamline states
// (device/PV/beamlinePV/src/pv_beamline_{extrapolate,histo,peak,calculate_denom,multi_fitter,
// cleanup}.cu). The reference runs six kernels with the z histogram in HBM and float atomics;
// here one workgroup per event runs them all with the histogram in LDS:
//
//   1. tracks: PVTrack (z, x, y, tx, ty, 1/c00, 1/c11) of each track, and its histogram window
//   2. histogram: each lane owns 14 consecutive bins and adds, track by track in track order,
//      the Gaussian-integral increments the reference adds with atomics, so every bin sums its
//      terms in the CPU target's order (no atomics, identical sums)
//   3. peaks: the CPU target's protocluster scan and extrema split (pv_beamline_peak.cu, the
//      branch without warp intrinsics) on one lane, reading the histogram from LDS
//   4. denominators: one lane per track, seeds in order
//   5. multi-fitter: one lane per seed, tracks in order (the CPU target's sums), vertices kept in
//      seed order by a ballot
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
constexpr unsigned max_vertices = 32, max_clusters = 200, max_clusteredges = 200, max_subclusters = 50;
constexpr unsigned block = 256, bins_per_lane = (nbins + block - 1) / block; // 14
constexpr unsigned track_chunk = 1024; // histogram windows staged per pass
constexpr uint32_t status_overflow = 1u << 16;

struct Extremum {
  unsigned short index;
  float value, integral;
};
struct Cluster {
  unsigned short izfirst, izlast, izmax;
};

__device__ __forceinline__ float gauss_integral(float x)
{
  const float a = sqrtf(float(2 * order_polynomial + 3));
  const float xi = x / a;
  const float eta = 1.f - xi * xi;
  return 0.5f + xi * (0.5f + eta * (0.25f + eta * 0.1875f));
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
  __shared__ float h[block * bins_per_lane];
  __shared__ uint32_t s_window[track_chunk]; // minbin | maxbin << 16, or 0 for no window
  __shared__ float s_zerr[track_chunk];
  __shared__ float s_z[track_chunk];
  __shared__ float s_zpeaks[max_vertices];
  __shared__ unsigned s_npeaks;
  __shared__ PVVertexOut s_fit[max_vertices];
  __shared__ unsigned s_nfit;
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
  for (unsigned b = t; b < block * bins_per_lane; b += block) h[b] = 0.f;
  __syncthreads(); // the tracks are read back by other lanes below (global memory, same workgroup)

  // 2. Histogram (pv_beamline_histo.cu:28-86), chunks of tracks in order.
  const unsigned lo = t * bins_per_lane, hi = min(lo + bins_per_lane, static_cast<unsigned>(nbins));
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
    for (unsigned i = 0; i < cn; ++i) {
      const uint32_t window = s_window[i];
      if (!window) continue;
      const unsigned minbin = window & 0xFFFFu, maxbin = window >> 16;
      if (maxbin < lo || minbin >= hi) continue;
      const float zerr = s_zerr[i], z = s_z[i];
      const unsigned b0 = max(minbin, lo), b1 = min(maxbin, hi - 1);
      for (unsigned b = b0; b <= b1; ++b) {
        // The increment of bin b: the integral up to its upper edge (1 for the last bin)
        // minus the integral up to its lower edge (0 for the first bin).
        const float upper = b < maxbin ? gauss_integral((zmin + (int) (b + 1) * dz - z) / zerr) : 1.f;
        const float lower = b > minbin ? gauss_integral((zmin + (int) b * dz - z) / zerr) : 0.f;
        h[b] += upper - lower;
      }
    }
    __syncthreads();
  }
  if (zhisto_out)
    for (unsigned b = t; b < nbins; b += block) zhisto_out[(size_t) nbins * event_number + b] = h[b];

  // 3. Peaks (pv_beamline_peak.cu:23-236, the CPU target's branch), on one lane.
  if (t == 0) {
    unsigned short clusteredges[max_clusteredges];
    int number_of_clusteredges = 0;
    bool prevempty = true;
    float integral = h[0];
    for (unsigned short i = 1; i < nbins; ++i) {
      const float zBin = zmin + i * dz;
      const float Z0Err = zBin < SMOG2_pp_separation ? SMOG2_maxTrackZ0Err : pp_maxTrackZ0Err;
      const float inv_maxTrackZ0Err = 1.f / (10.f * Z0Err);
      const float threshold = dz * inv_maxTrackZ0Err;
      const float v = h[i];
      integral += v;
      const bool empty = v < threshold;
      if (empty != prevempty) {
        const float minInSeed = zBin < SMOG2_pp_separation ? SMOG2_minTracksInSeed : pp_minTracksInSeed;
        if (prevempty || integral > minInSeed) {
          if (number_of_clusteredges < (int) max_clusteredges) clusteredges[number_of_clusteredges] = i;
          else st |= status_overflow;
          number_of_clusteredges++;
        }
        else {
          number_of_clusteredges--;
        }
        prevempty = empty;
        integral = v;
      }
    }
    number_of_clusteredges = min(number_of_clusteredges, (int) max_clusteredges);
    unsigned number_of_clusters = 0;
    float* zp = zpeaks + (size_t) max_vertices * event_number;
    // Only the first max_vertices clusters become seeds; the cluster mean needs izmax alone.
    auto add_cluster = [&](unsigned short izmax) {
      if (number_of_clusters < max_vertices) {
        const float* b = h + izmax;
        const float d1 = *b - *(b - 1);
        const float d2 = *b - *(b + 1);
        const float idz = d1 + d2 > 0 ? 0.5f * (d1 - d2) / (d1 + d2) : 0.0f;
        const float z = zmin + dz * (izmax + idz + 0.5f);
        zp[number_of_clusters] = z;
        s_zpeaks[number_of_clusters] = z;
      }
      else {
        st |= status_overflow;
      }
      ++number_of_clusters;
    };
    const unsigned Nproto = number_of_clusteredges / 2;
    for (unsigned short p = 0; p < Nproto; ++p) {
      const unsigned short ibegin = clusteredges[p * 2];
      const unsigned short iend = clusteredges[p * 2 + 1];
      const float mindip = minDipDensity * dz;
      const float minpeak = minDensity * dz;
      // One entry more than the capacity, zeroed: the split below reads extrema[N] when N is even
      // (one past the last extremum, unset in the reference); here it reads zeros.
      Extremum extrema[max_vertices + 1] = {};
      int number_of_extrema = 0;
      auto push = [&](unsigned short index, float value, float integ) {
        if (number_of_extrema < (int) max_vertices) extrema[number_of_extrema++] = Extremum {index, value, integ};
        else st |= status_overflow;
      };
      bool rising = true;
      float integ = h[ibegin];
      push(ibegin, h[ibegin], integ);
      for (unsigned short i = ibegin; i < iend; ++i) {
        const float value = h[i];
        const bool stillrising = h[i + 1] > value;
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
      }
      push(iend, h[iend], integ);
      const int N = number_of_extrema;
      unsigned number_of_subclusters = 0;
      if (N > 3) {
        for (int i = 1; i < (N / 2) + 1; ++i) {
          const float z_extrema = zmin + extrema[2 * i].index * dz;
          const float minInSeed = z_extrema < SMOG2_pp_separation ? SMOG2_minTracksInSeed : pp_minTracksInSeed;
          if (extrema[2 * i].integral - extrema[2 * i - 2].integral > minInSeed) {
            if (number_of_subclusters < max_subclusters) {
              ++number_of_subclusters;
              if (number_of_clusters < max_clusters) add_cluster(extrema[2 * i - 1].index);
              else st |= status_overflow;
            }
            else {
              st |= status_overflow;
            }
          }
        }
      }
      if (number_of_subclusters == 0 && extrema[1].value >= minpeak) {
        if (number_of_clusters < max_clusters) add_cluster(extrema[1].index);
        else st |= status_overflow;
      }
    }
    const unsigned n_peaks = min(number_of_clusters, max_vertices);
    number_of_zpeaks[event_number] = n_peaks;
    s_npeaks = n_peaks;
  }
  __syncthreads();
  const unsigned n_peaks = s_npeaks;

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

  // 5. Multi-fitter (pv_beamline_multi_fitter.cu:28-185): lane s of wave 0 fits seed s.
  if (t < wave_size) {
    bool ok = false;
    PVVertexOut vertex {};
    if (l < n_peaks) {
      bool converged = false, accept = true;
      float vtxcov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      float vx = blx, vy = bly;
      const float seed_z = s_zpeaks[l];
      float vz = seed_z;
      float chi2tot = 0.f, sum_weights = 0.f;
      unsigned nselected = 0;
      const unsigned minTracks = seed_z <= SMOG2_pp_separation ? SMOG2_minNumTracksPerVertex : pp_minNumTracksPerVertex;
      for (unsigned iter = 0; iter < maxFitIter && !converged; ++iter) {
        float a00 = 0.f, a11 = 0.f, a20 = 0.f, a21 = 0.f, a22 = 0.f;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f;
        float local_chi2tot = 0.f, local_sum_weights = 0.f;
        nselected = 0;
        for (unsigned i = 0; i < nt; ++i) {
          const PVTrackOut trk = pvtracks[t0 + i];
          if (trk.z < zmin || trk.z >= zmax) continue;
          const float H_20 = -trk.tx, H_21 = -trk.ty;
          const float HWH_20 = H_20 * trk.W_00, HWH_21 = H_21 * trk.W_11;
          const float HWH_22 = H_20 * H_20 * trk.W_00 + H_21 * H_21 * trk.W_11;
          const float dzv = vz - trk.z;
          const float rx = vx - (trk.x + trk.tx * dzv), ry = vy - (trk.y + trk.ty * dzv);
          const float chi2 = rx * rx * trk.W_00 + ry * ry * trk.W_11;
          if (chi2 < maxChi2) {
            ++nselected;
            const float dz_seed = seed_z - trk.z;
            const float sx = blx - (trk.x + trk.tx * dz_seed), sy = bly - (trk.y + trk.ty * dz_seed);
            const float chi2_seed = sx * sx * trk.W_00 + sy * sy * trk.W_11;
            const float exp_chi2_0 = expf(chi2_seed * (-0.5f));
            const float nom = expf(chi2 * (-0.5f));
            const float den = chi2CutExp + nom;
            const float w = nom / (den + pvtracks_denom[t0 + i] - exp_chi2_0);
            if (w > minWeight) {
              const float hx = rx * trk.W_00, hy = ry * trk.W_11, hz = -trk.tx * rx * trk.W_00 - trk.ty * ry * trk.W_11;
              g0 = g0 + hx * w;
              g1 = g1 + hy * w;
              g2 = g2 + hz * w;
              a00 += w * trk.W_00;
              a11 += w * trk.W_11;
              a20 += w * HWH_20;
              a21 += w * HWH_21;
              a22 += w * HWH_22;
              local_chi2tot += w * chi2;
              local_sum_weights += w;
            }
          }
        }
        chi2tot = 0.f + local_chi2tot;
        sum_weights = 0.f + local_sum_weights;
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
    }
    const uint64_t mask = __ballot(ok);
    const unsigned slot = __popcll(mask & ((1ull << l) - 1ull));
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
    if (unique) multi_final_vertices[(size_t) max_vertices * event_number + __popcll(mask & ((1ull << l) - 1ull))] = v1;
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
