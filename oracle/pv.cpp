// TEST INFRASTRUCTURE — CPU restatement of Allen's beamline primary-vertex finder, the consumer
// of the VELO beamline states (see common.hpp for the rules of use), with CPU-target semantics:
// one thread per block, tracks and seeds in order, atomics as plain increments, and the peak
// finder's non-vectorised branch (the one the CPU target compiles).
//
//   pv_beamline_extrapolate     device/PV/beamlinePV/src/pv_beamline_extrapolate.cu:31-41
//   pv_beamline_histo           device/PV/beamlinePV/src/pv_beamline_histo.cu:28-86
//   pv_beamline_peak            device/PV/beamlinePV/src/pv_beamline_peak.cu:23-236 (#else branch)
//   pv_beamline_calculate_denom device/PV/beamlinePV/src/pv_beamline_calculate_denom.cu:24-46
//   pv_beamline_multi_fitter    device/PV/beamlinePV/src/pv_beamline_multi_fitter.cu:28-185
//   pv_beamline_cleanup         device/PV/beamlinePV/src/pv_beamline_cleanup.cu:95-150
// Constants: device/PV/beamlinePV/include/BeamlinePVConstants.cuh; PVTrack, Extremum, Cluster:
// TrackBeamLineVertexFinder.cuh; PV::Vertex and capacities: PV_Definitions.cuh.
//
// Capacities the reference does not check (extrema per protocluster, clusters, vertices) are
// checked here: what exceeds them is dropped and status bit 16 is raised.
#include "oracle.hpp"

#include <cmath>

namespace vo {

namespace {

constexpr float pv_zmin = -541.f, pv_zmax = 307.f, pv_dz = 0.25f;
constexpr float pp_maxTrackZ0Err = 1.5f, SMOG2_maxTrackZ0Err = 10.f, SMOG2_pp_separation = -341.f;
constexpr float maxTrackBLChi2 = 10.f;
constexpr int order_polynomial = 2;
constexpr float minDensity = 0.f, minDipDensity = 3.f, pp_minTracksInSeed = 2.5f, SMOG2_minTracksInSeed = 1.75f;
constexpr float maxVertexRho2 = 0.3f, maxFitIterF = 7.f, maxChi2 = 12.f, minWeight = 0.f;
constexpr float chi2CutExp = 0.135335283f, maxDeltaZConverged = 0.0005f;
constexpr unsigned pp_minNumTracksPerVertex = 4, SMOG2_minNumTracksPerVertex = 3;
constexpr float minChi2Dist = 25.f;
constexpr unsigned max_vertices = pv_max_vertices, max_clusters = 200, max_clusteredges = 200, max_subclusters = 50;

float gauss_integral(float x)
{
  const float a = std::sqrt(float(2 * order_polynomial + 3));
  const float xi = x / a;
  const float eta = 1.f - xi * xi;
  constexpr float p[] = {0.5f, 0.25f, 0.1875f, 0.15625f};
  return 0.5f + xi * (p[0] + eta * (p[1] + eta * p[2]));
}

struct Extremum {
  unsigned short index;
  float value, integral;
};
struct Cluster {
  unsigned short izfirst, izlast, izmax;
};

} // namespace

void pv_beamline(
  unsigned n_events,
  const uint32_t* offsets_all_velo_tracks,
  const float* beamline_states,
  const float* beamline,
  PVTrack* pvtracks,
  float* zhisto,
  float* zpeaks,
  uint32_t* number_of_zpeaks,
  float* pvtracks_denom,
  PVVertex* multi_fit_vertices,
  uint32_t* number_of_multi_fit_vertices,
  PVVertex* multi_final_vertices,
  uint32_t* number_of_multi_final_vertices,
  uint32_t& status)
{
  const unsigned total = offsets_all_velo_tracks[n_events];
  // pv_beamline_extrapolate: the beamline state as (z, x, y, tx, ty, 1/c00, 1/c11).
  for (unsigned t = 0; t < total; ++t) {
    const float* s = beamline_states + 6 * (size_t) t;
    const float* c = beamline_states + 6 * (size_t) total + 8 * (size_t) t;
    pvtracks[t] = PVTrack {s[2], s[0], s[1], s[3], s[4], 1.f / c[0], 1.f / c[3]};
  }
  for (unsigned e = 0; e < n_events; ++e) {
    const unsigned t0 = offsets_all_velo_tracks[e], nt = offsets_all_velo_tracks[e + 1] - t0;
    const PVTrack* tracks = pvtracks + t0;

    // pv_beamline_histo
    float* h = zhisto + (size_t) pv_nbins * e;
    std::fill(h, h + pv_nbins, 0.f);
    for (unsigned i = 0; i < nt; ++i) {
      const PVTrack& trk = tracks[i];
      if (!(pv_zmin < trk.z && trk.z < pv_zmax)) continue;
      const float diffx2 = (trk.x - beamline[0]) * (trk.x - beamline[0]);
      const float diffy2 = (trk.y - beamline[1]) * (trk.y - beamline[1]);
      const float blchi2 = diffx2 * trk.W_00 + diffy2 * trk.W_11;
      if (blchi2 >= maxTrackBLChi2) continue;
      const float zbin = (trk.z - pv_zmin) / pv_dz;
      const float zweight = trk.tx * trk.tx * trk.W_00 + trk.ty * trk.ty * trk.W_11;
      const float zerr = 1.f / std::sqrt(zweight);
      const float maxTrackZ0Err = trk.z < SMOG2_pp_separation ? SMOG2_maxTrackZ0Err : pp_maxTrackZ0Err;
      if (zerr < maxTrackZ0Err) {
        const float a = std::sqrt(float(2 * order_polynomial + 3));
        const float halfwindow = a * zerr / pv_dz;
        const int minbin = std::max(int(zbin - halfwindow), 1);
        const int maxbin = std::min(int(zbin + halfwindow), int(pv_nbins) - 2);
        if (maxbin >= minbin) {
          float integral = 0;
          for (int b = minbin; b < maxbin; ++b) {
            const float relz = (pv_zmin + (b + 1) * pv_dz - trk.z) / zerr;
            const float thisintegral = gauss_integral(relz);
            h[b] += thisintegral - integral;
            integral = thisintegral;
          }
          h[maxbin] += 1.f - integral;
        }
      }
    }

    // pv_beamline_peak (the CPU target's branch)
    unsigned short clusteredges[max_clusteredges];
    int number_of_clusteredges = 0;
    {
      bool prevempty = true;
      float integral = h[0];
      for (unsigned short i = 1; i < pv_nbins; ++i) {
        const float zBin = pv_zmin + i * pv_dz;
        const float Z0Err = zBin < SMOG2_pp_separation ? SMOG2_maxTrackZ0Err : pp_maxTrackZ0Err;
        const float inv_maxTrackZ0Err = 1.f / (10.f * Z0Err);
        const float threshold = pv_dz * inv_maxTrackZ0Err;
        integral += h[i];
        const bool empty = h[i] < threshold;
        if (empty != prevempty) {
          const float minInSeed = zBin < SMOG2_pp_separation ? SMOG2_minTracksInSeed : pp_minTracksInSeed;
          if (prevempty || integral > minInSeed) {
            if (number_of_clusteredges < (int) max_clusteredges) clusteredges[number_of_clusteredges] = i;
            else status |= pv_status_overflow;
            number_of_clusteredges++;
          }
          else {
            number_of_clusteredges--;
          }
          prevempty = empty;
          integral = h[i];
        }
      }
      number_of_clusteredges = std::min(number_of_clusteredges, (int) max_clusteredges);
    }
    Cluster clusters[max_clusters];
    unsigned number_of_clusters = 0;
    auto add_cluster = [&](Cluster c) {
      if (number_of_clusters < max_clusters) clusters[number_of_clusters++] = c;
      else status |= pv_status_overflow;
    };
    const unsigned Nproto = number_of_clusteredges / 2;
    for (unsigned short p = 0; p < Nproto; ++p) {
      const unsigned short ibegin = clusteredges[p * 2];
      const unsigned short iend = clusteredges[p * 2 + 1];
      const float mindip = minDipDensity * pv_dz;
      const float minpeak = minDensity * pv_dz;
      // One entry more than the capacity, zeroed: the split below reads extrema[N] when N is even
      // (one past the last extremum, unset in the reference); here it reads zeros.
      Extremum extrema[max_vertices + 1] = {};
      int number_of_extrema = 0;
      auto push = [&](Extremum x) {
        if (number_of_extrema < (int) max_vertices) extrema[number_of_extrema++] = x;
        else status |= pv_status_overflow;
      };
      bool rising = true;
      float integral = h[ibegin];
      push({ibegin, h[ibegin], integral});
      for (unsigned short i = ibegin; i < iend; ++i) {
        const float value = h[i];
        const bool stillrising = h[i + 1] > value;
        if (rising && !stillrising && value >= minpeak) {
          const int n = number_of_extrema;
          if (n >= 2) {
            const float dv1 = extrema[n - 2].value - extrema[n - 1].value;
            const float dv2 = value - extrema[n - 1].value;
            if (dv1 > mindip && dv2 > mindip) {
              push({i, value, integral + 0.5f * value});
            }
            else if (dv1 > dv2) {
              number_of_extrema--;
            }
            else {
              number_of_extrema -= 2;
              push({i, value, integral + 0.5f * value});
            }
          }
          else {
            push({i, value, integral + 0.5f * value});
          }
        }
        else if (rising != stillrising) {
          push({i, value, integral + 0.5f * value});
        }
        rising = stillrising;
        integral += value;
      }
      push({iend, h[iend], integral});
      const int N = number_of_extrema;
      Cluster subclusters[max_subclusters];
      unsigned number_of_subclusters = 0;
      if (N > 3) {
        for (int i = 1; i < (N / 2) + 1; ++i) {
          const float z_extrema = pv_zmin + extrema[2 * i].index * pv_dz;
          const float minInSeed = z_extrema < SMOG2_pp_separation ? SMOG2_minTracksInSeed : pp_minTracksInSeed;
          if (extrema[2 * i].integral - extrema[2 * i - 2].integral > minInSeed) {
            if (number_of_subclusters < max_subclusters)
              subclusters[number_of_subclusters++] =
                Cluster {extrema[2 * i - 2].index, extrema[2 * i].index, extrema[2 * i - 1].index};
            else status |= pv_status_overflow;
          }
        }
      }
      if (number_of_subclusters == 0) {
        if (extrema[1].value >= minpeak)
          add_cluster(Cluster {extrema[0].index, extrema[number_of_extrema - 1].index, extrema[1].index});
      }
      else {
        subclusters[0].izfirst = ibegin;
        for (unsigned i = 0; i < number_of_subclusters; i++) add_cluster(subclusters[i]);
      }
    }
    float* zp = zpeaks + (size_t) max_vertices * e;
    const unsigned n_peaks = std::min(number_of_clusters, max_vertices);
    if (number_of_clusters > max_vertices) status |= pv_status_overflow;
    for (unsigned i = 0; i < n_peaks; ++i) {
      const float* b = h + clusters[i].izmax;
      const float d1 = *b - *(b - 1);
      const float d2 = *b - *(b + 1);
      const float idz = d1 + d2 > 0 ? 0.5f * (d1 - d2) / (d1 + d2) : 0.0f;
      zp[i] = pv_zmin + pv_dz * (clusters[i].izmax + idz + 0.5f);
    }
    number_of_zpeaks[e] = n_peaks;

    // pv_beamline_calculate_denom
    float* denom = pvtracks_denom + t0;
    for (unsigned i = 0; i < nt; ++i) {
      float d = 0.f;
      const PVTrack& trk = tracks[i];
      for (unsigned j = 0; j < n_peaks; ++j) {
        const float dz = zp[j] - trk.z;
        const float rx = trk.x + trk.tx * dz, ry = trk.y + trk.ty * dz;
        const float chi2 = rx * rx * trk.W_00 + ry * ry * trk.W_11;
        d += std::exp(chi2 * (-0.5f));
      }
      denom[i] = d;
    }

    // pv_beamline_multi_fitter
    PVVertex* vertices = multi_fit_vertices + (size_t) max_vertices * e;
    unsigned n_fit = 0;
    PVVertex vertex {};
    for (unsigned s = 0; s < n_peaks; ++s) {
      bool converged = false, accept = true;
      float vtxcov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      float vx = beamline[0], vy = beamline[1];
      const float seed_z = zp[s];
      float vz = seed_z;
      float chi2tot = 0.f, sum_weights = 0.f;
      unsigned nselected = 0;
      const unsigned minTracks = seed_z <= SMOG2_pp_separation ? SMOG2_minNumTracksPerVertex : pp_minNumTracksPerVertex;
      for (unsigned iter = 0; iter < maxFitIterF && !converged; ++iter) {
        float a00 = 0.f, a11 = 0.f, a20 = 0.f, a21 = 0.f, a22 = 0.f;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f;
        sum_weights = 0.f;
        nselected = 0;
        chi2tot = 0.f;
        float local_chi2tot = 0.f, local_sum_weights = 0.f;
        for (unsigned i = 0; i < nt; ++i) {
          const PVTrack& trk = tracks[i];
          if (trk.z < pv_zmin || trk.z >= pv_zmax) continue;
          // PVTrackInVertex (TrackBeamLineVertexFinder.cuh:60-81)
          const float H_20 = -trk.tx, H_21 = -trk.ty;
          const float HWH_00 = trk.W_00, HWH_20 = H_20 * trk.W_00, HWH_11 = trk.W_11, HWH_21 = H_21 * trk.W_11;
          const float HWH_22 = H_20 * H_20 * trk.W_00 + H_21 * H_21 * trk.W_11;
          const float dz = vz - trk.z;
          const float rx = vx - (trk.x + trk.tx * dz), ry = vy - (trk.y + trk.ty * dz);
          const float chi2 = rx * rx * trk.W_00 + ry * ry * trk.W_11;
          if (chi2 < maxChi2) {
            ++nselected;
            const float dz_seed = seed_z - trk.z;
            const float sx = beamline[0] - (trk.x + trk.tx * dz_seed), sy = beamline[1] - (trk.y + trk.ty * dz_seed);
            const float chi2_seed = sx * sx * trk.W_00 + sy * sy * trk.W_11;
            const float exp_chi2_0 = std::exp(chi2_seed * (-0.5f));
            const float nom = std::exp(chi2 * (-0.5f));
            const float den = chi2CutExp + nom;
            const float w = nom / (den + denom[i] - exp_chi2_0);
            if (w > minWeight) {
              const float hx = rx * trk.W_00, hy = ry * trk.W_11, hz = -trk.tx * rx * trk.W_00 - trk.ty * ry * trk.W_11;
              g0 = g0 + hx * w;
              g1 = g1 + hy * w;
              g2 = g2 + hz * w;
              a00 += w * HWH_00;
              a11 += w * HWH_11;
              a20 += w * HWH_20;
              a21 += w * HWH_21;
              a22 += w * HWH_22;
              local_chi2tot += w * chi2;
              local_sum_weights += w;
            }
          }
        }
        chi2tot += local_chi2tot;
        sum_weights += local_sum_weights;
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
          const float dzv = -1.f * (vtxcov[3] * g0 + vtxcov[4] * g1 + vtxcov[5] * g2);
          chi2tot += dx * g0 + dy * g1 + dzv * g2;
          vx = vx + dx;
          vy = vy + dy;
          vz = vz + dzv;
          converged = std::fabs(dzv) < maxDeltaZConverged;
        }
        else {
          converged = true;
          accept = false;
        }
      }
      if (accept) {
        vertex.chi2 = chi2tot;
        vertex.x = vx;
        vertex.y = vy;
        vertex.z = vz;
        vertex.cov00 = vtxcov[0];
        vertex.cov10 = vtxcov[1];
        vertex.cov11 = vtxcov[2];
        vertex.cov20 = vtxcov[3];
        vertex.cov21 = vtxcov[4];
        vertex.cov22 = vtxcov[5];
        vertex.nTracks = sum_weights;
        const float bdx = vertex.x - beamline[0], bdy = vertex.y - beamline[1];
        const float rho2 = bdx * bdx + bdy * bdy;
        const unsigned minT = vertex.z <= SMOG2_pp_separation ? SMOG2_minNumTracksPerVertex : pp_minNumTracksPerVertex;
        if (nselected >= minT && rho2 < maxVertexRho2) vertices[n_fit++] = vertex;
      }
    }
    number_of_multi_fit_vertices[e] = n_fit;

    // pv_beamline_cleanup
    PVVertex* finals = multi_final_vertices + (size_t) max_vertices * e;
    unsigned n_final = 0;
    for (unsigned i = 0; i < n_fit; ++i) {
      bool unique = true;
      const PVVertex& v1 = vertices[i];
      for (unsigned j = 0; j < n_fit; ++j) {
        if (i == j) continue;
        const PVVertex& v2 = vertices[j];
        float chi2_dist = (v1.z - v2.z) * (v1.z - v2.z);
        chi2_dist = chi2_dist / (v1.cov22 + v2.cov22);
        if (chi2_dist < minChi2Dist && v1.nTracks < v2.nTracks) unique = false;
      }
      if (unique) finals[n_final++] = v1;
    }
    number_of_multi_final_vertices[e] = n_final;
  }
}

} // namespace vo
