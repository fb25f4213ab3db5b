// TEST INFRASTRUCTURE — CPU oracle (see common.hpp for the rules of use).
// Restates, with CPU-target semantics:
//   host_prefix_sum                host/prefix_sum/src/HostPrefixSum.cpp:14-107 (portable loop form)
//   K7 velo_copy_track_hit_number  device/velo/consolidate_tracks/src/VeloCopyTrackHitNumber.cu:49-78
//   K8 velo_consolidate_tracks     device/velo/consolidate_tracks/src/VeloConsolidateTracks.cu:120-192
//   K9 velo_kalman_filter          device/velo/simplified_kalman_filter/src/VeloKalmanFilter.cu:335-418,
//                                  include/VeloKalmanFilter.cuh:25-158
// plus an FP64 restatement of K9 used to state the FP32 error.
#include "oracle.hpp"

namespace vo {

// HostPrefixSum.cpp:76-107: output[0] = 0, output[i+1] = sum(input[0..i]), total = output[n].
uint32_t prefix_sum(const uint32_t* input, size_t n, uint32_t* output)
{
  uint32_t sum = 0;
  output[0] = 0;
  for (size_t i = 0; i < n; ++i) {
    sum += input[i];
    output[i + 1] = sum;
  }
  return sum;
}

// VeloCopyTrackHitNumber.cu:49-78 (grid = number of events, not the event list).
void copy_track_hit_number(
  unsigned n_events,
  const uint32_t* offsets_estimated_input_size,
  const TrackHits* tracks,
  const uint32_t* offsets_velo_tracks,
  const uint32_t* offsets_three_hit,
  uint32_t* velo_track_hit_number,
  uint32_t* offsets_all_velo_tracks)
{
  offsets_all_velo_tracks[0] = 0;
  for (unsigned e = 0; e < n_events; ++e) {
    const unsigned toff = track_offset(offsets_estimated_input_size, e);
    const TrackHits* ev_tracks = tracks + toff;
    const unsigned n_tracks = offsets_velo_tracks[e + 1] - offsets_velo_tracks[e];
    const unsigned n_three = offsets_three_hit[e + 1] - offsets_three_hit[e];
    uint32_t* out = velo_track_hit_number + offsets_velo_tracks[e] + offsets_three_hit[e];
    for (unsigned i = 0; i < n_tracks; ++i) out[i] = ev_tracks[i].hitsNum;
    for (unsigned i = 0; i < n_three; ++i) out[n_tracks + i] = 3;
    offsets_all_velo_tracks[e + 1] = offsets_velo_tracks[e + 1] + offsets_three_hit[e + 1];
  }
}

// VeloConsolidateTracks.cu:120-192. The consolidated hits use the Clusters layout over
// the total number of hits in tracks (VeloConsolidated.cuh:310-335).
void consolidate_tracks(
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  const uint32_t* offsets_estimated_input_size,
  const TrackHits* tracks,
  const TrackletHits* three_hit_tracks_output,
  const uint32_t* offsets_three_hit,
  const uint32_t* offsets_all_velo_tracks,
  const uint32_t* offsets_velo_track_hit_number,
  const char* sorted_container,
  char* velo_track_hits)
{
  const unsigned total_clusters = offsets_estimated_input_size[n_module_pairs * n_events];
  const unsigned total_hits = offsets_velo_track_hit_number[offsets_all_velo_tracks[n_events]];
  for (unsigned block = 0; block < n_list; ++block) {
    const unsigned e = event_list[block];
    const unsigned toff = track_offset(offsets_estimated_input_size, e);
    const unsigned n_total = offsets_all_velo_tracks[e + 1] - offsets_all_velo_tracks[e];
    const unsigned n_three = offsets_three_hit[e + 1] - offsets_three_hit[e];
    const unsigned n_main = n_total - n_three;
    const unsigned hit_offset = offsets_estimated_input_size[e * n_module_pairs];
    const Clusters cc {const_cast<char*>(sorted_container), total_clusters, hit_offset};
    for (unsigned i = 0; i < n_total; ++i) {
      const unsigned global_track = offsets_all_velo_tracks[e] + i;
      const Clusters out {velo_track_hits, total_hits, offsets_velo_track_hit_number[global_track]};
      const uint16_t* hits;
      unsigned nh;
      if (i < n_main) {
        hits = tracks[toff + i].hits;
        nh = tracks[toff + i].hitsNum;
      }
      else {
        hits = three_hit_tracks_output[toff + i - n_main].hits;
        nh = 3;
      }
      for (unsigned h = 0; h < nh; ++h) {
        const unsigned hi = hits[h];
        out.xyz()[3 * (size_t)(out.offset + h)] = cc.x(hi);
        out.xyz()[3 * (size_t)(out.offset + h) + 1] = cc.y(hi);
        out.xyz()[3 * (size_t)(out.offset + h) + 2] = cc.z(hi);
        out.ids()[out.offset + h] = cc.id(hi);
      }
    }
  }
}

namespace {
template<typename F>
struct State {
  F x, y, z, tx, ty;
  F c00, c20, c22, c11, c31, c33;
};
template<typename F>
struct Mini {
  F x, y, z, tx, ty;
};

// VeloKalmanFilter.cuh:25-60
// Conditioning limit of the FP64 covariance fallback: the FP32 update (1 - Kx) * predcovXX loses
// about kappa * 2^-24 of relative accuracy, kappa = 1 + whit * predcovXX; above kappa = 100 the
// fallback recomputes the step's covariance prediction and update in FP64, the update in a form
// without that cancellation ((1 - Kx) = R / whit), and rounds the three terms to FP32. The state
// update keeps the reference's FP32 arithmetic (no reference counterpart).
constexpr float kalman_kappa_limit_minus_one = 99.f;

template<typename F>
inline void filter_step(
  F z,
  F zhit,
  F xhit,
  F whit,
  F& x,
  F& tx,
  F& covXX,
  F& covXTx,
  F& covTxTx,
  bool fp64_fallback = false)
{
  const F dz = zhit - z;
  const F predx = x + dz * tx;
  const F dz_t_covTxTx = dz * covTxTx;
  const F predcovXTx = covXTx + dz_t_covTxTx;
  const F dx_t_covXTx = dz * covXTx;
  const F predcovXX = covXX + 2 * dx_t_covXTx + dz * dz_t_covTxTx;
  const F predcovTxTx = covTxTx;
  const F R = F(1) / ((F(1) / whit) + predcovXX);
  const F Kx = predcovXX * R;
  const F KTx = predcovXTx * R;
  const F r = xhit - predx;
  x = predx + Kx * r;
  tx = tx + KTx * r;
  if (fp64_fallback && predcovXX * whit > F(kalman_kappa_limit_minus_one)) {
    // The prediction is ill-conditioned too (covXX is a few ulps of dz^2 * covTxTx), so the
    // whole covariance part of the step runs in FP64 from the stored FP32 covariance.
    const double dzd = dz, c00 = covXX, c20 = covXTx, c22 = covTxTx;
    const double pXTx = c20 + dzd * c22;
    const double pXX = c00 + 2.0 * (dzd * c20) + dzd * (dzd * c22);
    const double pTxTx = c22;
    const double iw = 1.0 / static_cast<double>(whit);
    const double Rd = 1.0 / (iw + pXX);
    covXX = static_cast<F>(pXX * iw * Rd);
    covXTx = static_cast<F>(pXTx * iw * Rd);
    covTxTx = static_cast<F>(pTxTx - pXTx * pXTx * Rd);
    return;
  }
  covXX = (1 - Kx) * predcovXX;
  covXTx = (1 - Kx) * predcovXTx;
  covTxTx = predcovTxTx - KTx * predcovXTx;
}

// VeloKalmanFilter.cuh:67-158
template<bool upstream, typename F>
State<F> simplified_fit(
  const float* xyz,
  unsigned nh,
  const Mini<F>& at_beamline,
  const F* beamline,
  bool backward,
  bool fp64_fallback)
{
  const int direction = (backward ? 1 : -1) * (upstream ? 1 : -1);
  const F noise2PerLayer = F(1e-8f) + F(7e-6f) * (at_beamline.tx * at_beamline.tx + at_beamline.ty * at_beamline.ty);
  int firsthit = 0;
  int lasthit = static_cast<int>(nh) - 1;
  int dhit = 1;
  if ((F(xyz[3 * lasthit + 2]) - F(xyz[3 * firsthit + 2])) * direction < 0) {
    std::swap(firsthit, lasthit);
    dhit = -1;
  }
  State<F> s;
  s.x = xyz[3 * firsthit];
  s.y = xyz[3 * firsthit + 1];
  s.z = xyz[3 * firsthit + 2];
  s.tx = at_beamline.tx;
  s.ty = at_beamline.ty;
  s.c00 = F(param_w_inverted);
  s.c11 = F(param_w_inverted);
  s.c20 = 0;
  s.c31 = 0;
  s.c22 = 1;
  s.c33 = 1;
  for (int i = firsthit + dhit; i != lasthit + dhit; i += dhit) {
    const F hx = xyz[3 * i], hy = xyz[3 * i + 1], hz = xyz[3 * i + 2];
    s.c22 += noise2PerLayer;
    s.c33 += noise2PerLayer;
    filter_step<F>(s.z, hz, hx, F(param_w), s.x, s.tx, s.c00, s.c20, s.c22, fp64_fallback);
    filter_step<F>(s.z, hz, hy, F(param_w), s.y, s.ty, s.c11, s.c31, s.c33, fp64_fallback);
    s.z = hz;
  }
  s.c22 += noise2PerLayer;
  s.c33 += noise2PerLayer;
  F delta_z = 0;
  if (upstream) {
    delta_z = (s.tx * (beamline[0] - s.x) + s.ty * (beamline[1] - s.y)) / (s.tx * s.tx + s.ty * s.ty);
  }
  else {
    delta_z = F(z_end_velo) - s.z;
  }
  s.x = s.x + s.tx * delta_z;
  s.y = s.y + s.ty * delta_z;
  s.z = s.z + delta_z;
  const F dz2 = delta_z * delta_z;
  s.c00 += dz2 * s.c22 + F(2) * delta_z * s.c20;
  s.c11 += dz2 * s.c33 + F(2) * delta_z * s.c31;
  s.c20 += s.c22 * delta_z;
  s.c31 += s.c33 * delta_z;
  return s;
}

// VeloKalmanFilter.cu:335-355
template<typename F>
Mini<F> linear_fit(const float* xyz, unsigned nh, const F* beamline)
{
  Mini<F> s;
  const F fx = xyz[0], fy = xyz[1], fz = xyz[2];
  const F lx = xyz[3 * (nh - 1)], ly = xyz[3 * (nh - 1) + 1], lz = xyz[3 * (nh - 1) + 2];
  s.tx = (lx - fx) / (lz - fz);
  s.ty = (ly - fy) / (lz - fz);
  const F delta_z = (s.tx * (beamline[0] - lx) + s.ty * (beamline[1] - ly)) / (s.tx * s.tx + s.ty * s.ty);
  s.x = lx + s.tx * delta_z;
  s.y = ly + s.ty * delta_z;
  s.z = lz + delta_z;
  return s;
}

// Velo::Consolidated::States_t::set (VeloConsolidated.cuh:578-594): x,y,z,tx,ty and the six
// covariance terms; qop, chi2 and ndof are not written by the kernel (left as they were).
template<typename F, typename Out>
inline void store_state(Out* base, unsigned total_tracks, unsigned i, const State<F>& s)
{
  Out* st = base + 6 * (size_t) i;
  st[0] = s.x;
  st[1] = s.y;
  st[2] = s.z;
  st[3] = s.tx;
  st[4] = s.ty;
  Out* cv = base + 6 * (size_t) total_tracks + 8 * (size_t) i;
  cv[0] = s.c00;
  cv[1] = s.c20;
  cv[2] = s.c22;
  cv[3] = s.c11;
  cv[4] = s.c31;
  cv[5] = s.c33;
}

// Per-track work of the kernel (VeloKalmanFilter.cu:386-417), in precision F.
template<typename F>
void fit_track(
  const float* xyz,
  unsigned nh,
  const F* beamline,
  bool& backward,
  State<F>& bl,
  State<F>& ev,
  bool fp64_fallback = false)
{
  const Mini<F> lin = linear_fit<F>(xyz, nh, beamline);
  backward = lin.z > F(xyz[2]);
  bl = simplified_fit<true, F>(xyz, nh, lin, beamline, backward, fp64_fallback);
  const Mini<F> blm {bl.x, bl.y, bl.z, bl.tx, bl.ty};
  ev = simplified_fit<false, F>(xyz, nh, blm, beamline, backward, fp64_fallback);
}
} // namespace

// VeloKalmanFilter.cu:357-418. The hits of track t are read from the consolidated
// container: [u32 id x H][f32 xyz x 3H], H = total hits in tracks.
void kalman_filter(
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  const uint32_t* offsets_all_velo_tracks,
  const uint32_t* offsets_velo_track_hit_number,
  const char* velo_track_hits,
  const float* beamline,
  bool* is_backward,
  float* beamline_states,
  float* endvelo_states,
  double* beamline_states_f64,
  double* endvelo_states_f64,
  bool fp64_fallback)
{
  const unsigned total_tracks = offsets_all_velo_tracks[n_events];
  const unsigned total_hits = offsets_velo_track_hit_number[total_tracks];
  const float* xyz_all = reinterpret_cast<const float*>(velo_track_hits) + total_hits;
  const double beamline_d[2] = {beamline[0], beamline[1]};
  for (unsigned block = 0; block < n_list; ++block) {
    const unsigned e = event_list[block];
    for (unsigned t = offsets_all_velo_tracks[e]; t < offsets_all_velo_tracks[e + 1]; ++t) {
      const unsigned h0 = offsets_velo_track_hit_number[t];
      const unsigned nh = offsets_velo_track_hit_number[t + 1] - h0;
      const float* xyz = xyz_all + 3 * (size_t) h0;
      bool backward;
      State<float> bl, ev;
      fit_track<float>(xyz, nh, beamline, backward, bl, ev, fp64_fallback);
      is_backward[t] = backward;
      store_state(beamline_states, total_tracks, t, bl);
      store_state(endvelo_states, total_tracks, t, ev);
      if (beamline_states_f64) {
        bool bwd_d;
        State<double> bld, evd;
        fit_track<double>(xyz, nh, beamline_d, bwd_d, bld, evd);
        store_state(beamline_states_f64, total_tracks, t, bld);
        store_state(endvelo_states_f64, total_tracks, t, evd);
      }
    }
  }
}

} // namespace vo
