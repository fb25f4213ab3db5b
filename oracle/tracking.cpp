// TEST INFRASTRUCTURE — CPU oracle (see common.hpp for the rules of use).
// Restates, with CPU-target semantics:
//   K4 velo_sort_by_phi              device/velo/search_by_triplet/src/SortByPhi.cu:49-153
//   K5 velo_search_by_triplet        device/velo/search_by_triplet/src/SearchByTriplet.cu:110-609
//      (scalar track_seeding, i.e. the default dispatch; SearchByTriplet.cu:163,229)
//   K6 velo_three_hit_tracks_filter  device/velo/search_by_triplet/src/ThreeHitTracksFilter.cu:48-169
#include "oracle.hpp"

#include <vector>

namespace vo {

// SortByPhi.cu:49-153. Gaps of the container are left untouched (their content is unspecified).
void sort_by_phi(
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* module_cluster_num,
  const char* cluster_container,
  char* sorted_container,
  uint32_t* hit_permutation)
{
  const unsigned total = offsets_estimated_input_size[n_module_pairs * n_events];
  const Clusters in {const_cast<char*>(cluster_container), total, 0};
  const Clusters out {sorted_container, total, 0};
  for (unsigned block = 0; block < n_list; ++block) {
    const unsigned event_number = event_list[block];
    const uint32_t* mp_start = offsets_estimated_input_size + event_number * n_module_pairs;
    const uint32_t* mp_num = module_cluster_num + event_number * n_module_pairs;
    for (unsigned mp = 0; mp < n_module_pairs; ++mp) {
      const unsigned start = mp_start[mp], num = mp_num[mp];
      for (unsigned r = 0; r < num; ++r) {
        const unsigned hi = start + r;
        const int16_t phi = in.phi(hi);
        unsigned position = 0;
        for (unsigned j = 0; j < num; ++j) {
          const unsigned oi = start + j;
          const int16_t ophi = in.phi(oi);
          position += phi > ophi || (hi != oi && phi == ophi && in.id(hi) > in.id(oi));
        }
        hit_permutation[start + position] = hi;
      }
      for (unsigned r = 0; r < num; ++r) {
        const unsigned src = hit_permutation[start + r];
        out.set(start + r, in.id(src), in.x(src), in.y(src), in.z(src), in.phi(src));
      }
    }
  }
}

void seed_features(const Clusters& cc, unsigned h0, unsigned h1, unsigned h2, unsigned module_pair, float* f)
{
  const float x0 = cc.x(h0), y0 = cc.y(h0), z0 = cc.z(h0);
  const float x1 = cc.x(h1), y1 = cc.y(h1), z1 = cc.z(h1);
  const float x2 = cc.x(h2), y2 = cc.y(h2), z2 = cc.z(h2);
  // h2 against the (h0, h1) line: the seeding scatter's components.
  const float td = 1.0f / (z1 - z0);
  const float tx = (x1 - x0) * td;
  const float ty = (y1 - y0) * td;
  const float dz2 = z2 - z0;
  const float dx2 = (x0 + tx * dz2) - x2;
  const float dy2 = (y0 + ty * dz2) - y2;
  // h1 against the (h0, h2) line.
  const float td02 = 1.0f / (z2 - z0);
  const float dz1 = z1 - z0;
  const float dx1 = (x0 + ((x2 - x0) * td02) * dz1) - x1;
  const float dy1 = (y0 + ((y2 - y0) * td02) * dz1) - y1;
  const int16_t dphi01 = static_cast<int16_t>(cc.phi(h1) - cc.phi(h0));
  const int16_t dphi12 = static_cast<int16_t>(cc.phi(h2) - cc.phi(h1));
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

namespace {
struct ModulePair {
  unsigned hit_start;
  unsigned hit_num;
  float z[2];
};

struct HitBase {
  float x, y, z;
};

// SearchByTriplet.cu:272-292
inline void find_forward_candidate(
  const ModulePair& mp,
  const int16_t* hit_phis,
  const HitBase& h0,
  float tx,
  float ty,
  float dz,
  int16_t phi_tolerance,
  int16_t& candidate,
  int16_t& extrapolated_phi,
  SearchStats& st)
{
  const float predx = tx * dz;
  const float predy = ty * dz;
  const float xp = h0.x + predx;
  const float yp = h0.y + predy;
  const int16_t phi = hit_phi_16(xp, yp);
  ++st.extrapolations;
  candidate = static_cast<int16_t>(binary_search_leftmost(
    hit_phis + mp.hit_start, mp.hit_num, static_cast<int16_t>(phi - phi_tolerance), &st.search_steps));
  extrapolated_phi = phi;
}

// SearchByTriplet.cu:298-446 with blockDim == 1.
void track_seeding(
  const Clusters& cc,
  const ModulePair* mpd,
  const bool* hit_used,
  TrackletHits* tracklets,
  uint32_t* tracks_to_follow,
  uint16_t* h1_indices,
  uint32_t* atomics,
  float max_scatter,
  int16_t phi_tolerance,
  unsigned h0_candidates_to_consider,
  unsigned module_pair,
  SearchStats& st,
  SeedSink* sink,
  const SearchParams& sp,
  const float* seed_logits,
  uint32_t& status)
{
  const ModulePair& prev = mpd[0];
  const ModulePair& cur = mpd[1];
  const ModulePair& next = mpd[2];
  // With a scorer, the seeds of a group are decided together (see SearchParams::seed_logits), so
  // they are held here until their group is complete: (rel, h0, h2) of each found seed.
  struct Held {
    unsigned rel;
    uint16_t h0, h2;
  };
  std::vector<Held> held;
  const unsigned group = seed_logits ? sp.seed_block / h0_candidates_to_consider : 0;
  auto append = [&](unsigned h0, unsigned h1, unsigned h2) {
    const unsigned tn = atomics[atomic_ttf]++ % max_tracks_to_follow;
    tracklets[tn] = TrackletHits {{static_cast<uint16_t>(h0), static_cast<uint16_t>(h1), static_cast<uint16_t>(h2)}};
    tracks_to_follow[tn] = bit_seed | tn;
  };
  auto decide = [&]() {
    bool any = false;
    for (const Held& s : held) {
      const float logit = seed_logits[h1_indices[s.rel]];
      if (logit != logit) status |= status_missing_seed_logit;
      any = any || logit >= sp.seed_threshold;
    }
#if VO_MUTANT == 20
    any = true;
#endif
    for (const Held& s : held)
      if (!any || seed_logits[h1_indices[s.rel]] >= sp.seed_threshold) append(s.h0, h1_indices[s.rel], s.h2);
    held.clear();
  };
  // Group number of each compacted h1: its run of seed_block positions and, within the run, its
  // rank among the run's unused hits divided by the group size.
  std::vector<unsigned> group_of;
  for (unsigned r = 0; r < cur.hit_num; ++r) {
    const unsigned h1_index = cur.hit_start + r;
    if (!hit_used[h1_index]) {
      if (seed_logits) {
        const unsigned run = r / sp.seed_block;
        unsigned in_run = 0;
#if VO_MUTANT == 21
        for (unsigned q = 0; q < r; ++q) in_run += !hit_used[cur.hit_start + q];
#else
        for (unsigned q = run * sp.seed_block; q < r; ++q) in_run += !hit_used[cur.hit_start + q];
#endif
        group_of.push_back(run << 16 | in_run / group);
      }
      h1_indices[atomics[atomic_local_hits]++] = h1_index;
    }
  }
  const unsigned n_h1 = atomics[atomic_local_hits];
  for (unsigned rel = 0; rel < n_h1; ++rel) {
    if (seed_logits && rel > 0 && group_of[rel] != group_of[rel - 1]) decide();
    const unsigned h1_index = h1_indices[rel];
    uint16_t best_h0 = 0, best_h2 = 0;
    float best_fit = max_scatter;
    const HitBase h1 {cc.x(h1_index), cc.y(h1_index), cc.z(h1_index)};
    const int16_t h1_phi = cc.phi(h1_index);
    unsigned best_h0s[max_h0_candidates];
    int phi_index = binary_search_leftmost(cc.phi_begin() + prev.hit_start, prev.hit_num, h1_phi, &st.search_steps);
    unsigned found = 0;
    for (unsigned i = 0; i < prev.hit_num && found < h0_candidates_to_consider; ++i) {
      const unsigned sign = i & 0x01;
#if VO_MUTANT == 12
      const int index_diff = sign ? -static_cast<int>(i) : static_cast<int>(i);
#else
      const int index_diff = sign ? i : -static_cast<int>(i);
#endif
      phi_index += index_diff;
      const int in_bounds = phi_index < 0 ? phi_index + static_cast<int>(prev.hit_num) :
                                            (phi_index >= static_cast<int>(prev.hit_num) ?
                                               phi_index - static_cast<int>(prev.hit_num) :
                                               phi_index);
      const unsigned h0_index = prev.hit_start + in_bounds;
      if (!hit_used[h0_index]) best_h0s[found++] = h0_index;
    }
    for (unsigned i = 0; i < found; ++i) {
      const unsigned h0_index = best_h0s[i];
      const HitBase h0 {cc.x(h0_index), cc.y(h0_index), cc.z(h0_index)};
      const float td = 1.0f / (h1.z - h0.z);
      const float txn = h1.x - h0.x;
      const float tyn = h1.y - h0.y;
      const float tx = txn * td;
      const float ty = tyn * td;
      int16_t cand, ephi;
#if VO_MUTANT == 16
      find_forward_candidate(next, cc.phi_begin(), h0, tx, ty, next.z[1] - prev.z[0], phi_tolerance, cand, ephi, st);
#else
      find_forward_candidate(next, cc.phi_begin(), h0, tx, ty, next.z[0] - prev.z[0], phi_tolerance, cand, ephi, st);
#endif
      for (unsigned j = 0; j < next.hit_num; ++j) {
        const unsigned in_bounds = (cand + j) % next.hit_num;
        const unsigned h2_index = next.hit_start + in_bounds;
        const int16_t phi_diff = cc.phi(h2_index) - ephi;
        const int16_t abs_phi_diff = phi_diff < 0 ? -phi_diff : phi_diff;
        if (abs_phi_diff > phi_tolerance) break;
        if (!hit_used[h2_index]) {
          ++st.seed_evaluations;
          const HitBase h2 {cc.x(h2_index), cc.y(h2_index), cc.z(h2_index)};
          const float dz = h2.z - h0.z;
          const float predx = h0.x + tx * dz;
          const float predy = h0.y + ty * dz;
          const float dx = predx - h2.x;
          const float dy = predy - h2.y;
          const float scatter = dx * dx + dy * dy;
          if (sink) sink->candidate(h0_index, h1_index, h2_index, scatter);
#if VO_MUTANT == 15
          if (scatter <= best_fit && scatter < max_scatter) {
#else
          if (scatter < best_fit) {
#endif
            best_fit = scatter;
            best_h0 = h0_index;
            best_h2 = h2_index;
          }
        }
      }
    }
    if (best_fit < max_scatter && sink) {
      float f[n_seed_features];
      seed_features(cc, best_h0, h1_index, best_h2, module_pair, f);
      sink->seed(module_pair, best_h0, h1_index, best_h2, f);
    }
    if (best_fit < max_scatter) {
#if VO_MUTANT == 18
      if (atomics[atomic_ttf] >= max_tracks_to_follow) continue;
#endif
      if (seed_logits)
        held.push_back(Held {rel, best_h0, best_h2});
      else
        append(best_h0, h1_index, best_h2);
    }
  }
  if (seed_logits) decide();
}

// SearchByTriplet.cu:616-818: the CPU target's seeding. The h1 hits are visited in order without
// the compaction into h1_indices (so neither dev_rel_indices nor the local-hits atomic is
// written), and the h0 candidates are scored four at a time in lock step: in each round every
// active lane takes its next unused h2 of its phi window, and the lane of least scatter among
// those below the best so far takes the best. Among equal scatters of one round the lowest lane
// wins (umesimd's imin over the mask; its tie rule is not pinned by anything in the reference).
void track_seeding_vectorized(
  const Clusters& cc,
  const ModulePair* mpd,
  const bool* hit_used,
  TrackletHits* tracklets,
  uint32_t* tracks_to_follow,
  uint32_t* atomics,
  float max_scatter,
  int16_t phi_tolerance,
  unsigned h0_candidates_to_consider,
  unsigned module_pair,
  SearchStats& st,
  SeedSink* sink)
{
  constexpr unsigned lanes = 4; // vector128_length() for float
  const ModulePair& prev = mpd[0];
  const ModulePair& cur = mpd[1];
  const ModulePair& next = mpd[2];
  for (unsigned r = 0; r < cur.hit_num; ++r) {
    const unsigned h1_index = cur.hit_start + r;
    if (hit_used[h1_index]) continue;
    uint16_t best_h0 = 0, best_h2 = 0;
    float best_fit = max_scatter;
    const HitBase h1 {cc.x(h1_index), cc.y(h1_index), cc.z(h1_index)};
    unsigned best_h0s[max_h0_candidates];
    int phi_index =
      binary_search_leftmost(cc.phi_begin() + prev.hit_start, prev.hit_num, cc.phi(h1_index), &st.search_steps);
    unsigned found = 0;
    for (unsigned i = 0; i < prev.hit_num && found < h0_candidates_to_consider; ++i) {
      phi_index += (i & 1) ? static_cast<int>(i) : -static_cast<int>(i);
      const int n = static_cast<int>(prev.hit_num);
      const int in_bounds = phi_index < 0 ? phi_index + n : (phi_index >= n ? phi_index - n : phi_index);
      const unsigned h0_index = prev.hit_start + in_bounds;
      if (!hit_used[h0_index]) best_h0s[found++] = h0_index;
    }
    for (unsigned batch = 0; batch < found; batch += lanes) {
      const unsigned len = batch + lanes < found ? lanes : found - batch;
      HitBase h0[lanes];
      float tx[lanes], ty[lanes];
      int16_t ephi[lanes];
      int cand[lanes];
      unsigned iter[lanes];
      bool active[lanes];
      for (unsigned l = 0; l < len; ++l) {
        const unsigned h0_index = best_h0s[batch + l];
        h0[l] = HitBase {cc.x(h0_index), cc.y(h0_index), cc.z(h0_index)};
        const float td = 1.0f / (h1.z - h0[l].z);
        tx[l] = (h1.x - h0[l].x) * td;
        ty[l] = (h1.y - h0[l].y) * td;
        int16_t c;
        find_forward_candidate(next, cc.phi_begin(), h0[l], tx[l], ty[l], next.z[0] - prev.z[0], phi_tolerance, c, ephi[l], st);
        cand[l] = c;
        iter[l] = ~0u;
        active[l] = true;
      }
      for (;;) {
        unsigned h2_of[lanes];
        bool any = false;
        for (unsigned l = 0; l < len; ++l) {
          while (active[l]) {
            ++iter[l];
            if (iter[l] == next.hit_num) {
              active[l] = false;
              break;
            }
            const unsigned h2_index = next.hit_start + (cand[l] + iter[l]) % next.hit_num;
            const int16_t d = static_cast<int16_t>(cc.phi(h2_index) - ephi[l]);
            if ((d < 0 ? -d : d) > phi_tolerance) {
              active[l] = false;
              break;
            }
            if (!hit_used[h2_index]) {
              h2_of[l] = h2_index;
              break;
            }
          }
          any = any || active[l];
        }
        if (!any) break;
        int pick = -1;
        float pick_scatter = 0.f;
        for (unsigned l = 0; l < len; ++l) {
          if (!active[l]) continue;
          ++st.seed_evaluations;
          const HitBase h2 {cc.x(h2_of[l]), cc.y(h2_of[l]), cc.z(h2_of[l])};
          const float dz = h2.z - h0[l].z;
          const float dx = (h0[l].x + tx[l] * dz) - h2.x;
          const float dy = (h0[l].y + ty[l] * dz) - h2.y;
          const float scatter = dx * dx + dy * dy;
          if (sink) sink->candidate(best_h0s[batch + l], h1_index, h2_of[l], scatter);
#if VO_MUTANT == 19
          if (scatter < best_fit && (pick < 0 || scatter <= pick_scatter)) {
#else
          if (scatter < best_fit && (pick < 0 || scatter < pick_scatter)) {
#endif
            pick = static_cast<int>(l);
            pick_scatter = scatter;
          }
        }
        if (pick >= 0) {
          best_fit = pick_scatter;
          best_h0 = best_h0s[batch + pick];
          best_h2 = h2_of[pick];
        }
      }
    }
    if (best_fit < max_scatter) {
      if (sink) {
        float f[n_seed_features];
        seed_features(cc, best_h0, h1_index, best_h2, module_pair, f);
        sink->seed(module_pair, best_h0, h1_index, best_h2, f);
      }
      const unsigned tn = atomics[atomic_ttf]++ % max_tracks_to_follow;
      tracklets[tn] = TrackletHits {{best_h0, static_cast<uint16_t>(h1_index), best_h2}};
      tracks_to_follow[tn] = bit_seed | tn;
    }
  }
}

// SearchByTriplet.cu:451-609 with blockDim == 1.
void track_forwarding(
  const Clusters& cc,
  bool* hit_used,
  const ModulePair* mpd,
  unsigned diff_ttf,
  uint32_t* tracks_to_follow,
  TrackletHits* three_hit_tracks,
  unsigned prev_ttf,
  TrackletHits* tracklets,
  TrackHits* tracks,
  uint32_t* atomics,
  uint32_t* number_of_velo_tracks_event,
  int16_t phi_tolerance,
  float max_scatter,
  unsigned max_skipped_modules,
  SearchStats& st)
{
  const ModulePair& next = mpd[2];
  for (unsigned e = 0; e < diff_ttf; ++e) {
    const uint32_t full = tracks_to_follow[(prev_ttf + e) % max_tracks_to_follow];
    const bool track_flag = (full & bit_seed) == bit_seed;
    const unsigned skipped = (full & bit_skipped) >> skipped_pos;
    unsigned track_number = full & bit_track_number;
    unsigned number_of_hits;
    TrackHits* t;
    if (track_flag) {
      t = reinterpret_cast<TrackHits*>(tracklets + track_number);
      number_of_hits = 3;
    }
    else {
      t = tracks + track_number;
      number_of_hits = t->hitsNum;
    }
    const unsigned h0_num = t->hits[number_of_hits - 2];
    const unsigned h1_num = t->hits[number_of_hits - 1];
    const HitBase h0 {cc.x(h0_num), cc.y(h0_num), cc.z(h0_num)};
#if VO_MUTANT == 17
    const unsigned h0_module = ((cc.id(h1_num) & 0x3fc0000u) >> 18) / 4;
#else
    const unsigned h0_module = ((cc.id(h0_num) & 0x3fc0000u) >> 18) / 4;
#endif
    const HitBase h1 {cc.x(h1_num), cc.y(h1_num), cc.z(h1_num)};
    const float td = 1.0f / (h1.z - h0.z);
    const float txn = h1.x - h0.x;
    const float tyn = h1.y - h0.y;
    const float tx = txn * td;
    const float ty = tyn * td;
    float best_fit = max_scatter;
    int best_h2 = -1;
    int16_t cand, ephi;
    find_forward_candidate(
      next, cc.phi_begin(), h0, tx, ty, next.z[h0_module % 2] - h0.z, phi_tolerance, cand, ephi, st);
    for (unsigned j = 0; j < next.hit_num; ++j) {
      const unsigned in_bounds = (cand + j) % next.hit_num;
      const unsigned h2_index = next.hit_start + in_bounds;
      const int16_t phi_diff = cc.phi(h2_index) - ephi;
      const int16_t abs_phi_diff = phi_diff < 0 ? -phi_diff : phi_diff;
      if (abs_phi_diff > phi_tolerance) break;
#if VO_MUTANT == 14
      if (hit_used[h2_index]) continue;
#endif
      ++st.forward_evaluations;
      const HitBase h2 {cc.x(h2_index), cc.y(h2_index), cc.z(h2_index)};
      const float dz = h2.z - h0.z;
      const float predx = h0.x + tx * dz;
      const float predy = h0.y + ty * dz;
      const float dx = predx - h2.x;
      const float dy = predy - h2.y;
      const float scatter = dx * dx + dy * dy;
      if (scatter < best_fit) {
        best_fit = scatter;
        best_h2 = h2_index;
      }
    }
    if (best_h2 != -1) {
      hit_used[best_h2] = true;
      if (number_of_hits == 3) {
        hit_used[t->hits[0]] = true;
        hit_used[t->hits[1]] = true;
        hit_used[t->hits[2]] = true;
        track_number = (*number_of_velo_tracks_event)++;
        tracks[track_number].hits[0] = t->hits[0];
        tracks[track_number].hits[1] = t->hits[1];
        tracks[track_number].hits[2] = t->hits[2];
        tracks[track_number].hits[3] = best_h2;
        tracks[track_number].hitsNum = 4;
      }
      else {
        t->hits[t->hitsNum++] = best_h2;
      }
      if (number_of_hits + 1 < max_track_size) {
        const unsigned p = atomics[atomic_ttf]++ % max_tracks_to_follow;
        tracks_to_follow[p] = track_number;
      }
    }
#if VO_MUTANT == 13
    else if (skipped <= max_skipped_modules) {
#else
    else if (skipped < max_skipped_modules) {
#endif
      track_number = ((skipped + 1) << skipped_pos) | (full & (bit_seed | bit_track_number));
      const unsigned p = atomics[atomic_ttf]++ % max_tracks_to_follow;
      tracks_to_follow[p] = track_number;
    }
    else if (number_of_hits == 3) {
      const unsigned p = atomics[atomic_three_hit]++;
      three_hit_tracks[p] = TrackletHits {{t->hits[0], t->hits[1], t->hits[2]}};
    }
  }
}
} // namespace

// SearchByTriplet.cu:110-265 (one event == one block).
void search_by_triplet_event(
  const Clusters& cc,
  const uint32_t* module_hit_start,
  const uint32_t* module_hit_num,
  const Geometry& g,
  TrackHits* tracks,
  TrackletHits* three_hit_tracks,
  TrackletHits* tracklets,
  uint32_t* tracks_to_follow,
  bool* hit_used,
  uint16_t* h1_rel_indices,
  uint32_t* atomics,
  uint32_t* number_of_velo_tracks_event,
  const SearchParams& p,
  SearchStats& st,
  uint32_t& status,
  SeedSink* sink)
{
  const unsigned hit_offset = module_hit_start[0];
  ModulePair mpd[3];
  auto load = [&](unsigned first) {
    for (unsigned i = 0; i < 3; ++i) {
      const unsigned m = first - i;
      mpd[i].hit_start = module_hit_start[m] - hit_offset;
      mpd[i].hit_num = module_hit_num[m];
      mpd[i].z[0] = g.module_zs[2 * m];
      mpd[i].z[1] = g.module_zs[2 * m + 1];
    }
  };
  unsigned first_module_pair = n_module_pairs - 1;
  load(first_module_pair);
  const int16_t phi_tolerance_i16 = hit_phi_float_to_16(p.phi_tolerance);
  // SearchByTriplet.cu:163,229: the CPU target dispatches to the vectorized seeding.
  auto seeding = [&](unsigned h0_candidates, unsigned module_pair) {
    if (p.vectorized_seeding)
      track_seeding_vectorized(
        cc, mpd, hit_used, tracklets, tracks_to_follow, atomics, p.max_scatter, phi_tolerance_i16, h0_candidates,
        module_pair, st, sink);
    else
      track_seeding(
        cc, mpd, hit_used, tracklets, tracks_to_follow, h1_rel_indices, atomics, p.max_scatter, phi_tolerance_i16,
        h0_candidates, module_pair, st, sink, p, p.seed_logits, status);
  };
#if VO_MUTANT == 10
  seeding(seeding_h0_candidates, first_module_pair - 1);
#else
  seeding(initial_seeding_h0_candidates, first_module_pair - 1);
#endif
  unsigned last_ttf = 0;
  --first_module_pair;
  while (first_module_pair > 1) {
    load(first_module_pair);
    const unsigned prev_ttf = last_ttf;
    last_ttf = atomics[atomic_ttf];
    const unsigned diff_ttf = last_ttf - prev_ttf;
    if (diff_ttf >= max_tracks_to_follow) status |= status_ttf_ring_overflow;
    atomics[atomic_local_hits] = 0;
    track_forwarding(
      cc, hit_used, mpd, diff_ttf, tracks_to_follow, three_hit_tracks, prev_ttf, tracklets, tracks, atomics,
      number_of_velo_tracks_event, phi_tolerance_i16, p.max_scatter, p.max_skipped_modules, st);
#if VO_MUTANT == 11
    seeding(initial_seeding_h0_candidates, first_module_pair - 1);
#else
    seeding(seeding_h0_candidates, first_module_pair - 1);
#endif
    --first_module_pair;
  }
  const unsigned prev_ttf = last_ttf;
  last_ttf = atomics[atomic_ttf];
  const unsigned diff_ttf = last_ttf - prev_ttf;
  if (diff_ttf >= max_tracks_to_follow) status |= status_ttf_ring_overflow;
  for (unsigned e = 0; e < diff_ttf; ++e) {
    const uint32_t full = tracks_to_follow[(prev_ttf + e) % max_tracks_to_follow];
    if ((full & bit_seed) == bit_seed) {
      const TrackletHits& t = tracklets[full & bit_track_number];
      three_hit_tracks[atomics[atomic_three_hit]++] = t;
    }
  }
}

void search_by_triplet(
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  const char* sorted_container,
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* module_cluster_num,
  const Geometry& g,
  TrackHits* tracks,
  TrackletHits* tracklets,
  uint32_t* tracks_to_follow,
  TrackletHits* three_hit_tracks,
  bool* hit_used,
  uint32_t* atomics_velo,
  uint16_t* rel_indices,
  uint32_t* number_of_velo_tracks,
  const SearchParams& p,
  SearchStats& st,
  uint32_t& status,
  SeedSink* sink)
{
  const unsigned total = offsets_estimated_input_size[n_module_pairs * n_events];
  const unsigned container = track_container_size(total, n_events);
  for (unsigned block = 0; block < n_list; ++block) {
    const unsigned event_number = event_list[block];
    const uint32_t* mhs = offsets_estimated_input_size + event_number * n_module_pairs;
    const uint32_t* mhn = module_cluster_num + event_number * n_module_pairs;
    const unsigned hit_offset = mhs[0];
    if (mhs[n_module_pairs] - hit_offset > 65536u) status |= status_hit_index_overflow;
    const Clusters cc {const_cast<char*>(sorted_container), total, hit_offset};
    const unsigned toff = track_offset(offsets_estimated_input_size, event_number);
    if (sink) sink->begin_event(event_number);
    SearchParams pe = p;
    if (pe.seed_logits) pe.seed_logits += hit_offset;
    search_by_triplet_event(
      cc, mhs, mhn, g, tracks + toff, three_hit_tracks + toff, tracklets + event_number * max_tracks_to_follow,
      tracks_to_follow + event_number * max_tracks_to_follow, hit_used + hit_offset, rel_indices + hit_offset,
      atomics_velo + event_number * num_atomics, number_of_velo_tracks + event_number, pe, st, status, sink);
    const unsigned cap = (event_number + 1 < n_events ? track_offset(offsets_estimated_input_size, event_number + 1) :
                                                        container) -
                         toff;
    if (number_of_velo_tracks[event_number] > cap) status |= status_track_slots_overflow;
    if (atomics_velo[event_number * num_atomics + atomic_three_hit] > cap) status |= status_three_hit_slots_overflow;
  }
}

// ThreeHitTracksFilter.cu:48-107
float means_square_fit_chi2(const Clusters& cc, const TrackletHits& t)
{
  float s0, sx, sz, sxz, sz2, u0, uy, uz, uyz, uz2;
  s0 = sx = sz = sxz = sz2 = 0.0f;
  u0 = uy = uz = uyz = uz2 = 0.0f;
  for (unsigned h = 0; h < 3; ++h) {
    const unsigned hn = t.hits[h];
    const float x = cc.x(hn), y = cc.y(hn), z = cc.z(hn);
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
  for (unsigned h = 0; h < 3; ++h) {
    const unsigned hn = t.hits[h];
    const float z = cc.z(hn);
    const float x = state_x + state_tx * z;
    const float y = state_y + state_ty * z;
    const float dx = x - cc.x(hn);
    const float dy = y - cc.y(hn);
    chi2 += dx * dx * param_w + dy * dy * param_w;
  }
  return chi2 * 0.5f;
}

// ThreeHitTracksFilter.cu:109-169
void three_hit_tracks_filter(
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  const char* sorted_container,
  const uint32_t* offsets_estimated_input_size,
  const TrackletHits* three_hit_tracks_input,
  const uint32_t* atomics_velo,
  const bool* hit_used_all,
  float max_chi2,
  TrackletHits* output,
  uint32_t* number_of_output)
{
  const unsigned total = offsets_estimated_input_size[n_module_pairs * n_events];
  for (unsigned block = 0; block < n_list; ++block) {
    const unsigned event_number = event_list[block];
    const unsigned hit_offset = offsets_estimated_input_size[event_number * n_module_pairs];
    const bool* hit_used = hit_used_all + hit_offset;
    const Clusters cc {const_cast<char*>(sorted_container), total, hit_offset};
    const unsigned toff = track_offset(offsets_estimated_input_size, event_number);
    const TrackletHits* in = three_hit_tracks_input + toff;
    const unsigned n_in = atomics_velo[event_number * num_atomics + atomic_three_hit];
    TrackletHits* out = output + toff;
    uint32_t* n_out = number_of_output + event_number;
    for (unsigned i = 0; i < n_in; ++i) {
      const TrackletHits& t = in[i];
      const bool any_used = hit_used[t.hits[0]] || hit_used[t.hits[1]] || hit_used[t.hits[2]];
      const float chi2 = means_square_fit_chi2(cc, t);
      if (!any_used && chi2 < max_chi2) out[(*n_out)++] = t;
    }
  }
}

} // namespace vo
