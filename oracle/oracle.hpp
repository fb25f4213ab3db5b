// TEST INFRASTRUCTURE — CPU oracle declarations (see common.hpp for the rules of use).
#pragma once
#include "common.hpp"
#include <utility>

namespace vo {

struct SearchParams {
  float phi_tolerance = 0.045f;   // SearchByTriplet.cuh:129
  float max_scatter = 0.08f;      // SearchByTriplet.cuh:130
  unsigned max_skipped_modules = 1; // velo_reconstruction.py:188
  // Seeding as the CPU target dispatches it (track_seeding_vectorized, SearchByTriplet.cu:616-818)
  // instead of the default track_seeding (:298-446).
  bool vectorized_seeding = false;
  // Seed pruning by a learned scorer (no reference counterpart). seed_logits holds one logit per
  // cluster slot of the event (the logit of the seed whose h1 is that hit; NaN where none was
  // scored). A seed is kept when its logit >= seed_threshold. Seeds are decided in groups: the
  // unused h1 hits of each run of seed_block positions of the module pair, compacted in order and
  // cut into groups of seed_block / K (K = the h0 candidates considered); a group none of whose
  // seeds reaches the threshold keeps them all (the plain search as fallback).
  const float* seed_logits = nullptr;
  float seed_threshold = 0.0f;
  unsigned seed_block = 256;
};

// Exact work counts of search_by_triplet, used as the algorithmic op count of the kernel.
struct SearchStats {
  uint64_t seed_evaluations = 0;    // (h0, h2) pairs scored in seeding
  uint64_t forward_evaluations = 0; // h2 candidates scored in forwarding
  uint64_t extrapolations = 0;      // find_forward_candidate calls (hit_phi_16 + binary search)
  uint64_t search_steps = 0;        // binary-search iterations
};

// Observer of every scored seeding triplet (used to extract learning features).
struct SeedSink {
  virtual ~SeedSink() = default;
  virtual void begin_event(unsigned) {}
  virtual void candidate(unsigned, unsigned, unsigned, float) {}
  // A seed written by seeding (h1's module pair `module_pair`), with its scorer features.
  virtual void seed(unsigned, unsigned, unsigned, unsigned, const float*) {}
};

// Scorer features of a seed (h0, h1, h2) whose h1 lies in module pair `module_pair`
// (no reference counterpart): h2's residual to the (h0, h1) line and h1's residual to the
// (h0, h2) line in x and y, the phi steps h0 -> h1 and h1 -> h2, the (h0, h1) slopes, the
// scatter and the module pair, each scaled to order one.
constexpr unsigned n_seed_features = 10;
void seed_features(const Clusters& cc, unsigned h0, unsigned h1, unsigned h2, unsigned module_pair, float* f);

void cache_sp_patterns(SpTables& t);

// Result of make_cluster (MaskedVeloClustering.cu:49-84).
struct ClusterGrow {
  uint64_t pixels;
  uint32_t x, y, n;
};
ClusterGrow make_cluster(uint64_t pixel_map, uint64_t start, unsigned col_lower, unsigned row_lower);

void calculate_number_of_candidates(
  const RawInput& in,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  unsigned block_dim_x,
  uint32_t* number_of_candidates,
  uint32_t* bank_index);

void estimate_input_size(
  const RawInput& in,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  const uint32_t* candidates_offsets,
  uint32_t* estimated_input_size,
  uint32_t* module_candidate_num,
  uint32_t* cluster_candidates);

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
  uint32_t* module_pair_cluster_num,
  char* cluster_container);

// Retina-cluster decoding (retina.cpp). Sensor-pair slots per event: 208 (v2/v3) or 104 (v4).
unsigned retina_sensor_slots(int version);
void calculate_number_of_retinaclusters_each_sensor_pair(
  const RawInput& in,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  uint32_t* each_sensor_pair_size,
  uint32_t* retina_bank_index);
void decode_retinaclusters(
  const RawInput& in,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  unsigned event_start,
  const Geometry& g,
  const uint32_t* offsets_each_sensor_pair_size,
  const uint32_t* retina_bank_index,
  uint32_t* module_pair_cluster_num,
  uint32_t* offsets_module_pair_cluster,
  char* cluster_container,
  uint32_t* hit_permutations,
  int64_t* hit_sorting_key);

void sort_by_phi(
  const uint32_t* event_list,
  unsigned n_list,
  unsigned n_events,
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* module_cluster_num,
  const char* cluster_container,
  char* sorted_container,
  uint32_t* hit_permutation);

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
  SeedSink* sink);

float means_square_fit_chi2(const Clusters& cc, const TrackletHits& t);

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
  uint32_t* number_of_output);

uint32_t prefix_sum(const uint32_t* input, size_t n, uint32_t* output);

// Beamline PV finder (pv.cpp). PVTrack (TrackBeamLineVertexFinder.cuh:14-31) and PV::Vertex
// (PV_Definitions.cuh:64-97), both plain floats (48-byte vertex).
struct PVTrack {
  float z, x, y, tx, ty, W_00, W_11;
};
struct PVVertex {
  float x, y, z, chi2;
  int32_t ndof;
  float cov00, cov10, cov11, cov20, cov21, cov22, nTracks;
};
static_assert(sizeof(PVVertex) == 48, "layout");
constexpr uint32_t pv_status_overflow = 1u << 16;
constexpr unsigned pv_nbins = 3392, pv_max_vertices = 32;
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
  uint32_t& status);

void copy_track_hit_number(
  unsigned n_events,
  const uint32_t* offsets_estimated_input_size,
  const TrackHits* tracks,
  const uint32_t* offsets_velo_tracks,
  const uint32_t* offsets_three_hit,
  uint32_t* velo_track_hit_number,
  uint32_t* offsets_all_velo_tracks);

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
  char* velo_track_hits);

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
  bool fp64_fallback);

bool load_geometry(const char* bytes, size_t n, Geometry& g);

} // namespace vo
