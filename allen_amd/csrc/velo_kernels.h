// Internal launch descriptors between the C ABI (velo_api.hip) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace velo {

struct CandidatesLaunch {
  int version;
  const char* raw;
  const uint32_t* offsets;
  const uint32_t* sizes;
  const uint32_t* types;
  bool mep = false; // MEP layout (MEPTools.h:59-200)
  const uint32_t* event_list; // NULL = identity
  unsigned n_list;
  unsigned event_start;
  unsigned block_dim_x;
  uint32_t* number_of_candidates;
  uint32_t* bank_index;
  // Whole-sequence path (identity event list only): also write the bank directory, per
  // (event, module pair) up to 8 BankEntry records in bank order and their number.
  void* directory = nullptr;
  uint32_t* directory_count = nullptr;
};

// Retina-cluster decoding (velo_retina.hip).
struct RetinaSizesLaunch {
  int version;
  const char* raw;
  const uint32_t* offsets;
  const uint32_t* sizes;
  const uint32_t* types;
  bool mep = false; // MEP layout (MEPTools.h:59-200)
  const uint32_t* event_list; // NULL = identity
  unsigned n_list;
  unsigned event_start;
  uint32_t* each_sensor_pair_size; // slots per event: 208 (v2/v3) or 104 (v4)
  uint32_t* retina_bank_index;     // slots
};
struct RetinaDecodeLaunch {
  int version;
  const char* raw;
  const uint32_t* offsets;
  const uint32_t* sizes;
  const uint32_t* types;
  bool mep = false; // MEP layout (MEPTools.h:59-200)
  const uint32_t* event_list; // NULL = identity
  unsigned n_list;
  unsigned event_start;
  const void* geometry;
  const uint32_t* offsets_each_sensor_pair_size;
  const uint32_t* retina_bank_index;
  const uint32_t* number_of_events;
  uint32_t* module_cluster_num;         // 26 per event
  uint32_t* offsets_module_pair_cluster; // 26 per event + 1
  char* cluster_container;
  uint32_t* hit_permutations;
  int64_t* hit_sorting_key;
  // Entries of offsets_module_pair_cluster written, by (event, module pair) index over all
  // events whether listed or not (DecodeRetinaClusters.cu:11-36): 26 x host_number_of_events.
  unsigned populate_items = 0;
};
unsigned retina_sensor_slots(int version);
hipError_t launch_retina_sensor_sizes(const RetinaSizesLaunch& a, hipStream_t s);
hipError_t launch_retina_decode(const RetinaDecodeLaunch& a, hipStream_t s);

// Beamline PV finder (velo_pv.hip). PVTrack (TrackBeamLineVertexFinder.cuh:14-31) with plain
// floats, PV::Vertex (PV_Definitions.cuh:64-97), 48 bytes.
struct PVTrackOut {
  float z, x, y, tx, ty, W_00, W_11;
};
struct PVVertexOut {
  float x, y, z, chi2;
  int32_t ndof;
  float cov00, cov10, cov11, cov20, cov21, cov22, nTracks;
};
static_assert(sizeof(PVVertexOut) == 48, "layout");
struct PVLaunch {
  const uint32_t* event_list; // NULL = identity
  unsigned n_list;
  const uint32_t* number_of_events;
  const uint32_t* offsets_all_velo_tracks;
  const float* beamline_states;
  const float* beamline;
  PVTrackOut* pvtracks;
  float* zhisto; // optional
  float* zpeaks;
  uint32_t* number_of_zpeaks;
  float* pvtracks_denom;
  PVVertexOut* multi_fit_vertices;
  uint32_t* number_of_multi_fit_vertices;
  PVVertexOut* multi_final_vertices;
  uint32_t* number_of_multi_final_vertices;
  uint32_t* status; // optional
};
hipError_t launch_pv_beamline(const PVLaunch& a, hipStream_t s);

struct EstimateLaunch {
  int version;
  const char* raw;
  const uint32_t* offsets;
  const uint32_t* sizes;
  const uint32_t* types;
  bool mep = false; // MEP layout (MEPTools.h:59-200)
  const uint32_t* event_list;
  unsigned n_list;
  unsigned event_start;
  const uint32_t* candidates_offsets;
  uint32_t* estimated_input_size;
  uint32_t* module_candidate_num;
  uint32_t* cluster_candidates;
};

struct MaskedClusteringLaunch {
  int version;
  const char* raw;
  const uint32_t* offsets;
  const uint32_t* sizes;
  const uint32_t* types;
  bool mep = false; // MEP layout (MEPTools.h:59-200)
  const uint32_t* event_list;
  unsigned n_list;
  unsigned event_start;
  const uint32_t* number_of_events;
  const uint32_t* offsets_estimated_input_size;
  const uint32_t* module_candidate_num;
  const uint32_t* cluster_candidates;
  const uint32_t* candidates_offsets;
  const uint32_t* bank_index;
  const void* geometry;
  const uint8_t* sp_patterns;
  const float* sp_fx;
  const float* sp_fy;
  uint32_t* module_cluster_num;
  char* cluster_container;
};

struct SortLaunch {
  const uint32_t* event_list;
  unsigned n_list;
  const uint32_t* number_of_events;
  const uint32_t* offsets_estimated_input_size;
  const uint32_t* module_cluster_num;
  const char* cluster_container;
  char* sorted_container;
  uint32_t* hit_permutation;
};

struct FusedClusterLaunch {
  int version;
  const char* raw;
  const uint32_t* offsets;
  const uint32_t* sizes;
  const uint32_t* types;
  bool mep = false; // MEP layout (MEPTools.h:59-200)
  unsigned n_events;
  const uint32_t* offsets_estimated_input_size;
  const void* geometry;
  const uint8_t* sp_patterns;
  const float* sp_fx;
  const float* sp_fy;
  uint32_t* module_cluster_num;
  char* sorted_container;
  char* unsorted_scratch;
  uint32_t* status;
};

struct EstimateMpLaunch {
  int version;
  bool resident_grid = false; // one wave per item (default) or a resident grid looping over items
  const char* raw;
  const void* directory;
  const uint32_t* directory_count;
  unsigned n_events;
  uint32_t* estimated_input_size;
  uint8_t* sp_flags; // one byte per 32-bit word of the raw buffer
};

struct ClusterMpLaunch {
  int version;
  bool resident_grid = false;
  const char* raw;
  const void* directory;
  const uint32_t* directory_count;
  unsigned n_events;
  const uint32_t* offsets_estimated_input_size;
  const uint8_t* sp_flags;
  const void* geometry;
  const uint8_t* sp_patterns;
  const float* sp_fx;
  const float* sp_fy;
  uint32_t* module_cluster_num;
  char* sorted_container;
  char* unsorted_scratch;
};

// What search_by_triplet does with the seeds it writes: nothing more, record them (SeedLog), or
// score them with the learned scorer and keep those at or above the threshold.
enum { seed_mode_plain = 0, seed_mode_log = 1, seed_mode_scored = 2 };

// Every seed written by seeding, appended in no particular order: records (u32 event, module
// pair of h1, h0, h1, h2; event-relative hit indices) and n_seed_features floats each.
constexpr unsigned n_seed_features = 10;
struct SeedLog {
  uint32_t* counter = nullptr; // zero at launch; NULL = no log
  uint32_t* records = nullptr;
  float* features = nullptr;
  unsigned capacity = 0;
  float* logits = nullptr; // scored mode: each seed's logit
};

struct SearchLaunch {
  const uint32_t* event_list;
  unsigned n_list;
  const uint32_t* number_of_events;
  const char* sorted_container;
  const uint32_t* offsets_estimated_input_size;
  const uint32_t* module_cluster_num;
  const void* geometry;
  void* tracks;
  void* tracklets;
  uint32_t* tracks_to_follow;
  void* three_hit_tracks;
  uint8_t* hit_used;
  uint32_t* atomics_velo;
  uint16_t* rel_indices;
  uint32_t* number_of_velo_tracks;
  uint32_t* status;
  float phi_tolerance;
  float max_scatter;
  unsigned max_skipped_modules;
  unsigned block_dim;          // 0 = 256
  unsigned lds_hits;           // event hits staged in LDS (16 B each); 0 = search_default_lds_hits
  unsigned long long* work_stats; // optional: seeding evals, forwarding evals, extrapolations, search steps
  unsigned long long* event_clock = nullptr; // optional: per event, event_clock_words u64 of wall clock ticks
  // Optional: persistent mode, a resident grid whose workgroups take event slots from this
  // counter (zero at launch) instead of one workgroup per event.
  uint32_t* work_counter = nullptr;
  SeedLog seed_log;
  // Scored mode: the scorer's FP32 parameters (mlp::n_params) and the logit threshold.
  const float* scorer_params = nullptr;
  float scorer_threshold = 0.f;
  // LDS window of the ring of tracks to follow (a power of two, 64..512) and capacity of the
  // seeding's list of unused hits (a multiple of 16, up to 1024); 0 = 512 and 1024.
  unsigned ring_window = 0;
  unsigned list_cap = 0;
};

struct ThreeHitFilterLaunch {
  const uint32_t* event_list;
  unsigned n_list;
  const uint32_t* number_of_events;
  const char* sorted_container;
  const uint32_t* offsets_estimated_input_size;
  const void* three_hit_tracks_input;
  const uint32_t* atomics_velo;
  const uint8_t* hit_used;
  void* three_hit_tracks_output;
  uint32_t* number_of_three_hit_tracks_output;
  float max_chi2;
};

struct CopyHitNumberLaunch {
  unsigned n_events;
  const void* tracks;
  const uint32_t* offsets_velo_tracks;
  const uint32_t* offsets_estimated_input_size;
  const uint32_t* offsets_three_hit;
  uint32_t* velo_track_hit_number;
  uint32_t* offsets_all_velo_tracks;
};

struct ConsolidateLaunch {
  const uint32_t* event_list;
  unsigned n_list;
  const uint32_t* number_of_events;
  const uint32_t* offsets_all_velo_tracks;
  const void* tracks;
  const uint32_t* offsets_velo_track_hit_number;
  const char* sorted_container;
  const uint32_t* offsets_estimated_input_size;
  const void* three_hit_tracks_output;
  const uint32_t* offsets_three_hit;
  char* velo_track_hits;
  uint8_t* accepted_velo_tracks; // optional
  bool lds_map = true;           // hits find their track through an LDS map (else a binary search)
};

struct KalmanLaunch {
  const uint32_t* event_list;
  unsigned n_list;
  const uint32_t* number_of_events;
  const uint32_t* offsets_all_velo_tracks;
  const uint32_t* offsets_velo_track_hit_number;
  const char* velo_track_hits;
  const float* beamline;
  uint8_t* is_backward;
  char* beamline_states;
  char* endvelo_states;
  unsigned fp64_fallback = 0; // FP64 covariance step where the FP32 update is ill-conditioned
  // The reference's view input and outputs (velo_amd.h layouts): with tracks_view the tracks and
  // their hits are read through the views; the KalmanStates views are written when set.
  const void* tracks_view = nullptr;
  void* beamline_states_view = nullptr;
  void* endvelo_states_view = nullptr;
};

// create_velo_views (VeloConsolidateTracks.cu:11-46): one workgroup per event of the batch.
struct VeloViewsLaunch {
  unsigned n_events; // host_number_of_events
  const uint32_t* number_of_events;
  const uint32_t* offsets_all_velo_tracks;
  const uint32_t* offsets_velo_track_hit_number;
  const char* velo_track_hits;
  void* hits_view;
  void* track_view;
  void* tracks_view;
  void* multi_event_tracks_view;
  void* imec_velo_tracks;
};

// The Velo::Clusters view of each listed event: {container, total, 0} with total =
// totals[totals_stride * number_of_events[0]]; with zero_all every one of n_events entries is
// zeroed first.
struct ClustersViewLaunch {
  const uint32_t* event_list;
  unsigned n_list;
  unsigned n_events;
  bool zero_all;
  const uint32_t* number_of_events;
  const uint32_t* totals;
  unsigned totals_stride;
  char* container;
  void* views;
};

hipError_t launch_number_of_candidates(const CandidatesLaunch& a, hipStream_t s);
hipError_t launch_estimate_input_size(const EstimateLaunch& a, hipStream_t s);
hipError_t launch_masked_clustering(const MaskedClusteringLaunch& a, hipStream_t s);
hipError_t launch_sort_by_phi(const SortLaunch& a, hipStream_t s);
hipError_t launch_cluster_sorted(const FusedClusterLaunch& a, hipStream_t s);
hipError_t launch_estimate_mp(const EstimateMpLaunch& a, hipStream_t s);
hipError_t launch_cluster_mp(const ClusterMpLaunch& a, hipStream_t s);
hipError_t launch_search_by_triplet(const SearchLaunch& a, hipStream_t s);
// The largest n_big events of an ordered list on `side` (fork/join with `s` through the two
// events), the others on `s`; a plain launch when there is nothing to split.
hipError_t launch_search_split(
  const SearchLaunch& a,
  unsigned n_big,
  unsigned big_block,
  unsigned big_lds_hits,
  hipStream_t s,
  hipStream_t side,
  hipEvent_t fork,
  hipEvent_t join);
constexpr unsigned event_clock_words = 10; // start, end, ticks in 8 phases
// Event hits staged in LDS by default: four workgroups of 256 lanes fit one CU's 160 KiB.
constexpr unsigned search_default_lds_hits = 1760;
size_t search_dynamic_lds(unsigned block, unsigned lds_hits, unsigned window, unsigned list_cap);
// order[0..n) = events by descending hit count (n read on the device).
hipError_t launch_order_events_by_hits(
  const uint32_t* offsets_estimated_input_size,
  const uint32_t* number_of_events,
  uint32_t* order,
  hipStream_t s);
hipError_t launch_three_hit_tracks_filter(const ThreeHitFilterLaunch& a, hipStream_t s);
hipError_t launch_copy_track_hit_number(const CopyHitNumberLaunch& a, hipStream_t s);
hipError_t launch_consolidate_tracks(const ConsolidateLaunch& a, hipStream_t s);
hipError_t launch_kalman_filter(const KalmanLaunch& a, hipStream_t s);

// Learned seed scorer (10 -> 32 -> 32 -> 1 MLP). Parameter layout (FP32): W1 [10][32], b1 [32],
// W2 [32][32], b2 [32], w3 [32], b3.
namespace mlp {
constexpr unsigned n_inputs = 10, n_hidden = 32;
constexpr unsigned param_w1 = 0, param_b1 = 320, param_w2 = 352, param_b2 = 1376, param_w3 = 1408, param_b3 = 1440;
constexpr unsigned n_params = 1441;
constexpr unsigned rows_per_block = 64; // seeds per workgroup of the training and scoring kernels (two waves)
inline unsigned mlp_blocks(unsigned n) { return (n + rows_per_block - 1) / rows_per_block; }
} // namespace mlp
hipError_t launch_mlp_train(
  const float* features,
  const uint8_t* labels,
  unsigned n,
  float inv_norm,
  const float* params,
  float* partial,
  float* loss_partial,
  hipStream_t s);
hipError_t launch_mlp_reduce_adam(
  const float* partial,
  const float* loss_partial,
  unsigned n_blocks,
  float* params,
  float* m,
  float* v,
  float* grad_out,
  float* loss_out,
  float lr,
  unsigned step,
  hipStream_t s);
hipError_t launch_mlp_score(const float* features, unsigned n, const float* params, float* logits, hipStream_t s);

// Device exclusive scan with a leading zero: out[0] = 0, out[i+1] = sum(in[0..i]).
size_t prefix_sum_scratch_words(unsigned n);
hipError_t launch_prefix_sum(const uint32_t* in, unsigned n, uint32_t* out, uint32_t* scratch, hipStream_t s);
// Same, with the length read on the device: min(*n_dev, n_cap).
hipError_t launch_prefix_sum_dyn(
  const uint32_t* in,
  const uint32_t* n_dev,
  unsigned n_cap,
  uint32_t* out,
  uint32_t* scratch,
  hipStream_t s);
// Same, with the tile sums in a static device array of the library (no allocation): at most
// prefix_sum_static_max_items items. prefix_sum_static_init() looks the current device's copy up.
constexpr unsigned prefix_sum_static_max_items = 1u << 28;
hipError_t prefix_sum_static_init();
hipError_t launch_prefix_sum_static(const uint32_t* in, unsigned n, uint32_t* out, hipStream_t s);
// Fill `n` u32 with `value` (a kernel, so the work can be captured in a graph).
hipError_t launch_create_velo_views(const VeloViewsLaunch& a, hipStream_t s);
hipError_t launch_clusters_views(const ClustersViewLaunch& a, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t value, hipStream_t s);
hipError_t launch_fill_u8(uint8_t* p, size_t n, uint8_t value, hipStream_t s);

// Launch-geometry queries of the current device, computed once per (device, query) and shared by
// every host thread (the lookups are serialised by a mutex).
// Dynamic LDS a workgroup may use (hipDeviceAttributeMaxSharedMemoryPerBlock), 0 on failure.
int device_max_lds_per_block();
// Measured peaks (velo_microbench.hip): kind 0 vector-ALU lane-ops/s, 1 LDS bytes/s, 2 HBM copy bytes/s.
hipError_t run_microbench(int kind, double* result, hipStream_t stream);
// Workgroups of `kernel` with `block` lanes and `dyn_lds` bytes of dynamic LDS that stay resident
// on the whole device (CUs x occupancy), 0 if the runtime cannot tell.
unsigned device_resident_blocks(const void* kernel, unsigned block, size_t dyn_lds);

} // namespace velo
