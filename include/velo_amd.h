/* velo_amd — MI355X (gfx950) VELO reconstruction path behind a C ABI.
 *
 * Each entry point replaces one algorithm of Allen's VELO-SP sequence
 * (configuration/python/AllenConf/velo_reconstruction.py:72-321) and takes a POD whose
 * fields mirror, name for name, that algorithm's `Parameters` struct: HOST_INPUTs become
 * scalars, DEVICE_INPUT/OUTPUTs and MASK_INPUTs become device pointers, PROPERTYs become
 * scalars.  Every entry point enqueues work on the given HIP stream and returns without
 * synchronising.  Entry points never allocate or free argument buffers; the only device
 * memory the library owns lives in a velo_constants (geometry, super-pixel tables,
 * beamline) or a velo_sequence (scratch and the intermediate buffers of a whole run).
 *
 * Data layouts are those of the reference with half_t == float (the CPU-target layout):
 *   raw input      per-event blocks [u32 n_banks][u32 bank_offsets[n_banks+1]][banks],
 *                  bank = [u32 sourceID][(v2/v3: u32 count)][u32 SP words] (Transpose.cpp:430-498)
 *   raw sizes      [u32 offsets_in_u16_units x (n+1)][u16 bank sizes]   (MEPTools.h:35-44)
 *   raw types      [u32 offsets_in_bytes x (n+1)][u8 bank types]        (MEPTools.h:46-55)
 *   MEP layout     (mep_layout != 0, MEPTools.h:59-200) offsets = [u32 n_banks][u32 n_events]
 *                  [u32 source IDs x n_banks][u32 fragment offsets x n_banks, per event]; a
 *                  fragment is a bank without its source ID; sizes and types indexed by bank, then event
 *   clusters       [u32 LHCbID x T][f32 x,y,z x T][i16 phi x T], 18 B each (VeloEventModel.cuh:92-183)
 *   track hits     Velo::TrackHits (u16 hits[26], u8 hitsNum; 54 B), Velo::TrackletHits (3 x u16)
 *   states         [x,y,z,tx,ty,qop x N][c00,c20,c22,c11,c31,c33,chi2,ndof x N] (VeloConsolidated.cuh:386-640)
 *
 * Return codes: VELO_OK, or one of the errors below; a HIP launch error is reported as
 * VELO_ERR_HIP (query it with velo_last_hip_error). */
#ifndef VELO_AMD_H
#define VELO_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VELO_OK = 0,
  VELO_ERR_INVALID_ARGUMENT = 1,
  VELO_ERR_BANK_VERSION = 2,    /* like StrException in EstimateInputSize.cu:265-268 */
  VELO_ERR_CAPACITY = 3,        /* a run exceeds what a velo_sequence was created for */
  VELO_ERR_HIP = 4,
  VELO_ERR_UNSUPPORTED_LAYOUT = 5, /* a compressed MDF record */
  VELO_ERR_IO = 6                  /* a file cannot be read or is not a well-formed MDF file */
};

/* Device status bits written by the kernels when a fixed-capacity structure of the
 * reference overflows (its results are then no longer defined by the reference). */
enum {
  VELO_STATUS_TTF_RING_OVERFLOW = 1u << 0,  /* > 2048 live tracks to follow (VeloDefinitions.cuh:26) */
  VELO_STATUS_TRACK_SLOTS_OVERFLOW = 1u << 1,/* more tracks than 0.34 x clusters slots (VeloEventModel.cuh:192-196) */
  VELO_STATUS_HIT_INDEX_OVERFLOW = 1u << 2,  /* > 65536 clusters in an event (uint16 hit indices) */
  VELO_STATUS_THREE_HIT_SLOTS_OVERFLOW = 1u << 3,
  VELO_STATUS_LDS_FALLBACK = 1u << 8         /* informational: an event took the global-memory path */
};

int velo_last_hip_error(void);
const char* velo_error_string(int code);
/* Tick rate (kHz) of the device wall clock used by the "search_event_clock" buffer. */
int velo_wall_clock_khz(unsigned* khz);
/* Measured peaks of the current device for roofline figures (no counterpart in the reference).
 * kind 0: vector-ALU FMA issue, lane operations per second (16 independent v_fma_f32 chains per
 * lane); 1: LDS read bandwidth, bytes per second (conflict-free 128-bit reads); 2: HBM copy
 * bandwidth, bytes read plus written per second (a 1 GiB device copy). Each times five launches
 * after one untimed launch with HIP events on `stream` and waits for them; it allocates its own
 * buffers (kind 2: two of 1 GiB) and frees them. */
int velo_microbench(int kind, double* result, hipStream_t stream);

/* ---- Constants: replaces Constants::dev_velo_geometry / dev_velo_sp_patterns / sp_fx / sp_fy /
 * dev_beamline filled by integration/non_event_data/src/VPGeometry.cpp:21-63 and Beamline.cpp:17-31. */
typedef struct velo_constants velo_constants;
int velo_constants_create(const char* velo_geometry_bin, size_t nbytes, const float* beamline_xy, velo_constants** out);
int velo_constants_destroy(velo_constants* c);
/* Device pointers held by the constants (for use in the *_params structs below). */
const void* velo_constants_dev_geometry(const velo_constants* c);
const uint8_t* velo_constants_dev_sp_patterns(const velo_constants* c);
const float* velo_constants_dev_sp_fx(const velo_constants* c);
const float* velo_constants_dev_sp_fy(const velo_constants* c);
const float* velo_constants_dev_beamline(const velo_constants* c);
/* cache_sp_patterns (host/velo/clustering/src/ClusteringCommon.cpp:6-104): the cluster table of
 * every 8-pixel super-pixel pattern, computed on the host (no device needed). */
void velo_cache_sp_patterns(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy);
/* Host copy of the super-pixel tables (ClusteringCommon.cpp:6-104): 256 + 256 B, 2 x 512 floats. */
int velo_constants_sp_tables(const velo_constants* c, uint8_t* patterns, uint8_t* sizes, float* fx, float* fy);

/* ---- MDF ingestion (host only; no device needed). Reads the raw banks of one type (63: VELO
 * super pixels, 73: retina clusters; as Allen's provider, main/src/Provider.cpp:240-248, takes one
 * of the two encodings) from an uncompressed MDF file (mdf/include/mdf_header.hpp:45-220,
 * mdf/lhcb/Event/RawBank.h:196-250) into Allen's event layout (main/src/Transpose.cpp:220-498,
 * banks sorted by sourceID): the "raw", "offsets", "sizes" and "types" buffers that
 * velo_sequence_input and the single algorithms take, in host memory owned by the handle.
 * max_events = 0 reads every record. All banks of the type must share one version. */
typedef struct velo_mdf_events velo_mdf_events;
int velo_mdf_read(const char* path, unsigned bank_type, unsigned max_events, velo_mdf_events** out);
int velo_mdf_events_info(const velo_mdf_events* e, unsigned* n_events, int* bank_version, size_t* raw_bytes);
int velo_mdf_events_buffer(const velo_mdf_events* e, const char* name, const void** host_ptr, size_t* bytes);
void velo_mdf_events_free(velo_mdf_events* e);

/* ---- host_prefix_sum_t (host/prefix_sum/src/HostPrefixSum.cpp:76-107), as a device scan:
 * dev_output_buffer[0] = 0, dev_output_buffer[i+1] = sum(dev_input_buffer[0..i]) (u32 wrap).
 * If host_total_sum_holder is not NULL the total is copied there (the one host<->device sync point
 * the reference has here); with NULL nothing waits on the host. */
typedef struct {
  const unsigned* dev_input_buffer;
  unsigned input_size;
  unsigned* dev_output_buffer;     /* input_size + 1 */
  unsigned* dev_scratch;           /* >= velo_prefix_sum_scratch_size(input_size) u32, or NULL for sizes <= 4096 */
  unsigned* host_total_sum_holder; /* optional */
} velo_prefix_sum_params;
size_t velo_prefix_sum_scratch_size(unsigned input_size);
int velo_prefix_sum(const velo_prefix_sum_params* p, hipStream_t stream);

/* ---- The same prefix sum without allocation or workspace. No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). dev_scratch must be NULL: the tile sums live in a static device array of
 * the library (256 KiB per device, held until the process ends), so input_size is at most
 * velo_prefix_sum_static_max_items() (2^28), else VELO_ERR_CAPACITY. Calls on one device share that
 * array: calls that overlap on one device (other streams, other threads) race. */
size_t velo_prefix_sum_static_max_items(void);
int velo_prefix_sum_static(const velo_prefix_sum_params* p, hipStream_t stream);

/* ---- velo_calculate_number_of_candidates_t
 * (device/velo/mask_clustering/include/VeloCalculateNumberOfCandidates.cuh:10-35) */
typedef struct {
  unsigned host_number_of_events;
  int host_raw_bank_version;
  const unsigned* dev_event_list; unsigned event_list_size; /* MASK_INPUT */
  const char* dev_velo_raw_input;
  const unsigned* dev_velo_raw_input_offsets;
  const unsigned* dev_velo_raw_input_sizes;
  const unsigned* dev_velo_raw_input_types;
  unsigned* dev_number_of_candidates;  /* host_number_of_events */
  unsigned* dev_velo_bank_index;       /* 208 */
  unsigned block_dim_x;                /* property, default 256 */
  unsigned event_start;                /* RuntimeOptions::event_interval first */
  int mep_layout;                      /* RuntimeOptions::mep_layout: raw input in the MEP layout */
} velo_calculate_number_of_candidates_params;
int velo_calculate_number_of_candidates(const velo_calculate_number_of_candidates_params* p, hipStream_t stream);

/* ---- velo_estimate_input_size_t (device/velo/mask_clustering/include/EstimateInputSize.cuh:10-36) */
typedef struct {
  unsigned host_number_of_events;
  unsigned host_number_of_cluster_candidates;
  int host_raw_bank_version;
  const unsigned* dev_event_list; unsigned event_list_size;
  const unsigned* dev_candidates_offsets;
  const char* dev_velo_raw_input;
  const unsigned* dev_velo_raw_input_offsets;
  const unsigned* dev_velo_raw_input_sizes;
  const unsigned* dev_velo_raw_input_types;
  unsigned* dev_estimated_input_size;  /* 26 x events */
  unsigned* dev_module_candidate_num;  /* events */
  unsigned* dev_cluster_candidates;    /* host_number_of_cluster_candidates */
  unsigned event_start;
  int mep_layout;
} velo_estimate_input_size_params;
int velo_estimate_input_size(const velo_estimate_input_size_params* p, hipStream_t stream);

/* ---- Allen's device view objects of this path, byte for byte (half_t == float, 64-bit pointers):
 * the objects the clustering kernels (MaskedVeloClustering.cu:519-522, DecodeRetinaClusters.cu:346),
 * create_velo_views (VeloConsolidateTracks.cu:11-46) and velo_kalman_filter (VeloKalmanFilter.cu:370-374)
 * construct in place. The reference's classes have no virtual functions, and their empty CRTP
 * bases take no bytes, so each is its data members in declaration order. Pads are written as 0. */
typedef struct {               /* Velo::Clusters = Clusters_t<char> (VeloEventModel.cuh:91-110), 16 B */
  char* base_pointer;
  unsigned total_number_of_hits;
  unsigned offset;
} velo_clusters_view;
typedef struct {               /* Allen::Views::Velo::Consolidated::Hits (VeloConsolidated.cuh:60-99), 24 B */
  const char* base_pointer;
  unsigned offset;
  unsigned size;
  unsigned total_number_of_hits;
  unsigned pad_;
} velo_hits_view;
typedef struct {               /* ...::Consolidated::Track (VeloConsolidated.cuh:199-259), 24 B */
  const velo_hits_view* hits;  /* the event's Hits view */
  unsigned track_index;
  unsigned track_container_offset;
  unsigned offset;             /* of the track's first hit within the event */
  unsigned number_of_hits;
} velo_track_view;
typedef struct {               /* ...::Consolidated::Tracks (VeloConsolidated.cuh:261-294), 16 B */
  const velo_track_view* track; /* the event's first Track view */
  unsigned size;
  unsigned offset;
} velo_tracks_view;
typedef struct {               /* Allen::MultiEventContainer<Tracks> (MultiEventContainer.cuh:38-65), 24 B */
  int type_id;                 /* Identifiable::m_type_id: VELO_TYPE_ID_VELO_TRACKS */
  unsigned pad_;
  const velo_tracks_view* container;
  unsigned number_of_events;
  unsigned pad2_;
} velo_multi_event_tracks_view;
typedef struct {               /* Allen::Views::Physics::KalmanStates (States.cuh:191-220), 24 B */
  const char* base_pointer;
  unsigned offset;
  unsigned size;
  unsigned total_number_of_tracks;
  unsigned pad_;
} velo_kalman_states_view;
#define VELO_TYPE_ID_VELO_TRACKS 1 /* Allen::TypeIDs::VeloTracks (backend/include/TypeID.h:23-34) */

/* ---- velo_masked_clustering_t (device/velo/mask_clustering/include/MaskedVeloClustering.cuh:14-37) */
typedef struct {
  unsigned host_total_number_of_velo_clusters;
  unsigned host_number_of_events;
  int host_raw_bank_version;
  const char* dev_velo_raw_input;
  const unsigned* dev_velo_raw_input_offsets;
  const unsigned* dev_velo_raw_input_sizes;
  const unsigned* dev_velo_raw_input_types;
  const unsigned* dev_offsets_estimated_input_size;
  const unsigned* dev_module_candidate_num;
  const unsigned* dev_cluster_candidates;
  const unsigned* dev_event_list; unsigned event_list_size;
  const unsigned* dev_candidates_offsets;
  const unsigned* dev_number_of_events;
  const unsigned* dev_velo_bank_index;
  unsigned* dev_module_cluster_num;        /* 26 x events */
  char* dev_velo_cluster_container;        /* 18 x host_total_number_of_velo_clusters bytes */
  const velo_constants* constants;
  unsigned event_start;
  int mep_layout;
  velo_clusters_view* dev_velo_clusters;   /* optional, events: written for the listed events */
} velo_masked_clustering_params;
int velo_masked_clustering(const velo_masked_clustering_params* p, hipStream_t stream);

/* ---- calculate_number_of_retinaclusters_each_sensor_pair_t
 * (device/velo/retinacluster_decoding/include/CalculateNumberOfRetinaClustersPerSensor.cuh:9-21):
 * the number of cluster words of every retina bank (type 73) per slot: one slot per sensor for
 * versions 2/3 (bank count word), per sensor pair for version 4 (bank size / 4); slots without
 * a retina bank get 0. The first event of the list also records each slot's bank number
 * (every bank type, sourceID & 0x1FF); other entries of dev_retina_bank_index are left as they
 * are, as in the reference. */
typedef struct {
  unsigned host_number_of_events;
  int host_raw_bank_version;
  const unsigned* dev_event_list; unsigned event_list_size; /* MASK_INPUT */
  const char* dev_velo_retina_raw_input;
  const unsigned* dev_velo_retina_raw_input_offsets;
  const unsigned* dev_velo_retina_raw_input_sizes;
  const unsigned* dev_velo_retina_raw_input_types;
  unsigned* dev_retina_bank_index;      /* 208 (v2/v3) or 104 (v4) */
  unsigned* dev_each_sensor_pair_size;  /* host_number_of_events x 208 (v2/v3) or 104 (v4) */
  unsigned event_start;
  int mep_layout;
} velo_calculate_number_of_retinaclusters_params;
int velo_calculate_number_of_retinaclusters_each_sensor_pair(
  const velo_calculate_number_of_retinaclusters_params* p,
  hipStream_t stream);
/* Slots per event of the retina path: 208 for bank versions 2/3, 104 for version 4. */
unsigned velo_retina_sensor_slots(int bank_version);

/* ---- decode_retinaclusters_t (device/velo/retinacluster_decoding/include/DecodeRetinaClusters.cuh:13-48):
 * decodes every word, keys it by (phi as u16) << 48 | LHCbID (the largest key for a zero word)
 * and writes the clusters of each module pair sorted by key from dev_offsets_module_pair_cluster
 * on; dev_module_cluster_num counts the non-zero words. Where words share a key the slot holds
 * the last of them in word order. No scratch is needed. */
typedef struct {
  unsigned host_total_number_of_velo_clusters;
  int host_raw_bank_version;
  unsigned host_number_of_events;
  const char* dev_velo_retina_raw_input;
  const unsigned* dev_velo_retina_raw_input_offsets;
  const unsigned* dev_velo_retina_raw_input_sizes;
  const unsigned* dev_velo_retina_raw_input_types;
  const unsigned* dev_offsets_each_sensor_pair_size; /* slots x events + 1 */
  const unsigned* dev_retina_bank_index;
  const unsigned* dev_event_list; unsigned event_list_size;
  const unsigned* dev_number_of_events;
  unsigned* dev_module_cluster_num;            /* 26 x events */
  unsigned* dev_offsets_module_pair_cluster;   /* 26 x events + 1 */
  char* dev_velo_cluster_container;            /* 18 x host_total_number_of_velo_clusters bytes */
  unsigned* dev_hit_permutations;              /* host_total_number_of_velo_clusters */
  int64_t* dev_hit_sorting_key;                /* host_total_number_of_velo_clusters */
  const velo_constants* constants;
  unsigned event_start;
  int mep_layout;
  velo_clusters_view* dev_velo_clusters;   /* optional, events: zeroed, then written for the listed events */
} velo_decode_retinaclusters_params;
int velo_decode_retinaclusters(const velo_decode_retinaclusters_params* p, hipStream_t stream);

/* ---- velo_sort_by_phi_t (device/velo/search_by_triplet/include/SortByPhi.cuh:14-28) */
typedef struct {
  unsigned host_number_of_events;
  unsigned host_total_number_of_velo_clusters;
  const unsigned* dev_event_list; unsigned event_list_size;
  const unsigned* dev_offsets_estimated_input_size;
  const unsigned* dev_module_cluster_num;
  const char* dev_velo_cluster_container;
  const unsigned* dev_number_of_events;
  char* dev_sorted_velo_cluster_container;
  unsigned* dev_hit_permutation;           /* host_total_number_of_velo_clusters */
} velo_sort_by_phi_params;
int velo_sort_by_phi(const velo_sort_by_phi_params* p, hipStream_t stream);

/* ---- velo_search_by_triplet_t (device/velo/search_by_triplet/include/SearchByTriplet.cuh:15-45) */
typedef struct {
  unsigned host_number_of_events;
  unsigned host_total_number_of_velo_clusters;
  const unsigned* dev_event_list; unsigned event_list_size;
  const unsigned* dev_number_of_events;
  const char* dev_sorted_velo_cluster_container;
  const unsigned* dev_offsets_estimated_input_size;
  const unsigned* dev_module_cluster_num;
  void* dev_tracks;                 /* Velo::TrackHits x track_container_size */
  void* dev_tracklets;              /* Velo::TrackletHits x 2048 x events */
  unsigned* dev_tracks_to_follow;   /* 2048 x events */
  void* dev_three_hit_tracks;       /* Velo::TrackletHits x track_container_size */
  uint8_t* dev_hit_used;            /* bool x clusters */
  unsigned* dev_atomics_velo;       /* 3 x events */
  uint16_t* dev_rel_indices;        /* clusters */
  unsigned* dev_number_of_velo_tracks; /* events */
  unsigned* dev_status;             /* optional: VELO_STATUS_* bits are OR-ed in */
  const velo_constants* constants;
  float phi_tolerance;              /* 0.045 */
  float max_scatter;                /* 0.08 */
  unsigned max_skipped_modules;     /* 1 */
  unsigned block_dim_x;             /* 0 = library default (256) */
} velo_search_by_triplet_params;
int velo_search_by_triplet(const velo_search_by_triplet_params* p, hipStream_t stream);
/* track_container_size of SearchByTriplet.cu:18-28 */
unsigned velo_track_container_size(unsigned total_clusters, unsigned number_of_events);

/* ---- velo_three_hit_tracks_filter_t (device/velo/search_by_triplet/include/ThreeHitTracksFilter.cuh:10-31) */
typedef struct {
  unsigned host_number_of_events;
  unsigned host_total_number_of_velo_clusters;
  const unsigned* dev_event_list; unsigned event_list_size;
  const char* dev_sorted_velo_cluster_container;
  const unsigned* dev_offsets_estimated_input_size;
  const void* dev_three_hit_tracks_input;
  const unsigned* dev_atomics_velo;
  const uint8_t* dev_hit_used;
  const unsigned* dev_number_of_events;
  void* dev_three_hit_tracks_output;
  unsigned* dev_number_of_three_hit_tracks_output;
  float max_chi2;                   /* 20 */
} velo_three_hit_tracks_filter_params;
int velo_three_hit_tracks_filter(const velo_three_hit_tracks_filter_params* p, hipStream_t stream);

/* ---- velo_copy_track_hit_number_t (device/velo/consolidate_tracks/include/VeloCopyTrackHitNumber.cuh) */
typedef struct {
  unsigned host_number_of_events;
  unsigned host_number_of_velo_tracks_at_least_four_hits;
  unsigned host_number_of_three_hit_tracks_filtered;
  const void* dev_tracks;
  const unsigned* dev_offsets_velo_tracks;
  const unsigned* dev_offsets_estimated_input_size;
  const unsigned* dev_offsets_number_of_three_hit_tracks_filtered;
  unsigned* dev_velo_track_hit_number;
  unsigned* dev_offsets_all_velo_tracks;   /* events + 1 */
  unsigned* host_number_of_reconstructed_velo_tracks; /* optional host copy (D2H + wait) */
} velo_copy_track_hit_number_params;
int velo_copy_track_hit_number(const velo_copy_track_hit_number_params* p, hipStream_t stream);

/* ---- velo_consolidate_tracks_t (device/velo/consolidate_tracks/include/VeloConsolidateTracks.cuh) */
typedef struct {
  unsigned host_accumulated_number_of_hits_in_velo_tracks;
  unsigned host_number_of_reconstructed_velo_tracks;
  unsigned host_number_of_three_hit_tracks_filtered;
  unsigned host_number_of_events;
  const unsigned* dev_event_list; unsigned event_list_size;
  const unsigned* dev_offsets_all_velo_tracks;
  const void* dev_tracks;
  const unsigned* dev_offsets_velo_track_hit_number;
  const char* dev_sorted_velo_cluster_container;
  const unsigned* dev_offsets_estimated_input_size;
  const void* dev_three_hit_tracks_output;
  const unsigned* dev_offsets_number_of_three_hit_tracks_filtered;
  const unsigned* dev_number_of_events;
  char* dev_velo_track_hits;                 /* 18 x accumulated hits bytes */
  uint8_t* dev_accepted_velo_tracks;         /* tracks; set to 1 */
  /* create_velo_views (VeloConsolidateTracks.cu:11-46), for every one of host_number_of_events
   * events: all five NULL (no views) or all five set. */
  velo_hits_view* dev_velo_hits_view;                          /* events */
  velo_track_view* dev_velo_track_view;                        /* tracks */
  velo_tracks_view* dev_velo_tracks_view;                      /* events */
  velo_multi_event_tracks_view* dev_velo_multi_event_tracks_view; /* 1 */
  const velo_multi_event_tracks_view** dev_imec_velo_tracks;   /* 1: the multi-event view's address */
} velo_consolidate_tracks_params;
int velo_consolidate_tracks(const velo_consolidate_tracks_params* p, hipStream_t stream);

/* ---- velo_kalman_filter_t (device/velo/simplified_kalman_filter/include/VeloKalmanFilter.cuh:160-284) */
typedef struct {
  unsigned host_number_of_reconstructed_velo_tracks;
  unsigned host_number_of_events;
  const unsigned* dev_event_list; unsigned event_list_size;
  const unsigned* dev_number_of_events;
  const unsigned* dev_offsets_all_velo_tracks;
  const unsigned* dev_offsets_velo_track_hit_number; /* with dev_velo_track_hits: the tracks, when */
  const char* dev_velo_track_hits;                   /* dev_velo_tracks_view is NULL */
  uint8_t* dev_is_backward;
  char* dev_velo_kalman_beamline_states;     /* 56 x tracks bytes */
  char* dev_velo_kalman_endvelo_states;      /* 56 x tracks bytes */
  const velo_constants* constants;
  /* 0: the reference's FP32 arithmetic. 1: a filter step whose FP32 covariance update is
   * ill-conditioned (kappa = 1 + 6400 * predcovXX > 100; every track's first step) has its
   * covariance prediction and update computed in FP64 and rounded to FP32 (no reference
   * counterpart; covariance within ~1e-6 of an FP64 fit instead of up to ~50 %). */
  unsigned covariance_fp64_fallback;
  /* The reference's view input: when set, tracks and their hits are read through these views
   * (velo_consolidate_tracks writes them) and the two fields above may be NULL. */
  const velo_tracks_view* dev_velo_tracks_view;             /* events */
  velo_kalman_states_view* dev_velo_kalman_beamline_states_view; /* optional, events: listed ones written */
  velo_kalman_states_view* dev_velo_kalman_endvelo_states_view;  /* optional, events: listed ones written */
} velo_kalman_filter_params;
int velo_kalman_filter(const velo_kalman_filter_params* p, hipStream_t stream);

/* ---- Beamline primary-vertex finder, the consumer of the beamline states: the six algorithms
 * pv_beamline_extrapolate_t, pv_beamline_histo_t, pv_beamline_peak_t,
 * pv_beamline_calculate_denom_t, pv_beamline_multi_fitter_t and pv_beamline_cleanup_t
 * (device/PV/beamlinePV/include/pv_beamline_*.cuh, wired in
 * configuration/python/AllenConf/primary_vertex_reconstruction.py) in one call, one workgroup per
 * event with the z histogram in LDS. Outputs carry the reference's names and layouts: PVTrack is
 * 7 floats (z, x, y, tx, ty, W_00, W_11), PV::Vertex 48 bytes (x, y, z, chi2, int ndof, cov00,
 * cov10, cov11, cov20, cov21, cov22, nTracks), 32 per event; dev_zhisto (3392 floats per event)
 * is written only when not NULL. Capacities the reference leaves unchecked (extrema, clusters,
 * vertices) raise VELO_STATUS_PV_OVERFLOW in dev_status (optional). */
#define VELO_PV_MAX_VERTICES 32
#define VELO_PV_HISTO_BINS 3392
#define VELO_STATUS_PV_OVERFLOW (1u << 16)
typedef struct {
  unsigned host_number_of_events;
  const unsigned* dev_event_list; unsigned event_list_size;
  const unsigned* dev_number_of_events;
  const unsigned* dev_offsets_all_velo_tracks;
  const float* dev_velo_kalman_beamline_states;
  float* dev_pvtracks;                        /* 7 floats per track */
  float* dev_zhisto;                          /* optional: 3392 per event */
  float* dev_zpeaks;                          /* 32 per event */
  unsigned* dev_number_of_zpeaks;
  float* dev_pvtracks_denom;                  /* 1 per track */
  void* dev_multi_fit_vertices;               /* 32 x 48 bytes per event */
  unsigned* dev_number_of_multi_fit_vertices;
  void* dev_multi_final_vertices;             /* 32 x 48 bytes per event */
  unsigned* dev_number_of_multi_final_vertices;
  unsigned* dev_status;                       /* optional */
  const velo_constants* constants;            /* beamline position */
} velo_pv_beamline_params;
int velo_pv_beamline(const velo_pv_beamline_params* p, hipStream_t stream);

/* ---- Whole sequence: the nine algorithms and five scans above, enqueued back to back on one
 * stream with every size read on the device (no host synchronisation inside a run), so the run
 * can be captured in a hipGraph.  Clustering and the phi sort are fused per module pair.
 * Buffers are sized at creation for at most max_events events and max_raw_bytes of raw input. */
typedef struct velo_sequence velo_sequence;
typedef struct {
  const char* dev_velo_raw_input;
  const unsigned* dev_velo_raw_input_offsets;
  const unsigned* dev_velo_raw_input_sizes;
  const unsigned* dev_velo_raw_input_types;
  size_t raw_input_bytes;          /* host-known size of dev_velo_raw_input */
  unsigned number_of_events;
  int bank_version;
  float phi_tolerance, max_scatter, max_chi2;
  unsigned max_skipped_modules;
  int retina;                      /* the banks are retina clusters (type 73): the retina decoding
                                      replaces candidates, estimate, masked clustering and sort */
  int mep_layout;                  /* raw input in the MEP layout (RuntimeOptions::mep_layout) */
} velo_sequence_input;
int velo_sequence_create(const velo_constants* c, unsigned max_events, size_t max_raw_bytes, velo_sequence** out);
int velo_sequence_destroy(velo_sequence* s);
void velo_sequence_default_input(velo_sequence_input* in);
int velo_sequence_run(velo_sequence* s, const velo_sequence_input* in, hipStream_t stream);
/* Device pointer and capacity (bytes) of a named buffer of the last run; names are the
 * reference's argument names without the dev_ prefix (e.g. "offsets_all_velo_tracks"). */
int velo_sequence_buffer(velo_sequence* s, const char* name, void** dev_ptr, size_t* capacity_bytes);
/* Waits for the stream and returns the OR of the VELO_STATUS_* bits of the last run. */
int velo_sequence_status(velo_sequence* s, hipStream_t stream, unsigned* status);
/* Knobs applied to subsequent runs:
 *   "search_block"      workgroup size of search_by_triplet (multiple of 64, <= 1024; 0 = 256)
 *   "search_lds_hits"   hits of an event staged in LDS by search_by_triplet (<= 8192, within the
 *                       device's LDS per workgroup; 0 = 1760); up to 5x that many hits keep phi
 *                       and flags in LDS and coordinates in global memory, larger events run on
 *                       global memory
 *   "search_largest_first"  1 (default): search_by_triplet takes the events in descending order
 *                       of their hit count; 0: in event order. Outputs do not depend on it
 *   "seed_log"          N > 0: search_by_triplet also records up to N seeds of each run, in no
 *                       particular order, into "seed_log_records" (u32 event, module pair of h1,
 *                       h0, h1, h2 as event-relative hit indices), "seed_log_features" (10 x f32,
 *                       the scorer's inputs) and "seed_log_counter" (seeds seen, may exceed N);
 *                       0 (default) frees the log
 *   "clustering_resident_grid"  1: the fused clustering kernels run as a resident grid whose
 *                       waves loop over (event, module pair) items; 0 (default): one wave per item
 *   "search_persistent" 1: search_by_triplet runs as a resident grid whose workgroups take the
 *                       next event from a device counter (work stealing) instead of one
 *                       workgroup per event; 0 (default). Outputs do not depend on it
 *   "fused_clustering"  1: clustering and phi sort fused per module pair (default); 0: the four
 *                       reference algorithms one after the other
 *   "timed_stage"       VELO_STAGE_* to bracket with HIP events on the run's stream (0 = none)
 *   "kalman_fp64_fallback"  1: velo_kalman_filter's covariance_fp64_fallback (default 0)
 *   "consolidate_lds_map"   1 (default): consolidation maps an event's track hits to their
 *                       tracks in LDS (events up to 2047 tracks and 4096 track hits); 0: every
 *                       hit finds its track by a binary search over the hit-number offsets, as
 *                       larger events always do. Outputs do not depend on it
 *   "count_work"        1: search_by_triplet also counts its work into the buffer "search_work"
 *                       (u64 x 4: seeding evaluations, forwarding evaluations, extrapolations,
 *                       phi-search probes) and, per event, the start and end wall-clock
 *                       ticks of its workgroup and the ticks spent in forwarding, h1 compaction,
 *                       seeding step 1 (h0 search, pendulum, barrier wait), seeding steps 2 and 3
 *                       and the rest into "search_event_clock" (u64 x 10 per event) */
int velo_sequence_set_tuning(velo_sequence* s, const char* knob, unsigned value);
enum {
  VELO_STAGE_ALL = 1,              /* the whole run */
  VELO_STAGE_CANDIDATES = 2,       /* velo_calculate_number_of_candidates */
  VELO_STAGE_ESTIMATE = 3,         /* velo_estimate_input_size */
  VELO_STAGE_CLUSTERING = 4,       /* masked clustering + sort by phi */
  VELO_STAGE_SEARCH = 5,           /* velo_search_by_triplet */
  VELO_STAGE_THREE_HIT_FILTER = 6, /* velo_three_hit_tracks_filter */
  VELO_STAGE_CONSOLIDATE = 7,      /* velo_consolidate_tracks */
  VELO_STAGE_KALMAN = 8,           /* velo_kalman_filter */
  VELO_STAGE_PV = 9,               /* the beamline PV finder (knob "pv_finder") */
  VELO_STAGE_COUNT = 9
};
/* Sum of the elapsed milliseconds of the timed stage over the runs since "timed_stage" was
 * last set, and the number of those runs (waits for them). */
int velo_sequence_stage_ms(velo_sequence* s, float* total_ms, unsigned* runs);
/* Captures one velo_sequence_run with these inputs on `stream` (not the null stream) into a
 * HIP graph held by the sequence, replacing any earlier one; velo_sequence_replay launches it.
 * The graph keeps the input pointers, event count and knobs of the capture; every size inside
 * the run is read on the device, and buffers are cleared by kernels, not memset nodes. */
int velo_sequence_capture(velo_sequence* s, const velo_sequence_input* in, hipStream_t stream);
int velo_sequence_replay(velo_sequence* s, hipStream_t stream);
/* NN-pruned seeding: with params (VELO_MLP_PARAMS host floats, copied) search_by_triplet scores
 * every seed with the learned scorer and drops those whose logit is below `threshold`, except
 * that a group of seeds (one seeding pass over up to 64 h1 hits) that would lose all its seeds
 * keeps them; NULL restores the plain search. Needs "search_block" 256. With "seed_log" on, the
 * log also holds each seed's logit ("seed_log_logits"). */
int velo_sequence_set_scorer(velo_sequence* s, const float* params, float threshold);

/* ---- Learned seed scorer (no reference counterpart): a 10 -> 32 -> 32 -> 1 MLP with ReLU on
 * BF16 MFMA with FP32 accumulation, over the 10 features of a seed (velo_seed_features).
 * Parameters: VELO_MLP_PARAMS FP32 values, W1 [10][32], b1 [32], W2 [32][32], b2 [32], w3 [32], b3;
 * they are rounded to bf16 where they enter a matrix product. */
#define VELO_MLP_PARAMS 1441
#define VELO_SEED_FEATURES 10
/* Scratch (floats) for velo_mlp_train_step over n rows. */
size_t velo_mlp_scratch_floats(unsigned n);
/* One step of full-batch gradient descent over n rows: mean binary cross-entropy of the logits
 * against labels (0/1); Adam (beta 0.9, 0.999, eps 1e-8) with learning rate lr at 1-based step
 * `step` updates dev_params, dev_adam_m, dev_adam_v; lr = 0 computes the gradient only.
 * Optional outputs: dev_grad (VELO_MLP_PARAMS floats), dev_loss (sum of the rows' losses). */
int velo_mlp_train_step(
  const float* dev_features,
  const uint8_t* dev_labels,
  unsigned n,
  float* dev_params,
  float* dev_adam_m,
  float* dev_adam_v,
  float* dev_scratch,
  float lr,
  unsigned step,
  float* dev_grad,
  float* dev_loss,
  hipStream_t stream);
/* Logits of n rows of features. */
int velo_mlp_score(const float* dev_features, unsigned n, const float* dev_params, float* dev_logits, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
