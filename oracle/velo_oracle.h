/* TEST INFRASTRUCTURE — C ABI of the CPU oracle of the VELO reconstruction path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (allen_amd/, include/velo_amd.h) never links or calls it. */
#ifndef VELO_ORACLE_H
#define VELO_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Allen event layout (main/src/Transpose.cpp:430-498) plus geometry and properties. */
typedef struct {
  const char* raw;            /* per-event blocks: [u32 n_banks][u32 offsets[n+1]][banks] */
  const uint32_t* offsets;    /* n_events+1 byte offsets of the event blocks */
  const uint32_t* sizes;      /* [u32 offs in u16 units x (n+1)][u16 bank sizes] */
  const uint32_t* types;      /* [u32 offs in bytes x (n+1)][u8 bank types] */
  unsigned n_events;
  unsigned event_start;
  int bank_version;           /* 2, 3 or 4 */
  const char* geometry;       /* velo_geometry.bin contents */
  size_t geometry_bytes;
  const float* beamline;      /* beamline.bin contents: x, y */
  float phi_tolerance;        /* velo_search_by_triplet properties */
  float max_scatter;
  unsigned max_skipped_modules;
  float max_chi2;             /* velo_three_hit_tracks_filter property */
  unsigned block_dim_x;       /* velo_calculate_number_of_candidates block size (decides grid) */
  int compute_f64;            /* also run the FP64 Kalman restatement */
  int kalman_fp64_fallback;   /* FP32 Kalman with the FP64 covariance update on ill-conditioned steps */
  int collect_seeds;          /* also return every seed: "seed_log" (u32 event, module pair, h0, h1, h2)
                                 and "seed_features" (f32 x 10) */
  int retina;                 /* the banks are retina clusters (type 73): decode_retinaclusters replaces
                                 the candidates, estimate, masked clustering and phi sort */
  int pv_finder;              /* also run the beamline PV finder on the beamline states */
  const uint32_t* event_list; /* MASK_INPUT dev_event_list (event numbers < n_events), or NULL for all */
  unsigned event_list_size;
  int vectorized_seeding;     /* seeding as the CPU target dispatches it (track_seeding_vectorized) */
  /* Optional: start at search_by_triplet from these sorted clusters (Velo::Clusters layout over
   * clusters_offsets[26 n_events] slots) instead of decoding raw banks. */
  const char* sorted_clusters;
  const uint32_t* clusters_offsets; /* offsets_estimated_input_size, 26 n_events + 1 */
  const uint32_t* clusters_counts;  /* module_cluster_num, 26 n_events */
  /* Optional seed pruning by a learned scorer (no reference counterpart): one logit per cluster
   * slot (offsets_estimated_input_size indexing: the logit of the seed whose h1 is that hit, NaN
   * where none), kept when >= seed_threshold, decided in groups of seed_block / K consecutive
   * unused h1 hits of each seed_block positions of a module pair (K = h0 candidates considered);
   * a group with no kept seed keeps all of its seeds. Scalar seeding only. */
  const float* seed_logits;
  float seed_threshold;
  unsigned seed_block;      /* 0: 256 */
  int mep_layout;           /* raw, offsets, sizes and types are in the MEP layout (MEPTools.h:59-200) */
} vo_inputs;

void vo_default_inputs(vo_inputs* a);
/* Runs the whole sequence; returns a handle holding every intermediate buffer, or NULL. */
void* vo_run(const vo_inputs* a);
/* Buffer names are the reference's argument names without the dev_/host_ prefix. */
int vo_get(void* handle, const char* name, const void** ptr, size_t* nbytes);
void vo_free(void* handle);
/* One event per task on n_threads std::threads; returns events/s. */
double vo_throughput(const vo_inputs* a, unsigned n_threads, unsigned repetitions, double* seconds);

void vo_cache_sp_patterns(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy);
float vo_fast_atan2f(float y, float x);
int16_t vo_hit_phi_16(float x, float y);
int16_t vo_hit_phi_float_to_16(float phi);
uint32_t vo_prefix_sum(const uint32_t* input, size_t n, uint32_t* output);
int vo_binary_search_leftmost_i16(const int16_t* array, unsigned n, int16_t value);
uint64_t vo_make_cluster(uint64_t pixel_map, uint64_t start, unsigned col_lower, unsigned row_lower, uint32_t* xyn);
float vo_three_hit_chi2(const float* xyz9);
int vo_load_geometry(const char* bytes, size_t n, float* module_zs, float* local_x, float* x_pitch, float* ltg);

#ifdef __cplusplus
}
#endif
#endif
