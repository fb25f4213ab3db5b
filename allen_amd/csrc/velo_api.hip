// C ABI of the MI355X VELO path (declarations and layouts in include/velo_amd.h).
#include "velo_amd.h"
#include "velo_common.cuh"
#include "velo_kernels.h"

#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local int g_last_hip_error = 0;

int hip_status(hipError_t e)
{
  if (e == hipSuccess) return VELO_OK;
  g_last_hip_error = static_cast<int>(e);
  return VELO_ERR_HIP;
}

// Super-pixel cluster table: for each 8-bit pattern (bit = 4*col + row of a 2x4 SP) the
// 8-connected clusters in order of their lowest pixel, each as centroid pixel (col << 2 | row)
// plus the fractional centroid (the contract of ClusteringCommon.cpp:6-104).
void build_sp_tables(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy)
{
  std::memset(patterns, 0, 256);
  std::memset(sizes, 0, 256);
  std::memset(fx, 0, 512 * sizeof(float));
  std::memset(fy, 0, 512 * sizeof(float));
  for (unsigned sp = 0; sp < 256; ++sp) {
    unsigned label[8];
    for (unsigned b = 0; b < 8; ++b) label[b] = (sp >> b) & 1u ? b : 8u;
    // Propagate the smallest pixel index through 8-neighbours until stable.
    bool changed = true;
    while (changed) {
      changed = false;
      for (unsigned a = 0; a < 8; ++a) {
        if (label[a] == 8u) continue;
        for (unsigned b = 0; b < 8; ++b) {
          if (label[b] == 8u || a == b) continue;
          const int dc = static_cast<int>(a / 4) - static_cast<int>(b / 4);
          const int dr = static_cast<int>(a % 4) - static_cast<int>(b % 4);
          if (dc >= -1 && dc <= 1 && dr >= -1 && dr <= 1 && label[b] < label[a]) {
            label[a] = label[b];
            changed = true;
          }
        }
      }
    }
    unsigned k = 0;
    for (unsigned root = 0; root < 8 && k < 2; ++root) {
      unsigned n = 0, sx = 0, sy = 0;
      for (unsigned b = 0; b < 8; ++b)
        if (label[b] == root) {
          ++n;
          sx += b / 4;
          sy += b % 4;
        }
      if (!n) continue;
      const unsigned cx = sx / n, cy = sy / n;
      patterns[sp] |= static_cast<uint8_t>(((cx << 2) | cy) << (4 * k));
      if (k == 1) patterns[sp] |= 8;
      sizes[sp] |= static_cast<uint8_t>(n << (4 * k));
      fx[2 * sp + k] = static_cast<float>(sx) / static_cast<float>(n) - static_cast<float>(cx);
      fy[2 * sp + k] = static_cast<float>(sy) / static_cast<float>(n) - static_cast<float>(cy);
      ++k;
    }
  }
}

bool parse_geometry(const char* p0, size_t n, velo::Geometry& g)
{
  // [u64 52][f32 x52][u64 768][f32 x768][u64 768][f32 x768][u64 208][u64 12][f32 x208x12]
  const char* p = p0;
  const char* end = p0 + n;
  auto arr = [&](uint64_t expected, float* d) {
    if (end - p < 8) return false;
    uint64_t m;
    std::memcpy(&m, p, 8);
    p += 8;
    if (m != expected || static_cast<uint64_t>(end - p) < 4 * m) return false;
    std::memcpy(d, p, 4 * m);
    p += 4 * m;
    return true;
  };
  if (!arr(velo::n_modules, g.module_zs) || !arr(velo::n_sensor_columns, g.local_x) ||
      !arr(velo::n_sensor_columns, g.x_pitch))
    return false;
  if (end - p < 16) return false;
  uint64_t n_ltg, n_trans;
  std::memcpy(&n_ltg, p, 8);
  std::memcpy(&n_trans, p + 8, 8);
  p += 16;
  if (n_ltg != velo::n_sensors || n_trans != 12 || static_cast<uint64_t>(end - p) != 4 * n_ltg * n_trans) return false;
  std::memcpy(g.ltg, p, 4 * n_ltg * n_trans);
  return true;
}

bool version_ok(int v) { return v == 2 || v == 3 || v == 4; }

} // namespace

struct velo_constants {
  velo::Geometry* dev_geometry = nullptr;
  uint8_t* dev_sp_patterns = nullptr;
  float* dev_sp_fx = nullptr;
  float* dev_sp_fy = nullptr;
  float* dev_beamline = nullptr;
  uint8_t patterns[256], sizes[256];
  float fx[512], fy[512];
};

extern "C" {

int velo_last_hip_error(void) { return g_last_hip_error; }

int velo_wall_clock_khz(unsigned* khz)
{
  if (!khz) return VELO_ERR_INVALID_ARGUMENT;
  int dev = 0, rate = 0;
  int rc = hip_status(hipGetDevice(&dev));
  if (rc == VELO_OK) rc = hip_status(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev));
  *khz = rc == VELO_OK ? static_cast<unsigned>(rate) : 0u;
  return rc;
}

int velo_microbench(int kind, double* result, hipStream_t stream)
{
  if (!result || kind < 0 || kind > 2) return VELO_ERR_INVALID_ARGUMENT;
  return hip_status(velo::run_microbench(kind, result, stream));
}

const char* velo_error_string(int code)
{
  switch (code) {
  case VELO_OK: return "ok";
  case VELO_ERR_INVALID_ARGUMENT: return "invalid argument";
  case VELO_ERR_BANK_VERSION: return "Velo SP bank version not supported";
  case VELO_ERR_CAPACITY: return "input exceeds the capacity of the sequence";
  case VELO_ERR_HIP: return "HIP error";
  case VELO_ERR_UNSUPPORTED_LAYOUT: return "compressed MDF record not supported";
  case VELO_ERR_IO: return "file cannot be read or is not a well-formed MDF file";
  default: return "unknown error";
  }
}

int velo_constants_create(const char* geometry_bin, size_t nbytes, const float* beamline_xy, velo_constants** out)
{
  if (!geometry_bin || !beamline_xy || !out) return VELO_ERR_INVALID_ARGUMENT;
  velo::Geometry g;
  if (!parse_geometry(geometry_bin, nbytes, g)) return VELO_ERR_INVALID_ARGUMENT;
  auto* c = new velo_constants;
  build_sp_tables(c->patterns, c->sizes, c->fx, c->fy);
  int rc = VELO_OK;
  auto alloc = [&](void** p, size_t b) {
    if (rc == VELO_OK) rc = hip_status(hipMalloc(p, b));
  };
  alloc(reinterpret_cast<void**>(&c->dev_geometry), sizeof(velo::Geometry));
  alloc(reinterpret_cast<void**>(&c->dev_sp_patterns), 256);
  alloc(reinterpret_cast<void**>(&c->dev_sp_fx), 512 * sizeof(float));
  alloc(reinterpret_cast<void**>(&c->dev_sp_fy), 512 * sizeof(float));
  alloc(reinterpret_cast<void**>(&c->dev_beamline), 2 * sizeof(float));
  if (rc == VELO_OK) rc = hip_status(hipMemcpy(c->dev_geometry, &g, sizeof(g), hipMemcpyHostToDevice));
  if (rc == VELO_OK) rc = hip_status(hipMemcpy(c->dev_sp_patterns, c->patterns, 256, hipMemcpyHostToDevice));
  if (rc == VELO_OK) rc = hip_status(hipMemcpy(c->dev_sp_fx, c->fx, 512 * sizeof(float), hipMemcpyHostToDevice));
  if (rc == VELO_OK) rc = hip_status(hipMemcpy(c->dev_sp_fy, c->fy, 512 * sizeof(float), hipMemcpyHostToDevice));
  if (rc == VELO_OK)
    rc = hip_status(hipMemcpy(c->dev_beamline, beamline_xy, 2 * sizeof(float), hipMemcpyHostToDevice));
  // The device address of the allocation-free prefix sum's static scratch on this device.
  if (rc == VELO_OK) rc = hip_status(velo::prefix_sum_static_init());
  if (rc != VELO_OK) {
    velo_constants_destroy(c);
    return rc;
  }
  *out = c;
  return VELO_OK;
}

int velo_constants_destroy(velo_constants* c)
{
  if (!c) return VELO_ERR_INVALID_ARGUMENT;
  (void) hipFree(c->dev_geometry);
  (void) hipFree(c->dev_sp_patterns);
  (void) hipFree(c->dev_sp_fx);
  (void) hipFree(c->dev_sp_fy);
  (void) hipFree(c->dev_beamline);
  delete c;
  return VELO_OK;
}

const void* velo_constants_dev_geometry(const velo_constants* c) { return c ? c->dev_geometry : nullptr; }
const uint8_t* velo_constants_dev_sp_patterns(const velo_constants* c) { return c ? c->dev_sp_patterns : nullptr; }
const float* velo_constants_dev_sp_fx(const velo_constants* c) { return c ? c->dev_sp_fx : nullptr; }
const float* velo_constants_dev_sp_fy(const velo_constants* c) { return c ? c->dev_sp_fy : nullptr; }
const float* velo_constants_dev_beamline(const velo_constants* c) { return c ? c->dev_beamline : nullptr; }

int velo_constants_sp_tables(const velo_constants* c, uint8_t* patterns, uint8_t* sizes, float* fx, float* fy)
{
  if (!c) return VELO_ERR_INVALID_ARGUMENT;
  if (patterns) std::memcpy(patterns, c->patterns, 256);
  if (sizes) std::memcpy(sizes, c->sizes, 256);
  if (fx) std::memcpy(fx, c->fx, sizeof(c->fx));
  if (fy) std::memcpy(fy, c->fy, sizeof(c->fy));
  return VELO_OK;
}

void velo_cache_sp_patterns(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy)
{
  build_sp_tables(patterns, sizes, fx, fy);
}

size_t velo_prefix_sum_scratch_size(unsigned input_size) { return velo::prefix_sum_scratch_words(input_size); }

int velo_prefix_sum(const velo_prefix_sum_params* p, hipStream_t stream)
{
  if (!p || !p->dev_output_buffer || (p->input_size && !p->dev_input_buffer)) return VELO_ERR_INVALID_ARGUMENT;
  if (velo::prefix_sum_scratch_words(p->input_size) > 1 && !p->dev_scratch) return VELO_ERR_INVALID_ARGUMENT;
  int rc = hip_status(
    velo::launch_prefix_sum(p->dev_input_buffer, p->input_size, p->dev_output_buffer, p->dev_scratch, stream));
  if (rc == VELO_OK && p->host_total_sum_holder) {
    rc = hip_status(hipMemcpyAsync(
      p->host_total_sum_holder, p->dev_output_buffer + p->input_size, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    if (rc == VELO_OK) rc = hip_status(hipStreamSynchronize(stream));
  }
  return rc;
}

size_t velo_prefix_sum_static_max_items(void) { return velo::prefix_sum_static_max_items; }

int velo_prefix_sum_static(const velo_prefix_sum_params* p, hipStream_t stream)
{
  if (!p || !p->dev_output_buffer || (p->input_size && !p->dev_input_buffer) || p->dev_scratch)
    return VELO_ERR_INVALID_ARGUMENT;
  if (p->input_size > velo::prefix_sum_static_max_items) return VELO_ERR_CAPACITY;
  int rc = hip_status(velo::launch_prefix_sum_static(p->dev_input_buffer, p->input_size, p->dev_output_buffer, stream));
  if (rc == VELO_OK && p->host_total_sum_holder) {
    rc = hip_status(hipMemcpyAsync(
      p->host_total_sum_holder, p->dev_output_buffer + p->input_size, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    if (rc == VELO_OK) rc = hip_status(hipStreamSynchronize(stream));
  }
  return rc;
}

int velo_calculate_number_of_candidates(const velo_calculate_number_of_candidates_params* p, hipStream_t stream)
{
  if (!p || !p->dev_number_of_candidates || !p->dev_velo_bank_index) return VELO_ERR_INVALID_ARGUMENT;
  if (p->host_raw_bank_version < 0) return VELO_OK; // no VP banks in the data
  if (!version_ok(p->host_raw_bank_version)) return VELO_ERR_BANK_VERSION;
  velo::CandidatesLaunch a;
  a.version = p->host_raw_bank_version;
  a.raw = p->dev_velo_raw_input;
  a.offsets = p->dev_velo_raw_input_offsets;
  a.sizes = p->dev_velo_raw_input_sizes;
  a.types = p->dev_velo_raw_input_types;
  a.mep = p->mep_layout != 0;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.event_start = p->event_start;
  a.block_dim_x = p->block_dim_x ? p->block_dim_x : 256;
  a.number_of_candidates = p->dev_number_of_candidates;
  a.bank_index = p->dev_velo_bank_index;
  int rc = hip_status(velo::launch_fill_u32(a.number_of_candidates, p->host_number_of_events, 0, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_number_of_candidates(a, stream));
  return rc;
}

unsigned velo_retina_sensor_slots(int bank_version) { return velo::retina_sensor_slots(bank_version); }

int velo_calculate_number_of_retinaclusters_each_sensor_pair(
  const velo_calculate_number_of_retinaclusters_params* p,
  hipStream_t stream)
{
  if (!p || !p->dev_each_sensor_pair_size || !p->dev_retina_bank_index) return VELO_ERR_INVALID_ARGUMENT;
  const unsigned slots = velo::retina_sensor_slots(p->host_raw_bank_version);
  int rc = hip_status(
    velo::launch_fill_u32(p->dev_each_sensor_pair_size, (size_t) p->host_number_of_events * slots, 0, stream));
  if (rc != VELO_OK) return rc;
  if (p->host_raw_bank_version < 0) return hip_status(velo::launch_fill_u32(p->dev_retina_bank_index, slots, 0, stream));
  if (!version_ok(p->host_raw_bank_version)) return VELO_ERR_BANK_VERSION;
  velo::RetinaSizesLaunch a;
  a.version = p->host_raw_bank_version;
  a.raw = p->dev_velo_retina_raw_input;
  a.offsets = p->dev_velo_retina_raw_input_offsets;
  a.sizes = p->dev_velo_retina_raw_input_sizes;
  a.types = p->dev_velo_retina_raw_input_types;
  a.mep = p->mep_layout != 0;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.event_start = p->event_start;
  a.each_sensor_pair_size = p->dev_each_sensor_pair_size;
  a.retina_bank_index = p->dev_retina_bank_index;
  return hip_status(velo::launch_retina_sensor_sizes(a, stream));
}

int velo_decode_retinaclusters(const velo_decode_retinaclusters_params* p, hipStream_t stream)
{
  if (!p || !p->constants || !p->dev_module_cluster_num || !p->dev_offsets_module_pair_cluster ||
      !p->dev_velo_cluster_container || !p->dev_hit_permutations || !p->dev_hit_sorting_key ||
      !p->dev_number_of_events)
    return VELO_ERR_INVALID_ARGUMENT;
  const size_t items = (size_t) p->host_number_of_events * velo::n_module_pairs;
  const size_t T = p->host_total_number_of_velo_clusters;
  // The reference's memsets (DecodeRetinaClusters.cu:428-447).
  int rc = hip_status(velo::launch_fill_u32(p->dev_module_cluster_num, items, 0, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_fill_u32(p->dev_offsets_module_pair_cluster, items + 1, 0, stream));
  if (rc == VELO_OK && p->host_raw_bank_version < 0) {
    rc = hip_status(velo::launch_fill_u8(reinterpret_cast<uint8_t*>(p->dev_velo_cluster_container), 18 * T, 0, stream));
    if (rc == VELO_OK)
      rc = hip_status(velo::launch_fill_u32(reinterpret_cast<uint32_t*>(p->dev_hit_sorting_key), 2 * T, 0, stream));
    return rc;
  }
  if (rc != VELO_OK) return rc;
  if (!version_ok(p->host_raw_bank_version)) return VELO_ERR_BANK_VERSION;
  if (items > 0xFFFFFFFFull / 2) return VELO_ERR_INVALID_ARGUMENT;
  rc = hip_status(velo::launch_fill_u32(p->dev_hit_permutations, T, 0, stream));
  if (rc != VELO_OK) return rc;
  velo::RetinaDecodeLaunch a;
  a.version = p->host_raw_bank_version;
  a.raw = p->dev_velo_retina_raw_input;
  a.offsets = p->dev_velo_retina_raw_input_offsets;
  a.sizes = p->dev_velo_retina_raw_input_sizes;
  a.types = p->dev_velo_retina_raw_input_types;
  a.mep = p->mep_layout != 0;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.event_start = p->event_start;
  a.geometry = p->constants->dev_geometry;
  a.offsets_each_sensor_pair_size = p->dev_offsets_each_sensor_pair_size;
  a.retina_bank_index = p->dev_retina_bank_index;
  a.number_of_events = p->dev_number_of_events;
  a.module_cluster_num = p->dev_module_cluster_num;
  a.offsets_module_pair_cluster = p->dev_offsets_module_pair_cluster;
  a.cluster_container = p->dev_velo_cluster_container;
  a.hit_permutations = p->dev_hit_permutations;
  a.hit_sorting_key = p->dev_hit_sorting_key;
  a.populate_items = static_cast<unsigned>(items);
  rc = hip_status(velo::launch_retina_decode(a, stream));
  if (rc == VELO_OK && p->dev_velo_clusters) {
    // DecodeRetinaClusters.cu:331-346,442: zeroed, then {container, total of the sensor-pair
    // offsets, 0} per listed event.
    velo::ClustersViewLaunch v;
    v.event_list = p->dev_event_list;
    v.n_list = p->event_list_size;
    v.n_events = p->host_number_of_events;
    v.zero_all = true;
    v.number_of_events = p->dev_number_of_events;
    v.totals = p->dev_offsets_each_sensor_pair_size;
    v.totals_stride = velo::retina_sensor_slots(p->host_raw_bank_version);
    v.container = p->dev_velo_cluster_container;
    v.views = p->dev_velo_clusters;
    rc = hip_status(velo::launch_clusters_views(v, stream));
  }
  return rc;
}

int velo_pv_beamline(const velo_pv_beamline_params* p, hipStream_t stream)
{
  if (!p || !p->constants || !p->dev_offsets_all_velo_tracks || !p->dev_velo_kalman_beamline_states ||
      !p->dev_number_of_events || !p->dev_pvtracks || !p->dev_zpeaks || !p->dev_number_of_zpeaks ||
      !p->dev_pvtracks_denom || !p->dev_multi_fit_vertices || !p->dev_number_of_multi_fit_vertices ||
      !p->dev_multi_final_vertices || !p->dev_number_of_multi_final_vertices)
    return VELO_ERR_INVALID_ARGUMENT;
  velo::PVLaunch a;
  a.event_list = p->dev_event_list;
  a.n_list = p->dev_event_list ? p->event_list_size : p->host_number_of_events;
  a.number_of_events = p->dev_number_of_events;
  a.offsets_all_velo_tracks = p->dev_offsets_all_velo_tracks;
  a.beamline_states = p->dev_velo_kalman_beamline_states;
  a.beamline = p->constants->dev_beamline;
  a.pvtracks = reinterpret_cast<velo::PVTrackOut*>(p->dev_pvtracks);
  a.zhisto = p->dev_zhisto;
  a.zpeaks = p->dev_zpeaks;
  a.number_of_zpeaks = p->dev_number_of_zpeaks;
  a.pvtracks_denom = p->dev_pvtracks_denom;
  a.multi_fit_vertices = static_cast<velo::PVVertexOut*>(p->dev_multi_fit_vertices);
  a.number_of_multi_fit_vertices = p->dev_number_of_multi_fit_vertices;
  a.multi_final_vertices = static_cast<velo::PVVertexOut*>(p->dev_multi_final_vertices);
  a.number_of_multi_final_vertices = p->dev_number_of_multi_final_vertices;
  a.status = p->dev_status;
  return hip_status(velo::launch_pv_beamline(a, stream));
}

int velo_estimate_input_size(const velo_estimate_input_size_params* p, hipStream_t stream)
{
  if (!p || !p->dev_estimated_input_size || !p->dev_module_candidate_num) return VELO_ERR_INVALID_ARGUMENT;
  int rc = hip_status(velo::launch_fill_u32(
    p->dev_estimated_input_size, (size_t) p->host_number_of_events * velo::n_module_pairs, 0, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_fill_u32(p->dev_module_candidate_num, p->host_number_of_events, 0, stream));
  if (rc != VELO_OK || p->host_raw_bank_version < 0) return rc;
  if (!version_ok(p->host_raw_bank_version)) return VELO_ERR_BANK_VERSION;
  velo::EstimateLaunch a;
  a.version = p->host_raw_bank_version;
  a.raw = p->dev_velo_raw_input;
  a.offsets = p->dev_velo_raw_input_offsets;
  a.sizes = p->dev_velo_raw_input_sizes;
  a.types = p->dev_velo_raw_input_types;
  a.mep = p->mep_layout != 0;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.event_start = p->event_start;
  a.candidates_offsets = p->dev_candidates_offsets;
  a.estimated_input_size = p->dev_estimated_input_size;
  a.module_candidate_num = p->dev_module_candidate_num;
  a.cluster_candidates = p->dev_cluster_candidates;
  return hip_status(velo::launch_estimate_input_size(a, stream));
}

int velo_masked_clustering(const velo_masked_clustering_params* p, hipStream_t stream)
{
  if (!p || !p->constants || !p->dev_module_cluster_num || !p->dev_velo_cluster_container)
    return VELO_ERR_INVALID_ARGUMENT;
  if (p->dev_velo_clusters && (!p->dev_number_of_events || !p->dev_offsets_estimated_input_size))
    return VELO_ERR_INVALID_ARGUMENT;
  int rc = hip_status(velo::launch_fill_u32(
    p->dev_module_cluster_num, (size_t) p->host_number_of_events * velo::n_module_pairs, 0, stream));
  if (rc != VELO_OK || p->host_raw_bank_version < 0) return rc;
  if (!version_ok(p->host_raw_bank_version)) return VELO_ERR_BANK_VERSION;
  velo::MaskedClusteringLaunch a;
  a.version = p->host_raw_bank_version;
  a.raw = p->dev_velo_raw_input;
  a.offsets = p->dev_velo_raw_input_offsets;
  a.sizes = p->dev_velo_raw_input_sizes;
  a.types = p->dev_velo_raw_input_types;
  a.mep = p->mep_layout != 0;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.event_start = p->event_start;
  a.number_of_events = p->dev_number_of_events;
  a.offsets_estimated_input_size = p->dev_offsets_estimated_input_size;
  a.module_candidate_num = p->dev_module_candidate_num;
  a.cluster_candidates = p->dev_cluster_candidates;
  a.candidates_offsets = p->dev_candidates_offsets;
  a.bank_index = p->dev_velo_bank_index;
  a.geometry = p->constants->dev_geometry;
  a.sp_patterns = p->constants->dev_sp_patterns;
  a.sp_fx = p->constants->dev_sp_fx;
  a.sp_fy = p->constants->dev_sp_fy;
  a.module_cluster_num = p->dev_module_cluster_num;
  a.cluster_container = p->dev_velo_cluster_container;
  rc = hip_status(velo::launch_masked_clustering(a, stream));
  if (rc == VELO_OK && p->dev_velo_clusters) {
    // MaskedVeloClustering.cu:519-522: {container, estimated total, 0} per listed event.
    velo::ClustersViewLaunch v;
    v.event_list = p->dev_event_list;
    v.n_list = p->event_list_size;
    v.n_events = p->host_number_of_events;
    v.zero_all = false;
    v.number_of_events = p->dev_number_of_events;
    v.totals = p->dev_offsets_estimated_input_size;
    v.totals_stride = velo::n_module_pairs;
    v.container = p->dev_velo_cluster_container;
    v.views = p->dev_velo_clusters;
    rc = hip_status(velo::launch_clusters_views(v, stream));
  }
  return rc;
}

int velo_sort_by_phi(const velo_sort_by_phi_params* p, hipStream_t stream)
{
  if (!p || !p->dev_sorted_velo_cluster_container || !p->dev_hit_permutation) return VELO_ERR_INVALID_ARGUMENT;
  velo::SortLaunch a;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.number_of_events = p->dev_number_of_events;
  a.offsets_estimated_input_size = p->dev_offsets_estimated_input_size;
  a.module_cluster_num = p->dev_module_cluster_num;
  a.cluster_container = p->dev_velo_cluster_container;
  a.sorted_container = p->dev_sorted_velo_cluster_container;
  a.hit_permutation = p->dev_hit_permutation;
  int rc = hip_status(velo::launch_fill_u32(p->dev_hit_permutation, p->host_total_number_of_velo_clusters, 0, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_sort_by_phi(a, stream));
  return rc;
}

unsigned velo_track_container_size(unsigned total_clusters, unsigned number_of_events)
{
  return velo::track_container_size(total_clusters, number_of_events);
}

int velo_search_by_triplet(const velo_search_by_triplet_params* p, hipStream_t stream)
{
  if (!p || !p->constants || !p->dev_tracks || !p->dev_tracklets || !p->dev_tracks_to_follow ||
      !p->dev_three_hit_tracks || !p->dev_hit_used || !p->dev_atomics_velo || !p->dev_rel_indices ||
      !p->dev_number_of_velo_tracks)
    return VELO_ERR_INVALID_ARGUMENT;
  if (p->block_dim_x && (p->block_dim_x % 64 || p->block_dim_x > 1024)) return VELO_ERR_INVALID_ARGUMENT;
  velo::SearchLaunch a;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.number_of_events = p->dev_number_of_events;
  a.sorted_container = p->dev_sorted_velo_cluster_container;
  a.offsets_estimated_input_size = p->dev_offsets_estimated_input_size;
  a.module_cluster_num = p->dev_module_cluster_num;
  a.geometry = p->constants->dev_geometry;
  a.tracks = p->dev_tracks;
  a.tracklets = p->dev_tracklets;
  a.tracks_to_follow = p->dev_tracks_to_follow;
  a.three_hit_tracks = p->dev_three_hit_tracks;
  a.hit_used = p->dev_hit_used;
  a.atomics_velo = p->dev_atomics_velo;
  a.rel_indices = p->dev_rel_indices;
  a.number_of_velo_tracks = p->dev_number_of_velo_tracks;
  a.status = p->dev_status;
  a.phi_tolerance = p->phi_tolerance;
  a.max_scatter = p->max_scatter;
  a.max_skipped_modules = p->max_skipped_modules;
  a.block_dim = p->block_dim_x;
  a.lds_hits = 0;
  a.work_stats = nullptr;
  // The reference's memsets (SearchByTriplet.cu:47-49): events left out of dev_event_list keep
  // zero track counts and atomics, and no hit is marked used.
  int rc = hip_status(velo::launch_fill_u32(
    p->dev_atomics_velo, (size_t) p->host_number_of_events * velo::num_atomics, 0, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_fill_u8(p->dev_hit_used, p->host_total_number_of_velo_clusters, 0, stream));
  if (rc == VELO_OK)
    rc = hip_status(velo::launch_fill_u32(p->dev_number_of_velo_tracks, p->host_number_of_events, 0, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_search_by_triplet(a, stream));
  return rc;
}

int velo_three_hit_tracks_filter(const velo_three_hit_tracks_filter_params* p, hipStream_t stream)
{
  if (!p || !p->dev_three_hit_tracks_output || !p->dev_number_of_three_hit_tracks_output)
    return VELO_ERR_INVALID_ARGUMENT;
  velo::ThreeHitFilterLaunch a;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.number_of_events = p->dev_number_of_events;
  a.sorted_container = p->dev_sorted_velo_cluster_container;
  a.offsets_estimated_input_size = p->dev_offsets_estimated_input_size;
  a.three_hit_tracks_input = p->dev_three_hit_tracks_input;
  a.atomics_velo = p->dev_atomics_velo;
  a.hit_used = p->dev_hit_used;
  a.three_hit_tracks_output = p->dev_three_hit_tracks_output;
  a.number_of_three_hit_tracks_output = p->dev_number_of_three_hit_tracks_output;
  a.max_chi2 = p->max_chi2;
  int rc = hip_status(
    velo::launch_fill_u32(p->dev_number_of_three_hit_tracks_output, p->host_number_of_events, 0, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_three_hit_tracks_filter(a, stream));
  return rc;
}

int velo_copy_track_hit_number(const velo_copy_track_hit_number_params* p, hipStream_t stream)
{
  if (!p || !p->dev_offsets_all_velo_tracks || !p->dev_velo_track_hit_number) return VELO_ERR_INVALID_ARGUMENT;
  velo::CopyHitNumberLaunch a;
  a.n_events = p->host_number_of_events;
  a.tracks = p->dev_tracks;
  a.offsets_velo_tracks = p->dev_offsets_velo_tracks;
  a.offsets_estimated_input_size = p->dev_offsets_estimated_input_size;
  a.offsets_three_hit = p->dev_offsets_number_of_three_hit_tracks_filtered;
  a.velo_track_hit_number = p->dev_velo_track_hit_number;
  a.offsets_all_velo_tracks = p->dev_offsets_all_velo_tracks;
  int rc = hip_status(velo::launch_fill_u32(p->dev_offsets_all_velo_tracks, p->host_number_of_events + 1, 0, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_copy_track_hit_number(a, stream));
  if (rc == VELO_OK && p->host_number_of_reconstructed_velo_tracks) {
    rc = hip_status(hipMemcpyAsync(
      p->host_number_of_reconstructed_velo_tracks, p->dev_offsets_all_velo_tracks + p->host_number_of_events,
      sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    if (rc == VELO_OK) rc = hip_status(hipStreamSynchronize(stream));
  }
  return rc;
}

int velo_consolidate_tracks(const velo_consolidate_tracks_params* p, hipStream_t stream)
{
  if (!p || !p->dev_velo_track_hits) return VELO_ERR_INVALID_ARGUMENT;
  velo::ConsolidateLaunch a;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.number_of_events = p->dev_number_of_events;
  a.offsets_all_velo_tracks = p->dev_offsets_all_velo_tracks;
  a.tracks = p->dev_tracks;
  a.offsets_velo_track_hit_number = p->dev_offsets_velo_track_hit_number;
  a.sorted_container = p->dev_sorted_velo_cluster_container;
  a.offsets_estimated_input_size = p->dev_offsets_estimated_input_size;
  a.three_hit_tracks_output = p->dev_three_hit_tracks_output;
  a.offsets_three_hit = p->dev_offsets_number_of_three_hit_tracks_filtered;
  a.velo_track_hits = p->dev_velo_track_hits;
  a.accepted_velo_tracks = p->dev_accepted_velo_tracks;
  // Every track is accepted (VeloConsolidateTracks.cu:84), listed event or not.
  const int n_views = !!p->dev_velo_hits_view + !!p->dev_velo_track_view + !!p->dev_velo_tracks_view +
                      !!p->dev_velo_multi_event_tracks_view + !!p->dev_imec_velo_tracks;
  if (n_views != 0 && n_views != 5) return VELO_ERR_INVALID_ARGUMENT;
  if (n_views && (!p->dev_number_of_events || !p->dev_offsets_all_velo_tracks || !p->dev_offsets_velo_track_hit_number))
    return VELO_ERR_INVALID_ARGUMENT;
  int rc = VELO_OK;
  if (p->dev_accepted_velo_tracks)
    rc = hip_status(
      velo::launch_fill_u8(p->dev_accepted_velo_tracks, p->host_number_of_reconstructed_velo_tracks, 1, stream));
  if (rc == VELO_OK) rc = hip_status(velo::launch_consolidate_tracks(a, stream));
  if (rc == VELO_OK && n_views) {
    velo::VeloViewsLaunch v;
    v.n_events = p->host_number_of_events;
    v.number_of_events = p->dev_number_of_events;
    v.offsets_all_velo_tracks = p->dev_offsets_all_velo_tracks;
    v.offsets_velo_track_hit_number = p->dev_offsets_velo_track_hit_number;
    v.velo_track_hits = p->dev_velo_track_hits;
    v.hits_view = p->dev_velo_hits_view;
    v.track_view = p->dev_velo_track_view;
    v.tracks_view = p->dev_velo_tracks_view;
    v.multi_event_tracks_view = p->dev_velo_multi_event_tracks_view;
    v.imec_velo_tracks = p->dev_imec_velo_tracks;
    rc = hip_status(velo::launch_create_velo_views(v, stream));
  }
  return rc;
}

int velo_kalman_filter(const velo_kalman_filter_params* p, hipStream_t stream)
{
  if (!p || !p->constants || !p->dev_velo_kalman_beamline_states || !p->dev_velo_kalman_endvelo_states ||
      !p->dev_is_backward || !p->dev_number_of_events || !p->dev_offsets_all_velo_tracks)
    return VELO_ERR_INVALID_ARGUMENT;
  if (!p->dev_velo_tracks_view && (!p->dev_offsets_velo_track_hit_number || !p->dev_velo_track_hits))
    return VELO_ERR_INVALID_ARGUMENT;
  velo::KalmanLaunch a;
  a.event_list = p->dev_event_list;
  a.n_list = p->event_list_size;
  a.number_of_events = p->dev_number_of_events;
  a.offsets_all_velo_tracks = p->dev_offsets_all_velo_tracks;
  a.offsets_velo_track_hit_number = p->dev_offsets_velo_track_hit_number;
  a.velo_track_hits = p->dev_velo_track_hits;
  a.beamline = p->constants->dev_beamline;
  a.is_backward = p->dev_is_backward;
  a.beamline_states = p->dev_velo_kalman_beamline_states;
  a.endvelo_states = p->dev_velo_kalman_endvelo_states;
  a.fp64_fallback = p->covariance_fp64_fallback != 0;
  a.tracks_view = p->dev_velo_tracks_view;
  a.beamline_states_view = p->dev_velo_kalman_beamline_states_view;
  a.endvelo_states_view = p->dev_velo_kalman_endvelo_states_view;
  return hip_status(velo::launch_kalman_filter(a, stream));
}

} // extern "C"

// ---------------------------------------------------------------------------------------
// Whole sequence.
struct velo_sequence {
  const velo_constants* constants = nullptr;
  unsigned max_events = 0;
  size_t max_raw_bytes = 0;
  size_t sp_cap = 0, cluster_cap = 0, track_cap = 0, hits_cap = 0;
  std::map<std::string, std::pair<void*, size_t>> bufs;
  unsigned search_block = 256;
  unsigned search_lds_hits = 0; // 0: the launcher's default for the kernel variant
  unsigned search_ring_window = 0; // LDS window of the ring of tracks to follow; 0: 512
  unsigned search_list_cap = 0;    // seeding's list of unused hits; 0: 1024
  bool search_largest_first = true;
  bool search_persistent = false;
  // The largest events (E / search_split of them, events by descending hit count) run on a
  // second stream of the sequence in wider workgroups with more LDS, beside the other events'
  // launch on the run's stream; 0 = one launch. The stream and its events are made when the
  // knob is first set (an idle stream still takes one of the process's hardware queues).
  unsigned search_split = 0;
  unsigned search_big_block = 256;
  unsigned search_big_lds_hits = 4096;
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  bool clustering_resident_grid = false;
  unsigned seed_log_capacity = 0; // seeds; buffers "seed_log_*" exist once the knob was set
  bool scorer_on = false;         // buffer "scorer_params" holds the scorer's parameters
  float scorer_threshold = 0.f;
  unsigned kalman_fp64_fallback = 0;
  bool consolidate_lds_map = true; // 0: consolidation takes the binary-search path for every event
  unsigned pv_finder = 0;         // 1: PV finder after the fit; 2: also the z histograms ("zhisto")
  bool fused_clustering = true;
  bool count_work = false;
  // HIP events bracketing one stage of the run (0 = none), see velo_sequence_stage_ms.
  // One (begin, end) event pair per timed run since the knob was last set.
  unsigned timed_stage = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pairs;
  size_t n_timed = 0;
  // One captured run (velo_sequence_capture / velo_sequence_replay).
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  void release_graph()
  {
    if (graph_exec) (void) hipGraphExecDestroy(graph_exec);
    if (graph) (void) hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }

  template<typename T>
  T* get(const char* name)
  {
    return reinterpret_cast<T*>(bufs.at(name).first);
  }
};

extern "C" {

int velo_sequence_create(const velo_constants* c, unsigned max_events, size_t max_raw_bytes, velo_sequence** out)
{
  if (!c || !out || !max_events || !max_raw_bytes) return VELO_ERR_INVALID_ARGUMENT;
  auto* s = new velo_sequence;
  s->constants = c;
  s->max_events = max_events;
  s->max_raw_bytes = max_raw_bytes;
  const size_t E = max_events;
  s->sp_cap = max_raw_bytes / 4 + 1;
  s->cluster_cap = 2 * s->sp_cap;
  const unsigned T = static_cast<unsigned>(std::min<size_t>(s->cluster_cap, 0xFFFFFFF0u));
  s->track_cap = velo::track_container_size(T, max_events) + 1;
  s->hits_cap = s->track_cap * (velo::max_track_size + 3);
  const size_t scan_words =
    std::max({velo::prefix_sum_scratch_words(26 * max_events), velo::prefix_sum_scratch_words(velo::n_sensors * max_events),
              velo::prefix_sum_scratch_words(2 * s->track_cap)}) +
    1;
  const std::vector<std::pair<const char*, size_t>> layout = {
    {"number_of_events", 4},
    {"status", 4},
    {"number_of_candidates", 4 * E},
    {"velo_bank_index", 4 * velo::n_sensors},
    {"candidates_offsets", 4 * (E + 1)},
    {"estimated_input_size", 4 * 26 * E},
    {"module_candidate_num", 4 * E},
    {"cluster_candidates", 4 * 2 * s->sp_cap},
    {"offsets_estimated_input_size", 4 * (26 * E + 1)},
    {"module_cluster_num", 4 * 26 * E},
    {"velo_cluster_container", velo::cluster_element_bytes() * s->cluster_cap},
    {"hit_permutation", 4 * s->cluster_cap},
    {"sorted_velo_cluster_container", velo::cluster_element_bytes() * s->cluster_cap},
    {"tracks", sizeof(velo::TrackHits) * s->track_cap},
    {"tracklets", sizeof(velo::TrackletHits) * velo::max_tracks_to_follow * E},
    {"tracks_to_follow", 4 * velo::max_tracks_to_follow * E},
    {"three_hit_tracks", sizeof(velo::TrackletHits) * s->track_cap},
    {"hit_used", s->cluster_cap},
    {"atomics_velo", 4 * velo::num_atomics * E},
    {"rel_indices", 2 * s->cluster_cap},
    {"number_of_velo_tracks", 4 * E},
    {"offsets_velo_tracks", 4 * (E + 1)},
    {"three_hit_tracks_output", sizeof(velo::TrackletHits) * s->track_cap},
    {"number_of_three_hit_tracks_output", 4 * E},
    {"offsets_three_hit_tracks_filtered", 4 * (E + 1)},
    {"velo_track_hit_number", 4 * 2 * s->track_cap},
    {"offsets_all_velo_tracks", 4 * (E + 1)},
    {"offsets_velo_track_hit_number", 4 * (2 * s->track_cap + 1)},
    {"velo_track_hits", velo::cluster_element_bytes() * s->hits_cap},
    {"accepted_velo_tracks", 2 * s->track_cap},
    {"is_backward", 2 * s->track_cap},
    {"velo_kalman_beamline_states", 56 * 2 * s->track_cap},
    {"velo_kalman_endvelo_states", 56 * 2 * s->track_cap},
    {"scan_scratch", 4 * scan_words},
    {"search_work", 4 * sizeof(unsigned long long)},
    {"search_event_clock", velo::event_clock_words * sizeof(unsigned long long) * E},
    {"search_order", 4 * E},
    {"search_counter", 4},
    {"sp_flags", s->sp_cap},
    {"bank_directory", 16 * 8 * 26 * E},
    {"bank_directory_count", 4 * 26 * E},
    {"each_sensor_pair_size", 4 * velo::n_sensors * E},
    {"offsets_each_sensor_pair_size", 4 * (velo::n_sensors * E + 1)},
    {"retina_bank_index", 4 * velo::n_sensors},
    {"hit_permutations", 4 * s->sp_cap},
    {"hit_sorting_key", 8 * s->sp_cap},
  };
  for (const auto& kv : layout) {
    void* p = nullptr;
    const size_t b = (kv.second + 255) & ~size_t(255);
    const int rc = hip_status(hipMalloc(&p, b));
    if (rc != VELO_OK) {
      velo_sequence_destroy(s);
      return rc;
    }
    s->bufs[kv.first] = {p, b};
  }
  *out = s;
  return VELO_OK;
}

int velo_sequence_destroy(velo_sequence* s)
{
  if (!s) return VELO_ERR_INVALID_ARGUMENT;
  s->release_graph();
  if (s->fork_ev) (void) hipEventDestroy(s->fork_ev);
  if (s->join_ev) (void) hipEventDestroy(s->join_ev);
  if (s->side) (void) hipStreamDestroy(s->side);
  for (auto& e : s->ev_pairs) {
    (void) hipEventDestroy(e.first);
    (void) hipEventDestroy(e.second);
  }
  for (auto& kv : s->bufs) (void) hipFree(kv.second.first);
  delete s;
  return VELO_OK;
}

void velo_sequence_default_input(velo_sequence_input* in)
{
  std::memset(in, 0, sizeof(*in));
  in->bank_version = 4;
  in->phi_tolerance = 0.045f;
  in->max_scatter = 0.08f;
  in->max_chi2 = 20.f;
  in->max_skipped_modules = 1;
}

int velo_sequence_buffer(velo_sequence* s, const char* name, void** dev_ptr, size_t* capacity_bytes)
{
  if (!s || !name) return VELO_ERR_INVALID_ARGUMENT;
  auto it = s->bufs.find(name);
  if (it == s->bufs.end()) return VELO_ERR_INVALID_ARGUMENT;
  if (dev_ptr) *dev_ptr = it->second.first;
  if (capacity_bytes) *capacity_bytes = it->second.second;
  return VELO_OK;
}

int velo_sequence_set_tuning(velo_sequence* s, const char* knob, unsigned value)
{
  if (!s || !knob) return VELO_ERR_INVALID_ARGUMENT;
  const std::string k(knob);
  if (k == "search_block") {
    if (value && (value % 64 || value > 1024)) return VELO_ERR_INVALID_ARGUMENT;
    s->search_block = value ? value : 256;
    return VELO_OK;
  }
  if (k == "search_lds_hits") {
    if (value > 8192) return VELO_ERR_INVALID_ARGUMENT;
    s->search_lds_hits = value;
    return VELO_OK;
  }
  if (k == "pv_finder") {
    // Buffers of the PV finder, made on first use (the z histograms only for value 2).
    if (value > 2) return VELO_ERR_INVALID_ARGUMENT;
    const size_t E = s->max_events, T = 2 * s->track_cap;
    const std::vector<std::pair<const char*, size_t>> layout = {
      {"pvtracks", 28 * T}, {"zpeaks", 4 * VELO_PV_MAX_VERTICES * E}, {"number_of_zpeaks", 4 * E},
      {"pvtracks_denom", 4 * T}, {"multi_fit_vertices", 48 * VELO_PV_MAX_VERTICES * E},
      {"number_of_multi_fit_vertices", 4 * E}, {"multi_final_vertices", 48 * VELO_PV_MAX_VERTICES * E},
      {"number_of_multi_final_vertices", 4 * E}, {"zhisto", value == 2 ? 4 * VELO_PV_HISTO_BINS * E : 0}};
    for (const auto& kv : layout) {
      if (!kv.second || s->bufs.count(kv.first)) continue;
      void* p = nullptr;
      const int rc = hip_status(hipMalloc(&p, kv.second));
      if (rc != VELO_OK) return rc;
      s->bufs[kv.first] = {p, kv.second};
    }
    s->pv_finder = value;
    return VELO_OK;
  }
  if (k == "kalman_fp64_fallback") {
    s->kalman_fp64_fallback = value != 0;
    return VELO_OK;
  }
  if (k == "consolidate_lds_map") {
    s->consolidate_lds_map = value != 0;
    return VELO_OK;
  }
  if (k == "seed_log") {
    // (Re)allocates the log: counter, records (5 x u32) and features (10 x f32) per seed.
    for (const char* name : {"seed_log_counter", "seed_log_records", "seed_log_features", "seed_log_logits"}) {
      auto it = s->bufs.find(name);
      if (it != s->bufs.end()) {
        (void) hipFree(it->second.first);
        s->bufs.erase(it);
      }
    }
    s->seed_log_capacity = 0;
    if (!value) return VELO_OK;
    const std::vector<std::pair<const char*, size_t>> layout = {
      {"seed_log_counter", 4}, {"seed_log_records", 20 * (size_t) value},
      {"seed_log_features", 4 * velo::n_seed_features * (size_t) value}, {"seed_log_logits", 4 * (size_t) value}};
    for (const auto& kv : layout) {
      void* p = nullptr;
      const int rc = hip_status(hipMalloc(&p, kv.second));
      if (rc != VELO_OK) return rc;
      s->bufs[kv.first] = {p, kv.second};
    }
    s->seed_log_capacity = value;
    return VELO_OK;
  }
  if (k == "clustering_resident_grid") {
    s->clustering_resident_grid = value != 0;
    return VELO_OK;
  }
  if (k == "search_persistent") {
    s->search_persistent = value != 0;
    return VELO_OK;
  }
  if (k == "search_largest_first") {
    s->search_largest_first = value != 0;
    return VELO_OK;
  }
  if (k == "search_ring_window") {
    if (value && (value < 64 || value > 512 || (value & (value - 1)))) return VELO_ERR_INVALID_ARGUMENT;
    s->search_ring_window = value;
    return VELO_OK;
  }
  if (k == "search_list_cap") {
    if (value % 16 || value > 1024) return VELO_ERR_INVALID_ARGUMENT;
    s->search_list_cap = value;
    return VELO_OK;
  }
  if (k == "search_split") {
    // The second stream of the split search and its fork/join events (default priority).
    if (value && !s->side) {
      int rc = hip_status(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
      if (rc == VELO_OK) rc = hip_status(hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming));
      if (rc == VELO_OK) rc = hip_status(hipEventCreateWithFlags(&s->join_ev, hipEventDisableTiming));
      if (rc != VELO_OK) return rc;
    }
    s->search_split = value;
    return VELO_OK;
  }
  if (k == "search_big_block") {
    if (!value || value % 64 || value > 1024) return VELO_ERR_INVALID_ARGUMENT;
    s->search_big_block = value;
    return VELO_OK;
  }
  if (k == "search_big_lds_hits") {
    if (value > 8192) return VELO_ERR_INVALID_ARGUMENT;
    s->search_big_lds_hits = value;
    return VELO_OK;
  }
  if (k == "fused_clustering") {
    s->fused_clustering = value != 0;
    return VELO_OK;
  }
  if (k == "count_work") {
    s->count_work = value != 0;
    return VELO_OK;
  }
  if (k == "timed_stage") {
    if (value > VELO_STAGE_COUNT) return VELO_ERR_INVALID_ARGUMENT;
    s->timed_stage = value;
    s->n_timed = 0;
    return VELO_OK;
  }
  return VELO_ERR_INVALID_ARGUMENT;
}

int velo_sequence_run(velo_sequence* s, const velo_sequence_input* in, hipStream_t stream)
{
  if (!s || !in || !in->dev_velo_raw_input || !in->dev_velo_raw_input_offsets || !in->dev_velo_raw_input_sizes ||
      !in->dev_velo_raw_input_types)
    return VELO_ERR_INVALID_ARGUMENT;
  if (!version_ok(in->bank_version)) return VELO_ERR_BANK_VERSION;
  if (in->number_of_events > s->max_events || in->raw_input_bytes > s->max_raw_bytes) return VELO_ERR_CAPACITY;
  const unsigned E = in->number_of_events;
  if (!E) return VELO_OK;
  auto* n_events_dev = s->get<uint32_t>("number_of_events");
  auto* status = s->get<uint32_t>("status");
  auto* scratch = s->get<uint32_t>("scan_scratch");
#define VELO_TRY(expr)                                                                                              \
  do {                                                                                                              \
    const int rc_ = hip_status(expr);                                                                               \
    if (rc_ != VELO_OK) return rc_;                                                                                 \
  } while (0)
  // Brackets stage `k` (1-based, VELO_STAGE_*) with the sequence's HIP events when it is the timed one.
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  if (s->timed_stage) {
    if (s->n_timed == s->ev_pairs.size()) {
      hipEvent_t a = nullptr, b = nullptr;
      int rc = hip_status(hipEventCreate(&a));
      if (rc == VELO_OK) rc = hip_status(hipEventCreate(&b));
      if (rc != VELO_OK) return rc;
      s->ev_pairs.push_back({a, b});
    }
    ev_begin = s->ev_pairs[s->n_timed].first;
    ev_end = s->ev_pairs[s->n_timed].second;
    ++s->n_timed;
  }
  auto mark_begin = [&](unsigned k) -> int {
    return s->timed_stage == k ? hip_status(hipEventRecord(ev_begin, stream)) : VELO_OK;
  };
  auto mark_end = [&](unsigned k) -> int {
    return s->timed_stage == k ? hip_status(hipEventRecord(ev_end, stream)) : VELO_OK;
  };
#define VELO_MARK(rc_expr)                                                                                          \
  do {                                                                                                              \
    const int rcm_ = (rc_expr);                                                                                     \
    if (rcm_ != VELO_OK) return rcm_;                                                                               \
  } while (0)
  VELO_MARK(mark_begin(VELO_STAGE_ALL));
  VELO_TRY(velo::launch_fill_u32(n_events_dev, 1, E, stream));
  VELO_TRY(velo::launch_fill_u32(status, 1, 0, stream));

  auto* est_off = s->get<uint32_t>("offsets_estimated_input_size");
  auto* mcn = s->get<uint32_t>("module_cluster_num");
  auto* sorted = s->get<char>("sorted_velo_cluster_container");
  if (in->retina) {
    // velo_reconstruction.py:18-71: retina cluster counts per slot, their prefix sum, and the
    // decoding, which writes the module-pair offsets, counts and the sorted container.
    const unsigned slots = velo::retina_sensor_slots(in->bank_version);
    velo::RetinaSizesLaunch r1;
    r1.version = in->bank_version;
    r1.raw = in->dev_velo_raw_input;
    r1.offsets = in->dev_velo_raw_input_offsets;
    r1.sizes = in->dev_velo_raw_input_sizes;
    r1.types = in->dev_velo_raw_input_types;
    r1.mep = in->mep_layout != 0;
    r1.event_list = nullptr;
    r1.n_list = E;
    r1.event_start = 0;
    r1.each_sensor_pair_size = s->get<uint32_t>("each_sensor_pair_size");
    r1.retina_bank_index = s->get<uint32_t>("retina_bank_index");
    VELO_MARK(mark_begin(VELO_STAGE_CANDIDATES));
    VELO_TRY(velo::launch_fill_u32(r1.retina_bank_index, slots, 0xFFFFFFFFu, stream));
    VELO_TRY(velo::launch_retina_sensor_sizes(r1, stream));
    VELO_MARK(mark_end(VELO_STAGE_CANDIDATES));
    auto* sp_off = s->get<uint32_t>("offsets_each_sensor_pair_size");
    VELO_MARK(mark_begin(VELO_STAGE_ESTIMATE));
    VELO_TRY(velo::launch_prefix_sum(r1.each_sensor_pair_size, slots * E, sp_off, scratch, stream));
    VELO_MARK(mark_end(VELO_STAGE_ESTIMATE));
    velo::RetinaDecodeLaunch r2;
    r2.version = in->bank_version;
    r2.raw = r1.raw;
    r2.offsets = r1.offsets;
    r2.sizes = r1.sizes;
    r2.types = r1.types;
    r2.mep = r1.mep;
    r2.event_list = nullptr;
    r2.n_list = E;
    r2.event_start = 0;
    r2.geometry = s->constants->dev_geometry;
    r2.offsets_each_sensor_pair_size = sp_off;
    r2.retina_bank_index = r1.retina_bank_index;
    r2.number_of_events = n_events_dev;
    r2.module_cluster_num = mcn;
    r2.offsets_module_pair_cluster = est_off;
    r2.cluster_container = sorted;
    r2.hit_permutations = s->get<uint32_t>("hit_permutations");
    r2.hit_sorting_key = s->get<int64_t>("hit_sorting_key");
    r2.populate_items = E * velo::n_module_pairs;
    VELO_MARK(mark_begin(VELO_STAGE_CLUSTERING));
    VELO_TRY(velo::launch_retina_decode(r2, stream));
    VELO_MARK(mark_end(VELO_STAGE_CLUSTERING));
  }
  else {
    velo::CandidatesLaunch k1;
    k1.version = in->bank_version;
    k1.raw = in->dev_velo_raw_input;
    k1.offsets = in->dev_velo_raw_input_offsets;
    k1.sizes = in->dev_velo_raw_input_sizes;
    k1.types = in->dev_velo_raw_input_types;
    k1.mep = in->mep_layout != 0;
    k1.event_list = nullptr;
    k1.n_list = E;
    k1.event_start = 0;
    k1.block_dim_x = 256;
    k1.number_of_candidates = s->get<uint32_t>("number_of_candidates");
    k1.bank_index = s->get<uint32_t>("velo_bank_index");
    if (s->fused_clustering) {
      k1.directory = s->get<void>("bank_directory");
      k1.directory_count = s->get<uint32_t>("bank_directory_count");
    }
    VELO_MARK(mark_begin(VELO_STAGE_CANDIDATES));
    VELO_TRY(velo::launch_number_of_candidates(k1, stream));
    VELO_MARK(mark_end(VELO_STAGE_CANDIDATES));
    VELO_TRY(velo::launch_prefix_sum(k1.number_of_candidates, E, s->get<uint32_t>("candidates_offsets"), scratch, stream));

    velo::EstimateLaunch k2;
    k2.version = in->bank_version;
    k2.raw = k1.raw;
    k2.offsets = k1.offsets;
    k2.sizes = k1.sizes;
    k2.types = k1.types;
    k2.mep = k1.mep;
    k2.event_list = nullptr;
    k2.n_list = E;
    k2.event_start = 0;
    k2.candidates_offsets = s->get<uint32_t>("candidates_offsets");
    k2.estimated_input_size = s->get<uint32_t>("estimated_input_size");
    k2.module_candidate_num = s->get<uint32_t>("module_candidate_num");
    k2.cluster_candidates = s->get<uint32_t>("cluster_candidates");
    auto* sp_flags = s->get<uint8_t>("sp_flags");
    VELO_MARK(mark_begin(VELO_STAGE_ESTIMATE));
    if (s->fused_clustering) {
      velo::EstimateMpLaunch e;
      e.version = in->bank_version;
      e.raw = k1.raw;
      e.directory = k1.directory;
      e.directory_count = k1.directory_count;
      e.n_events = E;
      e.estimated_input_size = k2.estimated_input_size;
      e.sp_flags = sp_flags;
      e.resident_grid = s->clustering_resident_grid;
      VELO_TRY(velo::launch_estimate_mp(e, stream));
    }
    else {
      VELO_TRY(velo::launch_fill_u32(k2.module_candidate_num, E, 0, stream));
      VELO_TRY(velo::launch_estimate_input_size(k2, stream));
    }
    VELO_MARK(mark_end(VELO_STAGE_ESTIMATE));
    VELO_TRY(velo::launch_prefix_sum(k2.estimated_input_size, 26 * E, est_off, scratch, stream));

    VELO_MARK(mark_begin(VELO_STAGE_CLUSTERING));
    if (s->fused_clustering) {
      velo::ClusterMpLaunch f;
      f.version = in->bank_version;
      f.raw = k1.raw;
      f.directory = k1.directory;
      f.directory_count = k1.directory_count;
      f.n_events = E;
      f.offsets_estimated_input_size = est_off;
      f.sp_flags = sp_flags;
      f.geometry = s->constants->dev_geometry;
      f.sp_patterns = s->constants->dev_sp_patterns;
      f.sp_fx = s->constants->dev_sp_fx;
      f.sp_fy = s->constants->dev_sp_fy;
      f.module_cluster_num = mcn;
      f.sorted_container = sorted;
      f.unsorted_scratch = s->get<char>("velo_cluster_container");
      f.resident_grid = s->clustering_resident_grid;
      VELO_TRY(velo::launch_cluster_mp(f, stream));
    }
    else {
      velo::MaskedClusteringLaunch k3;
      k3.version = in->bank_version;
      k3.raw = k1.raw;
      k3.offsets = k1.offsets;
      k3.sizes = k1.sizes;
      k3.types = k1.types;
      k3.mep = k1.mep;
      k3.event_list = nullptr;
      k3.n_list = E;
      k3.event_start = 0;
      k3.number_of_events = n_events_dev;
      k3.offsets_estimated_input_size = est_off;
      k3.module_candidate_num = k2.module_candidate_num;
      k3.cluster_candidates = k2.cluster_candidates;
      k3.candidates_offsets = k2.candidates_offsets;
      k3.bank_index = k1.bank_index;
      k3.geometry = s->constants->dev_geometry;
      k3.sp_patterns = s->constants->dev_sp_patterns;
      k3.sp_fx = s->constants->dev_sp_fx;
      k3.sp_fy = s->constants->dev_sp_fy;
      k3.module_cluster_num = mcn;
      k3.cluster_container = s->get<char>("velo_cluster_container");
      VELO_TRY(velo::launch_masked_clustering(k3, stream));
      velo::SortLaunch k4;
      k4.event_list = nullptr;
      k4.n_list = E;
      k4.number_of_events = n_events_dev;
      k4.offsets_estimated_input_size = est_off;
      k4.module_cluster_num = mcn;
      k4.cluster_container = k3.cluster_container;
      k4.sorted_container = sorted;
      k4.hit_permutation = s->get<uint32_t>("hit_permutation");
      VELO_TRY(velo::launch_sort_by_phi(k4, stream));
    }
    VELO_MARK(mark_end(VELO_STAGE_CLUSTERING));
  }

  velo::SearchLaunch k5;
  k5.event_list = nullptr;
  k5.n_list = E;
  k5.number_of_events = n_events_dev;
  k5.sorted_container = sorted;
  k5.offsets_estimated_input_size = est_off;
  k5.module_cluster_num = mcn;
  k5.geometry = s->constants->dev_geometry;
  k5.tracks = s->get<void>("tracks");
  k5.tracklets = s->get<void>("tracklets");
  k5.tracks_to_follow = s->get<uint32_t>("tracks_to_follow");
  k5.three_hit_tracks = s->get<void>("three_hit_tracks");
  k5.hit_used = s->get<uint8_t>("hit_used");
  k5.atomics_velo = s->get<uint32_t>("atomics_velo");
  k5.rel_indices = s->get<uint16_t>("rel_indices");
  k5.number_of_velo_tracks = s->get<uint32_t>("number_of_velo_tracks");
  k5.status = status;
  k5.phi_tolerance = in->phi_tolerance;
  k5.max_scatter = in->max_scatter;
  k5.max_skipped_modules = in->max_skipped_modules;
  k5.block_dim = s->search_block;
  k5.lds_hits = s->search_lds_hits;
  k5.ring_window = s->search_ring_window;
  k5.list_cap = s->search_list_cap;
  k5.work_stats = s->count_work ? s->get<unsigned long long>("search_work") : nullptr;
  k5.event_clock = s->count_work ? s->get<unsigned long long>("search_event_clock") : nullptr;
  if (k5.work_stats) VELO_TRY(velo::launch_fill_u32(reinterpret_cast<uint32_t*>(k5.work_stats), 8, 0, stream));
  VELO_MARK(mark_begin(VELO_STAGE_SEARCH));
  if (s->seed_log_capacity) {
    k5.seed_log.counter = s->get<uint32_t>("seed_log_counter");
    k5.seed_log.records = s->get<uint32_t>("seed_log_records");
    k5.seed_log.features = s->get<float>("seed_log_features");
    k5.seed_log.capacity = s->seed_log_capacity;
    k5.seed_log.logits = s->get<float>("seed_log_logits");
    VELO_TRY(velo::launch_fill_u32(k5.seed_log.counter, 1, 0, stream));
  }
  if (s->scorer_on) {
    k5.scorer_params = s->get<float>("scorer_params");
    k5.scorer_threshold = s->scorer_threshold;
  }
  if (s->search_persistent) {
    k5.work_counter = s->get<uint32_t>("search_counter");
    VELO_TRY(velo::launch_fill_u32(k5.work_counter, 1, 0, stream));
  }
  if (s->search_largest_first) {
    k5.event_list = s->get<uint32_t>("search_order");
    VELO_TRY(velo::launch_order_events_by_hits(est_off, n_events_dev, const_cast<uint32_t*>(k5.event_list), stream));
  }
  // The largest events on the sequence's second stream (fork and join through events, so the
  // run stays one stream to the caller and capturable); the persistent grid takes every event.
  const unsigned n_big = s->search_largest_first && !s->search_persistent && s->search_split ? E / s->search_split : 0u;
  if (n_big) {
    VELO_TRY(velo::launch_search_split(k5, n_big, s->search_big_block, s->search_big_lds_hits, stream, s->side,
                                       s->fork_ev, s->join_ev));
  }
  else {
    VELO_TRY(velo::launch_search_by_triplet(k5, stream));
  }
  VELO_MARK(mark_end(VELO_STAGE_SEARCH));
  auto* off_tracks = s->get<uint32_t>("offsets_velo_tracks");
  VELO_TRY(velo::launch_prefix_sum(k5.number_of_velo_tracks, E, off_tracks, scratch, stream));

  velo::ThreeHitFilterLaunch k6;
  k6.event_list = nullptr;
  k6.n_list = E;
  k6.number_of_events = n_events_dev;
  k6.sorted_container = sorted;
  k6.offsets_estimated_input_size = est_off;
  k6.three_hit_tracks_input = k5.three_hit_tracks;
  k6.atomics_velo = k5.atomics_velo;
  k6.hit_used = k5.hit_used;
  k6.three_hit_tracks_output = s->get<void>("three_hit_tracks_output");
  k6.number_of_three_hit_tracks_output = s->get<uint32_t>("number_of_three_hit_tracks_output");
  k6.max_chi2 = in->max_chi2;
  VELO_MARK(mark_begin(VELO_STAGE_THREE_HIT_FILTER));
  VELO_TRY(velo::launch_three_hit_tracks_filter(k6, stream));
  VELO_MARK(mark_end(VELO_STAGE_THREE_HIT_FILTER));
  auto* off_three = s->get<uint32_t>("offsets_three_hit_tracks_filtered");
  VELO_TRY(velo::launch_prefix_sum(k6.number_of_three_hit_tracks_output, E, off_three, scratch, stream));

  velo::CopyHitNumberLaunch k7;
  k7.n_events = E;
  k7.tracks = k5.tracks;
  k7.offsets_velo_tracks = off_tracks;
  k7.offsets_estimated_input_size = est_off;
  k7.offsets_three_hit = off_three;
  k7.velo_track_hit_number = s->get<uint32_t>("velo_track_hit_number");
  k7.offsets_all_velo_tracks = s->get<uint32_t>("offsets_all_velo_tracks");
  VELO_TRY(velo::launch_copy_track_hit_number(k7, stream));
  auto* off_hits = s->get<uint32_t>("offsets_velo_track_hit_number");
  VELO_TRY(velo::launch_prefix_sum_dyn(
    k7.velo_track_hit_number, k7.offsets_all_velo_tracks + E, 2 * s->track_cap, off_hits, scratch, stream));

  velo::ConsolidateLaunch k8;
  k8.lds_map = s->consolidate_lds_map;
  k8.event_list = nullptr;
  k8.n_list = E;
  k8.number_of_events = n_events_dev;
  k8.offsets_all_velo_tracks = k7.offsets_all_velo_tracks;
  k8.tracks = k5.tracks;
  k8.offsets_velo_track_hit_number = off_hits;
  k8.sorted_container = sorted;
  k8.offsets_estimated_input_size = est_off;
  k8.three_hit_tracks_output = k6.three_hit_tracks_output;
  k8.offsets_three_hit = off_three;
  k8.velo_track_hits = s->get<char>("velo_track_hits");
  k8.accepted_velo_tracks = s->get<uint8_t>("accepted_velo_tracks");
  VELO_MARK(mark_begin(VELO_STAGE_CONSOLIDATE));
  VELO_TRY(velo::launch_consolidate_tracks(k8, stream));
  VELO_MARK(mark_end(VELO_STAGE_CONSOLIDATE));

  velo::KalmanLaunch k9;
  k9.event_list = nullptr;
  k9.n_list = E;
  k9.number_of_events = n_events_dev;
  k9.offsets_all_velo_tracks = k7.offsets_all_velo_tracks;
  k9.offsets_velo_track_hit_number = off_hits;
  k9.velo_track_hits = k8.velo_track_hits;
  k9.beamline = s->constants->dev_beamline;
  k9.is_backward = s->get<uint8_t>("is_backward");
  k9.beamline_states = s->get<char>("velo_kalman_beamline_states");
  k9.fp64_fallback = s->kalman_fp64_fallback;
  k9.endvelo_states = s->get<char>("velo_kalman_endvelo_states");
  VELO_MARK(mark_begin(VELO_STAGE_KALMAN));
  VELO_TRY(velo::launch_kalman_filter(k9, stream));
  VELO_MARK(mark_end(VELO_STAGE_KALMAN));
  if (s->pv_finder) {
    velo::PVLaunch pv;
    pv.event_list = nullptr;
    pv.n_list = E;
    pv.number_of_events = n_events_dev;
    pv.offsets_all_velo_tracks = k7.offsets_all_velo_tracks;
    pv.beamline_states = s->get<float>("velo_kalman_beamline_states");
    pv.beamline = s->constants->dev_beamline;
    pv.pvtracks = s->get<velo::PVTrackOut>("pvtracks");
    pv.zhisto = s->pv_finder == 2 ? s->get<float>("zhisto") : nullptr;
    pv.zpeaks = s->get<float>("zpeaks");
    pv.number_of_zpeaks = s->get<uint32_t>("number_of_zpeaks");
    pv.pvtracks_denom = s->get<float>("pvtracks_denom");
    pv.multi_fit_vertices = s->get<velo::PVVertexOut>("multi_fit_vertices");
    pv.number_of_multi_fit_vertices = s->get<uint32_t>("number_of_multi_fit_vertices");
    pv.multi_final_vertices = s->get<velo::PVVertexOut>("multi_final_vertices");
    pv.number_of_multi_final_vertices = s->get<uint32_t>("number_of_multi_final_vertices");
    pv.status = status;
    VELO_MARK(mark_begin(VELO_STAGE_PV));
    VELO_TRY(velo::launch_pv_beamline(pv, stream));
    VELO_MARK(mark_end(VELO_STAGE_PV));
  }
  VELO_MARK(mark_end(VELO_STAGE_ALL));
#undef VELO_MARK
#undef VELO_TRY
  return VELO_OK;
}

int velo_sequence_capture(velo_sequence* s, const velo_sequence_input* in, hipStream_t stream)
{
  if (!s || !in || !stream) return VELO_ERR_INVALID_ARGUMENT;
  s->release_graph();
  // Stage timing records events on the host's schedule; it is off inside a captured run.
  const unsigned timed = s->timed_stage;
  s->timed_stage = 0;
  int rc = hip_status(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
  if (rc == VELO_OK) {
    const int run_rc = velo_sequence_run(s, in, stream);
    hipGraph_t g = nullptr;
    const int end_rc = hip_status(hipStreamEndCapture(stream, &g));
    rc = run_rc != VELO_OK ? run_rc : end_rc;
    if (rc == VELO_OK) rc = hip_status(hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
    if (rc == VELO_OK) s->graph = g;
    else if (g) (void) hipGraphDestroy(g);
  }
  s->timed_stage = timed;
  return rc;
}

size_t velo_mlp_scratch_floats(unsigned n)
{
  const size_t blocks = velo::mlp::mlp_blocks(n);
  return blocks * velo::mlp::n_params + blocks;
}

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
  hipStream_t stream)
{
  if (!dev_features || !dev_labels || !n || !dev_params || !dev_scratch) return VELO_ERR_INVALID_ARGUMENT;
  if (lr != 0.f && (!dev_adam_m || !dev_adam_v || step == 0)) return VELO_ERR_INVALID_ARGUMENT;
  const unsigned blocks = velo::mlp::mlp_blocks(n);
  float* partial = dev_scratch;
  float* loss_partial = dev_scratch + (size_t) blocks * velo::mlp::n_params;
  int rc = hip_status(velo::launch_mlp_train(
    dev_features, dev_labels, n, 1.f / static_cast<float>(n), dev_params, partial, loss_partial, stream));
  if (rc == VELO_OK)
    rc = hip_status(velo::launch_mlp_reduce_adam(
      partial, loss_partial, blocks, dev_params, dev_adam_m, dev_adam_v, dev_grad, dev_loss, lr, step, stream));
  return rc;
}

int velo_mlp_score(const float* dev_features, unsigned n, const float* dev_params, float* dev_logits, hipStream_t stream)
{
  if (!dev_features || !dev_params || !dev_logits) return VELO_ERR_INVALID_ARGUMENT;
  return hip_status(velo::launch_mlp_score(dev_features, n, dev_params, dev_logits, stream));
}

int velo_sequence_set_scorer(velo_sequence* s, const float* params, float threshold)
{
  if (!s) return VELO_ERR_INVALID_ARGUMENT;
  if (!params) {
    s->scorer_on = false;
    return VELO_OK;
  }
  if (!s->bufs.count("scorer_params")) {
    void* p = nullptr;
    const int rc = hip_status(hipMalloc(&p, 4 * velo::mlp::n_params));
    if (rc != VELO_OK) return rc;
    s->bufs["scorer_params"] = {p, 4 * velo::mlp::n_params};
  }
  const int rc = hip_status(
    hipMemcpy(s->get<float>("scorer_params"), params, 4 * velo::mlp::n_params, hipMemcpyHostToDevice));
  if (rc != VELO_OK) return rc;
  s->scorer_on = true;
  s->scorer_threshold = threshold;
  return VELO_OK;
}

int velo_sequence_replay(velo_sequence* s, hipStream_t stream)
{
  if (!s || !s->graph_exec) return VELO_ERR_INVALID_ARGUMENT;
  return hip_status(hipGraphLaunch(s->graph_exec, stream));
}

int velo_sequence_stage_ms(velo_sequence* s, float* total_ms, unsigned* runs)
{
  if (!s || !total_ms || !runs || !s->timed_stage) return VELO_ERR_INVALID_ARGUMENT;
  *total_ms = 0.f;
  *runs = static_cast<unsigned>(s->n_timed);
  for (size_t i = 0; i < s->n_timed; ++i) {
    float ms = 0.f;
    int rc = hip_status(hipEventSynchronize(s->ev_pairs[i].second));
    if (rc == VELO_OK) rc = hip_status(hipEventElapsedTime(&ms, s->ev_pairs[i].first, s->ev_pairs[i].second));
    if (rc != VELO_OK) return rc;
    *total_ms += ms;
  }
  return VELO_OK;
}

int velo_sequence_status(velo_sequence* s, hipStream_t stream, unsigned* status)
{
  if (!s || !status) return VELO_ERR_INVALID_ARGUMENT;
  int rc = hip_status(hipStreamSynchronize(stream));
  if (rc == VELO_OK)
    rc = hip_status(hipMemcpy(status, s->get<uint32_t>("status"), sizeof(unsigned), hipMemcpyDeviceToHost));
  return rc;
}

} // extern "C"
