// Shared device code of the MI355X (gfx950) VELO path: detector constants, the Allen
// event/cluster layouts, the phi arithmetic, and wave/block scans used for ordered
// compaction. Compiled with -ffp-contract=off: every float expression below is evaluated
// as written (no FMA contraction), which is what makes results bit-identical to the
// CPU target of the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace velo {

// device/event_model/velo/include/VeloDefinitions.cuh:16-57
constexpr unsigned n_modules = 52;
constexpr unsigned n_module_pairs = 26;
constexpr unsigned n_sensors = 208;
constexpr unsigned n_sensor_columns = 768;
constexpr float z_end_velo = 770.f;
constexpr float max_number_of_tracks_per_cluster = 0.34f;
constexpr unsigned minimum_container_size = 10;
constexpr unsigned max_track_size = 26;
constexpr unsigned max_tracks_to_follow = 2048;
constexpr float pixel_size = 0.055f;
constexpr float param_w = 6400.f;
constexpr float param_w_inverted = 0.00015625f;
constexpr unsigned max_h0_candidates = 8;
constexpr unsigned initial_seeding_h0_candidates = 8;
constexpr unsigned seeding_h0_candidates = 4;
constexpr unsigned num_atomics = 3;
enum { atomic_three_hit = 0, atomic_ttf = 1, atomic_local_hits = 2 };
constexpr uint32_t bit_seed = 0x80000000u;
constexpr uint32_t bit_track_number = 0x0FFFFFFFu;
constexpr uint32_t bit_skipped = 0x70000000u;
constexpr unsigned skipped_pos = 28;
constexpr unsigned bank_type_velo = 8;  // LHCb::RawBank::Velo
constexpr unsigned bank_type_vp = 63;   // LHCb::RawBank::VP
constexpr float pi_f = 3.141592654f;    // BackendCommonInterface.h:49
constexpr float convert_factor_i16 = 32768.0f / pi_f;

// Status bits raised when a fixed-capacity structure of the reference overflows.
enum : uint32_t {
  status_ttf_ring_overflow = 1u << 0,
  status_track_slots_overflow = 1u << 1,
  status_hit_index_overflow = 1u << 2,
  status_three_hit_slots_overflow = 1u << 3,
  status_lds_fallback = 1u << 8,
};

// VeloGeometry with fp32 fields (ClusteringDefinitions.cuh:154-198); n_trans is always 12.
struct Geometry {
  float module_zs[n_modules];
  float local_x[n_sensor_columns];
  float x_pitch[n_sensor_columns];
  float ltg[12 * n_sensors];
};

// Velo::TrackletHits / Velo::TrackHits (VeloEventModel.cuh:53-86)
struct TrackletHits {
  uint16_t hits[3];
};
struct TrackHits {
  uint16_t hits[max_track_size];
  uint8_t hitsNum;
};
static_assert(sizeof(TrackletHits) == 6, "layout");
static_assert(sizeof(TrackHits) == 54, "layout");

constexpr size_t cluster_element_bytes() { return 18; }

// Cluster SoA with half_t == float (VeloEventModel.cuh:92-183):
// [u32 id x T][f32 xyz x 3T][i16 phi x T]; T = total estimated clusters.
struct Clusters {
  char* base;
  unsigned total;
  __device__ __forceinline__ uint32_t* ids() const { return reinterpret_cast<uint32_t*>(base); }
  __device__ __forceinline__ float* xyz() const { return reinterpret_cast<float*>(base) + total; }
  __device__ __forceinline__ int16_t* phis() const { return reinterpret_cast<int16_t*>(base) + 8 * (size_t) total; }
};

// Raw-event accessors: Allen's layout (ClusteringDefinitions.cuh:100-146; MEPTools.h:25-55), or
// with mep the MEP layout (MEP::RawEvent, MEPTools.h:59-200): offsets[0] = banks, offsets[2 + b]
// = source ID of bank b, offsets[2 + banks (1 + event) + b] = byte offset of the event's fragment
// of bank b (a bank without its source ID word); sizes and types indexed by bank, then event.
struct RawEvent {
  uint32_t n_banks;
  const uint32_t* bank_offsets; // Allen: the event's bank offsets; MEP: the offsets array
  const char* payload;          // Allen: the event's payload; MEP: the fragments
  const uint16_t* bank_sizes;   // Allen: the event's sizes; MEP: the sizes array
  const uint8_t* bank_types;    // Allen: the event's types; MEP: the types array
  unsigned event;               // MEP: the event's index
  bool mep;
  __device__ __forceinline__ RawEvent(
    const char* raw,
    const uint32_t* offsets,
    const uint32_t* sizes,
    const uint32_t* types,
    unsigned mep_layout,
    unsigned event_number)
  {
    mep = mep_layout != 0;
    event = event_number;
    if (mep) {
      n_banks = offsets[0];
      bank_offsets = offsets;
      payload = raw;
      bank_sizes = reinterpret_cast<const uint16_t*>(sizes);
      bank_types = reinterpret_cast<const uint8_t*>(types);
      return;
    }
    const char* p = raw + offsets[event_number];
    n_banks = *reinterpret_cast<const uint32_t*>(p);
    bank_offsets = reinterpret_cast<const uint32_t*>(p + 4);
    payload = p + 4 + 4 * (size_t)(n_banks + 1);
    bank_sizes = reinterpret_cast<const uint16_t*>(sizes) + sizes[event_number];
    bank_types = reinterpret_cast<const uint8_t*>(types) + types[event_number];
  }
};

// One bank of a module pair in the bank directory (16 bytes, one 128-bit load).
struct BankEntry {
  uint32_t word_offset; // first SP word, as an index into the raw buffer in 32-bit words
  uint32_t n_sp;
  uint32_t source;
  uint32_t pad;
};

struct RawBank {
  uint32_t sourceID;
  uint32_t count;
  const uint32_t* word;
  uint16_t size;
  uint8_t type;
};

template<int version>
__device__ __forceinline__ RawBank raw_bank(const RawEvent& ev, unsigned i)
{
  RawBank b;
  const uint32_t* p;
  if (ev.mep) {
    // MEP::raw_bank (MEPTools.h:112-126): the fragment, its source ID from the offsets.
    const uint32_t* offs = ev.bank_offsets;
    b.sourceID = offs[2 + i];
    p = reinterpret_cast<const uint32_t*>(ev.payload + offs[2 + ev.n_banks * (1 + ev.event) + i]);
    b.size = (ev.bank_sizes + reinterpret_cast<const uint32_t*>(ev.bank_sizes)[i])[ev.event];
    b.type = (ev.bank_types + reinterpret_cast<const uint32_t*>(ev.bank_types)[i])[ev.event];
  }
  else {
    p = reinterpret_cast<const uint32_t*>(ev.payload + ev.bank_offsets[i]);
    b.sourceID = p[0];
    p += 1;
    b.size = ev.bank_sizes[i];
    b.type = ev.bank_types[i];
  }
  if constexpr (version == 2 || version == 3) {
    b.count = p[0];
    b.word = p + 1;
  }
  else {
    b.count = 0;
    b.word = p;
  }
  return b;
}

template<int version>
__device__ __forceinline__ uint32_t bank_n_sp(const RawBank& b)
{
  if constexpr (version == 2 || version == 3) return b.count;
  else return b.size / 4u;
}

__device__ __forceinline__ bool is_velo_bank(uint8_t t) { return t == bank_type_vp || t == bank_type_velo; }

// ClusteringDefinitions.cuh:200-216, LHCbID.cuh:38-41
__device__ __forceinline__ uint32_t lhcb_id_of(unsigned sensor, unsigned chip, unsigned col, unsigned row)
{
  const uint32_t cid = (sensor << 18) | (chip << 16) | (col << 8) | row;
  return (0x2u << 28) | (cid & 0x0FFFFFFFu);
}

// VeloTools.cuh:18-50 — octant approximation of atan2, then scaled to int16.
__device__ __forceinline__ float fast_atan2f(float y, float x)
{
  const float c1 = .25f * pi_f;
  const float c2 = .75f * pi_f;
  const float abs_y = fabsf(y);
  const float x_plus_y = x + abs_y;
  const float x_sub_y = x - abs_y;
  const float y_sub_x = abs_y - x;
  const bool pos = x > 0;
  const float nom = pos ? x_sub_y : x_plus_y;
  const float den = pos ? x_plus_y : y_sub_x;
  const float r = nom / den;
  const float angle = (pos ? c1 : c2) - c1 * r;
  return copysignf(angle, y);
}
// float -> int16 through int32; NaN maps to 0 (low half of what x86 cvttss2si yields).
__device__ __forceinline__ int16_t f32_to_i16(float f)
{
  const int32_t i = (f != f) ? 0 : static_cast<int32_t>(f);
  return static_cast<int16_t>(i);
}
__device__ __forceinline__ int16_t hit_phi_16(float x, float y)
{
  return f32_to_i16(fast_atan2f(y, x) * convert_factor_i16);
}

// BinarySearch.cuh:15-30
template<typename T>
__device__ __forceinline__ int binary_search_leftmost(const T* array, unsigned size, T value)
{
  int l = 0;
  int r = size;
  while (l < r) {
    const int m = (l + r) / 2;
    if (value > array[m]) l = m + 1;
    else r = m;
  }
  return l;
}

// VeloEventModel.cuh:192-196
__device__ __host__ __forceinline__ unsigned track_offset(const uint32_t* offsets_estimated, unsigned event)
{
  return static_cast<unsigned>(static_cast<float>(offsets_estimated[event * n_module_pairs]) *
                               max_number_of_tracks_per_cluster);
}

__host__ __device__ __forceinline__ unsigned track_container_size(unsigned total_clusters, unsigned n_events)
{
  const float a = static_cast<float>(total_clusters) * max_number_of_tracks_per_cluster;
  const unsigned b = n_events * minimum_container_size;
  return a < static_cast<float>(b) ? b : static_cast<unsigned>(a);
}

// ---------------------------------------------------------------------------------------
// Wave (64 lanes) and block scans.
constexpr unsigned wave_size = 64;

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanes_below_mask()
{
  const unsigned l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive wave scan of u32.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
  const unsigned l = lane_id();
#pragma unroll
  for (unsigned d = 1; d < wave_size; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, wave_size);
    if (l >= d) v += o;
  }
  return v;
}

// Block-wide exclusive scan of u32; `scratch` holds >= blockDim/64 + 1 u32 in LDS.
// Returns the exclusive prefix of this thread and writes the block total to *total.
// Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* scratch, uint32_t* total)
{
  const unsigned wave = threadIdx.x / wave_size;
  const unsigned n_waves = (blockDim.x + wave_size - 1) / wave_size;
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane_id() == wave_size - 1 || threadIdx.x == blockDim.x - 1) scratch[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (unsigned w = 0; w < n_waves; ++w) {
      const uint32_t t = scratch[w];
      scratch[w] = s;
      s += t;
    }
    scratch[n_waves] = s;
  }
  __syncthreads();
  const uint32_t r = scratch[wave] + inc - v;
  *total = scratch[n_waves];
  __syncthreads();
  return r;
}

// Block-wide exclusive scan of a 0/1 flag through wave ballots.
__device__ __forceinline__ uint32_t block_flag_scan(bool flag, uint32_t* scratch, uint32_t* total)
{
  const unsigned wave = threadIdx.x / wave_size;
  const unsigned n_waves = (blockDim.x + wave_size - 1) / wave_size;
  const uint64_t b = __ballot(flag);
  const uint32_t below = __popcll(b & lanes_below_mask());
  if (lane_id() == 0) scratch[wave] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (unsigned w = 0; w < n_waves; ++w) {
      const uint32_t t = scratch[w];
      scratch[w] = s;
      s += t;
    }
    scratch[n_waves] = s;
  }
  __syncthreads();
  const uint32_t r = scratch[wave] + below;
  *total = scratch[n_waves];
  __syncthreads();
  return r;
}

} // namespace velo
