// TEST INFRASTRUCTURE — CPU oracle for the VELO reconstruction path.
//
// This is a from-scratch restatement of Allen's VELO-SP path with the
// semantics of Allen's CPU target (every kernel runs with blockDim = 1,
// sequentially over blockIdx, with non-atomic "atomics"), the fp32 cluster
// layout (half_t == float) and NO floating-point contraction (built with
// -ffp-contract=off).  It is the checker for the HIP path: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// The product path (allen_amd/) never calls into it.
//
// Execution semantics followed: backend/include/CPUBackend.h:99-145,
// stream/gear/include/Invoke.cuh:40-52.
//
// What pins each part (tests/test_oracle_pins.py): the super-pixel tables by a textbook
// labelling; LHCbIDs, cluster membership and positions by a textbook 8-connected flood fill
// (float64 positions) on the bundled MDF's real banks and on synthetic banks; fast_atan2f /
// hit_phi_16 by special angles and the closed-form error bound; the binary search by bisect;
// the prefix sum by the reference's own test shape (TestPrefixSum.cu:9-33); the phi sort and
// search by the reference's contracts plus known-answer single tracks; the three-hit chi2 by
// the closed-form 3-point fit; the FP64 Kalman by a textbook matrix filter.
// Parity unpinned (the restatement follows the reference's code and nothing independent
// fixes the result): K2's estimate for non-isolated SPs (EstimateInputSize.cu:104-204),
// rest_of_clusters' decisions at the edges of its window, seeding tie-breaks of the CPU
// target's umesimd track_seeding_vectorized (SearchByTriplet.cu:616-818; the scalar
// track_seeding is followed), and the ring of tracks to follow beyond 2048 live entries.
#pragma once

#include <cstdint>
#include <cstring>
#include <cstddef>
#include <vector>

// Planted mistakes for the pin tests (tests/test_pins_mutants.py): VO_MUTANT = n selects one
// wrong variant of a rule (see each "#if VO_MUTANT" site). Only `make mutants` builds them, into
// separate libraries under build/mutants/; the oracle itself is built with none.
#ifndef VO_MUTANT
#define VO_MUTANT 0
#endif

namespace vo {

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

// LHCb raw bank types accepted by the VELO decoders (mdf/lhcb/Event/RawBank.h:69,124)
constexpr unsigned bank_type_velo = 8;
constexpr unsigned bank_type_vp = 63;
constexpr unsigned bank_type_retina = 73; // VPRetinaCluster (mdf/lhcb/Event/RawBank.h:134)

// backend/include/BackendCommonInterface.h:49
constexpr float pi_f = 3.141592654f;

// device/velo/common/include/VeloTools.cuh:11-14 (evaluated in float, as constexpr)
constexpr float convert_factor_i16 = 32768.0f / pi_f;

// ClusteringDefinitions.cuh:154-198 (VeloGeometry, fp32 fields)
struct Geometry {
  float module_zs[n_modules];
  float local_x[n_sensor_columns];
  float x_pitch[n_sensor_columns];
  float ltg[12 * n_sensors];
};

// host/velo/clustering/src/ClusteringCommon.cpp:6-104 output
struct SpTables {
  uint8_t patterns[256];
  uint8_t sizes[256];
  float fx[512];
  float fy[512];
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

// Clusters_t with half_t == float (VeloEventModel.cuh:92-183):
// [u32 id x T][f32 xyz x 3T][i16 phi x T], T = total estimated clusters.
struct Clusters {
  char* base;
  unsigned total;
  unsigned offset;
  uint32_t* ids() const { return reinterpret_cast<uint32_t*>(base); }
  float* xyz() const { return reinterpret_cast<float*>(base) + total; }
  int16_t* phis() const { return reinterpret_cast<int16_t*>(base) + 8 * (size_t) total; }
  uint32_t id(unsigned i) const { return ids()[offset + i]; }
  float x(unsigned i) const { return xyz()[3 * (size_t)(offset + i)]; }
  float y(unsigned i) const { return xyz()[3 * (size_t)(offset + i) + 1]; }
  float z(unsigned i) const { return xyz()[3 * (size_t)(offset + i) + 2]; }
  int16_t phi(unsigned i) const { return phis()[offset + i]; }
  const int16_t* phi_begin() const { return phis() + offset; }
  void set(unsigned i, uint32_t id_, float x_, float y_, float z_, int16_t phi_) const
  {
    ids()[offset + i] = id_;
    xyz()[3 * (size_t)(offset + i)] = x_;
    xyz()[3 * (size_t)(offset + i) + 1] = y_;
    xyz()[3 * (size_t)(offset + i) + 2] = z_;
    phis()[offset + i] = phi_;
  }
};
constexpr unsigned cluster_element_size = 18;

// Allen event layout (main/src/Transpose.cpp:430-498, main/include/MEPTools.h:25-55), or with
// mep the MEP layout (MEPTools.h:59-200).
struct RawInput {
  const char* raw;
  const uint32_t* offsets;
  const uint32_t* sizes;
  const uint32_t* types;
  int version;
  bool mep = false;
};

struct RawBank {
  uint32_t sourceID;
  uint32_t count;
  const uint32_t* word;
  uint16_t size;
  uint8_t type;
  uint32_t sensor_pair() const { return sourceID & 0x1FFu; }
  uint32_t sensor_index0() const { return (sourceID & 0x1FFu) << 1; }
  uint32_t sensor_index1() const { return ((sourceID & 0x1FFu) << 1) | 1u; }
  uint32_t n_sp(int version) const { return (version == 2 || version == 3) ? count : size / 4u; }
};

// Velo::RawEvent<version, mep_layout> (ClusteringDefinitions.cuh:100-146; MEP::RawEvent,
// MEPTools.h:117-178). Allen layout: per event [u32 n_banks][u32 offsets[n_banks + 1]][banks],
// a bank being [sourceID][fragment]. MEP layout: offsets[0] = n_banks, offsets[2 + b] = the
// source ID of bank b, offsets[2 + n_banks (1 + event) + b] = the byte offset of the event's
// fragment of bank b (MEP::offset_index, MEPTools.h:66-69); sizes and types are indexed by bank,
// then event (MEP::bank_size / bank_type, MEPTools.h:83-104).
struct RawEvent {
  uint32_t n_banks = 0;
  const uint32_t* bank_offsets = nullptr;
  const char* payload = nullptr;
  const uint16_t* bank_sizes = nullptr;
  const uint8_t* bank_types = nullptr;
  int version = 0;
  const RawInput* mep = nullptr;
  unsigned event = 0;
  RawEvent(const RawInput& in, unsigned event_number)
  {
    version = in.version;
    if (in.mep) {
      mep = &in;
      event = event_number;
      n_banks = in.offsets[0];
      return;
    }
    const char* p = in.raw + in.offsets[event_number];
    std::memcpy(&n_banks, p, 4);
    bank_offsets = reinterpret_cast<const uint32_t*>(p + 4);
    payload = p + 4 + 4 * (size_t)(n_banks + 1);
    bank_sizes = reinterpret_cast<const uint16_t*>(in.sizes) + in.sizes[event_number];
    bank_types = reinterpret_cast<const uint8_t*>(in.types) + in.types[event_number];
  }
  RawBank bank(unsigned i) const
  {
    RawBank b;
    const uint32_t* p;
    if (mep) {
      b.sourceID = mep->offsets[2 + i];
      p = reinterpret_cast<const uint32_t*>(mep->raw + mep->offsets[2 + n_banks * (1 + event) + i]);
      b.size = (reinterpret_cast<const uint16_t*>(mep->sizes) + mep->sizes[i])[event];
      b.type = (reinterpret_cast<const uint8_t*>(mep->types) + mep->types[i])[event];
    }
    else {
      p = reinterpret_cast<const uint32_t*>(payload + bank_offsets[i]);
      b.sourceID = p[0];
      p += 1;
      b.size = bank_sizes[i];
      b.type = bank_types[i];
    }
    b.count = 0;
    if (version == 2 || version == 3) {
      b.count = p[0];
      p += 1;
    }
    b.word = p;
    return b;
  }
};

inline bool is_velo_bank(uint8_t t) { return t == bank_type_vp || t == bank_type_velo; }

// ClusteringDefinitions.cuh:200-216 and LHCbID.cuh:38-41
inline uint32_t get_channel_id(unsigned sensor, unsigned chip, unsigned col, unsigned row)
{
  return (sensor << 18) | (chip << 16) | (col << 8) | row;
}
inline uint32_t get_lhcb_id(uint32_t cid) { return (0x2u << 28) | (cid & 0x0FFFFFFFu); }

// VeloTools.cuh:18-50; signselect from BackendCommon.h:183
inline float signselect(float s, float a, float b) { return (s > 0) ? a : b; }
inline float fast_atan2f(float y, float x)
{
  const float c1 = .25f * pi_f;
  const float c2 = .75f * pi_f;
  const float abs_y = __builtin_fabsf(y);
  const float x_plus_y = x + abs_y;
  const float x_sub_y = x - abs_y;
  const float y_sub_x = abs_y - x;
  const float nom = signselect(x, x_sub_y, x_plus_y);
  const float den = signselect(x, x_plus_y, y_sub_x);
  const float r = nom / den;
  const float angle = signselect(x, c1, c2) - c1 * r;
  return __builtin_copysignf(angle, y);
}
// float -> int16 through int32 (what gcc's static_cast<int16_t>(float) emits on x86-64;
// NaN converts to INT32_MIN there, whose low 16 bits are 0, as on the GPU).
inline int16_t f32_to_i16(float f)
{
  int32_t i;
  if (f != f) i = 0;
  else i = static_cast<int32_t>(f);
  return static_cast<int16_t>(i);
}
inline int16_t hit_phi_16(float x, float y) { return f32_to_i16(fast_atan2f(y, x) * convert_factor_i16); }
inline int16_t hit_phi_float_to_16(float phi) { return f32_to_i16(phi * convert_factor_i16); }

// device/utils/binary_search/include/BinarySearch.cuh:15-30
template<typename T>
inline int binary_search_leftmost(const T* array, unsigned array_size, T value, uint64_t* steps = nullptr)
{
  int l = 0;
  int r = array_size;
  while (l < r) {
    const int m = (l + r) / 2;
    if (steps) ++*steps;
    if (value > array[m]) l = m + 1;
    else r = m;
  }
  return l;
}

// Velo::track_offset (VeloEventModel.cuh:192-196)
inline unsigned track_offset(const uint32_t* offsets_estimated, unsigned event_number)
{
  const unsigned offset_event = offsets_estimated[event_number * n_module_pairs];
  return static_cast<unsigned>(static_cast<float>(offset_event) * max_number_of_tracks_per_cluster);
}

// set_arguments_size of velo_search_by_triplet (SearchByTriplet.cu:18-39)
inline unsigned track_container_size(unsigned total_clusters, unsigned n_events)
{
  const float a = static_cast<float>(total_clusters) * max_number_of_tracks_per_cluster;
  const unsigned b = n_events * minimum_container_size;
  return a < static_cast<float>(b) ? b : static_cast<unsigned>(a);
}

// Status bits raised when a fixed-capacity structure of the reference would overflow.
enum StatusBits : uint32_t {
  status_ttf_ring_overflow = 1u << 0,
  status_track_slots_overflow = 1u << 1,
  status_hit_index_overflow = 1u << 2,
  status_three_hit_slots_overflow = 1u << 3,
  status_missing_seed_logit = 1u << 4, // seed pruning: a seed had no logit (it was dropped)
};

} // namespace vo
