// Retina-cluster decoding for gfx950: the clusters come made by the front-end FPGAs, one 32-bit
// word each (bank type 73), and are decoded, keyed by (phi, id) and written sorted per module
// pair (device/velo/retinacluster_decoding/src/DecodeRetinaClusters.cu,
// CalculateNumberOfRetinaClustersPerSensor.cu).
//
//   k_retina_sensor_sizes   one wave per event: the cluster words of every sensor (v2/v3) or
//                           sensor pair (v4) slot, and (first event of the list) the bank of
//                           each slot
//   k_retina_decode         one wave per (event, module pair): decodes the module pair's words,
//                           ranks them by their 64-bit key in LDS and writes every cluster at
//                           its rank: the reference's sorting-key, permutation and sorted-decode
//                           kernels in one pass, with no atomics
//
// Where several words share a key (zero padding words, or identical clusters) they share a rank
// and the slot keeps the last of them in word order, which is what the reference's CPU target
// leaves there.
#include "velo_common.cuh"
#include "velo_kernels.h"

namespace velo {

namespace retina {
constexpr unsigned bank_type = 73; // LHCb::RawBank::VPRetinaCluster (mdf/lhcb/Event/RawBank.h:134)
constexpr unsigned waves_per_block = 4;
constexpr unsigned key_cap = 512;  // keys of a module pair held in the wave's LDS slice

template<int version>
constexpr unsigned slots()
{
  return (version == 2 || version == 3) ? n_sensors : n_sensors / 2;
}

__device__ __forceinline__ void wave_sync_lds()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Decoded {
  bool valid;
  uint32_t id;
  float x, y, z;
  int16_t phi;
  int64_t key;
};

// DecodeRetinaClusters.cu:101-149 (key) and :263-306 (cluster): centroid column and row with
// their fractions (v2: eighths, v3/v4: quarters, the row's bits kept in fy), the LHCbID with
// the fraction flags, the global position and phi; the key is (phi as u16) << 48 | id, or the
// largest key for a zero word.
template<int version>
__device__ __forceinline__ Decoded decode(const Geometry& g, const RawBank& bank, uint32_t word)
{
  Decoded d;
  unsigned sensor;
  if constexpr (version == 2 || version == 3) sensor = bank.sourceID;
  else sensor = (((bank.sourceID & 0x1FFu) << 1) | ((word >> 22) & 0x1u));
  d.valid = bank.type == bank_type && sensor < n_sensors;
  if (!d.valid) sensor = 0;
  const float* ltg = g.ltg + 12 * sensor;
  uint32_t cx, cy, or_fx, or_fy;
  float fx, fy;
  if constexpr (version == 2) {
    cx = (word >> 14) & 0x3FFu;
    fx = ((word >> 11) & 0x7u) / 8.f;
    cy = (word >> 3) & 0xFFu;
    fy = (word & 0x7FFu) / 8.f;
    or_fx = 0;
    or_fy = 0;
  }
  else {
    cx = (word >> 12) & 0x3FFu;
    fx = ((word >> 10) & 0x3u) / 4.f;
    cy = (word >> 2) & 0xFFu;
    fy = (word & 0x3FFu) / 4.f;
    or_fx = ((word >> 11) & 0x1u) | ((word >> 10) & 0x1u);
    or_fy = ((word >> 1) & 0x1u) | (word & 0x1u);
  }
  if (cx >= n_sensor_columns) {
    d.valid = false;
    cx = 0;
  }
  d.id = lhcb_id_of(sensor, cx >> 8, cx & 0xFFu, cy) | (or_fx << 27) | (or_fy << 26);
  const float local_x = g.local_x[cx] + fx * g.x_pitch[cx];
  const float local_y = (0.5f + fy) * pixel_size;
  d.x = ltg[0] * local_x + ltg[1] * local_y + ltg[9];
  d.y = ltg[3] * local_x + ltg[4] * local_y + ltg[10];
  d.z = ltg[6] * local_x + ltg[7] * local_y + ltg[11];
  d.phi = hit_phi_16(d.x, d.y);
  d.key = word == 0 ? 0x7FFFFFFFFFFFFFFFll
                    : static_cast<int64_t>(static_cast<uint64_t>(static_cast<uint16_t>(d.phi)) << 48) | d.id;
  return d;
}
} // namespace retina

using namespace retina;

// CalculateNumberOfRetinaClustersPerSensor.cu:12-53. The last bank of a slot in bank order
// wins, as on the CPU target; slots without a retina bank get 0.
template<int version>
__global__ void __launch_bounds__(wave_size) k_retina_sensor_sizes(
  const char* raw,
  const uint32_t* offsets,
  const uint32_t* sizes,
  const uint32_t* types,
  unsigned mep,
  const uint32_t* event_list,
  unsigned event_start,
  uint32_t* each_sensor_pair_size,
  uint32_t* retina_bank_index)
{
  constexpr unsigned S = slots<version>();
  __shared__ int s_retina[S];
  __shared__ int s_any[S];
  const unsigned l = lane_id();
  const unsigned event_number = event_list ? event_list[blockIdx.x] : blockIdx.x;
  const RawEvent ev(raw, offsets, sizes, types, mep, event_number + event_start);
  for (unsigned i = l; i < S; i += wave_size) {
    s_retina[i] = -1;
    s_any[i] = -1;
  }
  wave_sync_lds();
  for (unsigned i = l; i < ev.n_banks; i += wave_size) {
    const RawBank b = raw_bank<version>(ev, i);
    const unsigned slot = (version == 2 || version == 3) ? b.sourceID : (b.sourceID & 0x1FFu);
    if (b.type == bank_type && slot < S) atomicMax(&s_retina[slot], static_cast<int>(i));
    const unsigned pair = b.sourceID & 0x1FFu;
    if (blockIdx.x == 0 && pair < S) atomicMax(&s_any[pair], static_cast<int>(i));
  }
  wave_sync_lds();
  uint32_t* out = each_sensor_pair_size + static_cast<size_t>(event_number) * S;
  for (unsigned i = l; i < S; i += wave_size) {
    const int bi = s_retina[i];
    out[i] = bi >= 0 ? bank_n_sp<version>(raw_bank<version>(ev, static_cast<unsigned>(bi))) : 0u;
    if (blockIdx.x == 0 && s_any[i] >= 0) retina_bank_index[i] = static_cast<uint32_t>(s_any[i]);
  }
}

// DecodeRetinaClusters.cu:11-395 for one (event, module pair) per wave.
template<int version>
__global__ void __launch_bounds__(waves_per_block* wave_size) k_retina_decode(
  const char* raw,
  const uint32_t* offsets,
  const uint32_t* sizes,
  const uint32_t* types,
  unsigned mep,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  const Geometry* geometry,
  const uint32_t* offsets_each_sensor_pair_size,
  const uint32_t* retina_bank_index,
  const uint32_t* number_of_events,
  uint32_t* module_cluster_num,
  uint32_t* offsets_module_pair_cluster,
  char* cluster_container,
  uint32_t* hit_permutations,
  int64_t* hit_sorting_key,
  unsigned populate_items)
{
  constexpr unsigned S = slots<version>();
  constexpr unsigned step = S / n_module_pairs;
  __shared__ int64_t s_key[waves_per_block][key_cap];
  const unsigned wave = threadIdx.x / wave_size, l = lane_id();
  const unsigned item = blockIdx.x * waves_per_block + wave;
  // populate_module_pair_offsets_and_sizes (:11-36), element by element over every event's
  // module pairs (listed or not); the grid covers max(populate_items, listed items).
  if (l == 0 && item < populate_items) {
    offsets_module_pair_cluster[item + 1] = offsets_each_sensor_pair_size[static_cast<size_t>(item + 1) * step];
    if (item == 0) offsets_module_pair_cluster[0] = 0;
  }
  if (item >= n_list * n_module_pairs) return;
  const unsigned b = item / n_module_pairs, mp = item % n_module_pairs;
  const unsigned event_number = event_list ? event_list[b] : b;
  const uint32_t* sp_off = offsets_each_sensor_pair_size + static_cast<size_t>(event_number) * S;
  const uint32_t lo = sp_off[mp * step], hi = sp_off[(mp + 1) * step];
  const unsigned n = hi - lo;
  const unsigned total = offsets_each_sensor_pair_size[static_cast<size_t>(*number_of_events) * S];
  const Geometry& g = *geometry;
  const RawEvent ev(raw, offsets, sizes, types, mep, event_number + event_start);
  const bool in_lds = n <= key_cap;
  int64_t* keys = s_key[wave];

  // Decode every word of the module pair (its slots' words, in slot and word order).
  auto decode_at = [&](unsigned k) {
    const uint32_t c = lo + k;
    unsigned s = mp * step;
#pragma unroll
    for (unsigned t = 1; t < step; ++t)
      if (sp_off[mp * step + t] <= c) s = mp * step + t;
    // The slot's bank as the first event of the list placed it; a slot whose bank is missing
    // or shorter in this event yields an invalid word instead of a read past the bank.
    const uint32_t bi = retina_bank_index[s];
    const uint32_t w = c - sp_off[s];
    if (bi >= ev.n_banks) {
      Decoded d {};
      d.valid = false;
      return d;
    }
    const RawBank bank = raw_bank<version>(ev, bi);
    if (w >= bank_n_sp<version>(bank)) {
      Decoded d {};
      d.valid = false;
      return d;
    }
    return decode<version>(g, bank, bank.word[w]);
  };
  uint32_t nonzero = 0;
  Decoded first {}; // the lane's first word, kept for the ranking pass
  for (unsigned k = l; k < n; k += wave_size) {
    const Decoded d = decode_at(k);
    if (k == l) first = d;
    // A word of another bank type has no key in the reference (left unset there); it ranks
    // last here and is not written.
    const int64_t key = d.valid ? d.key : 0x7FFFFFFFFFFFFFFFll;
    hit_sorting_key[lo + k] = key;
    if (in_lds) keys[k] = key;
    if (d.valid) nonzero += d.key != 0x7FFFFFFFFFFFFFFFll;
    else hit_permutations[lo + k] = 0;
  }
#pragma unroll
  for (unsigned d = 32; d > 0; d >>= 1) nonzero += __shfl_xor(nonzero, d, wave_size);
  if (l == 0) module_cluster_num[event_number * n_module_pairs + mp] = nonzero;
  if (in_lds) {
    wave_sync_lds();
  }
  else {
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
  }
  const int64_t* kk = in_lds ? keys : hit_sorting_key + lo;

  // Rank by key; the last word of a key's run writes the slot.
  const Clusters out {cluster_container, total};
  for (unsigned i = l; i < n; i += wave_size) {
    const Decoded d = i == l ? first : decode_at(i);
    if (!d.valid) continue;
    // Every lane reads the same keys (n is wave-uniform), eight reads in flight at a time.
    unsigned position = 0;
    bool last = true;
    unsigned j = 0;
    for (; j + 8 <= n; j += 8) {
      int64_t o[8];
#pragma unroll
      for (unsigned u = 0; u < 8; ++u) o[u] = kk[j + u];
#pragma unroll
      for (unsigned u = 0; u < 8; ++u) {
        position += d.key > o[u];
        last = last && !(j + u > i && o[u] == d.key);
      }
    }
    for (; j < n; ++j) {
      const int64_t o = kk[j];
      position += d.key > o;
      last = last && !(j > i && o == d.key);
    }
    hit_permutations[lo + i] = lo + position;
    if (last) {
      const unsigned slot = lo + position;
      out.ids()[slot] = d.id;
      out.xyz()[3 * (size_t) slot] = d.x;
      out.xyz()[3 * (size_t) slot + 1] = d.y;
      out.xyz()[3 * (size_t) slot + 2] = d.z;
      out.phis()[slot] = d.phi;
    }
  }
}

hipError_t launch_retina_sensor_sizes(const RetinaSizesLaunch& a, hipStream_t s)
{
  if (!a.n_list) return hipSuccess;
#define VELO_LAUNCH_RSS(V)                                                                                          \
  hipLaunchKernelGGL(                                                                                               \
    k_retina_sensor_sizes<V>, dim3(a.n_list), dim3(wave_size), 0, s, a.raw, a.offsets, a.sizes, a.types, unsigned(a.mep),            \
    a.event_list, a.event_start, a.each_sensor_pair_size, a.retina_bank_index)
  switch (a.version) {
  case 2: VELO_LAUNCH_RSS(2); break;
  case 3: VELO_LAUNCH_RSS(3); break;
  default: VELO_LAUNCH_RSS(4); break;
  }
#undef VELO_LAUNCH_RSS
  return hipGetLastError();
}

hipError_t launch_retina_decode(const RetinaDecodeLaunch& a, hipStream_t s)
{
  const unsigned items = max(a.n_list * n_module_pairs, a.populate_items);
  if (!items) return hipSuccess;
  const unsigned grid = (items + waves_per_block - 1) / waves_per_block;
#define VELO_LAUNCH_RD(V)                                                                                           \
  hipLaunchKernelGGL(                                                                                               \
    k_retina_decode<V>, dim3(grid), dim3(waves_per_block * wave_size), 0, s, a.raw, a.offsets, a.sizes, a.types, unsigned(a.mep),    \
    a.event_list, a.n_list, a.event_start, reinterpret_cast<const Geometry*>(a.geometry),                         \
    a.offsets_each_sensor_pair_size, a.retina_bank_index, a.number_of_events, a.module_cluster_num,              \
    a.offsets_module_pair_cluster, a.cluster_container, a.hit_permutations, a.hit_sorting_key, a.populate_items)
  switch (a.version) {
  case 2: VELO_LAUNCH_RD(2); break;
  case 3: VELO_LAUNCH_RD(3); break;
  default: VELO_LAUNCH_RD(4); break;
  }
#undef VELO_LAUNCH_RD
  return hipGetLastError();
}

unsigned retina_sensor_slots(int version) { return (version == 2 || version == 3) ? n_sensors : n_sensors / 2; }

} // namespace velo
