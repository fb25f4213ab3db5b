// TEST INFRASTRUCTURE — CPU restatement of Allen's retina-cluster decoding (see common.hpp for
// the rules of use), with CPU-target semantics: one thread per block, blocks in order, atomics
// as plain increments, and the last of several writers to one slot is the one that stays.
//
//   calculate_number_of_retinaclusters_each_sensor_pair
//       device/velo/retinacluster_decoding/src/CalculateNumberOfRetinaClustersPerSensor.cu:12-53
//   decode_retinaclusters
//       device/velo/retinacluster_decoding/src/DecodeRetinaClusters.cu:11-36 (module-pair offsets),
//       :84-186, :188-238 (sorting key), :38-82 (permutations), :240-395 (sorted decode)
#include "oracle.hpp"

namespace vo {

namespace {

// Sensor-pair slots per event: one per sensor (v2/v3) or per sensor pair (v4).
unsigned retina_slots(int version) { return (version == 2 || version == 3) ? n_sensors : n_sensors / 2; }

// BinarySearch.cuh:33-47
int binary_search_rightmost(const uint32_t* array, unsigned array_size, uint32_t value)
{
  int l = 0;
  int r = array_size;
  while (l < r) {
    const int m = (l + r) / 2;
    if (array[m] > value) r = m;
    else l = m + 1;
  }
  return r - 1;
}

// One decoded retina word (DecodeRetinaClusters.cu:101-149 and :263-306).
struct RetinaCluster {
  unsigned sensor;
  uint32_t id;
  float x, y, z;
  int16_t phi;
};

RetinaCluster decode_word(int version, const Geometry& g, const RawBank& bank, uint32_t word)
{
  RetinaCluster c;
  if (version == 2 || version == 3) c.sensor = bank.sourceID;
  else c.sensor = ((word >> 22) & 0x1u) ? bank.sensor_index1() : bank.sensor_index0();
  const float* ltg = g.ltg + 12 * c.sensor;
  uint32_t cx, cy, or_fx, or_fy;
  float fx, fy;
  if (version == 2) {
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
    // The row's bits stay in: fy is the row plus its fraction, in quarter pixels.
    fy = (word & 0x3FFu) / 4.f;
    or_fx = ((word >> 11) & 0x1u) | ((word >> 10) & 0x1u);
    or_fy = ((word >> 1) & 0x1u) | (word & 0x1u);
  }
  const uint32_t chip = cx >> 8;
  const uint32_t cid =
    (or_fx << 27) | (or_fy << 26) | get_channel_id(c.sensor, chip, cx & 0xFFu, cy);
  c.id = get_lhcb_id(cid);
  const float local_x = g.local_x[cx] + fx * g.x_pitch[cx];
  const float local_y = (0.5f + fy) * pixel_size;
  c.x = ltg[0] * local_x + ltg[1] * local_y + ltg[9];
  c.y = ltg[3] * local_x + ltg[4] * local_y + ltg[10];
  c.z = ltg[6] * local_x + ltg[7] * local_y + ltg[11];
  c.phi = hit_phi_16(c.x, c.y);
  return c;
}

// DecodeRetinaClusters.cu:176-185: (phi as u16) << 48 | id, or the largest key for a zero word.
int64_t sorting_key(uint32_t word, const RetinaCluster& c)
{
  if (word == 0) return 0x7FFFFFFFFFFFFFFFll;
  return static_cast<int64_t>(static_cast<uint64_t>(static_cast<uint16_t>(c.phi)) << 48) | c.id;
}

} // namespace

unsigned retina_sensor_slots(int version) { return retina_slots(version); }

void calculate_number_of_retinaclusters_each_sensor_pair(
  const RawInput& in,
  const uint32_t* event_list,
  unsigned n_list,
  unsigned event_start,
  uint32_t* each_sensor_pair_size,
  uint32_t* retina_bank_index)
{
  const unsigned slots = retina_slots(in.version);
  for (unsigned b = 0; b < n_list; ++b) {
    const unsigned event_number = event_list[b];
    uint32_t* size = each_sensor_pair_size + (size_t) event_number * slots;
    const RawEvent ev(in, event_number + event_start);
    for (unsigned i = 0; i < ev.n_banks; ++i) {
      const RawBank bank = ev.bank(i);
      if (bank.type == bank_type_retina) {
        if (in.version == 2 || in.version == 3) size[bank.sourceID] = bank.count;
        else size[bank.sensor_pair()] = bank.size / 4u;
      }
      if (b == 0) retina_bank_index[bank.sensor_pair()] = i;
    }
  }
}

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
  int64_t* hit_sorting_key)
{
  const unsigned slots = retina_slots(in.version);
  const unsigned step = slots / n_module_pairs;
  const unsigned total = offsets_each_sensor_pair_size[(size_t) n_events * slots];
  std::vector<uint32_t> zero_num((size_t) n_events * n_module_pairs, 0);
  // The reference clears these three before its kernels.
  std::fill(module_pair_cluster_num, module_pair_cluster_num + (size_t) n_events * n_module_pairs, 0u);
  std::fill(offsets_module_pair_cluster, offsets_module_pair_cluster + (size_t) n_events * n_module_pairs + 1, 0u);
  std::fill(hit_permutations, hit_permutations + total, 0u);

  // populate_module_pair_offsets_and_sizes (:11-36)
  for (unsigned e = 0; e < n_events * n_module_pairs; ++e)
    offsets_module_pair_cluster[e + 1] = offsets_each_sensor_pair_size[(size_t)(e + 1) * step];

  // velo_calculate_sorting_key (:188-238)
  for (unsigned b = 0; b < n_list; ++b) {
    const unsigned event_number = event_list[b];
    const uint32_t* sp_off = offsets_each_sensor_pair_size + (size_t) event_number * slots;
    const uint32_t start = sp_off[0];
    const uint32_t n = sp_off[slots] - start;
    const RawEvent ev(in, event_number + event_start);
    for (uint32_t k = 0; k < n; ++k) {
      const unsigned sensor_pair = binary_search_rightmost(sp_off, slots, k + start);
      const RawBank bank = ev.bank(retina_bank_index[sensor_pair]);
      if (bank.type != bank_type_retina) continue;
      const uint32_t word = bank.word[k - (sp_off[sensor_pair] - start)];
      const RetinaCluster c = decode_word(in.version, g, bank, word);
      const unsigned mp = event_number * n_module_pairs + c.sensor / 8;
      if (word != 0) ++module_pair_cluster_num[mp];
      else ++zero_num[mp];
      hit_sorting_key[start + k] = sorting_key(word, c);
    }
  }

  // velo_calculate_permutations (:38-82)
  for (unsigned b = 0; b < n_list; ++b) {
    const unsigned event_number = event_list[b];
    for (unsigned mp = 0; mp < n_module_pairs; ++mp) {
      const unsigned idx = event_number * n_module_pairs + mp;
      const uint32_t hit_start = offsets_module_pair_cluster[idx];
      const uint32_t hit_num = module_pair_cluster_num[idx] + zero_num[idx];
      for (uint32_t i = 0; i < hit_num; ++i) {
        const int64_t key = hit_sorting_key[hit_start + i];
        uint32_t position = 0;
        for (uint32_t j = 0; j < hit_num; ++j) {
          if (i == j) continue;
          position += key > hit_sorting_key[hit_start + j];
        }
        hit_permutations[hit_start + i] = hit_start + position;
      }
    }
  }

  // decode_retinaclusters_sorted (:308-395)
  Clusters out {cluster_container, total, 0};
  for (unsigned b = 0; b < n_list; ++b) {
    const unsigned event_number = event_list[b];
    const uint32_t* sp_off = offsets_each_sensor_pair_size + (size_t) event_number * slots;
    const uint32_t start = sp_off[0];
    const uint32_t n = sp_off[slots] - start;
    const RawEvent ev(in, event_number + event_start);
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t cluster_number = start + k;
      const unsigned sensor_pair = binary_search_rightmost(sp_off, slots, cluster_number);
      const RawBank bank = ev.bank(retina_bank_index[sensor_pair]);
      if (bank.type != bank_type_retina) continue;
      const uint32_t word = bank.word[cluster_number - sp_off[sensor_pair]];
      const RetinaCluster c = decode_word(in.version, g, bank, word);
      out.set(hit_permutations[cluster_number], c.id, c.x, c.y, c.z, c.phi);
    }
  }
}

} // namespace vo
