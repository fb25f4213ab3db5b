// MDF ingestion for the VELO path (host code): reads the raw banks of one type from an MDF file
// and lays them out as Allen's event layout, the input of velo_sequence_run and of the single
// algorithms.
//
// The record format (mdf/include/mdf_header.hpp:45-220, packed structs): a 20-byte MDFHeader
// (u32 size x 3, u32 checksum, u8 compression, u8 hdr = version << 4 | sub-header words, u8 data
// type, u8 spare), the sub-header of its version (sizeOf: v0 40 bytes in all, v1/v3 48, v2 20),
// then the raw banks, each [u16 magic 0xCBCB][u16 length with the 8-byte header][u8 type]
// [u8 version][u16 sourceID][payload], padded to 4 bytes (mdf/lhcb/Event/RawBank.h:196-250).
// The layout (main/src/Transpose.cpp:220-498): per event [u32 n_banks][u32 offsets[n_banks+1]]
// then each bank as [u32 sourceID][payload padded to 4 bytes], banks sorted by sourceID; the
// sizes array holds n+1 u32 offsets in u16 units then the u16 payload sizes, the types array n+1
// u32 offsets in bytes then the u8 types. As in Allen's provider (main/src/Provider.cpp:240-248)
// only one of the two VELO encodings is taken: super pixels (type 63) or retina clusters (73).
#include "velo_amd.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct velo_mdf_events {
  unsigned n_events = 0;
  int bank_version = -1;
  std::vector<uint32_t> raw, offsets, sizes, types;
};

namespace {

unsigned mdf_header_size(unsigned version)
{
  switch (version) {
  case 0: return 20 + 22 - 2; // Header0: u8 event type, u8 trigger high, u32 trigger low, u32 mask x 4
  case 1:
  case 3: return 20 + 28; // Header1: u32 mask x 4, run, orbit, bunch
  case 2: return 20;
  default: return 0;
  }
}

struct Bank {
  unsigned type, version, source;
  const unsigned char* payload;
  unsigned bytes;
};

} // namespace

extern "C" {

int velo_mdf_read(const char* path, unsigned bank_type, unsigned max_events, velo_mdf_events** out)
{
  if (!path || !out) return VELO_ERR_INVALID_ARGUMENT;
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return VELO_ERR_IO;
  std::vector<unsigned char> data;
  unsigned char chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) data.insert(data.end(), chunk, chunk + n);
  const bool read_error = std::ferror(f) != 0;
  std::fclose(f);
  if (read_error) return VELO_ERR_IO;

  auto* ev = new velo_mdf_events;
  std::vector<std::vector<Bank>> events;
  size_t pos = 0;
  int rc = VELO_OK;
  while (pos + 20 <= data.size() && (max_events == 0 || events.size() < max_events)) {
    uint32_t size0;
    std::memcpy(&size0, data.data() + pos, 4);
    const unsigned compression = data[pos + 16], hdr = data[pos + 17];
    const unsigned hlen = mdf_header_size(hdr >> 4);
    if (hlen == 0 || size0 < hlen || pos + size0 > data.size()) {
      rc = VELO_ERR_IO;
      break;
    }
    if (compression & 0xF) {
      rc = VELO_ERR_UNSUPPORTED_LAYOUT; // compressed records are not read
      break;
    }
    std::vector<Bank> banks;
    size_t p = pos + hlen;
    const size_t end = pos + size0;
    while (p + 8 <= end) {
      uint16_t magic, length, source;
      std::memcpy(&magic, data.data() + p, 2);
      std::memcpy(&length, data.data() + p + 2, 2);
      std::memcpy(&source, data.data() + p + 6, 2);
      if (magic != 0xCBCB || length < 8 || p + length > end) {
        rc = VELO_ERR_IO;
        break;
      }
      const unsigned type = data[p + 4], version = data[p + 5];
      if (type == bank_type) banks.push_back({type, version, source, data.data() + p + 8, length - 8u});
      p += (length + 3u) & ~3u;
    }
    if (rc != VELO_OK) break;
    std::stable_sort(banks.begin(), banks.end(), [](const Bank& a, const Bank& b) { return a.source < b.source; });
    for (const Bank& b : banks) {
      if (ev->bank_version < 0) ev->bank_version = static_cast<int>(b.version);
      else if (ev->bank_version != static_cast<int>(b.version)) rc = VELO_ERR_BANK_VERSION;
    }
    if (rc != VELO_OK) break;
    events.push_back(std::move(banks));
    pos += size0;
  }
  if (rc != VELO_OK) {
    delete ev;
    return rc;
  }

  // Allen's layout.
  const unsigned ne = static_cast<unsigned>(events.size());
  ev->n_events = ne;
  size_t total_banks = 0;
  for (const auto& e : events) total_banks += e.size();
  ev->sizes.assign((ne + 1) + (total_banks + 1) / 2, 0);
  ev->types.assign((ne + 1) + (total_banks + 3) / 4, 0);
  ev->offsets.assign(ne + 1, 0);
  auto* s16 = reinterpret_cast<uint16_t*>(ev->sizes.data());
  auto* t8 = reinterpret_cast<uint8_t*>(ev->types.data());
  size_t si = 2 * (ne + 1), ti = 4 * (ne + 1);
  for (unsigned e = 0; e < ne; ++e) {
    const auto& banks = events[e];
    ev->offsets[e] = static_cast<uint32_t>(4 * ev->raw.size());
    ev->sizes[e] = static_cast<uint32_t>(si);
    ev->types[e] = static_cast<uint32_t>(ti);
    ev->raw.push_back(static_cast<uint32_t>(banks.size()));
    const size_t off_pos = ev->raw.size();
    ev->raw.resize(ev->raw.size() + banks.size() + 1, 0);
    uint32_t o = 0;
    for (size_t b = 0; b < banks.size(); ++b) {
      ev->raw[off_pos + b] = o;
      ev->raw.push_back(banks[b].source);
      const size_t words = (banks[b].bytes + 3) / 4;
      const size_t at = ev->raw.size();
      ev->raw.resize(at + words, 0);
      std::memcpy(ev->raw.data() + at, banks[b].payload, banks[b].bytes);
      o += static_cast<uint32_t>(4 * (1 + words));
      s16[si++] = static_cast<uint16_t>(banks[b].bytes);
      t8[ti++] = static_cast<uint8_t>(banks[b].type);
    }
    ev->raw[off_pos + banks.size()] = o;
  }
  ev->offsets[ne] = static_cast<uint32_t>(4 * ev->raw.size());
  ev->sizes[ne] = static_cast<uint32_t>(si);
  ev->types[ne] = static_cast<uint32_t>(ti);
  *out = ev;
  return VELO_OK;
}

int velo_mdf_events_info(const velo_mdf_events* e, unsigned* n_events, int* bank_version, size_t* raw_bytes)
{
  if (!e) return VELO_ERR_INVALID_ARGUMENT;
  if (n_events) *n_events = e->n_events;
  if (bank_version) *bank_version = e->bank_version;
  if (raw_bytes) *raw_bytes = 4 * e->raw.size();
  return VELO_OK;
}

int velo_mdf_events_buffer(const velo_mdf_events* e, const char* name, const void** host_ptr, size_t* bytes)
{
  if (!e || !name || !host_ptr || !bytes) return VELO_ERR_INVALID_ARGUMENT;
  const std::string s(name);
  const std::vector<uint32_t>* v = s == "raw"       ? &e->raw
                                   : s == "offsets" ? &e->offsets
                                   : s == "sizes"   ? &e->sizes
                                   : s == "types"   ? &e->types
                                                    : nullptr;
  if (!v) return VELO_ERR_INVALID_ARGUMENT;
  *host_ptr = v->data();
  *bytes = 4 * v->size();
  return VELO_OK;
}

void velo_mdf_events_free(velo_mdf_events* e) { delete e; }

} // extern "C"
