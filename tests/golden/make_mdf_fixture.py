"""Writes tests/golden/mdf_minbias_vp.npz and mdf_minbias_retina.npz: the VELO super-pixel raw
banks (type 63, VP) and the retina cluster banks (type 73, VPRetinaCluster) of the 10 events of
the reference's bundled MDF file, each in Allen's event layout.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_mdf_fixture.py
The MDF record layout read here: MDFHeader (3 x u32 size, u32 checksum, u8 compression,
u8 hdr (version << 4 | length), u8 data type, u8 spare) followed by a sub-header whose size
depends on the header version (mdf/include/mdf_header.hpp:45-160), then raw banks
[u16 magic 0xCBCB][u16 length incl. 8 B header][u8 type][u8 version][u16 sourceID][data],
each padded to 4 bytes (mdf/lhcb/Event/RawBank.h:196-250).
"""
import os
import struct
import sys

import numpy as np

SRC = "/root/reference/input/minbias/mdf/MiniBrunel_2018_MinBias_FTv4_DIGI_retinacluster_v1.mdf"
HERE = os.path.dirname(os.path.abspath(__file__))
OUTPUTS = {63: os.path.join(HERE, "mdf_minbias_vp.npz"), 73: os.path.join(HERE, "mdf_minbias_retina.npz")}
SUBHEADER = {0: 20 + 52 - 2, 1: 20 + 28, 2: 20, 3: 20 + 28}


def read_events(path):
    data = open(path, "rb").read()
    pos = 0
    events = []
    while pos + 20 <= len(data):
        size0, size1, size2, _chk = struct.unpack_from("<4I", data, pos)
        compression, hdr, _dtype, _spare = struct.unpack_from("<4B", data, pos + 16)
        if compression != 0:
            raise ValueError("compressed MDF records are not supported")
        hlen = SUBHEADER[hdr >> 4]
        body = data[pos + hlen:pos + size0]
        banks = []
        p = 0
        while p + 8 <= len(body):
            magic, length, btype, version, source = struct.unpack_from("<HHBBH", body, p)
            if magic != 0xCBCB:
                raise ValueError(f"bad bank magic at {pos + hlen + p}")
            payload = body[p + 8:p + length]
            banks.append((btype, version, source, payload))
            p += (length + 3) // 4 * 4
        events.append(banks)
        pos += size0
    return events


def allen_layout(events, bank_type):
    """Allen layout of the banks of one type (main/src/Transpose.cpp:430-498), sorted by sourceID."""
    blocks, sizes16, types8, offsets, sizes_off, types_off = [], [], [], [0], [], []
    versions = set()
    n = len(events)
    s16 = 2 * (n + 1)
    t8 = 4 * (n + 1)
    for banks in events:
        sel = sorted([b for b in banks if b[0] == bank_type], key=lambda b: b[2])
        versions |= {b[1] for b in sel}
        words = [len(sel)]
        offs = []
        o = 0
        body = []
        for btype, version, source, payload in sel:
            offs.append(o)
            padded = payload + b"\0" * ((-len(payload)) % 4)
            body += [source] + list(np.frombuffer(padded, "<u4"))
            o += 4 + len(padded)
            sizes16.append(len(payload))
            types8.append(btype)
        offs.append(o)
        block = np.array(words + offs + body, dtype="<u4")
        blocks.append(block)
        offsets.append(offsets[-1] + 4 * block.size)
        sizes_off.append(s16)
        types_off.append(t8)
        s16 += len(sel)
        t8 += len(sel)
    sizes_off.append(s16)
    types_off.append(t8)
    raw = np.concatenate(blocks).view(np.uint8)
    sz = np.zeros(n + 1 + (len(sizes16) + 1) // 2, np.uint32)
    sz[:n + 1] = sizes_off
    sz.view("<u2")[2 * (n + 1):2 * (n + 1) + len(sizes16)] = sizes16
    ty = np.zeros(n + 1 + (len(types8) + 3) // 4, np.uint32)
    ty[:n + 1] = types_off
    ty.view(np.uint8)[4 * (n + 1):4 * (n + 1) + len(types8)] = types8
    assert len(versions) == 1, versions
    return raw, np.array(offsets, np.uint32), sz, ty, versions.pop()


if __name__ == "__main__":
    evs = read_events(sys.argv[1] if len(sys.argv) > 1 else SRC)
    for bank_type, out in OUTPUTS.items():
        raw, offsets, sizes, types, version = allen_layout(evs, bank_type)
        np.savez_compressed(out, raw=raw, offsets=offsets, sizes=sizes, types=types, version=np.int32(version),
                            n_events=np.int32(len(evs)))
        print(f"{len(evs)} events, bank type {bank_type} version {version}, raw {raw.size} bytes -> {out}")
