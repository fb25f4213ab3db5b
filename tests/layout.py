"""Allen's event layout (main/src/Transpose.cpp:430-498) built from hand-made banks, for known-answer
and adversarial tests of the oracle and the GPU path.

An event is a list of banks (sourceID, words[, type]). Per event the raw block is
[u32 n_banks][u32 offsets[n_banks + 1]][banks], a bank being [u32 sourceID][(v2/v3) u32 count][words];
offsets count bytes from the end of the offset table. The sizes array starts with n_events + 1 u32
offsets in u16 units (the first is 2 (n_events + 1)) followed by one u16 bank size per bank (bytes
after the sourceID); the types array likewise with byte offsets and one u8 type per bank.
"""
from __future__ import annotations

import numpy as np

from allen_amd.synth import Events

VP_BANK_TYPE = 63


def sp_word(sp_col: int, sp_row: int, pixels: int, isolated: bool = False, sensor_bit: int = 0) -> int:
    """Super-pixel word: bit 31 isolated, bit 23 sensor in the pair (v4), bits 22..14 column,
    13..8 row, 7..0 pixels (bit 4 * col + row)."""
    assert 0 <= sp_col < 384 and 0 <= sp_row < 64 and 0 < pixels < 256
    return (int(isolated) << 31) | (sensor_bit << 23) | (sp_col << 14) | (sp_row << 8) | pixels


def make_events(events, version: int = 4) -> Events:
    blocks, sizes16, types8 = [], [], []
    for banks in events:
        words = [len(banks)]
        offsets = []
        payload = []
        ev_sizes, ev_types = [], []
        for bank in banks:
            source, body = bank[0], [int(w) for w in bank[1]]
            btype = bank[2] if len(bank) > 2 else VP_BANK_TYPE
            if version in (2, 3):
                body = [len(body)] + body
            offsets.append(4 * len(payload))
            payload += [source] + body
            ev_sizes.append(4 * len(body))
            ev_types.append(btype)
        offsets.append(4 * len(payload))
        blocks.append(np.array(words + offsets + payload, dtype="<u4"))
        sizes16.append(ev_sizes)
        types8.append(ev_types)
    n = len(events)
    raw = np.concatenate(blocks).view(np.uint8) if blocks else np.zeros(0, np.uint8)
    offsets = np.zeros(n + 1, np.uint32)
    offsets[1:] = np.cumsum([b.nbytes for b in blocks])
    s16 = [s for ev in sizes16 for s in ev]
    t8 = [t for ev in types8 for t in ev]
    sizes = np.zeros((n + 1) + (len(s16) + 1) // 2, np.uint32)
    types = np.zeros((n + 1) + (len(t8) + 3) // 4, np.uint32)
    sz = sizes.view("<u2")
    ty = types.view(np.uint8)
    p16, p8 = 2 * (n + 1), 4 * (n + 1)
    for e in range(n):
        sizes[e], types[e] = p16, p8
        for s in sizes16[e]:
            sz[p16] = s
            p16 += 1
        for t in types8[e]:
            ty[p8] = t
            p8 += 1
    sizes[n], types[n] = p16, p8
    return Events(n_events=n, bank_version=version, raw=raw, offsets=offsets, sizes=sizes, types=types)


def full_event(banks_by_source: dict, version: int = 4):
    """Every VP bank of an event (104 sensor-pair banks for v4, 208 sensor banks for v2), in
    sourceID order, with the given words and empty banks elsewhere (the bank index of the
    clustering comes from the first event, so every bank is present)."""
    n_banks = 104 if version == 4 else 208
    return [(s, list(banks_by_source.get(s, []))) for s in range(n_banks)]


def velo_lhcb_id(module: int, sensor_in_module: int, low: int) -> int:
    """LHCbID of a VELO hit: detector type 2 in bits 28-31, sensor (4 per module) in bits 18-25,
    chip, column and row below (VeloEventModel.cuh, LHCbID.cuh)."""
    return (0x2 << 28) | ((4 * module + sensor_in_module) << 18) | (low & 0x3FFFF)


def sorted_clusters(events, hit_phi_16, gaps=None):
    """Sorted cluster containers (Velo::Clusters layout, x/y/z as float32) from explicit hits.

    events: per event a list of (module pair, x, y, z, LHCbID). Within a module pair the hits are
    ordered by (16-bit phi, LHCbID) as velo_sort_by_phi leaves them. gaps: optional per event
    and module pair extra (unfilled) slots, as the estimate leaves them. Returns
    (container uint8, offsets_estimated_input_size, module_cluster_num)."""
    counts, slots, per_mp = [], [], []
    for e, hits in enumerate(events):
        for m in range(26):
            hs = [(np.float32(x), np.float32(y), np.float32(z), int(i)) for (mp, x, y, z, i) in hits if mp == m]
            hs = [(hit_phi_16(x, y), i, x, y, z) for (x, y, z, i) in hs]
            hs.sort(key=lambda h: (h[0], h[1]))
            per_mp.append(hs)
            counts.append(len(hs))
            slots.append(len(hs) + (int(gaps[e][m]) if gaps is not None else 0))
    offsets = np.zeros(len(slots) + 1, np.uint32)
    offsets[1:] = np.cumsum(slots)
    total = int(offsets[-1])
    ids = np.zeros(total, "<u4")
    xyz = np.zeros((total, 3), "<f4")
    phi = np.zeros(total, "<i2")
    for k, hs in enumerate(per_mp):
        for j, (p, i, x, y, z) in enumerate(hs):
            s = int(offsets[k]) + j
            ids[s], xyz[s], phi[s] = i, (x, y, z), p
    container = np.concatenate([ids.view(np.uint8), xyz.reshape(-1).view(np.uint8), phi.view(np.uint8)])
    return container, offsets, np.asarray(counts, np.uint32)
