"""Independent restatements used to pin the oracle: bank decoding to pixels and a textbook
8-connected flood fill (scipy.ndimage.label) with centroid positions in float64."""
from __future__ import annotations

import numpy as np
from scipy import ndimage

EIGHT = np.ones((3, 3), dtype=bool)


def iter_banks(raw, offsets, sizes, types, version, event):
    """(sourceID, words) of each VP/Velo bank of an event in Allen's layout."""
    base = int(offsets[event])
    n_banks = int(raw[base:base + 4].view("<u4")[0])
    bank_off = raw[base + 4:base + 4 + 4 * (n_banks + 1)].view("<u4")
    payload = base + 4 + 4 * (n_banks + 1)
    s16 = sizes.view("<u2")[int(sizes[event]):int(sizes[event]) + n_banks]
    t8 = types.view(np.uint8)[int(types[event]):int(types[event]) + n_banks]
    for b in range(n_banks):
        if t8[b] not in (8, 63):
            continue
        p = payload + int(bank_off[b])
        source = int(raw[p:p + 4].view("<u4")[0])
        if version in (2, 3):
            count = int(raw[p + 4:p + 8].view("<u4")[0])
            words = raw[p + 8:p + 8 + 4 * count].view("<u4")
        else:
            words = raw[p + 4:p + 4 + int(s16[b])].view("<u4")
        yield source, words


def event_pixels(ev, event):
    """{sensor: set((col, row))} of one event."""
    out = {}
    for source, words in iter_banks(ev.raw, ev.offsets, ev.sizes, ev.types, ev.bank_version, event):
        for w in words.tolist():
            if ev.bank_version in (2, 3):
                sensor = source & 0x1FF
            else:
                sensor = ((source & 0x1FF) << 1) | ((w >> 23) & 1)
            sp_col = (w >> 14) & 0x1FF
            sp_row = (w >> 8) & 0x3F
            for bit in range(8):
                if (w >> bit) & 1:
                    out.setdefault(sensor, set()).add((2 * sp_col + bit // 4, 4 * sp_row + bit % 4))
    return out


def lhcb_id(sensor, col, row):
    return (0x2 << 28) | (sensor << 18) | ((col >> 8) << 16) | ((col & 0xFF) << 8) | row


def components(pixels: set):
    """Connected components (8-connectivity) of a sensor's pixels: list of (cols, rows) arrays."""
    if not pixels:
        return []
    px = np.array(sorted(pixels))
    grid = np.zeros((768, 256), dtype=np.int32)
    grid[px[:, 0], px[:, 1]] = 1
    lab, n = ndimage.label(grid, structure=EIGHT)
    labels = lab[px[:, 0], px[:, 1]]
    return [(px[labels == k, 0], px[labels == k, 1]) for k in range(1, n + 1)]


def cluster_position(geometry, sensor, cols, rows):
    """Centroid channel and global position in float64 (fractional centroid within the pixel)."""
    n = cols.size
    sx, sy = int(cols.sum()), int(rows.sum())
    cx, cy = sx // n, sy // n
    fx, fy = sx / n - cx, sy / n - cy
    lx = float(geometry["local_x"][cx]) + fx * float(geometry["x_pitch"][cx])
    ly = (cy + 0.5 + fy) * 0.055
    m = geometry["ltg"][sensor].astype(np.float64)
    g = np.array([m[0] * lx + m[1] * ly + m[9], m[3] * lx + m[4] * ly + m[10], m[6] * lx + m[7] * ly + m[11]])
    return lhcb_id(sensor, cx, cy), g


def parse_geometry(blob: bytes):
    """velo_geometry.bin: [u64 n][f32 x n] for module z, local x, x pitch, then [u64 208][u64 12][f32 ...]."""
    a = np.frombuffer(blob, np.uint8)
    p = 0
    out = {}
    for key in ["module_zs", "local_x", "x_pitch"]:
        n = int(a[p:p + 8].view("<u8")[0])
        out[key] = a[p + 8:p + 8 + 4 * n].view("<f4")
        p += 8 + 4 * n
    n_ltg, n_trans = a[p:p + 16].view("<u8")
    out["ltg"] = a[p + 16:p + 16 + 4 * int(n_ltg * n_trans)].view("<f4").reshape(int(n_ltg), int(n_trans))
    return out


def load_mdf_fixture(path):
    from allen_amd.synth import Events
    z = np.load(path)
    return Events(n_events=int(z["n_events"]), bank_version=int(z["version"]), raw=z["raw"],
                  offsets=z["offsets"], sizes=z["sizes"], types=z["types"])


def iter_retina_banks(raw, offsets, sizes, types, version, event):
    """(sourceID, words) of each retina cluster bank (type 73) of an event in Allen's layout."""
    base = int(offsets[event])
    n_banks = int(raw[base:base + 4].view("<u4")[0])
    bank_off = raw[base + 4:base + 4 + 4 * (n_banks + 1)].view("<u4")
    payload = base + 4 + 4 * (n_banks + 1)
    s16 = sizes.view("<u2")[int(sizes[event]):int(sizes[event]) + n_banks]
    t8 = types.view(np.uint8)[int(types[event]):int(types[event]) + n_banks]
    for b in range(n_banks):
        if t8[b] != 73:
            continue
        p = payload + int(bank_off[b])
        source = int(raw[p:p + 4].view("<u4")[0])
        if version in (2, 3):
            count = int(raw[p + 4:p + 8].view("<u4")[0])
            words = raw[p + 8:p + 8 + 4 * count].view("<u4")
        else:
            words = raw[p + 4:p + 4 + int(s16[b])].view("<u4")
        yield source, words


def retina_clusters(ev, geometry, event):
    """Every non-zero retina word of an event as (sensor, LHCbID, float64 global position): the
    centroid column and row with their quarter-pixel fractions (v3/v4 word layout: column
    21..12, its fraction 11..10, row 9..2, its fraction 1..0; v4: bit 22 = sensor in pair);
    the LHCbID carries the fraction flags in bits 27 (column) and 26 (row)."""
    out = []
    for source, words in iter_retina_banks(ev.raw, ev.offsets, ev.sizes, ev.types, ev.bank_version, event):
        for w in words.tolist():
            if w == 0:
                continue
            sensor = source if ev.bank_version in (2, 3) else (((source & 0x1FF) << 1) | ((w >> 22) & 1))
            cx, qx, cy, qy = (w >> 12) & 0x3FF, (w >> 10) & 3, (w >> 2) & 0xFF, w & 3
            lx = float(geometry["local_x"][cx]) + qx / 4 * float(geometry["x_pitch"][cx])
            ly = (cy + qy / 4 + 0.5) * 0.055
            m = geometry["ltg"][sensor].astype(np.float64)
            g = np.array([m[0] * lx + m[1] * ly + m[9], m[3] * lx + m[4] * ly + m[10], m[6] * lx + m[7] * ly + m[11]])
            lhcb = lhcb_id(sensor, cx, cy) | (int(qx != 0) << 27) | (int(qy != 0) << 26)
            out.append((sensor, lhcb, g))
    return out
