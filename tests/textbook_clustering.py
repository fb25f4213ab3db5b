"""Pixel-coordinate statement of the super-pixel estimate and masked clustering, written from what
the reference's bit arithmetic means rather than from the arithmetic itself, to pin the oracle:

* velo_estimate_input_size (EstimateInputSize.cu:10-212): an isolated super pixel counts its
  8-connected components; every pixel of a non-isolated super pixel whose up, up-right, right and
  down-right neighbours are all empty is a candidate (at most one per lower and upper half of the
  super pixel). The neighbours outside the super pixel are looked up in every word of the bank,
  whatever its sensor or isolation flag (:111-141 test the current word's flag at :113 and no
  sensor bit). A candidate is encoded (word index << 11) | (sensor << 3) | (7 - pixel bit).
* rest_of_clusters (MaskedVeloClustering.cu:290-496): from a candidate, a directed 8-connected
  growth inside the four pixel columns of its super pixel and the one to its left, over the
  same-sensor non-isolated pixels of a window of 4 super-pixel rows x 5 super-pixel columns. The
  directed links are those of a 16-row window whose up-right step is missing from its two lowest
  rows and whose down-left step is missing from its two highest rows (from the highest row the
  up-right step lands in the lowest row two columns over, from the lowest row the down-left step
  in the highest row two columns back). The cluster is kept when the candidate is its right-most,
  then top-most pixel, no pixel of the column to the east touches its right-most column, and it
  has no pixel in the window's lowest or highest row; when it reaches the left-most of the four
  columns it grows on into the window's four west columns from their neighbours there, and the
  extension must not reach the window's highest row or its west-most column (its lowest row is
  not checked).
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

EIGHT = np.ones((3, 3), dtype=bool)


def decode(source: int, words, version: int):
    """Each word of a VP bank as (index, sensor, sp_col, sp_row, pixel byte, isolated, sensor bit)."""
    out = []
    for i, w in enumerate(int(x) for x in words):
        bit = (w >> 23) & 1
        sensor = (source & 0x1FF) if version in (2, 3) else (((source & 0x1FF) << 1) | bit)
        out.append((i, sensor, (w >> 14) & 0x1FF, (w >> 8) & 0x3F, w & 0xFF, bool(w >> 31), bit))
    return out


def module_pair(source: int, version: int) -> int:
    return (source & 0x1FF) // 8 if version in (2, 3) else ((source & 0x1FF) << 1) // 8


def sp_pixels(sp_col, sp_row, byte):
    """Absolute (column, row) of each set pixel of a super pixel (bit 4 * col + row)."""
    return [(2 * sp_col + b // 4, 4 * sp_row + b % 4) for b in range(8) if byte >> b & 1]


def sp_components(byte):
    """8-connected components of a super pixel's own pixels, each as a list of (col, row) in the SP."""
    grid = np.zeros((2, 4), dtype=np.int32)
    for b in range(8):
        if byte >> b & 1:
            grid[b // 4, b % 4] = 1
    lab, n = ndimage.label(grid, structure=EIGHT)
    return [list(zip(*np.nonzero(lab == k))) for k in range(1, n + 1)]


def estimate_bank(source, words, version):
    """(estimate of the bank's module pair, candidate words) of one bank."""
    dec = decode(source, words, version)
    by_addr = {}
    for (_, _, c, r, byte, _, _) in dec:
        by_addr[(c, r)] = by_addr.get((c, r), 0) | byte
    count, cands = 0, []
    for (i, sensor, c, r, byte, isolated, _) in dec:
        if isolated:
            count += len(sp_components(byte))
            continue

        def occupied(pc, pr):
            sc, sr = pc // 2, pr // 4
            if pr < 0:
                return False
            b = 4 * (pc % 2) + pr % 4
            src = byte if (sc, sr) == (c, r) else by_addr.get((sc, sr), 0)
            return bool(src >> b & 1)

        for half in (range(0, 2), range(2, 4)):
            found = []
            for lc in range(2):
                for lr in half:
                    if not byte >> (4 * lc + lr) & 1:
                        continue
                    pc, pr = 2 * c + lc, 4 * r + lr
                    if not any(occupied(pc + dc, pr + dr) for dc, dr in ((0, 1), (1, 1), (1, 0), (1, -1))):
                        found.append(4 * lc + lr)
            assert len(found) <= 1
            for b in found:
                cands.append((i << 11) | (sensor << 3) | (7 - b))
                count += 1
    return count, cands


def _grow(start, occupied, cols):
    """Directed growth from the pixels `start` (local (col, row) of the 16-row window) over the
    occupied local pixels within the column range `cols`."""
    lo, hi = cols
    cluster = set(p for p in start if p in occupied)
    todo = list(cluster)
    while todo:
        pc, lr = todo.pop()
        steps = []
        if lr < 15:
            steps += [(pc, lr + 1), (pc - 1, lr + 1)]
        if lr > 0:
            steps += [(pc, lr - 1), (pc + 1, lr - 1)]
        steps += [(pc + 1, lr), (pc - 1, lr)]
        if 2 <= lr <= 14:
            steps.append((pc + 1, lr + 1))
        if lr == 15:
            steps.append((pc + 2, 0))
        if 1 <= lr <= 13:
            steps.append((pc - 1, lr - 1))
        if lr == 0:
            steps.append((pc - 2, 15))
        for q in steps:
            if lo <= q[0] <= hi and q in occupied and q not in cluster:
                cluster.add(q)
                todo.append(q)
    return cluster


def rest_of_clusters(dec, candidate, version):
    """The cluster kept for one candidate as (sensor, sum of columns, sum of rows, pixels), or None."""
    i, sensor, loc = candidate >> 11, (candidate >> 3) & 0xFF, candidate & 7
    _, _, sc, sr, _, _, _ = dec[i]
    w0, r0 = 2 * (sc - 3), 4 * (sr - 2)
    occupied = set()
    for (_, s, c, r, byte, isolated, bit) in dec:
        if version > 3 and bit != (sensor & 1):
            continue
        if isolated or not (sc - 3 <= c <= sc + 1 and sr - 2 <= r <= sr + 1):
            continue
        for (pc, pr) in sp_pixels(c, r, byte):
            occupied.add((pc - w0, pr - r0))
    b = 7 - loc
    start = (2 * sc + b // 4 - w0, 4 * sr + b % 4 - r0)
    main = _grow([start], occupied, (4, 7))
    key = lambda p: (p[0], p[1])
    keep = all(key(p) <= key(start) for p in main)
    east = {(8, p[1] + d) for p in main if p[0] == 7 for d in (-1, 0, 1)}
    keep &= not (east & occupied)
    keep &= not any(p[1] in (0, 15) for p in main)
    pixels = set(main)
    if keep and any(p[0] == 4 for p in main):
        seeds = {(3, p[1] + d) for p in main if p[0] == 4 for d in (-1, 0, 1) if 0 <= p[1] + d <= 15}
        ext = _grow(sorted(seeds), occupied, (0, 3))
        keep &= not any(p[1] == 15 for p in ext) and not any(p[0] == 0 for p in ext)
        pixels |= ext
    if not keep:
        return None
    x = sum(p[0] + w0 for p in pixels)
    y = sum(p[1] + r0 for p in pixels)
    return sensor, x, y, len(pixels)


def cluster_of_sum(geometry, sensor, x, y, n):
    """LHCbID and float64 global position of a cluster from its column and row sums."""
    cx, cy = x // n, y // n
    fx, fy = x / n - cx, y / n - cy
    lx = float(geometry["local_x"][cx]) + fx * float(geometry["x_pitch"][cx])
    ly = (cy + 0.5 + fy) * 0.055
    m = geometry["ltg"][sensor].astype(np.float64)
    g = np.array([m[0] * lx + m[1] * ly + m[9], m[3] * lx + m[4] * ly + m[10], m[6] * lx + m[7] * ly + m[11]])
    lhcb = (0x2 << 28) | (sensor << 18) | ((cx >> 8) << 16) | ((cx & 0xFF) << 8) | cy
    return lhcb, g


def event(banks, version, geometry):
    """Per module pair: the estimate, the candidates and the clusters (LHCbID, float64 position) of
    one event's VP banks [(sourceID, words)]."""
    est = np.zeros(26, np.int64)
    cands, clusters = {}, {}
    for source, words in banks:
        mp = module_pair(source, version)
        n, cl = estimate_bank(source, words, version)
        est[mp] += n
        cands.setdefault(mp, []).extend(cl)
        dec = decode(source, words, version)
        for (_, sensor, c, r, byte, isolated, _) in dec:
            if not isolated:
                continue
            for comp in sp_components(byte):
                x = sum(2 * c + lc for lc, _ in comp)
                y = sum(4 * r + lr for _, lr in comp)
                clusters.setdefault(mp, []).append(cluster_of_sum(geometry, sensor, x, y, len(comp)))
        for cand in cl:
            got = rest_of_clusters(dec, cand, version)
            if got is not None:
                clusters.setdefault(mp, []).append(cluster_of_sum(geometry, *got))
    return est, cands, clusters
