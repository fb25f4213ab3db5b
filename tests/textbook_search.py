"""Statement of velo_search_by_triplet (SearchByTriplet.cu:110-609) from its rules, to pin the
oracle's restatement; it works on one event's sorted clusters.

Seeding, at each step: for every unused hit h1 of the current module pair (in hit order), take up
to K unused hits h0 of the previous module pair in pendulum order (the leftmost phi insertion
point of h1, then +1, -1, +2, -2, ... around the ring of that pair's hits; K = 8 at the first
step, 4 after). Every (h0, h2) pair with h2 an unused hit in the phi window of h0's extrapolation
into the next module pair is a candidate; the seed is the candidate of least scatter below
max_scatter. Equal scatters go to the earliest candidate: h0 order, then h2's order among the
window's unused hits (track_seeding); or, for the CPU target's lock-step track_seeding_vectorized,
batch of four h0, then that order, then the lane.

The phi window of an extrapolated phi e: from the leftmost insertion point of the 16-bit phi
e - tolerance, consecutive hits around the pair's ring while |phi - e| (16-bit) is within the
tolerance, at most all of them.

Forwarding: every entry of the ring of tracks to follow written since the last step is extended by
the hit of least scatter below max_scatter in the phi window of its last two hits' extrapolation
(used or not); the phi extrapolation goes to the module of the next pair with the parity of the
track's second-to-last hit. A found hit is marked used (and a tracklet's three hits, which then
become a track numbered in promotion order); tracks shorter than 26 hits go back to the ring.
A track without a hit goes back with its skip counter raised while that stays within
max_skipped_modules; otherwise a three-hit tracklet becomes a three-hit track. After the last
step the tracklets left in the ring become three-hit tracks. The ring has 2048 slots written
modulo 2048, and a tracklet lives in the slot of its seed.

Pruned seeding (a learned scorer's rule, no reference counterpart): each seed carries a logit
(looked up by its h1); the unused h1 hits of every run of `block` positions of a module pair, in
hit order, are cut into groups of block / K, and a group keeps its seeds whose logit reaches the
threshold, or all of them when none does.

Floating point is numpy float32 (IEEE single, one rounding per operation, as the reference
without contraction). The 16-bit phi of a point is the oracle's hit_phi_16, pinned on its own.
"""
from __future__ import annotations

import numpy as np

import oracle

F = np.float32
RING = 2048
MAX_TRACK_SIZE = 26
SEED_BIT = 0x80000000
SKIP_SHIFT = 28
SKIP_MASK = 0x70000000
NUMBER_MASK = 0x0FFFFFFF


def i16(v):
    return int(np.array(v, dtype=np.int64).astype(np.int16))


def hit_phi_16(x, y):
    return int(oracle.lib().vo_hit_phi_16(float(x), float(y)))


def leftmost(phis, value):
    """First index whose phi is not below value (phis sorted ascending)."""
    return int(np.searchsorted(np.asarray(phis, dtype=np.int64), value, side="left"))


class Event:
    """One event's sorted clusters: x, y, z (float32), phi (int16), LHCbID, per module pair start
    and count (event-relative)."""

    def __init__(self, x, y, z, phi, ids, mp_start, mp_num, module_zs):
        self.x, self.y, self.z = (np.asarray(a, dtype=F) for a in (x, y, z))
        self.phi = np.asarray(phi, dtype=np.int16)
        self.ids = np.asarray(ids, dtype=np.uint32)
        self.mp_start, self.mp_num = list(mp_start), list(mp_num)
        self.module_zs = np.asarray(module_zs, dtype=F)

    def pair(self, m):
        return self.mp_start[m], self.mp_num[m], (self.module_zs[2 * m], self.module_zs[2 * m + 1])

    def window(self, m, e, tol):
        """Hit indices of module pair m in the phi window of the extrapolated phi e."""
        start, n, _ = self.pair(m)
        if n == 0:
            return []
        phis = self.phi[start:start + n]
        first = leftmost(phis, i16(e - tol))
        out = []
        for j in range(n):
            k = (first + j) % n
            d = i16(int(phis[k]) - e)
            a = i16(-d if d < 0 else d)
            if a > tol:
                break
            out.append(start + k)
        return out


def search(ev: Event, phi_tolerance=0.045, max_scatter=0.08, max_skipped_modules=1, vectorized=False,
           logits=None, threshold=0.0, block=256):
    """logits: None (plain seeding) or a mapping from an event-relative h1 index to its seed's
    logit (pruned seeding)."""
    # Overwritten ring slots can pair a hit with itself: 1/0 and 0 * inf follow IEEE rules, as in
    # the reference, and compare false.
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return _search(ev, phi_tolerance, max_scatter, max_skipped_modules, vectorized, logits, threshold, block)


def _search(ev, phi_tolerance, max_scatter, max_skipped_modules, vectorized, logits, threshold, block):
    tol = int(oracle.lib().vo_hit_phi_float_to_16(float(phi_tolerance)))
    max_scatter = F(max_scatter)
    n_hits = ev.x.size
    used = np.zeros(n_hits, dtype=bool)
    ring = [0] * RING
    tracklets = [None] * RING
    tracks, three_hit = [], []
    counter = 0

    def seeding(prev_m, cur_m, next_m, k_h0):
        nonlocal counter
        ps, pn, pz = ev.pair(prev_m)
        cs, cn, _ = ev.pair(cur_m)
        dz_phi = F(ev.pair(next_m)[2][0] - pz[0])
        groups = {}  # (run, group) -> seeds (h0, h1, h2) in h1 order; dicts keep insertion order
        rank_in_run = {}
        for h1 in range(cs, cs + cn):
            if used[h1]:
                continue
            run = (h1 - cs) // block
            rank = rank_in_run.get(run, 0)
            rank_in_run[run] = rank + 1
            group = groups.setdefault((run, rank // (block // k_h0)), [])
            h0s = []
            if pn:
                idx = leftmost(ev.phi[ps:ps + pn], int(ev.phi[h1]))
                for i in range(pn):
                    if len(h0s) >= k_h0:
                        break
                    idx += i if i & 1 else -i
                    k = idx + pn if idx < 0 else (idx - pn if idx >= pn else idx)
                    if not used[ps + k]:
                        h0s.append(ps + k)
            cands = []
            for r0, h0 in enumerate(h0s):
                td = F(1) / (ev.z[h1] - ev.z[h0])
                tx = (ev.x[h1] - ev.x[h0]) * td
                ty = (ev.y[h1] - ev.y[h0]) * td
                e = hit_phi_16(ev.x[h0] + tx * dz_phi, ev.y[h0] + ty * dz_phi)
                r2 = 0
                for h2 in ev.window(next_m, e, tol):
                    if used[h2]:
                        continue
                    dz = ev.z[h2] - ev.z[h0]
                    dx = (ev.x[h0] + tx * dz) - ev.x[h2]
                    dy = (ev.y[h0] + ty * dz) - ev.y[h2]
                    s = dx * dx + dy * dy
                    if s < max_scatter:
                        key = (s, r0 // 4, r2, r0 % 4) if vectorized else (s, r0, r2)
                        cands.append((key, h0, h2))
                    r2 += 1
            if cands:
                _, h0, h2 = min(cands, key=lambda c: c[0])
                group.append((h0, h1, h2))
        for seeds in groups.values():
            if logits is not None:
                kept = [t for t in seeds if logits[t[1]] >= threshold]
                seeds = kept if kept else seeds
            for h0, h1, h2 in seeds:
                slot = counter % RING
                counter += 1
                tracklets[slot] = [h0, h1, h2]
                ring[slot] = SEED_BIT | slot

    def forwarding(first, next_m, n_entries):
        nonlocal counter
        for j in range(n_entries):
            full = ring[(first + j) % RING]
            number = full & NUMBER_MASK
            skipped = (full & SKIP_MASK) >> SKIP_SHIFT
            is_seed = bool(full & SEED_BIT)
            hits = tracklets[number] if is_seed else tracks[number]
            h0, h1 = hits[-2], hits[-1]
            parity = (((int(ev.ids[h0]) >> 18) & 0xFF) // 4) % 2
            td = F(1) / (ev.z[h1] - ev.z[h0])
            tx = (ev.x[h1] - ev.x[h0]) * td
            ty = (ev.y[h1] - ev.y[h0]) * td
            dz_phi = F(ev.pair(next_m)[2][parity] - ev.z[h0])
            e = hit_phi_16(ev.x[h0] + tx * dz_phi, ev.y[h0] + ty * dz_phi)
            best = None
            for h2 in ev.window(next_m, e, tol):
                dz = ev.z[h2] - ev.z[h0]
                dx = (ev.x[h0] + tx * dz) - ev.x[h2]
                dy = (ev.y[h0] + ty * dz) - ev.y[h2]
                s = dx * dx + dy * dy
                if s < max_scatter and (best is None or s < best[0]):
                    best = (s, h2)
            n = len(hits)
            if best is not None:
                used[best[1]] = True
                if is_seed:
                    used[hits[0]] = used[hits[1]] = used[hits[2]] = True
                    tracks.append(list(hits) + [best[1]])
                    number = len(tracks) - 1
                else:
                    hits.append(best[1])
                if n + 1 < MAX_TRACK_SIZE:
                    ring[counter % RING] = number
                    counter += 1
            elif skipped < max_skipped_modules:
                ring[counter % RING] = ((skipped + 1) << SKIP_SHIFT) | (full & (SEED_BIT | NUMBER_MASK))
                counter += 1
            elif is_seed:
                three_hit.append(list(hits))

    first = 25
    seeding(first, first - 1, first - 2, 8)
    last = 0
    first -= 1
    while first > 1:
        prev, last = last, counter
        forwarding(prev, first - 2, last - prev)
        seeding(first, first - 1, first - 2, 4)
        first -= 1
    prev, last = last, counter
    for j in range(last - prev):
        full = ring[(prev + j) % RING]
        if full & SEED_BIT:
            three_hit.append(list(tracklets[full & NUMBER_MASK]))
    return dict(tracks=tracks, three_hit=three_hit, used=used, ttf=counter, ring=ring)


def event_from_container(container, offsets, counts, e, module_zs):
    """The statement's Event of event e of a sorted container (Velo::Clusters layout)."""
    total = int(offsets[-1])
    b = np.ascontiguousarray(container[:18 * total])
    ids = b[:4 * total].view("<u4")
    xyz = b[4 * total:16 * total].view("<f4").reshape(-1, 3)
    phi = b[16 * total:18 * total].view("<i2")
    base = int(offsets[26 * e])
    end = int(offsets[26 * (e + 1)])
    mp_start = [int(offsets[26 * e + m]) - base for m in range(26)]
    mp_num = [int(counts[26 * e + m]) for m in range(26)]
    s = slice(base, end)
    return Event(xyz[s, 0], xyz[s, 1], xyz[s, 2], phi[s], ids[s], mp_start, mp_num, module_zs)
