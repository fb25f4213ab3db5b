"""VELO track checker against the synthetic generator's truth.

Follows the matching rules of Allen's TrackChecker for VELO tracks:
- **Match:** a track is matched to the particle that contributes at least 70 % of its hits (`checker/tracking/include/TrackChecker.h:45,264-266`).
- **Efficiency:** the fraction of reconstructible particles with at least one matched track.
- **Clones:** extra tracks matched to an already matched particle.
- **Ghosts:** tracks matched to no particle.
- **Hit purity / hit efficiency:** the matched particle's share of the track's hits, and the share of the particle's hits found on the track.

The reference's category "01_velo" (`TrackCheckerCategories.h:29-30`: has VELO hits, not an electron, 2 < eta < 5) is applied to the generator's particles as: hits in at least three modules and 2 < eta < 5. The generator makes no electrons.

Inputs are the consolidated outputs of a run (`offsets_all_velo_tracks`, `offsets_velo_track_hit_number`, and the LHCbIDs at the start of `velo_track_hits`), from the GPU sequence or any other implementation of the path.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MIN_WEIGHT = 0.7


@dataclass
class CheckResult:
    reconstructible: int      # 01_velo particles
    reconstructed: int        # of them, matched by at least one track
    tracks: int
    ghosts: int
    clones: int               # extra matched tracks on 01_velo particles
    matched_tracks: int       # tracks matched to a 01_velo particle
    hit_purity: float
    hit_efficiency: float
    # Per track: matched particle (global index, -1 for ghosts) and its hit share.
    track_particle: np.ndarray
    track_weight: np.ndarray

    @property
    def efficiency(self) -> float:
        return self.reconstructed / self.reconstructible if self.reconstructible else 0.0

    @property
    def ghost_rate(self) -> float:
        return self.ghosts / self.tracks if self.tracks else 0.0

    @property
    def clone_rate(self) -> float:
        return self.clones / self.matched_tracks if self.matched_tracks else 0.0

    def summary(self) -> str:
        return (f"VELO ghosts: {self.ghosts}/{self.tracks} {100 * self.ghost_rate:.2f}%\n"
                f"01_velo : {self.reconstructed}/{self.reconstructible} {100 * self.efficiency:.2f}% "
                f"({100 * self.clone_rate:.2f}% clones), hit purity {100 * self.hit_purity:.2f}%, "
                f"hit eff {100 * self.hit_efficiency:.2f}%")

    def as_dict(self) -> dict:
        return {"01_velo_efficiency": round(self.efficiency, 6), "ghost_rate": round(self.ghost_rate, 6),
                "clone_rate": round(self.clone_rate, 6), "reconstructible": self.reconstructible,
                "reconstructed": self.reconstructed, "tracks": self.tracks, "ghosts": self.ghosts,
                "hit_purity": round(self.hit_purity, 6), "hit_efficiency": round(self.hit_efficiency, 6)}


def reconstructible(events) -> np.ndarray:
    """Boolean per generated particle: category 01_velo."""
    p = events.particles
    return (p["n_modules"] >= 3) & (p["eta"] > 2.0) & (p["eta"] < 5.0)


def track_truth(events, offsets_all_velo_tracks, offsets_velo_track_hit_number, track_ids):
    """(particle, weight, n_hits) per track: the particle with the most hits on the track, the
    fraction of the track's hits it contributed, and the track's hit count. Hits without truth
    (noise) count towards n_hits only."""
    off_all = np.asarray(offsets_all_velo_tracks, dtype=np.int64)
    off_h = np.asarray(offsets_velo_track_hit_number, dtype=np.int64)
    n_tracks = int(off_all[-1])
    n_hits = np.diff(off_h[:n_tracks + 1])
    ids = np.asarray(track_ids, dtype=np.uint64)[:int(off_h[n_tracks])]
    track_event = np.repeat(np.arange(off_all.size - 1, dtype=np.uint64), np.diff(off_all))
    hit_track = np.repeat(np.arange(n_tracks, dtype=np.int64), n_hits)
    hit_event = track_event[hit_track]
    # Truth lookup keyed by (event, LHCbID).
    h = events.hits
    truth_event = np.repeat(np.arange(events.n_events, dtype=np.uint64), np.diff(events.hit_offsets.astype(np.int64)))
    has = h["particle"] != 0xFFFFFFFF
    keys = (truth_event[has] << np.uint64(32)) | h["lhcb_id"][has].astype(np.uint64)
    vals = h["particle"][has].astype(np.int64)
    order = np.argsort(keys, kind="stable")
    keys, vals = keys[order], vals[order]
    q = (hit_event << np.uint64(32)) | ids
    hit_particle = np.full(q.size, -1, np.int64)
    if keys.size:
        pos = np.minimum(np.searchsorted(keys, q), keys.size - 1)
        found = keys[pos] == q
        hit_particle[found] = vals[pos[found]]
    # Most frequent particle per track.
    particle = np.full(n_tracks, -1, np.int64)
    weight = np.zeros(n_tracks)
    sel = hit_particle >= 0
    if sel.any():
        pairs = np.stack([hit_track[sel], hit_particle[sel]], axis=1)
        uniq, counts = np.unique(pairs, axis=0, return_counts=True)
        # Per track the largest count (ties: the lowest particle index, as np.unique sorts).
        best = np.lexsort((-counts, uniq[:, 0]))
        first = np.ones(best.size, bool)
        first[1:] = uniq[best[1:], 0] != uniq[best[:-1], 0]
        b = best[first]
        particle[uniq[b, 0]] = uniq[b, 1]
        weight[uniq[b, 0]] = counts[b] / n_hits[uniq[b, 0]]
    return particle, weight, n_hits


def check(events, offsets_all_velo_tracks, offsets_velo_track_hit_number, track_ids) -> CheckResult:
    particle, weight, n_hits = track_truth(events, offsets_all_velo_tracks, offsets_velo_track_hit_number,
                                           track_ids)
    matched = (particle >= 0) & (weight >= MIN_WEIGHT)
    track_particle = np.where(matched, particle, -1)
    rec = reconstructible(events)
    on_rec = matched & rec[np.maximum(track_particle, 0)]
    matched_particles, n_per = np.unique(track_particle[on_rec], return_counts=True)
    p_hits = events.particles["n_hits"].astype(np.float64)
    purity = float(weight[on_rec].mean()) if on_rec.any() else 0.0
    hit_eff = float((weight[on_rec] * n_hits[on_rec] / p_hits[track_particle[on_rec]]).mean()) if on_rec.any() else 0.0
    return CheckResult(
        reconstructible=int(rec.sum()), reconstructed=int(matched_particles.size), tracks=int(particle.size),
        ghosts=int((~matched).sum()), clones=int((n_per - 1).sum()), matched_tracks=int(on_rec.sum()),
        hit_purity=purity, hit_efficiency=hit_eff, track_particle=track_particle,
        track_weight=np.where(matched, weight, 0.0))


def check_buffers(events, offsets_all_velo_tracks, offsets_velo_track_hit_number, velo_track_hits) -> CheckResult:
    """`velo_track_hits` as raw bytes of the consolidated container ([u32 id x H] first)."""
    h = int(np.asarray(offsets_velo_track_hit_number)[int(np.asarray(offsets_all_velo_tracks)[-1])])
    ids = np.ascontiguousarray(np.asarray(velo_track_hits, dtype=np.uint8)[:4 * h]).view("<u4")
    return check(events, offsets_all_velo_tracks, offsets_velo_track_hit_number, ids)
