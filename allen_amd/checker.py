"""VELO track checker against the synthetic generator's truth.

Follows the matching rules of Allen's TrackChecker for VELO tracks:
- **Match:** a track is matched to the particle that contributes at least 70 % of its hits (`checker/tracking/include/TrackChecker.h:45,264-266`).
- **Efficiency:** the fraction of reconstructible particles with at least one matched track.
- **Clones:** extra tracks matched to an already matched particle.
- **Ghosts:** tracks matched to no particle.
- **Hit purity / hit efficiency:** the matched particle's share of the track's hits, and the share of the particle's hits found on the track.

Categories (`TrackCheckerCategories.h:24-40`; the generator makes no electrons and no decays, so the categories of strange, beauty and signal decays accept nothing and, as in the reference's report, are not printed):
- **01_velo**: has VELO hits (hits in at least three modules) and 2 < eta < 5.
- **02_long**: a long track, and 2 < eta < 5. Allen takes `isLong` from the simulation (hits in the VELO and in the downstream SciFi stations). For these VELO-only synthetic events it is frozen as: has VELO hits, flies forward, lies in the downstream tracker's acceptance (|tx| < 0.30, |ty| < 0.25, 300 x 250 mrad) and has p > 2 GeV (slower particles leave that acceptance in the dipole field before the downstream stations).
- **03_long_P>5GeV**: 02_long with p > 5 GeV.

`report()` prints the reference's TrackChecker lines (`TrackChecker.h:76-101`, `TrackEffReport.cpp:83-108`): ghosts over all tracks, then per category found/accepted, efficiency, mean per-event efficiency, clones, hit purity and hit efficiency.

Inputs are the consolidated outputs of a run (`offsets_all_velo_tracks`, `offsets_velo_track_hit_number`, and the LHCbIDs at the start of `velo_track_hits`), from the GPU sequence or any other implementation of the path.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MIN_WEIGHT = 0.7


LONG_TX, LONG_TY, LONG_P_MEV = 0.30, 0.25, 2000.0
CATEGORIES = ("01_velo", "02_long", "03_long_P>5GeV")


@dataclass
class Category:
    name: str
    accepted: int
    found: int
    clones: int
    hit_purity: float
    hit_efficiency: float
    efficiency_per_event: float

    @property
    def efficiency(self) -> float:
        return self.found / self.accepted if self.accepted else 0.0

    @property
    def clone_rate(self) -> float:
        return self.clones / (self.found + self.clones) if self.found else 0.0

    def line(self) -> str:
        return (f"{self.name:<50s}: {self.found:9d}/{self.accepted:9d} {100 * self.efficiency:6.2f}% "
                f"({100 * self.efficiency_per_event:6.2f}%), {self.clones:9d} ({100 * self.clone_rate:6.2f}%) clones, "
                f"pur {100 * self.hit_purity:6.2f}%, hit eff {100 * self.hit_efficiency:6.2f}%")


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
    categories: dict = None   # name -> Category

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
        return self.report()

    def report(self) -> str:
        """The reference's VELO TrackChecker lines."""
        lines = [f"{'TrackChecker output':<50s}: {self.ghosts:9d}/{self.tracks:9d} {100 * self.ghost_rate:6.2f}% ghosts"]
        lines += [c.line() for c in (self.categories or {}).values() if c.accepted > 0]
        return "\n".join(lines)

    def as_dict(self) -> dict:
        out = {"01_velo_efficiency": round(self.efficiency, 6), "ghost_rate": round(self.ghost_rate, 6),
               "clone_rate": round(self.clone_rate, 6), "reconstructible": self.reconstructible,
               "reconstructed": self.reconstructed, "tracks": self.tracks, "ghosts": self.ghosts,
               "hit_purity": round(self.hit_purity, 6), "hit_efficiency": round(self.hit_efficiency, 6)}
        for name, c in (self.categories or {}).items():
            if name != "01_velo":
                out[f"{name}_efficiency"] = round(c.efficiency, 6)
                out[f"{name}_found"], out[f"{name}_accepted"] = c.found, c.accepted
        return out


def has_velo(events) -> np.ndarray:
    return events.particles["n_modules"] >= 3


def is_long(events) -> np.ndarray:
    """The frozen long-track definition of the synthetic events (see the module docstring)."""
    p = events.particles
    return has_velo(events) & (p["eta"] > 0) & (np.abs(p["tx"]) < LONG_TX) & (np.abs(p["ty"]) < LONG_TY) & \
        (p["p"] > LONG_P_MEV)


def category_masks(events) -> dict:
    p = events.particles
    in_eta = (p["eta"] > 2.0) & (p["eta"] < 5.0)
    long = is_long(events)
    return {"01_velo": has_velo(events) & in_eta, "02_long": long & in_eta,
            "03_long_P>5GeV": long & in_eta & (p["p"] > 5000.0)}


def reconstructible(events) -> np.ndarray:
    """Boolean per generated particle: category 01_velo."""
    return category_masks(events)["01_velo"]


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
        # (track, particle) pairs as one sortable key.
        n_part = int(hit_particle[sel].max()) + 1
        ukey, counts = np.unique(hit_track[sel] * n_part + hit_particle[sel], return_counts=True)
        u_track, u_part = ukey // n_part, ukey % n_part
        # Per track the largest count (ties: the lowest particle index; lexsort is stable and the
        # keys are sorted by particle within a track).
        best = np.lexsort((-counts, u_track))
        first = np.ones(best.size, bool)
        first[1:] = u_track[best[1:]] != u_track[best[:-1]]
        b = best[first]
        particle[u_track[b]] = u_part[b]
        weight[u_track[b]] = counts[b] / n_hits[u_track[b]]
    return particle, weight, n_hits


def _category(name, accept, matched, track_particle, weight, n_hits, events) -> Category:
    """TrackEffReport (TrackEffReport.cpp:38-81) of one category."""
    on = matched & accept[np.maximum(track_particle, 0)]
    found_particles, n_per = np.unique(track_particle[on], return_counts=True)
    p_hits = events.particles["n_hits"].astype(np.float64)
    purity = float(weight[on].mean()) if on.any() else 0.0
    hit_eff = float((weight[on] * n_hits[on] / p_hits[track_particle[on]]).mean()) if on.any() else 0.0
    # Mean of the per-event efficiencies over events with an accepted particle (:41-48).
    ev = events.particles["event"].astype(np.int64)
    acc_per = np.bincount(ev[accept], minlength=events.n_events)
    found_per = np.bincount(ev[found_particles], minlength=events.n_events) if found_particles.size else \
        np.zeros(events.n_events, np.int64)
    has = acc_per > 0
    per_event = float((found_per[has] / acc_per[has]).mean()) if has.any() else 0.0
    return Category(name, int(accept.sum()), int(found_particles.size), int((n_per - 1).sum()), purity, hit_eff,
                    per_event)


def check(events, offsets_all_velo_tracks, offsets_velo_track_hit_number, track_ids) -> CheckResult:
    particle, weight, n_hits = track_truth(events, offsets_all_velo_tracks, offsets_velo_track_hit_number,
                                           track_ids)
    matched = (particle >= 0) & (weight >= MIN_WEIGHT)
    track_particle = np.where(matched, particle, -1)
    cats = {name: _category(name, mask, matched, track_particle, weight, n_hits, events)
            for name, mask in category_masks(events).items()}
    velo = cats["01_velo"]
    return CheckResult(
        reconstructible=velo.accepted, reconstructed=velo.found, tracks=int(particle.size),
        ghosts=int((~matched).sum()), clones=velo.clones,
        matched_tracks=int((matched & category_masks(events)["01_velo"][np.maximum(track_particle, 0)]).sum()),
        hit_purity=velo.hit_purity, hit_efficiency=velo.hit_efficiency, track_particle=track_particle,
        track_weight=np.where(matched, weight, 0.0), categories=cats)


def check_buffers(events, offsets_all_velo_tracks, offsets_velo_track_hit_number, velo_track_hits) -> CheckResult:
    """`velo_track_hits` as raw bytes of the consolidated container ([u32 id x H] first)."""
    h = int(np.asarray(offsets_velo_track_hit_number)[int(np.asarray(offsets_all_velo_tracks)[-1])])
    ids = np.ascontiguousarray(np.asarray(velo_track_hits, dtype=np.uint8)[:4 * h]).view("<u4")
    return check(events, offsets_all_velo_tracks, offsets_velo_track_hit_number, ids)
