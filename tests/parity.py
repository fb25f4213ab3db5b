"""Comparison of the HIP path with the CPU oracle, buffer by buffer.

Exactness rules (the GPU path reproduces the CPU target's order, so most buffers are
compared element for element):
  * integer buffers: equal;
  * float buffers (cluster and track-hit coordinates, Kalman states): equal bit for bit;
  * the gapped cluster containers: only the filled part of each module pair
    ([start, start + count)); gap contents are unspecified in the reference;
  * unordered outputs of the clustering stages (candidate list, unsorted container):
    equal as sets per event / per module pair.
"""
from __future__ import annotations

import numpy as np

from allen_amd import velo

N_MP = 26
TRACK = velo.TRACK_DTYPE
TRACKLET = velo.TRACKLET_DTYPE


def cluster_views(buf: np.ndarray, total: int):
    """(ids u32, xyz f32 (T,3), phi i16) views of a cluster container with `total` slots."""
    b = np.ascontiguousarray(buf[:18 * total])
    ids = b[:4 * total].view("<u4")
    xyz = b[4 * total:16 * total].view("<f4").reshape(-1, 3)
    phi = b[16 * total:18 * total].view("<i2")
    return ids, xyz, phi


def filled_slots(offsets_est: np.ndarray, counts: np.ndarray) -> np.ndarray:
    idx = [np.arange(offsets_est[i], offsets_est[i] + counts[i]) for i in range(counts.size)]
    return np.concatenate(idx) if idx else np.zeros(0, np.int64)


class _Memo:
    """Reads each buffer of a result once. The oracle's result copies a whole buffer on every
    access, and the comparisons below index some buffers once per track or event: without the
    cache a 400-event comparison takes minutes of CPU time."""

    def __init__(self, src):
        self._src = src
        self._cache = {}

    def __getitem__(self, name):
        if name not in self._cache:
            self._cache[name] = self._src[name]
        return self._cache[name]


def _memo(ref):
    return ref if isinstance(ref, (dict, _Memo)) else _Memo(ref)


def is_retina(ref) -> bool:
    """Whether the oracle's run decoded retina cluster banks."""
    try:
        ref["each_sensor_pair_size"]
        return True
    except KeyError:
        return False


def gpu_buffers(seq: velo.VeloSequence, ref) -> dict:
    """Fetches the GPU sequence's buffers with the sizes of the oracle's run `ref`."""
    ref = _memo(ref)
    n_events = ref["number_of_velo_tracks"].size
    retina = is_retina(ref)
    out = {}

    def f(name, dtype, count):
        out[name] = seq.fetch(name, dtype, int(count))

    for name in ["number_of_velo_tracks", "number_of_three_hit_tracks_output"]:
        f(name, np.uint32, n_events)
    for name in ["offsets_velo_tracks", "offsets_three_hit_tracks_filtered", "offsets_all_velo_tracks"]:
        f(name, np.uint32, n_events + 1)
    f("offsets_estimated_input_size", np.uint32, N_MP * n_events + 1)
    f("module_cluster_num", np.uint32, N_MP * n_events)
    total = int(ref["offsets_estimated_input_size"][-1])
    f("sorted_velo_cluster_container", np.uint8, 18 * total)
    if retina:
        slots = ref["retina_bank_index"].size
        f("each_sensor_pair_size", np.uint32, slots * n_events)
        f("offsets_each_sensor_pair_size", np.uint32, slots * n_events + 1)
        f("retina_bank_index", np.uint32, slots)
        f("hit_permutations", np.uint32, total)
        f("hit_sorting_key", np.int64, total)
    else:
        for name in ["number_of_candidates", "module_candidate_num"]:
            f(name, np.uint32, n_events)
        f("candidates_offsets", np.uint32, n_events + 1)
        f("velo_bank_index", np.uint32, 208)
        f("estimated_input_size", np.uint32, N_MP * n_events)
        f("velo_cluster_container", np.uint8, 18 * total)
        f("cluster_candidates", np.uint32, ref["cluster_candidates"].size)
        f("hit_permutation", np.uint32, total)
    f("tracks", TRACK, ref["tracks"].size)
    f("tracklets", TRACKLET, ref["tracklets"].size)
    f("tracks_to_follow", np.uint32, ref["tracks_to_follow"].size)
    f("three_hit_tracks", TRACKLET, ref["three_hit_tracks"].size)
    f("three_hit_tracks_output", TRACKLET, ref["three_hit_tracks_output"].size)
    f("hit_used", np.bool_, total)
    f("rel_indices", np.uint16, total)
    f("atomics_velo", np.uint32, 3 * n_events)
    n_tracks = int(ref["offsets_all_velo_tracks"][-1])
    f("velo_track_hit_number", np.uint32, n_tracks)
    f("offsets_velo_track_hit_number", np.uint32, n_tracks + 1)
    f("velo_track_hits", np.uint8, ref["velo_track_hits"].size)
    f("is_backward", np.bool_, n_tracks)
    f("accepted_velo_tracks", np.uint8, n_tracks)
    for name in ["velo_kalman_beamline_states", "velo_kalman_endvelo_states"]:
        f(name, np.float32, 14 * n_tracks)
    return out


def _mismatch(name, a, b, limit=5):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"{name}: shape {a.shape} != {b.shape}"
    if a.size == 0:
        return None
    bad = np.nonzero(a != b)[0] if a.ndim == 1 else np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
    if bad.size == 0:
        return None
    return f"{name}: {bad.size} mismatches, first at {bad[:limit].tolist()}: gpu {a[bad[:limit]].tolist()} " \
           f"oracle {b[bad[:limit]].tolist()}"


def compare(gpu: dict, ref, fused: bool) -> list[str]:
    """Returns the list of mismatch descriptions (empty when everything agrees)."""
    ref = _memo(ref)
    errs = []
    n_events = ref["number_of_velo_tracks"].size
    retina = is_retina(ref)

    def eq(name, a=None, b=None):
        m = _mismatch(name, gpu[name] if a is None else a, ref[name] if b is None else b)
        if m:
            errs.append(m)

    names = ["each_sensor_pair_size", "offsets_each_sensor_pair_size", "retina_bank_index", "hit_sorting_key",
             "hit_permutations"] if retina else ["number_of_candidates", "candidates_offsets", "velo_bank_index",
                                                 "estimated_input_size"]
    for name in names + ["offsets_estimated_input_size", "module_cluster_num"]:
        eq(name)
    est = ref["offsets_estimated_input_size"]
    total = int(est[-1])
    counts = ref["module_cluster_num"]
    slots = filled_slots(est, counts)
    if retina:
        # Module pairs with zero words also fill the slot after their clusters (the last zero
        # word decoded), as the reference does.
        zero = (np.diff(est.astype(np.int64)) > counts)
        extra = (est[:-1].astype(np.int64) + counts)[zero]
        g_ids, g_xyz, g_phi = cluster_views(gpu["sorted_velo_cluster_container"], total)
        r_ids, r_xyz, r_phi = cluster_views(ref["sorted_velo_cluster_container"], total)
        eq("sorted.zero_word_slot", np.stack([g_ids[extra], g_xyz[extra].view("<u4")[:, 0]]),
           np.stack([r_ids[extra], r_xyz[extra].view("<u4")[:, 0]]))

    # Sorted container: filled slots equal, bit for bit.
    g_ids, g_xyz, g_phi = cluster_views(gpu["sorted_velo_cluster_container"], total)
    r_ids, r_xyz, r_phi = cluster_views(ref["sorted_velo_cluster_container"], total)
    eq("sorted.id", g_ids[slots], r_ids[slots])
    eq("sorted.phi", g_phi[slots], r_phi[slots])
    eq("sorted.xyz", g_xyz[slots].view("<u4"), r_xyz[slots].view("<u4"))

    if not fused and not retina:
        eq("module_candidate_num")
        # Candidate lists: same multiset per event.
        co = ref["candidates_offsets"]
        mcn = ref["module_candidate_num"]
        for e in range(n_events):
            a = np.sort(gpu["cluster_candidates"][co[e]:co[e] + mcn[e]])
            b = np.sort(ref["cluster_candidates"][co[e]:co[e] + mcn[e]])
            m = _mismatch(f"cluster_candidates[event {e}]", a, b)
            if m:
                errs.append(m)
                break
        # Unsorted container: same set of (id, phi, xyz) per module pair.
        u_ids, u_xyz, u_phi = cluster_views(gpu["velo_cluster_container"], total)
        o_ids, o_xyz, o_phi = cluster_views(ref["velo_cluster_container"], total)
        for k in range(counts.size):
            s = slice(est[k], est[k] + counts[k])
            ka = np.rec.fromarrays([u_ids[s], u_phi[s], *u_xyz[s].view("<u4").T]) if counts[k] else None
            kb = np.rec.fromarrays([o_ids[s], o_phi[s], *o_xyz[s].view("<u4").T]) if counts[k] else None
            if counts[k] and not np.array_equal(np.sort(ka), np.sort(kb)):
                errs.append(f"velo_cluster_container: module pair slot {k} differs as a set")
                break
        # Permutation: sorted[i] == unsorted[perm[i]] on the GPU side.
        perm = gpu["hit_permutation"][slots]
        if not (np.array_equal(u_ids[perm], g_ids[slots]) and np.array_equal(u_phi[perm], g_phi[slots])):
            errs.append("hit_permutation does not map the unsorted container onto the sorted one")

    # search_by_triplet outputs.
    eq("number_of_velo_tracks")
    eq("atomics_velo")
    eq("hit_used", gpu["hit_used"][slots], ref["hit_used"][slots])
    n_tr = ref["number_of_velo_tracks"]
    three_n = ref["atomics_velo"][0::3]
    ttf_n = ref["atomics_velo"][1::3]
    local_n = ref["atomics_velo"][2::3]
    tcs = ref["tracks"].size
    for e in range(n_events):
        toff = int(np.float32(est[e * N_MP]) * np.float32(0.34))
        for i in range(int(n_tr[e])):
            if toff + i >= tcs:
                break
            a, b = gpu["tracks"][toff + i], ref["tracks"][toff + i]
            if a["hitsNum"] != b["hitsNum"] or not np.array_equal(a["hits"][:b["hitsNum"]], b["hits"][:b["hitsNum"]]):
                errs.append(f"tracks[event {e}, {i}]: gpu {a['hits'][:a['hitsNum']].tolist()} "
                            f"oracle {b['hits'][:b['hitsNum']].tolist()}")
                break
        n3 = int(three_n[e])
        if n3:
            m = _mismatch(f"three_hit_tracks[event {e}]", gpu["three_hit_tracks"]["hits"][toff:toff + n3],
                          ref["three_hit_tracks"]["hits"][toff:toff + n3])
            if m:
                errs.append(m)
        written = min(int(ttf_n[e]), 2048)
        base = e * 2048
        m = _mismatch(f"tracks_to_follow[event {e}]", gpu["tracks_to_follow"][base:base + written],
                      ref["tracks_to_follow"][base:base + written])
        if m:
            errs.append(m)
        seeds = ref["tracks_to_follow"][base:base + written]
        live = np.unique(seeds[(seeds & 0x80000000) != 0] & 0x0FFFFFFF) if written else np.zeros(0, np.uint32)
        if live.size:
            m = _mismatch(f"tracklets[event {e}]", gpu["tracklets"]["hits"][base + live],
                          ref["tracklets"]["hits"][base + live])
            if m:
                errs.append(m)
        h0 = int(est[e * N_MP])
        m = _mismatch(f"rel_indices[event {e}]", gpu["rel_indices"][h0:h0 + int(local_n[e])],
                      ref["rel_indices"][h0:h0 + int(local_n[e])])
        if m:
            errs.append(m)
        if len(errs) > 20:
            break

    # Three-hit filter, consolidation, Kalman.
    for name in ["number_of_three_hit_tracks_output", "offsets_velo_tracks", "offsets_three_hit_tracks_filtered",
                 "offsets_all_velo_tracks", "velo_track_hit_number", "offsets_velo_track_hit_number",
                 "is_backward"]:
        eq(name)
    n3o = ref["number_of_three_hit_tracks_output"]
    for e in range(n_events):
        toff = int(np.float32(est[e * N_MP]) * np.float32(0.34))
        m = _mismatch(f"three_hit_tracks_output[event {e}]",
                      gpu["three_hit_tracks_output"]["hits"][toff:toff + n3o[e]],
                      ref["three_hit_tracks_output"]["hits"][toff:toff + n3o[e]])
        if m:
            errs.append(m)
            break
    # Consolidation writes id, x, y, z of each track hit (VeloConsolidateTracks.cu:180-190); the
    # phi words of the container are never written by the reference.
    n_hits = int(ref["offsets_velo_track_hit_number"][-1])
    eq("velo_track_hits", gpu["velo_track_hits"][:16 * n_hits], ref["velo_track_hits"][:16 * n_hits])
    if not np.all(gpu["accepted_velo_tracks"] == 1):
        errs.append("accepted_velo_tracks not all 1")
    n_tracks = int(ref["offsets_all_velo_tracks"][-1])
    for name in ["velo_kalman_beamline_states", "velo_kalman_endvelo_states"]:
        g = gpu[name].view("<u4")
        r = ref[name].view("<u4")
        st = np.r_[0:6 * n_tracks].reshape(-1, 6)[:, :5].ravel()        # x, y, z, tx, ty (not qop)
        cv = (6 * n_tracks + np.r_[0:8 * n_tracks]).reshape(-1, 8)[:, :6].ravel()  # six covariance terms
        eq(name + ".state", g[st], r[st])
        eq(name + ".cov", g[cv], r[cv])
    return errs
