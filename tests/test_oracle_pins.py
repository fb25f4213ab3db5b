"""Pins of the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible mistake in the oracle (a dropped term, a wrong sign,
index or shift, a transposed operand) fails it: closed forms, textbook algorithms, the
reference's runtime contracts, its own prefix-sum test, real detector banks from the
reference's bundled MDF file, and known-answer events.
"""
import bisect
import itertools
import math
import os

import numpy as np
import pytest
from scipy import ndimage

import oracle
from allen_amd import synth
from tests import parity, textbook

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def geo(geometry):
    return textbook.parse_geometry(geometry)


# ---------------------------------------------------------------------------- a2: SP tables
def test_sp_patterns_against_textbook_labelling():
    """cache_sp_patterns vs scipy 8-connected labelling of each 2x4 super pixel."""
    pat, sizes, fx, fy = oracle.sp_tables()
    for sp in range(1, 256):
        grid = np.zeros((2, 4), dtype=np.int32)  # [col][row], bit = 4 * col + row
        for b in range(8):
            if sp >> b & 1:
                grid[b // 4, b % 4] = 1
        lab, n = ndimage.label(grid, structure=np.ones((3, 3)))
        assert n <= 2
        # clusters in order of their lowest pixel index
        order = sorted(range(1, n + 1), key=lambda k: min(4 * c + r for c, r in zip(*np.nonzero(lab == k))))
        assert bool(pat[sp] & 8) == (n == 2)
        for slot, k in enumerate(order):
            cols, rows = np.nonzero(lab == k)
            cx, cy = cols.sum() // cols.size, rows.sum() // rows.size
            nib = (pat[sp] >> (4 * slot)) & 0x7
            assert nib == (cx << 2 | cy), (sp, slot)
            assert (sizes[sp] >> (4 * slot)) & 0xF == cols.size
            # evaluated in float32 as in the reference: x / float(n) - cx
            f32 = np.float32
            assert fx[2 * sp + slot] == f32(cols.sum()) / f32(cols.size) - f32(cx)
            assert fy[2 * sp + slot] == f32(rows.sum()) / f32(rows.size) - f32(cy)
            assert abs(fx[2 * sp + slot] - (cols.sum() / cols.size - cx)) < 1e-6


def test_isolated_estimate_rule_matches_table():
    """EstimateInputSize's two-cluster rule (pattern_0 | pattern_1) agrees with the flood fill."""
    pat = oracle.sp_tables()[0]
    for sp in range(256):
        p0 = bool(sp & 0x88) and not (sp & 0x44) and bool(sp & 0x33)
        p1 = bool(sp & 0xCC) and not (sp & 0x22) and bool(sp & 0x11)
        assert (p0 or p1) == bool(pat[sp] & 8), sp


def test_product_sp_tables_equal_oracle():
    from allen_amd import velo
    for a, b in zip(velo.cache_sp_patterns(), oracle.sp_tables()):
        np.testing.assert_array_equal(a, b)


# ---------------------------------------------------------------------------- a8: phi
def test_hit_phi_special_angles():
    """On the octant boundaries the approximation is exact in real arithmetic; in float32 the
    scaled value lands within one count of the ideal one (truncation toward zero)."""
    L = oracle.lib()
    assert L.vo_hit_phi_16(1.0, 0.0) == 0                    # r = 1: c1 - c1 * 1 is exactly 0
    assert L.vo_hit_phi_16(0.0, 1.0) in (16383, 16384)       # pi/2
    assert L.vo_hit_phi_16(0.0, -1.0) in (-16383, -16384)
    assert L.vo_hit_phi_16(1.0, 1.0) in (8191, 8192)         # pi/4
    assert L.vo_hit_phi_16(1.0, -1.0) == -L.vo_hit_phi_16(1.0, 1.0)
    assert L.vo_hit_phi_16(-1.0, 1.0) in (24575, 24576)      # 3pi/4
    assert L.vo_hit_phi_16(-1.0, 0.0) in (32767, -32768)     # pi (wraps through int32 -> int16)
    assert L.vo_hit_phi_float_to_16(0.045) == int(0.045 * 32768 / math.pi)  # default tolerance, 469


# Largest error of the octant approximation atan(r) ~ pi/4 * r over r in [-1, 1]: at
# r^2 = 4/pi - 1 (VeloTools.cuh:21 quotes it as 0.07).
R_STAR = math.sqrt(4 / math.pi - 1)
ATAN_BOUND = math.atan(R_STAR) - math.pi / 4 * R_STAR + 1e-6


def test_fast_atan2_error_bound_and_symmetry():
    """Error within the closed-form bound of the octant approximation, odd in y, monotone."""
    assert 0.0710 < ATAN_BOUND < 0.0712
    L = oracle.lib()
    rng = np.random.default_rng(0)
    ang = np.sort(rng.uniform(-math.pi + 1e-3, math.pi - 1e-3, 4000))
    r = rng.uniform(5, 60, ang.size)
    xs, ys = (r * np.cos(ang)).astype(np.float32), (r * np.sin(ang)).astype(np.float32)
    got = np.array([L.vo_fast_atan2f(float(y), float(x)) for x, y in zip(xs, ys)])
    assert np.max(np.abs(got - np.arctan2(ys, xs))) < ATAN_BOUND
    assert np.all(np.diff(got) >= -1e-6)
    for x, y in zip(xs[:200], ys[:200]):
        assert L.vo_fast_atan2f(float(-y), float(x)) == -L.vo_fast_atan2f(float(y), float(x))


# ---------------------------------------------------------------------------- a10, a5
def test_binary_search_leftmost_against_bisect():
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for n in [0, 1, 2, 7, 64, 333]:
        a = np.sort(rng.integers(-32768, 32767, n)).astype(np.int16)
        for v in rng.integers(-32768, 32767, 50).tolist() + (a[:5].tolist() if n else []):
            got = L.vo_binary_search_leftmost_i16(a.ctypes.data, n, int(v))
            assert got == bisect.bisect_left(a.tolist(), v)


def test_prefix_sum_reference_shape():
    """TestPrefixSum.cu:9-33: iota(1..1000) against a partial sum."""
    x = np.arange(1, 1001, dtype=np.uint32)
    out = np.zeros(1001, np.uint32)
    total = oracle.lib().vo_prefix_sum(x.ctypes.data, x.size, out.ctypes.data)
    assert out.tolist() == [0] + list(itertools.accumulate(x.tolist()))
    assert total == 500500


# ---------------------------------------------------------------------------- a7: mask growth
def _flood(pixel_map, start):
    grid = np.array([[(pixel_map >> (16 * c + r)) & 1 for r in range(16)] for c in range(4)])
    lab, _ = ndimage.label(grid, structure=np.ones((3, 3)))
    keep = set()
    for c in range(4):
        for r in range(16):
            if (start >> (16 * c + r)) & 1 and grid[c, r]:
                keep.add(lab[c, r])
    return sum(1 << (16 * c + r) for c in range(4) for r in range(16) if grid[c, r] and lab[c, r] in keep)


def test_make_cluster_is_8_connected_flood_fill():
    """Away from rows 0, 1, 14 and 15 of the 16-row window the mask growth is a textbook
    8-connected flood fill; x, y, n are the weighted sums of its pixels."""
    L = oracle.lib()
    rng = np.random.default_rng(2)
    xyn = np.zeros(3, np.uint32)
    inner_rows = sum(1 << (16 * c + r) for c in range(4) for r in range(2, 14))
    for _ in range(600):
        pm = int(rng.integers(0, 1 << 63)) & int(rng.integers(0, 1 << 63))
        pm &= inner_rows
        bits = [b for b in range(64) if pm >> b & 1]
        if not bits:
            continue
        start = 1 << int(rng.choice(bits))
        got = L.vo_make_cluster(pm, start, 10, 20, xyn.ctypes.data)
        want = _flood(pm, start)
        assert got == want
        cols = [b // 16 for b in range(64) if want >> b & 1]
        rows = [b % 16 for b in range(64) if want >> b & 1]
        assert xyn[2] == len(cols)
        assert xyn[0] == 10 * len(cols) + sum(cols)
        assert xyn[1] == 20 * len(rows) + sum(rows)


def test_make_cluster_diagonal_masks_of_the_reference():
    """make_8con_mask (MaskedVeloClustering.cu:21-26) masks the up-right diagonal with 0xFFFC
    (rows 0-1 cannot grow up-right) and the down-left diagonal with 0x3FFF (rows 14-15 cannot
    grow down-left); the up-right shift of row 15 lands in row 0 two columns over."""
    L = oracle.lib()
    xyn = np.zeros(3, np.uint32)
    bit = lambda c, r: 1 << (16 * c + r)
    # row 0 -> row 1 of the next column is not grown; row 2 -> row 3 is
    assert L.vo_make_cluster(bit(1, 0) | bit(2, 1), bit(1, 0), 0, 0, xyn.ctypes.data) == bit(1, 0)
    assert L.vo_make_cluster(bit(1, 2) | bit(2, 3), bit(1, 2), 0, 0, xyn.ctypes.data) == bit(1, 2) | bit(2, 3)
    # row 14 -> row 13 of the previous column is not grown; row 12 -> row 11 is
    assert L.vo_make_cluster(bit(2, 14) | bit(1, 13), bit(2, 14), 0, 0, xyn.ctypes.data) == bit(2, 14)
    assert L.vo_make_cluster(bit(2, 12) | bit(1, 11), bit(2, 12), 0, 0, xyn.ctypes.data) == bit(2, 12) | bit(1, 11)
    # row 15 of column 0 reaches row 0 of column 2
    assert L.vo_make_cluster(bit(0, 15) | bit(2, 0), bit(0, 15), 0, 0, xyn.ctypes.data) == bit(0, 15) | bit(2, 0)


# ---------------------------------------------------------------------------- a6/a7 on events
def _check_clusters(ev, geometry, geo, beamline, max_npix):
    r = oracle.run(ev, geometry, beamline)
    est, cnt = r["offsets_estimated_input_size"], r["module_cluster_num"]
    ids, xyz, phi = parity.cluster_views(r["velo_cluster_container"], int(est[-1]))
    n_checked = 0
    for e in range(ev.n_events):
        found = {}
        for mp in range(26):
            a = int(est[e * 26 + mp])
            for k in range(a, a + int(cnt[e * 26 + mp])):
                found[int(ids[k])] = (xyz[k].astype(np.float64), mp)
        for sensor, pixels in textbook.event_pixels(ev, e).items():
            for cols, rows in textbook.components(pixels):
                if cols.size > max_npix:
                    continue
                lhcb, g = textbook.cluster_position(geo, sensor, cols, rows)
                assert lhcb in found, (e, sensor, cols.tolist(), rows.tolist())
                pos, mp = found[lhcb]
                assert mp == sensor // 8
                np.testing.assert_allclose(pos, g, rtol=0, atol=1e-5)
                n_checked += 1
    return r, n_checked


def test_clusters_of_real_banks_against_textbook(geometry, geo, beamline):
    """Real v2 banks of the reference's bundled MDF (10 events): every connected component of
    up to 5 pixels is found with the same LHCbID and its float64 position to 1e-5 mm. Larger
    components exceed the masked window and are split or dropped by design."""
    ev = textbook.load_mdf_fixture(os.path.join(GOLDEN, "mdf_minbias_vp.npz"))
    assert ev.bank_version == 2
    _, n = _check_clusters(ev, geometry, geo, beamline, max_npix=5)
    assert n > 15000


@pytest.mark.parametrize("version", [4, 2])
def test_clusters_of_synthetic_banks_against_textbook(version, geometry, geo, beamline):
    ev = synth.generate(n_events=15, pileup=1, nu=7.6, seed=21, bank_version=version)
    r, n = _check_clusters(ev, geometry, geo, beamline, max_npix=64)
    assert n == int(r["module_cluster_num"].sum())  # and nothing else


def test_sorted_container_contracts(geometry, beamline):
    """SearchByTripletContracts.cu:6-81: per module pair sorted by (phi, id); |x|, |y| < 100 mm;
    the sorted container is a permutation of the unsorted one."""
    ev = synth.generate(n_events=20, pileup=1, nu=7.6, seed=22)
    r = oracle.run(ev, geometry, beamline)
    est, cnt = r["offsets_estimated_input_size"], r["module_cluster_num"]
    T = int(est[-1])
    s_ids, s_xyz, s_phi = parity.cluster_views(r["sorted_velo_cluster_container"], T)
    u_ids, u_xyz, u_phi = parity.cluster_views(r["velo_cluster_container"], T)
    for k in range(cnt.size):
        sl = slice(int(est[k]), int(est[k] + cnt[k]))
        key = s_phi[sl].astype(np.int64) * (1 << 33) + s_ids[sl].astype(np.int64)
        assert np.all(np.diff(key) > 0)
        assert np.all(np.abs(s_xyz[sl, :2]) < 100)
        assert sorted(zip(s_ids[sl].tolist(), s_phi[sl].tolist())) == sorted(zip(u_ids[sl].tolist(),
                                                                                u_phi[sl].tolist()))
    # phi is the octant approximation of the cluster's own position
    sl = parity.filled_slots(est, cnt)
    approx = s_phi[sl].astype(np.float64) * math.pi / 32768
    err = np.abs(np.angle(np.exp(1j * (approx - np.arctan2(s_xyz[sl, 1], s_xyz[sl, 0])))))
    assert err.max() < ATAN_BOUND + 32 / 32768 * math.pi / 32


# ---------------------------------------------------------------------------- a11: tracking
def _truth_map(ev, e):
    h = ev.hits[ev.hit_offsets[e]:ev.hit_offsets[e + 1]]
    return {int(x["lhcb_id"]): int(x["particle"]) for x in h if x["particle"] != 0xFFFFFFFF}


def _track_ids(r, t):
    off = r["offsets_velo_track_hit_number"]
    H = int(off[-1])
    ids = r["velo_track_hits"][:4 * H].view("<u4")
    return ids[off[t]:off[t + 1]].tolist()


def test_known_answer_single_tracks(geometry, beamline):
    """Events of 1-3 noiseless straight tracks with single-pixel, fully efficient hits: each
    particle crossing >= 4 modules comes back as exactly one track holding exactly its hits."""
    ev = synth.generate(n_events=150, mean_tracks=2.0, noise_fraction=0.0, hit_efficiency=1.0,
                        cluster_size_p=(1.0, 0.0, 0.0), seed=23)
    r = oracle.run(ev, geometry, beamline)
    off_all = r["offsets_all_velo_tracks"]
    n_checked = 0
    for e in range(ev.n_events):
        truth = _truth_map(ev, e)
        by_particle = {}
        for lhcb, p in truth.items():
            by_particle.setdefault(p, set()).add(lhcb)
        tracks = [set(_track_ids(r, t)) for t in range(off_all[e], off_all[e + 1])]
        # One hit per module pair, on contiguous module pairs: seeding needs three consecutive
        # pairs and the search walks only downstream, so hits above a gap are not collected, and
        # a particle crossing both modules of a pair can leave a clone.
        def contiguous(h):
            mps = sorted(((i >> 18) & 0xFF) // 8 for i in h)
            return len(set(mps)) == len(mps) and mps[-1] - mps[0] == len(mps) - 1
        simple = {p: h for p, h in by_particle.items() if contiguous(h)}
        if len(by_particle) > 3 or len(simple) != len(by_particle):
            continue
        for p, hits in simple.items():
            if ev.particles[p]["n_modules"] < 4:
                continue
            assert hits in tracks, (e, p)
            n_checked += 1
        assert len(tracks) <= len(by_particle)
    assert n_checked > 15


def test_search_contracts_and_efficiency_band(geometry, beamline):
    """SearchByTripletContracts.cu:83-120 (<= 26 hits, no repeated hit), consolidated ids are VELO
    LHCbIDs (VeloConsolidateTracks.cu:194-269), and the physics sanity band of the reference's
    published VELO-SP numbers: efficiency >= 98 %, ghost rate < 1 %."""
    ev = synth.generate(n_events=60, mean_tracks=120.0, seed=24)
    r = oracle.run(ev, geometry, beamline)
    assert int(r["status"][0]) == 0
    tr = r["tracks"]
    est = r["offsets_estimated_input_size"]
    for e in range(ev.n_events):
        toff = int(np.float32(est[e * 26]) * np.float32(0.34))
        for t in tr[toff:toff + int(r["number_of_velo_tracks"][e])]:
            n = int(t["hitsNum"])
            assert 4 <= n <= 26
            assert len(set(t["hits"][:n].tolist())) == n
    off_all = r["offsets_all_velo_tracks"]
    found, ghosts, n_tracks = set(), 0, int(off_all[-1])
    for e in range(ev.n_events):
        truth = _truth_map(ev, e)
        base = int(ev.particle_offsets[e])
        for t in range(off_all[e], off_all[e + 1]):
            ids = _track_ids(r, t)
            assert all((i >> 28) == 0x2 for i in ids) and len(set(ids)) == len(ids)
            ps = [truth.get(i, -1) for i in ids]
            best = max(set(ps), key=ps.count)
            if best >= 0 and ps.count(best) >= 0.7 * len(ps):
                found.add(best)
            else:
                ghosts += 1
        del base
    recon = [i for i, p in enumerate(ev.particles) if p["n_modules"] >= 3]
    eff = sum(1 for i in recon if i in found) / len(recon)
    assert eff >= 0.98, eff
    assert ghosts / n_tracks < 0.01


# ---------------------------------------------------------------------------- a12, a15
def test_three_hit_chi2_closed_form():
    """Weighted straight-line LSQ chi2 / 2 (ThreeHitTracksFilter.cu:48-107) for hits near z = 0.
    Far from z = 0 the float32 sums of 6400 * z^2 cancel (the reference's own arithmetic)."""
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for _ in range(300):
        z = np.sort(rng.choice(np.arange(-50.0, 50.0, 12.5), 3, replace=False))
        x0, y0, tx, ty = rng.normal(0, 0.05, 2).tolist() + rng.normal(0, 0.2, 2).tolist()
        pts = np.stack([x0 + tx * z + rng.normal(0, 0.02, 3), y0 + ty * z + rng.normal(0, 0.02, 3), z], 1)
        pts32 = pts.astype(np.float32)
        got = L.vo_three_hit_chi2(pts32.ctypes.data)
        p = pts32.astype(np.float64)
        chi2 = 0.0
        for k in (0, 1):
            A = np.stack([np.ones(3), p[:, 2]], 1)
            coef = np.linalg.lstsq(A, p[:, k], rcond=None)[0]
            chi2 += 6400.0 * np.sum((A @ coef - p[:, k]) ** 2)
        assert abs(got - chi2 / 2) <= 2e-3 * max(0.05, chi2 / 2), (got, chi2 / 2)


def _textbook_fit(xyz, bl, start_tx, start_ty, backward, upstream):
    """Matrix-form Kalman filter in float64, one 2-state filter per projection: state (p, t),
    F = [[1, dz], [0, 1]], H = [1, 0], measurement variance 1/6400, process noise on t added
    before each propagation (the convention of VeloKalmanFilter.cuh:67-158)."""
    hits = xyz.astype(np.float64)
    direction = (1 if backward else -1) * (1 if upstream else -1)
    order = list(range(len(hits)))
    if (hits[-1, 2] - hits[0, 2]) * direction < 0:
        order = order[::-1]
    noise = float(np.float32(1e-8)) + float(np.float32(7e-6)) * (start_tx ** 2 + start_ty ** 2)
    out = []
    for k, t0 in ((0, start_tx), (1, start_ty)):
        s = np.array([hits[order[0], k], t0])
        P = np.diag([1 / 6400.0, 1.0])
        z = hits[order[0], 2]
        for i in order[1:]:
            P = P + np.diag([0.0, noise])
            F = np.array([[1.0, hits[i, 2] - z], [0.0, 1.0]])
            s = F @ s
            P = F @ P @ F.T
            H = np.array([1.0, 0.0])
            K = P @ H / (H @ P @ H + 1 / 6400.0)
            s = s + K * (hits[i, k] - s[0])
            P = (np.eye(2) - np.outer(K, H)) @ P
            z = hits[i, 2]
        P = P + np.diag([0.0, noise])
        out.append((s, P))
    (sx, Px), (sy, Py) = out
    if upstream:
        dz = (sx[1] * (bl[0] - sx[0]) + sy[1] * (bl[1] - sy[0])) / (sx[1] ** 2 + sy[1] ** 2)
    else:
        dz = 770.0 - z
    F = np.array([[1.0, dz], [0.0, 1.0]])
    sx, sy = F @ sx, F @ sy
    Px, Py = F @ Px @ F.T, F @ Py @ F.T
    return np.array([sx[0], sy[0], z + dz, sx[1], sy[1]]), np.array([Px[0, 0], Px[0, 1], Px[1, 1], Py[0, 0],
                                                                       Py[0, 1], Py[1, 1]])


def test_kalman_fp64_against_textbook_matrix_filter(geometry, beamline):
    """The FP64 restatement of the simplified fit equals a matrix-form Kalman filter."""
    ev = synth.generate(n_events=6, mean_tracks=60.0, backward_fraction=0.2, seed=27)
    r = oracle.run(ev, geometry, beamline)
    n = int(r["offsets_all_velo_tracks"][-1])
    off = r["offsets_velo_track_hit_number"]
    H = int(off[-1])
    xyz_all = r["velo_track_hits"][4 * H:16 * H].view("<f4").reshape(-1, 3)
    b64 = r["velo_kalman_beamline_states_f64"]
    e64 = r["velo_kalman_endvelo_states_f64"]
    bl = beamline.astype(np.float64)
    for t in range(n):
        xyz = xyz_all[off[t]:off[t + 1]].astype(np.float64)
        first, last = xyz[0], xyz[-1]
        tx = (last[0] - first[0]) / (last[2] - first[2])
        ty = (last[1] - first[1]) / (last[2] - first[2])
        dz = (tx * (bl[0] - last[0]) + ty * (bl[1] - last[1])) / (tx * tx + ty * ty)
        backward = last[2] + dz > first[2]
        assert bool(r["is_backward"][t]) == backward
        st, cov = _textbook_fit(xyz, bl, tx, ty, backward, True)
        got_st = b64[6 * t:6 * t + 5]
        got_cov = b64[6 * n + 8 * t:6 * n + 8 * t + 6]
        np.testing.assert_allclose(got_st, st, rtol=1e-8, atol=1e-9)
        np.testing.assert_allclose(got_cov, cov, rtol=1e-7, atol=1e-15)
        st2, cov2 = _textbook_fit(xyz, bl, st[3], st[4], backward, False)
        np.testing.assert_allclose(e64[6 * t:6 * t + 5], st2, rtol=1e-8, atol=1e-9)
        np.testing.assert_allclose(e64[6 * n + 8 * t:6 * n + 8 * t + 6], cov2, rtol=1e-7, atol=1e-15)


def test_kalman_fp32_error_and_truth(geometry, beamline):
    """The FP32 fit against its FP64 restatement (the error of the reference's own FP32 arithmetic:
    (1 - Kx) cancels once predcovXX * 6400 >> 1), and the FP64 beamline state against the truth."""
    ev = synth.generate(n_events=20, mean_tracks=100.0, seed=25)
    r = oracle.run(ev, geometry, beamline)
    n = int(r["offsets_all_velo_tracks"][-1])
    f32 = r["velo_kalman_beamline_states"][:6 * n].reshape(-1, 6)[:, :5].astype(np.float64)
    f64 = r["velo_kalman_beamline_states_f64"][:6 * n].reshape(-1, 6)[:, :5]
    d = np.abs(f32 - f64)
    assert np.quantile(d[:, 3:], 0.999) < 5e-4       # slopes
    assert np.quantile(d[:, :2], 0.999) < 0.1        # x, y at the beamline (mm)
    bl = r["velo_kalman_beamline_states_f64"]
    off_all = r["offsets_all_velo_tracks"]
    dtx = []
    for e in range(ev.n_events):
        truth = _truth_map(ev, e)
        for t in range(off_all[e], off_all[e + 1]):
            ps = [truth.get(i, -1) for i in _track_ids(r, t)]
            best = max(set(ps), key=ps.count)
            if best >= 0 and ps.count(best) == len(ps):
                p = ev.particles[best]
                dtx.append(bl[6 * t + 3] - p["tx"])
                assert abs(bl[6 * t + 3] - p["tx"]) < 2e-3 and abs(bl[6 * t + 4] - p["ty"]) < 2e-3
    assert len(dtx) > 1000 and abs(np.mean(dtx)) < 1e-4


def _normalised_cov_error(c32, c64):
    """|c32 - c64| per term, diagonal terms relative, off-diagonal terms over sqrt(c_ii c_jj)."""
    pairs = [(0, 0), (0, 2), (2, 2), (3, 3), (3, 5), (5, 5)]
    e = np.empty_like(c64)
    for k, (i, j) in enumerate(pairs):
        e[:, k] = np.abs(c32[:, k] - c64[:, k]) / np.sqrt(np.abs(c64[:, i] * c64[:, j]))
    return e


def test_kalman_fp64_covariance_fallback_against_fp64(geometry, beamline):
    """The FP32 fit's covariance cancels on ill-conditioned steps (kappa = 1 + 6400 predcovXX up to
    ~1e7 on the first step): up to tens of percent against the FP64 fit. With the FP64 covariance
    fallback the FP32 fit stays within FP32 storage precision of the FP64 fit, and so do the states
    (the gains no longer carry the covariance error)."""
    ev = synth.generate(n_events=40, pileup=1, nu=7.6, seed=31)
    plain = oracle.run(ev, geometry, beamline)
    fb = oracle.run(ev, geometry, beamline, kalman_fp64_fallback=1)
    n = int(plain["offsets_all_velo_tracks"][-1])
    assert n > 1000
    for name in ["velo_kalman_beamline_states", "velo_kalman_endvelo_states"]:
        f64 = fb[name + "_f64"][:14 * n]
        c64 = f64[6 * n:].reshape(-1, 8)[:, :6]
        s64 = f64[:6 * n].reshape(-1, 6)[:, :5]
        e_plain = _normalised_cov_error(plain[name][6 * n:14 * n].reshape(-1, 8)[:, :6].astype(np.float64), c64)
        c_fb = fb[name][6 * n:14 * n].reshape(-1, 8)[:, :6].astype(np.float64)
        e_fb = _normalised_cov_error(c_fb, c64)
        assert np.quantile(e_plain, 0.99) > 1e-2          # the reference's FP32 arithmetic
        assert e_fb.max() < 1e-5, e_fb.max()              # with the fallback
        d = np.abs(fb[name][:6 * n].reshape(-1, 6)[:, :5].astype(np.float64) - s64)
        assert d[:, 3:].max() < 1e-6                      # slopes
        assert d[:, :2].max() < 2e-4                      # x, y (mm)
    # The fallback leaves the FP64 restatement and the track list untouched.
    np.testing.assert_array_equal(plain["velo_kalman_beamline_states_f64"], fb["velo_kalman_beamline_states_f64"])
    np.testing.assert_array_equal(plain["is_backward"], fb["is_backward"])


def test_kalman_backward_flag(geometry, beamline):
    ev = synth.generate(n_events=10, mean_tracks=100.0, backward_fraction=0.5, seed=26)
    r = oracle.run(ev, geometry, beamline)
    bwd = r["is_backward"]
    assert 0.2 < bwd.mean() < 0.8
    # backward tracks point to negative z: tx, ty of the state have the opposite sign of x, y at the
    # end-velo extrapolation ... checked through the direction of the hit ordering
    bl = r["velo_kalman_beamline_states"]
    ev_st = r["velo_kalman_endvelo_states"]
    np.testing.assert_allclose(ev_st[2:6 * bwd.size:6], 770.0, atol=1e-3)
    del bl


def test_mlp_restatement_gradient_by_finite_differences():
    """oracle/mlp.py's gradient (without bf16 rounding) against central differences of its loss,
    and its bf16 rounding on known values."""
    from oracle import mlp
    rng = np.random.default_rng(4)
    x = rng.normal(0, 1, (50, 10))
    y = (rng.random(50) < 0.4).astype(np.float64)
    p = rng.normal(0, 0.4, mlp.N_PARAMS)
    g = mlp.gradient(x, y, p, rounded=False)
    eps = 1e-6
    for q in list(range(0, 1441, 37)) + [1440, 1439, 1375, 351]:
        d = np.zeros(mlp.N_PARAMS)
        d[q] = eps
        fd = (mlp.loss(x, y, p + d, False) - mlp.loss(x, y, p - d, False)) / (2 * eps) / x.shape[0]
        assert abs(fd - g[q]) <= 1e-6 + 1e-5 * abs(fd), (q, fd, g[q])
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1 + 2 ** -9], np.float32)
    np.testing.assert_array_equal(mlp.bf16(vals), [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0])


def test_seed_features_against_geometry(geometry, beamline):
    """The scorer's ten seed features, restated from their meaning in float64 on the hits the
    seed log names: h2's residual to the line through h0 and h1 and h1's residual to the line
    through h0 and h2 (written with interpolation weights, predicted minus measured), the 16-bit
    phi steps with wrap-around, the (h0, h1) slopes, the squared h2 residual and h1's module
    pair (which must contain h1)."""
    ev = synth.generate(n_events=30, pileup=1, nu=7.6, seed=62)
    r = oracle.run(ev, geometry, beamline, collect_seeds=1)
    rec = r["seed_log"].reshape(-1, 5).astype(np.int64)
    got = r["seed_features"].reshape(-1, 10).astype(np.float64)
    assert rec.shape[0] > 3000
    n = ev.n_events
    est = r["offsets_estimated_input_size"].astype(np.int64)
    t = int(est[-1])
    cont = np.asarray(r["sorted_velo_cluster_container"]).view(np.uint8)
    xyz = cont[4 * t:16 * t].view("<f4").reshape(-1, 3).astype(np.float64)
    phi = cont[16 * t:18 * t].view("<i2").astype(np.int64)
    base = est[0:26 * n:26][rec[:, 0]]
    h0, h1, h2 = (base + rec[:, k] for k in (2, 3, 4))
    # h1 lies in the module pair the record names.
    mp_lo = est[26 * rec[:, 0] + rec[:, 1]]
    mp_hi = est[26 * rec[:, 0] + rec[:, 1] + 1]
    assert np.all((mp_lo <= h1) & (h1 < mp_hi))
    p0, p1, p2 = xyz[h0], xyz[h1], xyz[h2]
    z0, z1, z2 = p0[:, 2], p1[:, 2], p2[:, 2]
    w2 = ((z2 - z0) / (z1 - z0))[:, None]
    r2 = (1 - w2) * p0[:, :2] + w2 * p1[:, :2] - p2[:, :2]
    w1 = ((z1 - z0) / (z2 - z0))[:, None]
    r1 = (1 - w1) * p0[:, :2] + w1 * p2[:, :2] - p1[:, :2]
    wrap = lambda d: (d + 32768) % 65536 - 32768
    slopes = (p1[:, :2] - p0[:, :2]) / (z1 - z0)[:, None]
    want = np.column_stack([4 * r2, 4 * r1, wrap(phi[h1] - phi[h0]) / 512, wrap(phi[h2] - phi[h1]) / 512,
                            4 * slopes, 12.5 * (r2 ** 2).sum(axis=1), rec[:, 1] / 32])
    np.testing.assert_allclose(got[:, :4], want[:, :4], rtol=0, atol=2e-4)
    np.testing.assert_array_equal(got[:, 4:6], want[:, 4:6])
    np.testing.assert_allclose(got[:, 6:8], want[:, 6:8], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(got[:, 8], want[:, 8], rtol=1e-3, atol=2e-4)
    np.testing.assert_array_equal(got[:, 9], want[:, 9])


# ---------------------------------------------------------------------------- retina decoding
def _check_retina(ev, geometry, geo, beamline):
    """Every module pair of the oracle's retina decoding against the textbook decoding of the
    banks: the same LHCbIDs (one cluster per non-zero word), float64 positions to 1e-5 mm, the
    (phi, id) order, and module-pair ranges that hold every word of the pair's banks (zero
    padding words included)."""
    r = oracle.run(ev, geometry, beamline)
    est, cnt = r["offsets_estimated_input_size"], r["module_cluster_num"]
    ids, xyz, phi = parity.cluster_views(r["sorted_velo_cluster_container"], int(est[-1]))
    n_checked = 0
    words_per_mp = np.zeros(ev.n_events * 26, np.int64)
    for e in range(ev.n_events):
        for source, words in textbook.iter_retina_banks(ev.raw, ev.offsets, ev.sizes, ev.types, ev.bank_version, e):
            sensor0 = source if ev.bank_version in (2, 3) else (source & 0x1FF) << 1
            words_per_mp[e * 26 + sensor0 // 8] += words.size
        by_mp = {}
        for sensor, lhcb, g in textbook.retina_clusters(ev, geo, e):
            by_mp.setdefault(sensor // 8, []).append((lhcb, g))
        for mp in range(26):
            k = e * 26 + mp
            a, c = int(est[k]), int(cnt[k])
            want = by_mp.get(mp, [])
            assert c == len(want), (e, mp)
            key = phi[a:a + c].astype(np.int64) * (1 << 33) + ids[a:a + c].astype(np.int64)
            assert np.all(np.diff(key) > 0)
            got = {int(i): p.astype(np.float64) for i, p in zip(ids[a:a + c], xyz[a:a + c])}
            for lhcb, g in want:
                np.testing.assert_allclose(got[lhcb], g, rtol=0, atol=1e-5)
            n_checked += c
    np.testing.assert_array_equal(est, np.concatenate([[0], np.cumsum(words_per_mp)]))
    return r, n_checked


@pytest.mark.parametrize("version", [4, 3])
def test_retina_decoding_of_synthetic_banks_against_textbook(version, geometry, geo, beamline):
    ev = synth.generate(n_events=12, pileup=1, nu=7.6, seed=41, retina=1, bank_version=version)
    r, n = _check_retina(ev, geometry, geo, beamline)
    assert n > 10000
    if version == 4:
        assert int(r["offsets_estimated_input_size"][-1]) > n  # zero padding words present


def test_retina_decoding_of_real_banks_against_textbook(geometry, geo, beamline):
    """The retina cluster banks (type 73, version 3) of the reference's bundled MDF file."""
    ev = textbook.load_mdf_fixture(os.path.join(GOLDEN, "mdf_minbias_retina.npz"))
    ev.retina = True
    assert ev.bank_version == 3
    _, n = _check_retina(ev, geometry, geo, beamline)
    assert n > 20000


def test_retina_clusters_of_real_data_match_super_pixel_clusters(geometry, beamline):
    """The same 10 real events carry both encodings of their pixels: clusters the FPGA made
    (retina banks) and raw super pixels clustered here. Decoded retina clusters sit on super-pixel
    clusters of the same sensor: the median distance is zero, 94 % are within 0.03 mm (about half
    a pixel); the rest are clusters the two clusterings split differently."""
    ret = textbook.load_mdf_fixture(os.path.join(GOLDEN, "mdf_minbias_retina.npz"))
    ret.retina = True
    sp = textbook.load_mdf_fixture(os.path.join(GOLDEN, "mdf_minbias_vp.npz"))
    rr, rs = oracle.run(ret, geometry, beamline), oracle.run(sp, geometry, beamline)

    def filled(r, e):
        est, cnt = r["offsets_estimated_input_size"], r["module_cluster_num"]
        ids, xyz, _ = parity.cluster_views(r["sorted_velo_cluster_container"], int(est[-1]))
        idx = np.concatenate([np.arange(est[e * 26 + m], est[e * 26 + m] + cnt[e * 26 + m]) for m in range(26)])
        idx = idx.astype(np.int64)
        return (ids[idx] >> 18) & 0xFF, xyz[idx].astype(np.float64)

    dist = []
    for e in range(ret.n_events):
        sr, xr = filled(rr, e)
        ss, xs = filled(rs, e)
        for k in range(sr.size):
            on = ss == sr[k]
            if on.any():
                dist.append(np.min(np.linalg.norm(xs[on, :2] - xr[k, :2], axis=1)))
    dist = np.array(dist)
    assert dist.size > 20000
    assert np.median(dist) == 0.0
    assert (dist < 0.03).mean() > 0.93
    # Both give the same number of tracks to within a few per cent.
    n_ret, n_sp = int(rr["offsets_all_velo_tracks"][-1]), int(rs["offsets_all_velo_tracks"][-1])
    assert abs(n_ret - n_sp) < 0.03 * n_sp


def test_retina_synthetic_tracking_efficiency(geometry, beamline):
    """The whole path on synthetic retina banks finds the generated particles (01_velo)."""
    from allen_amd import checker
    ev = synth.generate(n_events=20, pileup=0, mean_tracks=120.0, seed=42, retina=1)
    r = oracle.run(ev, geometry, beamline)
    res = checker.check_buffers(ev, r["offsets_all_velo_tracks"], r["offsets_velo_track_hit_number"],
                                r["velo_track_hits"])
    assert res.efficiency > 0.99 and res.ghost_rate < 0.005


# ---------------------------------------------------------------------------- beamline PV finder
def _windowed_tracks(pvt, beamline):
    """Tracks the histogram takes (pv_beamline_histo.cu:40-62), restated in numpy."""
    z, x, y, tx, ty, w0, w1 = pvt.T.astype(np.float64)
    ok = (-541.0 < z) & (z < 307.0)
    ok &= (x - beamline[0]) ** 2 * w0 + (y - beamline[1]) ** 2 * w1 < 10.0
    ok &= 1.0 / np.sqrt(tx * tx * w0 + ty * ty * w1) < np.where(z < -341.0, 10.0, 1.5)
    return ok


def test_pv_histogram_is_a_sum_of_unit_gaussians(geometry, beamline):
    """Each accepted track adds increments that telescope to exactly one (the integral of its
    approximate Gaussian over its window), all non-negative, centred on its z: the histogram's
    total is the number of accepted tracks, and its first moment their mean z."""
    ev = synth.generate(n_events=30, pileup=1, nu=7.6, seed=82)
    r = oracle.run(ev, geometry, beamline, pv_finder=1)
    h = r["zhisto"].reshape(ev.n_events, 3392).astype(np.float64)
    pvt = r["pvtracks"].reshape(-1, 7)
    off = r["offsets_all_velo_tracks"].astype(np.int64)
    centres = -541.0 + (np.arange(3392) + 0.5) * 0.25
    assert h.min() >= 0.0
    for e in range(ev.n_events):
        t = pvt[off[e]:off[e + 1]]
        ok = _windowed_tracks(t, beamline)
        assert abs(h[e].sum() - ok.sum()) < 1e-3
        if ok.sum():
            mean_z = (h[e] * centres).sum() / h[e].sum()
            assert abs(mean_z - t[ok, 0].astype(np.float64).mean()) < 0.2


def test_pv_finder_finds_the_generated_vertices(geometry, beamline):
    """Pile-up events: generated vertices with at least five reconstructible tracks are found
    within 1 mm in z (94 %), with a median |dz| below 0.1 mm, on the beamline, and almost every
    reconstructed vertex sits on a generated one."""
    ev = synth.generate(n_events=60, pileup=1, nu=7.6, seed=83)
    r = oracle.run(ev, geometry, beamline, pv_finder=1)
    n = ev.n_events
    nfin = r["number_of_multi_final_vertices"]
    fin = r["multi_final_vertices"].reshape(n, 32)
    p, po = ev.particles, ev.particle_offsets
    found, total, dz, fakes, n_rec = 0, 0, [], 0, 0
    for e in range(n):
        pe = p[po[e]:po[e + 1]]
        sel = (pe["n_modules"] >= 3) & (pe["eta"] > 2) & (pe["eta"] < 5)
        pvs, counts = np.unique(pe["pv"][sel], return_counts=True)
        zt = np.array([pe["z0"][pe["pv"] == v][0] for v in pvs[counts >= 5]], np.float64)
        zall = np.unique(pe["z0"]).astype(np.float64)
        rec = fin[e][:nfin[e]]
        n_rec += rec.size
        total += zt.size
        for z in zt:
            d = np.abs(rec["z"] - z)
            if d.size and d.min() < 1.0:
                found += 1
                dz.append(d.min())
        for v in rec:
            if zall.size == 0 or np.abs(zall - v["z"]).min() > 1.0:
                fakes += 1
            assert v["x"] ** 2 + v["y"] ** 2 < 0.3
    assert total > 300
    assert found / total > 0.9
    assert np.median(dz) < 0.1
    assert fakes < 0.03 * n_rec
    assert r["status"][0] == 0
