"""Pins of the oracle's super-pixel estimate (K2) and masked clustering (K3) on adversarial banks
(runs without a GPU).

The oracle restates EstimateInputSize.cu and MaskedVeloClustering.cu bit for bit. These tests
check it against tests/textbook_clustering.py, which states the same rules in pixel coordinates,
on banks built to reach every branch: known-answer banks for the estimate's quirks, every
8-connected shape of up to four pixels at every alignment to the super-pixel grid, long shapes
that reach the edges of the clustering window and the rows where its growth masks are
asymmetric, and random pixel maps with wrong isolation flags and both sensors of a pair
overlapping.
"""
import numpy as np
import pytest

import oracle
from allen_amd import synth
from tests import layout, parity, textbook, textbook_clustering as tc


@pytest.fixture(scope="module")
def geo(geometry):
    return textbook.parse_geometry(geometry)


def _order(c):
    """Clusters by LHCbID, ties (same centroid pixel) by position rounded well above float32 error."""
    return (c[0],) + tuple(np.round(np.asarray(c[1], dtype=np.float64), 3))


def _oracle_event_views(r, e):
    est = r["estimated_input_size"].reshape(-1, 26)[e]
    co, mcn = r["candidates_offsets"], r["module_candidate_num"]
    cands = sorted(r["cluster_candidates"][co[e]:co[e] + mcn[e]].tolist())
    off, cnt = r["offsets_estimated_input_size"], r["module_cluster_num"]
    ids, xyz, _ = parity.cluster_views(r["velo_cluster_container"], int(off[-1]))
    clusters = {}
    for mp in range(26):
        a, n = int(off[e * 26 + mp]), int(cnt[e * 26 + mp])
        clusters[mp] = sorted(zip(ids[a:a + n].tolist(), map(tuple, xyz[a:a + n].astype(np.float64))), key=_order)
    return est, cands, clusters, cnt[e * 26:(e + 1) * 26]


def check_against_statement(ev, geometry, geo, beamline):
    """Estimates, candidates and clusters of every event equal the pixel-coordinate statement;
    returns the number of clusters checked."""
    r = oracle.run(ev, geometry, beamline)
    n_checked = 0
    for e in range(ev.n_events):
        banks = list(textbook.iter_banks(ev.raw, ev.offsets, ev.sizes, ev.types, ev.bank_version, e))
        est, cands, clusters = tc.event(banks, ev.bank_version, geo)
        o_est, o_cands, o_clusters, o_cnt = _oracle_event_views(r, e)
        np.testing.assert_array_equal(o_est, est, err_msg=f"estimate, event {e}")
        assert o_cands == sorted(c for v in cands.values() for c in v), f"candidates, event {e}"
        assert np.all(o_cnt <= o_est), "a module pair holds more clusters than its estimate"
        for mp in range(26):
            want = sorted(clusters.get(mp, []), key=_order)
            got = o_clusters[mp]
            assert [g[0] for g in got] == [w[0] for w in want], f"cluster ids, event {e}, module pair {mp}"
            if want:
                np.testing.assert_allclose(np.array([g[1] for g in got]), np.array([w[1] for w in want]),
                                           rtol=0, atol=2e-5)
            n_checked += len(want)
    return n_checked


def encode(pixels_by_sensor, version, flag="true", rng=None):
    """SP words of absolute pixels {sensor: set((col, row))}: banks {sourceID: words}. flag "true"
    sets the isolation bit exactly when no 8-neighbouring SP of the same sensor exists; "never"
    and "random" give the adversarial flags."""
    banks = {}
    for sensor, pixels in pixels_by_sensor.items():
        sps = {}
        for (c, r) in pixels:
            key = (c // 2, r // 4)
            sps[key] = sps.get(key, 0) | (1 << (4 * (c % 2) + r % 4))
        for (sc, sr), byte in sps.items():
            has_nb = any((sc + dc, sr + dr) in sps for dc in (-1, 0, 1) for dr in (-1, 0, 1) if dc or dr)
            iso = {"true": not has_nb, "never": False, "random": bool(rng.integers(2)) if rng is not None else False}[flag]
            source = sensor if version == 2 else sensor >> 1
            w = layout.sp_word(sc, sr, byte, iso, (sensor & 1) if version == 4 else 0)
            banks.setdefault(source, []).append(w)
    if rng is not None:
        for s in banks:
            rng.shuffle(banks[s])
    return banks


# ---------------------------------------------------------------------------- K2 known answers
def _run_single(banks_by_source, version, geometry, beamline):
    ev = layout.make_events([layout.full_event(banks_by_source, version)], version)
    return oracle.run(ev, geometry, beamline)


def _candidates(r):
    return sorted(r["cluster_candidates"][:int(r["module_candidate_num"][0])].tolist())


def test_estimate_scans_isolated_words_as_neighbours(geometry, beamline):
    """EstimateInputSize.cu:113 tests the current word's isolation flag (always clear on that
    branch), so an isolated-flagged word directly above a candidate pixel still suppresses it."""
    low = layout.sp_word(10, 5, 0x08)                 # pixel (20, 23): col 0, row 3 of SP (10, 5)
    top = layout.sp_word(10, 6, 0x01, isolated=True)  # pixel (20, 24): the pixel above it
    r = _run_single({0: [low, top]}, 4, geometry, beamline)
    # the low word yields no candidate (:162-171, its up neighbour is set); the isolated word
    # counts its one cluster (:58-71): 1 in module pair 0 (a scan skipping isolated words gives 2)
    assert r["estimated_input_size"][0] == 1
    assert r["module_candidate_num"][0] == 0


def test_estimate_has_no_sensor_filter(geometry, beamline):
    """EstimateInputSize.cu:111-141 scans both sensors of a v4 bank: a pixel of the other sensor
    to the right of a candidate suppresses it. The clustering (MaskedVeloClustering.cu:364-368)
    filters by sensor, so that cluster is lost."""
    s0 = layout.sp_word(10, 5, 0x80, sensor_bit=0)    # pixel (21, 23) of sensor 0
    s1 = layout.sp_word(11, 5, 0x08, sensor_bit=1)    # pixel (22, 23) of sensor 1, right of it
    r = _run_single({0: [s0, s1]}, 4, geometry, beamline)
    # only sensor 1's pixel is a candidate (word 1, sensor 1, pixel bit 3 -> code 4): estimate 1
    assert r["estimated_input_size"][0] == 1
    assert _candidates(r) == [(1 << 11) | (1 << 3) | 4]
    assert r["module_cluster_num"][0] == 1


def test_estimate_candidate_encoding_and_halves(geometry, beamline):
    """One candidate per half of a super pixel (EstimateInputSize.cu:173-204), encoded as
    (word << 11) | (sensor << 3) | (clz(bit) - 24), i.e. 7 - pixel bit; a non-isolated word
    with two separate pixels yields two candidates and two clusters."""
    pad = layout.sp_word(30, 40, 0x01)
    w = layout.sp_word(12, 7, 0x81)                   # pixels (24, 28) and (25, 31)
    r = _run_single({5: [pad, w]}, 4, geometry, beamline)
    sensor = 5 << 1
    mp = sensor // 8
    assert _candidates(r) == sorted(
        [(0 << 11) | (sensor << 3) | 7, (1 << 11) | (sensor << 3) | 7, (1 << 11) | (sensor << 3) | 0])
    assert r["estimated_input_size"][mp] == 3
    assert r["module_cluster_num"][mp] == 3


def test_estimate_row_does_not_wrap(geometry, beamline):
    """sp_row - 1 wraps to a large unsigned value at row 0 (EstimateInputSize.cu:128-130): a
    word at row 63 of the next column is not a right-bottom neighbour."""
    w0 = layout.sp_word(10, 0, 0x10)                  # pixel (21, 0)
    w1 = layout.sp_word(11, 63, 0x08)                 # pixel (22, 255)
    r = _run_single({0: [w0, w1]}, 4, geometry, beamline)
    assert r["estimated_input_size"][0] == 2 and r["module_candidate_num"][0] == 2


def test_estimate_v2_per_sensor_banks(geometry, beamline):
    """v2 banks hold one sensor each: the module pair is sourceID / 8 and the sensor is the
    sourceID (EstimateInputSize.cu:19-22, 36-38)."""
    w = layout.sp_word(100, 20, 0x33)                 # a 2x2 block: one candidate, one cluster
    r = _run_single({37: [w, layout.sp_word(200, 50, 0x01)]}, 2, geometry, beamline)
    assert r["estimated_input_size"][37 // 8] == 2
    assert _candidates(r) == sorted(
        [(0 << 11) | (37 << 3) | (7 - 5), (1 << 11) | (37 << 3) | 7])


# ---------------------------------------------------------------------------- shapes
def polyplets(n_max):
    """Every 8-connected pixel shape of 1..n_max pixels, normalised to start at (0, 0)."""
    shapes = {frozenset([(0, 0)])}
    out = set(shapes)
    for _ in range(n_max - 1):
        grown = set()
        for s in shapes:
            for (c, r) in s:
                for dc in (-1, 0, 1):
                    for dr in (-1, 0, 1):
                        q = (c + dc, r + dr)
                        if q in s:
                            continue
                        t = s | {q}
                        mc, mr = min(p[0] for p in t), min(p[1] for p in t)
                        grown.add(frozenset((p[0] - mc, p[1] - mr) for p in t))
        shapes = grown
        out |= grown
    return sorted(out, key=lambda s: (len(s), sorted(s)))


def place(shapes, version, flag, rng=None, spacing=(8, 6)):
    """Events holding the shapes on a grid of super-pixel cells (one shape per cell, cells far
    enough apart that no clustering window sees two shapes), at every alignment of the shape's
    origin within a super pixel."""
    sensors = list(range(0, 208, 5))
    cells = [(s, cc, rr) for s in sensors for cc in range(2, 384 // spacing[0] - 1) for rr in range(1, 64 // spacing[1] - 1)]
    items = [(shape, dc, dr) for shape in shapes for dc in range(2) for dr in range(4)]
    events, k = [], 0
    while k < len(items):
        pix = {}
        for (s, cc, rr) in cells:
            if k >= len(items):
                break
            shape, dc, dr = items[k]
            k += 1
            c0, r0 = cc * spacing[0] * 2 + dc, rr * spacing[1] * 4 + dr
            pix.setdefault(s, set()).update((c0 + c, r0 + r) for c, r in shape)
        by_source = encode(pix, version, flag, rng)
        events.append(layout.full_event(by_source, version))
    return layout.make_events(events, version)


@pytest.mark.parametrize("flag", ["true", "never"])
def test_every_small_shape_at_every_alignment(flag, geometry, geo, beamline):
    """All 135 shapes of up to four pixels at the 8 alignments of their origin in a super pixel;
    with correct isolation flags and with every word flagged non-isolated (single-SP shapes then
    go through rest_of_clusters). Every shape is also exactly one textbook 8-connected cluster."""
    shapes = polyplets(4)
    assert len(shapes) == 1 + 4 + 20 + 110
    ev = place(shapes, 4, flag)
    n = check_against_statement(ev, geometry, geo, beamline)
    r = oracle.run(ev, geometry, beamline)
    assert n == int(r["module_cluster_num"].sum()) == len(shapes) * 8


def _line(dc, dr, n):
    return frozenset((i * dc - min(0, (n - 1) * dc), i * dr - min(0, (n - 1) * dr)) for i in range(n))


def test_long_shapes_reach_the_window_edges(geometry, geo, beamline):
    """Lines in the eight directions and hooks of up to 12 pixels reach the window's top and bottom
    rows, its west-most column, the east column and the rows with asymmetric growth masks; the
    oracle's keep/drop decisions and the clusters it keeps equal the statement's."""
    shapes = set()
    for n in range(2, 13):
        for d in ((1, 0), (0, 1), (1, 1), (1, -1)):
            shapes.add(_line(*d, n))
        for m in range(1, 5):  # hooks: a vertical bar with a foot to the west or east
            bar = [(m, r) for r in range(n)]
            shapes.add(frozenset(bar + [(m - k, 0) for k in range(1, m + 1)]))
            shapes.add(frozenset(bar + [(m + k, n - 1) for k in range(1, 4)]))
            shapes.add(frozenset(bar + [(m - k, n - 1 - k) for k in range(1, m + 1)]))
    shapes = sorted(shapes, key=lambda s: (len(s), sorted(s)))
    for flag in ("true", "never"):
        ev = place(shapes, 4, flag, spacing=(10, 8))
        n = check_against_statement(ev, geometry, geo, beamline)
        assert n > 0


@pytest.mark.parametrize("version", [4, 2])
def test_random_maps_with_adversarial_flags(version, geometry, geo, beamline):
    """Dense random pixel maps on patches of a few super pixels, both sensors of a pair
    overlapping (v4), with correct, absent and random isolation flags and shuffled word order."""
    rng = np.random.default_rng(7 + version)
    events = []
    for e in range(12):
        pix = {}
        for _ in range(30):
            sensor = int(rng.integers(0, 208))
            c0, r0 = int(rng.integers(0, 760)), int(rng.integers(0, 250))
            w, h = int(rng.integers(2, 10)), int(rng.integers(2, 16))
            density = float(rng.uniform(0.1, 0.6))
            for c in range(c0, min(c0 + w, 768)):
                for r in range(r0, min(r0 + h, 256)):
                    if rng.random() < density:
                        pix.setdefault(sensor, set()).add((c, r))
                        if version == 4 and rng.random() < 0.3:
                            pix.setdefault(sensor ^ 1, set()).add((c, r))
        flag = ["true", "never", "random"][e % 3]
        events.append(layout.full_event(encode(pix, version, flag, rng), version))
    ev = layout.make_events(events, version)
    n = check_against_statement(ev, geometry, geo, beamline)
    assert n > 1000


def test_synthetic_events_against_statement(geometry, geo, beamline):
    """The generator's pile-up events (real flags, noise, cluster shapes of the digitisation)."""
    ev = synth.generate(n_events=6, pileup=1, nu=7.6, seed=5, bank_version=4)
    assert check_against_statement(ev, geometry, geo, beamline) > 4000
