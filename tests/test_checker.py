"""The VELO track checker (allen_amd/checker.py) against a per-track loop restatement of the same
rules, and on events whose tracks are known exactly."""
import numpy as np

import oracle
from allen_amd import checker, synth


def _loop_check(ev, r):
    """Plain-Python restatement: majority particle per track, 70 % rule, 01_velo particles."""
    off_all = r["offsets_all_velo_tracks"]
    off_h = r["offsets_velo_track_hit_number"]
    n = int(off_all[-1])
    ids = r["velo_track_hits"][:4 * int(off_h[n])].view("<u4")
    rec = {i for i, p in enumerate(ev.particles) if p["n_modules"] >= 3 and 2.0 < p["eta"] < 5.0}
    matched, ghosts = {}, 0
    for e in range(ev.n_events):
        h = ev.hits[ev.hit_offsets[e]:ev.hit_offsets[e + 1]]
        truth = {int(x["lhcb_id"]): int(x["particle"]) for x in h if x["particle"] != 0xFFFFFFFF}
        for t in range(off_all[e], off_all[e + 1]):
            ps = [truth.get(int(i), -1) for i in ids[off_h[t]:off_h[t + 1]]]
            cand = [p for p in ps if p >= 0]
            if not cand:
                ghosts += 1
                continue
            counts = {p: cand.count(p) for p in set(cand)}
            best = max(counts.values())
            p = min(q for q, c in counts.items() if c == best)
            if best / len(ps) >= 0.7:
                matched.setdefault(p, []).append(t)
            else:
                ghosts += 1
    reco = [p for p in matched if p in rec]
    clones = sum(len(matched[p]) - 1 for p in reco)
    return len(rec), len(reco), n, ghosts, clones


def test_checker_against_loop(geometry, beamline):
    ev = synth.generate(n_events=25, pileup=1, nu=7.6, seed=41)
    r = oracle.run(ev, geometry, beamline)
    res = checker.check_buffers(ev, r["offsets_all_velo_tracks"], r["offsets_velo_track_hit_number"],
                                r["velo_track_hits"])
    n_rec, n_reco, n_tr, ghosts, clones = _loop_check(ev, r)
    assert (res.reconstructible, res.reconstructed, res.tracks, res.ghosts, res.clones) == \
        (n_rec, n_reco, n_tr, ghosts, clones)
    assert 0.9 < res.efficiency <= 1.0 and res.ghost_rate < 0.1


def test_checker_on_clean_single_tracks(geometry, beamline):
    """Noiseless single-pixel tracks: everything reconstructible is found, no ghosts, pure hits."""
    ev = synth.generate(n_events=100, mean_tracks=2.0, noise_fraction=0.0, hit_efficiency=1.0,
                        cluster_size_p=(1.0, 0.0, 0.0), seed=23)
    r = oracle.run(ev, geometry, beamline)
    res = checker.check_buffers(ev, r["offsets_all_velo_tracks"], r["offsets_velo_track_hit_number"],
                                r["velo_track_hits"])
    assert res.ghosts == 0 and res.hit_purity == 1.0
    assert res.efficiency > 0.97


class _OracleBuffers:
    """The fetch() of a GPU sequence, served from an oracle run (same buffer names)."""

    def __init__(self, r):
        self.r = r

    def fetch(self, name, dtype, count):
        a = np.ascontiguousarray(self.r[name]).view(np.uint8).view(dtype)
        return a[:count]


def test_seed_labels_against_loop(geometry, beamline):
    """scorer.seed_labels on the oracle's own seed log: a seed is positive iff a consolidated
    track starts with its three hits and is matched (>= 70 % of hits from one particle)."""
    from allen_amd import scorer
    ev = synth.generate(n_events=12, pileup=1, nu=7.6, seed=43)
    r = oracle.run(ev, geometry, beamline, collect_seeds=1)
    rec = r["seed_log"].reshape(-1, 5).astype(np.int64)
    lab = scorer.seed_labels(ev, _OracleBuffers(r), rec)
    # Loop restatement.
    est = r["offsets_estimated_input_size"].astype(np.int64)
    total = int(est[-1])
    ids = r["sorted_velo_cluster_container"][:4 * total].view("<u4")
    off_all = r["offsets_all_velo_tracks"]
    off_h = r["offsets_velo_track_hit_number"]
    tids = r["velo_track_hits"][:4 * int(off_h[-1])].view("<u4")
    res = checker.check(ev, off_all, off_h, tids)
    want = []
    for e, _, h0, h1, h2 in rec:
        trip = [int(ids[est[26 * e] + h]) for h in (h0, h1, h2)]
        hit = 0
        for t in range(off_all[e], off_all[e + 1]):
            if tids[off_h[t]:off_h[t] + 3].tolist() == trip and res.track_particle[t] >= 0:
                hit = 1
        want.append(hit)
    np.testing.assert_array_equal(lab, want)
    assert 0.3 < lab.mean() <= 1.0 and lab.sum() > 100


def _loop_categories(ev, r):
    """Per category: (accepted, found, clones, mean per-event efficiency), by plain loops."""
    off_all = r["offsets_all_velo_tracks"]
    off_h = r["offsets_velo_track_hit_number"]
    ids = r["velo_track_hits"][:4 * int(off_h[int(off_all[-1])])].view("<u4")
    matched = {}
    for e in range(ev.n_events):
        h = ev.hits[ev.hit_offsets[e]:ev.hit_offsets[e + 1]]
        truth = {int(x["lhcb_id"]): int(x["particle"]) for x in h if x["particle"] != 0xFFFFFFFF}
        for t in range(off_all[e], off_all[e + 1]):
            ps = [truth.get(int(i), -1) for i in ids[off_h[t]:off_h[t + 1]]]
            for p in set(q for q in ps if q >= 0):
                if ps.count(p) / len(ps) >= 0.7:
                    matched.setdefault(p, []).append(t)
    out = {}
    for name in checker.CATEGORIES:
        acc = []
        for i, p in enumerate(ev.particles):
            velo = p["n_modules"] >= 3 and 2.0 < p["eta"] < 5.0
            long = velo and abs(p["tx"]) < 0.30 and abs(p["ty"]) < 0.25 and p["p"] > 2000.0
            ok = {"01_velo": velo, "02_long": long, "03_long_P>5GeV": long and p["p"] > 5000.0}[name]
            if ok:
                acc.append(i)
        found = [i for i in acc if i in matched]
        clones = sum(len(matched[i]) - 1 for i in found)
        per = {}
        for i in acc:
            e = int(ev.particles[i]["event"])
            a, f = per.get(e, (0, 0))
            per[e] = (a + 1, f + (i in matched))
        eff_evt = sum(f / a for a, f in per.values()) / len(per) if per else 0.0
        out[name] = (len(acc), len(found), clones, eff_evt)
    return out


def test_long_categories_against_loop(geometry, beamline):
    ev = synth.generate(n_events=25, pileup=1, nu=7.6, seed=42)
    r = oracle.run(ev, geometry, beamline)
    res = checker.check_buffers(ev, r["offsets_all_velo_tracks"], r["offsets_velo_track_hit_number"],
                                r["velo_track_hits"])
    want = _loop_categories(ev, r)
    for name, (acc, found, clones, eff_evt) in want.items():
        c = res.categories[name]
        assert (c.accepted, c.found, c.clones) == (acc, found, clones), name
        assert abs(c.efficiency_per_event - eff_evt) < 1e-12, name
    assert 0 < want["03_long_P>5GeV"][0] < want["02_long"][0] < want["01_velo"][0]
    lines = res.report().splitlines()
    assert lines[0].startswith("TrackChecker output") and len(lines) == 4


def test_report_lines_in_the_reference_format():
    """The reference's veloSP validation log on an EPYC 7502
    (test/reference/SciFiv6_upgrade_DC19_01_Bs2PhiPhiMD_1k_newLHCbID_hlt1_pp_veloSP_validation_epyc7502.txt:2-4)
    printed from its own counts."""
    want = [
        "01_velo                                           :    113217/   115180  98.30% ( 98.38%),      2117 (  1.84%) "
        "clones, pur  99.72%, hit eff  96.97%",
        "02_long                                           :     67859/    68385  99.23% ( 99.26%),       882 (  1.28%) "
        "clones, pur  99.81%, hit eff  98.02%",
    ]
    got = [checker.Category("01_velo", 115180, 113217, 2117, 0.99721, 0.96971, 0.98381).line(),
           checker.Category("02_long", 68385, 67859, 882, 0.99811, 0.98021, 0.99261).line()]
    assert got == want
    ghost = checker.CheckResult(0, 0, 264911, 1842, 0, 0, 0.0, 0.0, np.zeros(0), np.zeros(0), {}).report()
    assert ghost == "TrackChecker output                               :      1842/   264911   0.70% ghosts"
