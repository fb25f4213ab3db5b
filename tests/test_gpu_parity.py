"""The HIP path against the CPU oracle on seeded synthetic events (runs on an MI355X)."""
import numpy as np
import pytest

import oracle
from allen_amd import synth, velo
from tests import parity

pytestmark = pytest.mark.gpu

# Small, seeded datasets that the oracle finishes in seconds.
DATASETS = {
    "single_pv": dict(n_events=30, pileup=0, mean_tracks=120.0, seed=11),
    "pileup": dict(n_events=120, pileup=1, nu=7.6, seed=12),
    "shuffled_backward": dict(n_events=30, pileup=0, mean_tracks=150.0, backward_fraction=0.3, shuffle_sp=1,
                              seed=13),
    "v2_banks": dict(n_events=20, pileup=0, mean_tracks=100.0, bank_version=2, seed=14),
    "empty_events": dict(n_events=6, pileup=0, mean_tracks=0.0, noise_fraction=0.0, seed=15),
    "stress": dict(n_events=12, pileup=1, nu=15.0, noise_fraction=0.05, dead_asic_fraction=0.01, seed=16),
    # The same kind of events in the MEP layout (MEPTools.h:59-200), super pixels and retina clusters.
    "mep_pileup": dict(n_events=40, pileup=1, nu=7.6, seed=18),
    "mep_v2_banks": dict(n_events=20, pileup=0, mean_tracks=100.0, bank_version=2, seed=19),
    "mep_retina_v3": dict(n_events=30, pileup=1, nu=7.6, retina=1, bank_version=3, seed=20),
}


@pytest.fixture(scope="module")
def constants(geometry, beamline):
    c = velo.VeloConstants(geometry, beamline)
    yield c
    c.close()


def run_both(name, geometry, beamline, constants, fused, tuning=None, oracle_props=None, events=None):
    ev = events if events is not None else synth.generate(**DATASETS[name])
    if name.startswith("mep_"):
        ev = synth.to_mep(ev)
    print(f"[parity] dataset={name} events={ev.n_events} raw_bytes={ev.raw.size} fused={fused} tuning={tuning}",
          flush=True)
    ref = oracle.run(ev, geometry, beamline, **(oracle_props or {}))
    dev = velo.DeviceEvents(ev)
    seq = velo.VeloSequence(constants, ev.n_events, max(ev.raw.size, 4096))
    seq.set_tuning("fused_clustering", int(fused))
    for k, v in (tuning or {}).items():
        seq.set_tuning(k, v)
    seq.run(dev)
    status = seq.status()
    gpu = parity.gpu_buffers(seq, ref)
    seq.close()
    return ref, gpu, status


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per_algorithm"])
@pytest.mark.parametrize("name", list(DATASETS))
def test_sequence_matches_oracle(name, fused, geometry, beamline, constants):
    ref, gpu, status = run_both(name, geometry, beamline, constants, fused)
    assert int(ref["status"][0]) == 0, "the oracle flags a capacity overflow; pick a smaller dataset"
    assert status & velo.STATUS_OVERFLOW_MASK == 0
    errs = parity.compare(gpu, ref, fused)
    assert not errs, "\n".join(errs[:20])
    if name != "empty_events":
        assert ref["offsets_all_velo_tracks"][-1] > 0


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per_algorithm"])
def test_real_sp_banks_match_oracle(fused, geometry, beamline, constants):
    """The reference's bundled MDF file's real super-pixel banks (10 minimum-bias events, bank
    version 2, 208 banks per event): pixel maps of real data, with clusters at the edges of the
    masked window and larger than it, through the whole HIP path against the oracle."""
    import os
    from tests import textbook
    ev = textbook.load_mdf_fixture(os.path.join(os.path.dirname(__file__), "golden", "mdf_minbias_vp.npz"))
    assert ev.bank_version == 2 and ev.n_events == 10
    ref, gpu, status = run_both("mdf_minbias_vp", geometry, beamline, constants, fused, events=ev)
    assert int(ref["status"][0]) == 0
    assert status & velo.STATUS_OVERFLOW_MASK == 0
    errs = parity.compare(gpu, ref, fused)
    assert not errs, "\n".join(errs[:20])
    # The same real banks in the MEP layout.
    ref, gpu, status = run_both("mep_mdf_minbias_vp", geometry, beamline, constants, fused, events=ev)
    assert int(ref["status"][0]) == 0
    assert status & velo.STATUS_OVERFLOW_MASK == 0
    errs = parity.compare(gpu, ref, fused)
    assert not errs, "\n".join(errs[:20])
    assert ref["module_cluster_num"].sum() > 15000 and ref["offsets_all_velo_tracks"][-1] > 1000


def test_batched_persistent_graph_replays(geometry, beamline, constants):
    """The batched workload's event mix (one vertex with Poisson(120) tracks, 5 % of the events at
    nu = 20 pile-up) through the persistent work-stealing search, the whole sequence captured as
    one HIP graph and replayed three times with eager runs of other events in between: every
    replay equals the oracle bit for bit."""
    import torch
    cfg = dict(n_events=400, pileup=0, mean_tracks=120.0, high_occupancy_fraction=0.05, high_nu=20.0, seed=23)
    ev = synth.generate(**cfg)
    other = synth.generate(**dict(cfg, n_events=100, seed=24))
    print(f"[batched-graph] events={ev.n_events} raw_bytes={ev.raw.size}", flush=True)
    ref = oracle.run(ev, geometry, beamline)
    assert int(ref["status"][0]) == 0
    dev, dev_other = velo.DeviceEvents(ev), velo.DeviceEvents(other)
    seq = velo.VeloSequence(constants, ev.n_events, max(ev.raw.size, other.raw.size))
    seq.set_tuning("search_persistent", 1)
    stream = torch.cuda.Stream()
    seq.capture(dev, stream)
    stream.synchronize()
    for replay in range(3):
        seq.run(dev_other)
        torch.cuda.synchronize()
        seq.replay(stream)
        stream.synchronize()
        assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
        errs = parity.compare(parity.gpu_buffers(seq, ref), ref, True)
        assert not errs, f"replay {replay}:\n" + "\n".join(errs[:20])
    seq.close()


def test_batched_mix_graph_replays(geometry, beamline, constants):
    """The batched workload's event mix (one vertex with Poisson(120) tracks, 5 % of the events at
    nu = 20 pile-up; events of up to ~5000 hits take the mixed LDS layout) through the whole
    sequence captured as one HIP graph, as bench.py --graph runs it (one workgroup per event), and
    replayed three times with eager runs of other events in between: every replay equals the
    oracle bit for bit."""
    import torch
    cfg = dict(n_events=400, pileup=0, mean_tracks=120.0, high_occupancy_fraction=0.05, high_nu=20.0, seed=23)
    ev = synth.generate(**cfg)
    other = synth.generate(**dict(cfg, n_events=100, seed=24))
    print(f"[batched-graph] events={ev.n_events} raw_bytes={ev.raw.size}", flush=True)
    ref = oracle.run(ev, geometry, beamline)
    assert int(ref["status"][0]) == 0
    dev, dev_other = velo.DeviceEvents(ev), velo.DeviceEvents(other)
    seq = velo.VeloSequence(constants, ev.n_events, max(ev.raw.size, other.raw.size))
    seq.run(dev)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    seq.capture(dev, stream)
    stream.synchronize()
    for replay in range(3):
        seq.run(dev_other)
        torch.cuda.synchronize()
        seq.replay(stream)
        stream.synchronize()
        assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
        errs = parity.compare(parity.gpu_buffers(seq, ref), ref, True)
        assert not errs, f"replay {replay}:\n" + "\n".join(errs[:20])
    seq.close()


TUNINGS = {
    "block64": {"search_block": 64},
    "block1024": {"search_block": 1024, "search_lds_hits": 2048},
    "mixed_lds": {"search_block": 256, "search_lds_hits": 600},   # phi + used in LDS, xyz global
    "global_hits": {"search_block": 128, "search_lds_hits": 64},  # everything in global memory
    "event_order": {"search_largest_first": 0},                   # events in their own order
    "persistent": {"search_persistent": 1},                       # resident grid, work stealing
    "persistent_block64": {"search_persistent": 1, "search_block": 64, "search_largest_first": 0},
    "clustering_resident": {"clustering_resident_grid": 1},       # clustering waves loop over items
    "split": {"search_split": 32},                                # largest 1/32 of the events on a second stream
    # the smallest LDS window of the ring and list of unused hits: forwarding reads the ring back
    # from HBM and seeding walks the pendulum in module pairs of more than 64 hits
    "small_window_list": {"search_ring_window": 64, "search_list_cap": 64},
    # half the events in the side stream's launch, small workgroups and LDS there (mixed layout)
    "split_half": {"search_split": 2, "search_big_block": 128, "search_big_lds_hits": 600},
    # consolidation's binary-search path (events beyond 2047 tracks or 4096 track hits take it)
    "consolidate_search": {"consolidate_lds_map": 0},
}


@pytest.mark.parametrize("tuning", list(TUNINGS))
def test_sequence_tuning_variants(tuning, geometry, beamline, constants):
    """Launch geometries and LDS layouts do not change a single bit of the result."""
    ref, gpu, status = run_both("pileup", geometry, beamline, constants, True, TUNINGS[tuning])
    assert status & velo.STATUS_OVERFLOW_MASK == 0
    errs = parity.compare(gpu, ref, True)
    assert not errs, "\n".join(errs[:20])


def test_kalman_fp64_covariance_fallback(geometry, beamline, constants):
    """With the FP64 covariance fallback on, the GPU states and covariances equal the oracle's
    fallback bit for bit (same FP64 operations, no contraction on either side)."""
    ref, gpu, status = run_both("pileup", geometry, beamline, constants, True, {"kalman_fp64_fallback": 1},
                                {"kalman_fp64_fallback": 1})
    assert status & velo.STATUS_OVERFLOW_MASK == 0
    errs = parity.compare(gpu, ref, True)
    assert not errs, "\n".join(errs[:20])


def _normalised_cov_error(c32, c64):
    """|c32 - c64| per stored term; diagonal terms relative, off-diagonal ones over sqrt(c_ii c_jj)."""
    pairs = [(0, 0), (0, 2), (2, 2), (3, 3), (3, 5), (5, 5)]
    e = np.empty_like(c64)
    for k, (i, j) in enumerate(pairs):
        e[:, k] = np.abs(c32[:, k] - c64[:, k]) / np.sqrt(np.abs(c64[:, i] * c64[:, j]))
    return e


def test_stress_kalman_against_fp64(geometry, beamline, constants):
    """High-occupancy events (nu = 15, 5 % noise pixels, 1 % dead ASICs). The GPU fit equals the
    oracle's bit for bit without and with the FP64 covariance fallback. Against the FP64
    restatement of the fit, the FP32 covariance of the plain fit is off by percents on
    ill-conditioned steps (three-hit tracks included); with the fallback every track's covariance
    is within FP32 storage precision, and the states within 2e-4 mm and 1e-6 in slope."""
    cfg = dict(n_events=40, pileup=1, nu=15.0, noise_fraction=0.05, dead_asic_fraction=0.01, seed=17)
    ev = synth.generate(**cfg)
    print(f"[stress-kalman] events={ev.n_events} raw_bytes={ev.raw.size}", flush=True)
    dev = velo.DeviceEvents(ev)
    out = {}
    for fb in (0, 1):
        ref = oracle.run(ev, geometry, beamline, kalman_fp64_fallback=fb)
        seq = velo.VeloSequence(constants, ev.n_events, ev.raw.size)
        seq.set_tuning("kalman_fp64_fallback", fb)
        seq.run(dev)
        assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
        gpu = parity.gpu_buffers(seq, ref)
        seq.close()
        errs = parity.compare(gpu, ref, True)
        assert not errs, "\n".join(errs[:20])
        out[fb] = ref
    ref = out[1]
    n = int(ref["offsets_all_velo_tracks"][-1])
    hits = np.diff(ref["offsets_velo_track_hit_number"][:n + 1].astype(np.int64))
    three = hits == 3
    assert n > 5000 and three.sum() > 50
    for name in ["velo_kalman_beamline_states", "velo_kalman_endvelo_states"]:
        f64 = ref[name + "_f64"][:14 * n]
        c64 = f64[6 * n:].reshape(-1, 8)[:, :6]
        s64 = f64[:6 * n].reshape(-1, 6)[:, :5]
        e = {fb: _normalised_cov_error(out[fb][name][6 * n:14 * n].reshape(-1, 8)[:, :6].astype(np.float64), c64)
             for fb in (0, 1)}
        d = np.abs(out[1][name][:6 * n].reshape(-1, 6)[:, :5].astype(np.float64) - s64)
        print(f"[stress-kalman] {name}: {n} tracks ({three.sum()} three-hit); covariance error vs FP64 "
              f"plain q99 {np.quantile(e[0], 0.99):.3g} max {e[0].max():.3g} (three-hit max {e[0][three].max():.3g}); "
              f"fallback max {e[1].max():.3g} (three-hit {e[1][three].max():.3g}); state error with fallback: "
              f"x,y {d[:, :2].max():.3g} mm, slopes {d[:, 3:].max():.3g}", flush=True)
        assert e[1].max() < 1e-5
        assert e[0][three].max() > 1e-3 > e[1][three].max()
        assert d[:, 3:].max() < 1e-6 and d[:, :2].max() < 2e-4


def test_captured_sequence_replays(geometry, beamline, constants):
    """The whole sequence captured as one HIP graph: three replays, with an eager run on other
    events in between (which overwrites the buffers), each reproduce the eager result and the
    oracle bit for bit."""
    import torch
    ev = synth.generate(**DATASETS["pileup"])
    other = synth.generate(**dict(DATASETS["pileup"], n_events=40, seed=99))
    print(f"[graph] events={ev.n_events} other={other.n_events}", flush=True)
    ref = oracle.run(ev, geometry, beamline)
    dev, dev_other = velo.DeviceEvents(ev), velo.DeviceEvents(other)
    seq = velo.VeloSequence(constants, ev.n_events, max(ev.raw.size, other.raw.size))
    seq.run(dev)
    torch.cuda.synchronize()
    eager = parity.gpu_buffers(seq, ref)
    stream = torch.cuda.Stream()
    seq.capture(dev, stream)
    stream.synchronize()
    for replay in range(3):
        seq.run(dev_other)
        torch.cuda.synchronize()
        seq.replay(stream)
        stream.synchronize()
        got = parity.gpu_buffers(seq, ref)
        # Same rules as the oracle comparison: filled slots only (gap contents are unspecified and
        # keep what the run on the other events left there).
        errs = parity.compare(got, eager, True)
        assert not errs, f"replay {replay} against the eager run:\n" + "\n".join(errs[:20])
        errs = parity.compare(got, ref, True)
        assert not errs, f"replay {replay} against the oracle:\n" + "\n".join(errs[:20])
    assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
    seq.close()


def test_autotune_keeps_outputs(geometry, beamline, constants):
    """A reduced autotuner sweep: every candidate reproduces the default outputs (the tuner raises
    otherwise) and the chosen knobs are applied to the sequence."""
    from allen_amd import autotune
    ev = synth.generate(**DATASETS["pileup"])
    dev = velo.DeviceEvents(ev)
    seq = velo.VeloSequence(constants, ev.n_events, ev.raw.size)
    res = autotune.tune(seq, dev, reps=1, blocks=(64, 256), lds_hits=(512, 1760))
    assert len(res.trials) == 16 and res.search_ms <= res.default_search_ms
    assert res.knobs["search_block"] in (64, 256)
    seq.run(dev)
    ref = oracle.run(ev, geometry, beamline)
    errs = parity.compare(parity.gpu_buffers(seq, ref), ref, True)
    assert not errs, "\n".join(errs[:20])
    seq.close()


def test_throughput_autotune_on_two_streams(geometry, beamline, constants):
    """The throughput tuner over two sequences on two streams: every candidate keeps the outputs,
    the winner is applied to both, and both then reproduce the oracle."""
    import torch
    from allen_amd import autotune
    ev = synth.generate(**DATASETS["pileup"])
    print(f"[autotune-throughput] events={ev.n_events}", flush=True)
    dev = velo.DeviceEvents(ev)
    seqs = [velo.VeloSequence(constants, ev.n_events, ev.raw.size) for _ in range(2)]
    streams = [torch.cuda.Stream() for _ in range(2)]

    def measure():
        torch.cuda.synchronize()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        main = torch.cuda.current_stream()
        start.record(main)
        for st in streams:
            st.wait_event(start)
        for i in range(4):
            seqs[i % 2].run(dev, stream=streams[i % 2])
        for st in streams:
            main.wait_stream(st)
        end.record(main)
        torch.cuda.synchronize()
        return start.elapsed_time(end) / 4

    cands = ({"search_block": 128, "search_lds_hits": 512}, {"search_block": 192, "search_lds_hits": 800})
    res = autotune.tune_throughput(seqs, measure, ev.n_events, candidates=cands)
    assert len(res.trials) == 2 and res.search_ms <= res.default_search_ms
    ref = oracle.run(ev, geometry, beamline)
    for sq, st in zip(seqs, streams):
        sq.run(dev, stream=st)
        torch.cuda.synchronize()
        errs = parity.compare(parity.gpu_buffers(sq, ref), ref, True)
        assert not errs, "\n".join(errs[:20])
        sq.close()


def test_dense_module_pairs_take_the_tiled_paths(geometry, beamline, constants):
    """Events dense enough that module pairs exceed the per-wave LDS slices of clustering."""
    DATASETS["dense"] = dict(n_events=4, pileup=1, nu=60.0, noise_fraction=0.3, seed=17)
    ref, gpu, status = run_both("dense", geometry, beamline, constants, True)
    est = ref["offsets_estimated_input_size"]
    assert np.diff(est).max() > 320  # some module pair is above the cluster slice
    if int(ref["status"][0]) == 0:
        errs = parity.compare(gpu, ref, True)
        assert not errs, "\n".join(errs[:20])
    else:  # capacity overflow in the search: compare the clustering stages only
        errs = [e for e in parity.compare(gpu, ref, True) if e.split(":")[0].startswith(("sorted", "module_cluster",
                                                                                          "estimated", "offsets_est"))]
        assert not errs, "\n".join(errs[:20])


def test_prefix_sum_iota(constants):
    """The reference's own prefix-sum test shape (TestPrefixSum.cu:9-33): iota 1..1000."""
    import torch
    x = torch.arange(1, 1001, dtype=torch.int32, device="cuda")
    out = torch.empty(1001, dtype=torch.int32, device="cuda")
    total = velo.prefix_sum(x, 1000, out, want_total=True)
    expect = np.concatenate([[0], np.cumsum(np.arange(1, 1001))])
    assert total == expect[-1]
    np.testing.assert_array_equal(out.cpu().numpy(), expect)


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 26 * 1000, 32768, 32769, 26 * 20000 + 3])
def test_prefix_sum_sizes(n, constants):
    import torch
    rng = np.random.default_rng(n)
    host = rng.integers(0, 1 << 20, size=n, dtype=np.int64).astype(np.uint32)
    x = torch.from_numpy(host.view(np.int32)).cuda()
    out = torch.full((n + 1,), -1, dtype=torch.int32, device="cuda")
    total = velo.prefix_sum(x, n, out, want_total=True)
    expect = np.zeros(n + 1, np.uint32)
    oracle.lib().vo_prefix_sum(host.ctypes.data, n, expect.ctypes.data)
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), expect)
    assert total == int(expect[-1])


def test_constants_tables_match_oracle(constants):
    mine = constants.sp_tables()
    ref = oracle.sp_tables()
    for a, b in zip(mine, ref):
        np.testing.assert_array_equal(a, b)
