"""The seed scorer's MFMA kernels against the CPU restatement (oracle/mlp.py)."""
import numpy as np
import pytest

from oracle import mlp as ref

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 0.7, (n, 10)).astype(np.float32)
    y = (x[:, 0] + 0.5 * x[:, 3] - x[:, 7] > 0.1).astype(np.uint8)
    return x, y


def test_mlp_logits_match_reference():
    """Forward pass (the MFMA operand/result lane layout included) on 1000 rows, the last block
    partial."""
    import torch
    from allen_amd import scorer
    x, _ = _data(1000, 1)
    p = scorer.init_params(3)
    p[320:352] = 0.1
    p[1376:1408] = -0.05
    p[1440] = 0.2
    net = scorer.MLP(p)
    got = net.score(torch.from_numpy(x).cuda()).cpu().numpy()
    want = ref.forward(x, p)["z3"]
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)


def test_mlp_gradient_matches_reference():
    """Gradient of the mean cross-entropy (lr = 0: no update) against the restatement."""
    import torch
    from allen_amd import scorer
    x, y = _data(777, 2)
    p = scorer.init_params(5)
    net = scorer.MLP(p)
    loss, g = net.train_step(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), 0.0, want_grad=True)
    want = ref.gradient(x, y.astype(np.float64), p)
    scale = np.abs(want).max()
    np.testing.assert_allclose(g, want, rtol=2e-3, atol=2e-4 * scale)
    np.testing.assert_allclose(loss, ref.loss(x, y, p) / x.shape[0], rtol=1e-4)
    np.testing.assert_array_equal(net.host_params(), p)  # lr = 0 leaves the parameters


def test_mlp_training_reduces_loss():
    import torch
    from allen_amd import scorer
    x, y = _data(20000, 3)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    net = scorer.MLP(seed=7)
    first, _ = net.train_step(xd, yd, 3e-3)
    for _ in range(300):
        last, _ = net.train_step(xd, yd, 3e-3)
    assert last < 0.35 * first, (first, last)
    acc = ((net.score(xd).cpu().numpy() > 0) == (y == 1)).mean()
    assert acc > 0.97


def test_seed_log_matches_oracle(geometry, beamline):
    """The GPU seed log holds exactly the oracle's seeds (as a set: the GPU appends them in no
    particular order) with bit-identical features, and the plain outputs are unchanged."""
    import oracle
    from allen_amd import synth, velo
    from tests import parity
    ev = synth.generate(n_events=60, pileup=1, nu=7.6, seed=61)
    ref = oracle.run(ev, geometry, beamline, collect_seeds=1)
    constants = velo.VeloConstants(geometry, beamline)
    seq = velo.VeloSequence(constants, ev.n_events, ev.raw.size)
    n_ref = ref["seed_log"].size // 5
    seq.set_tuning("seed_log", n_ref + 100)
    seq.run(velo.DeviceEvents(ev))
    rec, feat, seen = seq.seed_log()
    assert seen == n_ref and rec.shape[0] == n_ref
    want_rec = ref["seed_log"].reshape(-1, 5)
    want_feat = ref["seed_features"].reshape(-1, 10)
    key = lambda r: np.lexsort(r.T[::-1])
    gi, wi = key(rec), key(want_rec)
    np.testing.assert_array_equal(rec[gi], want_rec[wi])
    np.testing.assert_array_equal(feat[gi].view(np.uint32), want_feat[wi].view(np.uint32))
    errs = parity.compare(parity.gpu_buffers(seq, ref), ref, True)
    assert not errs, "\n".join(errs[:20])
    seq.close()
    constants.close()


def _scored_run(geometry, beamline, ev, params, threshold, log_cap=0, block=0):
    from allen_amd import velo
    constants = velo.VeloConstants(geometry, beamline)
    seq = velo.VeloSequence(constants, ev.n_events, ev.raw.size)
    if block:
        seq.set_tuning("search_block", block)
    if log_cap:
        seq.set_tuning("seed_log", log_cap)
    seq.set_scorer(params, threshold)
    seq.run(velo.DeviceEvents(ev))
    return seq, constants


def test_scored_search_extreme_thresholds_equal_plain(geometry, beamline):
    """Threshold -inf keeps every seed; +inf would drop every group of seeds, which the plain
    fallback keeps: both reproduce the plain search (and the oracle) exactly."""
    import oracle
    from allen_amd import scorer, synth
    from tests import parity
    ev = synth.generate(n_events=40, pileup=1, nu=7.6, seed=71)
    ref = oracle.run(ev, geometry, beamline)
    p = scorer.init_params(9)
    for thr in (-1e30, 1e30):
        seq, constants = _scored_run(geometry, beamline, ev, p, thr)
        errs = parity.compare(parity.gpu_buffers(seq, ref), ref, True)
        assert not errs, f"threshold {thr}:\n" + "\n".join(errs[:20])
        seq.close()
        constants.close()


def test_scored_search_logits_match_score_kernel(geometry, beamline):
    """The logits computed inside search_by_triplet equal k_mlp_score's on the same features."""
    import torch
    from allen_amd import scorer, synth
    ev = synth.generate(n_events=40, pileup=1, nu=7.6, seed=72)
    p = scorer.init_params(11)
    seq, constants = _scored_run(geometry, beamline, ev, p, -1e30, log_cap=40000)
    rec, feat, seen = seq.seed_log()
    assert 0 < seen == rec.shape[0]
    logits = seq.seed_logits(seen)
    want = scorer.MLP(p).score(torch.from_numpy(feat).cuda()).cpu().numpy()
    np.testing.assert_allclose(logits, want, rtol=1e-5, atol=1e-5)
    seq.close()
    constants.close()


def test_scored_search_prunes(geometry, beamline):
    """A threshold at the median logit drops seeds: fewer tracks to follow, no overflow."""
    from allen_amd import scorer, synth, velo
    ev = synth.generate(n_events=40, pileup=1, nu=7.6, seed=73)
    p = scorer.init_params(12)
    seq, constants = _scored_run(geometry, beamline, ev, p, -1e30, log_cap=40000)
    _, _, seen = seq.seed_log()
    med = float(np.median(seq.seed_logits(seen)))
    plain_ttf = int(seq.fetch("atomics_velo", np.uint32, 3 * ev.n_events)[1::3].sum())
    seq.close()
    constants.close()
    seq, constants = _scored_run(geometry, beamline, ev, p, med)
    pruned_ttf = int(seq.fetch("atomics_velo", np.uint32, 3 * ev.n_events)[1::3].sum())
    assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
    assert pruned_ttf < plain_ttf
    seq.close()
    constants.close()


def _physics(seq, ev):
    from allen_amd import checker
    off = seq.fetch("offsets_all_velo_tracks", np.uint32, ev.n_events + 1)
    hn = seq.fetch("offsets_velo_track_hit_number", np.uint32, int(off[-1]) + 1)
    return checker.check_buffers(ev, off, hn, seq.fetch("velo_track_hits", np.uint8, 4 * int(hn[-1])))


def test_trained_scorer_gain(geometry, beamline):
    """The shipped scorer at its threshold, on events it was not trained on: the long-track
    efficiency stays within 0.2 % of the plain search's and the ghost rate drops. A network
    without training (random initialisation) pruning as many seeds removes fewer ghosts."""
    from allen_amd import scorer, synth, velo
    params, thr = scorer.trained()
    ev = synth.generate(n_events=2000, pileup=0, mean_tracks=120.0, seed=4242)
    constants = velo.VeloConstants(geometry, beamline)
    dev = velo.DeviceEvents(ev)
    seq = velo.VeloSequence(constants, ev.n_events, ev.raw.size)
    seq.run(dev)
    plain = _physics(seq, ev)
    n_seeds = int(seq.fetch("atomics_velo", np.uint32, 3 * ev.n_events)[1::3].sum())
    seq.set_scorer(params, thr)
    seq.run(dev)
    assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
    trained = _physics(seq, ev)
    kept = int(seq.fetch("atomics_velo", np.uint32, 3 * ev.n_events)[1::3].sum())
    assert kept < n_seeds
    assert trained.efficiency >= plain.efficiency - 0.002
    assert trained.ghosts < plain.ghosts
    # Random network: its logits of every seed (threshold -inf keeps them all), then the
    # threshold that prunes the same number of seeds.
    rnd = scorer.init_params(5)
    seq.set_tuning("seed_log", 200 * ev.n_events)
    seq.set_scorer(rnd, -1e30)
    seq.run(dev)
    _, _, seen = seq.seed_log()
    logits = np.sort(seq.seed_logits(seen))
    seq.set_tuning("seed_log", 0)
    seq.set_scorer(rnd, float(logits[n_seeds - kept]))
    seq.run(dev)
    untrained = _physics(seq, ev)
    assert plain.ghosts - untrained.ghosts < plain.ghosts - trained.ghosts
    print(f"plain {plain.summary()}\ntrained {trained.summary()}\nuntrained {untrained.summary()}")
    seq.close()
    constants.close()


def _logits_by_slot(seq, ref, n_events):
    """The scored search's logged logit of every seed, placed at its h1's cluster slot (NaN
    elsewhere), the oracle's input for the pruning rule."""
    rec, _, seen = seq.seed_log()
    assert rec.shape[0] == seen, "the seed log overflowed"
    logits = seq.seed_logits(seen)
    off = ref["offsets_estimated_input_size"]
    out = np.full(int(off[-1]), np.nan, np.float32)
    slot = off[26 * rec[:, 0].astype(np.int64)].astype(np.int64) + rec[:, 3]
    assert np.unique(slot).size == slot.size, "two seeds share an h1"
    out[slot] = logits
    return out


@pytest.mark.parametrize("n_events,seed,block", [(300, 81, 256), (120, 82, 256), (300, 83, 128), (120, 84, 64)])
def test_pruned_search_matches_oracle_at_shipped_threshold(n_events, seed, block, geometry, beamline):
    """Config-4 events (one vertex, Poisson(120) tracks) through the NN-pruned search at the shipped
    scorer's threshold: with the logits the search computed (logged beside it), the oracle's pruning
    rule (groups of block / K unused h1 hits per `block` positions, keep-all fallback) gives every
    buffer of the sequence bit for bit, for workgroups of 256, 128 and 64 lanes."""
    import oracle
    from allen_amd import scorer, synth, velo
    from tests import parity
    params, thr = scorer.trained()
    ev = synth.generate(n_events=n_events, pileup=0, mean_tracks=120.0, seed=seed)
    plain = oracle.run(ev, geometry, beamline, compute_f64=0)
    n_seeds = int(plain["atomics_velo"][1::3].sum())
    seq, constants = _scored_run(geometry, beamline, ev, params, thr, log_cap=2 * n_seeds + 1000, block=block)
    assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
    logits = _logits_by_slot(seq, plain, ev.n_events)
    ref = oracle.run(ev, geometry, beamline, seed_logits=logits, seed_threshold=thr, seed_block=block)
    assert int(ref["status"][0]) == 0, f"oracle status {int(ref['status'][0]):#x}"
    pruned = int(ref["atomics_velo"][1::3].sum())
    print(f"[pruned] {n_events} events: {n_seeds} plain seeds, {pruned} kept at threshold {thr}")
    assert pruned < n_seeds
    errs = parity.compare(parity.gpu_buffers(seq, ref), ref, True)
    assert not errs, "\n".join(errs[:20])
    seq.close()
    constants.close()


def test_pruned_search_physics_on_20k_events(geometry, beamline):
    """Config 4's evaluation size: 20 000 one-vertex events, plain and NN-pruned at the shipped
    threshold. The long-track efficiency loss stays within 0.2 % absolute and the ghost rate does
    not rise."""
    from allen_amd import checker, scorer, synth, velo
    params, thr = scorer.trained()
    ev = synth.generate(n_events=20000, pileup=0, mean_tracks=120.0, seed=2020, keep_truth=True, n_threads=16)
    constants = velo.VeloConstants(geometry, beamline)
    dev = velo.DeviceEvents(ev)
    seq = velo.VeloSequence(constants, ev.n_events, ev.raw.size)

    def physics():
        off = seq.fetch("offsets_all_velo_tracks", np.uint32, ev.n_events + 1)
        hn = seq.fetch("offsets_velo_track_hit_number", np.uint32, int(off[-1]) + 1)
        return checker.check_buffers(ev, off, hn, seq.fetch("velo_track_hits", np.uint8, 4 * int(hn[-1]))).as_dict()

    seq.run(dev)
    assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
    plain = physics()
    seq.set_scorer(params, thr)
    seq.run(dev)
    assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
    pruned = physics()
    print(f"[pruned 20k] plain {plain}\n[pruned 20k] pruned {pruned}")
    assert plain["02_long_efficiency"] - pruned["02_long_efficiency"] <= 0.002
    assert pruned["ghost_rate"] <= plain["ghost_rate"]
    seq.close()
    constants.close()
