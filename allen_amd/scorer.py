"""The learned seed scorer: host side of the MLP kernels (training and scoring on the GPU).

The network is 10 -> 32 -> 32 -> 1 with ReLU. Its matrix products run on BF16 MFMA with FP32 accumulation; the parameters are kept as FP32 masters (`allen_amd/csrc/velo_scorer.hip`). The 10 inputs of a seed are those of `velo_seed_features`.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import velo

# The parameters trained by main() with its defaults, and their evaluation (threshold included).
TRAINED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "seed_scorer.npz")

N_PARAMS = 1441
N_FEATURES = 10
LAYOUT = {"w1": (0, (10, 32)), "b1": (320, (32,)), "w2": (352, (32, 32)), "b2": (1376, (32,)),
          "w3": (1408, (32,)), "b3": (1440, ())}


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation of the FP32 parameter vector."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    p[0:320] = rng.normal(0, np.sqrt(2 / 10), 320)
    p[352:1376] = rng.normal(0, np.sqrt(2 / 32), 1024)
    p[1408:1440] = rng.normal(0, np.sqrt(1 / 32), 32)
    return p


class MLP:
    """Parameters and Adam state on the device, with training and scoring calls."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0):
        self.torch = velo._torch()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        assert p.size == N_PARAMS
        t = self.torch
        self.params = t.from_numpy(p.copy()).cuda()
        self.m = t.zeros(N_PARAMS, dtype=t.float32, device="cuda")
        self.v = t.zeros(N_PARAMS, dtype=t.float32, device="cuda")
        self.step_count = 0
        self._scratch = None

    def _scratch_for(self, n: int):
        need = int(velo.lib().velo_mlp_scratch_floats(n))
        if self._scratch is None or self._scratch.numel() < need:
            self._scratch = self.torch.empty(need, dtype=self.torch.float32, device="cuda")
        return self._scratch

    def train_step(self, features, labels, lr: float, want_grad: bool = False):
        """One Adam step on the full batch (device tensors: features float32 [n, 10], labels uint8
        [n]); returns (mean loss, gradient or None) for the parameters before the step."""
        t = self.torch
        n = int(labels.numel())
        grad = t.empty(N_PARAMS, dtype=t.float32, device="cuda") if want_grad else None
        loss = t.empty(1, dtype=t.float32, device="cuda")
        if lr:
            self.step_count += 1
        velo.check(velo.lib().velo_mlp_train_step(
            velo._ptr(features), velo._ptr(labels), n, velo._ptr(self.params), velo._ptr(self.m), velo._ptr(self.v),
            velo._ptr(self._scratch_for(n)), float(lr), self.step_count if lr else 0, velo._ptr(grad),
            velo._ptr(loss), velo.stream_handle()), "velo_mlp_train_step")
        return float(loss.item()) / n, (grad.cpu().numpy() if want_grad else None)

    def score(self, features):
        """Logits of the rows of a device tensor float32 [n, 10]."""
        t = self.torch
        n = int(features.shape[0])
        out = t.empty(n, dtype=t.float32, device="cuda")
        velo.check(velo.lib().velo_mlp_score(velo._ptr(features), n, velo._ptr(self.params), velo._ptr(out),
                                             velo.stream_handle()), "velo_mlp_score")
        return out

    def host_params(self) -> np.ndarray:
        return self.params.cpu().numpy().copy()


# ---------------------------------------------------------------------------------------------
# Training data: the seeds the plain search writes, labelled by whether they end in a track
# that the checker matches to a generated particle (>= 70 % of its hits).

def seed_labels(events, seq, records: np.ndarray) -> np.ndarray:
    """1 for each seed (rows of `records`: event, module pair, h0, h1, h2) whose three hits open a
    consolidated track of the last run that the checker matches to a particle, else 0."""
    from . import checker
    n = events.n_events
    est = seq.fetch("offsets_estimated_input_size", np.uint32, 26 * n + 1).astype(np.int64)
    ids = seq.fetch("sorted_velo_cluster_container", np.uint8, 4 * int(est[-1])).view("<u4")
    seed_ids = ids[est[0:26 * n:26][records[:, 0]][:, None] + records[:, 2:5].astype(np.int64)]
    off_all = seq.fetch("offsets_all_velo_tracks", np.uint32, n + 1)
    n_tracks = int(off_all[-1])
    off_h = seq.fetch("offsets_velo_track_hit_number", np.uint32, n_tracks + 1).astype(np.int64)
    th = seq.fetch("velo_track_hits", np.uint8, 4 * int(off_h[-1])).view("<u4")
    res = checker.check(events, off_all, off_h, th)
    track_event = np.repeat(np.arange(n), np.diff(off_all.astype(np.int64)))
    first3 = th[off_h[:-1, None] + np.arange(3)[None, :]]
    good = {(int(e), int(a), int(b), int(c))
            for e, (a, b, c), m in zip(track_event, first3, res.track_particle >= 0) if m}
    return np.fromiter(((int(e), int(a), int(b), int(c)) in good
                        for e, (a, b, c) in zip(records[:, 0], seed_ids)), dtype=np.uint8, count=records.shape[0])


def collect(constants, batches, log=None):
    """Runs the plain GPU search over each batch of events (an iterable of synth.Events with
    truth) with the seed log on; returns features (float32 [n, 10]) and labels (uint8 [n])."""
    feats, labs = [], []
    for ev in batches:
        seq = velo.VeloSequence(constants, ev.n_events, ev.raw.size)
        seq.set_tuning("seed_log", 400 * ev.n_events)
        seq.run(velo.DeviceEvents(ev))
        rec, f, seen = seq.seed_log()
        if seen > rec.shape[0]:
            raise velo.VeloError(f"seed log overflow: {seen} seeds, capacity {rec.shape[0]}")
        lab = seed_labels(ev, seq, rec)
        seq.close()
        feats.append(f)
        labs.append(lab)
        if log:
            log(f"[scorer] {ev.n_events} events: {rec.shape[0]} seeds, {lab.mean():.3f} positive")
    return np.concatenate(feats), np.concatenate(labs)


def train(features: np.ndarray, labels: np.ndarray, epochs: int = 20, batch: int = 65536, lr: float = 3e-3,
          seed: int = 0, log=None) -> MLP:
    """Adam on shuffled minibatches on the GPU."""
    torch = velo._torch()
    xd = torch.from_numpy(np.ascontiguousarray(features, np.float32)).cuda()
    yd = torch.from_numpy(np.ascontiguousarray(labels, np.uint8)).cuda()
    net = MLP(seed=seed)
    rng = np.random.default_rng(seed)
    n = labels.size
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).cuda()
        xs, ys = xd[perm].contiguous(), yd[perm].contiguous()
        tot = 0.0
        for b0 in range(0, n, batch):
            loss, _ = net.train_step(xs[b0:b0 + batch], ys[b0:b0 + batch], lr)
            tot += loss * min(batch, n - b0)
        if log:
            log(f"[scorer] epoch {ep}: mean loss {tot / n:.4f}")
    return net


def roc(scores: np.ndarray, labels: np.ndarray) -> dict:
    """ROC area, and for the pruning thresholds the fraction of seeds removed and of positive seeds
    lost (a seed is kept when its logit >= threshold)."""
    order = np.argsort(-scores, kind="stable")
    y = labels[order].astype(np.float64)
    pos, neg = y.sum(), (1 - y).sum()
    tpr = np.concatenate([[0], np.cumsum(y) / max(pos, 1)])
    fpr = np.concatenate([[0], np.cumsum(1 - y) / max(neg, 1)])
    auc = float(np.trapezoid(tpr, fpr)) if neg and pos else float("nan")
    table = []
    for lost in (0.0005, 0.001, 0.002, 0.005, 0.01):
        k = int(np.ceil((1 - lost) * pos))
        kept_pos = np.cumsum(y)
        idx = int(np.searchsorted(kept_pos, k)) if k > 0 else 0
        thr = float(scores[order][min(idx, y.size - 1)])
        keep = scores >= thr
        table.append({"positive_lost": round(float(1 - (labels[keep] == 1).sum() / max(pos, 1)), 5),
                      "seeds_removed": round(float(1 - keep.mean()), 5), "threshold": thr})
    return {"auc": auc, "positives": int(pos), "negatives": int(neg), "thresholds": table}


def save(path: str, net: MLP, meta: dict) -> None:
    np.savez(path, params=net.host_params(), **{k: np.asarray(v) for k, v in meta.items()})


def trained() -> tuple[np.ndarray, float]:
    """The shipped parameters and their operating threshold (a seed is kept at logit >= it)."""
    with open(os.path.splitext(TRAINED)[0] + ".json") as f:
        meta = json.load(f)
    return load_params(TRAINED), float(meta["threshold"])


def load_params(path: str) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def main(argv=None):
    """Trains the scorer on seeds of the plain GPU search and prints its evaluation as JSON."""
    import argparse
    import json
    import sys
    from . import synth
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--train-events", type=int, default=16000)
    ap.add_argument("--test-events", type=int, default=2000)
    ap.add_argument("--chunk", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    constants = velo.VeloConstants(synth.read_geometry(), synth.read_beamline())
    # Config-1-style events (one PV, Poisson(120) tracks, 3 % noise), disjoint seeds for testing.
    batches = lambda n, s0: (synth.generate(n_events=min(args.chunk, n - c), pileup=0, mean_tracks=120.0,
                                            seed=s0 + i) for i, c in enumerate(range(0, n, args.chunk)))
    x, y = collect(constants, batches(args.train_events, 1000 * args.seed), log)
    xt, yt = collect(constants, batches(args.test_events, 1000 * args.seed + 500), log)
    log(f"[scorer] training on {y.size} seeds ({y.mean():.4f} positive), testing on {yt.size}")
    net = train(x, y, epochs=args.epochs, seed=args.seed, log=log)
    torch = velo._torch()
    st = net.score(torch.from_numpy(xt).cuda()).cpu().numpy()
    out = {"train_seeds": int(y.size), "test_seeds": int(yt.size), "test": roc(st, yt)}
    if args.out:
        save(args.out, net, {"train_seeds": y.size})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
