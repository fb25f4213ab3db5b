"""The learned seed scorer: host side of the MLP kernels (training and scoring on the GPU).

The network is 10 -> 32 -> 32 -> 1 with ReLU. Its matrix products run on BF16 MFMA with FP32 accumulation; the parameters are kept as FP32 masters (`allen_amd/csrc/velo_scorer.hip`). The 10 inputs of a seed are those of `velo_seed_features`.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import velo

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
