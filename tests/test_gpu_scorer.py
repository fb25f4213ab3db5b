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
