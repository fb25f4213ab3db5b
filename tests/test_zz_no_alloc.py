"""velo_prefix_sum_static (the allocation-free prefix sum) against velo_prefix_sum, through the C ABI."""
import ctypes

import numpy as np
import pytest

from allen_amd import velo

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    import torch
    rng = np.random.default_rng(seed)
    host = rng.integers(0, 1 << 31, size=n, dtype=np.int64).astype(np.uint32)
    return torch.from_numpy(host.view(np.int32)).cuda()


def test_static_scratch_equals_workspace_scan_on_consecutive_calls():
    """Two calls in a row, the second smaller (fewer tile sums than the first left in the static
    array), each bit-identical to the workspace entry point on the same input, totals included;
    a third, one-pass size uses no tile sums."""
    import torch
    for n, seed in [(26 * 20000 + 3, 1), (40961, 2), (5000, 3)]:
        x = _inputs(n, seed)
        want = torch.full((n + 1,), -1, dtype=torch.int32, device="cuda")
        got = torch.full((n + 1,), -1, dtype=torch.int32, device="cuda")
        t_want = velo.prefix_sum(x, n, want, want_total=True)
        t_got = velo.prefix_sum_static(x, n, got, want_total=True)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy())
        assert t_got == t_want
        host = x.cpu().numpy().view(np.uint32).astype(np.uint64)
        assert t_got == int(host.sum() % (1 << 32))


def test_static_scratch_rejects_bad_arguments():
    """Sizes beyond the static array's capacity and a workspace pointer are refused before any launch."""
    import torch
    lib = velo.lib()
    cap = int(lib.velo_prefix_sum_static_max_items())
    assert cap == 1 << 28
    out = torch.empty(8, dtype=torch.int32, device="cuda")
    x = torch.zeros(8, dtype=torch.int32, device="cuda")
    p = velo.PrefixSumParams(x.data_ptr(), cap + 1, out.data_ptr(), 0, 0)
    assert lib.velo_prefix_sum_static(ctypes.byref(p), velo.stream_handle()) == 3  # VELO_ERR_CAPACITY
    p = velo.PrefixSumParams(x.data_ptr(), 8, out.data_ptr(), x.data_ptr(), 0)
    assert lib.velo_prefix_sum_static(ctypes.byref(p), velo.stream_handle()) == 1  # VELO_ERR_INVALID_ARGUMENT
