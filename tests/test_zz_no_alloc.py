"""The allocation-free entry point gvc_volume_metrics_noalloc, through the C ABI.

Its block partials live in the library's static device array, not in a
caller's workspace.  Bars: bit-identical to gvc_volume_metrics (same kernels,
same summation order) on two calls in a row, the second with fewer samples (the
array keeps the first call's partials: the second must not read them); sizes
beyond its 2^24-sample limit are rejected with GVC_ERR_CAPACITY before any work.
"""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

from gvc import _lib  # noqa: E402

DEV = "cuda:0"


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def with_workspace(cells, gt):
    L = _lib.lib()
    S = cells.numel()
    n = L.gvc_volume_metrics_bytes(S)
    ws = torch.empty(n, dtype=torch.uint8, device=DEV)
    out = torch.empty(_lib.GVC_METRIC_SUMS, dtype=torch.float64, device=DEV)
    _lib.check(L.gvc_volume_metrics(S, _p(cells), _p(gt), _p(out), _p(ws), n, _stream()), "gvc_volume_metrics")
    return out


def no_alloc(cells, gt):
    L = _lib.lib()
    out = torch.full((_lib.GVC_METRIC_SUMS,), float("nan"), dtype=torch.float64, device=DEV)
    _lib.check(L.gvc_volume_metrics_noalloc(cells.numel(), _p(cells), _p(gt), _p(out), _stream()),
               "gvc_volume_metrics_noalloc")
    return out


def inputs(S, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, S).astype(np.float32)
    g = rng.uniform(0, 1, S).astype(np.float32)
    c[rng.random(S) < 0.1] = -1.0
    g[rng.random(S) < 0.1] = -1.0
    return torch.tensor(c, device=DEV), torch.tensor(g, device=DEV)


def test_no_alloc_equals_workspace_call_twice():
    for S, seed in ((3_000_001, 1), (70_000, 2), (1, 3)):
        cells, gt = inputs(S, seed)
        a = no_alloc(cells, gt)
        b = with_workspace(cells, gt)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (S, a, b)
        assert int(a[7].item()) + int(a[8].item()) == S


def test_no_alloc_limit():
    L = _lib.lib()
    cells, gt = inputs(16, 4)
    out = torch.zeros(_lib.GVC_METRIC_SUMS, dtype=torch.float64, device=DEV)
    st = L.gvc_volume_metrics_noalloc((1 << 24) + 1, _p(cells), _p(gt), _p(out), _stream())
    assert st == _lib.GVC_ERR_CAPACITY
    assert b"2^24" in L.gvc_last_error()
    assert torch.all(out == 0)
