"""The plain GPU sequence (benchmark baseline) computes the same mixture.

It is only a baseline, so its bar is looser than the product's: the same pair
membership (same predicate, exact pair count), the same sentinel decisions away
from the cutoff, the field within 1e-4 relative and gradients within 1e-3 of
the largest magnitude, all against the FP64 oracle (float atomics and no
compensation make it non-deterministic in the last bits).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

import oracle  # noqa: E402
from gvc import workloads  # noqa: E402
from plain_baseline import plain  # noqa: E402

f32 = np.float32
DEV = "cuda:0"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, torch.float32)


@pytest.mark.parametrize("n_samples", [32768, 1000, 0])
def test_plain_sequence_matches_oracle(n_samples):
    g = workloads.config0(0)
    scales = np.exp(g["log_scales"]).astype(f32)
    samples = g["samples"][:n_samples].copy()
    S = samples.shape[0]
    rng = np.random.default_rng(5)
    gV = rng.normal(size=S).astype(f32)
    gW = rng.normal(size=S).astype(f32)
    tree = plain.PlainSampleTree(T(samples))
    pairs = torch.zeros(1, dtype=torch.int64, device=DEV)
    args = (T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]), T(g["weights"]))
    cells, cw, conics, boxes = plain.plain_forward(*args, 1.0, tree, pairs)
    grads = plain.plain_backward(*args, 1.0, conics, boxes, tree, cells, cw, T(gV), T(gW))
    torch.cuda.synchronize()
    conic_ref, box_ref, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    assert np.array_equal(conics.cpu().numpy(), conic_ref)
    cd = oracle.conic_fp64(scales, g["rotations"])
    Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, samples, cd)
    assert int(pairs.item()) == int(cnt.sum())
    if S == 0:
        return
    V, W = cells.cpu().numpy(), cw.cpu().numpy()
    far = np.abs(Wr - 1e-2) > 1e-4
    assert np.array_equal((W > 0)[far], (Wr > 0)[far])
    ok = (W > 0) & (Wr > 0)
    assert (np.abs(V[ok] - Vr[ok]) / np.maximum(np.abs(Vr[ok]), 1e-3)).max() <= 1e-4
    assert (np.abs(W[ok] - Wr[ok]) / np.maximum(np.abs(Wr[ok]), 1e-3)).max() <= 1e-4
    ref = oracle.backward(g["means"], scales, g["rotations"], g["values"], g["weights"], conic_ref, box_ref, samples,
                          Vr, Wr, gV.astype(np.float64), gW.astype(np.float64))
    for name, t in zip(["means", "scales", "rotations", "values", "weights"], grads):
        a = t.cpu().numpy().reshape(-1).astype(np.float64)
        b = ref[name].reshape(-1)
        assert np.abs(a - b).max() / max(np.abs(b).max(), 1e-30) <= 1e-3, name
