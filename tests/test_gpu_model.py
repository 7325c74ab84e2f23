"""The mixture on a model the reference trained (tests/golden/mito1024wveg_iter1000.ply,
5,258 Gaussians with rotated, anisotropic covariances and a saturated value
logit), its raw parameters activated as the reference does
(R/scene/gaussian_model.py:31-48, :114-132), evaluated on points drawn in the
model's bounding box, through the C ABI, against the oracle.

Bars: conics bit-exact, pair membership exact, sentinel decisions equal, field
within 1e-5 relative of the FP64 oracle; gradients within 1e-5 of the FP64
oracle relative to the largest magnitude of each parameter tensor.
"""
import os

import numpy as np
import pytest

from conftest import assert_grads, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

import oracle  # noqa: E402
from gvc import mixture as mx  # noqa: E402
from gvc import model_io as mio  # noqa: E402
from gvc import training  # noqa: E402

DEV = "cuda:0"
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mito1024wveg_iter1000.ply")


def activated_model():
    p = mio.read_ply(GOLDEN)
    means, vals, wts, log_s, rots = mio.to_torch(p, DEV)
    with torch.no_grad():
        scales, rq, w, v = training.activate(log_s, rots, wts, vals)
    cpu = lambda t: t.detach().cpu().numpy().astype(np.float32)
    return {"means": p.means, "scales": cpu(scales), "rotations": cpu(rq), "values": cpu(v).reshape(-1, 1),
            "weights": cpu(w).reshape(-1, 1)}


def box_points(g, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = g["means"].min(0), g["means"].max(0)
    return rng.uniform(lo, hi, (n, 3)).astype(np.float32)


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, torch.float32)


def test_realistic_model_forward():
    g = activated_model()
    assert np.all(np.isfinite(g["values"])) and np.all(g["values"] <= 1.0)
    samples = box_points(g, 60000, 0)
    ss = mx.SampleSet(T(samples))
    cells, cw, conics, st = mx.mixture_forward(T(g["means"]), T(g["scales"]), T(g["rotations"]), T(g["values"]),
                                               T(g["weights"]), 1.0, ss)
    nrec, npairs, err = st.stats()
    assert err == 0
    conic_ref, box_ref, _ = oracle.preprocess(g["means"], g["scales"], g["rotations"], g["weights"])
    assert np.array_equal(conics.cpu().numpy().reshape(-1, 6), conic_ref)
    Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, samples)
    assert npairs == int(cnt.sum()) > 0
    V, W = cells.cpu().numpy(), cw.cpu().numpy()
    assert np.array_equal(W == 0, Wr == 0)
    assert np.all(V[W == 0] == -1.0)
    valid = Wr > 0
    assert valid.sum() > 1000
    ev = np.abs(V[valid] - Vr[valid]) / np.maximum(np.abs(Vr[valid]), 1e-3)
    ew = np.abs(W[valid] - Wr[valid]) / np.maximum(np.abs(Wr[valid]), 1e-3)
    assert max(ev.max(), ew.max()) <= 1e-5


def test_realistic_model_gradients():
    g = activated_model()
    samples = box_points(g, 12000, 1)
    rng = np.random.default_rng(2)
    gV = rng.normal(size=samples.shape[0]).astype(np.float32)
    gW = rng.normal(size=samples.shape[0]).astype(np.float32)
    ss = mx.SampleSet(T(samples))
    args = [T(g[k]) for k in ("means", "scales", "rotations", "values", "weights")]
    cells, cw, _, st = mx.mixture_forward(*args, 1.0, ss)
    grads = mx.mixture_backward(st, *args, 1.0, ss, cells, cw, T(gV), T(gW))
    conic_ref, box_ref, _ = oracle.preprocess(g["means"], g["scales"], g["rotations"], g["weights"])
    cd = oracle.conic_fp64(g["scales"], g["rotations"])
    Vr, Wr, _ = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, samples, cd)
    ref = oracle.backward(g["means"], g["scales"], g["rotations"], g["values"], g["weights"], conic_ref, box_ref,
                          samples, Vr, Wr, gV.astype(np.float64), gW.astype(np.float64))
    for name in ("means", "scales", "rotations", "values", "weights"):
        assert np.abs(ref[name]).max() > 0
    assert_grads(grads, ref)
