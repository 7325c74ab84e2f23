"""The drop-in ``bvh_diff_gaussian_rasterization`` surface against the oracle.

The autograd path a caller uses (GaussianRasterizer -> rasterize_gaussians ->
_RasterizeGaussians, DGR/bvh_diff_gaussian_rasterization/__init__.py:15-140)
must give all five gradients (means, scales, rotations, values, weights) of the
FP64 oracle within 1e-5 relative (normwise per tensor), with both traversal
variants (use_gaussian_bvh), and its debug mode must print the reference's
intersection statistics (CR/rasterizer_impl.cu:175-216) with counts equal to the
oracle's pair membership.
"""
import numpy as np
import pytest

from conftest import assert_grads, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

import oracle  # noqa: E402
from bvh_diff_gaussian_rasterization import GaussianRasterizationSettings, GaussianRasterizer  # noqa: E402
from gvc import workloads  # noqa: E402

f32 = np.float32
DEV = "cuda:0"


def settings(use_gaussian_bvh=False, debug=False):
    return GaussianRasterizationSettings(volume_mins=(0, 0, 0), volume_maxes=(1, 1, 1), cell_count=32, bg=-1.0,
                                         scale_modifier=1.0, use_gaussian_bvh=use_gaussian_bvh, debug=debug)


def config0_inputs(seed=0, stride=4):
    g = workloads.config0(seed)
    return g, np.exp(g["log_scales"]).astype(f32), g["samples"][::stride].copy()


@pytest.mark.parametrize("use_gaussian_bvh", [False, True])
def test_autograd_five_gradients(use_gaussian_bvh):
    g, scales, samples = config0_inputs(0)
    rng = np.random.default_rng(9)
    gV = rng.normal(size=samples.shape[0]).astype(f32)
    gW = rng.normal(size=samples.shape[0]).astype(f32)
    r = GaussianRasterizer(settings(use_gaussian_bvh))
    r.build_bvh(torch.tensor(samples, device=DEV), use_gaussian_bvh=use_gaussian_bvh)
    t = {k: torch.tensor(v, device=DEV, requires_grad=True)
         for k, v in dict(means=g["means"], scales=scales, rots=g["rotations"], values=g["values"],
                          weights=g["weights"]).items()}
    cells, cw = r(t["means"], t["scales"], t["rots"], t["values"], t["weights"])
    loss = (cells * torch.tensor(gV, device=DEV)).sum() + (cw * torch.tensor(gW, device=DEV)).sum()
    loss.backward()
    conic, box, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    cd = oracle.conic_fp64(scales, g["rotations"])
    Vd, Wd, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic, box, samples, cd)
    valid = Wd > 0
    V, W = cells.detach().cpu().numpy(), cw.detach().cpu().numpy()
    assert np.array_equal(W > 0, valid)
    assert (np.abs(V[valid] - Vd[valid]) / np.maximum(np.abs(Vd[valid]), 1e-3)).max() <= 1e-5
    ref = oracle.backward(g["means"], scales, g["rotations"], g["values"], g["weights"], conic, box, samples, Vd,
                          Wd, gV.astype(np.float64), gW.astype(np.float64))
    grads = [t["means"].grad, t["scales"].grad, t["rots"].grad, t["values"].grad, t["weights"].grad]
    for gr, x in zip(grads, (t["means"], t["scales"], t["rots"], t["values"], t["weights"])):
        assert gr is not None and gr.shape == x.shape
    assert_grads(grads, ref)


@pytest.mark.parametrize("use_gaussian_bvh", [False, True])
def test_debug_intersection_counts(use_gaussian_bvh, capsys):
    """debug=True prints the reference's line: per sample (use_gaussian_bvh) or
    per Gaussian, the largest count, its first index and the mean; the counts
    are the oracle's accepted pairs."""
    from bvh_diff_gaussian_rasterization import _C
    g, scales, samples = config0_inputs(3, stride=8)
    r = GaussianRasterizer(settings(use_gaussian_bvh, debug=True))
    r.build_bvh(torch.tensor(samples, device=DEV), use_gaussian_bvh=use_gaussian_bvh)
    args = [torch.tensor(v, device=DEV) for v in (g["means"], scales, g["rotations"], g["values"], g["weights"])]
    _C.rasterize_gaussians(*args, 1.0, 0, 0, 0, 1, 1, 1, -1.0, use_gaussian_bvh, False)  # no print
    capsys.readouterr()
    cells, cw, conics, state = _C.rasterize_gaussians(*args, 1.0, 0, 0, 0, 1, 1, 1, -1.0, use_gaussian_bvh, True,
                                                      return_state=True)
    out = capsys.readouterr().out
    per_sample, per_gaussian = [c.cpu().numpy() for c in state.pair_counts(r.sample_set)]
    conic, box, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    mask, _ = oracle.pair_mask(g["means"], conic, box, samples)
    assert np.array_equal(per_sample, mask.sum(axis=0))
    assert np.array_equal(per_gaussian, mask.sum(axis=1))
    counts, what = (mask.sum(axis=0), "sample") if use_gaussian_bvh else (mask.sum(axis=1), "gaussian")
    line = (f"Intersections: max={counts.max()} ({what} {int(np.argmax(counts))}), "
            f"avg={counts.sum() / counts.size:.3f} over {counts.size} {what}s")
    assert line in out, out
