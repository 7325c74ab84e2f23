"""GPU parity of the fused training-step glue against the reference formulas in
float64 torch (CPU autograd): activations (R/scene/gaussian_model.py:31-48,
:71-77, :114-132) and the fitting loss (R/train.py:389-420).

Bars: forward values within 1e-6 relative, gradients within 1e-5 relative to
the largest magnitude of each tensor (fp32 kernels against float64).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

from gvc import training  # noqa: E402

DEV = "cuda:0"


def ref_activate(ls, q, wl, vl, max_scale):
    s = torch.exp(ls)
    if max_scale > 0:
        r = torch.linalg.norm(s, dim=1, keepdim=True) + 1e-8          # R/scene/gaussian_model.py:71-77
        r_soft = max_scale * torch.tanh(r / max_scale)
        s = s * (r_soft / r)
    return s, torch.nn.functional.normalize(q), torch.sigmoid(wl), torch.sigmoid(vl)


def rel(a, b):
    a = a.detach().cpu().double()
    b = b.detach().cpu().double()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


@pytest.mark.parametrize("max_scale", [0.02, 0.0])
def test_activations_match_reference(max_scale):
    rng = np.random.default_rng(0)
    P = 5000
    ls = rng.normal(np.log(0.01), 1.0, (P, 3))
    ls[:10] = np.log(0.5)                                              # deep in the cap
    q = rng.normal(size=(P, 4))
    q[10] = [1e-3, 0, 0, 0]
    wl = rng.normal(size=(P, 1))
    vl = rng.normal(size=(P, 1))
    raw64 = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in (ls, q, wl, vl)]
    ref = ref_activate(*raw64, max_scale)
    ups = [torch.tensor(rng.normal(size=t.shape), dtype=torch.float64) for t in ref]
    sum(((r * u).sum() for r, u in zip(ref, ups))).backward()
    raw32 = [torch.tensor(a, dtype=torch.float32, device=DEV, requires_grad=True) for a in (ls, q, wl, vl)]
    out = training.activate(*raw32, max_scale=max_scale)
    for o, r in zip(out, ref):
        assert rel(o, r) <= 1e-6
    sum(((o * u.to(DEV, torch.float32)).sum() for o, u in zip(out, ups))).backward()
    for a, b in zip(raw32, raw64):
        assert rel(a.grad, b.grad) <= 1e-5


def ref_loss(cells, weights, gt, fn_reg, fp_reg, encode_surface):
    """R/train.py:389-420 with float64 values; the masks are evaluated on the
    float32 tensors, as the reference's fp32 training tensors compare."""
    c32, w32, g32 = cells.detach().float(), weights.detach().float(), gt.float()
    recon = torch.logical_and(g32 != -1, c32 != -1)
    if encode_surface:
        l1 = torch.abs(cells - gt).mean()
    else:
        l1 = torch.abs(cells[recon] - gt[recon]).mean()
    t = 0.0105 if encode_surface else 0.011
    fn_mask = torch.logical_and(torch.logical_and(g32 != -1, w32 > 0.0), w32 < t)
    fn_vals = torch.clamp(0.011 - weights[fn_mask], min=0)
    loss = l1 + fn_reg * fn_vals.sum() / ((fn_vals > 0).sum().double() + 1e-8)
    if encode_surface:
        fp_mask = torch.logical_and(g32 == -1, w32 > 0.0)
        fp_vals = weights[fp_mask]
        loss = loss + fp_reg * fp_vals.sum() / ((fp_vals > 0).sum().double() + 1e-8)
    return loss


@pytest.mark.parametrize("surface", [False, True])
def test_loss_matches_reference(surface):
    rng = np.random.default_rng(1)
    S = 300_001
    cells = rng.uniform(0, 1, S).astype(np.float32)
    gt = rng.uniform(0, 1, S).astype(np.float32)
    w = rng.uniform(0, 0.05, S).astype(np.float32)
    cells[::7] = -1.0
    gt[::11] = -1.0
    w[::13] = 0.0
    w[5::17] = 0.0105
    c64 = torch.tensor(cells, dtype=torch.float64, requires_grad=True)
    w64 = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    ref = ref_loss(c64, w64, torch.tensor(gt, dtype=torch.float64), 1.0, 0.5, surface)
    (ref * 3.0).backward()
    c32 = torch.tensor(cells, device=DEV, requires_grad=True)
    w32 = torch.tensor(w, device=DEV, requires_grad=True)
    out = training.fit_loss(c32, w32, torch.tensor(gt, device=DEV), 1.0, 0.5, surface)
    assert abs(float(out) - float(ref)) <= 1e-6 * abs(float(ref))
    (out * 3.0).backward()
    assert rel(c32.grad, c64.grad) <= 1e-5
    assert rel(w32.grad, w64.grad) <= 1e-5
    # deterministic
    out2 = training.fit_loss(c32.detach(), w32.detach(), torch.tensor(gt, device=DEV), 1.0, 0.5, surface)
    assert float(out2) == float(out)
