"""The evaluation harness on the GPU (gvc.evaluation over gvc_volume_metrics)
against the oracle's restatement of R/volume_metrics.py:331-369.

Bars: counts exact; FP64 sums within 1e-12 relative of numpy's float64 sums;
metrics within 1e-9 relative; the end-to-end evaluation of a model the reference
trained equal to the oracle's metrics on the same rendered cells."""
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

from oracle import eval_ref  # noqa: E402
from gvc import evaluation as ev  # noqa: E402
from gvc import model_io as mio  # noqa: E402
from gvc import training  # noqa: E402

DEV = "cuda:0"
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mito1024wveg_iter1000.ply")


def same_metrics(a, b, rel=1e-9):
    assert set(a) == set(b)
    for k in a:
        x, y = float(a[k]), float(b[k])
        if np.isnan(y):
            assert np.isnan(x), k
        else:
            assert x == pytest.approx(y, rel=rel, abs=1e-300), k


@pytest.mark.parametrize("S", [1, 1000, (1 << 21) + 7])
def test_metric_sums_and_metrics(S):
    rng = np.random.default_rng(S)
    g = rng.uniform(0, 1, S).astype(np.float32)
    c = (g + rng.normal(0, 0.05, S)).astype(np.float32)
    g[rng.random(S) < 0.2] = -1.0
    c[rng.random(S) < 0.1] = -1.0
    tc, tg = torch.as_tensor(c, device=DEV), torch.as_tensor(g, device=DEV)
    s = ev.metric_sums(tc, tg).cpu().numpy()
    c64, g64 = c.astype(np.float64), g.astype(np.float64)
    c0 = np.where(c == -1, 0.0, c64)
    both = (g != -1) & (c != -1)
    want = [np.abs(c64 - g64).sum(), ((c64 - g64) ** 2).sum(), ((c64 - g64)[both] ** 2).sum(), both.sum(),
            (g64 ** 2).sum(), ((c0 - g64) ** 2).sum(), ((c0 - g64)[g != -1] ** 2).sum(), (g != -1).sum(),
            (g == -1).sum(), ((c == -1) & (g != -1)).sum(), ((c != -1) & (g == -1)).sum()]
    for k in (3, 7, 8, 9, 10):
        assert s[k] == want[k], k
    for k in (0, 1, 2, 4, 5, 6):
        assert s[k] == pytest.approx(want[k], rel=1e-12, abs=1e-300), k
    for grid in (True, False):
        same_metrics(ev.volume_metrics(tc, tg, grid=grid), eval_ref.metrics(c, g, grid=grid))


def test_evaluate_reference_model_on_grid():
    p = mio.read_ply(GOLDEN)
    means, vals, wts, log_s, rots = mio.to_torch(p, DEV)
    with torch.no_grad():
        scales, rq, w, v = training.activate(log_s, rots, wts, vals)
    mins, maxes = p.means.min(0), p.means.max(0)
    pts, spacing = ev.evaluation_grid(mins, maxes, 48, np.random.RandomState(0))
    # ground truth: a smooth field inside an ellipsoid of the box, -1 outside
    ctr, half = (mins + maxes) / 2, (maxes - mins) / 2
    r = np.linalg.norm((pts - ctr) / half, axis=1)
    gt = np.where(r < 0.9, 0.5 + 0.4 * np.cos(3 * r), -1.0).astype(np.float32)
    samples = torch.as_tensor(pts, dtype=torch.float32, device=DEV)
    m, cells = ev.evaluate(means, scales, rq, v.reshape(-1, 1), w.reshape(-1, 1), samples,
                           torch.as_tensor(gt, device=DEV))
    assert m["samples"] == 48 ** 3 and m["invalid"] == int((gt == -1).sum())
    assert 0 < m["false_negative"] + m["false_positive"] < m["samples"]
    same_metrics(m, eval_ref.metrics(cells.cpu().numpy(), gt))
    # the grid lies in the model's box and keeps the reference's order
    assert np.all(pts >= mins - 1e-12) and np.all(pts <= maxes + 1e-12)


def test_evaluate_mesh_mode():
    """The mesh branch (R/volume_metrics.py:291-312): points drawn inside the mesh
    (sample_mesh_points), ground truth located and interpolated (-1 where no
    tetrahedron holds a point), metrics with the mesh-mode false-positive share."""
    from gvc import workloads
    from gvc.sampling import TetMeshSampler, sample_mesh_points
    rng = np.random.default_rng(3)
    n = 10
    pts = workloads.jittered_lattice(rng, n, 0.3)
    tets = workloads.kuhn_tets(n)
    vals = (0.5 + 0.4 * np.sin(3 * pts[:, 0]) * pts[:, 1]).astype(np.float64)
    sampler = TetMeshSampler(pts, tets, vals, device=DEV)
    g = torch.Generator(device=DEV).manual_seed(4)
    samples = sample_mesh_points(torch.as_tensor(pts, device=DEV), torch.as_tensor(tets, device=DEV), 1, 20000,
                                 generator=g).reshape(-1, 3)
    # half of them pushed out of the unit cube: the probe marks those -1
    samples[::2] += 2.0
    gt = sampler.sample(samples)[0].float()
    assert int((gt == -1).sum()) == 10000
    P = 3000
    means = torch.as_tensor(rng.uniform(0, 1, (P, 3)), dtype=torch.float32, device=DEV)
    scales = torch.full((P, 3), 0.05, device=DEV)
    rots = torch.zeros((P, 4), device=DEV)
    rots[:, 0] = 1
    values = torch.as_tensor(rng.uniform(0.1, 0.9, (P, 1)), dtype=torch.float32, device=DEV)
    weights = torch.full((P, 1), 0.3, device=DEV)
    m, cells = ev.evaluate(means, scales, rots, values, weights, samples, gt, grid=False)
    assert m["invalid"] == 10000 and "percent_invalid" not in m
    same_metrics(m, eval_ref.metrics(cells.cpu().numpy(), gt.cpu().numpy(), grid=False))
