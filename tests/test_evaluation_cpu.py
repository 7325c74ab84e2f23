"""The evaluation harness's host parts (gvc.evaluation, R/volume_metrics.py:186-370)
and its oracle (oracle/eval_ref.py), without a GPU."""
import math

import numpy as np
import pytest

from gvc import evaluation as ev
from oracle import eval_ref


@pytest.mark.parametrize("n", [2, 5, 8])
def test_grid_is_the_reference_grid(n):
    mins, maxes = [-0.3, 0.1, 2.0], [0.7, 0.4, 5.0]
    ref = eval_ref.reference_grid(mins, maxes, n, np.random.RandomState(11))
    got, spacing = ev.evaluation_grid(mins, maxes, n, np.random.RandomState(11))
    assert np.array_equal(got, ref)
    assert np.allclose(spacing, (np.array(maxes) - np.array(mins)) / (n - 1))
    # without jitter the order is x fastest: (x_c, y_b, z_a) at (a*n + b)*n + c
    class Zero:
        def uniform(self, lo, hi, shape):
            return np.zeros(shape)
    plain, _ = ev.evaluation_grid(mins, maxes, n, Zero())
    a, b, c = n - 1, 0, 1 % n
    assert np.allclose(plain[(a * n + b) * n + c], [np.linspace(mins[0], maxes[0], n)[c],
                                                    np.linspace(mins[1], maxes[1], n)[b],
                                                    np.linspace(mins[2], maxes[2], n)[a]])


def test_metrics_oracle_closed_forms():
    rng = np.random.default_rng(0)
    g = rng.uniform(0, 1, 1000)
    d = 0.03
    m = eval_ref.metrics(g + d, g)
    assert m["L1"] == pytest.approx(d) and m["L2"] == pytest.approx(d * d)
    assert m["PSNR"] == pytest.approx(-10 * math.log10(d * d + 1e-8))
    assert m["PSNR of original"] == pytest.approx(-10 * math.log10(np.mean(g * g) + 1e-8))
    # masks with known counts: 100 false negatives, 50 false positives, 70 true invalid
    c, gt = g + d, g.copy()
    c[:100] = -1
    gt[100:150] = -1
    c[150:220] = -1
    gt[150:220] = -1
    m = eval_ref.metrics(c, gt)
    assert (m["false_negative"], m["false_positive"], m["invalid"]) == (100, 50, 120)
    assert m["false_negative_percent"] == pytest.approx(0.1)
    ok = np.arange(1000) >= 220
    assert m["PSNR without false positives/negatives"] == pytest.approx(-10 * math.log10(d * d + 1e-8))
    # with -1 read as 0: the false negatives contribute g^2
    mse4 = (np.sum(g[:100] ** 2) + np.sum((g[100:150] + d + 1) ** 2) + 0 + np.sum(np.full(ok.sum(), d * d)) +
            np.sum((0 + 1) ** 2 * np.ones(70))) / 1000
    assert m["PSNR with 0s for fn"] == pytest.approx(-10 * math.log10(mse4 + 1e-8))
    mesh = eval_ref.metrics(c, gt, grid=False)
    assert mesh["false_positive_percent"] == pytest.approx(50 / 120)


def test_vtk_writer(tmp_path):
    a = np.arange(24, dtype=np.float64).reshape(2, 3, 4) / 8
    path = tmp_path / "t.vtk"
    ev.write_vtk_structured(str(path), a, spacing=(0.5, 0.25, 1.0))
    lines = path.read_text().splitlines()
    assert lines[:10] == ["# vtk DataFile Version 3.0", "3D Scalar Data", "ASCII", "DATASET STRUCTURED_POINTS",
                          "DIMENSIONS 2 3 4", "ORIGIN 0.0 0.0 0.0", "SPACING 0.5 0.25 1.0", "POINT_DATA 24",
                          "SCALARS values double", "LOOKUP_TABLE default"]
    assert [float(v) for v in lines[10:]] == list(a.flatten())
