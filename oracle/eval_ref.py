"""CPU restatement of the volume evaluation's metrics -- TEST INFRASTRUCTURE ONLY
(imported by tests/, never by the product path).

metrics(cells, gt, grid) restates R/volume_metrics.py:331-369 on float64 numpy
arrays: L1 (l1_loss = mean |c - g|, R/utils/loss_utils.py:19-20), L2, the five
PSNR variants 20 log10(1) - 10 log10(mse + 1e-8) of the mean squared errors
over all samples / over g != -1 & c != -1 / of g itself / with c = -1 read as 0
/ that over g != -1, and the false-negative / false-positive counts and
percentages.  Pinned by closed forms (a constant offset gives L1 = |d|,
L2 = d^2, PSNR = -10 log10(d^2 + 1e-8)) and by masks built with known counts.

reference_grid(mins, maxes, n, rng) is the reference's grid construction itself
(np.meshgrid 'ij', np.rot90(k=1, axes=(2, 0)), np.flip(axis=2), jitter, clip),
R/volume_metrics.py:207-227, to pin the direct x-fastest construction.
"""
import numpy as np


def _psnr(mse):
    return 20 * np.log10(1.0) - 10 * np.log10(mse + 1e-8)


def metrics(cells, gt, grid=True):
    c = np.asarray(cells, dtype=np.float64).reshape(-1)
    g = np.asarray(gt, dtype=np.float64).reshape(-1)
    c2 = c.copy()
    c2[c2 == -1] = 0
    both = (g != -1) & (c != -1)
    out = {
        "L1": np.abs(c - g).mean(),
        "L2": np.mean((c - g) ** 2),
        "PSNR": _psnr(np.mean((c - g) ** 2)),
        "PSNR without false positives/negatives": _psnr(np.mean((c[both] - g[both]) ** 2)) if both.any() else np.nan,
        "PSNR of original": _psnr(np.mean(g ** 2)),
        "PSNR with 0s for fn": _psnr(np.mean((c2 - g) ** 2)),
        "PSNR with 0s for fn and no fp": _psnr(np.mean((c2[g != -1] - g[g != -1]) ** 2)) if (g != -1).any() else np.nan,
        "samples": c.size,
        "invalid": int(np.count_nonzero(g == -1)),
        "false_negative": int(np.count_nonzero((c == -1) & (g != -1))),
        "false_positive": int(np.count_nonzero((c != -1) & (g == -1))),
    }
    if grid:
        out["percent_invalid"] = out["invalid"] / c.size
        out["false_negative_percent"] = out["false_negative"] / c.size
        out["false_positive_percent"] = out["false_positive"] / c.size
    else:
        inv = out["invalid"]
        out["false_positive_percent"] = out["false_positive"] / inv if inv else np.nan
    return out


def reference_grid(mins, maxes, cell_count, rng):
    n = cell_count
    spacing = [(maxes[k] - mins[k]) / (n - 1) for k in range(3)]
    x = np.linspace(mins[0], maxes[0], n)
    y = np.linspace(mins[1], maxes[1], n)
    z = np.linspace(mins[2], maxes[2], n)
    x, y, z = np.meshgrid(x, y, z, indexing="ij")
    samples = np.vstack([x.ravel(), y.ravel(), z.ravel()]).T
    samples_3d = samples.reshape(n, n, n, 3)
    rot = np.rot90(samples_3d, k=1, axes=(2, 0))
    flat = np.flip(rot, axis=2).reshape(-1, 3)
    jitter = rng.uniform(-0.5, 0.5, flat.shape) * np.array(spacing)[None, :]
    return np.clip(flat + jitter, np.array(mins)[None, :], np.array(maxes)[None, :])
