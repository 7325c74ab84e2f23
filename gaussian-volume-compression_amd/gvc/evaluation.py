"""Volume evaluation of a fitted model (R/volume_metrics.py:186-370).

The reference evaluates a model once: it samples either a jittered 128^3 grid of
the model's bounding box (ground truth from the mesh sampler) or 1,000,000 points
inside the mesh (ground truth from a probe, -1 where the probe is invalid),
renders the mixture there, and prints L1, L2, five PSNR variants and the false
negative / false positive counts.  Here the points and the ground truth come from
the libgvc samplers, the mixture from the libgvc forward, and every metric from
one FP64 pass (``gvc_volume_metrics``).

``evaluation_grid`` reproduces the reference's grid: np.linspace per axis, the
``rot90(k=1, axes=(2, 0))`` + ``flip(axis=2)`` reordering -- which is the
x-fastest order, flat index ``(a*n + b)*n + c`` at ``(x_c, y_b, z_a)`` -- then
``np.random.uniform(-0.5, 0.5)`` jitter in units of the spacing, clipped to the
box.  With the same NumPy random state it draws the reference's points.
"""
import math

import numpy as np
import torch

from ._lib import call, lib
from .mixture import _as_f32, _ptr, _stream

METRIC_SUMS = 11


def evaluation_grid(mins, maxes, cell_count=128, rng=None):
    """[cell_count^3, 3] float64 points (R/volume_metrics.py:207-227) and the grid
    spacing.  ``rng``: a numpy RandomState / Generator-like object with
    ``uniform`` (default: the global np.random, as the reference)."""
    rng = np.random if rng is None else rng
    n = int(cell_count)
    mins = np.asarray(mins, dtype=np.float64)
    maxes = np.asarray(maxes, dtype=np.float64)
    spacing = (maxes - mins) / (n - 1)
    axes = [np.linspace(mins[k], maxes[k], n) for k in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    jitter = rng.uniform(-0.5, 0.5, pts.shape) * spacing[None, :]
    return np.clip(pts + jitter, mins[None, :], maxes[None, :]), spacing


def metric_sums(cells, gt):
    """The 11 FP64 sums of ``gvc_volume_metrics`` as a float64 device tensor."""
    c = _as_f32(cells.detach().reshape(-1).contiguous(), "cells")
    g = gt.reshape(-1).to(device=c.device, dtype=torch.float32).contiguous()
    S = int(c.numel())
    if g.numel() != S:
        raise ValueError("cells and gt must have one entry per sample")
    out = torch.empty(METRIC_SUMS, dtype=torch.float64, device=c.device)
    nb = lib().gvc_volume_metrics_bytes(S)
    ws = torch.empty(nb, dtype=torch.uint8, device=c.device)
    call("gvc_volume_metrics", S, _ptr(c), _ptr(g), _ptr(out), _ptr(ws), nb, _stream(c.device))
    return out


def _psnr(mse):
    # 20 log10(1) - 10 log10(mse + 1e-8)
    return 20 * math.log10(1.0) - 10 * math.log10(mse + 1e-8)


def volume_metrics(cells, gt, grid=True):
    """The metrics R/volume_metrics.py:331-369 prints, as a dict (one small read).
    ``grid``: the grid-mode percentages (of all samples) instead of the mesh-mode
    counts and false-positive share of the invalid samples."""
    s = metric_sums(cells, gt).cpu().numpy()
    S = int(cells.numel())
    mean = lambda x, n: x / n if n else float("nan")
    l1, l2 = mean(s[0], S), mean(s[1], S)
    mse2, mse3, mse4, mse5 = mean(s[2], s[3]), mean(s[4], S), mean(s[5], S), mean(s[6], s[7])
    out = {
        "L1": l1, "L2": l2,
        "PSNR": _psnr(l2),
        "PSNR without false positives/negatives": _psnr(mse2),
        "PSNR of original": _psnr(mse3),
        "PSNR with 0s for fn": _psnr(mse4),
        "PSNR with 0s for fn and no fp": _psnr(mse5),
        "samples": S, "invalid": int(s[8]), "false_negative": int(s[9]), "false_positive": int(s[10]),
    }
    if grid:
        out["percent_invalid"] = s[8] / S
        out["false_negative_percent"] = s[9] / S
        out["false_positive_percent"] = s[10] / S
    else:
        out["false_positive_percent"] = s[10] / s[8] if s[8] else float("nan")
    return out


def evaluate(means3D, scales, rotations, values, weights, samples, gt, grid=True, scale_modifier=1.0, params=None):
    """Render the mixture (activated parameters) at ``samples`` [S,3] and compare
    with ``gt`` [S] (-1 outside the volume).  Returns (metrics dict, cells)."""
    from . import mixture as mx
    with torch.no_grad():
        ss = mx.SampleSet(samples.to(torch.float32).contiguous())
        cells, _, _, _ = mx.mixture_forward(means3D, scales, rotations, values, weights, scale_modifier, ss,
                                            params=params)
    return volume_metrics(cells, gt, grid=grid), cells


def write_vtk_structured(path, values, spacing=(1.0, 1.0, 1.0)):
    """ASCII legacy-VTK structured points, as tensor_to_vtk writes them
    (R/utils/debug_utils.py:7-37): DIMENSIONS from the array's shape, origin 0,
    the values in C order, one per line, declared double."""
    a = np.asarray(values)
    nx, ny, nz = a.shape
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\n3D Scalar Data\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {nx} {ny} {nz}\nORIGIN 0.0 0.0 0.0\n")
        f.write(f"SPACING {spacing[0]} {spacing[1]} {spacing[2]}\n")
        f.write(f"POINT_DATA {nx * ny * nz}\nSCALARS values double\nLOOKUP_TABLE default\n")
        for v in a.flatten():
            f.write(f"{v}\n")
