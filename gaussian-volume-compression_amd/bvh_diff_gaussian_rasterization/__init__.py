"""Drop-in for the reference's ``bvh_diff_gaussian_rasterization`` package
(DGR/bvh_diff_gaussian_rasterization/__init__.py:1-230), running on libgvc.

Public surface kept as in the reference: ``GaussianRasterizationSettings``,
``GaussianRasterizer`` (``build_bvh``, ``forward``, ``intersect``),
``rasterize_gaussians``, ``_RasterizeGaussians``, ``cpu_deep_copy_tuple``, ``_C``.

Differences a caller can see:
  * ``GaussianRasterizer.build_bvh`` keeps its sample set on the rasterizer (and as
    ``_C``'s current one); the autograd node keeps the sample set and the Gaussian
    BVH of its forward, so rebuilding between forward and backward is safe.
  * Gradients are those of the forward formula (the reference's scale/rotation
    chain, CR/backward.cu:68-124, is not reproduced).
  * Pair membership uses the exact box predicate of every Gaussian, independent of
    how the BVH is partitioned.
"""
from typing import NamedTuple

import torch
import torch.nn as nn

from . import _C


def cpu_deep_copy_tuple(input_tuple):
    """DGR/bvh_diff_gaussian_rasterization/__init__.py:7-12."""
    copied_tensors = [item.cpu().clone() if isinstance(item, torch.Tensor) else item for item in input_tuple]
    return tuple(copied_tensors)


def rasterize_gaussians(means3D, scales, rotations, values, weights, raster_settings, sample_set=None):
    """DGR/bvh_diff_gaussian_rasterization/__init__.py:15-30.  ``sample_set``
    defaults to the one built by the last ``build_bvh``."""
    return _RasterizeGaussians.apply(means3D, scales, rotations, values, weights, raster_settings, sample_set)


class _RasterizeGaussians(torch.autograd.Function):
    """DGR/bvh_diff_gaussian_rasterization/__init__.py:33-140."""

    @staticmethod
    def forward(ctx, means3D, scales, rotations, values, weights, raster_settings, sample_set=None):
        ss = sample_set if sample_set is not None else _C.current_sample_set()
        vmin = raster_settings.volume_mins
        vmax = raster_settings.volume_maxes
        cells, cell_weights, conics, state = _C.rasterize_gaussians(
            means3D, scales, rotations, values, weights, raster_settings.scale_modifier, vmin[0], vmin[1], vmin[2],
            vmax[0], vmax[1], vmax[2], raster_settings.bg, raster_settings.use_gaussian_bvh, raster_settings.debug,
            sample_set=ss, return_state=True)
        ctx.raster_settings = raster_settings
        ctx.sample_set = ss
        ctx.forward_state = state
        ctx.save_for_backward(means3D, scales, rotations, values, weights, cells, cell_weights, conics)
        return cells, cell_weights

    @staticmethod
    def backward(ctx, grad_out_cells, grad_out_cell_weights):
        rs = ctx.raster_settings
        means3D, scales, rotations, values, weights, cells, cell_weights, conics = ctx.saved_tensors
        if grad_out_cells is None:
            grad_out_cells = torch.zeros_like(cells)
        if grad_out_cell_weights is None:
            grad_out_cell_weights = torch.zeros_like(cell_weights)
        vmin, vmax = rs.volume_mins, rs.volume_maxes
        gm, gs, gr, gv, gw = _C.rasterize_gaussians_backward(
            means3D, scales, rotations, conics, values, weights, cells, cell_weights, rs.scale_modifier, vmin[0],
            vmin[1], vmin[2], vmax[0], vmax[1], vmax[2], rs.bg, grad_out_cells, grad_out_cell_weights,
            rs.use_gaussian_bvh, rs.debug, sample_set=ctx.sample_set, forward_state=ctx.forward_state)
        ctx.forward_state = None
        return (gm.view_as(means3D), gs.view_as(scales), gr.view_as(rotations), gv.view(values.shape),
                gw.view(weights.shape), None, None)


class GaussianRasterizationSettings(NamedTuple):
    """DGR/bvh_diff_gaussian_rasterization/__init__.py:143-150."""
    volume_mins: tuple
    volume_maxes: tuple
    cell_count: int
    bg: float
    scale_modifier: float
    use_gaussian_bvh: bool
    debug: bool


class GaussianRasterizer(nn.Module):
    """DGR/bvh_diff_gaussian_rasterization/__init__.py:153-230."""

    def __init__(self, raster_settings):
        super().__init__()
        self.raster_settings = raster_settings
        self.sample_set = None

    def build_bvh(self, samples, force_debug=False, use_gaussian_bvh=False):
        debug = True if force_debug else self.raster_settings.debug
        _C.build_bvh(samples, debug, use_gaussian_bvh)
        self.sample_set = _C.current_sample_set()

    def forward(self, means3D, scales=None, rotations=None, values=None, weights=None, debug=False):
        raster_settings = self.raster_settings._replace(debug=debug) if debug else self.raster_settings
        if scales is None or rotations is None:
            raise Exception("Please provide scale/rotation pair!")
        if values is None:
            raise Exception("Please provide scalar values for each Gaussian!")
        if weights is None:
            raise Exception("Please provide scalar weights for each Gaussian!")
        return rasterize_gaussians(means3D, scales, rotations, values, weights, raster_settings,
                                   sample_set=self.sample_set)

    def intersect(self, means3D, scales=None, rotations=None, debug=False):
        """Dead in the reference (it calls an undefined ``intersect_gaussians``,
        DGR/bvh_diff_gaussian_rasterization/__init__.py:207-230); same error here."""
        if scales is None or rotations is None:
            raise Exception("Please provide scale/rotation pair!")
        raise NameError("name 'intersect_gaussians' is not defined")
