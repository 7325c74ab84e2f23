"""Batches of independent volumes fitted in one launch per pass.

Each volume has its own Gaussians, samples, Morton order and BVH; a 64-point
sample group never mixes volumes.  Inputs are concatenated volume-major.  The
reference fits one volume per process (R/train.py:135-190, one mesh per run);
a batch runs V such fits side by side, and with the same inputs every volume's
outputs and gradients equal the single-volume calls of ``gvc.mixture`` bit for
bit.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import mixture as mx
from ._lib import call, gvc_batch, lib
from .mixture import _as_f32, _ptr, _stream


class Batch:
    """Offsets of V volumes (Gaussians and samples per volume) on the device.
    Describing a batch copies a small table to the device and synchronises:
    do it once per batch shape, outside captured regions."""

    def __init__(self, gaussians_per_volume, samples_per_volume, leaf_size=mx.DEFAULT_LEAF_SIZE, device="cuda:0"):
        P = np.ascontiguousarray(gaussians_per_volume, dtype=np.int32)
        S = np.ascontiguousarray(samples_per_volume, dtype=np.int32)
        if P.ndim != 1 or P.shape != S.shape or P.size == 0:
            raise ValueError("need one Gaussian count and one sample count per volume")
        self.device = torch.device(device)
        self.V = int(P.size)
        self.gaussians = P.astype(np.int64)
        self.samples = S.astype(np.int64)
        self.P_off = np.concatenate([[0], np.cumsum(self.gaussians)])
        self.S_off = np.concatenate([[0], np.cumsum(self.samples)])
        nbytes = lib().gvc_batch_table_bytes(self.V)
        self.table = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self.c = gvc_batch()
        call("gvc_batch_describe", self.V, P.ctypes.data_as(ctypes.c_void_p), S.ctypes.data_as(ctypes.c_void_p),
             int(leaf_size), _ptr(self.table), nbytes, ctypes.byref(self.c), _stream(self.device))
        self.leaf_size = int(leaf_size)
        self.P = int(self.c.P_total)
        self.S = int(self.c.S_total)

    def gaussian_slice(self, v):
        return slice(int(self.P_off[v]), int(self.P_off[v + 1]))

    def sample_slice(self, v):
        return slice(int(self.S_off[v]), int(self.S_off[v + 1]))


class BatchedSampleSet:
    """The samples of every volume, each volume in its own Morton-sorted groups."""

    def __init__(self, batch, samples):
        samples = _as_f32(samples, "samples", 3)
        if samples.shape[0] != batch.S:
            raise ValueError(f"batch has {batch.S} samples, got {samples.shape[0]}")
        self.batch = batch
        self.samples = samples
        self.device = samples.device
        self.S = batch.S
        nbytes = lib().gvc_sample_set_bytes_batched(ctypes.byref(batch.c))
        self.buffer = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        call("gvc_sample_set_build_batched", ctypes.byref(batch.c), _ptr(samples), _ptr(self.buffer), nbytes,
             _stream(self.device))


class BatchedForwardState:
    """Records, counts and candidate log of a batched forward."""

    def __init__(self, batch, buffer, cparams, params):
        self.batch = batch
        self.buffer = buffer
        self.params = params
        self._c = cparams
        self.P = batch.P
        self.S = batch.S
        self.chunk_slots_used = 0
        self.lane_groups = 0

    def stats(self):
        """(n_records, n_pairs, error_flags) over the whole batch; synchronises;
        also sets ``chunk_slots_used`` and ``lane_groups``."""
        nr = ctypes.c_uint64(0)
        npairs = ctypes.c_uint64(0)
        err = ctypes.c_uint32(0)
        used = ctypes.c_uint64(0)
        lanes = ctypes.c_uint64(0)
        call("gvc_forward_stats_batched", ctypes.byref(self.batch.c), _ptr(self.buffer), ctypes.byref(self._c),
             ctypes.byref(nr), ctypes.byref(npairs), ctypes.byref(err), ctypes.byref(used), ctypes.byref(lanes),
             _stream(self.buffer.device))
        self.chunk_slots_used = int(used.value)
        self.lane_groups = int(lanes.value)
        return int(nr.value), int(npairs.value), int(err.value)

    def pair_counts(self, sample_set):
        """Accepted pairs per sample [S_total] and per Gaussian [P_total] (int32
        device tensors, volume-major positions; needs the candidate log)."""
        dev = self.buffer.device
        sc = torch.empty(self.S, dtype=torch.int32, device=dev)
        gc = torch.empty(self.P, dtype=torch.int32, device=dev)
        call("gvc_forward_pair_counts_batched", ctypes.byref(self.batch.c), _ptr(self.buffer),
             ctypes.byref(self._c), _ptr(sample_set.buffer), _ptr(sc), _ptr(gc), _stream(dev))
        return sc, gc

    def accumulate(self, accum):
        """Stream-ordered accum (int64 [3]) += (pairs, records), accum[2] |= flags."""
        call("gvc_forward_accumulate_batched", ctypes.byref(self.batch.c), _ptr(self.buffer), ctypes.byref(self._c),
             _ptr(accum), _stream(self.buffer.device))


def _check_inputs(batch, means3D, scales, rotations, values, weights):
    means3D = _as_f32(means3D, "means3D", 3)
    scales = _as_f32(scales, "scales", 3)
    rotations = _as_f32(rotations, "rotations", 4)
    values = _as_f32(values, "values", 1)
    weights = _as_f32(weights, "weights", 1)
    for name, t in (("means3D", means3D), ("scales", scales), ("rotations", rotations), ("values", values),
                    ("weights", weights)):
        if t.shape[0] != batch.P:
            raise ValueError(f"{name} has {t.shape[0]} rows, the batch has {batch.P} Gaussians")
    return means3D, scales, rotations, values, weights


def _forward(batch, cp, params, inputs, scale_modifier, sample_set, want_conics, outputs=None):
    means3D, scales, rotations, values, weights = inputs
    dev = means3D.device
    nbytes = lib().gvc_forward_state_bytes_batched(ctypes.byref(batch.c), ctypes.byref(cp))
    if nbytes == 0:
        raise _lib.GvcError("invalid mixture parameters")
    state = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    if outputs is None:
        cells = torch.empty(batch.S, dtype=torch.float32, device=dev)
        cw = torch.empty(batch.S, dtype=torch.float32, device=dev)
    else:
        cells, cw = outputs
    conics = torch.empty(6 * batch.P, dtype=torch.float32, device=dev) if want_conics else None
    mx._staged_call("gvc_mixture_forward_batched",
                    (ctypes.byref(batch.c), _ptr(means3D), _ptr(scales), _ptr(rotations), _ptr(values),
                     _ptr(weights), float(scale_modifier), _ptr(sample_set.buffer)), cp,
                    (_ptr(cells), _ptr(cw), _ptr(conics), _ptr(state), nbytes, _stream(dev)),
                    ((_lib.GVC_STAGE_BUILD, "fwd_build"), (_lib.GVC_STAGE_WALK, "fwd_walk"),
                     (_lib.GVC_STAGE_COUNTS, "fwd_counts")))
    fs = BatchedForwardState(batch, state, cp, params)
    if mx._audit is not None:
        mx._audit.append(fs)
    return cells, cw, conics, fs


def mixture_forward_batched(batch, means3D, scales, rotations, values, weights, scale_modifier, sample_set,
                            params=None, want_conics=True):
    """Returns (cells [S_total], cell_weights [S_total], conics [6 P_total] or None, state)."""
    params = params or mx.default_params()
    if params.leaf_size != batch.leaf_size:
        raise ValueError(f"params.leaf_size {params.leaf_size} differs from the batch's {batch.leaf_size}")
    if sample_set.batch is not batch:
        raise ValueError("the sample set was built for another batch")
    inputs = _check_inputs(batch, means3D, scales, rotations, values, weights)
    return _forward(batch, params.c(batch.c.groups), params, inputs, scale_modifier, sample_set, want_conics)


def mixture_backward_batched(fstate, means3D, scales, rotations, values, weights, scale_modifier, sample_set,
                             out_cells, out_weights, grad_cells, grad_weights, record_capacity=None):
    """Gradients (means3D, scales, rotations, values, weights) over the whole batch."""
    batch = fstate.batch
    P, S = batch.P, batch.S
    dev = fstate.buffer.device
    inputs = _check_inputs(batch, means3D, scales, rotations, values, weights)
    out_cells = _as_f32(out_cells, "out_cells")
    out_weights = _as_f32(out_weights, "out_weights")
    grad_cells = _as_f32(grad_cells, "grad_cells")
    grad_weights = _as_f32(grad_weights, "grad_weights")
    if grad_cells.numel() != S or grad_weights.numel() != S:
        raise ValueError("upstream gradients must have one entry per sample")
    gm = torch.empty((P, 3), dtype=torch.float32, device=dev)
    gs = torch.empty((P, 3), dtype=torch.float32, device=dev)
    gr = torch.empty((P, 4), dtype=torch.float32, device=dev)
    gv = torch.empty((P, 1), dtype=torch.float32, device=dev)
    gw = torch.empty((P, 1), dtype=torch.float32, device=dev)
    cap = record_capacity if record_capacity is not None else mx._record_capacity
    if cap is None:
        scratch = []

        def rerun(cp2):
            if not scratch:
                scratch.extend(torch.empty(S, dtype=torch.float32, device=dev) for _ in range(2))
            return _forward(batch, cp2, fstate.params, inputs, scale_modifier, sample_set, False,
                            outputs=tuple(scratch))[3]

        fstate, cap = mx._redo_until_log_fits(fstate, batch.c.groups, rerun)
    cp = fstate._c
    nbytes = lib().gvc_backward_bytes(P, cap)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    _, scales, rotations, _, _ = inputs
    mx._staged_call("gvc_mixture_backward_batched",
                    (ctypes.byref(batch.c), _ptr(scales), _ptr(rotations), float(scale_modifier),
                     _ptr(sample_set.buffer)), cp,
                    (_ptr(out_cells), _ptr(out_weights), _ptr(grad_cells), _ptr(grad_weights), _ptr(fstate.buffer),
                     fstate.buffer.numel(), ctypes.c_uint64(cap), _ptr(ws), nbytes, _ptr(gm), _ptr(gs), _ptr(gr),
                     _ptr(gv), _ptr(gw), _stream(dev)),
                    ((_lib.GVC_STAGE_WALK, "bwd_walk"), (_lib.GVC_STAGE_REDUCE, "bwd_reduce")))
    return gm, gs, gr, gv, gw


class _RasterizeBatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, scales, rotations, values, weights, scale_modifier, sample_set, params):
        cells, cw, _, fs = mixture_forward_batched(sample_set.batch, means3D, scales, rotations, values, weights,
                                                   scale_modifier, sample_set, params, want_conics=False)
        ctx.fs = fs
        ctx.scale_modifier = scale_modifier
        ctx.sample_set = sample_set
        ctx.save_for_backward(means3D, scales, rotations, values, weights, cells, cw)
        return cells, cw

    @staticmethod
    def backward(ctx, g_cells, g_cw):
        means3D, scales, rotations, values, weights, cells, cw = ctx.saved_tensors
        if g_cells is None:
            g_cells = torch.zeros_like(cells)
        if g_cw is None:
            g_cw = torch.zeros_like(cw)
        gm, gs, gr, gv, gw = mixture_backward_batched(ctx.fs, means3D, scales, rotations, values, weights,
                                                      ctx.scale_modifier, ctx.sample_set, cells, cw,
                                                      g_cells.contiguous(), g_cw.contiguous())
        return gm, gs, gr, gv.view_as(values), gw.view_as(weights), None, None, None


def rasterize_batch(means3D, scales, rotations, values, weights, sample_set, scale_modifier=1.0, params=None):
    """Differentiable batched mixture: (cells [S_total], cell_weights [S_total])."""
    return _RasterizeBatch.apply(means3D, scales, rotations, values, weights, float(scale_modifier), sample_set,
                                 params)
