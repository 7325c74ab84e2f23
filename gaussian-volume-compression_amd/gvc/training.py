"""Fused training-step glue of libgvc as autograd Functions.

``activate`` mirrors the reference model's parameter activations
(R/scene/gaussian_model.py:31-48, :71-77, :114-132: capped exp scales,
normalised quaternions, sigmoid weights and values); ``fit_loss`` mirrors the
fitting loss of R/train.py:389-420.  Both run as single HIP kernels per
direction, with deterministic FP64 reductions and no host synchronisation, so a
training step built from them can be captured in a graph.
"""
import torch

from ._lib import call, lib
from .mixture import _as_f32, _ptr, _stream

REFERENCE_MAX_SCALE = 0.02   # R/scene/gaussian_model.py:64 (self.max_scale)


class _Activate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_scales, rotations, weight_logits, value_logits, max_scale):
        ls = _as_f32(log_scales, "log_scales", 3)
        q = _as_f32(rotations, "rotations", 4)
        wl = _as_f32(weight_logits, "weight_logits", 1)
        vl = _as_f32(value_logits, "value_logits", 1)
        P = int(ls.shape[0])
        dev = ls.device
        scales = torch.empty((P, 3), dtype=torch.float32, device=dev)
        qn = torch.empty((P, 4), dtype=torch.float32, device=dev)
        w = torch.empty((P, 1), dtype=torch.float32, device=dev)
        v = torch.empty((P, 1), dtype=torch.float32, device=dev)
        call("gvc_activate_forward", P, _ptr(ls), _ptr(q), _ptr(wl), _ptr(vl), float(max_scale), _ptr(scales),
             _ptr(qn), _ptr(w), _ptr(v), _stream(dev))
        ctx.save_for_backward(ls, q, wl, vl)
        ctx.max_scale = float(max_scale)
        ctx.shapes = (weight_logits.shape, value_logits.shape)
        return scales, qn, w.reshape(weight_logits.shape), v.reshape(value_logits.shape)

    @staticmethod
    def backward(ctx, gs, gq, gw, gv):
        ls, q, wl, vl = ctx.saved_tensors
        P = int(ls.shape[0])
        dev = ls.device

        def g(t, shape):
            return torch.zeros(shape, dtype=torch.float32, device=dev) if t is None else t.float().contiguous()

        gs, gq = g(gs, (P, 3)), g(gq, (P, 4))
        gw, gv = g(gw, ctx.shapes[0]), g(gv, ctx.shapes[1])
        dls = torch.empty_like(ls)
        dq = torch.empty_like(q)
        dwl = torch.empty((P,), dtype=torch.float32, device=dev)
        dvl = torch.empty((P,), dtype=torch.float32, device=dev)
        call("gvc_activate_backward", P, _ptr(ls), _ptr(q), _ptr(wl), _ptr(vl), ctx.max_scale, _ptr(gs), _ptr(gq),
             _ptr(gw), _ptr(gv), _ptr(dls), _ptr(dq), _ptr(dwl), _ptr(dvl), _stream(dev))
        return dls, dq, dwl.reshape(ctx.shapes[0]), dvl.reshape(ctx.shapes[1]), None


def activate(log_scales, rotations, weight_logits, value_logits, max_scale=REFERENCE_MAX_SCALE):
    """(scales [P,3], rotations [P,4], weights, values) from the raw parameters."""
    return _Activate.apply(log_scales, rotations, weight_logits, value_logits, max_scale)


class _FitLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cells, weights, gt, fn_reg, fp_reg, encode_surface):
        c = _as_f32(cells, "cells")
        w = _as_f32(weights, "weights")
        g = gt.to(torch.float32).contiguous()
        S = int(c.numel())
        if w.numel() != S or g.numel() != S:
            raise ValueError("cells, weights and gt must have one entry per sample")
        dev = c.device
        nbytes = lib().gvc_loss_state_bytes(S)
        state = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        loss = torch.empty((), dtype=torch.float32, device=dev)
        call("gvc_loss_forward", S, _ptr(c), _ptr(w), _ptr(g), float(fn_reg), float(fp_reg), int(bool(encode_surface)),
             _ptr(loss), _ptr(state), nbytes, _stream(dev))
        ctx.save_for_backward(c, w, g, state)
        ctx.args = (float(fn_reg), float(fp_reg), int(bool(encode_surface)), cells.shape, weights.shape)
        return loss

    @staticmethod
    def backward(ctx, up):
        c, w, g, state = ctx.saved_tensors
        fn_reg, fp_reg, surf, cshape, wshape = ctx.args
        S = int(c.numel())
        dc = torch.empty_like(c)
        dw = torch.empty_like(w)
        up = up.to(torch.float32).contiguous()
        call("gvc_loss_backward", S, _ptr(c), _ptr(w), _ptr(g), fn_reg, fp_reg, surf, _ptr(up), _ptr(state),
             state.numel(), _ptr(dc), _ptr(dw), _stream(c.device))
        return dc.reshape(cshape), dw.reshape(wshape), None, None, None, None


def fit_loss(cells, weights, gt, fn_reg=1.0, fp_reg=0.0, encode_surface=False):
    """Scalar loss of R/train.py:389-420 (gt: -1 outside the volume)."""
    return _FitLoss.apply(cells, weights, gt, fn_reg, fp_reg, encode_surface)
