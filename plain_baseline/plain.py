"""torch-level driver of the plain GPU sequence (one volume at a time)."""
import ctypes

import torch

from gvc import mixture as mx

from . import lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _st(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hip error {rc}")


class PlainSampleTree:
    """Sample BVH of one volume: points in Morton order, 256-point leaves, an
    implicit binary tree of boxes over the leaves (rebuilt every step, as the
    reference rebuilds its sample BVH)."""

    def __init__(self, samples):
        dev = samples.device
        S = int(samples.shape[0])
        self.S = S
        self.samples = samples
        L = lib()
        NP = L.plain_leaf_slots(S)
        codes = torch.empty(max(S, 1), dtype=torch.int64, device=dev)
        idx = torch.empty(max(S, 1), dtype=torch.int32, device=dev)
        self.pts = torch.empty((max(S, 1), 4), dtype=torch.float32, device=dev)
        self.lo = torch.empty((2 * NP - 1, 4), dtype=torch.float32, device=dev)
        self.hi = torch.empty((2 * NP - 1, 4), dtype=torch.float32, device=dev)
        order = idx
        if S > 0:
            _ok(L.plain_codes(S, _p(samples), _p(codes), _p(idx), _st(dev)), "plain_codes")
            _, order = mx.sort_pairs_u64(codes[:S], idx[:S], 63)
        _ok(L.plain_build(S, _p(samples), _p(order), _p(self.pts), _p(self.lo), _p(self.hi), _st(dev)), "plain_build")


def plain_forward(means, scales, rots, values, weights, k, tree, pairs):
    """(cells, cell_weights, conics, boxes); pairs (int64 [1]) += accepted pairs."""
    dev = means.device
    P, S = int(means.shape[0]), tree.S
    conics = torch.empty((P, 6), dtype=torch.float32, device=dev)
    boxes = torch.empty((P, 6), dtype=torch.float32, device=dev)
    cells = torch.empty(S, dtype=torch.float32, device=dev)
    cw = torch.empty(S, dtype=torch.float32, device=dev)
    L = lib()
    _ok(L.plain_preprocess(P, _p(means), _p(scales), _p(rots), _p(weights), float(k), _p(conics), _p(boxes),
                           _st(dev)), "plain_preprocess")
    _ok(L.plain_forward(P, _p(means), _p(values), _p(weights), _p(conics), _p(boxes), S, _p(tree.pts), _p(tree.lo),
                        _p(tree.hi), _p(cells), _p(cw), _p(pairs), _st(dev)), "plain_forward")
    return cells, cw, conics, boxes


def plain_backward(means, scales, rots, values, weights, k, conics, boxes, tree, cells, cw, g_cells, g_cw):
    dev = means.device
    P = int(means.shape[0])
    gm = torch.empty((P, 3), dtype=torch.float32, device=dev)
    gs = torch.empty((P, 3), dtype=torch.float32, device=dev)
    gr = torch.empty((P, 4), dtype=torch.float32, device=dev)
    gv = torch.empty((P, 1), dtype=torch.float32, device=dev)
    gw = torch.empty((P, 1), dtype=torch.float32, device=dev)
    _ok(lib().plain_backward(P, _p(means), _p(scales), _p(rots), _p(values), _p(weights), float(k), _p(conics),
                             _p(boxes), tree.S, _p(tree.pts), _p(tree.lo), _p(tree.hi), _p(cells), _p(cw),
                             _p(g_cells), _p(g_cw), _p(gm), _p(gs), _p(gr), _p(gv), _p(gw), _st(dev)),
        "plain_backward")
    return gm, gs, gr, gv, gw


class _PlainSequence(torch.autograd.Function):
    """V volumes one after another: per volume preprocess, forward walk with
    atomics and normalise; the backward walks each volume again."""

    @staticmethod
    def forward(ctx, means, scales, rots, values, weights, g_off, trees, pairs):
        outs, saved = [], []
        for v, tree in enumerate(trees):
            a, b = int(g_off[v]), int(g_off[v + 1])
            cells, cw, conics, boxes = plain_forward(means[a:b], scales[a:b], rots[a:b], values[a:b], weights[a:b],
                                                     1.0, tree, pairs)
            outs.append((cells, cw))
            saved.append((conics, boxes))
        cells = torch.cat([o[0] for o in outs])
        cw = torch.cat([o[1] for o in outs])
        ctx.trees, ctx.g_off, ctx.conics = trees, g_off, saved
        ctx.save_for_backward(means, scales, rots, values, weights, cells, cw)
        return cells, cw

    @staticmethod
    def backward(ctx, g_cells, g_cw):
        means, scales, rots, values, weights, cells, cw = ctx.saved_tensors
        g_cells = g_cells.contiguous()
        g_cw = g_cw.contiguous()
        grads = [[] for _ in range(5)]
        s0 = 0
        for v, tree in enumerate(ctx.trees):
            a, b = int(ctx.g_off[v]), int(ctx.g_off[v + 1])
            s1 = s0 + tree.S
            conics, boxes = ctx.conics[v]
            gv = plain_backward(means[a:b], scales[a:b], rots[a:b], values[a:b], weights[a:b], 1.0, conics, boxes,
                                tree, cells[s0:s1], cw[s0:s1], g_cells[s0:s1], g_cw[s0:s1])
            for k in range(5):
                grads[k].append(gv[k])
            s0 = s1
        gm, gs, gr, gvv, gw = (torch.cat(g) for g in grads)
        return gm, gs, gr, gvv.view_as(values), gw.view_as(weights), None, None, None


def plain_sequence(means, scales, rots, values, weights, g_off, trees, pairs):
    """Differentiable: (cells, cell_weights) of every volume, concatenated.
    g_off: Gaussian offsets per volume (V + 1 ints); trees: one PlainSampleTree
    per volume; pairs (int64 [1]) accumulates the accepted pairs."""
    return _PlainSequence.apply(means, scales, rots, values, weights, g_off, trees, pairs)
