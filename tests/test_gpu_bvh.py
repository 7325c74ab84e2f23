"""The Gaussian BVH a forward builds, read back through gvc_forward_tree, against
the oracle: bit-exact Morton keys and order, topology and node boxes.

The reference builds its Gaussian tree with cuBQL's radixBuilder
(DGR/ext.cu:39-42, CR/rasterizer_impl.cu:78-83); the canonical tree here is the
Karras radix tree over leaves of L consecutive Gaussians in Morton order.  Bars:
  * order and keys: the stable sort of the top 24 bits of oracle.morton3d_unit
    of the means (equal keys keep the caller's order), bit for bit;
  * links and leaf ranges: oracle.karras_build over the leaves' first keys
    (pinned in tests/test_oracle.py against a recursive top-down split);
  * node boxes: the brute-force union of the oracle's Gaussian boxes
    (oracle.preprocess) of every leaf under each child, with the kernels' NaN
    rule (fminf / fmaxf skip NaN; a leaf whose Gaussians all have a NaN lower x
    bound has a NaN box), bit for bit.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

import oracle  # noqa: E402
from gvc import mixture as mx  # noqa: E402

f32 = np.float32
DEV = "cuda:0"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV)


def gaussians(P, dup, seed):
    rng = np.random.default_rng(seed)
    means = rng.uniform(0.0, 1.0, (P, 3)).astype(f32)
    if dup and P > 2:
        k = max(P // 5, 1)
        src = rng.integers(0, P, k)
        dst = rng.integers(0, P, k)
        means[dst] = means[src]                       # identical means: equal keys
        near = rng.integers(0, P, k)
        means[near] = (np.floor(means[near] * 2048.0) / 2048.0 + 1e-6).astype(f32)  # same 2^-11 cell
    if P >= 17:
        means[3] = [-0.05, 0.5, 1.2]                  # outside the unit cube
        means[5] = [np.nan, 0.25, 0.75]               # NaN coordinate: NaN box
    scales = np.exp(rng.normal(np.log(0.01), 0.3, (P, 3))).astype(f32)
    q = rng.normal(size=(P, 4))
    rots = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(f32)
    values = rng.uniform(0, 1, (P, 1)).astype(f32)
    weights = rng.uniform(0.01, 1, (P, 1)).astype(f32)
    return means, scales, rots, values, weights


def leaf_boxes(box_sorted, L):
    """Per leaf: NaN box when every Gaussian's lower x bound is NaN, else the
    per-coordinate fmin / fmax over its Gaussians (NaN skipped)."""
    P = box_sorted.shape[0]
    NL = (P + L - 1) // L
    out = np.empty((NL, 6), f32)
    starts = np.arange(NL) * L
    lo = np.fmin.reduceat(box_sorted[:, :3], starts, axis=0)
    hi = np.fmax.reduceat(box_sorted[:, 3:], starts, axis=0)
    out[:, :3], out[:, 3:] = lo, hi
    anyx = np.logical_or.reduceat(~np.isnan(box_sorted[:, 0]), starts)
    out[~anyx] = np.nan
    return out


def child_range(c, first, last):
    return (~c, ~c) if c < 0 else (int(first[c]), int(last[c]))


@pytest.mark.parametrize("P,leaf,dup", [(2, 1, False), (17, 1, True), (17, 4, False), (1000, 1, True),
                                        (1000, 4, True), (1000, 16, False), (1_000_000, 4, True)])
def test_tree_bit_exact(P, leaf, dup):
    means, scales, rots, values, weights = gaussians(P, dup, seed=P + leaf)
    samples = np.random.default_rng(1).uniform(0, 1, (64, 3)).astype(f32)
    ss = mx.SampleSet(T(samples))
    _, _, _, st = mx.mixture_forward(T(means), T(scales), T(rots), T(values), T(weights), 1.0, ss,
                                     mx.MixtureParams(leaf_size=leaf), want_conics=False)
    tree = st.tree()
    order = tree["order"].cpu().numpy()
    keys = tree["keys"].cpu().numpy().astype(np.uint64)
    nodes = tree["nodes"].cpu().numpy()
    boxes = tree["boxes"].cpu().numpy()

    codes = oracle.morton3d_unit(means) >> np.uint64(39)
    ref_order = np.argsort(codes, kind="stable")
    assert np.array_equal(order, ref_order)
    assert np.array_equal(keys, codes[ref_order])
    if dup:
        assert np.unique(keys).size < P            # ties were exercised

    leaf_keys = keys[::leaf]
    NL = leaf_keys.size
    left, right, first, last = oracle.karras_build(leaf_keys)
    assert nodes.shape == (max(NL - 1, 0), 4)
    assert np.array_equal(nodes[:, 0], left) and np.array_equal(nodes[:, 1], right)
    assert np.array_equal(nodes[:, 2], first) and np.array_equal(nodes[:, 3], last)
    if NL < 2:
        return
    assert (first[0], last[0]) == (0, NL - 1)

    _, box, _ = oracle.preprocess(means, scales, rots, weights)
    lb = leaf_boxes(box[ref_order], leaf)
    NI = NL - 1
    check = np.arange(NI)
    if NI > 30000:  # every node near the root (long ranges) and a random sample of the rest
        rng = np.random.default_rng(7)
        check = np.unique(np.concatenate([np.arange(2000), rng.integers(0, NI, 20000)]))
    for i in check:
        for side, c in ((0, int(left[i])), (1, int(right[i]))):
            a, b = child_range(c, first, last)
            want = np.concatenate([np.fmin.reduce(lb[a:b + 1, :3], axis=0), np.fmax.reduce(lb[a:b + 1, 3:], axis=0)])
            got = boxes[i, 6 * side:6 * side + 6]
            assert np.array_equal(got, want.astype(f32), equal_nan=True), (i, side, got, want)
