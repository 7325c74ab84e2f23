"""Autotuner for the mixture kernels.

Sweeps BVH leaf size x points per walk x block shape on the caller's own data,
times a forward + backward per configuration with HIP events, and keeps the
fastest.  Points per walk is how a wavefront's 64 points (one per lane) share
the tree walk: 1 = every lane walks for its own point (lane mode), k = runs of
up to k consecutive points share one walk of their box (run_points).  The
reference's launch shape has no counterpart (one warp per Gaussian over a
sample BVH); the sweep's axes are leaf size {2, 4, 8, 16}, points per wavefront
{1, 4, 16} and block size {256, 512, 1024}; the 64-point shared walk is added
because it is the measured default.  Results
are cached per (log2 P, log2 S) bucket in memory and, if ``GVC_TUNE_CACHE``
names a file, in that JSON file.  Every configuration accepts exactly the
same (Gaussian, point) pairs; the order in which a point's compensated sum
visits its Gaussians follows the traversal, so fields and gradients of two
configurations agree to the accuracy of the compensated sums, not bit for bit.
"""
import json
import math
import os

import torch

from . import mixture as mx
from ._lib import GvcError

LEAF_SIZES = (2, 4, 8, 16)
POINTS_PER_WALK = (1, 4, 16, 64)
BLOCKS = (256, 512, 1024)

_cache = {}


def _bucket(P, S):
    return f"{int(round(math.log2(max(P, 1))))}:{int(round(math.log2(max(S, 1))))}"


def _load_file():
    path = os.environ.get("GVC_TUNE_CACHE")
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}
    return {}


def _save_file(entries):
    path = os.environ.get("GVC_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data.update(entries)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def _time(run, reps):
    """Best of ``reps`` timed calls of run() (after one untimed call), in ms."""
    ev = []
    for r in range(reps + 1):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        if r > 0:
            ev.append((e0, e1))
    torch.cuda.synchronize()
    return min(a.elapsed_time(b) for a, b in ev)


def params_for(base, leaf_size, points_per_walk, block_threads):
    """MixtureParams of one configuration of the sweep (base's precision and
    candidate-log choices kept)."""
    lane = "always" if points_per_walk == 1 else base.lane_mode
    rp = 64 if points_per_walk == 1 else int(points_per_walk)
    return mx.MixtureParams(leaf_size=leaf_size, block_threads=block_threads, run_points=rp, kahan=base.kahan,
                            fp64_cond=base.fp64_cond, candidate_log=base.candidate_log, lane_mode=lane,
                            fp64_scale_ratio=base.fp64_scale_ratio, factor_scale_ratio=base.factor_scale_ratio)


def _sweep(make_run, base, leaf_sizes, points_per_walk, blocks, reps):
    """Times make_run(params)() for every configuration; returns (best, its ms, table)."""
    table = {}
    best, best_ms = None, float("inf")
    for L in leaf_sizes:
        for ppw in points_per_walk:
            for bt in blocks:
                prm = params_for(base, L, ppw, bt)
                label = f"leaf{L}/points{ppw}/block{bt}"
                try:
                    ms = _time(make_run(prm), reps)
                except GvcError:
                    table[label] = None      # infeasible (LDS)
                    continue
                table[label] = ms
                if ms < best_ms:
                    best, best_ms = prm, ms
    if best is None:
        raise GvcError("autotune: no feasible configuration")
    return best, best_ms, table


def _cached(key, base, use_cache):
    if not use_cache:
        return None
    if key in _cache:
        return _cache[key]
    disk = _load_file().get(key)
    if disk and "points_per_walk" in disk:
        best = params_for(base, disk["leaf_size"], disk["points_per_walk"], disk["block_threads"])
        _cache[key] = best
        return best
    return None


def _remember(key, best, best_ms):
    _cache[key] = best
    _save_file({key: dict(best.describe(), ms=best_ms)})


def autotune(means, scales, rots, values, weights, sample_set, base=None, leaf_sizes=LEAF_SIZES,
             points_per_walk=POINTS_PER_WALK, blocks=BLOCKS, reps=2, use_cache=True):
    """Returns (best MixtureParams, {config label: ms}).  ``base`` supplies the
    precision and candidate-log choices, which are not tuned."""
    base = base or mx.default_params()
    P, S = int(means.shape[0]), int(sample_set.S)
    key = _bucket(P, S)
    hit = _cached(key, base, use_cache)
    if hit is not None:
        return hit, {}
    g = torch.Generator(device=means.device)
    g.manual_seed(1234)
    gV = torch.randn(S, device=means.device, generator=g)
    gW = torch.randn(S, device=means.device, generator=g)

    def make_run(prm):
        def run():
            cells, cw, _, st = mx.mixture_forward(means, scales, rots, values, weights, 1.0, sample_set, prm,
                                                  want_conics=False)
            mx.mixture_backward(st, means, scales, rots, values, weights, 1.0, sample_set, cells, cw, gV, gW)
        return run

    best, best_ms, table = _sweep(make_run, base, leaf_sizes, points_per_walk, blocks, reps)
    _remember(key, best, best_ms)
    return best, table


def autotune_batched(gaussians_per_volume, samples_per_volume, means, scales, rots, values, weights, samples,
                     base=None, leaf_sizes=LEAF_SIZES, points_per_walk=POINTS_PER_WALK, blocks=BLOCKS, reps=2,
                     use_cache=True):
    """The same sweep for a batch of volumes (inputs concatenated volume-major);
    one Batch and sample set per leaf size.  Returns (best, table)."""
    from . import batched as bx
    base = base or mx.default_params()
    key = "batch:" + _bucket(int(means.shape[0]), int(samples.shape[0])) + f":{len(gaussians_per_volume)}"
    hit = _cached(key, base, use_cache)
    if hit is not None:
        return hit, {}
    S = int(samples.shape[0])
    g = torch.Generator(device=means.device)
    g.manual_seed(1234)
    gV = torch.randn(S, device=means.device, generator=g)
    gW = torch.randn(S, device=means.device, generator=g)
    sets = {}

    def make_run(prm):
        if prm.leaf_size not in sets:
            b = bx.Batch(gaussians_per_volume, samples_per_volume, leaf_size=prm.leaf_size, device=means.device)
            sets[prm.leaf_size] = (b, bx.BatchedSampleSet(b, samples))
        b, ss = sets[prm.leaf_size]

        def run():
            cells, cw, _, st = bx.mixture_forward_batched(b, means, scales, rots, values, weights, 1.0, ss, prm,
                                                          want_conics=False)
            bx.mixture_backward_batched(st, means, scales, rots, values, weights, 1.0, ss, cells, cw, gV, gW)
        return run

    best, best_ms, table = _sweep(make_run, base, leaf_sizes, points_per_walk, blocks, reps)
    _remember(key, best, best_ms)
    return best, table
