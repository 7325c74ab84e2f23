"""Drop-in for the reference's compiled ``_C`` module (DGR/ext.cu:150-154).

Same three functions and argument lists.  The reference keeps the sample set and
its BVH in C++ statics (DGR/ext.cu:8-10); here ``build_bvh`` builds a
``gvc.mixture.SampleSet`` and keeps it as this module's current sample set, which
``rasterize_gaussians`` and ``rasterize_gaussians_backward`` use unless a
``sample_set`` keyword is given.  ``use_gaussian_bvh`` selects between two
traversal strategies in the reference that compute the same mixture; both map
to the one canonical evaluator here.  ``min_*``, ``max_*`` and ``background``
are accepted and unused, as in the reference kernels.
"""
import time

import torch

from gvc import mixture as _mx

_current = None


def current_sample_set():
    if _current is None:
        raise RuntimeError("build_bvh(samples) must be called before rasterizing")
    return _current


def _set_current(ss):
    global _current
    _current = ss


def build_bvh(samples, debug, use_gaussian_bvh):
    """DGR/ext.cu:28-65.  Sorts the samples into 64-point groups."""
    t0 = time.perf_counter()
    ss = _mx.SampleSet(samples)
    _set_current(ss)
    if debug:
        torch.cuda.synchronize(ss.device)
        print(f"Sample BVH time: {1e3 * (time.perf_counter() - t0):.3f} ms")
    return None


def rasterize_gaussians(means3D, scales, rotations, values, weights, scale_modifier, min_x, min_y, min_z, max_x,
                        max_y, max_z, background, use_gaussian_bvh, debug, sample_set=None, params=None,
                        return_state=False):
    """DGR/ext.cu:67-105 -> (cells [S], cell_weights [S], conics [6P])."""
    ss = sample_set if sample_set is not None else current_sample_set()
    t0 = time.perf_counter()
    cells, cw, conics, state = _mx.mixture_forward(means3D, scales, rotations, values, weights, scale_modifier, ss,
                                                   params=params, want_conics=True)
    if debug:
        torch.cuda.synchronize(ss.device)
        nrec, npairs, err = state.stats()
        print(f"Forward time: {1e3 * (time.perf_counter() - t0):.3f} ms, pairs {npairs}, partial records {nrec}")
        if state.params.candidate_log:
            _print_intersections(state, ss, use_gaussian_bvh)
    if return_state:
        return cells, cw, conics, state
    return cells, cw, conics


def _print_intersections(state, ss, use_gaussian_bvh):
    """The reference's debug statistics of accepted pairs (CR/rasterizer_impl.cu:
    175-216): per sample with use_gaussian_bvh, else per Gaussian -- the maximum,
    its first index and the mean."""
    per_sample, per_gaussian = state.pair_counts(ss)
    counts, what = (per_sample, "samples") if use_gaussian_bvh else (per_gaussian, "gaussians")
    n = counts.numel()
    if n == 0:
        print(f"Intersections: max=0 ({what[:-1]} -1), avg=0.000 over 0 {what}")
        return
    mx = int(counts.max())
    idx = int(torch.argmax(counts)) if mx > 0 else -1
    avg = float(counts.to(torch.float64).sum()) / n
    print(f"Intersections: max={mx} ({what[:-1]} {idx}), avg={avg:.3f} over {n} {what}")


def rasterize_gaussians_backward(means3D, scales, rotations, conics, values, weights, out_cells, out_weights,
                                 scale_modifier, min_x, min_y, min_z, max_x, max_y, max_z, background,
                                 dL_dout_cells, dL_dout_cell_weights, use_gaussian_bvh, debug, sample_set=None,
                                 forward_state=None, params=None):
    """DGR/ext.cu:107-147 -> (dmeans3D [P,3], dscales [P,3], drotations [P,4],
    dvalues [P,1], dweights [P,1]).  Without a forward_state the Gaussian BVH and
    the per-group counts are rebuilt (a forward walk) before the adjoint walk."""
    ss = sample_set if sample_set is not None else current_sample_set()
    t0 = time.perf_counter()
    if forward_state is None:
        _, _, _, forward_state = _mx.mixture_forward(means3D, scales, rotations, values, weights, scale_modifier, ss,
                                                     params=params, want_conics=False)
    grads = _mx.mixture_backward(forward_state, means3D, scales, rotations, values, weights, scale_modifier, ss,
                                 out_cells, out_weights, dL_dout_cells, dL_dout_cell_weights)
    if debug:
        torch.cuda.synchronize(ss.device)
        print(f"Backward time: {1e3 * (time.perf_counter() - t0):.3f} ms")
    return grads
