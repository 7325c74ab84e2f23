"""CPU oracle for the Gaussian-mixture hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the timed CPU
baseline).  The product path (``gaussian-volume-compression_amd/``) never imports
it and fails loudly when its HIP library is missing.

Layout:
  gvc_oracle.c     C restatement (canonical fp32 predicate, FP64 contributions,
                   Morton codes, Karras tree, plain fp32 CPU baseline); built by
                   ``Makefile`` into ``_build/liboracle.so``.
  autograd_ref.py  torch float64 restatement of the forward formula, used for
                   autograd gradients and frozen-predicate central differences.
  mesh_ref.py      exact tet location (rational orient3d + simulation of
                   simplicity), trilinear grid sampling, brute-force 3-NN.
  train_ref.py     the reference's training-loop procedures on torch tensors
                   (learning-rate schedule, densify/prune surgery).
  eval_ref.py      the reference's evaluation grid construction and metrics.

The seeded synthetic inputs of the benchmark configurations are product code
(``gvc/workloads.py``), shared by the tests and ``bench.py``.

The reference itself cannot be built here (it needs nvcc, cuBQL, glm and
Viskores), so there is no ``_ref/`` build.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    """Compile the C oracle (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        F = ctypes.c_float
        L.orc_log.restype = ctypes.c_double
        L.orc_log.argtypes = [ctypes.c_double]
        L.orc_box_radius.restype = F
        L.orc_box_radius.argtypes = [F]
        L.orc_preprocess.restype = None
        L.orc_preprocess.argtypes = [I, P, P, P, P, F, I, F, F, P, P, P, P]
        L.orc_conic_factor_ff.restype = I
        L.orc_conic_factor_ff.argtypes = [P, P, F, P, P]
        L.orc_forward.restype = None
        L.orc_forward.argtypes = [I, I, P, P, P, P, P, P, P, P, P, P, P]
        L.orc_forward_binned.restype = I
        L.orc_forward_binned.argtypes = [I, I, P, P, P, P, P, P, P, P, I, ctypes.c_longlong, P, P, P]
        L.orc_precise_power.restype = F
        L.orc_precise_power.argtypes = [P, P, P]
        L.orc_conic_factor.restype = I
        L.orc_conic_factor.argtypes = [P, P, P]
        L.orc_conic_fp64_n.restype = None
        L.orc_conic_fp64_n.argtypes = [I, P, P, ctypes.c_float, P]
        L.orc_backward.restype = None
        L.orc_backward.argtypes = [I, I, P, P, P, P, P, F, P, P, P, P, P, P, P, P, P, P, P, P, P, P]
        L.orc_cov_chain.restype = None
        L.orc_cov_chain.argtypes = [P, P, F, P, P, P, P]
        L.orc_conic_fp64.restype = None
        L.orc_conic_fp64.argtypes = [P, P, F, P]
        L.orc_morton3d_unit.restype = None
        L.orc_morton3d_unit.argtypes = [I, P, I, P]
        L.orc_karras_build.restype = None
        L.orc_karras_build.argtypes = [I, P, P, P, P, P]
        L.orc_cpu_plain_fwdbwd.restype = ctypes.c_longlong
        L.orc_cpu_plain_fwdbwd.argtypes = [I, I, P, P, P, P, P, P, P, P, P, P, P]
        L.orc_pair_accept_ext.restype = I
        L.orc_pair_accept_ext.argtypes = [P, P, P, P, P, P]
        L.orc_pair_mask.restype = None
        L.orc_pair_mask.argtypes = [I, I, P, P, P, P, P, P, P]
        L.orc_num_threads.restype = I
        L.orc_num_threads.argtypes = []
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def log(x):
    return lib().orc_log(float(x))


def box_radius(w):
    return np.float32(lib().orc_box_radius(float(np.float32(w))))


def preprocess(means, scales, rots, weights, scale_modifier=1.0, fp64_cond=False, fp64_ratio=0.0,
               factor=False, factor_ratio=0.0):
    """Canonical conic (P,6) and box (P,6: lo.xyz hi.xyz) -- CR/forward.cu:10-114 --
    and the FP64-path mask (scale ratio above fp64_ratio, 0 = 1000).  factor=True
    also returns the precise-path factors (P,12: hi[6] lo[6] of the Cholesky factor
    of the FP64 conic, or below the FP64 threshold and above factor_ratio (0 = 10)
    the float-float factor; NaN rows off the path), for forward/backward/pair_mask."""
    means, scales, rots, weights = _f32(means), _f32(scales), _f32(rots), _f32(weights).reshape(-1)
    P = means.shape[0]
    conic = np.zeros((P, 6), np.float32)
    box = np.zeros((P, 6), np.float32)
    ill = np.zeros(P, np.int32)
    fac = np.zeros((P, 12), np.float32) if factor else None
    lib().orc_preprocess(P, _p(means), _p(scales), _p(rots), _p(weights), float(scale_modifier),
                         int(bool(fp64_cond)), float(fp64_ratio), float(factor_ratio), _p(conic), _p(box), _p(ill),
                         _p(fac))
    if factor:
        return conic, box, ill.astype(bool), fac
    return conic, box, ill.astype(bool)


def precise_power(fac, mean, x):
    """fp32 power of one pair on the precise path (fac: 12 floats)."""
    return float(lib().orc_precise_power(_p(_f32(fac)), _p(_f32(mean)), _p(_f32(x))))


def conic_factor(conic6_d):
    """(ok, hi[6], lo[6]) Cholesky factor of one FP64 conic."""
    hi = np.zeros(6, np.float32); lo = np.zeros(6, np.float32)
    ok = lib().orc_conic_factor(_p(_f64(conic6_d)), _p(hi), _p(lo))
    return bool(ok), hi, lo


def conic_factor_ff(scales3, q4, scale_modifier=1.0):
    """(ok, hi[6], lo[6]) float-float Cholesky factor of one Gaussian's conic."""
    hi = np.zeros(6, np.float32); lo = np.zeros(6, np.float32)
    ok = lib().orc_conic_factor_ff(_p(_f32(scales3)), _p(_f32(q4)), float(scale_modifier), _p(hi), _p(lo))
    return bool(ok), hi, lo


def conic_fp64(scales, rots, scale_modifier=1.0):
    scales, rots = _f32(scales), _f32(rots)
    scales, rots = _f32(scales), _f32(rots)
    P = scales.shape[0]
    out = np.zeros((P, 6), np.float64)
    lib().orc_conic_fp64_n(P, _p(scales), _p(rots), float(scale_modifier), _p(out))
    return out


def forward(means, values, weights, conic, box, samples, conic_d=None, binned=None, factor=None):
    """FP64 forward with the canonical predicate.  Returns V, W, counts.  Brute
    force, or (binned=True, default for large problems) the same sums with the
    Gaussians binned in a uniform grid -- identical results, every accepted pair
    visited in ascending Gaussian order."""
    means, values, weights = _f32(means), _f32(values).reshape(-1), _f32(weights).reshape(-1)
    conic, box, samples = _f32(conic), _f32(box), _f32(samples)
    cd = _f64(conic_d) if conic_d is not None else None
    fac = _f32(factor) if factor is not None else None
    P, S = means.shape[0], samples.shape[0]
    V = np.zeros(S, np.float64)
    W = np.zeros(S, np.float64)
    cnt = np.zeros(S, np.int32)
    if binned is None:
        binned = P * S > 200_000_000
    if binned and P > 0:
        ext = (box[:, 3:] - box[:, :3]).astype(np.float64)
        fin = np.all(np.isfinite(box), axis=1)
        span = np.nanmax(np.where(fin[:, None], box[:, 3:], -np.inf), axis=0) - \
            np.nanmin(np.where(fin[:, None], box[:, :3], np.inf), axis=0) if fin.any() else np.ones(3)
        width = np.median(ext[fin], axis=0) if fin.any() else np.ones(3)
        G = int(np.clip(np.min(np.asarray(span) / np.maximum(width, 1e-12)), 1, 256))
        rc = lib().orc_forward_binned(P, S, _p(means), _p(values), _p(weights), _p(conic), _p(box), _p(cd),
                                      _p(fac), _p(samples), G, 1_500_000_000, _p(V), _p(W), _p(cnt))
        if rc == 0:
            return V, W, cnt
    lib().orc_forward(P, S, _p(means), _p(values), _p(weights), _p(conic), _p(box), _p(cd), _p(fac),
                      _p(samples), _p(V), _p(W), _p(cnt))
    return V, W, cnt


def backward(means, scales, rots, values, weights, conic, box, samples, V, W, gV, gW,
             scale_modifier=1.0, factor=None):
    """Brute-force FP64 backward (frozen canonical predicate -- the precise path's
    where factor is given --, FP64 conic)."""
    fac = _f32(factor) if factor is not None else None
    means, scales, rots = _f32(means), _f32(scales), _f32(rots)
    values, weights = _f32(values).reshape(-1), _f32(weights).reshape(-1)
    conic, box, samples = _f32(conic), _f32(box), _f32(samples)
    V, W, gV, gW = _f64(V), _f64(W), _f64(gV), _f64(gW)
    P, S = means.shape[0], samples.shape[0]
    dm = np.zeros((P, 3)); ds = np.zeros((P, 3)); dr = np.zeros((P, 4))
    dv = np.zeros(P); dw = np.zeros(P); dc = np.zeros((P, 6))
    lib().orc_backward(P, S, _p(means), _p(scales), _p(rots), _p(values), _p(weights),
                       float(scale_modifier), _p(conic), _p(box), _p(fac), _p(samples), _p(V), _p(W),
                       _p(gV), _p(gW), _p(dm), _p(ds), _p(dr), _p(dv), _p(dw), _p(dc))
    return dict(means=dm, scales=ds, rotations=dr, values=dv, weights=dw, conic=dc)


def cov_chain(scales, rots, conic6_d, gconic6, scale_modifier=1.0):
    scales, rots = _f32(scales), _f32(rots)
    C, G = _f64(conic6_d), _f64(gconic6)
    P = scales.shape[0]
    ds = np.zeros((P, 3)); dq = np.zeros((P, 4))
    L = lib()
    for i in range(P):
        L.orc_cov_chain(_p(scales[i]), _p(rots[i]), float(scale_modifier), _p(C[i]), _p(G[i]),
                        _p(ds[i]), _p(dq[i]))
    return ds, dq


def morton3d_unit(xyz, clamp=False):
    xyz = _f32(xyz)
    n = xyz.shape[0]
    out = np.zeros(n, np.uint64)
    lib().orc_morton3d_unit(n, _p(xyz), int(bool(clamp)), _p(out))
    return out


def karras_build(sorted_keys):
    k = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    n = k.shape[0]
    m = max(n - 1, 0)
    left = np.zeros(m, np.int32); right = np.zeros(m, np.int32)
    first = np.zeros(m, np.int32); last = np.zeros(m, np.int32)
    if n >= 2:
        lib().orc_karras_build(n, _p(k), _p(left), _p(right), _p(first), _p(last))
    return left, right, first, last


def cpu_plain_fwdbwd(means, values, weights, conic, box, samples, gV, gW):
    means, values, weights = _f32(means), _f32(values).reshape(-1), _f32(weights).reshape(-1)
    conic, box, samples, gV, gW = _f32(conic), _f32(box), _f32(samples), _f32(gV), _f32(gW)
    P, S = means.shape[0], samples.shape[0]
    V = np.zeros(S, np.float32); W = np.zeros(S, np.float32)
    g = np.zeros((P, 11), np.float32)
    pairs = lib().orc_cpu_plain_fwdbwd(P, S, _p(means), _p(values), _p(weights), _p(conic),
                                       _p(box), _p(samples), _p(gV), _p(gW), _p(V), _p(W), _p(g))
    return V, W, g, int(pairs)


def pair_mask(means, conic, box, samples, factor=None):
    """Dense canonical membership mask (P, S) bool and the fp32 powers."""
    means, conic, box, samples = _f32(means), _f32(conic), _f32(box), _f32(samples)
    fac = _f32(factor) if factor is not None else None
    P, S = means.shape[0], samples.shape[0]
    mask = np.zeros((P, S), np.uint8)
    power = np.zeros((P, S), np.float32)
    lib().orc_pair_mask(P, S, _p(means), _p(conic), _p(box), _p(fac), _p(samples), _p(mask), _p(power))
    return mask.astype(bool), power


def num_threads():
    return lib().orc_num_threads()
