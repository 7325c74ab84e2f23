"""CPU oracle for the mesh sampler and the 3-NN initialiser -- TEST INFRASTRUCTURE ONLY.

Tet location (GS/sample_mesh.cu:242-414 semantics; GS/ = R/gpu_sampling/): a point
is in tet (v0,v1,v2,v3) iff every face orientation of the point agrees with the
tet's orientation, with exact zeros resolved by simulation of simplicity (the point
perturbed to p + (e, e^2, e^3)); the lowest containing tet index wins; -1 outside.
Orientation signs are exact: Python ``fractions.Fraction`` over the float32 inputs
(filtered by a float64 evaluation with a generous error bound).  This is an
independent implementation of the predicate the GPU evaluates with double
expansions.

Uniform-grid probe (GS/sample_mesh.cu:26-240 semantics): t_a = (p_a - o_a) / s_a in
float32, valid iff 0 <= t_a <= n_a - 1, cell min(floor(t), n-2), trilinear weights
in float64; -1 outside (all-ones bits for unsigned types).

3-NN: brute force over all pairs, float32 squared distances in the order
(dx*dx + dy*dy) + dz*dz, mean ((d1 + d2) + d3) / 3 in float32.
"""
from fractions import Fraction

import numpy as np


def _orient_exact(a, b, c, d):
    A = [Fraction(float(x)) for x in a]
    B = [Fraction(float(x)) for x in b]
    Cc = [Fraction(float(x)) for x in c]
    D = [Fraction(float(x)) for x in d]
    ad = [A[i] - D[i] for i in range(3)]
    bd = [B[i] - D[i] for i in range(3)]
    cd = [Cc[i] - D[i] for i in range(3)]
    det = (ad[0] * (bd[1] * cd[2] - bd[2] * cd[1]) - ad[1] * (bd[0] * cd[2] - bd[2] * cd[0]) +
           ad[2] * (bd[0] * cd[1] - bd[1] * cd[0]))
    return det


def _sign(x):
    return (x > 0) - (x < 0)


def orient_sign(a, b, c, d):
    """Exact sign of det[a-d; b-d; c-d] for float32 coordinates."""
    a64, b64, c64, d64 = (np.asarray(v, dtype=np.float64) for v in (a, b, c, d))
    m = np.stack([a64 - d64, b64 - d64, c64 - d64])
    det = np.linalg.det(m)
    bound = 1e-12 * np.prod(np.abs(m).sum(axis=1)) + 1e-300
    if abs(det) > bound:
        return int(np.sign(det))
    return _sign(_orient_exact(a, b, c, d))


def _normal_exact(f0, f1, f2):
    F0 = [Fraction(float(x)) for x in f0]
    F1 = [Fraction(float(x)) for x in f1]
    F2 = [Fraction(float(x)) for x in f2]
    u = [F1[i] - F0[i] for i in range(3)]
    v = [F2[i] - F0[i] for i in range(3)]
    return [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]


def tet_contains(P, tet, p):
    """(inside, barycentric float64[4]) under the simulation-of-simplicity rule."""
    V = [P[int(i)] for i in tet]
    sT = orient_sign(V[0], V[1], V[2], V[3])
    if sT == 0:
        return False, None
    signs = []
    for i in range(4):
        F = [V[k] for k in range(4) if k != i]
        si = -1 if ((3 - i) & 1) else 1
        sg = orient_sign(F[0], F[1], F[2], p)
        if sg == 0:
            n = _normal_exact(*F)
            for comp in n:
                if comp != 0:
                    sg = -_sign(comp)
                    break
        if si * sg != sT:
            return False, None
        signs.append(si)
    vol = _orient_exact(V[0], V[1], V[2], V[3])
    bary = []
    for i in range(4):
        F = [V[k] for k in range(4) if k != i]
        bary.append(float(signs[i] * _orient_exact(F[0], F[1], F[2], p) / vol))
    return True, np.array(bary)


def locate(points, tets, samples, values=None):
    """Brute force over all tets (bbox prefilter).  Returns (cell_ids, values)."""
    P = np.asarray(points, dtype=np.float32)
    T = np.asarray(tets, dtype=np.int64)
    S = np.asarray(samples, dtype=np.float32)
    lo = P[T].min(axis=1)
    hi = P[T].max(axis=1)
    ids = np.full(S.shape[0], -1, np.int64)
    out = np.full(S.shape[0], -1.0)
    for s in range(S.shape[0]):
        p = S[s]
        cand = np.nonzero(np.all((lo <= p) & (p <= hi), axis=1))[0]
        for t in cand:  # ascending: the first hit is the lowest index
            ok, bary = tet_contains(P, T[t], p)
            if ok:
                ids[s] = t
                if values is not None:
                    vv = np.asarray(values, dtype=np.float64)[T[t]]
                    out[s] = float(np.dot(bary, vv))
                break
    return ids, out


def grid_sample(dims, origin, spacing, values, samples):
    dims = np.asarray(dims, dtype=np.int64)
    o = np.asarray(origin, dtype=np.float32)
    sp = np.asarray(spacing, dtype=np.float32)
    vals = np.asarray(values)
    S = np.asarray(samples, dtype=np.float32)
    t = ((S - o) / sp).astype(np.float32)
    ok = np.all((t >= 0) & (t <= (dims - 1).astype(np.float32)), axis=1)
    i = np.floor(np.where(np.isfinite(t), t, 0)).astype(np.int64)
    i = np.clip(i, 0, np.maximum(dims - 2, 0))
    f = (t - i.astype(np.float32)).astype(np.float64)
    f[:, dims == 1] = 0.0
    flat = vals.reshape(-1).astype(np.float64)
    acc = np.zeros(S.shape[0])
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = ((f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) *
                     (f[:, 2] if dz else 1 - f[:, 2]))
                xi = np.minimum(i[:, 0] + dx, dims[0] - 1)
                yi = np.minimum(i[:, 1] + dy, dims[1] - 1)
                zi = np.minimum(i[:, 2] + dz, dims[2] - 1)
                idx = xi + dims[0] * (yi + dims[1] * zi)
                acc += np.where(w != 0, w * flat[np.clip(idx, 0, flat.size - 1)], 0.0)
    if vals.dtype in (np.uint8, np.uint16):
        out = np.trunc(acc).astype(vals.dtype)
        out[~ok] = np.iinfo(vals.dtype).max
    else:
        out = acc.astype(vals.dtype)
        out[~ok] = -1
    return out


def knn3(points):
    P = np.asarray(points, dtype=np.float32)
    n = P.shape[0]
    out = np.zeros(n, np.float32)
    for i in range(n):
        d = P - P[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        d2[i] = np.inf
        b = np.sort(d2)[:3].astype(np.float32)
        while b.size < 3:
            b = np.append(b, np.float32(np.inf))
        out[i] = np.float32((b[0] + b[1]) + b[2]) / np.float32(3.0)
    return out
