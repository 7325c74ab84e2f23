"""CPU oracle for the mesh sampler and the 3-NN initialiser -- TEST INFRASTRUCTURE ONLY.

Tet location (GS/sample_mesh.cu:242-414 semantics; GS/ = R/gpu_sampling/): a point
is in the closed tet (v0,v1,v2,v3) iff every face orientation of the point is zero
or agrees with the tet's orientation (degenerate tets hold nothing); the lowest
containing tet index wins; -1 outside.
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


def tet_contains(P, tet, p):
    """(inside, barycentric float64[4]) for the closed tet: every face orientation
    zero or of the tet's sign; a degenerate tet holds nothing."""
    V = [P[int(i)] for i in tet]
    sT = orient_sign(V[0], V[1], V[2], V[3])
    if sT == 0:
        return False, None
    signs = []
    for i in range(4):
        F = [V[k] for k in range(4) if k != i]
        si = -1 if ((3 - i) & 1) else 1
        sg = orient_sign(F[0], F[1], F[2], p)
        if sg != 0 and si * sg != sT:
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


# ---------------------------------------------------------------------------
# Mixed VTK cells (GS/sample_mesh.cu:300-336 cell table, Viskores Probe
# semantics): a sample is in a cell iff its parametric coordinates lie in the
# cell's parametric domain (tetrahedra: barycentrics >= 0; hexahedra: [0,1]^3;
# wedges: r, s >= 0, r + s <= 1, 0 <= t <= 1; pyramids: [0,1]^3), within a
# tolerance; the lowest containing cell id wins; the value is the cell's
# interpolant sum_k N_k v_k.  Parametric coordinates by Newton's method on the
# isoparametric map (numpy, float64, vectorised over candidate pairs).  This is
# an independent restatement: it decides containment in parametric space, not
# through the tetrahedral splits the GPU locates with.  Pinned by: exact
# reproduction of linear fields (the isoparametric completeness property), by
# cells known by construction, and on pure-tetrahedron meshes by the exact
# orientation oracle above.
# ---------------------------------------------------------------------------
_HEX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                        np.float64)


def _shape(kind, xi):
    """N [m, k] and dN/dxi [m, k, 3] for parametric points xi [m, 3]."""
    r, s, t = xi[:, 0], xi[:, 1], xi[:, 2]
    one = np.ones_like(r)
    if kind == 12:
        # tensor-product linear: N_k = prod_a (c_ka xi_a + (1 - c_ka)(1 - xi_a))
        c = _HEX_CORNERS[None, :, :]
        f = c * xi[:, None, :] + (1 - c) * (1 - xi[:, None, :])          # [m, 8, 3]
        df = 2 * c - 1
        N = f.prod(axis=2)
        dN = np.stack([df[..., a] * np.prod(np.delete(f, a, axis=2), axis=2) for a in range(3)], axis=2)
        return N, dN
    if kind == 13:
        tri = np.stack([1 - r - s, r, s], 1)
        dtri = np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])
        N = np.concatenate([tri * (1 - t)[:, None], tri * t[:, None]], 1)
        dN = np.zeros((xi.shape[0], 6, 3))
        for k in range(3):
            for h, w, dw in ((0, 1 - t, -one), (1, t, one)):
                dN[:, 3 * h + k, 0] = dtri[k, 0] * w
                dN[:, 3 * h + k, 1] = dtri[k, 1] * w
                dN[:, 3 * h + k, 2] = tri[:, k] * dw
        return N, dN
    if kind == 14:
        quad = np.stack([(1 - r) * (1 - s), r * (1 - s), r * s, (1 - r) * s], 1)
        dq_r = np.stack([-(1 - s), 1 - s, s, -s], 1)
        dq_s = np.stack([-(1 - r), -r, r, 1 - r], 1)
        N = np.concatenate([quad * (1 - t)[:, None], t[:, None]], 1)
        dN = np.zeros((xi.shape[0], 5, 3))
        dN[:, :4, 0] = dq_r * (1 - t)[:, None]
        dN[:, :4, 1] = dq_s * (1 - t)[:, None]
        dN[:, :4, 2] = -quad
        dN[:, 4, 2] = 1.0
        return N, dN
    raise ValueError(kind)


def _inside(kind, xi, tol):
    r, s, t = xi[:, 0], xi[:, 1], xi[:, 2]
    if kind == 13:
        return (r >= -tol) & (s >= -tol) & (r + s <= 1 + tol) & (t >= -tol) & (t <= 1 + tol)
    return np.all((xi >= -tol) & (xi <= 1 + tol), axis=1)


def _parametric(kind, X, p, iters=60):
    """Newton on sum_k N_k(xi) X_k = p; X [m, k, 3], p [m, 3] -> xi [m, 3]."""
    xi = np.tile({12: [0.5, 0.5, 0.5], 13: [1 / 3, 1 / 3, 0.5], 14: [0.5, 0.5, 0.2]}[kind], (p.shape[0], 1))
    for _ in range(iters):
        N, dN = _shape(kind, xi)
        F = np.einsum("mk,mka->ma", N, X) - p
        J = np.einsum("mkb,mka->mab", dN, X)
        ok = np.abs(np.linalg.det(J)) > 1e-300
        step = np.zeros_like(xi)
        step[ok] = np.linalg.solve(J[ok], F[ok][..., None])[..., 0]
        xi = xi - step
        if np.max(np.abs(step), initial=0.0) < 1e-15:
            break
    return xi


def locate_cells(pts, conn, cell_types, cell_offsets, samples, values=None, tol=1e-9):
    """(cell ids [S] int32, values [S] float64 or None, parametric margin [S]):
    margin = how far inside the winning cell's domain the sample lies (small =
    near a face, where two cells may both claim it)."""
    P = np.asarray(pts, np.float64)
    conn = np.asarray(conn, np.int64).reshape(-1)
    types = np.asarray(cell_types, np.int64).reshape(-1)
    offs = np.asarray(cell_offsets, np.int64).reshape(-1)[:types.shape[0]]
    smp = np.asarray(samples, np.float64)
    S = smp.shape[0]
    npts = {10: 4, 12: 8, 13: 6, 14: 5}
    best = np.full(S, np.iinfo(np.int64).max)
    val = np.full(S, -1.0)
    margin = np.zeros(S)
    for kind, k in npts.items():
        cells = np.nonzero(types == kind)[0]
        if cells.size == 0:
            continue
        ids = conn[offs[cells][:, None] + np.arange(k)[None, :]]
        X = P[ids]                                             # [c, k, 3]
        lo, hi = X.min(axis=1), X.max(axis=1)
        inbox = np.all((smp[:, None, :] >= lo[None] - 1e-12) & (smp[:, None, :] <= hi[None] + 1e-12), axis=2)
        si, ci = np.nonzero(inbox)
        if si.size == 0:
            continue
        Xp, pp = X[ci], smp[si]
        if kind == 10:
            M = np.stack([Xp[:, 1] - Xp[:, 0], Xp[:, 2] - Xp[:, 0], Xp[:, 3] - Xp[:, 0]], 2)
            lam = np.linalg.solve(M, (pp - Xp[:, 0])[..., None])[..., 0]
            bary = np.concatenate([1 - lam.sum(1, keepdims=True), lam], 1)
            ins = np.all(bary >= -tol, axis=1)
            N = bary
            mg = bary.min(axis=1)
        else:
            xi = _parametric(kind, Xp, pp)
            ins = _inside(kind, xi, tol)
            N, _ = _shape(kind, xi)
            if kind == 13:
                mg = np.minimum.reduce([xi[:, 0], xi[:, 1], 1 - xi[:, 0] - xi[:, 1], xi[:, 2], 1 - xi[:, 2]])
            else:
                mg = np.minimum(xi, 1 - xi).min(axis=1)
        for q in np.nonzero(ins)[0]:
            s_, c_ = si[q], cells[ci[q]]
            if c_ < best[s_]:
                best[s_] = c_
                margin[s_] = mg[q]
                if values is not None:
                    val[s_] = float(np.dot(N[q], np.asarray(values, np.float64)[ids[ci[q]]]))
    out_ids = np.where(best == np.iinfo(np.int64).max, -1, best).astype(np.int32)
    return out_ids, (val if values is not None else None), margin
