"""Pins the mesh / 3-NN oracle (oracle/mesh_ref.py) on the CPU.

Tet location: by construction (a point drawn inside tet t with barycentrics b
lies in t and interpolates to sum b v -- the construction of R/train.py:34-61),
closed cells (every point of the closed domain, boundary included, is located;
the lowest containing tet wins; shared faces, edges and vertices interpolate an
affine field exactly).
Grid probe: trilinear interpolation reproduces affine functions exactly.
3-NN: against scipy's KD-tree.
"""
import numpy as np
import pytest

from oracle import mesh_ref
from gvc import workloads

f32 = np.float32


def small_mesh(n=3, jitter=0.3, seed=0):
    rng = np.random.default_rng(seed)
    pts = workloads.jittered_lattice(rng, n, jitter)
    tets = workloads.kuhn_tets(n)
    return pts, tets


def test_by_construction_and_values():
    pts, tets = small_mesh(3, 0.3, 1)
    rng = np.random.default_rng(2)
    vals = rng.normal(size=pts.shape[0])
    t = rng.integers(0, tets.shape[0], 300)
    u = np.sort(rng.uniform(0.02, 0.98, (300, 3)), axis=1)
    b = np.stack([u[:, 0], u[:, 1] - u[:, 0], u[:, 2] - u[:, 1], 1 - u[:, 2]], 1)
    keep = b.min(1) > 0.01
    t, b = t[keep], b[keep]
    samples = (b[:, :, None] * pts[tets[t]].astype(np.float64)).sum(1).astype(f32)
    ids, out = mesh_ref.locate(pts, tets, samples, vals)
    assert np.array_equal(ids, t)
    expect = (b * vals[tets[t]]).sum(1)
    np.testing.assert_allclose(out, expect, rtol=1e-5, atol=1e-5)


def test_closed_tets_cover_the_domain():
    pts, tets = small_mesh(2, 0.0, 0)  # exact lattice: many points on shared faces, edges, vertices
    g = np.linspace(0.0, 1.0, 9, dtype=f32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    samples = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    P = pts
    vals = 1.0 + 2.0 * pts[:, 0] - pts[:, 1] + 0.5 * pts[:, 2]
    ids, out = mesh_ref.locate(P, tets, samples, vals)
    # every point of the closed domain is located, boundary points included (the
    # reference's evaluation grid clips samples onto the boundary planes)
    assert np.all(ids >= 0)
    # the lowest containing tet wins, and it really contains the point
    for s_, p in enumerate(samples[::7]):
        holders = [t for t in range(tets.shape[0]) if mesh_ref.tet_contains(P, tets[t], p)[0]]
        assert holders and ids[7 * s_] == holders[0]
    # points on shared faces, edges and vertices interpolate the (continuous) affine field exactly
    s64 = samples.astype(np.float64)
    np.testing.assert_allclose(out, 1.0 + 2.0 * s64[:, 0] - s64[:, 1] + 0.5 * s64[:, 2], rtol=0, atol=1e-12)
    # an interior lattice vertex is held by exactly the tets it is a vertex of
    centre = np.array([0.5, 0.5, 0.5], f32)
    vid = int(np.nonzero(np.all(P == centre, axis=1))[0][0])
    around = int(np.any(tets == vid, axis=1).sum())
    assert around > 6
    assert sum(mesh_ref.tet_contains(P, tets[t], centre)[0] for t in range(tets.shape[0])) == around


def test_outside_points():
    pts, tets = small_mesh(2, 0.2, 3)
    samples = np.array([[-0.1, 0.5, 0.5], [0.5, 1.2, 0.5], [np.nan, 0.5, 0.5]], f32)
    ids, out = mesh_ref.locate(pts, tets, samples, np.zeros(pts.shape[0]))
    assert np.all(ids == -1) and np.all(out == -1.0)


def test_grid_probe_reproduces_affine():
    dims = (5, 7, 4)
    origin = (0.1, -0.2, 0.3)
    spacing = (0.25, 0.125, 0.5)
    ii, jj, kk = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    x = origin[0] + ii * spacing[0]
    y = origin[1] + jj * spacing[1]
    z = origin[2] + kk * spacing[2]
    f = 2.0 * x - 3.0 * y + 0.5 * z + 1.0
    values = np.transpose(f, (2, 1, 0)).ravel()  # x fastest
    rng = np.random.default_rng(4)
    samples = np.stack([rng.uniform(origin[a] - 0.1, origin[a] + (dims[a] - 1) * spacing[a] + 0.1, 2000)
                        for a in range(3)], 1).astype(f32)
    out = mesh_ref.grid_sample(dims, origin, spacing, values, samples)
    s = samples.astype(np.float64)
    inside = np.all([(samples[:, a] >= f32(origin[a])) &
                     (((samples[:, a] - f32(origin[a])) / f32(spacing[a])).astype(f32) <= dims[a] - 1)
                     for a in range(3)], axis=0)
    expect = 2.0 * s[:, 0] - 3.0 * s[:, 1] + 0.5 * s[:, 2] + 1.0
    np.testing.assert_allclose(out[inside], expect[inside], rtol=0, atol=2e-6)
    assert np.all(out[~inside] == -1)
    u8 = mesh_ref.grid_sample(dims, origin, spacing, (values * 10).astype(np.uint8), samples)
    assert np.all(u8[~inside] == 255)


def test_knn3_against_kdtree():
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(5)
    P = rng.uniform(0, 1, (1500, 3)).astype(f32)
    P[10] = P[11]  # a duplicate pair: distance 0
    got = mesh_ref.knn3(P)
    d, _ = cKDTree(P.astype(np.float64)).query(P.astype(np.float64), k=4)
    expect = (d[:, 1:] ** 2).mean(1)
    np.testing.assert_allclose(got, expect, rtol=1e-5, atol=1e-9)
    assert got[10] == got[10] and got[10] < expect.mean()


def test_mixed_cells_reproduce_linear_fields():
    """Isoparametric completeness: every cell type interpolates a linear field
    exactly, wherever the sample is located."""
    p, c, t, o = workloads.mixed_cell_mesh(3, 0)
    assert set(np.unique(t)) == {10, 12, 13, 14}
    rng = np.random.default_rng(1)
    smp = rng.uniform(-0.2, 1.3, (3000, 3)).astype(np.float32)
    a = np.array([0.3, -1.2, 0.7])
    f = p.astype(np.float64) @ a + 0.25
    ids, vals, _ = mesh_ref.locate_cells(p, c, t, o, smp, f)
    ok = ids >= 0
    assert ok.sum() > 500 and (~ok).sum() > 500
    np.testing.assert_allclose(vals[ok], smp[ok].astype(np.float64) @ a + 0.25, atol=1e-12)


def test_mixed_cells_by_construction():
    """Points made inside a known cell (trilinear / wedge / pyramid / barycentric
    maps written out here) are located in that cell."""
    p, c, t, o = workloads.mixed_cell_mesh(3, 4)
    P = p.astype(np.float64)
    rng = np.random.default_rng(2)
    pts, want = [], []
    for cell in range(t.shape[0]):
        k = {10: 4, 12: 8, 13: 6, 14: 5}[int(t[cell])]
        X = P[c[o[cell]:o[cell] + k]]
        for _ in range(3):
            r, s, u = rng.uniform(0.05, 0.95, 3)
            if t[cell] == 10:
                w = rng.dirichlet(np.ones(4)) * 0.8 + 0.05
                x = w @ X
            elif t[cell] == 12:
                x = (X[0] * (1 - r) * (1 - s) * (1 - u) + X[1] * r * (1 - s) * (1 - u) + X[2] * r * s * (1 - u) +
                     X[3] * (1 - r) * s * (1 - u) + X[4] * (1 - r) * (1 - s) * u + X[5] * r * (1 - s) * u +
                     X[6] * r * s * u + X[7] * (1 - r) * s * u)
            elif t[cell] == 13:
                r, s = r * 0.45, s * 0.45
                bot = X[0] * (1 - r - s) + X[1] * r + X[2] * s
                top = X[3] * (1 - r - s) + X[4] * r + X[5] * s
                x = bot * (1 - u) + top * u
            else:
                u = u * 0.9
                base = X[0] * (1 - r) * (1 - s) + X[1] * r * (1 - s) + X[2] * r * s + X[3] * (1 - r) * s
                x = base * (1 - u) + X[4] * u
            pts.append(x)
            want.append(cell)
    ids, _, margin = mesh_ref.locate_cells(p, c, t, o, np.asarray(pts, np.float32))
    assert np.array_equal(ids, np.asarray(want, np.int32))
    assert margin.min() > 1e-4


def test_mixed_oracle_agrees_with_exact_tets():
    """On a tetrahedra-only mesh the parametric oracle and the exact orientation
    oracle pick the same cells away from faces."""
    rng = np.random.default_rng(3)
    n = 4
    pts = workloads.jittered_lattice(rng, n, 0.25)
    tets = workloads.kuhn_tets(n)
    conn = tets.reshape(-1)
    types = np.full(tets.shape[0], 10)
    offs = np.arange(tets.shape[0]) * 4
    smp = rng.uniform(-0.05, 1.05, (1500, 3)).astype(np.float32)
    ids, _, margin = mesh_ref.locate_cells(pts, conn, types, offs, smp)
    ref, _ = mesh_ref.locate(pts, tets, smp, np.zeros(pts.shape[0]))
    clear = (margin > 1e-6) | (ids < 0)
    assert np.array_equal(ids[clear], ref[clear])
