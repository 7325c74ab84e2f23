"""GPU parity of the mesh sampler and distCUDA2 against the oracle.

Bars: cell ids bit-exact (int32), interpolated values within 1e-9 relative of the
exact rational barycentrics; grid probe bit-exact for unsigned types and within
1e-12 for floating types; distCUDA2 bit-exact (same float operation order).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

from gvc import workloads  # noqa: E402
from gvc.sampling import TetMeshSampler  # noqa: E402
from oracle import mesh_ref  # noqa: E402

f32 = np.float32
DEV = "cuda:0"


def lattice_samples(n, rng, count):
    """Random points plus points exactly on lattice vertices, edges and faces."""
    g = np.arange(n + 1, dtype=np.float64) / n
    special = []
    for _ in range(count // 4):
        i, j, k = rng.integers(0, n + 1, 3)
        a, b = rng.integers(0, n, 2)
        special.append([g[i], g[j], g[k]])                          # vertex
        special.append([g[i], (g[a] + g[a + 1]) / 2, g[k]])        # edge midpoint
        special.append([(g[a] + g[a + 1]) / 2, (g[b] + g[b + 1]) / 2, g[k]])  # face point
        t = rng.uniform()
        special.append([g[a] + t / n, g[a] + t / n, g[k]])          # on the x = y diagonal plane
    rand = rng.uniform(-0.05, 1.05, (count, 3))
    return np.concatenate([rand, np.array(special)]).astype(f32)


def test_tet_locate_subcell_boundaries():
    """Points on and one ulp either side of the locator's sub-cell boundaries.
    An 8^3 Kuhn mesh on [0, 1]^3 (3,072 tets) gets a 15^3 grid (about one cell
    per tet) split 3 x 3 x 3, so the sub-cell planes sit at multiples of 1/45:
    ids must stay the exact predicate's wherever rounding puts the point's
    sub-cell."""
    rng = np.random.default_rng(36)
    pts = workloads.jittered_lattice(rng, 8, 0.3)
    tets = workloads.kuhn_tets(8)
    vals = rng.normal(size=pts.shape[0])
    g = np.arange(46) / 45.0
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(f32)
    X = X[rng.choice(X.shape[0], 1500, replace=False)]
    up = np.nextafter(X, np.float32(np.inf))
    down = np.nextafter(X, np.float32(-np.inf))
    mixed = np.where(rng.uniform(size=X.shape) < 0.5, up, down)
    samples = np.concatenate([X, up, down, mixed]).astype(f32)
    sampler = TetMeshSampler(torch.tensor(pts, device=DEV), tets, vals)
    gv, gid = sampler.sample(torch.tensor(samples, device=DEV))
    ids, out = mesh_ref.locate(pts, tets, samples, vals)
    assert np.array_equal(gid.cpu().numpy(), ids)
    ok = ids >= 0
    assert ok.mean() > 0.9
    np.testing.assert_allclose(gv.cpu().numpy()[ok], out[ok], rtol=1e-9, atol=1e-12)


def test_bound_values_match_vertex_interpolation():
    """Bound per-tet interpolants (gvc_tet_locate_bound) against the per-call
    vertex values path (gvc_tet_locate_interp[_set]) and the oracle: same ids and
    barycentrics, values within 1e-12 relative; rebinding new values follows
    them (an affine map of the field maps the samples' values)."""
    from gvc.mixture import SampleSet
    rng = np.random.default_rng(12)
    pts = workloads.jittered_lattice(rng, 10, 0.3)
    tets = workloads.kuhn_tets(10)
    vals = rng.normal(size=pts.shape[0])
    samples = lattice_samples(10, rng, 4000)
    st = torch.tensor(samples, device=DEV)
    bound = TetMeshSampler(torch.tensor(pts, device=DEV), tets, vals)
    plain = TetMeshSampler(torch.tensor(pts, device=DEV), tets, vals, bind=False)
    assert bound.bound and not plain.bound
    ids, out = mesh_ref.locate(pts, tets, samples, vals)
    ok = ids >= 0
    for ss in (None, SampleSet(st)):
        bv, bid, bb = bound.sample(st, with_bary=True, sample_set=ss)
        pv, pid, pb = plain.sample(st, with_bary=True, sample_set=ss)
        assert np.array_equal(bid.cpu().numpy(), ids) and np.array_equal(pid.cpu().numpy(), ids)
        assert torch.equal(bb, pb)
        b, p = bv.cpu().numpy(), pv.cpu().numpy()
        np.testing.assert_allclose(b[ok], p[ok], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(b[ok], out[ok], rtol=1e-9, atol=1e-12)
        assert np.all(b[~ok] == -1.0)
    bound.bind_values(torch.tensor(2.0 * vals + 1.0, device=DEV))
    b2, _ = bound.sample(st)
    np.testing.assert_allclose(b2.cpu().numpy()[ok], 2.0 * out[ok] + 1.0, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("n,jitter", [(4, 0.0), (6, 0.3)])
def test_tet_locate_bit_exact(n, jitter):
    rng = np.random.default_rng(n)
    pts = workloads.jittered_lattice(rng, n, jitter)
    tets = workloads.kuhn_tets(n)
    vals = rng.normal(size=pts.shape[0])
    samples = lattice_samples(n, rng, 1200)
    sampler = TetMeshSampler(torch.tensor(pts, device=DEV), tets, vals)
    gv, gid = sampler.sample(torch.tensor(samples, device=DEV))
    ids, out = mesh_ref.locate(pts, tets, samples, vals)
    assert np.array_equal(gid.cpu().numpy(), ids)
    ok = ids >= 0
    got = gv.cpu().numpy()
    np.testing.assert_allclose(got[ok], out[ok], rtol=1e-9, atol=1e-12)
    assert np.all(got[~ok] == -1.0)
    # same answers when the samples are processed in a sample set's Morton order
    from gvc.mixture import SampleSet
    st = torch.tensor(samples, device=DEV)
    gv2, gid2 = sampler.sample(st, sample_set=SampleSet(st))
    assert np.array_equal(gid2.cpu().numpy(), ids)
    assert np.array_equal(gv2.cpu().numpy(), got)


def test_gpu_sampleu_dropin():
    from gpu_mesh_sampling import gpu_sampleu
    rng = np.random.default_rng(7)
    n = 5
    pts = workloads.jittered_lattice(rng, n, 0.25)
    tets = workloads.kuhn_tets(n)
    C = tets.shape[0]
    conn = tets.reshape(-1).astype(np.int64)
    types = np.full(C, 10, np.int64)
    offsets = np.arange(C + 1, dtype=np.int64) * 4
    vals = rng.normal(size=pts.shape[0])
    samples = rng.uniform(-0.1, 1.1, (3000, 3)).astype(f32)
    out = gpu_sampleu(pts, conn, types, offsets, vals, samples)
    assert isinstance(out, np.ndarray) and out.dtype == np.float64
    ids, ref = mesh_ref.locate(pts, tets, samples, vals)
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-12)
    # the native module's name (GS/ext.cpp:6, re-exported by GS/gpu_mesh_sampling/__init__.py:1)
    from gpu_mesh_sampling import sample_meshu
    assert np.array_equal(sample_meshu(pts, conn, types, offsets, vals, samples), out)
    with pytest.raises(RuntimeError):
        gpu_sampleu(pts, conn, np.full(C, 7, np.int64), offsets, vals, samples)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.uint16, np.uint8])
def test_grid_sample(dtype):
    from gpu_mesh_sampling import gpu_sample
    rng = np.random.default_rng(8)
    dims = np.array([9, 6, 11], np.int64)
    origin = np.array([0.1, 0.0, -0.3], f32)
    spacing = np.array([0.1, 0.2, 0.05], f32)
    n = int(dims.prod())
    if dtype in (np.uint8, np.uint16):
        values = rng.integers(0, np.iinfo(dtype).max, n).astype(dtype)
    else:
        values = rng.normal(size=n).astype(dtype)
    samples = np.stack([rng.uniform(origin[a] - 0.1, origin[a] + (dims[a] - 1) * spacing[a] + 0.1, 5000)
                        for a in range(3)], 1).astype(f32)
    samples[:8] = origin  # exact grid points and corners
    samples[8] = origin + (dims - 1) * spacing
    out = gpu_sample(dims, origin, spacing, values, samples)
    ref = mesh_ref.grid_sample(dims, origin, spacing, values, samples)
    from gpu_mesh_sampling import sample_mesh  # the native module's name (GS/ext.cpp:5)
    assert np.array_equal(sample_mesh(dims, origin, spacing, values, samples), out)
    assert out.dtype == values.dtype
    if dtype in (np.uint8, np.uint16):
        assert np.array_equal(out, ref)
    else:
        np.testing.assert_allclose(out, ref, rtol=1e-12 if dtype == np.float64 else 1e-6, atol=1e-12)
    with pytest.raises(RuntimeError):
        gpu_sample(dims, origin, spacing, values[:-1], samples)


def test_distcuda2_bit_exact():
    from simple_knn._C import distCUDA2
    rng = np.random.default_rng(9)
    for P in (rng.uniform(0, 1, (5000, 3)), rng.normal(size=(3000, 3)) * [1, 1, 0.001],
              np.concatenate([rng.uniform(0, 1, (500, 3))] * 2)):
        P = P.astype(f32)
        got = distCUDA2(torch.tensor(P, device=DEV)).cpu().numpy()
        ref = mesh_ref.knn3(P)
        assert np.array_equal(got, ref)



def test_distcuda2_few_and_coincident_points():
    """Fewer than four points (a missing neighbour counts +inf, as an unfilled
    3-best list) and coincident points (zero extent on some or every axis)."""
    from simple_knn._C import distCUDA2
    rng = np.random.default_rng(10)
    cases = [rng.uniform(0, 1, (k, 3)) for k in (1, 2, 3, 4)]
    cases += [np.full((5, 3), 0.25), np.concatenate([np.full((4, 3), 0.5), [[0.5, 0.5, 0.75]]])]
    for P in cases:
        P = np.asarray(P, f32)
        got = distCUDA2(torch.tensor(P, device=DEV)).cpu().numpy()
        assert np.array_equal(got, mesh_ref.knn3(P)), P.shape

def test_mixed_cells_against_oracle():
    """Hexahedra, wedges, pyramids and tetrahedra (planar faces): the GPU picks the
    oracle's cell wherever the oracle's choice is clear of a face, interpolates
    within 1e-9 of it, and reproduces a linear field exactly."""
    from gvc.sampling import CellMeshSampler
    p, c, t, o = workloads.mixed_cell_mesh(4, 2)
    rng = np.random.default_rng(11)
    smp = rng.uniform(-0.2, 1.3, (6000, 3)).astype(f32)
    vals = rng.normal(size=p.shape[0])
    sampler = CellMeshSampler(p, c, t, o, vals, device=DEV)
    gv, gid = sampler.sample(torch.tensor(smp, device=DEV))
    gv, gid = gv.cpu().numpy(), gid.cpu().numpy()
    ids, ref, margin = mesh_ref.locate_cells(p, c, t, o, smp, vals)
    clear = (margin > 1e-6) | (ids < 0)
    assert clear.mean() > 0.99
    assert np.array_equal(gid[clear], ids[clear])
    ok = clear & (ids >= 0)
    np.testing.assert_allclose(gv[ok], ref[ok], rtol=1e-9, atol=1e-12)
    assert np.all(gv[gid < 0] == -1.0)
    # linear field: exact up to rounding for every located sample, clear or not
    a = np.array([0.3, -1.2, 0.7])
    lin = CellMeshSampler(p, c, t, o, p.astype(np.float64) @ a + 0.25, device=DEV)
    lv, lid = lin.sample(torch.tensor(smp, device=DEV))
    lv, lid = lv.cpu().numpy(), lid.cpu().numpy()
    inside = lid >= 0
    np.testing.assert_allclose(lv[inside], smp[inside].astype(np.float64) @ a + 0.25, atol=1e-9)
    # the numpy drop-in gives the same values
    from gpu_mesh_sampling import gpu_sampleu
    out = gpu_sampleu(p, c, t, o, vals, smp)
    assert np.array_equal(out, gv)


def test_sample_mesh_points_by_construction():
    """R/train.py:34-61 sampling: each point lies in the tetrahedron that made it;
    the locator agrees wherever the point is not within rounding of that
    tetrahedron's faces, and interpolates its barycentric combination."""
    from gvc.sampling import sample_mesh_points
    rng = np.random.default_rng(12)
    n = 6
    pts = workloads.jittered_lattice(rng, n, 0.3)
    tets = workloads.kuhn_tets(n)
    vals = rng.normal(size=pts.shape[0])
    gen = torch.Generator(device=DEV)
    gen.manual_seed(5)
    P = torch.tensor(pts, device=DEV)
    s, made = sample_mesh_points(P, torch.tensor(tets, device=DEV), 3, 4000, generator=gen, return_cells=True)
    assert s.shape == (3, 4000, 3) and s.dtype == torch.float32
    s = s.reshape(-1, 3)
    made = made.reshape(-1).cpu().numpy()
    sampler = TetMeshSampler(P, tets, vals)
    gv, gid = sampler.sample(s)
    gv, gid = gv.cpu().numpy(), gid.cpu().numpy()
    X = pts.astype(np.float64)[tets[made]]
    x = s.cpu().numpy().astype(np.float64)
    M = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], 2)
    lam = np.linalg.solve(M, (x - X[:, 0])[..., None])[..., 0]
    bary = np.concatenate([1 - lam.sum(1, keepdims=True), lam], 1)
    clear = bary.min(axis=1) > 1e-5
    assert clear.mean() > 0.99
    assert np.array_equal(gid[clear], made[clear].astype(np.int32))
    ref = (bary * vals[tets[made]]).sum(1)
    np.testing.assert_allclose(gv[clear], ref[clear], rtol=1e-9, atol=1e-12)


def test_exterior_points_locate_outside():
    """R/train.py:62-92 sampling on the device: points lifted off the mesh's
    boundary along the outward normals are outside the mesh for the exact
    locator (id -1, value -1); points left on a boundary face are inside it or
    on it."""
    from gvc.sampling import BoundarySurface
    rng = np.random.default_rng(13)
    n = 6
    pts = workloads.jittered_lattice(rng, n, 0.3)
    tets = workloads.kuhn_tets(n)
    vals = rng.normal(size=pts.shape[0])
    P = torch.tensor(pts, device=DEV)
    surf = BoundarySurface(P, torch.tensor(tets, device=DEV))
    assert surf.faces.device.type == "cuda" and surf.count == 6 * n * n * 2
    gen = torch.Generator(device=DEV)
    gen.manual_seed(9)
    s, f = surf.sample(2, 5000, offset=0.01, generator=gen, return_faces=True)
    s = s.reshape(-1, 3)
    f = f.reshape(-1)
    tri = surf.points[surf.faces[f]]
    h = ((s - tri[:, 0]) * surf.normals[f]).sum(1).cpu().numpy()
    gv, gid = TetMeshSampler(P, tets, vals).sample(s)
    gv, gid = gv.cpu().numpy(), gid.cpu().numpy()
    lifted = h > 1e-5
    assert lifted.mean() > 0.99
    assert np.all(gid[lifted] == -1) and np.all(gv[lifted] == -1.0)


def test_boundary_points_are_inside():
    """Points clipped onto the mesh's boundary planes -- what the reference's
    evaluation grid does (np.clip to the box, R/volume_metrics.py:222-226) -- lie
    in closed boundary tets and are located, as Probe's closed cells locate them."""
    from gvc.evaluation import evaluation_grid
    rng = np.random.default_rng(21)
    n = 6
    pts = workloads.jittered_lattice(rng, n, 0.3)
    tets = workloads.kuhn_tets(n)
    vals = 2.0 * pts[:, 0] - pts[:, 1] + 0.25 * pts[:, 2]
    grid, _ = evaluation_grid([0, 0, 0], [1, 1, 1], 24, np.random.RandomState(1))
    g32 = grid.astype(f32)
    on_face = np.any((g32 == 0) | (g32 == 1), axis=1)
    assert on_face.sum() > 1000
    gv, gid = TetMeshSampler(torch.tensor(pts, device=DEV), tets, vals).sample(torch.tensor(g32, device=DEV))
    gid = gid.cpu().numpy()
    assert np.all(gid >= 0)
    ids, ref = mesh_ref.locate(pts, tets, g32[on_face][::5], vals)
    assert np.array_equal(gid[on_face][::5], ids)
    np.testing.assert_allclose(gv.cpu().numpy()[on_face][::5], ref, rtol=1e-9, atol=1e-12)
