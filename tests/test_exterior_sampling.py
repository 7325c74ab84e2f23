"""Exterior samples (R/train.py:62-92) from the boundary of a tetrahedral mesh.

Pinned by: the boundary of a known lattice (count, faces on the cube's faces,
axis-aligned outward normals); two closed-surface identities that catch a
missing, doubled or flipped face (sum of area * normal = 0, and sum of
area * (normal . centroid) / 3 = the mesh volume, the divergence theorem); and
each sample's construction (on its triangle's outward side, within the offset,
its projection inside the triangle)."""
import numpy as np
import pytest
import torch

from gvc import workloads
from gvc.sampling import BoundarySurface, sample_exterior_points


def mesh(n, jitter, seed=0):
    rng = np.random.default_rng(seed)
    return workloads.jittered_lattice(rng, n, jitter), workloads.kuhn_tets(n)


def signed_volumes(pts, tets):
    p = pts.astype(np.float64)[tets]
    return np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0]) / 6


def tet_volume(pts, tets, n):
    """Volume of the domain: each tetrahedron's volume signed by its orientation
    in the unjittered lattice (the jitter inverts a few, which then overlap
    their neighbours; the signed sum still equals the volume the boundary
    encloses)."""
    ref = signed_volumes(workloads.jittered_lattice(np.random.default_rng(0), n, 0.0), tets)
    return (signed_volumes(pts, tets) * np.sign(ref)).sum()


@pytest.mark.parametrize("n,jitter", [(4, 0.0), (6, 0.3)])
def test_boundary_of_lattice(n, jitter):
    pts, tets = mesh(n, jitter)
    surf = BoundarySurface(torch.as_tensor(pts), torch.as_tensor(tets))
    assert surf.count == 6 * n * n * 2                  # two triangles per boundary square
    tri = pts[surf.faces.numpy()].astype(np.float64)    # (F, 3, 3)
    nrm = surf.normals.numpy().astype(np.float64)
    # every boundary triangle lies in one face of the unit cube, normal pointing out of it
    for f in range(surf.count):
        on = [(a, v) for a in range(3) for v in (0.0, 1.0) if np.all(tri[f, :, a] == v)]
        assert len(on) == 1
        a, v = on[0]
        expect = np.zeros(3)
        expect[a] = 1.0 if v == 1.0 else -1.0
        assert np.allclose(nrm[f], expect, atol=1e-6)
    # closed-surface identities
    cr = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    assert np.abs((area[:, None] * nrm).sum(0)).max() < 1e-9 * max(1.0, area.sum())
    flux = (area * (nrm * tri.mean(1)).sum(1)).sum() / 3.0
    assert abs(flux - tet_volume(pts, tets, n)) < 1e-9


def test_closed_surface_identities_irregular_mesh():
    # a mesh whose boundary is not a cube: keep the lattice's tetrahedra whose
    # centroid lies inside a ball (a staircase of faces in every direction)
    n = 8
    pts, tets = mesh(n, 0.25, seed=3)
    c = pts[tets].mean(1)
    keep = np.linalg.norm(c - 0.5, axis=1) < 0.42
    tets = tets[keep]
    surf = BoundarySurface(torch.as_tensor(pts), torch.as_tensor(tets))
    tri = pts[surf.faces.numpy()].astype(np.float64)
    nrm = surf.normals.numpy().astype(np.float64)
    cr = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    assert surf.count > 100
    assert np.abs((area[:, None] * nrm).sum(0)).max() < 1e-6
    flux = (area * (nrm * tri.mean(1)).sum(1)).sum() / 3.0
    assert abs(flux - tet_volume(pts, tets, n)) < 1e-9
    # the normal of each boundary face points away from its own tetrahedron
    assert np.all(np.abs(np.linalg.norm(nrm, axis=1) - 1) < 1e-6)


def test_samples_by_construction():
    pts, tets = mesh(5, 0.3, seed=1)
    surf = BoundarySurface(torch.as_tensor(pts), torch.as_tensor(tets))
    g = torch.Generator().manual_seed(7)
    off = 0.01
    out, fidx = surf.sample(3, 4000, offset=off, generator=g, return_faces=True)
    assert out.shape == (3, 4000, 3) and out.dtype == torch.float32
    x = out.reshape(-1, 3).numpy().astype(np.float64)
    f = fidx.reshape(-1).numpy()
    tri = pts[surf.faces.numpy()[f]].astype(np.float64)
    nrm = surf.normals.numpy()[f].astype(np.float64)
    h = ((x - tri[:, 0]) * nrm).sum(1)                  # height above the triangle's plane
    assert h.min() >= -1e-6 and h.max() <= off + 1e-6
    assert h.mean() == pytest.approx(off / 2, rel=0.05)  # U[0, offset)
    foot = x - h[:, None] * nrm
    # barycentrics of the foot point in its triangle
    e1, e2, r = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0], foot - tri[:, 0]
    d11, d12, d22 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
    r1, r2 = (r * e1).sum(1), (r * e2).sum(1)
    den = d11 * d22 - d12 * d12
    b1 = (d22 * r1 - d12 * r2) / den
    b2 = (d11 * r2 - d12 * r1) / den
    assert min(b1.min(), b2.min(), (1 - b1 - b2).min()) > -1e-5
    # the domain is the unit cube (boundary vertices stay on its faces): a sample
    # lifted off its face lies outside
    lifted = h > 1e-5
    outside = np.any((x < 0.0) | (x > 1.0), axis=1)
    assert np.all(outside[lifted])
    # same generator seed, same draws
    again = sample_exterior_points(torch.as_tensor(pts), torch.as_tensor(tets), 3, 4000, offset=off,
                                   generator=torch.Generator().manual_seed(7))
    assert torch.equal(again, out)
