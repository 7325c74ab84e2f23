"""torch-level wrappers of the mesh sampler and the 3-NN initialiser of libgvc."""
import ctypes

import numpy as np
import torch

from ._lib import GvcError, call, lib
from .mixture import _ptr, _stream

_DTYPES = {torch.float64: 0, torch.float32: 1, torch.uint16: 2, torch.uint8: 3}


class TetMeshSampler:
    """Device-resident tetrahedral mesh with its point locator.

    points [N,3] float32, tets [C,4] int (vertex indices), values [N] (optional,
    interpolated as float64).  ``sample(samples)`` returns (values [S] float64 with
    -1 outside, cell_ids [S] int32 with -1 outside) and never leaves the device.
    With ``bind=True`` (default) the values are bound into the locator as per-tet
    interpolants (gvc_tet_locator_bind_values); call ``bind_values`` again after
    changing them.
    """

    def __init__(self, points, tets, values=None, device="cuda", bind=True):
        dev = torch.device(device) if not isinstance(points, torch.Tensor) else points.device
        if dev.type != "cuda":
            raise ValueError("TetMeshSampler needs a GPU device; there is no CPU path")
        self.device = dev
        self.points = torch.as_tensor(points, dtype=torch.float32).to(dev).contiguous()
        if self.points.dim() != 2 or self.points.shape[1] != 3:
            raise ValueError("points must have shape (N, 3)")
        t = torch.as_tensor(tets)
        if t.dim() != 2 or t.shape[1] != 4:
            raise ValueError("tets must have shape (C, 4)")
        N = int(self.points.shape[0])
        if t.numel() and (int(t.min()) < 0 or int(t.max()) >= N):
            raise ValueError("tet vertex index out of range")
        self.tets = t.to(dev, torch.int32).contiguous()
        self.values = None
        if values is not None:
            self.values = torch.as_tensor(values, dtype=torch.float64).to(dev).contiguous().reshape(-1)
            if self.values.numel() != N:
                raise ValueError("values must have one entry per point")
        self.N, self.C = N, int(self.tets.shape[0])
        cap = max(24 * self.C, 1024)
        for _ in range(2):
            nbytes = lib().gvc_tet_locator_bytes(self.N, self.C, cap)
            buf = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            need = ctypes.c_uint64(0)
            st = lib().gvc_tet_locator_build(self.N, _ptr(self.points), self.C, _ptr(self.tets), cap, _ptr(buf),
                                             nbytes, ctypes.byref(need), _stream(dev))
            if st == 5 and need.value > cap:  # GVC_ERR_CAPACITY
                cap = int(need.value)
                continue
            if st != 0:
                from ._lib import check
                check(st, "gvc_tet_locator_build")
            self.locator, self.locator_bytes, self.capacity = buf, nbytes, cap
            self.entries = int(need.value)
            self.bound = False
            if bind and self.values is not None:
                self.bind_values(self.values)
            return
        raise GvcError("tet locator build did not converge")

    def bind_values(self, values):
        """Stores per-tet interpolants of ``values`` [N] (float64) in the locator;
        later samples use them (stream-ordered, no host sync)."""
        v = torch.as_tensor(values, dtype=torch.float64).to(self.device).contiguous().reshape(-1)
        if v.numel() != self.N:
            raise ValueError("values must have one entry per point")
        self.values = v
        call("gvc_tet_locator_bind_values", self.N, _ptr(self.points), self.C, _ptr(self.locator),
             self.locator_bytes, self.capacity, _ptr(v), _stream(self.device))
        self.bound = True

    def sample(self, samples, with_values=True, with_bary=False, sample_set=None):
        """Locate ``samples`` [S,3]; with ``sample_set`` (a gvc.mixture.SampleSet of
        the same points) they are processed in its Morton order."""
        S = int(samples.shape[0]) if sample_set is None else sample_set.S
        ids = torch.empty(S, dtype=torch.int32, device=self.device)
        vals = torch.empty(S, dtype=torch.float64, device=self.device) if (with_values and self.values is not None) \
            else None
        bary = torch.empty((S, 4), dtype=torch.float32, device=self.device) if with_bary else None
        if self.bound and vals is not None:
            s = None
            if sample_set is not None:
                if sample_set.device != self.device:
                    raise ValueError("sample set on another device")
            else:
                s = torch.as_tensor(samples, dtype=torch.float32).to(self.device).contiguous()
                if s.dim() != 2 or s.shape[1] != 3:
                    raise ValueError("samples must have shape (S, 3)")
            call("gvc_tet_locate_bound", self.N, _ptr(self.points), self.C, _ptr(self.locator), self.locator_bytes,
                 self.capacity, S, _ptr(s), _ptr(sample_set.buffer if sample_set is not None else None), _ptr(ids),
                 _ptr(vals), _ptr(bary), _stream(self.device))
        elif sample_set is not None:
            if sample_set.device != self.device:
                raise ValueError("sample set on another device")
            call("gvc_tet_locate_interp_set", self.N, _ptr(self.points), self.C, _ptr(self.locator),
                 self.locator_bytes, self.capacity, _ptr(self.values), _ptr(sample_set.buffer), S, _ptr(ids),
                 _ptr(vals), _ptr(bary), _stream(self.device))
        else:
            s = torch.as_tensor(samples, dtype=torch.float32).to(self.device).contiguous()
            if s.dim() != 2 or s.shape[1] != 3:
                raise ValueError("samples must have shape (S, 3)")
            call("gvc_tet_locate_interp", self.N, _ptr(self.points), self.C, _ptr(self.locator), self.locator_bytes,
                 self.capacity, _ptr(self.values), S, _ptr(s), _ptr(ids), _ptr(vals), _ptr(bary),
                 _stream(self.device))
        if with_bary:
            return vals, ids, bary
        return vals, ids


# Tetrahedral splits used to locate the other 3D VTK cells (local point indices).
# Planar faces make the split exact; a point on a shared face goes to one cell.
_SPLITS = {
    12: np.array([[0, 1, 3, 4], [1, 2, 3, 6], [1, 3, 4, 6], [1, 4, 5, 6], [3, 4, 6, 7]], np.int64),
    13: np.array([[0, 1, 2, 3], [1, 2, 3, 4], [2, 3, 4, 5]], np.int64),
    14: np.array([[0, 1, 2, 4], [0, 2, 3, 4]], np.int64),
}
_NPTS = {1: 1, 3: 2, 5: 3, 9: 4, 10: 4, 12: 8, 13: 6, 14: 5}
VTK_CELL_TYPES = frozenset(_NPTS)


def split_cells(conn, cell_types, cell_offsets):
    """VTK cells -> (tets [T,4] int64, parent cell [T] int64).  Tetrahedra map to
    themselves; hexahedra, wedges and pyramids to tetrahedral splits; vertices,
    lines, triangles and quads (no volume) to nothing.  Raises RuntimeError for a
    type outside the reference's table or a cell with the wrong point count."""
    types = np.asarray(cell_types).astype(np.int64).reshape(-1)
    conn = np.asarray(conn).astype(np.int64).reshape(-1)
    offs = np.asarray(cell_offsets).astype(np.int64).reshape(-1)
    C = types.shape[0]
    bad = types[~np.isin(types, list(_NPTS))]
    if bad.size:
        raise RuntimeError(f"Unsupported VTK cell type: {int(bad[0])}")
    start = offs[:C]
    end = np.append(offs[1:C], conn.shape[0]) if offs.shape[0] <= C else offs[1:C + 1]
    count = end - start
    want = np.vectorize(_NPTS.get, otypes=[np.int64])(types) if C else np.zeros(0, np.int64)
    if C and not np.array_equal(count, want):
        k = int(np.nonzero(count != want)[0][0])
        raise RuntimeError(f"cell {k} of VTK type {int(types[k])} has {int(count[k])} point ids, expected {int(want[k])}")
    tets, parent = [], []
    for t, local in ((10, np.arange(4)[None, :]),) + tuple(_SPLITS.items()):
        cells = np.nonzero(types == t)[0]
        if cells.size == 0:
            continue
        ids = conn[start[cells][:, None, None] + local[None, :, :]]          # [n, k, 4]
        tets.append(ids.reshape(-1, 4))
        parent.append(np.repeat(cells, local.shape[0]))
    if not tets:
        return np.zeros((0, 4), np.int64), np.zeros(0, np.int64)
    tets = np.concatenate(tets)
    parent = np.concatenate(parent)
    order = np.argsort(parent, kind="stable")      # tets in cell order: lower cell ids win ties
    return tets[order], parent[order]


class CellMeshSampler:
    """Device-resident mixed VTK mesh (types 1, 3, 5, 9, 10, 12, 13, 14) with its
    locator: the reference's gpu_sampleu over Viskores Probe (GS/sample_mesh.cu:
    242-414).  A sample is located in the cells' tetrahedral splits; tetrahedra
    interpolate by exact barycentrics, hexahedra / wedges / pyramids by their VTK
    shape functions at the sample's parametric coordinates.  Cells without volume
    never hold a sample.  ``sample(samples)`` -> (values [S] float64 with -1
    outside, cell ids [S] int32 with -1 outside)."""

    def __init__(self, points, conn, cell_types, cell_offsets, values=None, device="cuda"):
        tets, parent = split_cells(conn, cell_types, cell_offsets)
        self.tet_sampler = TetMeshSampler(points, tets, values, device=device)
        dev = self.tet_sampler.device
        self.device = dev
        types = np.asarray(cell_types).astype(np.int64).reshape(-1)
        self.mixed = bool(np.any(np.isin(types, (12, 13, 14))))
        self.parent = torch.as_tensor(parent, dtype=torch.int32).to(dev)
        self.types = torch.as_tensor(types.astype(np.uint8)).to(dev)
        self.offsets = torch.as_tensor(np.asarray(cell_offsets).astype(np.int64).reshape(-1)[:types.shape[0]]
                                       .astype(np.int32)).to(dev)
        self.conn = torch.as_tensor(np.asarray(conn).astype(np.int64).reshape(-1).astype(np.int32)).to(dev)

    def sample(self, samples, with_values=True):
        s = torch.as_tensor(samples, dtype=torch.float32).to(self.device).contiguous()
        vals, tet_ids = self.tet_sampler.sample(s, with_values=with_values)
        ids = torch.where(tet_ids >= 0, self.parent[tet_ids.clamp_min(0).long()], tet_ids).contiguous()
        if with_values and vals is not None and self.mixed:
            t = self.tet_sampler
            call("gvc_cell_interp", int(s.shape[0]), _ptr(s), _ptr(ids), _ptr(self.types), _ptr(self.offsets),
                 _ptr(self.conn), _ptr(t.points), _ptr(t.values), _ptr(vals), _stream(self.device))
        return vals, ids


def sample_mesh_points(points, tets, num_batches, batch_size, generator=None, return_cells=False):
    """R/train.py:34-61: random points inside a tetrahedral mesh -- a uniformly
    chosen tetrahedron, barycentrics (u0, u1 - u0, u2 - u1, 1 - u2) from three
    sorted uniforms, the point their combination of its vertices.  Returns
    [num_batches, batch_size, 3] float32 (and, with return_cells, the chosen
    tetrahedron per point [num_batches, batch_size] int64: the cell that holds
    the point by construction)."""
    pts = torch.as_tensor(points, dtype=torch.float32)
    cells = torch.as_tensor(tets, dtype=torch.long).to(pts.device)
    total = int(num_batches) * int(batch_size)
    idx = torch.randint(0, cells.shape[0], (total,), device=pts.device, generator=generator)
    verts = pts[cells[idx]]
    u = torch.rand(total, 3, device=pts.device, generator=generator).sort(dim=1).values
    bary = torch.stack([u[:, 0], u[:, 1] - u[:, 0], u[:, 2] - u[:, 1], 1.0 - u[:, 2]], dim=1)
    out = (bary.unsqueeze(2) * verts).sum(dim=1).reshape(num_batches, batch_size, 3)
    if return_cells:
        return out, idx.reshape(num_batches, batch_size)
    return out


class BoundarySurface:
    """The boundary triangles of a tetrahedral mesh with outward unit normals,
    for R/train.py:62-92 (sample_exterior_points), which takes them from
    pyvista's extract_surface().triangulate().compute_normals().  Here: a face
    of a tetrahedron is on the boundary iff no other tetrahedron has it (the 4C
    faces sorted by their sorted vertex triples, equal neighbours removed); its
    normal, computed in float64 from the vertices, points away from the owning
    tetrahedron's opposite vertex.  Faces are ordered by their sorted vertex
    triples (pyvista's order differs, so samples agree in distribution, not draw
    for draw).  Everything stays on the points' device."""

    def __init__(self, points, tets):
        pts = torch.as_tensor(points, dtype=torch.float32)
        dev = pts.device
        cells = torch.as_tensor(tets, dtype=torch.long).to(dev)
        if cells.dim() != 2 or cells.shape[1] != 4:
            raise ValueError("tets must have shape (C, 4)")
        N = int(pts.shape[0])
        # face k of a tetrahedron is opposite its vertex k
        opp = torch.tensor([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]], device=dev)
        faces = cells[:, opp].reshape(-1, 3)                       # (4C, 3), row 4t+k
        apex = cells.reshape(-1)                                   # opposite vertex of row 4t+k
        key = faces.sort(dim=1).values
        # lexicographic order of (k0, k1, k2): stable sorts on (k1, k2), then k0
        order = torch.sort(key[:, 1] * N + key[:, 2], stable=True).indices
        order = order[torch.sort(key[order, 0], stable=True).indices]
        ks = key[order]
        same_next = torch.zeros(ks.shape[0], dtype=torch.bool, device=dev)
        if ks.shape[0] > 1:
            eq = (ks[1:] == ks[:-1]).all(dim=1)
            same_next[:-1] |= eq
            same_next[1:] |= eq
        rows = order[~same_next]
        tri = faces[rows]
        a, b, c = (pts[tri[:, i]].double() for i in range(3))
        d = pts[apex[rows]].double()
        n = torch.linalg.cross(b - a, c - a)
        flip = ((n * (d - a)).sum(dim=1) > 0).unsqueeze(1)
        n = torch.where(flip, -n, n)
        n = n / n.norm(dim=1, keepdim=True).clamp_min(1e-300)
        self.points = pts
        self.faces = tri
        self.normals = n.float()

    @property
    def count(self):
        return int(self.faces.shape[0])

    def sample(self, num_batches, batch_size, offset=0.01, generator=None, return_faces=False):
        """Points just outside the surface, as R/train.py:62-92 draws them: a
        uniformly chosen boundary triangle, barycentrics (1 - sqrt(u0),
        sqrt(u0)(1 - u1), sqrt(u0) u1), then a step of U[0, offset) along the
        outward normal.  [num_batches, batch_size, 3] float32 (and, with
        return_faces, the chosen triangle per point)."""
        dev = self.points.device
        total = int(num_batches) * int(batch_size)
        idx = torch.randint(0, self.count, (total,), device=dev, generator=generator)
        v = self.points[self.faces[idx]]
        u = torch.rand(total, 2, device=dev, generator=generator)
        su = u[:, 0].sqrt()
        bary = torch.stack([1 - su, su * (1 - u[:, 1]), su * u[:, 1]], dim=1)
        surf = (bary.unsqueeze(2) * v).sum(dim=1)
        dist = torch.rand(total, 1, device=dev, generator=generator) * offset
        out = (surf + self.normals[idx] * dist).reshape(num_batches, batch_size, 3)
        if return_faces:
            return out, idx.reshape(num_batches, batch_size)
        return out


def sample_exterior_points(points, tets, num_batches, batch_size, offset=0.01, generator=None):
    """R/train.py:62-92 for a tetrahedral mesh (points [N,3], tets [C,4]); see
    BoundarySurface.  Build the surface once and call its sample() to draw
    repeatedly."""
    return BoundarySurface(points, tets).sample(num_batches, batch_size, offset, generator)


def grid_sample(dims, origin, spacing, values, samples):
    """Trilinear probe of a uniform point grid on the GPU (torch tensors in/out)."""
    dev = samples.device
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int64).reshape(3))
    o = np.ascontiguousarray(np.asarray(origin, dtype=np.float32).reshape(3))
    sp = np.ascontiguousarray(np.asarray(spacing, dtype=np.float32).reshape(3))
    if values.dtype not in _DTYPES:
        raise RuntimeError("Unsupported values dtype. Expected float64, float32, uint16, or uint8.")
    values = values.contiguous()
    if values.numel() != int(d.prod()):
        raise RuntimeError("Values array size does not match dims.")
    s = samples.to(torch.float32).contiguous()
    S = int(s.shape[0])
    out = torch.empty(S, dtype=values.dtype, device=dev)
    call("gvc_grid_sample", d.ctypes.data_as(ctypes.c_void_p), o.ctypes.data_as(ctypes.c_void_p),
         sp.ctypes.data_as(ctypes.c_void_p), _DTYPES[values.dtype], _ptr(values), S, _ptr(s), _ptr(out), _stream(dev))
    return out


def knn3_mean_sq(points):
    """distCUDA2: mean squared distance to the 3 nearest other points, [N] float32."""
    p = points
    if p.device.type != "cuda":
        raise ValueError("distCUDA2 needs a GPU tensor")
    p = p.to(torch.float32).contiguous()
    n = int(p.shape[0])
    out = torch.empty(n, dtype=torch.float32, device=p.device)
    nbytes = lib().gvc_knn3_bytes(n)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=p.device)
    call("gvc_knn3", n, _ptr(p), _ptr(out), _ptr(ws), nbytes, _stream(p.device))
    return out
