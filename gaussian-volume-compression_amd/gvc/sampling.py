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
    """

    def __init__(self, points, tets, values=None, device="cuda"):
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
        cap = max(16 * self.C, 1024)
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
            return
        raise GvcError("tet locator build did not converge")

    def sample(self, samples, with_values=True, with_bary=False, sample_set=None):
        """Locate ``samples`` [S,3]; with ``sample_set`` (a gvc.mixture.SampleSet of
        the same points) they are processed in its Morton order."""
        S = int(samples.shape[0]) if sample_set is None else sample_set.S
        ids = torch.empty(S, dtype=torch.int32, device=self.device)
        vals = torch.empty(S, dtype=torch.float64, device=self.device) if (with_values and self.values is not None) \
            else None
        bary = torch.empty((S, 4), dtype=torch.float32, device=self.device) if with_bary else None
        if sample_set is not None:
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
