"""The reference's native module ``_gpu_mesh_sampling`` (GS/ext.cpp:4-7: pybind
exports ``sample_mesh`` and ``sample_meshu``), on libgvc instead of Viskores.

Same positional arguments and numpy-in/numpy-out behaviour as the pybind
functions (GS/sample_mesh.h); torch tensors are accepted too and give torch
tensors back without leaving the device.  Errors mirror GS/sample_mesh.cu:
RuntimeError for a values size mismatch, an unsupported values dtype or an
unsupported cell type.
"""
import numpy as np
import torch

from gvc.sampling import CellMeshSampler, grid_sample


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("gpu_mesh_sampling needs a ROCm GPU")
    return torch.device("cuda", torch.cuda.current_device())


def sample_mesh(dims, origin, spacing, values, samples):
    """GS/sample_mesh.cu:26-240: trilinear probe of a uniform point grid; -1 outside.
    Output dtype follows ``values`` (float64, float32, uint16, uint8)."""
    as_torch = isinstance(samples, torch.Tensor)
    dev = samples.device if as_torch and samples.is_cuda else _device()
    if isinstance(values, torch.Tensor):
        v = values.reshape(-1).to(dev)
    else:
        va = np.ascontiguousarray(values)
        if va.dtype not in (np.float64, np.float32, np.uint16, np.uint8):
            raise RuntimeError("Unsupported values dtype. Expected float64, float32, uint16, or uint8.")
        v = torch.from_numpy(va.reshape(-1)).to(dev)
    s = samples if as_torch else torch.from_numpy(np.ascontiguousarray(samples, dtype=np.float32))
    out = grid_sample(np.asarray(dims), np.asarray(origin), np.asarray(spacing), v, s.to(dev))
    return out if as_torch else out.cpu().numpy()


def sample_meshu(pts, conn, cell_types, cell_offsets, values, samples):
    """GS/sample_mesh.cu:242-414: probe an unstructured mesh at the samples;
    float64 values, -1 outside."""
    as_torch = isinstance(samples, torch.Tensor)
    dev = samples.device if as_torch and samples.is_cuda else _device()
    p = pts if isinstance(pts, torch.Tensor) else np.ascontiguousarray(pts, dtype=np.float32)
    v = values if isinstance(values, torch.Tensor) else np.ascontiguousarray(values, dtype=np.float64)
    sampler = CellMeshSampler(p, conn, cell_types, cell_offsets, v, device=dev)
    s = samples if as_torch else torch.from_numpy(np.ascontiguousarray(samples, dtype=np.float32))
    out, _ = sampler.sample(s.to(dev))
    return out if as_torch else out.cpu().numpy()
