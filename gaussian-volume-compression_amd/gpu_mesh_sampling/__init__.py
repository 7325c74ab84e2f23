"""Drop-in for the reference's ``gpu_mesh_sampling`` (GS/gpu_mesh_sampling/__init__.py:1-20),
running on libgvc instead of Viskores.

``gpu_sample`` and ``gpu_sampleu``, and the native module's ``sample_mesh`` and
``sample_meshu`` (``_gpu_mesh_sampling``), keep their signatures and numpy-in/numpy-out
behaviour (torch tensors are accepted too and give torch tensors back, without
leaving the device).  Errors mirror GS/sample_mesh.cu: RuntimeError for a values
size mismatch, an unsupported values dtype or an unsupported cell type.  Nothing
is printed.  ``TetMeshSampler`` is the device-resident form for a training loop:
it keeps the locator and also returns int32 cell ids.

Cell types: the reference's table (GS/sample_mesh.cu:300-336).  Tetrahedra (VTK
10) are located exactly and interpolated by barycentrics; hexahedra (12), wedges
(13) and pyramids (14) are located through tetrahedral splits (exact for planar
faces) and interpolated by their VTK shape functions; vertices, lines, triangles
and quads (1, 3, 5, 9) have no volume and hold no sample.  Any other type raises.
"""
from gvc.sampling import CellMeshSampler, TetMeshSampler, grid_sample  # noqa: F401

from ._gpu_mesh_sampling import sample_mesh, sample_meshu  # noqa: F401  (GS/gpu_mesh_sampling/__init__.py:1)


def gpu_sample(dims, origin, spacing, values, samples):
    """GS/gpu_mesh_sampling/__init__.py:4-11: ``sample_mesh`` of the native module."""
    return sample_mesh(dims, origin, spacing, values, samples)


def gpu_sampleu(pts, conn, cell_types, cell_offsets, values, samples):
    """GS/gpu_mesh_sampling/__init__.py:13-21: ``sample_meshu`` of the native module."""
    return sample_meshu(pts, conn, cell_types, cell_offsets, values, samples)
