"""The C ABI library loads and exports every entry point include/gvc.h declares.

Runs without a GPU: only symbol lookup and the pure host-side queries (status
strings, workspace sizes, argument validation) are called.
"""
import ctypes
import os
import re

import pytest

import conftest

HEADER = os.path.join(conftest.ROOT, "include", "gvc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gvc_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    from gvc import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libgvc.so is not built (run __graft_entry__.build())")
    return _lib.lib()


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("gvc_preprocess", "gvc_sample_set_build", "gvc_mixture_forward", "gvc_mixture_backward",
              "gvc_forward_stats", "gvc_morton3d_unit", "gvc_sort_pairs_u64"):
        assert n in names


def test_every_declared_symbol_is_exported(L):
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_dropin_package_names():
    """The names the reference's three Python packages export (DGR/bvh_diff_gaussian_
    rasterization/__init__.py, GS/gpu_mesh_sampling/__init__.py:1-21, GS/ext.cpp:4-7,
    simple-knn's _C.distCUDA2) import without a GPU."""
    import bvh_diff_gaussian_rasterization as dgr
    import gpu_mesh_sampling as gms
    from gpu_mesh_sampling import _gpu_mesh_sampling
    from simple_knn._C import distCUDA2  # noqa: F401
    for n in ("GaussianRasterizationSettings", "GaussianRasterizer", "rasterize_gaussians"):
        assert hasattr(dgr, n)
    for n in ("gpu_sample", "gpu_sampleu", "sample_mesh", "sample_meshu"):
        assert callable(getattr(gms, n))
    assert gms.sample_mesh is _gpu_mesh_sampling.sample_mesh
    assert gms.sample_meshu is _gpu_mesh_sampling.sample_meshu


def test_every_declared_symbol_has_a_binding():
    from gvc import _lib
    missing = [n for n in declared_functions() if n not in _lib.EXPORTED]
    assert not missing, missing


def test_host_side_queries(L):
    from gvc import _lib
    assert L.gvc_version() >= 1
    assert L.gvc_status_string(0) == b"ok"
    assert L.gvc_status_string(1) == b"invalid argument"
    assert L.gvc_sample_set_bytes(1000) > 1000 * 16
    p = _lib.gvc_params(4, 256, 1, 0, 0, 0)
    assert L.gvc_forward_state_bytes(1000, 1000, ctypes.byref(p)) > 1000 * 80
    bad = _lib.gvc_params(0, 256, 1, 0, 0, 0)
    assert L.gvc_forward_state_bytes(1000, 1000, ctypes.byref(bad)) == 0
    assert L.gvc_backward_bytes(1000, 5000) > 5000 * 48


def test_invalid_arguments_rejected_without_gpu(L):
    # argument validation happens before any HIP call
    st = L.gvc_preprocess(-1, None, None, None, None, 1.0, 0, None, None, None)
    assert st == 1
    assert b"P < 0" in L.gvc_last_error()
    st = L.gvc_sample_set_build(10, None, None, 0, None)
    assert st == 1


def test_training_structs_and_validation(L):
    from gvc import _lib
    # the C layouts: 8 + 8 + 8*8 + 8*4 + 4 + 7*4 bytes; pointers after the int capacity
    assert ctypes.sizeof(_lib.gvc_train_clock) == 144
    assert ctypes.sizeof(_lib.gvc_lr_schedule) == 40
    assert _lib.gvc_store.xyz.offset == 8 and ctypes.sizeof(_lib.gvc_store) == 8 + 15 * 8
    assert L.gvc_error_order_bytes(100000) > 100000 * 12
    assert L.gvc_prune_bytes(1000) > 1000 * 12 * 3 * 4
    assert L.gvc_densify_bytes(1000) > 1000 * 4
    many = (_lib.gvc_adam_tensor * 9)()
    assert L.gvc_adam_step(9, many, None, 0.9, 0.999, 1e-15, None) == 1
    assert b"n_tensors" in L.gvc_last_error()
    one = (_lib.gvc_adam_tensor * 1)()
    assert L.gvc_adam_step(1, one, ctypes.c_void_p(8), 1.0, 0.999, 1e-15, None) == 1
    assert b"betas" in L.gvc_last_error()
    s = _lib.gvc_store()
    s.capacity = 10
    s.xyz = s.weight = s.scaling = s.rotation = s.value = 256
    assert L.gvc_densify_in_empty(8, 3, ctypes.byref(s), None, None, None, 0.1, 0.01, None, 0, None) == 1
    assert b"capacity" in L.gvc_last_error()
    s.exp_avg[0] = 256          # moments must be all present or all absent
    assert L.gvc_prune(5, ctypes.byref(s), None, 0.005, ctypes.c_void_p(8), ctypes.c_void_p(8), 0, None) == 1
    assert b"gvc_store" in L.gvc_last_error()
    sch = _lib.gvc_lr_schedule(1e-4, 1e-6, 1.0, 0, 0)
    assert L.gvc_lr_advance(ctypes.c_void_p(8), ctypes.byref(sch), 0, -1, None) == 1
    assert b"max_steps" in L.gvc_last_error()
