"""ctypes binding of libgvc.so (the C ABI in include/gvc.h).

There is no fallback: if the library is missing or was built for another GPU,
every entry point raises.  Build it with ``python -m gvc._build`` (or
``__graft_entry__.build()``).
"""
import ctypes
import os

from ._build import LIB_PATH

GVC_ERR_INVALID_ARGUMENT = 1
GVC_ERR_WORKSPACE_TOO_SMALL = 2
GVC_ERR_HIP = 3
GVC_ERR_UNSUPPORTED = 4
GVC_ERR_CAPACITY = 5
GVC_METRIC_SUMS = 11
GVC_FLAG_KAHAN = 1
GVC_FLAG_FP64_COND = 2
GVC_FLAG_CANDIDATE_LOG = 4
GVC_ERRFLAG_STACK_OVERFLOW = 1
GVC_ERRFLAG_RECORD_OVERFLOW = 2
GVC_ERRFLAG_CHUNK_OVERFLOW = 4
GVC_FLAG_LANE_NEVER = 8
GVC_FLAG_LANE_ALWAYS = 16
GVC_STAGE_BUILD = 1
GVC_STAGE_WALK = 2
GVC_STAGE_COUNTS = 4
GVC_STAGE_REDUCE = 1


class GvcError(RuntimeError):
    pass


class gvc_batch(ctypes.Structure):
    _fields_ = [("volumes", ctypes.c_int), ("leaf_size", ctypes.c_int), ("P_total", ctypes.c_int),
                ("S_total", ctypes.c_int), ("groups", ctypes.c_int), ("leaves", ctypes.c_int),
                ("internal_nodes", ctypes.c_int), ("table", ctypes.c_void_p)]


class gvc_params(ctypes.Structure):
    _fields_ = [("leaf_size", ctypes.c_int), ("block_threads", ctypes.c_int),
                ("flags", ctypes.c_uint), ("stages", ctypes.c_uint), ("chunk_slots", ctypes.c_uint),
                ("run_points", ctypes.c_uint), ("fp64_scale_ratio", ctypes.c_float),
                ("factor_scale_ratio", ctypes.c_float)]


GVC_ADAM_MAX_TENSORS = 8


class gvc_train_clock(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_int64), ("step", ctypes.c_int64),
                ("lr", ctypes.c_double * GVC_ADAM_MAX_TENSORS), ("step_size", ctypes.c_float * GVC_ADAM_MAX_TENSORS),
                ("bc2_sqrt", ctypes.c_float), ("reserved", ctypes.c_float * 7)]


class gvc_lr_schedule(ctypes.Structure):
    _fields_ = [("lr_init", ctypes.c_double), ("lr_final", ctypes.c_double), ("delay_mult", ctypes.c_double),
                ("delay_steps", ctypes.c_int64), ("max_steps", ctypes.c_int64)]


class gvc_adam_tensor(ctypes.Structure):
    _fields_ = [("param", ctypes.c_void_p), ("grad", ctypes.c_void_p), ("exp_avg", ctypes.c_void_p),
                ("exp_avg_sq", ctypes.c_void_p), ("numel", ctypes.c_uint64), ("lr_index", ctypes.c_int)]


class gvc_store(ctypes.Structure):
    _fields_ = [("capacity", ctypes.c_int), ("xyz", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("scaling", ctypes.c_void_p), ("rotation", ctypes.c_void_p), ("value", ctypes.c_void_p),
                ("exp_avg", ctypes.c_void_p * 5), ("exp_avg_sq", ctypes.c_void_p * 5)]


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_U = ctypes.c_uint
_F = ctypes.c_float
_SZ = ctypes.c_size_t
_U64 = ctypes.c_uint64
_PP = ctypes.POINTER(gvc_params)
_PB = ctypes.POINTER(gvc_batch)

# name: (restype, argtypes)
_SIGS = {
    "gvc_version": (_I, []),
    "gvc_status_string": (ctypes.c_char_p, [_I]),
    "gvc_last_error": (ctypes.c_char_p, []),
    "gvc_preprocess": (_I, [_I, _P, _P, _P, _P, _F, _U, _P, _P, _P]),
    "gvc_preprocess_factors": (_I, [_I, _P, _P, _P, _P, _F, _PP, _P, _P, _P, _P]),
    "gvc_morton3d_unit": (_I, [_I, _P, _I, _P, _P]),
    "gvc_sort_pairs_u64_bytes": (_SZ, [_I]),
    "gvc_sort_pairs_u64": (_I, [_I, _P, _P, _P, _P, _I, _P, _SZ, _P]),
    "gvc_sample_set_bytes": (_SZ, [_I]),
    "gvc_sample_set_build": (_I, [_I, _P, _P, _SZ, _P]),
    "gvc_forward_state_bytes": (_SZ, [_I, _I, _PP]),
    "gvc_mixture_forward": (_I, [_I, _P, _P, _P, _P, _P, _F, _P, _I, _PP, _P, _P, _P, _P, _SZ, _P]),
    "gvc_forward_stats": (_I, [_P, _I, _I, _PP, ctypes.POINTER(_U64), ctypes.POINTER(_U64),
                               ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(_U64), ctypes.POINTER(_U64), _P]),
    "gvc_forward_tree": (_I, [_P, _I, _I, _PP, _P, _P, _P, _P, _P]),
    "gvc_forward_pair_counts": (_I, [_P, _I, _I, _PP, _P, _P, _P, _P]),
    "gvc_backward_bytes": (_SZ, [_I, _U64]),
    "gvc_forward_accumulate": (_I, [_P, _I, _I, _PP, _P, _P]),
    "gvc_mixture_backward": (_I, [_I, _P, _P, _P, _P, _P, _F, _P, _I, _PP, _P, _P, _P, _P, _P, _SZ, _U64,
                                  _P, _SZ, _P, _P, _P, _P, _P, _P]),
    "gvc_tet_locator_bytes": (_SZ, [_I, _I, _U64]),
    "gvc_tet_locator_build": (_I, [_I, _P, _I, _P, _U64, _P, _SZ, ctypes.POINTER(_U64), _P]),
    "gvc_tet_locate_interp": (_I, [_I, _P, _I, _P, _SZ, _U64, _P, _I, _P, _P, _P, _P, _P]),
    "gvc_tet_locate_interp_set": (_I, [_I, _P, _I, _P, _SZ, _U64, _P, _P, _I, _P, _P, _P, _P]),
    "gvc_tet_locator_bind_values": (_I, [_I, _P, _I, _P, _SZ, _U64, _P, _P]),
    "gvc_tet_locate_bound": (_I, [_I, _P, _I, _P, _SZ, _U64, _I, _P, _P, _P, _P, _P, _P]),
    "gvc_grid_sample": (_I, [_P, _P, _P, _I, _P, _I, _P, _P, _P]),
    "gvc_knn3_bytes": (_SZ, [_I]),
    "gvc_knn3": (_I, [_I, _P, _P, _P, _SZ, _P]),
    "gvc_cell_interp": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gvc_batch_table_bytes": (_SZ, [_I]),
    "gvc_batch_describe": (_I, [_I, _P, _P, _I, _P, _SZ, _PB, _P]),
    "gvc_sample_set_bytes_batched": (_SZ, [_PB]),
    "gvc_sample_set_build_batched": (_I, [_PB, _P, _P, _SZ, _P]),
    "gvc_forward_state_bytes_batched": (_SZ, [_PB, _PP]),
    "gvc_mixture_forward_batched": (_I, [_PB, _P, _P, _P, _P, _P, _F, _P, _PP, _P, _P, _P, _P, _SZ, _P]),
    "gvc_forward_stats_batched": (_I, [_PB, _P, _PP, ctypes.POINTER(_U64), ctypes.POINTER(_U64),
                                       ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(_U64), ctypes.POINTER(_U64),
                                       _P]),
    "gvc_forward_accumulate_batched": (_I, [_PB, _P, _PP, _P, _P]),
    "gvc_forward_pair_counts_batched": (_I, [_PB, _P, _PP, _P, _P, _P, _P]),
    "gvc_mixture_backward_batched": (_I, [_PB, _P, _P, _F, _P, _PP, _P, _P, _P, _P, _P, _SZ, _U64, _P, _SZ,
                                          _P, _P, _P, _P, _P, _P]),
    "gvc_activate_forward": (_I, [_I, _P, _P, _P, _P, _F, _P, _P, _P, _P, _P]),
    "gvc_activate_backward": (_I, [_I, _P, _P, _P, _P, _F, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gvc_loss_state_bytes": (_SZ, [_I]),
    "gvc_loss_forward": (_I, [_I, _P, _P, _P, _F, _F, _I, _P, _P, _SZ, _P]),
    "gvc_loss_backward": (_I, [_I, _P, _P, _P, _F, _F, _I, _P, _P, _SZ, _P, _P, _P]),
    "gvc_lr_advance": (_I, [_P, ctypes.POINTER(gvc_lr_schedule), _I, ctypes.c_int64, _P]),
    "gvc_adam_step": (_I, [_I, ctypes.POINTER(gvc_adam_tensor), _P, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, _P]),
    "gvc_error_order_bytes": (_SZ, [_I]),
    "gvc_error_order": (_I, [_I, _P, _P, _I, _P, _F, _P, _P, _P, _SZ, _P]),
    "gvc_densify_bytes": (_SZ, [_I]),
    "gvc_densify_in_empty": (_I, [_I, _I, ctypes.POINTER(gvc_store), _P, _P, _P, _F, _F, _P, _SZ, _P]),
    "gvc_prune_bytes": (_SZ, [_I]),
    "gvc_prune": (_I, [_I, ctypes.POINTER(gvc_store), _P, _F, _P, _P, _SZ, _P]),
    "gvc_volume_metrics_bytes": (_SZ, [_I]),
    "gvc_volume_metrics": (_I, [_I, _P, _P, _P, _P, _SZ, _P]),
    "gvc_volume_metrics_noalloc": (_I, [_I, _P, _P, _P, _P]),
}

EXPORTED = tuple(_SIGS.keys())


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GvcError(f"libgvc.so not built ({LIB_PATH}); run __graft_entry__.build() "
                           "or python -m gvc._build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, what=""):
    if status != 0:
        L = lib()
        msg = L.gvc_status_string(status).decode()
        detail = L.gvc_last_error().decode()
        raise GvcError(f"{what}: {msg} ({detail})" if what else f"{msg} ({detail})")


def call(name, *args):
    f = getattr(lib(), name)
    st = f(*args)
    check(st, name)
