"""The plain GPU sequence: naive per-volume HIP kernels mirroring the reference's
sample-BVH evaluation (one wavefront per Gaussian, float atomics, a separate
normalise, no compensation, no batching).  Benchmark baseline only: bench.py
times it beside the fused batched path.  Not part of the library.
"""
import ctypes
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "gaussian-volume-compression_amd", "csrc")
LIB_PATH = os.path.join(HERE, "libgvcplain.so")

_lib = None


def build(verbose=False):
    from gvc import _build
    tmp = LIB_PATH + ".tmp"
    cmd = [_build.hipcc(), *_build.HIPCC_FLAGS, "-I" + CSRC, "-x", "hip", os.path.join(HERE, "plain.hip"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgvcplain.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        sigs = {
            "plain_leaf_slots": [I],
            "plain_preprocess": [I, P, P, P, P, F, P, P, P],
            "plain_codes": [I, P, P, P, P],
            "plain_build": [I, P, P, P, P, P, P],
            "plain_forward": [I, P, P, P, P, P, I, P, P, P, P, P, P, P],
            "plain_backward": [I, P, P, P, P, P, F, P, P, I, P, P, P, P, P, P, P, P, P, P, P, P, P],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = I
        _lib = L
    return _lib
