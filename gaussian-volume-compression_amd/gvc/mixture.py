"""torch-level wrappers of the mixture entry points of libgvc.

Device memory comes from torch (caching allocator, so the calls are capturable
in a torch.cuda.CUDAGraph); every kernel runs on torch's current stream.  No CPU
fallback exists: tensors must be float32 on a ROCm device.
"""
import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import GVC_FLAG_CANDIDATE_LOG, GVC_FLAG_FP64_COND, GVC_FLAG_KAHAN, call, gvc_params, lib

DEFAULT_LEAF_SIZE = 4
DEFAULT_BLOCK_THREADS = 256
DEFAULT_FLAGS = GVC_FLAG_KAHAN


@dataclass(frozen=True)
class MixtureParams:
    """Launch parameters (the autotuner's search space) and precision flags.
    fp64_cond: Gaussians whose scale ratio exceeds fp64_scale_ratio (0 = 1000,
    i.e. cond(Sigma) > 1e6) take the FP64 covariance inverse, and their pairs are
    evaluated from the Cholesky factor of that conic in compensated fp32; those
    above factor_scale_ratio (0 = 10) but below the FP64 threshold get the same
    factor computed in float-float arithmetic (no FP64).
    candidate_log: the forward keeps the candidate chunks it visited and the
    backward replays them; False re-walks the tree in the backward (no
    candidate storage).  chunk_slots: candidate-log slots per 64-point group;
    0 = the module's high-water mark for that number of groups, which grows
    whenever a forward ran out.  lane_mode: "auto" walks a group per lane (each
    lane its own point) where its points are sparse next to its Gaussians;
    "never" / "always" force one way for every group."""
    leaf_size: int = DEFAULT_LEAF_SIZE
    block_threads: int = DEFAULT_BLOCK_THREADS
    kahan: bool = True
    fp64_cond: bool = False
    candidate_log: bool = True
    chunk_slots: int = 0
    run_points: int = 64
    lane_mode: str = "auto"
    fp64_scale_ratio: float = 0.0
    factor_scale_ratio: float = 0.0

    def describe(self):
        """The launch shape the autotuner chooses: BVH leaf size, points walked
        together per wavefront walk (1 = every lane walks the tree for its own
        point, lane_mode "always"; else up to run_points consecutive points share
        one walk) and threads per block."""
        return {"leaf_size": int(self.leaf_size),
                "points_per_walk": 1 if self.lane_mode == "always" else int(self.run_points),
                "block_threads": int(self.block_threads)}

    def c(self, groups=0):
        flags = ((GVC_FLAG_KAHAN if self.kahan else 0) | (GVC_FLAG_FP64_COND if self.fp64_cond else 0) |
                 (GVC_FLAG_CANDIDATE_LOG if self.candidate_log else 0) |
                 {"auto": 0, "never": _lib.GVC_FLAG_LANE_NEVER, "always": _lib.GVC_FLAG_LANE_ALWAYS}[self.lane_mode])
        return gvc_params(int(self.leaf_size), int(self.block_threads), flags, 0,
                          int(self.chunk_slots or chunk_slots_for(groups)), int(self.run_points),
                          float(self.fp64_scale_ratio), float(self.factor_scale_ratio))


DEFAULT_CHUNK_SLOTS = 16
_chunk_slots = {}          # bit length of the group count -> candidate-log slots per group (high-water mark)
_default_params = MixtureParams()
_record_capacity = None   # fixed backward capacity (graph capture); None = read it
_stage_timer = None       # list collecting (stage, start_event, end_event) when set
_audit = None             # list collecting every ForwardState when set


def set_stage_timer(lst):
    """When ``lst`` is a list, every mixture call runs stage by stage and appends
    (stage name, start event, end event); events are recorded on the launch stream."""
    global _stage_timer
    _stage_timer = lst


def set_audit(lst):
    """When ``lst`` is a list, every ForwardState is appended (kept alive) so its
    counters and error flags can be read after a timed region."""
    global _audit
    _audit = lst


def _staged_call(name, args_before, cp, args_after, stages):
    """Runs one mixture entry point, split into its stages when timing."""
    if _stage_timer is None:
        cp.stages = 0
        call(name, *args_before, ctypes.byref(cp), *args_after)
        return
    for mask, label in stages:
        cp.stages = mask
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        call(name, *args_before, ctypes.byref(cp), *args_after)
        e1.record()
        _stage_timer.append((label, e0, e1))
    cp.stages = 0


def set_default_params(params):
    global _default_params
    _default_params = params


def default_params():
    return _default_params


def set_record_capacity(n):
    """Fix the backward's record capacity so no host read is needed (for graph
    capture).  None restores the exact, synchronising count."""
    global _record_capacity
    _record_capacity = None if n is None else int(n)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _as_f32(t, name, shape_tail=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a GPU tensor (got {t.device}); there is no CPU path")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32 (got {t.dtype})")
    if shape_tail is not None and (t.dim() != 2 or t.shape[1] != shape_tail):
        if not (shape_tail == 1 and t.dim() == 1):
            raise ValueError(f"{name} must have shape (P, {shape_tail}), got {tuple(t.shape)}")
    return t.contiguous()


class SampleSet:
    """Sample points in Morton-sorted 64-point groups (replaces the reference's
    static sample BVH, DGR/ext.cu:8-65).  Holds a reference to ``samples``."""

    def __init__(self, samples):
        samples = _as_f32(samples, "samples", 3)
        self.samples = samples
        self.device = samples.device
        self.S = int(samples.shape[0])
        nbytes = lib().gvc_sample_set_bytes(self.S)
        self.buffer = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        call("gvc_sample_set_build", self.S, _ptr(samples), _ptr(self.buffer), nbytes, _stream(self.device))


class ForwardState:
    """Records, per-group counts and the candidate log kept from a forward for its
    backward."""

    def __init__(self, buffer, cparams, params, P, S):
        self.buffer = buffer
        self.params = params
        self.P = P
        self.S = S
        self._c = cparams
        self.chunk_slots_used = 0
        self.lane_groups = 0

    def stats(self):
        """(n_records, n_pairs, error_flags) -- synchronises with the host; also
        sets ``chunk_slots_used`` and ``lane_groups`` (groups walked per lane)."""
        nr = ctypes.c_uint64(0)
        npairs = ctypes.c_uint64(0)
        err = ctypes.c_uint32(0)
        used = ctypes.c_uint64(0)
        lanes = ctypes.c_uint64(0)
        call("gvc_forward_stats", _ptr(self.buffer), self.P, self.S, ctypes.byref(self._c), ctypes.byref(nr),
             ctypes.byref(npairs), ctypes.byref(err), ctypes.byref(used), ctypes.byref(lanes),
             _stream(self.buffer.device))
        self.chunk_slots_used = int(used.value)
        self.lane_groups = int(lanes.value)
        return int(nr.value), int(npairs.value), int(err.value)

    def pair_counts(self, sample_set):
        """Accepted pairs per sample [S] and per Gaussian [P] (int32 device
        tensors; gvc_forward_pair_counts, needs the candidate log).  Debug only."""
        dev = self.buffer.device
        sc = torch.empty(self.S, dtype=torch.int32, device=dev)
        gc = torch.empty(self.P, dtype=torch.int32, device=dev)
        call("gvc_forward_pair_counts", _ptr(self.buffer), self.P, self.S, ctypes.byref(self._c),
             _ptr(sample_set.buffer), _ptr(sc), _ptr(gc), _stream(dev))
        return sc, gc

    def tree(self):
        """The Gaussian BVH of this forward (gvc_forward_tree): dict of device
        tensors order [P] int32 (caller index per BVH position), keys [P] int64
        (the 24-bit Morton keys), nodes [NL-1, 4] int32 (left, right, first leaf,
        last leaf; a child < 0 is leaf ~c), boxes [NL-1, 12] float32 (left child's
        lo.xyz hi.xyz, right child's lo.xyz hi.xyz).  Stream-ordered."""
        dev = self.buffer.device
        NL = (self.P + self._c.leaf_size - 1) // self._c.leaf_size if self.P > 0 else 0
        NI = max(NL - 1, 0)
        order = torch.empty(self.P, dtype=torch.int32, device=dev)
        keys = torch.empty(self.P, dtype=torch.int32, device=dev)
        nodes = torch.empty((NI, 4), dtype=torch.int32, device=dev)
        boxes = torch.empty((NI, 12), dtype=torch.float32, device=dev)
        call("gvc_forward_tree", _ptr(self.buffer), self.P, self.S, ctypes.byref(self._c), _ptr(order), _ptr(keys),
             _ptr(nodes), _ptr(boxes), _stream(dev))
        return dict(order=order, keys=keys.to(torch.int64) & 0xffffffff, nodes=nodes, boxes=boxes)

    def accumulate(self, accum):
        """Stream-ordered: accum (int64 [3] on the device) += (pairs, records),
        accum[2] |= error flags.  No host synchronisation (graph-capturable)."""
        call("gvc_forward_accumulate", _ptr(self.buffer), self.P, self.S, ctypes.byref(self._c), _ptr(accum),
             _stream(self.buffer.device))


def chunk_slots_for(groups):
    """Candidate-log slots per group used for a forward over ``groups`` groups
    when params.chunk_slots is 0."""
    return _chunk_slots.get(max(int(groups), 1).bit_length(), DEFAULT_CHUNK_SLOTS)


def _grow_chunk_slots(fstate, groups):
    """After a forward whose candidate log ran out: a larger slots-per-group
    value (at least double), also kept as the high-water mark for this many
    groups.  The slots counted by an overflowing forward stop where groups gave
    up logging, so the count is a lower bound: hence the doubling."""
    NG = max(1, int(groups))
    need = (fstate.chunk_slots_used + NG - 1) // NG
    new = max(2 * int(fstate._c.chunk_slots), need + need // 4 + 2)
    key = NG.bit_length()
    _chunk_slots[key] = max(_chunk_slots.get(key, DEFAULT_CHUNK_SLOTS), new)
    return new


def _redo_until_log_fits(fstate, groups, rerun, attempts=12):
    """Returns (state, n_records) of a forward whose candidate log held every
    chunk: ``fstate`` itself, or the state of rerun(cparams) with a grown log
    (same inputs, same results)."""
    for _ in range(attempts):
        cap, _, err = fstate.stats()
        if not err & _lib.GVC_ERRFLAG_CHUNK_OVERFLOW:
            return fstate, cap
        slots = _grow_chunk_slots(fstate, groups)
        cp2 = fstate.params.c(groups)
        cp2.chunk_slots = max(int(cp2.chunk_slots), slots)
        fstate = rerun(cp2)
    raise _lib.GvcError("candidate log still too small after growing it")


def mixture_forward(means3D, scales, rotations, values, weights, scale_modifier, sample_set, params=None,
                    want_conics=True):
    """Returns (cells [S], cell_weights [S], conics [6P] or None, ForwardState)."""
    params = params or _default_params
    means3D = _as_f32(means3D, "means3D", 3)
    scales = _as_f32(scales, "scales", 3)
    rotations = _as_f32(rotations, "rotations", 4)
    values = _as_f32(values, "values", 1)
    weights = _as_f32(weights, "weights", 1)
    P = int(means3D.shape[0])
    for name, t in (("scales", scales), ("rotations", rotations), ("values", values), ("weights", weights)):
        if t.shape[0] != P:
            raise ValueError(f"{name} has {t.shape[0]} rows, means3D has {P}")
    dev = means3D.device
    if sample_set.device != dev:
        raise ValueError("sample set and Gaussians are on different devices")
    S = sample_set.S
    cp = params.c((S + 63) // 64)
    return _forward(P, S, dev, cp, params, means3D, scales, rotations, values, weights, scale_modifier, sample_set,
                    want_conics)


def _forward(P, S, dev, cp, params, means3D, scales, rotations, values, weights, scale_modifier, sample_set,
             want_conics, outputs=None):
    nbytes = lib().gvc_forward_state_bytes(P, S, ctypes.byref(cp))
    if nbytes == 0:
        raise _lib.GvcError("invalid mixture parameters")
    state = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    if outputs is None:
        cells = torch.empty(S, dtype=torch.float32, device=dev)
        cw = torch.empty(S, dtype=torch.float32, device=dev)
    else:
        cells, cw = outputs
    conics = torch.empty(6 * P, dtype=torch.float32, device=dev) if want_conics else None
    _staged_call("gvc_mixture_forward",
                 (P, _ptr(means3D), _ptr(scales), _ptr(rotations), _ptr(values), _ptr(weights),
                  float(scale_modifier), _ptr(sample_set.buffer), S), cp,
                 (_ptr(cells), _ptr(cw), _ptr(conics), _ptr(state), nbytes, _stream(dev)),
                 ((_lib.GVC_STAGE_BUILD, "fwd_build"), (_lib.GVC_STAGE_WALK, "fwd_walk"),
                  (_lib.GVC_STAGE_COUNTS, "fwd_counts")))
    fs = ForwardState(state, cp, params, P, S)
    if _audit is not None:
        _audit.append(fs)
    return cells, cw, conics, fs


def mixture_backward(fstate, means3D, scales, rotations, values, weights, scale_modifier, sample_set, out_cells,
                     out_weights, grad_cells, grad_weights, record_capacity=None):
    """Gradients (means3D [P,3], scales [P,3], rotations [P,4], values [P,1], weights [P,1])."""
    P, S = fstate.P, fstate.S
    dev = fstate.buffer.device
    means3D = _as_f32(means3D, "means3D", 3)
    scales = _as_f32(scales, "scales", 3)
    rotations = _as_f32(rotations, "rotations", 4)
    values = _as_f32(values, "values", 1)
    weights = _as_f32(weights, "weights", 1)
    out_cells = _as_f32(out_cells, "out_cells")
    out_weights = _as_f32(out_weights, "out_weights")
    grad_cells = _as_f32(grad_cells, "grad_cells")
    grad_weights = _as_f32(grad_weights, "grad_weights")
    if grad_cells.numel() != S or grad_weights.numel() != S:
        raise ValueError("upstream gradients must have one entry per sample")
    gm = torch.empty((P, 3), dtype=torch.float32, device=dev)
    gs = torch.empty((P, 3), dtype=torch.float32, device=dev)
    gr = torch.empty((P, 4), dtype=torch.float32, device=dev)
    gv = torch.empty((P, 1), dtype=torch.float32, device=dev)
    gw = torch.empty((P, 1), dtype=torch.float32, device=dev)
    if P == 0:
        return gm, gs, gr, gv, gw
    cap = record_capacity if record_capacity is not None else _record_capacity
    if cap is None:
        # a forward whose candidate log was too small is redone (same inputs,
        # same results) into a larger state; its outputs go to scratch tensors
        scratch = []

        def rerun(cp2):
            if not scratch:
                scratch.extend(torch.empty(S, dtype=torch.float32, device=dev) for _ in range(2))
            return _forward(P, S, dev, cp2, fstate.params, means3D, scales, rotations, values, weights,
                            scale_modifier, sample_set, False, outputs=tuple(scratch))[3]

        fstate, cap = _redo_until_log_fits(fstate, (S + 63) // 64, rerun)
    cp = fstate._c
    nbytes = lib().gvc_backward_bytes(P, cap)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    _staged_call("gvc_mixture_backward",
                 (P, _ptr(means3D), _ptr(scales), _ptr(rotations), _ptr(values), _ptr(weights),
                  float(scale_modifier), _ptr(sample_set.buffer), S), cp,
                 (_ptr(out_cells), _ptr(out_weights), _ptr(grad_cells), _ptr(grad_weights), _ptr(fstate.buffer),
                  fstate.buffer.numel(), ctypes.c_uint64(cap), _ptr(ws), nbytes, _ptr(gm), _ptr(gs), _ptr(gr),
                  _ptr(gv), _ptr(gw), _stream(dev)),
                 ((_lib.GVC_STAGE_WALK, "bwd_walk"), (_lib.GVC_STAGE_REDUCE, "bwd_reduce")))
    return gm, gs, gr, gv, gw


def preprocess(means3D, scales, rotations, weights, scale_modifier=1.0, fp64_cond=False):
    """Canonical conic [P,6] and box [P,6] (lo.xyz, hi.xyz)."""
    means3D = _as_f32(means3D, "means3D", 3)
    scales = _as_f32(scales, "scales", 3)
    rotations = _as_f32(rotations, "rotations", 4)
    weights = _as_f32(weights, "weights", 1)
    P = int(means3D.shape[0])
    dev = means3D.device
    conic = torch.empty((P, 6), dtype=torch.float32, device=dev)
    box = torch.empty((P, 6), dtype=torch.float32, device=dev)
    call("gvc_preprocess", P, _ptr(means3D), _ptr(scales), _ptr(rotations), _ptr(weights), float(scale_modifier),
         GVC_FLAG_FP64_COND if fp64_cond else 0, _ptr(conic), _ptr(box), _stream(dev))
    return conic, box


def preprocess_factors(means3D, scales, rotations, weights, params, scale_modifier=1.0):
    """Canonical conic [P,6], box [P,6] and the precise path's factor [P,12]
    (hi[6] lo[6], NaN rows off the path) under ``params`` (fp64_cond and ratios)."""
    means3D = _as_f32(means3D, "means3D", 3)
    scales = _as_f32(scales, "scales", 3)
    rotations = _as_f32(rotations, "rotations", 4)
    weights = _as_f32(weights, "weights", 1)
    P = int(means3D.shape[0])
    dev = means3D.device
    conic = torch.empty((P, 6), dtype=torch.float32, device=dev)
    box = torch.empty((P, 6), dtype=torch.float32, device=dev)
    fac = torch.empty((P, 12), dtype=torch.float32, device=dev)
    cp = params.c()
    call("gvc_preprocess_factors", P, _ptr(means3D), _ptr(scales), _ptr(rotations), _ptr(weights),
         float(scale_modifier), ctypes.byref(cp), _ptr(conic), _ptr(box), _ptr(fac), _stream(dev))
    return conic, box, fac


def morton3d_unit(xyz, clamp=False):
    """63-bit Morton codes (int64 tensor holding the uint64 bit pattern)."""
    xyz = _as_f32(xyz, "xyz", 3)
    n = int(xyz.shape[0])
    out = torch.empty(n, dtype=torch.int64, device=xyz.device)
    call("gvc_morton3d_unit", n, _ptr(xyz), int(bool(clamp)), _ptr(out), _stream(xyz.device))
    return out


def sort_pairs_u64(keys, vals, key_bits=64):
    """Stable sort of (uint64 key as int64 bits, int32 value) pairs."""
    if keys.dtype != torch.int64 or vals.dtype != torch.int32:
        raise TypeError("keys must be int64 (uint64 bits), vals int32")
    n = int(keys.numel())
    dev = keys.device
    ko = torch.empty_like(keys)
    vo = torch.empty_like(vals)
    nbytes = lib().gvc_sort_pairs_u64_bytes(n)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    call("gvc_sort_pairs_u64", n, _ptr(keys.contiguous()), _ptr(vals.contiguous()), _ptr(ko), _ptr(vo), int(key_bits),
         _ptr(ws), nbytes, _stream(dev))
    return ko, vo
