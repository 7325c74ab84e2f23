"""The trainable Gaussian model of the training loop, in fixed-capacity buffers.

Mirrors the parts of the reference's ``GaussianModel`` that the training loop
drives (R/scene/gaussian_model.py): the initial cloud (``create_from_pcd``,
:139-205), the optimiser (``training_setup`` / ``update_learning_rate``,
:207-255), the activated parameters (``get_*``, :114-132), densification and
pruning with the Adam state surgery (:396-612, :727-799) and PLY save/load
(:265-394).  The reference rebuilds every parameter and both Adam moments with
``torch.cat`` / boolean indexing at each densification; here they live in buffers
of a fixed capacity, rows ``[0, count)`` live, and the surgery runs as HIP kernels
in place (libgvc ``gvc_densify_in_empty``, ``gvc_prune``); the optimiser is one
fused HIP launch (``gvc_adam_step``) whose step count and learning rates sit in a
device clock, so a training step can be captured in a graph and replayed.

Parameter tensors (``_xyz``, ``_weight``, ``_scaling``, ``_rotation``,
``_values``) are leaf views of the buffers, recreated when the count changes (the
reference likewise hands out new ``nn.Parameter`` objects after each surgery).
"""
import ctypes
import dataclasses

import numpy as np
import torch

from . import model_io
from ._lib import call, gvc_adam_tensor, gvc_lr_schedule, gvc_store, gvc_train_clock, lib
from .mixture import _ptr, _stream
from .training import REFERENCE_MAX_SCALE, activate

GROUPS = ("xyz", "weight", "scaling", "rotation", "value")   # the reference optimiser's groups, in order
WIDTHS = (3, 1, 3, 4, 1)
ADAM_BETAS = (0.9, 0.999)                                   # torch.optim.Adam defaults
ADAM_EPS = 1e-15                                            # R/scene/gaussian_model.py:240


@dataclasses.dataclass
class OptimizationParams:
    """The training arguments the loop reads (R/arguments/__init__.py:82-96)."""
    iterations: int = 16000
    position_lr_init: float = 0.00016
    position_lr_final: float = 0.0000016
    position_lr_delay_mult: float = 0.01
    position_lr_max_steps: int = 30_000
    values_lr: float = 0.0025
    weight_lr: float = 0.025
    scaling_lr: float = 0.001
    rotation_lr: float = 0.0001
    densification_interval: int = 100
    densify_from_iter: int = 500
    densify_until_iter: int = 16_000
    densify_grad_threshold: float = 0.000002


class GaussianModel:
    """Fixed-capacity model: ``capacity`` rows of raw parameters (and, after
    ``training_setup``, of Adam moments) on one GPU."""

    def __init__(self, capacity, device="cuda:0", max_scale=REFERENCE_MAX_SCALE):
        device = torch.device(device)
        if device.type != "cuda":
            raise ValueError(f"GaussianModel needs a GPU device (got {device}); there is no CPU path")
        if capacity <= 0:
            raise ValueError("capacity must be positive")
        self.capacity = int(capacity)
        self.device = device
        self.max_scale = max_scale
        self.count = 0
        self._buf = {g: torch.zeros((self.capacity, w), dtype=torch.float32, device=device)
                     for g, w in zip(GROUPS, WIDTHS)}
        self._m = None
        self._v = None
        self._clock = None
        self._schedule = None
        self._count_dev = torch.zeros(1, dtype=torch.int32, device=device)
        self._views()

    # ---- parameters ---------------------------------------------------------------

    def _views(self):
        n = self.count
        self._params = {}
        for g in GROUPS:
            t = self._buf[g][:n]
            t.requires_grad_(True)
            self._params[g] = t
        self._xyz = self._params["xyz"]
        self._weight = self._params["weight"]
        self._scaling = self._params["scaling"]
        self._rotation = self._params["rotation"]
        self._values = self._params["value"]

    def parameters(self):
        return [self._params[g] for g in GROUPS]

    def activated(self):
        """(scales, rotations, weights, values), one fused kernel per direction."""
        return activate(self._scaling, self._rotation, self._weight, self._values, self.max_scale)

    @property
    def get_xyz(self):
        return self._xyz

    @property
    def get_scaling(self):
        return self.activated()[0]

    @property
    def get_rotation(self):
        return self.activated()[1]

    @property
    def get_weight(self):
        return self.activated()[2]

    @property
    def get_values(self):
        return self.activated()[3]

    def _store(self):
        s = gvc_store()
        s.capacity = self.capacity
        s.xyz, s.weight, s.scaling, s.rotation, s.value = (self._buf[g].data_ptr() for g in GROUPS)
        for k, g in enumerate(GROUPS):
            s.exp_avg[k] = self._m[g].data_ptr() if self._m is not None else None
            s.exp_avg_sq[k] = self._v[g].data_ptr() if self._v is not None else None
        return s

    def _check_rows(self, extra):
        if self.count + extra > self.capacity:
            raise RuntimeError(f"model capacity {self.capacity} exceeded ({self.count} + {extra} rows)")

    # ---- construction ---------------------------------------------------------------

    def create_from_pcd(self, points, values):
        """Initial cloud (R/scene/gaussian_model.py:139-205): one Gaussian per point,
        log-scale log(sqrt(max(knn3, 1e-7))), identity rotation, weight 0.1, value
        logit of the point's value."""
        self.count = 0
        self._append(points, values, None, dist_scale=1.0, init_weight=0.1)

    def load_raw(self, means, weights, scales, rotations, values):
        """Rows from raw parameter tensors (as ``load_ply`` gives them)."""
        n = int(means.shape[0])
        if n > self.capacity:
            raise RuntimeError(f"model capacity {self.capacity} exceeded ({n} rows)")
        with torch.no_grad():
            for g, t in zip(GROUPS, (means, weights, scales, rotations, values)):
                self._buf[g][:n].copy_(torch.as_tensor(t, dtype=torch.float32).reshape(n, -1))
                if self._m is not None:
                    self._m[g][:n].zero_()
                    self._v[g][:n].zero_()
        self.count = n
        self._views()

    def load_ply(self, path):
        p = model_io.read_ply(path)
        self.load_raw(*(torch.as_tensor(a) for a in (p.means, p.weights, p.scales, p.rotations, p.values)))

    def save_ply(self, path):
        n = self.count
        cpu = {g: self._buf[g][:n].detach().cpu().numpy() for g in GROUPS}
        model_io.write_ply(path, model_io.GaussianParams(means=cpu["xyz"], values=cpu["value"],
                                                         weights=cpu["weight"], scales=cpu["scaling"],
                                                         rotations=cpu["rotation"]))

    # ---- optimiser ---------------------------------------------------------------------

    def training_setup(self, args):
        """Adam moments (zero) and the device clock: learning rates per group as
        R/scene/gaussian_model.py:212-247 sets them, the position rate on the
        exponential schedule."""
        self._m = {g: torch.zeros_like(self._buf[g]) for g in GROUPS}
        self._v = {g: torch.zeros_like(self._buf[g]) for g in GROUPS}
        c = gvc_train_clock()
        lrs = (args.position_lr_init, args.weight_lr, args.scaling_lr, args.rotation_lr, args.values_lr)
        for k, lr in enumerate(lrs):
            c.lr[k] = lr
        raw = np.frombuffer(bytes(c), dtype=np.uint8).copy()
        self._clock = torch.as_tensor(raw).to(self.device)
        self._schedule = gvc_lr_schedule(args.position_lr_init, args.position_lr_final,
                                         args.position_lr_delay_mult, 0, args.position_lr_max_steps)

    def update_learning_rate(self, iteration=None):
        """Position learning rate for ``iteration`` (R/scene/gaussian_model.py:249-255),
        or for the clock's next iteration when None (what a replayed graph does).
        Stream-ordered: nothing is read back."""
        it = -1 if iteration is None else int(iteration)
        call("gvc_lr_advance", _ptr(self._clock), ctypes.byref(self._schedule), 0, ctypes.c_int64(it),
             _stream(self.device))

    def load_state(self, other):
        """Copy another model's live rows, Adam moments and clock into this one
        (same capacity, both after training_setup): both then continue from the
        same training state."""
        if other.capacity != self.capacity or self._m is None or other._m is None:
            raise ValueError("load_state needs two set-up models of the same capacity")
        with torch.no_grad():
            for g in GROUPS:
                self._buf[g].copy_(other._buf[g])
                self._m[g].copy_(other._m[g])
                self._v[g].copy_(other._v[g])
            self._clock.copy_(other._clock)
        if self.count != other.count:
            self.count = other.count
            self._views()

    def clock(self):
        """(iteration, adam steps, learning rates) read from the device clock."""
        c = gvc_train_clock.from_buffer_copy(self._clock.cpu().numpy().tobytes())
        return int(c.iteration), int(c.step), [c.lr[k] for k in range(len(GROUPS))]

    def optimizer_step(self):
        """optimizer.step(): one fused launch over the parameters with a gradient."""
        if self._m is None:
            raise RuntimeError("training_setup() has not been called")
        arr = (gvc_adam_tensor * len(GROUPS))()
        for k, g in enumerate(GROUPS):
            p = self._params[g]
            gr = p.grad
            if gr is not None:
                if gr.dtype != torch.float32 or gr.shape != p.shape:
                    raise TypeError(f"gradient of {g} must be float32 of shape {tuple(p.shape)}")
                gr = gr.contiguous()
                p.grad = gr
            arr[k] = gvc_adam_tensor(p.data_ptr(), gr.data_ptr() if gr is not None else None,
                                     self._m[g].data_ptr(), self._v[g].data_ptr(), p.numel(), k)
        call("gvc_adam_step", len(GROUPS), arr, _ptr(self._clock), ADAM_BETAS[0], ADAM_BETAS[1], ADAM_EPS,
             _stream(self.device))

    def zero_grad(self):
        for p in self.parameters():
            p.grad = None

    # ---- surgery ------------------------------------------------------------------------

    def _append(self, samples, gt, idx, dist_scale, init_weight, trusted=False):
        k = int(idx.numel()) if idx is not None else int(samples.shape[0])
        self._check_rows(k)
        samples = samples.to(self.device, torch.float32).contiguous()
        gt = gt.to(self.device, torch.float32).contiguous().reshape(-1)
        if idx is not None:
            idx = idx.to(self.device, torch.int32).contiguous()
            if k and not trusted and not bool(((idx >= 0) & (idx < samples.shape[0])).all().item()):
                raise IndexError("densify indices outside the sample array")
        if gt.numel() != samples.shape[0]:
            raise ValueError(f"{gt.numel()} values for {samples.shape[0]} samples")
        total = self.count + k
        nb = lib().gvc_densify_bytes(total)
        ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        call("gvc_densify_in_empty", self.count, k, ctypes.byref(self._store()), _ptr(samples), _ptr(gt),
             _ptr(idx), float(dist_scale), float(init_weight), _ptr(ws), nb, _stream(self.device))
        self.count = total
        self._views()

    def densify_in_empty(self, empty_points, empty_values, new_scale=None, idx=None):
        """R/scene/gaussian_model.py:727-763.  With ``idx``, the new rows are
        ``empty_points[idx]`` and ``empty_values[idx]`` (no gathered copies);
        ``new_scale`` is unused, as in the reference."""
        self._append(empty_points, empty_values, idx, dist_scale=0.1, init_weight=0.01)

    def prune_points(self, mask=None, min_weight=None):
        """prune_points(mask) (R/scene/gaussian_model.py:546-554), or with
        ``min_weight`` the weight rule of densify_and_prune (:786).  Reads the new
        count back (one small copy)."""
        if (mask is None) == (min_weight is None):
            raise ValueError("give exactly one of mask and min_weight")
        if mask is not None:
            mask = mask.reshape(-1).to(self.device, torch.uint8).contiguous()
            if mask.numel() != self.count:
                raise ValueError(f"mask has {mask.numel()} entries for {self.count} Gaussians")
        nb = lib().gvc_prune_bytes(self.count)
        ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        call("gvc_prune", self.count, ctypes.byref(self._store()), _ptr(mask),
             float(min_weight) if min_weight is not None else 0.0, _ptr(self._count_dev), _ptr(ws), nb,
             _stream(self.device))
        self.count = int(self._count_dev.item())
        self._views()

    def densify_and_prune(self, max_grad, min_weight, new_scale, empty_points, empty_values, num_densify=None,
                          idx=None):
        """R/scene/gaussian_model.py:765-799: densify_in_empty, then prune the
        Gaussians whose weight is below ``min_weight`` (max_grad, new_scale and
        num_densify are unused, as in the reference)."""
        self.densify_in_empty(empty_points, empty_values, new_scale, idx=idx)
        self.prune_points(min_weight=min_weight)

    def error_order(self, cells, gt, low_weight=0.005):
        """(order, counts): sample indices with gt != -1 by descending
        |cells - gt| (ties in ascending index), then the rest; counts (device
        int32 [2]) = the valid samples and the Gaussians of weight <= low_weight."""
        cells = cells.detach().reshape(-1).to(torch.float32).contiguous()
        gt = gt.reshape(-1).to(torch.float32).contiguous()
        S = int(cells.numel())
        order = torch.empty(S, dtype=torch.int32, device=self.device)
        counts = torch.empty(2, dtype=torch.int32, device=self.device)
        nb = lib().gvc_error_order_bytes(S)
        ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        call("gvc_error_order", S, _ptr(cells), _ptr(gt), self.count, _ptr(self._buf["weight"]), float(low_weight),
             _ptr(order), _ptr(counts), _ptr(ws), nb, _stream(self.device))
        return order, counts

    def densify_from_errors(self, cells, gt, samples, cap_max, min_weight, max_k=500000):
        """The densification step of R/train.py:562-590: k = min(max_k, cap_max -
        count + #(weight <= 0.005) + 1000, #(gt != -1)) new Gaussians at the samples
        of largest |cells - gt|, then the weight prune.  Returns k.  Reads the two
        counts back (one small copy), then the pruned count, as the reference reads
        its shapes and counts."""
        with torch.no_grad():
            order, counts = self.error_order(cells, gt)
            valid, low = (int(x) for x in counts.cpu().tolist())
            k = min(int(max_k), int(cap_max) - self.count + low + 1000)
            k = min(k, valid)
            if k < 0:
                raise RuntimeError(f"selected index k out of range (k = {k}; is cap_max set?)")
            # indices from the error order are in range by construction
            self._append(samples.reshape(-1, 3), gt.reshape(-1), order[:k], dist_scale=0.1, init_weight=0.01,
                         trusted=True)
            self.prune_points(min_weight=min_weight)
        return k
