"""CPU restatement of the training loop's optimiser and model surgery -- TEST
INFRASTRUCTURE ONLY (imported by tests/, never by the product path).

  expon_lr       get_expon_lr_func (R/utils/general_utils.py:35-69), float64.
                 Pinned by closed forms: lr_init at step 0, lr_final at max_steps,
                 the geometric mean half-way, clamping past max_steps, the delay
                 factor at step 0.
  adam_step      one step of torch.optim.Adam as the reference configures it
                 (R/scene/gaussian_model.py:240: betas (0.9, 0.999), eps 1e-15),
                 float32 with scalars from float64, in torch's multi-tensor order.
                 Pinned against torch.optim.Adam itself.
  error_order    the densification pick of R/train.py:562-578: valid samples
                 (gt != -1) by descending |cells - gt|, ties in ascending index.
                 Pinned against torch.topk (values, and indices where no ties).
  new_rows       the rows densify_in_empty / create_from_pcd append
                 (R/scene/gaussian_model.py:727-763, :175-199), with the 3-NN of
                 mesh_ref.knn3 (pinned against scipy's KD-tree).
  prune          boolean-mask removal of rows (_prune_optimizer, :524-554).
  RefModel       the reference's surgery on torch tensors and a real
                 torch.optim.Adam (param_groups / state replaced as
                 cat_tensors_to_optimizer and _prune_optimizer replace them), for
                 loop-level comparisons.
"""
import math

import numpy as np
import torch

from . import mesh_ref

GROUPS = ("xyz", "weight", "scaling", "rotation", "value")


def expon_lr(step, lr_init, lr_final, lr_delay_steps=0, lr_delay_mult=1.0, max_steps=1000000):
    """R/utils/general_utils.py:53-69."""
    if step < 0 or (lr_init == 0.0 and lr_final == 0.0):
        return 0.0
    if lr_delay_steps > 0:
        delay_rate = lr_delay_mult + (1 - lr_delay_mult) * np.sin(0.5 * np.pi * np.clip(step / lr_delay_steps, 0, 1))
    else:
        delay_rate = 1.0
    t = np.clip(step / max_steps, 0, 1)
    return float(delay_rate * np.exp(np.log(lr_init) * (1 - t) + np.log(lr_final) * t))


def adam_step(p, g, m, v, lr, step, beta1=0.9, beta2=0.999, eps=1e-15):
    """New (p, m, v), float32, for Adam step number ``step`` (>= 1)."""
    f = np.float32
    p, g, m, v = (np.asarray(a, dtype=f) for a in (p, g, m, v))
    bc1 = 1 - beta1 ** step
    bc2 = 1 - beta2 ** step
    step_size = f(-(lr / bc1))
    bc2_sqrt = f(math.sqrt(bc2))
    # lerp as torch computes it: fma(weight, g - m, m), one rounding
    m = (np.float64(f(1 - beta1)) * (g - m).astype(np.float64) + m.astype(np.float64)).astype(f)
    v = (v * f(beta2) + f(1 - beta2) * g * g).astype(f)
    den = (np.sqrt(v) / bc2_sqrt + f(eps)).astype(f)
    p = (p + step_size * (m / den)).astype(f)
    return p, m, v


def error_order(cells, gt):
    """(order of the valid samples, number valid)."""
    cells = np.asarray(cells, dtype=np.float32).reshape(-1)
    gt = np.asarray(gt, dtype=np.float32).reshape(-1)
    valid = np.flatnonzero(gt != -1.0)
    err = np.abs(cells[valid] - gt[valid]).astype(np.float64)
    err = np.where(np.isnan(err), np.inf, err)
    rank_nan = np.isnan(np.abs(cells[valid] - gt[valid]))
    # primary: NaN first, then descending error; secondary: ascending index
    order = np.lexsort((valid, -err, ~rank_nan))
    return valid[order], valid.size


def new_rows(all_xyz, first, gt_values, dist_scale, init_weight):
    """Rows [first, len(all_xyz)) as the reference initialises them: dict of
    float32 arrays (xyz, weight, scaling, rotation, value)."""
    f = np.float32
    all_xyz = np.asarray(all_xyz, dtype=f)
    d2 = mesh_ref.knn3(all_xyz).astype(f)[first:]
    s = np.sqrt(np.maximum(d2 * f(dist_scale), f(1e-7))).astype(f)
    ls = np.log(s).astype(f)
    k = all_xyz.shape[0] - first
    rot = np.zeros((k, 4), f)
    rot[:, 0] = 1
    w0 = f(init_weight)
    g = np.asarray(gt_values, dtype=f).reshape(-1)
    with np.errstate(divide="ignore"):
        value = np.log(g / (f(1) - g)).astype(f)
    return {"xyz": all_xyz[first:], "weight": np.full((k, 1), np.log(w0 / (f(1) - w0)), f),
            "scaling": np.repeat(ls[:, None], 3, axis=1), "rotation": rot, "value": value.reshape(-1, 1)}


def prune(arrays, keep):
    return {k: np.asarray(a)[np.asarray(keep, dtype=bool)] for k, a in arrays.items()}


class RefModel:
    """The reference's parameters and optimiser, on torch tensors: Adam with the
    five groups, state surgery as R/scene/gaussian_model.py:524-612 does it."""

    def __init__(self, params, lrs, device):
        self.p = {g: torch.nn.Parameter(torch.as_tensor(params[g], dtype=torch.float32, device=device).clone())
                  for g in GROUPS}
        self.opt = torch.optim.Adam([{"params": [self.p[g]], "lr": lrs[i], "name": g} for i, g in enumerate(GROUPS)],
                                    lr=0.0, eps=1e-15)

    def _replace(self, fn_param, fn_state):
        for group in self.opt.param_groups:
            old = group["params"][0]
            st = self.opt.state.get(old, None)
            new = torch.nn.Parameter(fn_param(group["name"], old.detach()).requires_grad_(True))
            if st is not None:
                st["exp_avg"] = fn_state(group["name"], st["exp_avg"])
                st["exp_avg_sq"] = fn_state(group["name"], st["exp_avg_sq"])
                del self.opt.state[old]
                self.opt.state[new] = st
            group["params"][0] = new
            self.p[group["name"]] = new

    def cat(self, rows):
        dev = self.p["xyz"].device
        ext = {g: torch.as_tensor(rows[g], dtype=torch.float32, device=dev) for g in GROUPS}
        self._replace(lambda g, t: torch.cat((t, ext[g]), dim=0),
                      lambda g, s: torch.cat((s, torch.zeros_like(ext[g])), dim=0))

    def prune(self, keep):
        keep = torch.as_tensor(keep, dtype=torch.bool, device=self.p["xyz"].device)
        self._replace(lambda g, t: t[keep], lambda g, s: s[keep])

    def moments(self):
        out = {}
        for group in self.opt.param_groups:
            st = self.opt.state.get(group["params"][0], None)
            out[group["name"]] = None if st is None else (st["exp_avg"], st["exp_avg_sq"], float(st["step"]))
        return out
