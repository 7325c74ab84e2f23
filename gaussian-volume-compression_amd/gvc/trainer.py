"""The fitting loop of R/train.py:95-616 on libgvc, with the steady iterations
replayed from a graph.

Per iteration, as the reference: pick one of the precomputed sample batches at
random (``np.random.randint``, :354-362), build the sample set (``build_bvh``),
advance the position learning rate, evaluate the mixture, take the fitting loss
(L1 over the reconstructed samples -- over every sample when encoding the
surface -- + fn_reg x the false-negative term, + fp_reg x the false-positive term
when encoding the surface), back-propagate, switch fn_reg
to fn_reg2 after densify_from_iter, densify and prune at the densification
interval (:551-590), else take an Adam step (:593-595; at a densification
iteration the reference's parameters were replaced, so its step has no gradients
and does nothing).

Iterations without densification are captured once per (Gaussian count, fn_reg)
and replayed: batch copy into static buffers, sample set, learning-rate advance,
activations, forward, loss, backward and Adam all run from the graph with no host
synchronisation.  A densification iteration runs eagerly (it reads counts back, as
the reference does), and the next iteration captures again.  The captured
backward uses a record capacity fixed from the last eager iteration with a
margin; the mixture's device error flags are accumulated on the device and
checked at every densification and at the end (an overflow raises).
"""
import dataclasses

import numpy as np
import torch

from bvh_diff_gaussian_rasterization import GaussianRasterizationSettings, GaussianRasterizer

from . import mixture as mx
from .training import fit_loss


@dataclasses.dataclass
class LoopArgs:
    """The loop's own arguments (R/train.py:640-681, R/arguments/__init__.py:60-70)."""
    cap_max: int = 1_000_000
    min_weight: float = 0.005
    fn_reg: float = 1.0
    fn_reg2: float = 1.0
    fp_reg: float = 0.0
    encode_surface: bool = False
    saving_iterations: tuple = ()
    testing_iterations: tuple = ()


class Trainer:
    """Fits ``model`` (a trained-ready gvc.model.GaussianModel) to batches of
    samples [B, S, 3] with targets [B, S] (gt -1 outside the volume)."""

    def __init__(self, model, samples, gt, opt, loop=None, graph=True, seed=0, record_margin=1.25):
        if samples.dim() != 3 or samples.shape[2] != 3 or gt.shape != samples.shape[:2]:
            raise ValueError("samples must be [B, S, 3] and gt [B, S]")
        self.model = model
        self.samples = samples.to(model.device, torch.float32).contiguous()
        self.gt = gt.to(model.device, torch.float32).contiguous()
        self.opt = opt
        self.loop = loop or LoopArgs()
        self.use_graph = graph
        self.rng = np.random.RandomState(seed)
        self.record_margin = record_margin
        settings = GaussianRasterizationSettings(volume_mins=(0, 0, 0), volume_maxes=(1, 1, 1), cell_count=128,
                                                 bg=-1.0, scale_modifier=1.0, use_gaussian_bvh=False, debug=False)
        self.rast = GaussianRasterizer(settings)
        self.fn_reg = self.loop.fn_reg
        dev = model.device
        S = self.samples.shape[1]
        self.static_samples = torch.empty((S, 3), dtype=torch.float32, device=dev)
        self.static_gt = torch.empty((S,), dtype=torch.float32, device=dev)
        self.flags = torch.zeros(3, dtype=torch.int64, device=dev)   # pairs, records, error flags
        self.graph = None
        self.graph_key = None
        self.graph_out = None
        self.audit = []
        self.last_records = None
        self.densifications = 0
        self.replays = 0
        self.captures = 0

    def _densify_now(self, it):
        o = self.opt
        return (o.densify_from_iter <= it <= o.densify_until_iter and it % o.densification_interval == 0
                and it not in self.loop.saving_iterations and it not in self.loop.testing_iterations)

    def _body(self, samples, gt, step, iteration):
        m = self.model
        self.audit.clear()
        self.rast.build_bvh(samples)
        m.update_learning_rate(iteration)
        scales, rots, w, v = m.activated()
        cells, cw = self.rast(m.get_xyz, scales, rots, v, w)
        loss = fit_loss(cells, cw, gt, self.fn_reg, self.loop.fp_reg, self.loop.encode_surface)
        loss.backward()
        self.audit[-1].accumulate(self.flags)
        if step:
            m.optimizer_step()
        m.zero_grad()
        return cells.detach(), cw.detach(), loss.detach()

    def _eager(self, j, it, step):
        mx.set_record_capacity(None)
        out = self._body(self.samples[j], self.gt[j], step, it)
        self.last_records = self.audit[-1].stats()[0]
        return out

    def _capture(self, it):
        cap = int(self.last_records * self.record_margin) + 4096
        mx.set_record_capacity(cap)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.graph_out = self._body(self.static_samples, self.static_gt, True, None)
        self.graph_key = (self.model.count, self.fn_reg)
        self.captures += 1

    def check_flags(self):
        errs = int(self.flags[2].item())
        if errs:
            raise RuntimeError(f"mixture error flags raised during the loop: {errs}")

    def run(self, first_iteration=1, last_iteration=None, on_iteration=None):
        """Iterations first_iteration .. last_iteration (default opt.iterations).
        ``on_iteration(it, cells, weights, loss)`` sees each iteration's outputs
        (device tensors; replayed iterations hand out the graph's buffers)."""
        last = self.opt.iterations if last_iteration is None else last_iteration
        mx.set_audit(self.audit)
        try:
            for it in range(first_iteration, last + 1):
                j = self.rng.randint(0, self.samples.shape[0])
                densify = self._densify_now(it)
                step = it < self.opt.iterations and not densify
                key = (self.model.count, self.fn_reg)
                if densify or not self.use_graph or not step or self.last_records is None:
                    cells, cw, loss = self._eager(j, it, step)
                else:
                    self.static_samples.copy_(self.samples[j])
                    self.static_gt.copy_(self.gt[j])
                    if self.graph is None or self.graph_key != key:
                        # the clock's iteration is set here, the graph advances it
                        self.model.update_learning_rate(it - 1)
                        self._capture(it)
                    self.graph.replay()
                    self.replays += 1
                    cells, cw, loss = self.graph_out
                if on_iteration is not None:
                    on_iteration(it, cells, cw, loss)
                if it > self.opt.densify_from_iter:
                    self.fn_reg = self.loop.fn_reg2
                if densify:
                    self.check_flags()
                    self.model.densify_from_errors(cells, self.gt[j], self.samples[j], self.loop.cap_max,
                                                   self.loop.min_weight)
                    self.densifications += 1
            torch.cuda.synchronize()
            self.check_flags()
        finally:
            mx.set_audit(None)
            mx.set_record_capacity(None)
