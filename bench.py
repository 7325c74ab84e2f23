#!/usr/bin/env python3
"""Benchmark: one training step of the unstructured-mesh fit on one MI355X.

Workload (default ``--config mesh``, the 1M-Gaussian configuration the metric is
quoted on): a 129^3 jittered lattice split into 6 Kuhn tetrahedra per cube
(12,582,912 tets) carrying a 32-blob scalar field; 1,000,000 Gaussians at random
distinct mesh vertices, initialised like the reference (3-NN scales via
distCUDA2, identity rotations, weight 0.1, value = vertex value).  Every step:

  1. 2^22 uniform sample points (torch.rand on the device);
  2. the sample set (Morton sort into 64-point groups);
  3. exact tet location + barycentric interpolation of the ground truth (libgvc);
  4. the reference model's activations (soft-capped exp scales, max_scale 0.02,
     normalised quaternions, sigmoid weights and values; fused libgvc kernel);
  5. the differentiable mixture forward (libgvc, drop-in
     bvh_diff_gaussian_rasterization API);
  6. masked L1 + false-negative loss (R/train.py:389-420, fused libgvc kernel),
     backward through loss, mixture (deterministic reduction) and activations;
  7. the position learning-rate schedule and Adam with the reference's per-group
     rates (R/scene/gaussian_model.py:207-255): one fused libgvc launch over the
     fixed-capacity parameter buffers of gvc.model.GaussianModel, step count and
     rates in a device clock.
With --graph, steps 2-7 are captured once in a hipGraph and replayed (the
samples are drawn into a static buffer before each replay).

value = Gaussian-point evaluations per second, forward + backward
(2 x accepted pairs per step, summed over the timed steps / wall time).
The roofline object is the dominant kernel (the larger of the forward and
backward walks), which is bound by VALU issue: achieved = the pair math's
lane-ops x accepted pairs / its mean duration, measured with HIP events on the
launch stream inside the timed steps, against the non-packed fp32 VALU peak;
the 44-B-per-pair streaming model and the committed rocprofv3 counters (HBM
bytes, VALU issue) are reported beside it.  cpu_baseline is the plain fp32
reference formula (brute force, box predicate) from the oracle, timed on the
host cores on a bounded subset of the same step's points (its rate is
extrapolated to the step), plus one full step of the 2,048-Gaussian CPU
configuration.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gaussian-volume-compression_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PAIR = 44  # mean 12 + conic 24 + weight 4 + value 4
# VALU issue peak with packed fp32 off: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
# (one wave64 instruction per SIMD every 4 cycles), in 1e12 lane-ops/s.
VALU_PEAK_TLANE = 256 * 4 * 16 * 2.4e9 / 1e12
# VALU lane-ops of one accepted pair in the walks' pair math: forward d (3),
# conic.d (9), quad (3), power (1), window (2), exp (2), w e and w e v (2), two
# Kahan sums (8); backward the same quadratic form and exp (20) plus the 11
# partial sums (20).
VALU_OPS_PER_PAIR = {"fwd_walk": 30, "bwd_walk": 40}
PMC_FILE = os.path.join(ROOT, "profiles", "pmc_counters.json")
ADAM_BYTES_PER_GAUSSIAN = 12 * 7 * 4  # 12 floats: p, g, m, v read; p, m, v written


def make_model(W, max_scale, dev, capacity=None):
    """A gvc.model.GaussianModel holding the workload's raw parameters, with the
    reference's optimiser settings."""
    import torch
    from gvc.model import GaussianModel, OptimizationParams
    model = GaussianModel(capacity or W["means"].shape[0], dev, max_scale=max_scale)
    model.load_raw(W["means"], torch.logit(W["weights"]), W["log_scales"], W["rotations"],
                   torch.logit(W["values"].clamp(1e-4, 1 - 1e-4)))
    model.training_setup(OptimizationParams())
    return model


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="mesh", choices=["mesh", "grid", "batch32", "precision", "train", "eval"],
                    help="mesh: the 1M-Gaussian mesh fit (the headline line); grid: the 200k-Gaussian grid fit; "
                         "batch32: 32 uneven volumes in one batched launch per pass beside the plain sequence; "
                         "precision: 4M Gaussians on a 512^3 field, naive / Kahan / FP64-cond against the FP64 "
                         "oracle on 65,536 points, plus timed steps; train: the fitting loop of the mesh "
                         "workload with densification every 100 iterations (gvc.trainer); eval: the volume "
                         "evaluation of the 1M-Gaussian mesh model on a jittered 128^3 grid (gvc.evaluation)")
    ap.add_argument("--plain-steps", type=int, default=3, help="batch32: timed steps of the plain sequence")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--fp64-ratio", type=float, default=0.0,
                    help="precision: scale ratio above which Gaussians take the FP64 path (0 = 1000, cond 1e6)")
    ap.add_argument("--leaf", type=int, default=None)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--graph", action="store_true", help="capture the training step in a hipGraph (default for "
                                                         "--config grid, whose eager step is bound by launches)")
    ap.add_argument("--eager", action="store_true", help="grid: run the step eagerly instead of from a graph")
    ap.add_argument("--max-scale", type=float, default=0.02,
                    help="soft scale cap of the reference model (R/scene/gaussian_model.py:64); 0 disables it")
    ap.add_argument("--autotune", action="store_true",
                    help="choose leaf size x run points x block shape on the workload before timing")
    ap.add_argument("--recompute", action="store_true",
                    help="backward walks the tree again instead of replaying the forward's candidate log")
    ap.add_argument("--no-batch32", action="store_true",
                    help="mesh: skip the 32-volume batch legs that the default line reports beside the headline")
    ap.add_argument("--no-batch-autotune", action="store_true",
                    help="mesh: run the 32-volume batch leg at the default launch parameters instead of the "
                         "autotuner's choice (leaf size x points per walk x block size, swept on the batch)")
    ap.add_argument("--no-side-configs", action="store_true",
                    help="mesh: skip the structured-grid (200k Gaussians, 2^20 points) and precision (4M "
                         "Gaussians, 2^23 points, FP64 path) legs the default line reports beside the headline")
    ap.add_argument("--no-study", action="store_true",
                    help="precision: skip the five-mode precision study (the step and its error check still run)")
    ap.add_argument("--batch32-steps", type=int, default=10,
                    help="mesh: timed fused steps of the 32-volume batch leg (the plain sequence takes 2)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (x steps): the calibration launch for the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; --gpus above 1 is not supported: "
                                   "the path runs on one GPU and is not sharded across several", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    if args.copy_probe:
        n = args.copy_probe << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "launches": args.steps}))
        return

    if args.config == "batch32":
        return main_batch(args)
    if args.config == "train":
        return main_train(args)
    if args.config == "eval":
        return main_eval(args)

    from bvh_diff_gaussian_rasterization import GaussianRasterizationSettings, GaussianRasterizer
    from gvc import mixture as mx
    from gvc import training
    from gvc import workloads
    from gvc.sampling import TetMeshSampler

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t_setup = time.perf_counter()
    if args.config == "mesh":
        W = workloads.mesh_workload(dev, seed=args.seed)
        S = 1 << 22
        sampler = TetMeshSampler(W["mesh_points"], W["tets"], W["mesh_values"], device=dev)
        workload_name = ("unstructured mesh fit: 1,000,000 Gaussians, 2^22 mesh-sampled points/step, "
                         "129^3 jittered lattice, 12,582,912 tets")
    elif args.config == "grid":
        W = workloads.grid_workload(dev, seed=args.seed)
        S = 1 << 20
        sampler = None
        workload_name = "structured 128^3 grid fit: 200,000 Gaussians, 2^20 points/step"
    else:
        W = workloads.precision_workload(dev, seed=args.seed)
        W["grid_fn"] = W["field_fn"]
        S = 1 << 23
        sampler = None
        workload_name = ("precision study fit: 4,000,000 Gaussians (2% with one log-scale ln 1e-4, 0.5% ln 1e-6), "
                         "512^3 field of 256 blobs + 3 sinusoid octaves, 2^23 points/step")
    log(f"setup {time.perf_counter() - t_setup:.1f}s P={W['P']} S={S}")
    P = W["P"]

    params = mx.default_params()
    if args.leaf or args.block or args.recompute or args.config == "precision":
        params = mx.MixtureParams(leaf_size=args.leaf or params.leaf_size,
                                  block_threads=args.block or params.block_threads,
                                  candidate_log=not args.recompute, fp64_cond=args.config == "precision",
                                  fp64_scale_ratio=args.fp64_ratio if args.config == "precision" else 0.0)
        mx.set_default_params(params)

    study = precision_study(args, W, dev) if args.config == "precision" and not args.no_study else None

    tune = None
    if args.autotune:
        from gvc.autotune import autotune
        with torch.no_grad():
            s0 = torch.rand((S, 3), device=dev, generator=torch.Generator(device=dev).manual_seed(args.seed + 7))
            ss0 = mx.SampleSet(s0)
            wts = W["weights"].reshape(-1, 1).contiguous()
            vls = W["values"].reshape(-1, 1).contiguous()
            best, table = autotune(W["means"], torch.exp(W["log_scales"]), W["rotations"], vls, wts, ss0,
                                   base=params, use_cache=False)
        params = best
        mx.set_default_params(params)
        tune = {"chosen": best.describe(),
                "ms_fwd_bwd": {k: (round(v, 3) if v is not None else None) for k, v in table.items()}}
        log(f"autotune: {tune['chosen']}")

    # raw parameters as the reference model keeps them (log-scales of the capped
    # scale, logits of weight and value), in the model's fixed-capacity buffers
    model = make_model(W, args.max_scale, dev)
    settings = GaussianRasterizationSettings(volume_mins=(0, 0, 0), volume_maxes=(1, 1, 1), cell_count=128, bg=-1.0,
                                             scale_modifier=1.0, use_gaussian_bvh=False, debug=False)
    rast = GaussianRasterizer(settings)
    fn_reg = 1.0
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed + 1)
    stage_ev = []
    audit = []
    acc = torch.zeros(3, dtype=torch.int64, device=dev)   # pairs, records, error flags
    timing_stages = [False]

    def stamp(name, fn):
        if not timing_stages[0]:
            return fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        out = fn()
        e1.record()
        stage_ev.append((name, e0, e1))
        return out

    def body(samples):
        audit.clear()
        stamp("sample_set", lambda: rast.build_bvh(samples))
        if sampler is not None:
            gt = stamp("tet_locate", lambda: sampler.sample(samples, sample_set=rast.sample_set)[0])
        else:
            gt = W["grid_fn"](samples)
        scales, rots, w, v = stamp("activate", lambda: model.activated())
        cells, cw = rast(model.get_xyz, scales, rots, v, w)
        loss = stamp("loss", lambda: training.fit_loss(cells, cw, gt, fn_reg))
        loss.backward()
        model.update_learning_rate()
        stamp("adam", lambda: model.optimizer_step())
        model.zero_grad()
        audit[-1].accumulate(acc)
        return samples, cells, cw

    def eager_step():
        return body(torch.rand((S, 3), device=dev, generator=gen))

    # warmup: sizes the backward's record capacity and the candidate log (no host
    # reads afterwards)
    mx.set_record_capacity(None)
    mx.set_audit(audit)
    max_rec = 0
    lane_groups = 0
    for _ in range(args.warmup):
        eager_step()
        torch.cuda.synchronize()
        max_rec = max(max_rec, audit[-1].stats()[0])
        lane_groups = audit[-1].lane_groups
    cap = int(max_rec * 1.25) + 4096
    mx.set_record_capacity(cap)
    log(f"warmup done, record capacity {cap}, candidate-log slots/group {mx.chunk_slots_for((S + 63) // 64)}")
    # stage profile (untimed steps, HIP events between the stages)
    timer = []
    mx.set_stage_timer(timer)
    timing_stages[0] = True
    for _ in range(3):
        eager_step()
    torch.cuda.synchronize()
    timing_stages[0] = False
    mx.set_stage_timer(None)
    stage_ms = {}
    for name, e0, e1 in timer + stage_ev:
        stage_ms.setdefault(name, []).append(e0.elapsed_time(e1))
    stage_avg = {k: float(np.mean(v)) for k, v in stage_ms.items()}

    graph = None
    if args.graph or (args.config == "grid" and not args.eager):
        args.graph = True
        static = torch.empty((S, 3), device=dev)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                torch.rand((S, 3), device=dev, generator=gen, out=static)
                body(static)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        model.zero_grad()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            graph_out = body(static)
        torch.cuda.synchronize()
        log("training step captured in a graph")

    acc.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if graph is None:
            last = eager_step()
        else:
            torch.rand((S, 3), device=dev, generator=gen, out=static)
            graph.replay()
            last = graph_out
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mx.set_audit(None)
    wall = t1 - t0
    total_pairs, total_records, errs = (int(x) for x in acc.cpu().tolist())
    if errs:
        raise RuntimeError(f"mixture error flags during the timed steps: {errs}")
    pairs_per_step = total_pairs / args.steps
    value = 2.0 * total_pairs / wall
    roofline = walk_roofline(stage_avg, pairs_per_step, args.config == "mesh")

    out = {
        "metric": "Gaussian-point evals/s fwd+bwd @1M Gaussians",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload_name, "gaussians": P, "points_per_step": S,
                   "leaf_size": params.leaf_size, "block_threads": params.block_threads,
                   "run_points": params.run_points, "kahan": params.kahan, "candidate_log": params.candidate_log,
                   "fp64_cond": params.fp64_cond,
                   "fp64_scale_ratio": (params.fp64_scale_ratio or 1000.0) if params.fp64_cond else None,
                   "graph": bool(args.graph), "parallelism": "single GPU"},
        "pairs_per_step": pairs_per_step,
        "records_per_step": total_records / args.steps,
        "groups_walked_per_lane": lane_groups,
        "precision_study": study,
        "autotune": tune,
        "stage_ms": stage_avg,
        # fused Adam: p, g, m, v read and p, m, v written, 12 floats per Gaussian
        "optimizer": {"kernel": "adam", "bytes_per_launch": ADAM_BYTES_PER_GAUSSIAN * P,
                      "ms": stage_avg.get("adam"),
                      "GB/s": (ADAM_BYTES_PER_GAUSSIAN * P / (stage_avg["adam"] * 1e6)
                               if stage_avg.get("adam") else None)},
        "roofline": roofline,
    }

    if not args.no_error or not args.no_cpu_baseline:
        import oracle
        samples, cells, cw = last
        with torch.no_grad():
            m = model.get_xyz.detach()
            sc, rq, wv, vv = (t.detach() for t in model.activated())
            # the last step's forward ran before its Adam update: re-evaluate at
            # the current parameters for the error comparison
            ss = mx.SampleSet(samples)
            c2, w2, _, _ = mx.mixture_forward(m, sc, rq, vv, wv, 1.0, ss)
        mh, sh, rh, wh, vh = (t.cpu().numpy() for t in (m, sc, rq, wv, vv))
        # the oracle in the step's own mode (FP64 path and precise evaluation when on)
        conic, box, _, fac = oracle.preprocess(mh, sh, rh, wh, 1.0, params.fp64_cond, params.fp64_scale_ratio,
                                               factor=True, factor_ratio=params.factor_scale_ratio)
        fac = fac if params.fp64_cond else None
        sub = np.random.default_rng(0).choice(S, 1024, replace=False)
        pts = samples.cpu().numpy()[sub]
        if not args.no_error:
            Vr, Wr, _ = oracle.forward(mh, vh, wh, conic, box, pts, oracle.conic_fp64(sh, rh), factor=fac)
            V, Wg = c2.cpu().numpy()[sub], w2.cpu().numpy()[sub]
            ok = Wr > 0
            flips = int(np.sum((Wg > 0) != ok))
            err = max(float(np.max(np.abs(V[ok] - Vr[ok]) / np.maximum(np.abs(Vr[ok]), 1e-3), initial=0)),
                      float(np.max(np.abs(Wg[ok] - Wr[ok]) / np.maximum(np.abs(Wr[ok]), 1e-3), initial=0)))
            out["rel_err_vs_fp64"] = {"field_max_rel": err, "sentinel_flips": flips, "points": int(len(sub))}
            out["rel_err_vs_fp64"]["gradients"] = step_gradients((m, sc, rq, wv, vv), samples,
                                                                 (mh, sh, rh, wh, vh), conic, box, fac, params)
        if not args.no_cpu_baseline:
            rng = np.random.default_rng(1)
            allpts = samples.cpu().numpy()

            def run_cpu(n):
                t2 = time.perf_counter()
                _, _, _, pr = oracle.cpu_plain_fwdbwd(mh, vh, wh, conic, box, allpts[:n], rng.normal(size=n),
                                                     rng.normal(size=n))
                return time.perf_counter() - t2, pr

            n = 512
            while True:
                dt, pr = run_cpu(n)
                if dt >= 0.5 * args.cpu_seconds or n >= S:
                    break
                n = int(min(S, n * min(8.0, max(1.5, args.cpu_seconds / max(dt, 1e-3)))))
            c0 = cpu_config0_step()
            out["cpu_baseline"] = {"value": 2.0 * pr / dt, "unit": "evals/s", "cores": oracle.num_threads(),
                                   "cpu_model": cpu_model(), "kind": "port", "extrapolated": n < S,
                                   "sample": f"{n} of the step's {S} points against all {P} Gaussians, plain fp32 "
                                             f"brute force fwd+bwd with the box predicate ({dt:.1f} s); the rate "
                                             f"is extrapolated to the whole step",
                                   "step_s_extrapolated": dt * S / n,
                                   "config0_full_step": c0}
    if args.config == "mesh" and not args.no_batch32:
        # the 32-volume batch (fused batched step vs the plain per-volume
        # sequence, both from the same training state) beside the headline
        del model, rast, last
        mx.set_default_params(mx.MixtureParams())
        torch.cuda.empty_cache()
        b = batch32_legs(args, args.batch32_steps, 2, 3)
        out["batch32"] = {k: b[k] for k in ("ms_per_step", "pairs_per_step", "records_per_step", "evals_per_s",
                                            "speedup_vs_plain", "stage_ms", "same_start_state", "autotune")}
        out["batch32"].update(workload=b["config"]["workload"], leaf_size=b["config"]["leaf_size"],
                              block_threads=b["config"]["block_threads"], run_points=b["config"]["run_points"],
                              plain_ms_per_step=b["plain_sequence"]["ms_per_step"],
                              plain_pairs_per_step=b["plain_sequence"]["pairs_per_step"], steps=b["steps"],
                              plain_steps=b["plain_sequence"]["steps"])
    if args.config == "mesh" and not args.no_side_configs:
        # the structured-grid fit and the precision workload, each in a child
        # process of its own, reported beside the headline
        out["grid"] = side_config(["--config", "grid", "--steps", "20", "--warmup", "5", "--no-cpu-baseline"])
        out["precision"] = side_config(["--config", "precision", "--graph", "--steps", "10", "--warmup", "3",
                                        "--no-cpu-baseline", "--no-study"])
    print(json.dumps(out))


def side_config(extra):
    """Runs this script on another workload in a child process (its progress
    lines go to this process's stderr) and returns the main figures of its line,
    or the error it ended with."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__)] + extra
    log("side leg: " + " ".join(extra))
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=600)
    except subprocess.TimeoutExpired:
        return {"error": "timed out", "args": extra}
    lines = [l for l in r.stdout.strip().splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit status {r.returncode}", "args": extra}
    d = json.loads(lines[-1])
    keep = ("ms_per_step", "value", "unit", "pairs_per_step", "records_per_step", "stage_ms", "roofline",
            "rel_err_vs_fp64")
    res = {k: d[k] for k in keep if k in d}
    res["workload"] = d["config"]["workload"]
    res["graph"] = d["config"].get("graph")
    res["fp64_cond"] = d["config"].get("fp64_cond")
    res["args"] = " ".join(extra)
    return res


def walk_roofline(stage_avg, pairs_per_step, counters_apply):
    """Roofline of the dominant kernel (the larger of the two walks).  The walks
    are bound by VALU issue (their counters show the issue slots nearly full and
    HBM at a tenth of its peak): achieved = the pair math's algorithmic lane-ops
    (VALU_OPS_PER_PAIR x accepted pairs) / the kernel's mean duration (HIP events
    on the launch stream), against the non-packed fp32 VALU peak.  Beside it: the
    streaming model (44 B per pair against HBM peak: above 1 when records are
    reused from LDS and L2), and, for the mesh workload, the committed rocprofv3
    counters of that kernel (HBM bytes per launch from FETCH_SIZE / WRITE_SIZE,
    VALU issue and lane activity), labelled with the file and the commit they were
    measured at."""
    dom = max(("fwd_walk", "bwd_walk"), key=lambda k: stage_avg.get(k, 0.0))
    dom_s = stage_avg[dom] / 1e3
    ops = VALU_OPS_PER_PAIR[dom] * pairs_per_step
    achieved = ops / dom_s / 1e12
    stream = BYTES_PER_PAIR * pairs_per_step / dom_s / 1e9
    pmc = None
    if counters_apply and os.path.exists(PMC_FILE):
        try:
            allc = json.load(open(PMC_FILE))
            k = allc["kernels"].get(dom)
            if k:
                pmc = dict(k, source=os.path.relpath(PMC_FILE, ROOT), measured_at=allc.get("measured_at"),
                           hbm_GBps=k["hbm_bytes"] / dom_s / 1e9, hbm_frac=k["hbm_bytes"] / dom_s / 1e9 / HBM_PEAK_GBS)
        except (OSError, ValueError, KeyError):
            pmc = None
    return {"bound": "valu", "kernel": dom, "achieved": achieved, "peak": VALU_PEAK_TLANE, "unit": "Tlane-op/s",
            "frac": achieved / VALU_PEAK_TLANE, "traffic": pmc["hbm_bytes"] if pmc else None,
            "ops_per_launch": ops, "ops_per_pair": VALU_OPS_PER_PAIR[dom], "mean_ms": stage_avg[dom],
            "streaming_model": {"achieved": stream, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": stream / HBM_PEAK_GBS, "bytes_per_launch": BYTES_PER_PAIR * pairs_per_step},
            "counters": pmc}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_config0_step():
    """The CPU baseline at full size on the CPU configuration (2,048 Gaussians,
    the 32^3 grid): one plain fp32 forward + backward, seconds."""
    import numpy as np

    import oracle
    from gvc import workloads
    g = workloads.config0(0)
    scales = np.exp(g["log_scales"]).astype(np.float32)
    conic, box, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    S = g["samples"].shape[0]
    rng = np.random.default_rng(3)
    t0 = time.perf_counter()
    _, _, _, pairs = oracle.cpu_plain_fwdbwd(g["means"], g["values"], g["weights"], conic, box, g["samples"],
                                             rng.normal(size=S), rng.normal(size=S))
    dt = time.perf_counter() - t0
    return {"step_s": dt, "pairs": pairs, "evals_per_s": 2.0 * pairs / dt, "gaussians": int(g["means"].shape[0]),
            "points": int(S)}


def precision_study(args, W, dev, n_points=65536, n_stress=16384):
    """Field error of the accumulation / inversion modes against the FP64 oracle
    (FP64 conics, the mode's own canonical predicate) at the workload's initial
    parameters, on two fixed point sets: n_points uniform points, and n_stress
    points drawn inside the boxes of the Gaussians whose covariance takes the FP64
    path at the default threshold (scale ratio above 1000, cond(Sigma) > 1e6),
    where the fp32 inverse is least accurate.  Modes: fp32 conic form with naive
    or Kahan sums; Kahan with the precise path at the default thresholds (pairs
    evaluated from the conic's Cholesky factor in compensated fp32: the factor of
    the FP64 inverse above scale ratio 1000, the float-float factor above 10);
    the same with the float-float band switched off (FP64 factor above 1000
    only); and the FP64 factor above ratio 10."""
    import numpy as np
    import torch

    import oracle
    from gvc import mixture as mx
    from gvc import training

    t0 = time.perf_counter()
    with torch.no_grad():
        sc, rq, wv, vv = training.activate(W["log_scales"], W["rotations"], torch.logit(W["weights"]),
                                           torch.logit(W["values"].clamp(1e-4, 1 - 1e-4)), args.max_scale)
        m = W["means"]
    mh, sh, rh, wh, vh = (t.detach().cpu().numpy() for t in (m, sc, rq, wv, vv))
    cd = oracle.conic_fp64(sh, rh)
    modes = {"fp32_naive": mx.MixtureParams(kahan=False), "fp32_kahan": mx.MixtureParams(kahan=True),
             "fp32_kahan_fp64_cond": mx.MixtureParams(kahan=True, fp64_cond=True),
             "fp32_kahan_fp64_only": mx.MixtureParams(kahan=True, fp64_cond=True, factor_scale_ratio=1000.0),
             "fp32_kahan_fp64_ratio10": mx.MixtureParams(kahan=True, fp64_cond=True, fp64_scale_ratio=10.0)}
    pre = {name: oracle.preprocess(mh, sh, rh, wh, 1.0, prm.fp64_cond, prm.fp64_scale_ratio, factor=True,
                                   factor_ratio=prm.factor_scale_ratio)
           for name, prm in modes.items()}
    ill = np.asarray(pre["fp32_kahan_fp64_cond"][2]).reshape(-1).astype(bool)
    rng = np.random.default_rng(args.seed + 12)
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed + 11)
    uniform = torch.rand((n_points, 3), device=dev, generator=gen)
    box_ill = pre["fp32_kahan_fp64_cond"][1][ill]
    pick = rng.integers(0, max(int(ill.sum()), 1), n_stress)
    stress = (box_ill[pick, :3] + rng.uniform(0, 1, (n_stress, 3)) * (box_ill[pick, 3:] - box_ill[pick, :3])
              if ill.any() else rng.uniform(0, 1, (n_stress, 3))).astype(np.float32)
    fac_def = pre["fp32_kahan_fp64_cond"][3]
    out = {"gaussians": int(mh.shape[0]), "fp64_path_gaussians": int(ill.sum()),
           "float_float_factor_gaussians": int((np.isfinite(fac_def[:, 0]) & ~ill).sum()),
           "fp64_path_gaussians_ratio10": int(np.asarray(pre["fp32_kahan_fp64_ratio10"][2]).sum()),
           "oracle": "FP64 sums with FP64 conics, the mode's own (bit-exact) predicate", "sets": {}}
    for set_name, pts_t in (("uniform", uniform), ("inside_fp64_path_boxes", torch.tensor(stress, device=dev))):
        pts = pts_t.cpu().numpy()
        with torch.no_grad():
            ss = mx.SampleSet(pts_t)
            gpu = {}
            for name, prm in modes.items():
                cells, cw, _, st = mx.mixture_forward(m, sc, rq, vv, wv, 1.0, ss, prm, want_conics=False)
                gpu[name] = (cells.cpu().numpy(), cw.cpu().numpy(), st.stats()[1])
        ref = {name: oracle.forward(mh, vh, wh, pre[name][0], pre[name][1], pts, cd,
                                    factor=pre[name][3] if modes[name].fp64_cond else None) for name in modes}
        res = {"points": int(pts.shape[0]), "modes": {}}
        for name, (V, Wg, pairs) in gpu.items():
            Vr, Wr, cnt = ref[name]
            ok = (Wr > 0) & (Wg > 0)
            err = np.maximum(np.abs(V - Vr) / np.maximum(np.abs(Vr), 1e-3),
                             np.abs(Wg - Wr) / np.maximum(np.abs(Wr), 1e-3))
            err = np.where(ok, err, 0.0)
            res["modes"][name] = {
                "field_max_rel": float(err.max(initial=0.0)),
                "field_p999_rel": float(np.quantile(err[ok], 0.999)) if ok.any() else 0.0,
                "sentinel_flips": int(np.sum((Wg > 0) != (Wr > 0))),
                "pairs": int(pairs), "pairs_match_oracle": int(pairs) == int(cnt.sum())}
        out["sets"][set_name] = res
    out["gradients"] = precision_gradients(W, modes, pre, cd, (mh, sh, rh, wh, vh), (m, sc, rq, wv, vv), uniform,
                                           ill, args.seed)
    out["seconds"] = time.perf_counter() - t0
    log(f"precision study: {json.dumps(out['sets'])}")
    log(f"precision study gradients: {json.dumps(out['gradients'])}")
    return out


def step_gradients(dev_params, samples, host, conic, box, fac, params, n_check=256, seed=2):
    """Gradient error of one forward + backward over the step's full point set
    (random upstream gradients) against the FP64 oracle backward, for n_check
    random Gaussians.  A Gaussian's gradient sees only the points inside its box,
    so the oracle's V and W are evaluated at every step point inside the checked
    Gaussians' boxes (gathered on the device), against all Gaussians.  Error per
    parameter tensor: max |g - g*| / max |g*| over the checked Gaussians."""
    import numpy as np
    import torch

    import oracle
    from gvc import mixture as mx

    m, sc, rq, wv, vv = dev_params
    mh, sh, rh, wh, vh = host
    P, S, dev = mh.shape[0], samples.shape[0], samples.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    gV = torch.randn(S, device=dev, generator=gen)
    gW = torch.randn(S, device=dev, generator=gen)
    with torch.no_grad():
        ss = mx.SampleSet(samples)
        cells, cw, _, st = mx.mixture_forward(m, sc, rq, vv, wv, 1.0, ss, params, want_conics=False)
        grads = mx.mixture_backward(st, m, sc, rq, vv, wv, 1.0, ss, cells, cw, gV, gW)
        sub = np.random.default_rng(seed).choice(P, n_check, replace=False)
        bx = torch.as_tensor(box[sub], device=dev)
        inside = torch.zeros(S, dtype=torch.bool, device=dev)
        for k in range(n_check):
            inside |= ((samples >= bx[k, :3]) & (samples <= bx[k, 3:])).all(dim=1)
        idx = inside.nonzero().squeeze(1)
    pts = samples[idx].cpu().numpy()
    gv, gw = gV[idx].double().cpu().numpy(), gW[idx].double().cpu().numpy()
    Vr, Wr, _ = oracle.forward(mh, vh, wh, conic, box, pts, oracle.conic_fp64(sh, rh), factor=fac)
    ref = oracle.backward(mh[sub], sh[sub], rh[sub], vh[sub], wh[sub], conic[sub], box[sub], pts, Vr, Wr, gv, gw,
                          factor=fac[sub] if fac is not None else None)
    errs = {}
    for n, t in zip(["means", "scales", "rotations", "values", "weights"], grads):
        a = t.detach().cpu().numpy().reshape(P, -1)[sub].astype(np.float64)
        b = ref[n].reshape(n_check, -1)
        errs[n] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
    return {"gaussians": int(n_check), "points": int(pts.shape[0]), "max_rel_by_tensor": errs,
            "max_rel": max(errs.values())}


def precision_gradients(W, modes, pre, cd, host, dev_params, pts_t, ill, seed, n_each=256):
    """Gradient error of each Kahan mode against the FP64 oracle backward (the
    mode's own predicate, FP64 conics) for a forward + backward over the uniform
    point set with random upstream gradients.  A Gaussian's gradient depends only
    on its own parameters and on V, W and the upstream gradients at the points, so
    the oracle evaluates a subset: n_each Gaussians each from the FP64-path set
    (scale ratio above 1000), the ratio-35 set and the rest.  Error per parameter
    tensor: max |g - g*| / max |g*| over the subset, as in the parity tests."""
    import numpy as np
    import torch

    import oracle
    from gvc import mixture as mx

    mh, sh, rh, wh, vh = host
    m, sc, rq, wv, vv = dev_params
    rng = np.random.default_rng(seed + 13)
    deg = np.zeros(mh.shape[0], bool)
    deg[np.asarray(W["degenerate"])] = True
    sub = np.concatenate([rng.choice(np.flatnonzero(mask), n_each, replace=False)
                          for mask in (ill, deg, ~(ill | deg)) if mask.any()])
    pts = pts_t.cpu().numpy()
    gV = rng.normal(size=pts.shape[0]).astype(np.float32)
    gW = rng.normal(size=pts.shape[0]).astype(np.float32)
    res = {"gaussians": int(sub.size), "points": int(pts.shape[0]), "modes": {}}
    names = ["means", "scales", "rotations", "values", "weights"]
    for name, prm in modes.items():
        if not prm.kahan:
            continue
        fac = pre[name][3] if prm.fp64_cond else None
        with torch.no_grad():
            ss = mx.SampleSet(pts_t)
            cells, cw, _, st = mx.mixture_forward(m, sc, rq, vv, wv, 1.0, ss, prm, want_conics=False)
            grads = mx.mixture_backward(st, m, sc, rq, vv, wv, 1.0, ss, cells, cw,
                                        torch.as_tensor(gV, device=pts_t.device),
                                        torch.as_tensor(gW, device=pts_t.device))
        Vr, Wr, _ = oracle.forward(mh, vh, wh, pre[name][0], pre[name][1], pts, cd, factor=fac)
        ref = oracle.backward(mh[sub], sh[sub], rh[sub], vh[sub], wh[sub], pre[name][0][sub], pre[name][1][sub], pts,
                              Vr, Wr, gV.astype(np.float64), gW.astype(np.float64),
                              factor=fac[sub] if fac is not None else None)
        errs = {}
        for n, t in zip(names, grads):
            a = t.detach().cpu().numpy().reshape(mh.shape[0], -1)[sub].astype(np.float64)
            b = ref[n].reshape(sub.size, -1)
            errs[n] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
        res["modes"][name] = {"max_rel_by_tensor": errs, "max_rel": max(errs.values())}
    return res


def main_eval(args):
    """The volume evaluation of R/volume_metrics.py:186-370 (grid branch) on the
    mesh workload's 1,000,000-Gaussian model: a jittered 128^3 grid of the
    model's box drawn on the host as the reference draws it, the ground truth by
    exact tet location, the mixture forward, and every metric from one FP64 pass.
    Timed per evaluation (the grid is drawn once; its upload, the locate, the
    sample set, the forward and the metrics are in the timed region)."""
    import numpy as np
    import torch

    from gvc import evaluation as ev
    from gvc import workloads
    from gvc.sampling import TetMeshSampler

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    W = workloads.mesh_workload(dev, seed=args.seed)
    sampler = TetMeshSampler(W["mesh_points"], W["tets"], W["mesh_values"], device=dev)
    model = make_model(W, args.max_scale, dev)
    with torch.no_grad():
        sc, rq, w, v = model.activated()
    mins = model.get_xyz.detach().min(0).values.cpu().numpy()
    maxes = model.get_xyz.detach().max(0).values.cpu().numpy()
    pts, spacing = ev.evaluation_grid(mins, maxes, 128, np.random.RandomState(args.seed))
    pts32 = torch.as_tensor(pts.astype(np.float32))
    stages = {}

    def one():
        ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        ev0[0].record()
        samples = pts32.to(dev, non_blocking=False)
        ev0[1].record()
        gt = sampler.sample(samples)[0].float()
        ev0[2].record()
        cells_m, cells = ev.evaluate(model.get_xyz.detach(), sc, rq, v, w, samples, gt)
        ev0[3].record()
        torch.cuda.synchronize()
        for name, a, b in (("upload", 0, 1), ("ground_truth", 1, 2), ("forward_and_metrics", 2, 3)):
            stages.setdefault(name, []).append(ev0[a].elapsed_time(ev0[b]))
        return cells_m

    for _ in range(max(args.warmup, 1)):
        one()
    stages.clear()
    steps = args.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m = one()
    wall = (time.perf_counter() - t0) / steps
    S = pts.shape[0]
    out = {"metric": "volume evaluation, sampled points per second (128^3 grid, locate + forward + metrics)",
           "value": S / wall, "unit": "points/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * wall, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "evaluation of the 1,000,000-Gaussian mesh model on a jittered 128^3 grid",
                      "points": S, "parallelism": "single GPU"},
           "stage_ms": {k: float(np.mean(x)) for k, x in stages.items()},
           "metrics": {k: (float(x) if isinstance(x, (float, np.floating)) else x) for k, x in m.items()}}
    print(json.dumps(out))


def main_train(args):
    """The fitting loop (gvc.trainer.Trainer, R/train.py:343-597) on the mesh
    workload: 1,000,000 Gaussians, 4 precomputed batches of 2^22 mesh-sampled
    points, the reference's learning rates, densification (error-ordered
    densify_in_empty + weight prune) every 100 iterations from iteration 100 with
    cap_max 1,100,000.  Steady iterations are graph replays; a densification
    iteration runs eagerly and the next one captures again.  Timed: iterations
    warmup+1 .. warmup+steps (the default 300 include 3 densifications)."""
    import torch

    from gvc import workloads
    from gvc.model import OptimizationParams
    from gvc.sampling import TetMeshSampler
    from gvc.trainer import LoopArgs, Trainer

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t_setup = time.perf_counter()
    W = workloads.mesh_workload(dev, seed=args.seed)
    S, B = 1 << 22, 4
    sampler = TetMeshSampler(W["mesh_points"], W["tets"], W["mesh_values"], device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed + 1)
    samples = torch.rand((B, S, 3), device=dev, generator=gen)
    gt = torch.stack([sampler.sample(samples[b])[0].float() for b in range(B)])
    del sampler
    P0 = W["P"]
    cap = 1_100_000
    model = make_model(W, args.max_scale, dev, capacity=cap + 500_000)
    warm = max(args.warmup, 2)
    steps = args.steps if args.steps != 20 else 300
    opt = OptimizationParams(iterations=warm + steps + 1, densify_from_iter=100, densification_interval=100,
                             densify_until_iter=warm + steps)
    model.training_setup(opt)
    trainer = Trainer(model, samples, gt, opt, LoopArgs(cap_max=cap, min_weight=0.005), graph=True, seed=args.seed)
    log(f"setup {time.perf_counter() - t_setup:.1f}s P={P0} S={S} batches={B}")
    dens_ms = []
    orig = model.densify_from_errors

    def timed_densify(*a, **k):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = orig(*a, **k)
        torch.cuda.synchronize()
        dens_ms.append(1e3 * (time.perf_counter() - t))
        log(f"densified: k={out}, Gaussians {model.count}")
        return out

    model.densify_from_errors = timed_densify
    trainer.run(1, warm)
    torch.cuda.synchronize()
    dens_ms.clear()
    r0, c0 = trainer.replays, trainer.captures
    t0 = time.perf_counter()
    trainer.run(warm + 1, warm + steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {
        "metric": "fitting-loop iterations/s (fwd+bwd+Adam, densification every 100 iterations)",
        "value": steps / wall,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "fitting loop on the unstructured mesh fit: 1,000,000 initial Gaussians, 4 batches of "
                               "2^22 mesh-sampled points, densify every 100 iterations, cap_max 1,100,000",
                   "parallelism": "single GPU"},
        "densifications": len(dens_ms),
        "densify_ms": [round(x, 2) for x in dens_ms],
        "graph_replays": trainer.replays - r0,
        "graph_captures": trainer.captures - c0,
        "gaussians_final": model.count,
    }
    print(json.dumps(out))


def main_batch(args):
    """32 independent volumes of uneven size (sides log-uniform in [32, 256],
    Gaussians log-uniform in [1e4, 2e6], 2^24 points per step split in
    proportion to voxel count).  The fused path fits all of them per step with
    one batched forward and one batched backward (Kahan sums, deterministic
    reduction); the plain sequence runs the same steps with the naive per-volume
    kernels (one wavefront per Gaussian over a 256-point-leaf sample BVH, float
    atomics, separate normalise, no compensation), volume after volume.  Both
    use the same activations, loss and Adam, and both legs are timed from the
    same training state; the report gives both steps/s."""
    out = batch32_legs(args, args.steps, args.plain_steps, args.warmup)
    print(json.dumps(out))


def batch32_legs(args, steps, plain_steps, warmup):
    import numpy as np
    import torch

    from gvc import batched as bx
    from gvc import mixture as mx
    from gvc import training
    from gvc import workloads
    import plain_baseline.plain as plain

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t_setup = time.perf_counter()
    W = workloads.batch_workload(dev, seed=args.seed)
    V = len(W["gaussians"])
    P, S = W["P"], W["S"]
    log(f"batch32 setup {time.perf_counter() - t_setup:.1f}s volumes={V} P={P} S={S}")
    params = mx.MixtureParams()
    if args.leaf or args.block or args.recompute:
        params = mx.MixtureParams(leaf_size=args.leaf or params.leaf_size,
                                  block_threads=args.block or params.block_threads,
                                  candidate_log=not args.recompute)
    tune = None
    if args.autotune if args.config == "batch32" else not args.no_batch_autotune:
        from gvc.autotune import autotune_batched
        with torch.no_grad():
            wts, vls = W["weights"].reshape(-1, 1).contiguous(), W["values"].reshape(-1, 1).contiguous()
            best, table = autotune_batched(W["gaussians"], W["samples"], W["means"], torch.exp(W["log_scales"]),
                                           W["rotations"], vls, wts, W["point_batches"][0], base=params,
                                           use_cache=False)
        params = best
        tune = {"chosen": best.describe(), "ms_fwd_bwd": {k: (round(v, 3) if v is not None else None)
                                                          for k, v in table.items()}}
        log(f"autotune: {tune['chosen']}")
    batch = bx.Batch(W["gaussians"], W["samples"], leaf_size=params.leaf_size, device=dev)
    g_off = batch.P_off
    nb = len(W["point_batches"])

    audit = []
    acc = torch.zeros(3, dtype=torch.int64, device=dev)
    plain_pairs = torch.zeros(1, dtype=torch.int64, device=dev)

    def fused_step(k, model):
        pts, gt = W["point_batches"][k % nb], W["ground_truth"][k % nb]
        audit.clear()
        ss = bx.BatchedSampleSet(batch, pts)
        scales, rots, w, v = model.activated()
        cells, cw = bx.rasterize_batch(model.get_xyz, scales, rots, v, w, ss, params=params)
        training.fit_loss(cells, cw, gt, 1.0).backward()
        model.update_learning_rate()
        model.optimizer_step()
        model.zero_grad()
        audit[-1].accumulate(acc)

    def plain_step(k, model):
        pts, gt = W["point_batches"][k % nb], W["ground_truth"][k % nb]
        trees = [plain.PlainSampleTree(pts[batch.sample_slice(v)]) for v in range(V)]
        scales, rots, w, v = model.activated()
        cells, cw = plain.plain_sequence(model.get_xyz, scales, rots, v, w, g_off, trees, plain_pairs)
        training.fit_loss(cells, cw, gt, 1.0).backward()
        model.update_learning_rate()
        model.optimizer_step()
        model.zero_grad()

    # fused: warm-up sizes the record capacity, then a stage profile; the
    # training state after it is where both legs' timed steps start
    fp = make_model(W, args.max_scale, dev)
    mx.set_audit(audit)
    mx.set_record_capacity(None)
    max_rec = 0
    lane_groups = 0
    for k in range(warmup):
        fused_step(k, fp)
        torch.cuda.synchronize()
        max_rec = max(max_rec, audit[-1].stats()[0])
        lane_groups = audit[-1].lane_groups
    mx.set_record_capacity(int(max_rec * 1.25) + 4096)
    timer = []
    mx.set_stage_timer(timer)
    for k in range(2):
        fused_step(k, fp)
    torch.cuda.synchronize()
    mx.set_stage_timer(None)
    stage = {}
    for name, e0, e1 in timer:
        stage.setdefault(name, []).append(e0.elapsed_time(e1))
    stage = {k: float(np.mean(v)) for k, v in stage.items()}
    pp = make_model(W, args.max_scale, dev)
    pp.load_state(fp)                      # the plain leg starts from the same state
    acc.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        fused_step(k, fp)
    torch.cuda.synchronize()
    fused_s = (time.perf_counter() - t0) / steps
    mx.set_audit(None)
    pairs, records, errs = (int(x) for x in acc.cpu().tolist())
    if errs:
        raise RuntimeError(f"mixture error flags during the timed steps: {errs}")
    fused_pairs = pairs / steps
    log(f"fused {1e3 * fused_s:.2f} ms/step, {fused_pairs:.4g} pairs/step")
    del fp
    mx.set_record_capacity(None)
    torch.cuda.empty_cache()

    plain_pairs.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(plain_steps):
        plain_step(k, pp)
        torch.cuda.synchronize()
        log(f"plain step {k} done")
    plain_s = (time.perf_counter() - t0) / plain_steps
    plain_pairs_step = int(plain_pairs.item()) / plain_steps
    log(f"plain {1e3 * plain_s:.2f} ms/step, {plain_pairs_step:.4g} pairs/step")
    del pp
    torch.cuda.empty_cache()

    return {
        "metric": "steps/s, 32-volume batched fit fwd+bwd",
        "value": 1.0 / fused_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * fused_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{V} volumes, sides log-uniform in [32, 256], Gaussians log-uniform in "
                               f"[1e4, 2e6] (realised {P}), {S} points per step in proportion to voxel count",
                   "sides": [int(x) for x in W["sides"]], "gaussians": [int(x) for x in W["gaussians"]],
                   "points": [int(x) for x in W["samples"]], "leaf_size": params.leaf_size,
                   "block_threads": params.block_threads, "run_points": params.run_points, "kahan": params.kahan,
                   "candidate_log": params.candidate_log, "parallelism": "single GPU"},
        "pairs_per_step": fused_pairs,
        "records_per_step": records / steps,
        "groups": int(batch.c.groups),
        "groups_walked_per_lane": lane_groups,
        "evals_per_s": 2.0 * fused_pairs / fused_s,
        "stage_ms": stage,
        "autotune": tune,
        "plain_sequence": {"steps_per_s": 1.0 / plain_s, "ms_per_step": 1e3 * plain_s, "steps": plain_steps,
                           "pairs_per_step": plain_pairs_step,
                           "kernels": "per volume: preprocess, Morton sort, 256-point-leaf sample tree, one "
                                      "wavefront per Gaussian with float atomics, normalise; backward walks again"},
        "same_start_state": True,
        "speedup_vs_plain": plain_s / fused_s,
    }


if __name__ == "__main__":
    main()
