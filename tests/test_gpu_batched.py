"""Batches of independent volumes: one launch per pass over V volumes.

Bars: bit-exact.  Every volume's cells, cell weights, conics and gradients in a
batch equal the single-volume calls on that volume's inputs, bit for bit, and
pair counts add up exactly.  Each volume's batched results are also checked
against the FP64 oracle directly (the precise path's factors included): pair
count exact, field within 1e-5 relative, gradients within 1e-5 normwise.
"""
import numpy as np
import pytest

from conftest import assert_grads, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

import oracle  # noqa: E402

from gvc import batched as bx  # noqa: E402
from gvc import mixture as mx  # noqa: E402
from gvc import workloads  # noqa: E402

f32 = np.float32
DEV = "cuda:0"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, torch.float32)


def bits(t):
    return t.detach().cpu().numpy().view(np.uint32)


def make_volume(seed, P, S, scale_mul=1.0, shift=0.0):
    rng = np.random.default_rng(seed)
    g = workloads.random_gaussians(rng, P)
    g["scales"] = (np.exp(g["log_scales"]) * scale_mul).astype(f32)
    g["means"] = (g["means"] + shift).astype(f32)
    g["samples"] = rng.uniform(0.0, 1.0, (S, 3)).astype(f32)
    return g


VOLUMES = [  # (seed, P, S, scale multiplier, shift)
    (1, 3000, 5000, 1.0, 0.0),
    (2, 1, 70, 8.0, 0.0),
    (3, 800, 0, 1.0, 0.0),
    (4, 5000, 12000, 0.5, 0.0),
    (5, 1200, 3000, 2.0, 0.25),
]


def single(g, params, gV, gW):
    ss = mx.SampleSet(T(g["samples"]))
    args = (T(g["means"]), T(g["scales"]), T(g["rotations"]), T(g["values"]), T(g["weights"]))
    cells, cw, conics, st = mx.mixture_forward(*args, 1.0, ss, params)
    nrec, npairs, err = st.stats()
    assert err == 0
    grads = mx.mixture_backward(st, *args, 1.0, ss, cells, cw, gV, gW)
    return [cells, cw, conics] + list(grads), npairs


def batched(vols, params, gVs, gWs):
    batch = bx.Batch([g["means"].shape[0] for g in vols], [g["samples"].shape[0] for g in vols],
                     leaf_size=params.leaf_size)
    cat = {k: np.concatenate([g[k] for g in vols]) for k in ("means", "scales", "rotations", "values", "weights",
                                                              "samples")}
    ss = bx.BatchedSampleSet(batch, T(cat["samples"]))
    args = (T(cat["means"]), T(cat["scales"]), T(cat["rotations"]), T(cat["values"]), T(cat["weights"]))
    cells, cw, conics, st = bx.mixture_forward_batched(batch, *args, 1.0, ss, params)
    nrec, npairs, err = st.stats()
    assert err == 0
    gV = torch.cat(gVs) if gVs else None
    gW = torch.cat(gWs) if gWs else None
    grads = bx.mixture_backward_batched(st, *args, 1.0, ss, cells, cw, gV, gW)
    return batch, [cells, cw, conics] + list(grads), npairs


def upstream(vols, seed):
    rng = np.random.default_rng(seed)
    gVs = [T(rng.normal(size=g["samples"].shape[0])) for g in vols]
    gWs = [T(rng.normal(size=g["samples"].shape[0])) for g in vols]
    return gVs, gWs


def check_against_oracle(g, params, outs, gV, gW):
    """One volume's batched cells and gradients against the FP64 oracle (with the
    precise path's factors when params.fp64_cond)."""
    if g["means"].shape[0] == 0:
        return
    conic, box, _, fac = oracle.preprocess(g["means"], g["scales"], g["rotations"], g["weights"], 1.0,
                                           params.fp64_cond, params.fp64_scale_ratio, factor=True,
                                           factor_ratio=params.factor_scale_ratio)
    fac = fac if params.fp64_cond else None
    cd = oracle.conic_fp64(g["scales"], g["rotations"])
    Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic, box, g["samples"], cd, factor=fac)
    V, W = outs[0].cpu().numpy(), outs[1].cpu().numpy()
    assert np.array_equal(W > 0, Wr > 0)
    ok = Wr > 0
    err = np.maximum(np.abs(V[ok] - Vr[ok]) / np.maximum(np.abs(Vr[ok]), 1e-3),
                     np.abs(W[ok] - Wr[ok]) / np.maximum(np.abs(Wr[ok]), 1e-3))
    assert err.max(initial=0.0) <= 1e-5
    ref = oracle.backward(g["means"], g["scales"], g["rotations"], g["values"], g["weights"], conic, box,
                          g["samples"], Vr, Wr, gV.double().cpu().numpy(), gW.double().cpu().numpy(), factor=fac)
    if any(np.abs(ref[n]).max() > 0 for n in ("means", "values")):
        assert_grads(outs[2:], ref)
    return int(cnt.sum())


def check_against_singles(vols, params, seed=11):
    gVs, gWs = upstream(vols, seed)
    batch, outs, npairs = batched(vols, params, gVs, gWs)
    total = 0
    for v, g in enumerate(vols):
        ss_, gs_ = batch.sample_slice(v), batch.gaussian_slice(v)
        n_or = check_against_oracle(g, params, [outs[0][ss_], outs[1][ss_]] + [t[gs_] for t in outs[3:]],
                                    gVs[v], gWs[v])
        ref, np_v = single(g, params, gVs[v], gWs[v])
        assert n_or is None or n_or == np_v, f"pairs of volume {v}"
        total += np_v
        ss, gs = batch.sample_slice(v), batch.gaussian_slice(v)
        assert np.array_equal(bits(outs[0][ss]), bits(ref[0])), f"cells of volume {v}"
        assert np.array_equal(bits(outs[1][ss]), bits(ref[1])), f"cell weights of volume {v}"
        assert np.array_equal(bits(outs[2].view(-1, 6)[gs]), bits(ref[2].view(-1, 6))), f"conics of volume {v}"
        for i, name in enumerate(("means", "scales", "rotations", "values", "weights")):
            assert np.array_equal(bits(outs[3 + i][gs]), bits(ref[3 + i])), f"d{name} of volume {v}"
    assert npairs == total


@pytest.mark.parametrize("log", [True, False])
def test_batch_equals_single_calls(log):
    vols = [make_volume(*spec) for spec in VOLUMES]
    check_against_singles(vols, mx.MixtureParams(candidate_log=log))


@pytest.mark.parametrize("fp64_ratio", [10.0, 0.0])
def test_batch_precise_path_equals_single_calls(fp64_ratio):
    """FP64 mode with Gaussians on the precise path in every volume (a tenth made
    anisotropic, one scale divided by 3500 or 35, randomly rotated; FP64
    threshold at scale ratio 10, or the default 1000 with the float-float factor
    for the ratio-35 ones): the batched launch reads each volume's factors at its
    own Gaussians' entries, bit-identical to one call per volume."""
    vols = [make_volume(*spec) for spec in VOLUMES]
    for k, g in enumerate(vols):
        rng = np.random.default_rng(100 + k)
        P = g["means"].shape[0]
        pick = rng.permutation(P)[:max(1, P // 10)]
        ax = rng.integers(0, 3, pick.size)
        div = np.where(np.arange(pick.size) % 2 == 0, 3500.0, 35.0)
        g["scales"][pick, ax] = (g["scales"][pick, ax] / div).astype(f32)
        q = rng.normal(size=(pick.size, 4))
        g["rotations"][pick] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(f32)
    check_against_singles(vols, mx.MixtureParams(fp64_cond=True, fp64_scale_ratio=fp64_ratio))


@pytest.mark.parametrize("leaf,block", [(2, 512), (8, 1024)])
def test_batch_launch_shapes(leaf, block):
    vols = [make_volume(*spec) for spec in VOLUMES[:2] + VOLUMES[3:]]
    check_against_singles(vols, mx.MixtureParams(leaf_size=leaf, block_threads=block))


def test_batch_of_one_equals_single():
    g = workloads.config0(2)
    g["scales"] = np.exp(g["log_scales"]).astype(f32)
    check_against_singles([g], mx.MixtureParams())


def test_batch_of_many_small_volumes():
    """32 volumes of uneven size, as in a batched fit of many small meshes."""
    rng = np.random.default_rng(21)
    vols = [make_volume(100 + v, int(rng.integers(1, 600)), int(rng.integers(0, 900)),
                        float(np.exp(rng.uniform(-0.5, 1.0)))) for v in range(32)]
    check_against_singles(vols, mx.MixtureParams())


def test_batch_autograd_and_errors():
    vols = [make_volume(*spec) for spec in VOLUMES[:2]]
    batch = bx.Batch([g["means"].shape[0] for g in vols], [g["samples"].shape[0] for g in vols])
    cat = {k: np.concatenate([g[k] for g in vols]) for k in ("means", "scales", "rotations", "values", "weights",
                                                              "samples")}
    ss = bx.BatchedSampleSet(batch, T(cat["samples"]))
    leaves = [T(cat[k]).requires_grad_() for k in ("means", "scales", "rotations", "values", "weights")]
    cells, cw = bx.rasterize_batch(*leaves, ss)
    (cells.square().sum() + cw.sum()).backward()
    gV, gW = 2 * cells.detach(), torch.ones_like(cw)
    _, _, _, st = bx.mixture_forward_batched(batch, *[t.detach() for t in leaves], 1.0, ss)
    ref = bx.mixture_backward_batched(st, *[t.detach() for t in leaves], 1.0, ss, cells.detach(), cw.detach(), gV, gW)
    for t, r in zip(leaves, ref):
        assert np.array_equal(bits(t.grad.view(-1)), bits(r.view(-1)))
    with pytest.raises(ValueError):
        bx.mixture_forward_batched(batch, *[t.detach() for t in leaves], 1.0, ss, mx.MixtureParams(leaf_size=8))
    with pytest.raises(ValueError):
        bx.mixture_forward_batched(batch, *[t.detach()[:-1] for t in leaves], 1.0, ss)
