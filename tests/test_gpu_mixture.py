"""GPU parity of the mixture path (libgvc through its C ABI) against the oracle.

Bars (stated per test): bit-exact for the integer and predicate work (Morton
codes, conics and boxes, pair membership counts, sentinel decisions); field
within 1e-5 relative (max |V - V*| / max(|V*|, 1e-3)) of the FP64 oracle;
gradients within 1e-5 of the FP64 oracle, measured against the largest
gradient magnitude of each parameter tensor (conftest.grad_errors also reports
the elementwise error, floored at 1e-3 of that magnitude).
"""
import numpy as np
import pytest

from conftest import GRAD_BAR, assert_grads, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

import oracle  # noqa: E402
from gvc import mixture as mx  # noqa: E402
from gvc import workloads  # noqa: E402

f32 = np.float32
DEV = "cuda:0"


def T(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, dtype)


def field_err(V, W, Vr, Wr):
    valid = Wr > 0
    ev = np.abs(V[valid] - Vr[valid]) / np.maximum(np.abs(Vr[valid]), 1e-3)
    ew = np.abs(W[valid] - Wr[valid]) / np.maximum(np.abs(Wr[valid]), 1e-3)
    return max(ev.max(initial=0), ew.max(initial=0))


def edge_gaussians(seed=0, P=4000):
    rng = np.random.default_rng(seed)
    g = workloads.random_gaussians(rng, P)
    scales = np.exp(g["log_scales"]).astype(f32)
    w = g["weights"].copy()
    w[:20, 0] = [0.0, 1e-4, 9.99e-4, 1e-3, 1.0001e-3, 1.0, 2.0, 0.011, 0.0105, 1e-2,
                 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5]
    scales[10] = 0.0                       # degenerate: zero covariance
    scales[11] = [0.03, 1e-5, 0.02]        # cond ~ 1e7
    scales[12] = [1e-6, 1e-6, 1e-6]
    rots = g["rotations"].copy()
    rots[13] = [0, 0, 0, 0]                # zero quaternion
    rots[14] = [2.0, 0.5, -1.0, 0.1]       # unnormalised
    means = g["means"].copy()
    means[15] = [-0.2, 1.3, 0.5]           # outside the unit cube
    means[16] = [np.nan, 0.5, 0.5]
    scales[17] = [np.inf, 0.01, 0.01]
    return means, scales, rots, g["values"], w


@pytest.mark.parametrize("fp64_cond", [False, True])
def test_preprocess_bit_exact(fp64_cond):
    means, scales, rots, values, w = edge_gaussians(1)
    c_ref, b_ref, _ = oracle.preprocess(means, scales, rots, w, 1.0, fp64_cond)
    c, b = mx.preprocess(T(means), T(scales), T(rots), T(w), 1.0, fp64_cond)
    c, b = c.cpu().numpy(), b.cpu().numpy()
    # bit-exact (NaN positions must agree too)
    assert np.array_equal(c.view(np.uint32), c_ref.view(np.uint32)) or np.array_equal(c, c_ref, equal_nan=True)
    assert np.array_equal(b, b_ref, equal_nan=True)
    c2, b2 = mx.preprocess(T(means), T(scales), T(rots), T(w), 0.7, fp64_cond)
    c2r, b2r, _ = oracle.preprocess(means, scales, rots, w, 0.7, fp64_cond)
    assert np.array_equal(c2.cpu().numpy(), c2r, equal_nan=True)
    assert np.array_equal(b2.cpu().numpy(), b2r, equal_nan=True)


@pytest.mark.parametrize("fp64_ratio,factor_ratio", [(0.0, 0.0), (10.0, 0.0), (0.0, 1000.0), (200.0, 3.0)])
def test_precise_factors_bit_exact(fp64_ratio, factor_ratio):
    """The precise path's conic factors (FP64 above fp64_ratio, float-float
    between factor_ratio and fp64_ratio) bit-identical to the oracle's, NaN rows
    and edge cases included."""
    means, scales, rots, values, w = edge_gaussians(7)
    rng = np.random.default_rng(8)
    pick = rng.permutation(np.arange(30, means.shape[0]))[:1200]
    ax = rng.integers(0, 3, pick.size)
    div = np.asarray([4.0, 12.0, 35.0, 300.0, 999.0, 3500.0])[np.arange(pick.size) % 6]
    scales[pick, ax] = (scales[pick, ax] / div).astype(f32)
    prm = mx.MixtureParams(fp64_cond=True, fp64_scale_ratio=fp64_ratio, factor_scale_ratio=factor_ratio)
    c, b, fac = mx.preprocess_factors(T(means), T(scales), T(rots), T(w), prm)
    c_ref, b_ref, ill, fac_ref = oracle.preprocess(means, scales, rots, w, 1.0, True, fp64_ratio, factor=True,
                                                   factor_ratio=factor_ratio)
    assert np.array_equal(c.cpu().numpy(), c_ref, equal_nan=True)
    assert np.array_equal(b.cpu().numpy(), b_ref, equal_nan=True)
    got = fac.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), fac_ref.view(np.uint32)) or np.array_equal(got, fac_ref,
                                                                                            equal_nan=True)
    on = np.isfinite(fac_ref[:, 0])
    band = on & ~ill
    print(f"factor rows: FP64 {int((on & ill).sum())}, float-float {int(band.sum())}")
    if (factor_ratio or 10.0) < (fp64_ratio or 1000.0):
        assert band.sum() > 100
    else:
        assert band.sum() == 0


def test_morton_bit_exact():
    rng = np.random.default_rng(2)
    pts = np.concatenate([rng.uniform(-0.1, 1.1, (100000, 3)),
                          [[0, 0, 0], [1, 1, 1], [np.nan, 0.5, 2.0], [1e30, -1e30, 0.5],
                           [np.nextafter(1, 0), 0.5, 0.25]]]).astype(f32)
    for clamp in (False, True):
        ref = oracle.morton3d_unit(pts, clamp)
        got = mx.morton3d_unit(T(pts), clamp).cpu().numpy().view(np.uint64)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("n,bits", [(1, 64), (1000, 64), (70001, 64), (300000, 21), (5000, 8)])
def test_sort_pairs_stable(n, bits):
    rng = np.random.default_rng(n)
    if bits == 64:
        keys = rng.integers(0, 2**63 - 1, n, dtype=np.int64)
        keys[: n // 3] = keys[0]  # many equal keys: stability matters
    else:
        keys = rng.integers(0, 1 << bits, n, dtype=np.int64)
    vals = np.arange(n, dtype=np.int32)
    ko, vo = mx.sort_pairs_u64(T(keys, torch.int64), T(vals, torch.int32), bits)
    order = np.argsort(keys.view(np.uint64) & np.uint64((1 << bits) - 1 if bits < 64 else 2**64 - 1),
                       kind="stable")
    assert np.array_equal(vo.cpu().numpy(), vals[order])
    assert np.array_equal(ko.cpu().numpy(), keys[order])


def run_gpu(means, scales, rots, values, w, samples, params, k=1.0):
    ss = mx.SampleSet(T(samples))
    cells, cw, conics, st = mx.mixture_forward(T(means), T(scales), T(rots), T(values), T(w), k, ss, params)
    return ss, cells, cw, conics, st


@pytest.mark.parametrize("leaf,lane", [(1, "auto"), (2, "auto"), (4, "auto"), (8, "auto"), (16, "auto"),
                                       (1, "always"), (4, "always"), (16, "always"), (4, "never")])
def test_forward_config0_parity(leaf, lane):
    g = workloads.config0(0)
    scales = np.exp(g["log_scales"]).astype(f32)
    params = mx.MixtureParams(leaf_size=leaf, block_threads=256, lane_mode=lane)
    ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"],
                                        g["samples"], params)
    V, W = cells.cpu().numpy(), cw.cpu().numpy()
    conic_ref, box_ref, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    assert np.array_equal(conics.cpu().numpy().reshape(-1, 6), conic_ref)
    Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, g["samples"])
    nrec, npairs, err = st.stats()
    assert err == 0
    assert npairs == int(cnt.sum())  # exact pair membership
    assert np.array_equal(W == 0, Wr == 0)  # sentinel agreement
    assert np.all(V[W == 0] == -1.0)
    assert field_err(V, W, Vr, Wr) <= 1e-5


def test_forward_edge_cases():
    means, scales, rots, values, w = edge_gaussians(3, P=3000)
    rng = np.random.default_rng(4)
    samples = np.concatenate([rng.uniform(-0.05, 1.05, (10007, 3)), means[:40] ,
                              [[np.nan, 0.5, 0.5], [0.5, 0.5, 0.5]]]).astype(f32)
    samples[-2] = [np.nan, 0.5, 0.5]
    for leaf, lane in ((1, "auto"), (3, "auto"), (16, "auto"), (3, "always"), (16, "always")):
        params = mx.MixtureParams(leaf_size=leaf, lane_mode=lane)
        ss, cells, cw, conics, st = run_gpu(means, scales, rots, values, w, samples, params)
        conic_ref, box_ref, _ = oracle.preprocess(means, scales, rots, w)
        Vr, Wr, cnt = oracle.forward(means, values, w, conic_ref, box_ref, samples)
        V, W = cells.cpu().numpy(), cw.cpu().numpy()
        nrec, npairs, err = st.stats()
        assert npairs == int(cnt.sum())
        finite = np.isfinite(Vr) & np.isfinite(V)
        assert np.array_equal(np.isfinite(V), np.isfinite(Vr))
        assert np.array_equal((W == 0)[finite], (Wr == 0)[finite])
        assert field_err(V[finite], W[finite], Vr[finite], Wr[finite]) <= 1e-5


def test_empty_and_tiny():
    rng = np.random.default_rng(5)
    samples = rng.uniform(0, 1, (100, 3)).astype(f32)
    z3 = np.zeros((0, 3), f32)
    ss = mx.SampleSet(T(samples))
    cells, cw, conics, st = mx.mixture_forward(T(z3), T(z3), T(np.zeros((0, 4), f32)), T(np.zeros((0, 1), f32)),
                                               T(np.zeros((0, 1), f32)), 1.0, ss)
    assert torch.all(cells == 0) and torch.all(cw == 0)  # P == 0 leaves zeros (reference)
    # one Gaussian, samples not a multiple of 64, no samples
    m = np.array([[0.5, 0.5, 0.5]], f32)
    s = np.array([[0.2, 0.2, 0.2]], f32)
    q = np.array([[1, 0, 0, 0]], f32)
    v = np.array([[0.25]], f32)
    w = np.array([[1.0]], f32)
    ss, cells, cw, _, st = run_gpu(m, s, q, v, w, samples, mx.MixtureParams())
    cr, br, _ = oracle.preprocess(m, s, q, w)
    Vr, Wr, _ = oracle.forward(m, v, w, cr, br, samples)
    assert field_err(cells.cpu().numpy(), cw.cpu().numpy(), Vr, Wr) <= 1e-5
    # the adjoint of the one-Gaussian problem (a one-element Gaussian sort and
    # its inverse permutation; 100 samples: a partly filled second group)
    gV = rng.normal(size=samples.shape[0]).astype(f32)
    gW = rng.normal(size=samples.shape[0]).astype(f32)
    grads = mx.mixture_backward(st, T(m), T(s), T(q), T(v), T(w), 1.0, ss, cells, cw, T(gV), T(gW))
    cd = oracle.conic_fp64(s, q)
    Vd, Wd, _ = oracle.forward(m, v, w, cr, br, samples, cd)
    ref = oracle.backward(m, s, q, v, w, cr, br, samples, Vd, Wd, gV.astype(np.float64), gW.astype(np.float64))
    assert_grads(grads, ref, GRAD_BAR)
    ss0 = mx.SampleSet(T(np.zeros((0, 3), f32)))
    c0, w0, _, _ = mx.mixture_forward(T(m), T(s), T(q), T(v), T(w), 1.0, ss0)
    assert c0.numel() == 0


def _grad_check(g, scales, samples, params, tol=GRAD_BAR, seed=0):
    rng = np.random.default_rng(seed)
    S = samples.shape[0]
    gV = rng.normal(size=S).astype(f32)
    gW = rng.normal(size=S).astype(f32)
    ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"], samples,
                                        params)
    grads = mx.mixture_backward(st, T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]), T(g["weights"]),
                                1.0, ss, cells, cw, T(gV), T(gW))
    conic_ref, box_ref, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    cd = oracle.conic_fp64(scales, g["rotations"])
    Vr, Wr, _ = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, samples, cd)
    ref = oracle.backward(g["means"], scales, g["rotations"], g["values"], g["weights"], conic_ref, box_ref,
                          samples, Vr, Wr, gV.astype(np.float64), gW.astype(np.float64))
    assert_grads(grads, ref, tol)
    return grads


@pytest.mark.parametrize("leaf,log,lane", [(1, True, "auto"), (4, True, "auto"), (16, True, "auto"),
                                           (1, False, "auto"), (4, False, "auto"), (4, True, "always"),
                                           (16, False, "always"), (2, True, "never")])
def test_backward_config0_parity(leaf, log, lane):
    g = workloads.config0(0)
    scales = np.exp(g["log_scales"]).astype(f32)
    # subsample points so the FP64 brute-force oracle stays fast
    samples = g["samples"][::4].copy()
    _grad_check(g, scales, samples, mx.MixtureParams(leaf_size=leaf, candidate_log=log, lane_mode=lane))


def ill_conditioned_set(seed=31, frac=0.15, divisors=(35.0, 3500.0)):
    """Config-0 Gaussians with a fraction made anisotropic (one scale divided by
    each of ``divisors`` in turn) and randomly rotated, and points half from the
    regular grid, half inside the anisotropic Gaussians' boxes."""
    g = workloads.config0(0)
    rng = np.random.default_rng(seed)
    P = g["means"].shape[0]
    scales = np.exp(g["log_scales"]).astype(f32)
    pick = rng.permutation(P)[:int(frac * P)]
    div = np.asarray(divisors)[np.arange(pick.size) % len(divisors)]
    ax = rng.integers(0, 3, pick.size)
    scales[pick, ax] = (scales[pick, ax] / div).astype(f32)
    q = rng.normal(size=(pick.size, 4))
    rots = g["rotations"].copy()
    rots[pick] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(f32)
    _, box, _ = oracle.preprocess(g["means"], scales, rots, g["weights"])
    k = pick[rng.integers(0, pick.size, 4096)]
    inside = box[k, :3] + rng.uniform(0, 1, (4096, 3)) * (box[k, 3:] - box[k, :3])
    samples = np.concatenate([g["samples"][::8], inside]).astype(f32)
    return dict(g, rotations=rots), scales, samples


@pytest.mark.parametrize("ratio,leaf,log,lane", [(0.0, 4, True, "auto"), (0.0, 1, False, "always"),
                                                 (10.0, 4, True, "auto"), (10.0, 1, False, "auto"),
                                                 (10.0, 16, True, "always"), (10.0, 2, True, "never")])
def test_precise_path_parity(ratio, leaf, log, lane):
    """FP64 mode: Gaussians above the scale ratio take the FP64 inverse and their
    pairs are evaluated from the Cholesky factor (compensated fp32); at the
    default threshold (ratio 1000) the ratio-35 Gaussians below it take the same
    path with the factor computed in float-float arithmetic.  Bars: conics
    bit-exact, pair count equal to the oracle's (which evaluates the same factor
    form), field within 1e-5 and gradients within 1e-5 of the FP64 oracle."""
    g, scales, samples = ill_conditioned_set(divisors=(35.0, 3500.0))
    params = mx.MixtureParams(leaf_size=leaf, candidate_log=log, lane_mode=lane, fp64_cond=True,
                              fp64_scale_ratio=ratio)
    ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"], samples,
                                        params)
    conic_ref, box_ref, ill, fac = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"], 1.0, True,
                                                     ratio, factor=True)
    assert np.isfinite(fac[:, 0]).sum() == int(0.15 * 2048)
    assert ill.sum() == (int(0.15 * 2048) if ratio == 10.0 else int(0.15 * 2048) // 2)
    assert np.array_equal(conics.cpu().numpy().reshape(-1, 6), conic_ref)
    cd = oracle.conic_fp64(scales, g["rotations"])
    Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, samples, cd,
                                 factor=fac)
    V, W = cells.cpu().numpy(), cw.cpu().numpy()
    nrec, npairs, err = st.stats()
    assert err == 0
    assert npairs == int(cnt.sum())
    assert np.array_equal(W == 0, Wr == 0)
    assert field_err(V, W, Vr, Wr) <= 1e-5
    rng = np.random.default_rng(1)
    gV = rng.normal(size=samples.shape[0]).astype(f32)
    gW = rng.normal(size=samples.shape[0]).astype(f32)
    grads = mx.mixture_backward(st, T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]), T(g["weights"]),
                                1.0, ss, cells, cw, T(gV), T(gW))
    ref = oracle.backward(g["means"], scales, g["rotations"], g["values"], g["weights"], conic_ref, box_ref,
                          samples, Vr, Wr, gV.astype(np.float64), gW.astype(np.float64), factor=fac)
    assert_grads(grads, ref)


def sparse_and_dense(seed=12, P=40000, dense=1500, sparse=300):
    """Small Gaussians over the unit cube; points in one dense cluster (groups
    walked together) and scattered thinly (groups walked per lane)."""
    rng = np.random.default_rng(seed)
    g = workloads.random_gaussians(rng, P)
    scales = (np.exp(g["log_scales"]) * 0.3).astype(f32)
    samples = np.concatenate([rng.uniform(0.40, 0.45, (dense, 3)), rng.uniform(0.0, 1.0, (sparse, 3))]).astype(f32)
    return g, scales, samples


def test_lane_mode_sparse_groups():
    """Automatic lane mode: the pair set is the same under every policy, fields
    and gradients meet the oracle bars, replay and re-walk agree bit for bit, and
    lane-walked groups write one backward record per pair."""
    g, scales, samples = sparse_and_dense()
    conic_ref, box_ref, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, samples)
    recs, lanes = {}, {}
    NG = (samples.shape[0] + 63) // 64
    for lane in ("never", "auto", "always"):
        ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"], samples,
                                            mx.MixtureParams(lane_mode=lane))
        nrec, npairs, err = st.stats()
        # walked together, the sparse groups outgrow the default candidate log
        # (the eager backward redoes such a forward with a larger one)
        assert (err & ~mx._lib.GVC_ERRFLAG_CHUNK_OVERFLOW) == 0 and npairs == int(cnt.sum())
        V, W = cells.cpu().numpy(), cw.cpu().numpy()
        assert np.array_equal(W == 0, Wr == 0)
        assert field_err(V, W, Vr, Wr) <= 1e-5
        recs[lane], lanes[lane] = nrec, st.lane_groups
    assert recs["always"] == int(cnt.sum())            # per lane: one record per pair
    assert lanes["never"] == 0 and lanes["always"] == NG
    assert 0 < lanes["auto"] < NG                      # only the sparse groups are walked per lane
    for lane in ("auto", "always"):
        _grad_check(g, scales, samples, mx.MixtureParams(lane_mode=lane))
        rng = np.random.default_rng(3)
        gV = T(rng.normal(size=samples.shape[0]))
        gW = T(rng.normal(size=samples.shape[0]))
        outs = []
        for log in (True, False):
            ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"],
                                                samples, mx.MixtureParams(lane_mode=lane, candidate_log=log))
            grads = mx.mixture_backward(st, T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]),
                                        T(g["weights"]), 1.0, ss, cells, cw, gV, gW)
            assert st.stats()[2] == 0
            outs.append([cells.cpu().numpy()] + [t.cpu().numpy() for t in grads])
        for a, b in zip(*outs):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_backward_replay_equals_rewalk():
    """The candidate-log replay and the re-walk give bit-identical gradients."""
    g = workloads.config0(4)
    scales = np.exp(g["log_scales"]).astype(f32)
    samples = g["samples"]
    rng = np.random.default_rng(6)
    gV = T(rng.normal(size=samples.shape[0]))
    gW = T(rng.normal(size=samples.shape[0]))
    outs = []
    for log in (True, False):
        ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"], samples,
                                            mx.MixtureParams(candidate_log=log))
        grads = mx.mixture_backward(st, T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]),
                                    T(g["weights"]), 1.0, ss, cells, cw, gV, gW)
        assert st.stats()[2] == 0
        outs.append([cells.cpu().numpy()] + [t.cpu().numpy() for t in grads])
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_backward_deterministic():
    g = workloads.config0(1)
    scales = np.exp(g["log_scales"]).astype(f32)
    samples = g["samples"]
    rng = np.random.default_rng(3)
    gV = T(rng.normal(size=samples.shape[0]))
    gW = T(rng.normal(size=samples.shape[0]))
    outs = []
    for _ in range(2):
        ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"], samples,
                                            mx.MixtureParams())
        grads = mx.mixture_backward(st, T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]),
                                    T(g["weights"]), 1.0, ss, cells, cw, gV, gW)
        outs.append([cells.cpu().numpy(), cw.cpu().numpy()] + [t.cpu().numpy() for t in grads])
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_groups_straddling_clusters():
    """Samples in 8 tight corner clusters of 37 points: Morton groups straddle
    clusters, so groups are cut into runs; pair membership stays exact."""
    rng = np.random.default_rng(11)
    corners = np.array([[x, y, z] for x in (0.05, 0.95) for y in (0.05, 0.95) for z in (0.05, 0.95)])
    samples = np.concatenate([c + rng.normal(scale=0.01, size=(37, 3)) for c in corners]).astype(f32)
    g = workloads.random_gaussians(rng, 5000)
    scales = np.exp(g["log_scales"]).astype(f32)
    for leaf in (1, 4):
        ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"], samples,
                                            mx.MixtureParams(leaf_size=leaf))
        conic_ref, box_ref, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
        Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, samples)
        nrec, npairs, err = st.stats()
        assert err == 0
        assert npairs == int(cnt.sum())
        assert field_err(cells.cpu().numpy(), cw.cpu().numpy(), Vr, Wr) <= 1e-5
    _grad_check(g, scales, samples, mx.MixtureParams())


def test_candidate_log_overflow_recovers():
    """A forward whose candidate log is too small is redone by the (eager)
    backward with a larger log; the gradients equal those of a roomy log."""
    g = workloads.config0(2)
    scales = np.exp(g["log_scales"]).astype(f32)
    samples = g["samples"]
    rng = np.random.default_rng(8)
    gV = T(rng.normal(size=samples.shape[0]))
    gW = T(rng.normal(size=samples.shape[0]))
    saved = dict(mx._chunk_slots)
    try:
        outs = []
        for slots in (64, 1):
            ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"],
                                                samples, mx.MixtureParams(chunk_slots=slots))
            if slots == 1:
                assert st.stats()[2] & mx._lib.GVC_ERRFLAG_CHUNK_OVERFLOW
            grads = mx.mixture_backward(st, T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]),
                                        T(g["weights"]), 1.0, ss, cells, cw, gV, gW)
            outs.append([t.cpu().numpy() for t in grads])
        for a, b in zip(*outs):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        # with a fixed record capacity nothing is re-run: the overflow is reported
        ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"], samples,
                                            mx.MixtureParams(chunk_slots=1))
        mx.mixture_backward(st, T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]), T(g["weights"]), 1.0,
                            ss, cells, cw, gV, gW, record_capacity=1 << 20)
        assert st.stats()[2] & mx._lib.GVC_ERRFLAG_CHUNK_OVERFLOW
    finally:
        mx._chunk_slots.clear()
        mx._chunk_slots.update(saved)


@pytest.mark.parametrize("leaf,block,run", [(4, 512, 32), (8, 1024, 16), (2, 256, 16), (4, 256, 4), (16, 512, 8)])
def test_launch_shapes_parity(leaf, block, run):
    """Block shape, run length and leaf size change the schedule, never the
    predicate: exact pair counts and oracle-level fields and gradients."""
    g = workloads.config0(5)
    scales = np.exp(g["log_scales"]).astype(f32)
    params = mx.MixtureParams(leaf_size=leaf, block_threads=block, run_points=run)
    ss, cells, cw, conics, st = run_gpu(g["means"], scales, g["rotations"], g["values"], g["weights"],
                                        g["samples"], params)
    conic_ref, box_ref, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, g["samples"])
    nrec, npairs, err = st.stats()
    # short runs log more chunks per group than the default candidate log holds:
    # the eager backward redoes such a forward with a larger log
    assert (err & ~mx._lib.GVC_ERRFLAG_CHUNK_OVERFLOW) == 0 and npairs == int(cnt.sum())
    assert field_err(cells.cpu().numpy(), cw.cpu().numpy(), Vr, Wr) <= 1e-5
    _grad_check(g, scales, g["samples"][::4].copy(), params)


def test_autotune_small_sweep():
    """The autotuner times every configuration of its sweep (leaf size x points
    per walk x block shape), returns the fastest of its own table, and every
    configuration accepts exactly the oracle's pairs."""
    from gvc.autotune import autotune, params_for
    g = workloads.config0(6)
    scales = np.exp(g["log_scales"]).astype(f32)
    ss = mx.SampleSet(T(g["samples"]))
    args = (T(g["means"]), T(scales), T(g["rotations"]), T(g["values"]), T(g["weights"]))
    best, table = autotune(*args, ss, leaf_sizes=(2, 4), points_per_walk=(1, 16, 64), blocks=(256, 512), reps=1,
                           use_cache=False)
    assert len(table) == 12 and all(v is not None and v > 0 for v in table.values())
    label = min(table, key=table.get)
    d = best.describe()
    assert label == f"leaf{d['leaf_size']}/points{d['points_per_walk']}/block{d['block_threads']}"
    conic_ref, box_ref, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    _, _, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic_ref, box_ref, g["samples"])
    for L in (2, 4):
        for ppw in (1, 4, 16, 64):
            prm = params_for(mx.MixtureParams(), L, ppw, 256)
            _, _, _, st = mx.mixture_forward(*args, 1.0, ss, prm, want_conics=False)
            nrec, npairs, err = st.stats()
            assert npairs == int(cnt.sum()), (L, ppw)
            assert err & ~mx._lib.GVC_ERRFLAG_CHUNK_OVERFLOW == 0
