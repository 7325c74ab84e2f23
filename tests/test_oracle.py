"""Pins the CPU oracle (oracle/) before anything is checked against it.

Every function of the oracle is checked here against something other than its
own formula: closed forms, library routines (math.log, numpy.linalg.inv), torch
float64 autograd, central differences, brute force, and an independently written
tree builder.  One test plants the reference's own (defective) covariance chain
(CR/backward.cu:68-124) in the oracle's place and checks that the autograd
comparison rejects it -- so the gradient check would catch that class of error.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import autograd_ref
from gvc import workloads

f32 = np.float32


def ulp32(x):
    return np.spacing(np.float32(abs(x)))


# ---------------------------------------------------------------- canonical log
def test_log_matches_math_log():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(0.5, 2.0, 20000),
                         [1.0, 2.0, 0.5, 1e-3, 1e300, 5e-324, 2.2250738585072014e-308]])
    worst = 0.0
    for x in xs:
        a, b = oracle.log(x), math.log(x)
        if b != 0.0:
            worst = max(worst, abs(a - b) / math.ulp(b))
        else:
            assert a == 0.0
    assert worst <= 2.0
    assert oracle.log(0.0) == -math.inf
    assert math.isnan(oracle.log(-1.0)) and math.isnan(oracle.log(math.nan))
    assert oracle.log(math.inf) == math.inf


def test_box_radius_closed_form():
    rng = np.random.default_rng(2)
    ws = np.concatenate([rng.uniform(1.001e-3, 1.0, 5000), [0.1, 0.5, 1.0, 0.011, 2.0, 10.0]]).astype(f32)
    for w in ws:
        m = oracle.box_radius(w)
        exact = math.sqrt(-2.0 * math.log(1e-3 / float(w)))
        assert abs(float(m) - exact) <= 3 * ulp32(exact), (w, m, exact)
    # w < 1e-3 -> NaN radius (CR/forward.cu:79)
    for w in (0.0, 1e-4, 9.9e-4, -0.5):
        assert math.isnan(float(oracle.box_radius(w)))
    # known values quoted for the reference: 2.19 at w = 0.011, 3.72 at w = 1
    assert abs(float(oracle.box_radius(0.011)) - 2.19) < 5e-3
    assert abs(float(oracle.box_radius(1.0)) - 3.717) < 5e-3


# ---------------------------------------------------------------- preprocess
def _rot(q):
    return autograd_ref.rotation(__import__("torch").tensor(q, dtype=__import__("torch").float64)).numpy()


def test_identity_rotation_box_and_conic():
    rng = np.random.default_rng(3)
    P = 500
    means = rng.uniform(0, 1, (P, 3)).astype(f32)
    scales = np.exp(rng.normal(np.log(0.03), 0.3, (P, 3))).astype(f32)
    rots = np.tile(np.array([[1, 0, 0, 0]], f32), (P, 1))
    w = rng.uniform(0.01, 1.0, (P, 1)).astype(f32)
    k = 1.0
    conic, box, _ = oracle.preprocess(means, scales, rots, w, k)
    for i in range(P):
        m = oracle.box_radius(w[i, 0])
        # closed form: R = I, so the box is exactly fl(mu -+ fl(s m)) in fp32
        half = (scales[i] * m).astype(f32)
        assert np.array_equal(box[i, :3], (means[i] - half).astype(f32))
        assert np.array_equal(box[i, 3:], (means[i] + half).astype(f32))
        s = scales[i].astype(np.float64)
        var = s * s
        eps = var.max() * 1e-5
        expect = 1.0 / (var + eps)
        got = conic[i, [0, 3, 5]].astype(np.float64)
        np.testing.assert_allclose(got, expect, rtol=2e-6)
        assert np.all(conic[i, [1, 2, 4]] == 0.0)


def test_box_uses_rotation_rows():
    # 90 degrees about z: Rstd = [[0,-1,0],[1,0,0],[0,0,1]]; the box half-extent
    # along x is sum_k |Rstd[k][x]| s_k m = s_1 m  (rows, CR/forward.cu:95-97).
    c = np.float32(np.cos(np.pi / 4))
    q = np.array([[c, 0, 0, c]], f32)
    s = np.array([[0.01, 0.03, 0.02]], f32)
    mu = np.array([[0.5, 0.5, 0.5]], f32)
    w = np.array([[1.0]], f32)
    _, box, _ = oracle.preprocess(mu, s, q, w)
    m = float(oracle.box_radius(1.0))
    half = (box[0, 3:] - box[0, :3]) / 2
    np.testing.assert_allclose(half, [0.03 * m, 0.01 * m, 0.02 * m], rtol=1e-5)
    # 30 degrees about z: rows give |cos| s0 + |sin| s1 along x, which is NOT the
    # tight box of the m-sigma ellipse (fact recorded for the reference).
    th = np.pi / 6
    q = np.array([[np.cos(th / 2), 0, 0, np.sin(th / 2)]], f32)
    _, box, _ = oracle.preprocess(mu, s, q, w)
    half = (box[0, 3:] - box[0, :3]) / 2
    R = _rot(q)[0]
    expect = np.abs(R).T @ (s[0].astype(np.float64) * m)
    np.testing.assert_allclose(half, expect, rtol=1e-5)


def test_conic_inverts_covariance():
    rng = np.random.default_rng(4)
    P = 300
    q = rng.normal(size=(P, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(f32)
    s = np.exp(rng.normal(np.log(0.03), 0.3, (P, 3))).astype(f32)
    mu = rng.uniform(0, 1, (P, 3)).astype(f32)
    w = np.full((P, 1), 0.5, f32)
    conic, _, _ = oracle.preprocess(mu, s, q, w)
    R = _rot(q.astype(np.float64))
    for i in range(P):
        cov = R[i] @ np.diag(s[i].astype(np.float64) ** 2) @ R[i].T
        cov += np.eye(3) * np.max(np.abs(np.diag(cov))) * 1e-5
        inv = np.linalg.inv(cov)
        C = conic[i]
        Cm = np.array([[C[0], C[1], C[2]], [C[1], C[3], C[4]], [C[2], C[4], C[5]]], np.float64)
        np.testing.assert_allclose(Cm, inv, rtol=5e-4, atol=5e-4 * np.abs(inv).max())


def test_fp64_path_for_ill_conditioned():
    rng = np.random.default_rng(5)
    P = 200
    q = rng.normal(size=(P, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(f32)
    s = np.exp(rng.normal(np.log(0.03), 0.3, (P, 3))).astype(f32)
    s[:, 1] = 1e-5 * s[:, 0] / 0.03 * 0.03  # ratio ~3000 -> cond ~1e7
    s[:, 1] = np.float32(1e-5)
    mu = rng.uniform(0, 1, (P, 3)).astype(f32)
    w = np.full((P, 1), 0.5, f32)
    c32, _, ill32 = oracle.preprocess(mu, s, q, w, fp64_cond=False)
    c64, _, ill64 = oracle.preprocess(mu, s, q, w, fp64_cond=True)
    assert not ill32.any() and ill64.all()
    ref = oracle.conic_fp64(s, q)
    err64 = np.abs(c64 - ref).max(axis=1) / np.abs(ref).max(axis=1)
    err32 = np.abs(c32 - ref).max(axis=1) / np.abs(ref).max(axis=1)
    assert err64.max() < 1e-6          # only the final float rounding
    assert err32.max() > 10 * err64.max()  # the fp32 adjugate loses digits here


# ---------------------------------------------------------------- precise path
def _ill_set(P, ratio_axis, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(P, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(f32)
    s = np.exp(rng.normal(np.log(0.005), 0.2, (P, 3))).astype(f32)
    s[np.arange(P), rng.integers(0, 3, P)] = (s[:, 0] / ratio_axis).astype(f32)
    mu = rng.uniform(0.1, 0.9, (P, 3)).astype(f32)
    w = rng.uniform(0.2, 1.0, (P, 1)).astype(f32)
    return mu, s, q, w


def _sym(c):
    return np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]], np.float64)


def test_conic_factor_reconstructs_and_rejects():
    # U^T U with U = hi + lo reproduces the FP64 conic to double-float accuracy
    mu, s, q, w = _ill_set(300, 3000.0, 21)
    cd = oracle.conic_fp64(s, q)
    for c in cd:
        ok, hi, lo = oracle.conic_factor(c)
        assert ok
        U = np.zeros((3, 3))
        U[np.triu_indices(3)] = hi.astype(np.float64) + lo.astype(np.float64)
        C = _sym(c)
        assert np.abs(U.T @ U - C).max() <= 1e-13 * np.abs(C).max()
        assert np.all(np.abs(lo) <= np.spacing(np.abs(hi)))   # lo below half an ulp of hi
        assert U[0, 0] > 0 and U[1, 1] > 0 and U[2, 2] > 0
    # diagonal conic: U = diag(sqrt(c)), lo = the float rounding error of sqrt
    ok, hi, lo = oracle.conic_factor([4.0, 0, 0, 9.0, 0, 2.0])
    assert ok and np.array_equal(hi, np.array([2, 0, 0, 3, 0, np.float32(math.sqrt(2))], f32))
    assert lo[5] == np.float32(math.sqrt(2) - float(np.float32(math.sqrt(2))))
    # not positive definite
    assert not oracle.conic_factor([1.0, 2.0, 0, 1.0, 0, 1.0])[0]
    assert not oracle.conic_factor([-1.0, 0, 0, 1.0, 0, 1.0])[0]
    assert not oracle.conic_factor([np.nan, 0, 0, 1.0, 0, 1.0])[0]


@pytest.mark.parametrize("ratio", [35.0, 3500.0])
def test_precise_power_against_exact_quadratic_form(ratio):
    # power from the factor agrees with -0.5 d^T C d evaluated exactly (double,
    # exact d) to a few fp32 ulps of the power; the fp32 conic form of the same
    # Gaussian is orders of magnitude off, and dropping the lo parts (a planted
    # mistake) breaks the bound.
    mu, s, q, w = _ill_set(60, ratio, 22)
    conic, box, ill, fac = oracle.preprocess(mu, s, q, w, fp64_cond=True, fp64_ratio=10.0, factor=True)
    assert ill.all() and np.isfinite(fac).all()
    cd = oracle.conic_fp64(s, q)
    rng = np.random.default_rng(23)
    worst = worst_conic = worst_nolo = 0.0
    fac_nolo = fac.copy()
    fac_nolo[:, 6:] = 0.0
    for i in range(len(mu)):
        C = _sym(cd[i])
        K = np.linalg.cholesky(np.linalg.inv(C))       # points with quad in [0, 28]
        y = rng.normal(size=(200, 3))
        y *= (rng.uniform(0, 28, 200) ** 0.5 / np.linalg.norm(y, axis=1))[:, None]
        X = (mu[i].astype(np.float64) + y @ K.T).astype(f32)
        d = X.astype(np.float64) - mu[i].astype(np.float64)
        exact = -0.5 * np.einsum('ni,ij,nj->n', d, C, d)
        for x, e in zip(X, exact):
            tol = 4 * float(np.spacing(f32(max(abs(e), 1.0))))
            worst = max(worst, abs(oracle.precise_power(fac[i], mu[i], x) - e) / tol)
            worst_nolo = max(worst_nolo, abs(oracle.precise_power(fac_nolo[i], mu[i], x) - e) / tol)
            dx = (x - mu[i]).astype(f32)
            c = conic[i]
            r = [c[0] * dx[0] + c[1] * dx[1] + c[2] * dx[2], c[1] * dx[0] + c[3] * dx[1] + c[4] * dx[2],
                 c[2] * dx[0] + c[4] * dx[1] + c[5] * dx[2]]
            worst_conic = max(worst_conic, abs(-0.5 * float(np.dot(dx, r)) - e) / tol)
    assert worst <= 1.0
    assert worst_nolo > 1.0
    assert worst_conic > 10.0


@pytest.mark.parametrize("ratio", [12.0, 35.0, 300.0, 999.0])
def test_float_float_factor(ratio):
    # below the FP64 threshold the precise path's factor is computed in
    # float-float arithmetic: it agrees with the FP64 factor of the FP64 conic
    # (relative, normwise) and U^T U reproduces that conic, the rows are flagged
    # as off the FP64 path, and the conic output stays the fp32 one
    mu, s, q, w = _ill_set(200, ratio, 27)
    conic, box, ill, fac = oracle.preprocess(mu, s, q, w, fp64_cond=True, factor=True)
    conic32, box32, _ = oracle.preprocess(mu, s, q, w)
    r = s.max(axis=1) / s.min(axis=1)
    band = (r > 10.0) & (r <= 1000.0)
    assert band.sum() >= 60
    assert np.array_equal(ill, r > 1000.0)
    assert np.isfinite(fac[band]).all() and np.isnan(fac[~band & ~ill]).all()
    assert np.array_equal(conic[~ill], conic32[~ill]) and np.array_equal(box, box32)
    cd = oracle.conic_fp64(s, q)
    tol = 1e-12 * max(ratio, 10.0) ** 1.5
    for i in np.flatnonzero(band):
        ok, h2, l2 = oracle.conic_factor(cd[i])
        assert ok
        u = fac[i, :6].astype(np.float64) + fac[i, 6:]
        u2 = h2.astype(np.float64) + l2
        assert np.abs(u - u2).max() <= tol * np.abs(u2).max(), (i, np.abs(u - u2).max() / np.abs(u2).max())
        U = np.zeros((3, 3))
        U[np.triu_indices(3)] = u
        C = _sym(cd[i])
        assert np.abs(U.T @ U - C).max() <= 2 * tol * np.abs(C).max()
        assert np.all(np.abs(fac[i, 6:]) <= np.spacing(np.abs(fac[i, :6])))
    # above the factor threshold only; raising it to the FP64 threshold empties the band
    _, _, _, fac_off = oracle.preprocess(mu, s, q, w, fp64_cond=True, factor=True, factor_ratio=1000.0)
    assert np.isnan(fac_off[~ill]).all()
    # the precise power from this factor meets the exact quadratic form as the FP64 factor's does
    rng = np.random.default_rng(28)
    worst = 0.0
    for i in np.flatnonzero(band)[::10]:
        C = _sym(cd[i])
        K = np.linalg.cholesky(np.linalg.inv(C))
        y = rng.normal(size=(100, 3))
        y *= (rng.uniform(0, 28, 100) ** 0.5 / np.linalg.norm(y, axis=1))[:, None]
        X = (mu[i].astype(np.float64) + y @ K.T).astype(f32)
        d = X.astype(np.float64) - mu[i].astype(np.float64)
        exact = -0.5 * np.einsum('ni,ij,nj->n', d, C, d)
        for x, e in zip(X, exact):
            tol_p = 4 * float(np.spacing(f32(max(abs(e), 1.0))))
            worst = max(worst, abs(oracle.precise_power(fac[i], mu[i], x) - e) / tol_p)
    assert worst <= 1.0


def test_float_float_factor_special_cases():
    # identity rotation: the factor is diag(1 / sqrt(s^2 + eps)) to float-float accuracy
    s = np.array([0.02, 0.001, 0.004], f32)
    ok, hi, lo = oracle.conic_factor_ff(s, np.array([1, 0, 0, 0], f32))
    assert ok
    s2 = s.astype(np.float64) ** 2
    eps = s2.max() * 1e-5
    want = np.zeros(6)
    want[[0, 3, 5]] = 1.0 / np.sqrt(s2 + eps)
    got = hi.astype(np.float64) + lo
    assert np.abs(got - want).max() <= 1e-12 * want.max()
    # zero scales: not positive definite, no factor
    assert not oracle.conic_factor_ff(np.zeros(3, f32), np.array([1, 0, 0, 0], f32))[0]
    # the zero quaternion gives the identity rotation of the reference's formula
    ok0, hi0, lo0 = oracle.conic_factor_ff(s, np.zeros(4, f32))
    assert ok0 and np.array_equal(hi0, hi) and np.array_equal(lo0, lo)


def test_precise_path_forward_and_mask():
    # the forward with the factor matches FP64 sums at the same pairs, and the
    # pair set is exactly the points whose precise power lies in the window
    mu, s, q, w = _ill_set(40, 3000.0, 24)
    v = np.random.default_rng(25).uniform(0, 1, (40, 1)).astype(f32)
    conic, box, ill, fac = oracle.preprocess(mu, s, q, w, fp64_cond=True, factor=True)
    assert ill.all()
    rng = np.random.default_rng(26)
    k = rng.integers(0, 40, 3000)
    X = (box[k, :3] + rng.uniform(0, 1, (3000, 3)) * (box[k, 3:] - box[k, :3])).astype(f32)
    mask, power = oracle.pair_mask(mu, conic, box, X, factor=fac)
    for i in range(0, 40, 7):
        for j in np.flatnonzero(mask[i])[:20]:
            assert power[i, j] == np.float32(oracle.precise_power(fac[i], mu[i], X[j]))
    V, W, cnt = oracle.forward(mu, v, w, conic, box, X, oracle.conic_fp64(s, q), factor=fac)
    assert np.array_equal(cnt, mask.sum(axis=0))
    cd = oracle.conic_fp64(s, q)
    Wr = np.zeros(len(X)); Ur = np.zeros(len(X))
    for i in range(40):
        d = X[mask[i]].astype(np.float64) - mu[i].astype(np.float64)
        e = np.exp(-0.5 * np.einsum('ni,ij,nj->n', d, _sym(cd[i]), d))
        Wr[mask[i]] += float(w[i, 0]) * e
        Ur[mask[i]] += float(w[i, 0]) * e * float(v[i, 0])
    on = Wr > 1e-2
    assert np.allclose(W[on], Wr[on], rtol=1e-12) and np.allclose(V[on], Ur[on] / Wr[on], rtol=1e-12)


# ---------------------------------------------------------------- Morton
def test_morton_known_answers():
    pts = np.array([[0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                    [0.5, 0, 0], [np.nextafter(f32(1), f32(0)), 0, 0],
                    [-1.0 / 2097151.0, 0, 0], [2.0, 0, 0], [np.nan, 0, 0]], f32)
    c = oracle.morton3d_unit(pts)
    spread_all = 0x1249249249249249
    assert c[0] == 0
    assert c[1] == (1 << 63) - 1
    assert c[2] == spread_all and c[3] == spread_all << 1 and c[4] == spread_all << 2
    # 0.5 * 2097151 = 1048575.5 exactly in fp32 -> truncates to 1048575 = 2^20 - 1
    assert c[5] == sum(1 << (3 * b) for b in range(20))
    # (1 - 2^-24) * 2097151 = 2097150.875 exactly -> 2097150
    assert c[6] == spread_all & ~1
    # -1 wraps to 0x1fffff (the reference masks without clamping)
    assert c[7] == spread_all
    # 2.0 * 2097151 = 4194302 -> & 0x1fffff = 0x1ffffe
    assert c[8] == spread_all & ~1
    # NaN -> INT64_MIN & 0x1fffff = 0
    assert c[9] == 0
    cc = oracle.morton3d_unit(np.array([[-0.5, 1.5, np.nan]], f32), clamp=True)
    assert cc[0] == (spread_all << 1)


# ---------------------------------------------------------------- forward
def _small_problem(seed, P=48, S=3000, scale=0.05):
    rng = np.random.default_rng(seed)
    means = rng.uniform(0.2, 0.8, (P, 3)).astype(f32)
    scales = np.exp(rng.normal(np.log(scale), 0.3, (P, 3))).astype(f32)
    q = rng.normal(size=(P, 4))
    rots = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(f32)
    weights = rng.uniform(0.05, 1.0, (P, 1)).astype(f32)
    values = rng.uniform(0, 1, (P, 1)).astype(f32)
    samples = rng.uniform(0, 1, (S, 3)).astype(f32)
    return means, scales, rots, values, weights, samples


def test_forward_single_gaussian_closed_form():
    mu = np.array([[0.5, 0.5, 0.5]], f32)
    s = np.array([[0.05, 0.1, 0.07]], f32)
    q = np.array([[1, 0, 0, 0]], f32)
    w = np.array([[0.8]], f32)
    v = np.array([[0.3]], f32)
    xs = np.linspace(0.3, 0.7, 41).astype(f32)
    samples = np.stack([xs, np.full_like(xs, 0.52), np.full_like(xs, 0.49)], 1)
    conic, box, _ = oracle.preprocess(mu, s, q, w)
    V, W, cnt = oracle.forward(mu, v, w, conic, box, samples)
    var = s[0].astype(np.float64) ** 2
    var = var + var.max() * 1e-5
    d = samples.astype(np.float64) - mu[0]
    quad = (d * d / var).sum(1)
    expect = 0.8 * np.exp(-0.5 * quad)
    inside = (samples >= box[0, :3]).all(1) & (samples <= box[0, 3:]).all(1) & (quad <= 28)
    expect = np.where(inside, expect, 0.0)
    valid = expect > 1e-2
    np.testing.assert_allclose(W[valid], expect[valid], rtol=1e-6)
    assert np.all(V[valid] == np.float64(np.float32(0.3)))
    assert np.all(V[~valid] == -1.0) and np.all(W[~valid] == 0.0)


def test_forward_convexity_permutation_and_torch():
    import torch
    means, scales, rots, values, weights, samples = _small_problem(6)
    conic, box, _ = oracle.preprocess(means, scales, rots, weights)
    V, W, cnt = oracle.forward(means, values, weights, conic, box, samples, oracle.conic_fp64(scales, rots))
    ok = W > 0
    assert ok.sum() > 500
    assert V[ok].min() >= values.min() - 1e-12 and V[ok].max() <= values.max() + 1e-12
    perm = np.random.default_rng(0).permutation(means.shape[0])
    c2, b2, _ = oracle.preprocess(means[perm], scales[perm], rots[perm], weights[perm])
    V2, W2, cnt2 = oracle.forward(means[perm], values[perm], weights[perm], c2, b2, samples,
                                  oracle.conic_fp64(scales[perm], rots[perm]))
    np.testing.assert_allclose(V2, V, rtol=1e-12, atol=1e-12)
    assert np.array_equal(cnt, cnt2)
    mask, _ = oracle.pair_mask(means, conic, box, samples)
    assert np.array_equal(mask.sum(0), cnt)
    tV, tW = autograd_ref.forward(*(torch.tensor(x, dtype=torch.float64) for x in
                                    (means, scales, rots, values, weights, samples)), torch.tensor(mask))
    np.testing.assert_allclose(tV.numpy(), V, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(tW.numpy(), W, rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- backward
def _grads_problem(seed):
    means, scales, rots, values, weights, samples = _small_problem(seed, P=40, S=2500)
    conic, box, _ = oracle.preprocess(means, scales, rots, weights)
    cd = oracle.conic_fp64(scales, rots)
    V, W, _ = oracle.forward(means, values, weights, conic, box, samples, cd)
    rng = np.random.default_rng(seed + 100)
    gV = rng.normal(size=V.shape)
    gW = rng.normal(size=W.shape)
    mask, _ = oracle.pair_mask(means, conic, box, samples)
    return means, scales, rots, values, weights, samples, conic, box, V, W, gV, gW, mask


def test_backward_matches_autograd():
    (means, scales, rots, values, weights, samples, conic, box, V, W, gV, gW, mask) = _grads_problem(7)
    g = oracle.backward(means, scales, rots, values, weights, conic, box, samples, V, W, gV, gW)
    ref = autograd_ref.grads(means, scales, rots, values, weights, samples, mask, gV, gW)
    for name in ("means", "scales", "rotations", "values", "weights"):
        a = g[name].reshape(-1)
        b = ref[name].reshape(-1)
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-9 * np.abs(b).max())


def test_backward_central_differences():
    (means, scales, rots, values, weights, samples, conic, box, V, W, gV, gW, mask) = _grads_problem(8)
    g = oracle.backward(means, scales, rots, values, weights, conic, box, samples, V, W, gV, gW)
    params = [means, scales, rots, values, weights]
    names = ["means", "scales", "rotations", "values", "weights"]
    rng = np.random.default_rng(9)
    for _ in range(24):
        which = int(rng.integers(0, 5))
        idx = int(rng.integers(0, params[which].size))
        fd = autograd_ref.central_difference(params, which, idx, 1e-6, samples, mask, gV, gW)
        an = g[names[which]].reshape(-1)[idx]
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (names[which], idx, fd, an)


def _reference_chain(scale, q, conic6, gconic6):
    """CR/backward.cu:68-124 restated (planted mistake: not the forward's gradient)."""
    R = _rot(q[None].astype(np.float64))[0]
    S = np.diag(scale.astype(np.float64))
    M = S @ R.T  # glm M = S * R(glm) with R(glm) = Rstd^T
    C = np.array([[conic6[0], conic6[1], conic6[2]], [conic6[1], conic6[3], conic6[4]],
                  [conic6[2], conic6[4], conic6[5]]])
    G = np.array([[gconic6[0], gconic6[1], gconic6[2]], [gconic6[1], gconic6[3], gconic6[4]],
                  [gconic6[2], gconic6[4], gconic6[5]]])
    dC = -C @ G @ C  # off-diagonals of G not halved (CR/backward.cu:75-79)
    Sig = M.T @ M
    d = np.abs(np.diag(Sig))
    eps_d = np.where(d >= d.max(), 1e-5 * np.sign(np.diag(Sig)), 0.0)
    dSig = dC + np.diag(np.trace(dC) * eps_d)
    dM = 2.0 * dSig @ M  # glm "2 * dL_dSigma * M" (CR/backward.cu:105); the forward needs 2 M dSig
    return np.diag(dM @ R)  # dL_dS = dL_dM * R(glm)^T (CR/backward.cu:106-112)


def test_autograd_check_rejects_reference_chain():
    (means, scales, rots, values, weights, samples, conic, box, V, W, gV, gW, mask) = _grads_problem(10)
    g = oracle.backward(means, scales, rots, values, weights, conic, box, samples, V, W, gV, gW)
    ref = autograd_ref.grads(means, scales, rots, values, weights, samples, mask, gV, gW)
    cd = oracle.conic_fp64(scales, rots)
    bad = np.stack([_reference_chain(scales[i], rots[i], cd[i], g["conic"][i]) for i in range(len(scales))])
    rel_bad = np.abs(bad - ref["scales"]).max() / np.abs(ref["scales"]).max()
    rel_good = np.abs(g["scales"] - ref["scales"]).max() / np.abs(ref["scales"]).max()
    assert rel_good < 1e-8
    assert rel_bad > 1e-2


# ---------------------------------------------------------------- Karras tree
def _recursive_tree(keys):
    """Binary radix tree by recursive top-down splitting (independent of Karras's
    bottom-up search): a node over [lo, hi] splits after the last index whose key
    shares the node's longest common prefix plus a 0 bit; equal keys split by index
    bits.  Returns {(lo, hi): (left_range, right_range)}."""
    n = len(keys)
    ext = [(int(k) << 32) | i for i, k in enumerate(keys)]  # key then index (96 bits)
    out = {}

    def rec(lo, hi):
        if lo == hi:
            return
        a, b = ext[lo], ext[hi]
        bit = (a ^ b).bit_length() - 1
        split = lo
        while split + 1 <= hi and not ((ext[split + 1] >> bit) & 1):
            split += 1
        out[(lo, hi)] = ((lo, split), (split + 1, hi))
        rec(lo, split)
        rec(split + 1, hi)

    rec(0, n - 1)
    return out


@pytest.mark.parametrize("seed,n,dup", [(0, 2, False), (1, 17, False), (2, 300, True), (3, 512, True)])
def test_karras_matches_recursive_split(seed, n, dup):
    rng = np.random.default_rng(seed)
    if dup:
        keys = np.sort(rng.integers(0, 40, n).astype(np.uint64) << np.uint64(50))
    else:
        keys = np.sort(rng.integers(0, 2**63 - 1, n, dtype=np.int64).astype(np.uint64))
    left, right, first, last = oracle.karras_build(keys)
    ref = _recursive_tree(keys)
    assert len(ref) == n - 1

    def rng_of(c):
        return (~c, ~c) if c < 0 else (int(first[c]), int(last[c]))

    seen = set()
    for i in range(n - 1):
        node = (int(first[i]), int(last[i]))
        assert node in ref
        assert (rng_of(int(left[i])), rng_of(int(right[i]))) == ref[node]
        seen.add(node)
    assert seen == set(ref)
    assert (first[0], last[0]) == (0, n - 1)


def test_binned_forward_equals_brute_force():
    """The grid-binned FP64 forward visits the same pairs in the same order as the
    brute force: bitwise equal V, W and counts, also with NaN / infinite boxes,
    zero weights, samples outside every box and a NaN sample."""
    rng = np.random.default_rng(21)
    P, S = 6000, 3000
    means = rng.uniform(0, 1, (P, 3)).astype(np.float32)
    scales = np.exp(rng.normal(np.log(0.02), 0.5, (P, 3))).astype(np.float32)
    rots = rng.normal(size=(P, 4)).astype(np.float32)
    w = rng.uniform(0, 1, (P, 1)).astype(np.float32)
    v = rng.uniform(0, 1, (P, 1)).astype(np.float32)
    means[0] = [np.nan, 0.5, 0.5]
    scales[1] = [np.inf, 0.01, 0.01]
    w[2] = 0.0
    scales[3] = [0.02, 1e-6, 0.02]
    conic, box, _ = oracle.preprocess(means, scales, rots, w)
    cd = oracle.conic_fp64(scales, rots)
    pts = rng.uniform(-0.2, 1.2, (S, 3)).astype(np.float32)
    pts[0] = [np.nan, 0.5, 0.5]
    a = oracle.forward(means, v, w, conic, box, pts, cd, binned=False)
    b = oracle.forward(means, v, w, conic, box, pts, cd, binned=True)
    assert a[2].sum() > 10 * S
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint64) if x.dtype == np.float64 else x,
                              y.view(np.uint64) if y.dtype == np.float64 else y)


def test_cpu_plain_baseline_matches_oracle():
    # The CPU baseline bench.py times (plain fp32 brute force, no compensation)
    # computes the same thing as the oracle: the same pair set, the field within
    # 1e-4 relative of the FP64 forward, and its per-Gaussian partial gradients
    # (dv, dw, dmu, packed dconic) within 1e-3 normwise of the FP64 backward's.
    g = workloads.config0(0)
    scales = np.exp(g["log_scales"]).astype(f32)
    samples = g["samples"][::8].copy()
    conic, box, _ = oracle.preprocess(g["means"], scales, g["rotations"], g["weights"])
    rng = np.random.default_rng(29)
    gV = rng.normal(size=samples.shape[0]).astype(f32)
    gW = rng.normal(size=samples.shape[0]).astype(f32)
    V, W, g11, pairs = oracle.cpu_plain_fwdbwd(g["means"], g["values"], g["weights"], conic, box, samples, gV, gW)
    Vr, Wr, cnt = oracle.forward(g["means"], g["values"], g["weights"], conic, box, samples)
    assert pairs == int(cnt.sum()) > 0
    ok = Wr > 0
    assert np.array_equal(W > 0, ok)
    assert np.abs(V[ok] - Vr[ok]).max() <= 1e-4 * max(np.abs(Vr[ok]).max(), 1e-3)
    assert (np.abs(W[ok] - Wr[ok]) / Wr[ok]).max() <= 1e-4
    ref = oracle.backward(g["means"], scales, g["rotations"], g["values"], g["weights"], conic, box, samples, Vr,
                          Wr, gV.astype(np.float64), gW.astype(np.float64))
    for cols, name in ((slice(0, 1), "values"), (slice(1, 2), "weights"), (slice(2, 5), "means"),
                       (slice(5, 11), "conic")):
        a = g11[:, cols].astype(np.float64).reshape(-1)
        b = np.asarray(ref[name], np.float64).reshape(-1)
        assert np.abs(a - b).max() <= 1e-3 * np.abs(b).max(), name
