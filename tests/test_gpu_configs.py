"""The benchmark workloads at full size on the HIP path, checked against the oracle.

Each test builds one workload exactly as bench.py does (gvc.workloads), runs one
forward + backward over the whole step's points through the C ABI, and checks:
  * pair membership per sample (gvc_forward_pair_counts) equal to the oracle's on
    a subset of the points, and the pair totals consistent (per sample, per
    Gaussian and the forward's own count);
  * the field (V, W) within 1e-5 relative of the FP64 oracle on that subset, with
    the same sentinel decisions;
  * the gradients of 256 random Gaussians within 1e-5 (normwise per parameter
    tensor) of the FP64 oracle's: a Gaussian's gradient sees only the points
    inside its box, so the oracle's V and W are evaluated at every step point in
    the checked boxes, against all Gaussians.
Workloads: the structured fit (200k Gaussians, 2^20 points on a 128^3 grid
field), the mesh fit (1M Gaussians, 2^22 points drawn inside a 12.6M-tet mesh,
cell ids exact, plus central differences on 256 parameters), the precision
workload (4M Gaussians, 2^23 points, FP64-cond at the default threshold) and the
32-volume batch (9.3M Gaussians, 2^24 points).
"""
import numpy as np
import pytest

from conftest import assert_grads, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

import torch  # noqa: E402

import oracle  # noqa: E402
from gvc import batched as bx  # noqa: E402
from gvc import mixture as mx  # noqa: E402
from gvc import training  # noqa: E402
from gvc import workloads  # noqa: E402
from oracle import autograd_ref, mesh_ref  # noqa: E402

DEV = torch.device("cuda:0")
f32 = np.float32
FIELD_BAR = 1e-5


def activated(W, max_scale=0.02):
    """The reference model's activations of the raw parameters
    (R/scene/gaussian_model.py:31-48, :114-132), on the device and on the host."""
    with torch.no_grad():
        sc, rq, wv, vv = training.activate(W["log_scales"], W["rotations"], torch.logit(W["weights"]),
                                           torch.logit(W["values"].clamp(1e-4, 1 - 1e-4)), max_scale)
    dev = (W["means"].contiguous(), sc, rq, vv, wv)
    host = tuple(t.detach().cpu().numpy() for t in dev)
    return dev, host


def inside_boxes(samples, boxes):
    """Indices of the samples inside any of the boxes [k, 6] (device)."""
    bx_ = torch.as_tensor(boxes, device=samples.device)
    inside = torch.zeros(samples.shape[0], dtype=torch.bool, device=samples.device)
    for k in range(bx_.shape[0]):
        inside |= ((samples >= bx_[k, :3]) & (samples <= bx_[k, 3:])).all(dim=1)
    return inside.nonzero().squeeze(1)


def field_check(V, W, Vr, Wr):
    assert np.array_equal(W > 0, Wr > 0), "sentinel decisions differ"
    ok = Wr > 0
    err = np.maximum(np.abs(V[ok] - Vr[ok]) / np.maximum(np.abs(Vr[ok]), 1e-3),
                     np.abs(W[ok] - Wr[ok]) / np.maximum(np.abs(Wr[ok]), 1e-3))
    e = float(err.max(initial=0.0))
    print(f"field max rel err {e:.2e} over {int(ok.sum())} points")
    assert e <= FIELD_BAR
    return e


def run_step(dev_params, samples, params, seed=5):
    """One forward + backward over all samples with random upstream gradients."""
    m, sc, rq, vv, wv = dev_params
    gen = torch.Generator(device=DEV)
    gen.manual_seed(seed)
    gV = torch.randn(samples.shape[0], device=DEV, generator=gen)
    gW = torch.randn(samples.shape[0], device=DEV, generator=gen)
    with torch.no_grad():
        ss = mx.SampleSet(samples)
        cells, cw, _, st = mx.mixture_forward(m, sc, rq, vv, wv, 1.0, ss, params, want_conics=False)
        grads = mx.mixture_backward(st, m, sc, rq, vv, wv, 1.0, ss, cells, cw, gV, gW)
        per_sample, per_gaussian = st.pair_counts(ss)
    nrec, npairs, err = st.stats()
    assert err == 0
    ps, pg = per_sample.cpu().numpy(), per_gaussian.cpu().numpy()
    assert int(ps.sum()) == npairs == int(pg.sum())
    return dict(cells=cells, cw=cw, grads=grads, per_sample=ps, per_gaussian=pg, gV=gV, gW=gW, npairs=npairs)


def oracle_checks(host, samples, out, pre, fac, n_field=1024, n_grad=256, seed=3, extra_field=None):
    """Pair counts and field on a random subset, gradients of random Gaussians."""
    mh, sh, rh, vh, wh = host
    conic, box = pre
    P, S = mh.shape[0], samples.shape[0]
    rng = np.random.default_rng(seed)
    cd = oracle.conic_fp64(sh, rh)
    sub = rng.choice(S, n_field, replace=False)
    if extra_field is not None:
        sub = np.concatenate([sub, extra_field])
    pts = samples[torch.as_tensor(sub, device=DEV)].cpu().numpy()
    Vr, Wr, cnt = oracle.forward(mh, vh, wh, conic, box, pts, cd, factor=fac)
    assert np.array_equal(out["per_sample"][sub], cnt), "pair membership differs from the oracle"
    field_check(out["cells"].cpu().numpy()[sub], out["cw"].cpu().numpy()[sub], Vr, Wr)
    gsub = rng.choice(P, n_grad, replace=False)
    idx = inside_boxes(samples, box[gsub])
    gpts = samples[idx].cpu().numpy()
    Vg, Wg, _ = oracle.forward(mh, vh, wh, conic, box, gpts, cd, factor=fac)
    gv, gw = out["gV"][idx].double().cpu().numpy(), out["gW"][idx].double().cpu().numpy()
    ref = oracle.backward(mh[gsub], sh[gsub], rh[gsub], vh[gsub], wh[gsub], conic[gsub], box[gsub], gpts, Vg, Wg,
                          gv, gw, factor=fac[gsub] if fac is not None else None)
    grads = [t.detach().cpu().numpy().reshape(P, -1)[gsub] for t in out["grads"]]
    errs = assert_grads(grads, ref)
    return dict(gsub=gsub, idx=idx, gpts=gpts, Vg=Vg, Wg=Wg, gv=gv, gw=gw, ref=ref, grads=grads, errs=errs)


def test_structured_fit_200k():
    """200k Gaussians at random voxel centres of a 128^3 grid field (64 blobs +
    noise), 3-NN scales, 2^20 uniform points; ground truth by the grid probe."""
    W = workloads.grid_workload(DEV)
    dev_params, host = activated(W)
    gen = torch.Generator(device=DEV)
    gen.manual_seed(11)
    samples = torch.rand((1 << 20, 3), device=DEV, generator=gen)
    gt = W["grid_fn"](samples)
    sub = np.random.default_rng(0).choice(samples.shape[0], 4096, replace=False)
    n = 128
    f_xfast = W["field"].reshape(n, n, n).permute(2, 1, 0).contiguous().reshape(-1).cpu().numpy()
    ref_gt = mesh_ref.grid_sample(np.array([n, n, n]), np.zeros(3, f32), np.full(3, 1.0 / (n - 1), f32), f_xfast,
                                  samples[torch.as_tensor(sub, device=DEV)].cpu().numpy())
    np.testing.assert_allclose(gt.cpu().numpy()[sub], ref_gt, rtol=1e-6, atol=1e-6)
    out = run_step(dev_params, samples, mx.MixtureParams())
    mh, sh, rh, vh, wh = host
    conic, box, _ = oracle.preprocess(mh, sh, rh, wh)
    oracle_checks(host, samples, out, (conic, box), None)


def _near_face_truth(pts, tets, p, n):
    """Lowest index of a closed tet holding p (exact predicate) among the Kuhn tets
    of the 27 lattice cubes around p: with vertices jittered by at most 0.3 of a
    cell, no other tet can hold it."""
    c = np.floor(p.astype(np.float64) * n).astype(np.int64)
    best = -1
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for dk in (-1, 0, 1):
                i, j, k = c[0] + di, c[1] + dj, c[2] + dk
                if not (0 <= i < n and 0 <= j < n and 0 <= k < n):
                    continue
                cube = (i * n + j) * n + k
                for t in range(6 * cube, 6 * cube + 6):
                    if (best < 0 or t < best) and mesh_ref.tet_contains(pts, tets[t], p)[0]:
                        best = t
    return best


def test_mesh_fit_1m():
    """1M Gaussians on a 128^3 jittered lattice of 12.6M tets; 2^22 points drawn
    inside the mesh (R/train.py:34-61).  Cell ids: a point clear of every face of
    its construction tet is located in exactly that tet; near-face points in the
    lowest-index closed tet that holds them (exact predicate).  Then the field,
    pair and gradient checks, and central differences on 256 parameters."""
    from gvc.sampling import TetMeshSampler, sample_mesh_points
    W = workloads.mesh_workload(DEV)
    n = 128
    pts_m, tets = W["mesh_points"], W["tets"]
    gen = torch.Generator(device=DEV)
    gen.manual_seed(21)
    pts_d = torch.as_tensor(pts_m, device=DEV)
    samples, cell = sample_mesh_points(pts_d, tets, 1, 1 << 22, generator=gen, return_cells=True)
    samples, cell = samples[0].contiguous(), cell[0]
    sampler = TetMeshSampler(pts_d, tets, W["mesh_values"])
    gt, ids = sampler.sample(samples)
    ids = ids.long()
    # barycentrics of every point in its construction tet (float64, Cramer's rule)
    tl = torch.as_tensor(tets, device=DEV).long()
    tv = pts_d.double()[tl[cell]]                                             # [S, 4, 3]
    a, b, c, d = (tv[:, i] - tv[:, 3] for i in range(4))
    x = samples.double() - tv[:, 3]

    def det3(u, v, w):
        return (u * torch.cross(v, w, dim=1)).sum(1)

    D = det3(a, b, c)
    lam = torch.stack([det3(x, b, c) / D, det3(a, x, c) / D, det3(a, b, x) / D], 1)
    lam = torch.cat([lam, 1.0 - lam.sum(1, keepdim=True)], 1)
    # the +-0.3-cell jitter inverts some tets (their hulls overlap neighbours):
    # a point near one of them may lie in several tets, and the lowest index wins
    # the unjittered lattice: vertex i + m (j + m k) at (i, j, k) / n
    ref_lat = torch.stack(torch.meshgrid(*(torch.arange(n + 1, device=DEV, dtype=torch.float64) / n,) * 3,
                                         indexing="ij"), -1)[..., [2, 1, 0]].reshape(-1, 3)
    v0 = pts_d.double()[tl]
    vl = ref_lat[tl]

    def vol(V):
        return det3(V[:, 0] - V[:, 3], V[:, 1] - V[:, 3], V[:, 2] - V[:, 3])

    inverted = torch.sign(vol(v0)) != torch.sign(vol(vl))
    del v0, vl
    bad_cube = inverted.reshape(-1, 6).any(1).reshape(n, n, n).float()[None, None]
    near_bad = torch.nn.functional.max_pool3d(bad_cube, 3, stride=1, padding=1)[0, 0].reshape(-1) > 0
    clear = (lam.min(1).values > 1e-5) & ~near_bad[cell // 6]
    print(f"inverted tets {int(inverted.sum())}, points checked against their construction tet "
          f"{int(clear.sum())} of {samples.shape[0]}")
    assert int(clear.sum()) > 0.8 * samples.shape[0]
    assert torch.equal(ids[clear], cell[clear]), "cell ids of points clear of faces and overlaps differ"
    rest = (~clear).nonzero().squeeze(1)
    pick = rest[torch.as_tensor(np.random.default_rng(4).choice(rest.numel(), min(300, rest.numel()), replace=False),
                                device=DEV)]
    for i in pick.cpu().numpy():
        p = samples[i].cpu().numpy()
        assert int(ids[i]) == _near_face_truth(pts_m, tets, p, n), f"point {i}"
    # the interpolated ground truth: barycentric combination of the vertex values
    vals = W["mesh_values"].double()[torch.as_tensor(tets, device=DEV).long()[cell[clear]]]
    want = (vals * lam[clear]).sum(1)
    assert torch.allclose(gt.double()[clear], want, rtol=1e-9, atol=1e-12)

    dev_params, host = activated(W)
    out = run_step(dev_params, samples, mx.MixtureParams())
    mh, sh, rh, vh, wh = host
    conic, box, _ = oracle.preprocess(mh, sh, rh, wh)
    chk = oracle_checks(host, samples, out, (conic, box), None)
    central_differences(host, samples, out, conic, box, chk)


def central_differences(host, samples, out, conic, box, chk, n_params=256, rel_h=1e-3):
    """Frozen-predicate central differences in FP64 on n_params random parameters
    (one per checked Gaussian: a mean, scale, quaternion, value or weight
    component), step h = rel_h x the parameter's own scale (the Gaussian's
    smallest scale for a mean, the scale itself, 1 for a quaternion component or
    a value, the weight for a weight).  L = sum gV V + gW W over the step's points;
    a parameter of Gaussian k moves only k's contributions at the points of its
    pair set, where the other Gaussians' FP64 sums are held fixed.  Bars: the
    oracle's analytic gradient and the GPU's within 1e-4 of the differences (per
    parameter kind, normwise; the differences' own truncation error is O(h^2),
    about 1e-6 times the window's largest quadratic form, 28)."""
    mh, sh, rh, vh, wh = host
    gsub, gpts = chk["gsub"], chk["gpts"]
    Vg, Wg, gv, gw = chk["Vg"], chk["Wg"], chk["gv"], chk["gw"]
    rng = np.random.default_rng(17)
    kinds = [("means", 3), ("scales", 3), ("rotations", 4), ("values", 1), ("weights", 1)]
    fd, an, gp, kind_of = [], [], [], []
    for j in range(min(n_params, gsub.size)):
        k = int(gsub[j])
        mask, _ = oracle.pair_mask(mh[k:k + 1], conic[k:k + 1], box[k:k + 1], gpts)
        sel = mask[0] & (Wg > 1e-2)     # sentinel points are constants of the loss
        if not sel.any():
            continue
        x = torch.as_tensor(gpts[sel], dtype=torch.float64)
        cdk = torch.as_tensor(oracle.conic_fp64(sh[k:k + 1], rh[k:k + 1]))[0]
        d0 = x - torch.as_tensor(mh[k], dtype=torch.float64)
        C0 = torch.stack([torch.stack([cdk[0], cdk[1], cdk[2]]), torch.stack([cdk[1], cdk[3], cdk[4]]),
                          torch.stack([cdk[2], cdk[4], cdk[5]])])
        own0 = float(wh[k, 0]) * torch.exp(-0.5 * ((d0 @ C0) * d0).sum(1))
        W_o = torch.as_tensor(Wg[sel]) - own0
        U_o = torch.as_tensor(Vg[sel] * Wg[sel]) - own0 * float(vh[k, 0])
        gVk, gWk = torch.as_tensor(gv[sel]), torch.as_tensor(gw[sel])

        def loss(params):
            m_, s_, q_, v_, w_ = params
            C = autograd_ref.conic(s_[None], q_[None])[0]
            Cm = torch.stack([torch.stack([C[0], C[1], C[2]]), torch.stack([C[1], C[3], C[4]]),
                              torch.stack([C[2], C[4], C[5]])])
            d = x - m_
            own = w_ * torch.exp(-0.5 * ((d @ Cm) * d).sum(1))
            Wt = W_o + own
            Vt = (U_o + own * v_) / Wt
            return float((gVk * Vt).sum() + (gWk * Wt).sum())

        base = [torch.as_tensor(np.asarray(a, np.float64).reshape(-1)) for a in
                (mh[k], sh[k], rh[k], vh[k, 0], wh[k, 0])]
        base[3], base[4] = base[3][0], base[4][0]
        which = int(rng.integers(0, 5))
        comp = int(rng.integers(0, kinds[which][1]))
        scale = {0: float(sh[k].min()), 1: float(sh[k, comp]) if which == 1 else 1.0, 2: 1.0, 3: 1.0,
                 4: float(wh[k, 0])}[which]
        h = rel_h * scale
        plus = [b.clone() for b in base]
        minus = [b.clone() for b in base]
        if which <= 2:
            plus[which][comp] += h
            minus[which][comp] -= h
        else:
            plus[which] = plus[which] + h
            minus[which] = minus[which] - h
        fd.append((loss(plus) - loss(minus)) / (2 * h))
        an.append(float(np.asarray(chk["ref"][kinds[which][0]]).reshape(gsub.size, -1)[j, comp]))
        gp.append(float(chk["grads"][which].reshape(gsub.size, -1)[j, comp]))
        kind_of.append(which)
    fd, an, gp, kind_of = np.array(fd), np.array(an), np.array(gp), np.array(kind_of)
    assert fd.size >= 0.9 * min(n_params, gsub.size)
    for w in range(5):
        sel = kind_of == w
        if not sel.any():
            continue
        top = max(np.abs(fd[sel]).max(), 1e-300)
        e_or = np.abs(an[sel] - fd[sel]).max() / top
        e_gpu = np.abs(gp[sel] - fd[sel]).max() / top
        print(f"central differences, {kinds[w][0]}: {int(sel.sum())} params, oracle {e_or:.2e}, GPU {e_gpu:.2e}")
        assert e_or <= 1e-4, (kinds[w][0], e_or)
        assert e_gpu <= 1e-4, (kinds[w][0], e_gpu)


def test_precision_workload_4m():
    """4M Gaussians on the 512^3 precision field, 2% with one log-scale at ln 1e-4
    (scale ratio ~35) and 0.5% at ln 1e-6 (ratio > 1000, cond > 1e6), rotated;
    2^23 uniform points, FP64-cond at the default thresholds (FP64 factor above
    scale ratio 1000, float-float factor above 10).  Oracle on a
    65,536-point subset plus 4,096 points inside the ill-conditioned Gaussians'
    boxes; gradients of 256 Gaussians each from the ratio > 1000, ratio-35 and
    remaining sets."""
    W = workloads.precision_workload(DEV)
    dev_params, host = activated(W)
    gen = torch.Generator(device=DEV)
    gen.manual_seed(31)
    samples = torch.rand((1 << 23, 3), device=DEV, generator=gen)
    params = mx.MixtureParams(fp64_cond=True)
    out = run_step(dev_params, samples, params)
    mh, sh, rh, vh, wh = host
    conic, box, ill, fac = oracle.preprocess(mh, sh, rh, wh, 1.0, True, 0.0, factor=True)
    rng = np.random.default_rng(8)
    deg = np.zeros(mh.shape[0], bool)
    deg[np.asarray(W["degenerate"])] = True
    aniso = np.flatnonzero(ill | deg)
    stress_idx = inside_boxes(samples, box[rng.choice(aniso, 512, replace=False)]).cpu().numpy()
    stress_idx = rng.choice(stress_idx, min(4096, stress_idx.size), replace=False)
    chk = oracle_checks(host, samples, out, (conic, box), fac, n_field=65536, n_grad=256, extra_field=stress_idx)
    # gradients of the two anisotropic sets
    for mask in (ill, deg):
        gsub = rng.choice(np.flatnonzero(mask), 256, replace=False)
        idx = inside_boxes(samples, box[gsub])
        gpts = samples[idx].cpu().numpy()
        cd = oracle.conic_fp64(sh, rh)
        Vg, Wg, _ = oracle.forward(mh, vh, wh, conic, box, gpts, cd, factor=fac)
        ref = oracle.backward(mh[gsub], sh[gsub], rh[gsub], vh[gsub], wh[gsub], conic[gsub], box[gsub], gpts, Vg,
                              Wg, out["gV"][idx].double().cpu().numpy(), out["gW"][idx].double().cpu().numpy(),
                              factor=fac[gsub])
        assert_grads([t.detach().cpu().numpy().reshape(mh.shape[0], -1)[gsub] for t in out["grads"]], ref)
    del chk


def test_batch32_uneven_volumes():
    """32 volumes (sides log-uniform in [32, 256], Gaussians log-uniform in
    [1e4, 2e6], 2^24 points in proportion to voxel count) in one batched forward
    and backward; every volume checked against the oracle on its own Gaussians:
    pair counts and field at 64 of its points, gradients of 8 of its Gaussians."""
    W = workloads.batch_workload(DEV)
    dev_params, host = activated(W)
    batch = bx.Batch(W["gaussians"], W["samples"], device=DEV)
    samples = W["point_batches"][0]
    gen = torch.Generator(device=DEV)
    gen.manual_seed(41)
    gV = torch.randn(samples.shape[0], device=DEV, generator=gen)
    gW = torch.randn(samples.shape[0], device=DEV, generator=gen)
    m, sc, rq, vv, wv = dev_params
    with torch.no_grad():
        ss = bx.BatchedSampleSet(batch, samples)
        cells, cw, _, st = bx.mixture_forward_batched(batch, m, sc, rq, vv, wv, 1.0, ss)
        grads = bx.mixture_backward_batched(st, m, sc, rq, vv, wv, 1.0, ss, cells, cw, gV, gW)
        per_sample, per_gaussian = st.pair_counts(ss)
    nrec, npairs, err = st.stats()
    assert err == 0
    ps = per_sample.cpu().numpy()
    assert int(ps.sum()) == npairs == int(per_gaussian.sum())
    mh, sh, rh, vh, wh = host
    rng = np.random.default_rng(12)
    Vall, Wall = cells.cpu().numpy(), cw.cpu().numpy()
    gall = [t.detach().cpu().numpy() for t in grads]
    for v in range(len(W["gaussians"])):
        gs, sl = batch.gaussian_slice(v), batch.sample_slice(v)
        hv = tuple(a[gs] for a in (mh, sh, rh, vh, wh))
        conic, box, _ = oracle.preprocess(hv[0], hv[1], hv[2], hv[4])
        cd = oracle.conic_fp64(hv[1], hv[2])
        S_v = sl.stop - sl.start
        sub = rng.choice(S_v, min(64, S_v), replace=False)
        pts = samples[sl][torch.as_tensor(sub, device=DEV)].cpu().numpy()
        Vr, Wr, cnt = oracle.forward(hv[0], hv[3], hv[4], conic, box, pts, cd)
        assert np.array_equal(ps[sl][sub], cnt), f"pairs of volume {v}"
        field_check(Vall[sl][sub], Wall[sl][sub], Vr, Wr)
        gsub = rng.choice(hv[0].shape[0], 8, replace=False)
        idx = inside_boxes(samples[sl], box[gsub])
        gpts = samples[sl][idx].cpu().numpy()
        Vg, Wg, _ = oracle.forward(hv[0], hv[3], hv[4], conic, box, gpts, cd)
        ref = oracle.backward(hv[0][gsub], hv[1][gsub], hv[2][gsub], hv[3][gsub], hv[4][gsub], conic[gsub],
                              box[gsub], gpts, Vg, Wg, gV[sl][idx].double().cpu().numpy(),
                              gW[sl][idx].double().cpu().numpy())
        if all(np.abs(ref[n]).max() > 0 for n in ("means", "values")):
            assert_grads([g[gs].reshape(hv[0].shape[0], -1)[gsub] for g in gall], ref)
