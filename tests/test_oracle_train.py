"""The training-loop oracle (oracle/train_ref.py) pinned against torch's own Adam
and topk, and the learning-rate schedule against its closed forms."""
import math

import numpy as np
import pytest
import torch

from oracle import train_ref


def test_expon_lr_closed_forms():
    a, b, n = 1.6e-4, 1.6e-6, 30000
    assert train_ref.expon_lr(0, a, b, max_steps=n) == pytest.approx(a, rel=1e-15)
    assert train_ref.expon_lr(n, a, b, max_steps=n) == pytest.approx(b, rel=1e-12)
    assert train_ref.expon_lr(n // 2, a, b, max_steps=n) == pytest.approx(math.sqrt(a * b), rel=1e-12)
    assert train_ref.expon_lr(3 * n, a, b, max_steps=n) == pytest.approx(b, rel=1e-12)   # clamped
    assert train_ref.expon_lr(-1, a, b, max_steps=n) == 0.0
    assert train_ref.expon_lr(5, 0.0, 0.0, max_steps=n) == 0.0
    # delay: factor delay_mult at step 0, 1 from delay_steps on
    assert train_ref.expon_lr(0, a, b, 100, 0.01, n) == pytest.approx(0.01 * a, rel=1e-12)
    assert train_ref.expon_lr(100, a, b, 100, 0.01, n) == pytest.approx(train_ref.expon_lr(100, a, b, max_steps=n),
                                                                      rel=1e-12)
    # one tenth of the way: lr_init^(0.9) lr_final^(0.1)
    assert train_ref.expon_lr(n // 10, a, b, max_steps=n) == pytest.approx(a ** 0.9 * b ** 0.1, rel=1e-12)


def test_adam_step_matches_torch_adam():
    rng = np.random.default_rng(0)
    shapes = [(300, 3), (300, 1), (300, 3), (300, 4), (300, 1)]
    lrs = [1.6e-4, 0.025, 0.001, 0.0001, 0.0025]
    params = [rng.normal(size=s).astype(np.float32) for s in shapes]
    tp = [torch.nn.Parameter(torch.tensor(p)) for p in params]
    opt = torch.optim.Adam([{"params": [t], "lr": lr} for t, lr in zip(tp, lrs)], lr=0.0, eps=1e-15)
    mine = [(p.copy(), np.zeros_like(p), np.zeros_like(p)) for p in params]
    for step in range(1, 8):
        grads = [(rng.normal(size=s) * 10.0 ** rng.integers(-6, 2)).astype(np.float32) for s in shapes]
        for t, g in zip(tp, grads):
            t.grad = torch.tensor(g)
        opt.step()
        mine = [train_ref.adam_step(p, g, m, v, lr, step) for (p, m, v), g, lr in zip(mine, grads, lrs)]
        for (p, m, v), t in zip(mine, tp):
            st = opt.state[t]
            # same formula, same float32 rounding up to contraction: a few ulp
            np.testing.assert_allclose(m, st["exp_avg"].numpy(), rtol=2e-6, atol=1e-30)
            np.testing.assert_allclose(v, st["exp_avg_sq"].numpy(), rtol=2e-6, atol=1e-37)
            np.testing.assert_allclose(p, t.detach().numpy(), rtol=1e-6, atol=1e-7)
    # the update is not a plain SGD step: with a tiny gradient, Adam still moves by ~lr
    p0 = np.ones((4, 1), np.float32)
    p1, _, _ = train_ref.adam_step(p0, np.full((4, 1), 1e-12, np.float32), np.zeros_like(p0), np.zeros_like(p0),
                                   0.01, 1)
    np.testing.assert_allclose(p0 - p1, 0.01, rtol=1e-3)


def test_error_order_matches_topk():
    rng = np.random.default_rng(1)
    S = 5000
    cells = rng.normal(size=S).astype(np.float32)
    gt = rng.normal(size=S).astype(np.float32)
    gt[rng.random(S) < 0.3] = -1.0
    order, nvalid = train_ref.error_order(cells, gt)
    assert nvalid == int((gt != -1).sum())
    mask = torch.tensor(gt) != -1
    diff = torch.abs(torch.tensor(cells)[mask] - torch.tensor(gt)[mask])
    k = 700
    top = torch.topk(diff, k)
    loss_idx = torch.nonzero(mask).squeeze(1)[top.indices].numpy()
    assert np.array_equal(order[:k], loss_idx)       # distinct values: same picks, same order


def test_error_order_ties_and_nan():
    cells = np.array([0.5, 0.0, 0.5, 3.0, np.nan, 0.5, 2.0], np.float32)
    gt = np.array([0.0, -1.0, 0.0, -1.0, 0.0, 0.0, 0.0], np.float32)
    order, nvalid = train_ref.error_order(cells, gt)
    assert nvalid == 5
    # NaN first (torch.topk ranks NaN above every number), then 2.0, then the tied 0.5s by index
    assert order.tolist() == [4, 6, 0, 2, 5]
    top = torch.topk(torch.tensor([0.5, 0.5, np.nan, 0.5, 2.0]), 2)
    assert math.isnan(top.values[0].item()) and top.values[1].item() == 2.0


def test_new_rows_known_answer():
    # three points on a line plus one new point: its 3-NN mean square distance is known
    xyz = np.array([[0, 0, 0], [1, 0, 0], [3, 0, 0], [0, 2, 0]], np.float32)
    rows = train_ref.new_rows(xyz, 3, [0.25], 0.1, 0.01)
    d = (4 + 5 + 13) / 3.0
    assert rows["scaling"][0, 0] == pytest.approx(0.5 * math.log(0.1 * d), rel=1e-6)
    assert np.all(rows["scaling"][0] == rows["scaling"][0, 0])
    assert rows["rotation"].tolist() == [[1, 0, 0, 0]]
    assert rows["weight"][0, 0] == pytest.approx(math.log(0.01 / 0.99), rel=1e-6)
    assert rows["value"][0, 0] == pytest.approx(math.log(0.25 / 0.75), rel=1e-6)
    # the 1e-7 floor
    close = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]], np.float32)
    assert train_ref.new_rows(close, 0, np.full(4, 0.5), 1.0, 0.1)["scaling"][0, 0] == pytest.approx(
        0.5 * math.log(1e-7), rel=1e-6)


def test_ref_model_surgery_is_the_reference_procedure():
    rng = np.random.default_rng(2)
    n = 50
    params = {g: rng.normal(size=(n, w)).astype(np.float32) for g, w in zip(train_ref.GROUPS, (3, 1, 3, 4, 1))}
    ref = train_ref.RefModel(params, [0.1] * 5, "cpu")
    for g in train_ref.GROUPS:
        ref.p[g].grad = torch.ones_like(ref.p[g])
    ref.opt.step()
    before = {g: (ref.p[g].detach().clone(), ref.moments()[g][0].clone()) for g in train_ref.GROUPS}
    rows = {g: rng.normal(size=(5, w)).astype(np.float32) for g, w in zip(train_ref.GROUPS, (3, 1, 3, 4, 1))}
    ref.cat(rows)
    keep = rng.random(n + 5) < 0.6
    ref.prune(keep)
    for g in train_ref.GROUPS:
        full = torch.cat((before[g][0], torch.tensor(rows[g])))
        assert torch.equal(ref.p[g].detach(), full[torch.tensor(keep)])
        m = torch.cat((before[g][1], torch.zeros_like(torch.tensor(rows[g]))))
        assert torch.equal(ref.moments()[g][0], m[torch.tensor(keep)])
        assert ref.moments()[g][2] == 1.0      # the step count survives the surgery
