"""torch float64 restatement of the forward formula -- TEST INFRASTRUCTURE ONLY.

Differentiating this with autograd is the gradient oracle.  It follows
CR/forward.cu:10-76 (covariance, epsilon regularisation, adjugate inverse),
CR/forward.cu:229-239 (quadratic form, window) and CR/forward.cu:249-256
(normalise / sentinel), with CR/ = R/diff-gaussian-rasterization/cuda_rasterizer/.

The pair-membership predicate is FROZEN: ``mask`` (P, S) comes from the canonical
fp32 predicate (``oracle.pair_mask``) and is treated as a constant, exactly as the
hand-written adjoint treats the box and the window.  The epsilon = 1e-5 max|Sigma_ii|
term picks the first index on ties (its subgradient convention).
"""
import numpy as np
import torch


def rotation(q):
    """Standard rotation matrix of unnormalised q = (r, x, y, z), CR/forward.cu:39-44."""
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y),
        2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x),
        2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y),
    ], dim=1).reshape(-1, 3, 3)
    return R


def conic(scales, rots, scale_modifier=1.0):
    """Packed inverse of Sigma + eps I through the adjugate (CR/forward.cu:46-76)."""
    R = rotation(rots)
    s = scale_modifier * scales
    M = s[:, :, None] * R.transpose(1, 2)          # M[a, i] = s_a R[i, a]
    Sig = M.transpose(1, 2) @ M
    diag = torch.stack([Sig[:, 0, 0], Sig[:, 1, 1], Sig[:, 2, 2]], dim=1)
    j = torch.argmax(diag.detach().abs(), dim=1, keepdim=True)  # first index on ties
    eps = diag.abs().gather(1, j)[:, 0] * 1e-5
    a = Sig[:, 0, 0] + eps
    b = Sig[:, 0, 1]
    c = Sig[:, 0, 2]
    d = Sig[:, 1, 1] + eps
    e = Sig[:, 1, 2]
    f = Sig[:, 2, 2] + eps
    A00 = d * f - e * e
    A01 = c * e - b * f
    A02 = b * e - c * d
    A11 = a * f - c * c
    A12 = b * c - a * e
    A22 = a * d - b * b
    det = a * A00 + b * A01 + c * A02
    return torch.stack([A00, A01, A02, A11, A12, A22], dim=1) / det[:, None]


def forward(means, scales, rots, values, weights, samples, mask, scale_modifier=1.0):
    """Field (V, W) at the samples; sentinel (-1, 0) where W <= 1e-2."""
    C = conic(scales, rots, scale_modifier)
    d = samples[None, :, :] - means[:, None, :]                       # (P, S, 3)
    r0 = C[:, None, 0] * d[..., 0] + C[:, None, 1] * d[..., 1] + C[:, None, 2] * d[..., 2]
    r1 = C[:, None, 1] * d[..., 0] + C[:, None, 3] * d[..., 1] + C[:, None, 4] * d[..., 2]
    r2 = C[:, None, 2] * d[..., 0] + C[:, None, 4] * d[..., 1] + C[:, None, 5] * d[..., 2]
    quad = d[..., 0] * r0 + d[..., 1] * r1 + d[..., 2] * r2
    wgt = weights.reshape(-1, 1) * torch.exp(-0.5 * quad) * mask
    W = wgt.sum(0)
    U = (wgt * values.reshape(-1, 1)).sum(0)
    valid = W > 1e-2
    Wsafe = torch.where(valid, W, torch.ones_like(W))
    V = torch.where(valid, U / Wsafe, torch.full_like(W, -1.0))
    Wout = torch.where(valid, W, torch.zeros_like(W))
    return V, Wout


def _t(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64))


def grads(means, scales, rots, values, weights, samples, mask, gV, gW, scale_modifier=1.0):
    """Autograd gradients of L = sum(gV V + gW W) (float64, frozen predicate)."""
    params = [_t(x).clone().requires_grad_(True) for x in (means, scales, rots, values, weights)]
    V, W = forward(*params, _t(samples), _t(mask), scale_modifier)
    L = (V * _t(gV)).sum() + (W * _t(gW)).sum()
    L.backward()
    names = ("means", "scales", "rotations", "values", "weights")
    out = {n: p.grad.numpy() for n, p in zip(names, params)}
    out["V"] = V.detach().numpy()
    out["W"] = W.detach().numpy()
    return out


def loss_value(means, scales, rots, values, weights, samples, mask, gV, gW, scale_modifier=1.0):
    with torch.no_grad():
        V, W = forward(_t(means), _t(scales), _t(rots), _t(values), _t(weights), _t(samples),
                       _t(mask), scale_modifier)
        return float((V * _t(gV)).sum() + (W * _t(gW)).sum())


def central_difference(params, which, index, h, samples, mask, gV, gW, scale_modifier=1.0):
    """Frozen-predicate central difference dL/dparams[which][index] with step h."""
    p_plus = [np.array(p, dtype=np.float64, copy=True) for p in params]
    p_minus = [np.array(p, dtype=np.float64, copy=True) for p in params]
    p_plus[which].reshape(-1)[index] += h
    p_minus[which].reshape(-1)[index] -= h
    lp = loss_value(*p_plus, samples, mask, gV, gW, scale_modifier)
    lm = loss_value(*p_minus, samples, mask, gV, gW, scale_modifier)
    return (lp - lm) / (2 * h)
