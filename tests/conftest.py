import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gaussian-volume-compression_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libgvc.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


GRAD_NAMES = ("means", "scales", "rotations", "values", "weights")
GRAD_BAR = 1e-5  # gradients within 1e-5 relative of the FP64 oracle, normwise per parameter tensor


def grad_errors(grads, ref, names=GRAD_NAMES):
    """Relative errors of GPU gradients against the FP64 oracle's, per parameter
    tensor: {name: (normwise, elementwise)} with
      normwise    = max |g - g*| / max |g*|                       (the asserted bar);
      elementwise = max |g - g*| / max(|g*|, 1e-3 max |g*|)       (every entry
                    relative to itself, floored at 1e-3 of the tensor's largest
                    magnitude, as the field's error is floored at 1e-3)."""
    import numpy as np
    out = {}
    for n, t in zip(names, grads):
        a = (t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)).reshape(-1).astype(np.float64)
        b = np.asarray(ref[n], dtype=np.float64).reshape(-1)
        top = max(np.abs(b).max(initial=0.0), 1e-300)
        d = np.abs(a - b)
        out[n] = (float(d.max(initial=0.0) / top), float((d / np.maximum(np.abs(b), 1e-3 * top)).max(initial=0.0)))
    return out


def assert_grads(grads, ref, bar=GRAD_BAR, names=GRAD_NAMES):
    errs = grad_errors(grads, ref, names)
    worst = max(e[0] for e in errs.values())
    print("gradient errors (normwise, elementwise):", {k: (f"{v[0]:.2e}", f"{v[1]:.2e}") for k, v in errs.items()})
    assert worst <= bar, errs
    return errs
