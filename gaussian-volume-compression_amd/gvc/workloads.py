"""Seeded synthetic workloads of the benchmark configurations (numpy, host side).

Draw order is part of the definition (numpy.random.default_rng(seed)):

config0  2,048 Gaussians: means U[0,1)^3, log-scales N(ln 0.03, 0.3^2) per axis,
         rotations = normalised N(0,1)^4 (uniform on S^3, stored r,x,y,z),
         weights U(0,1), values U(0,1); samples = the 32^3 'ij' grid of [0,1]^3
         (as R/train.py:151-155); target f(x) = sum_k sin(2 pi a_k . x),
         a_k ~ U{1..4}^3; masked L1 loss (R/train.py:389-394).
blob_field  sum of random anisotropic blobs with the config0 distributions.
config1  128^3 blob field (64 blobs) + N(0, 0.01^2) noise; 200,000 Gaussians at
         random distinct voxel centres (R/scene/dataset_readers.py:29-53),
         identity rotations, weight 0.1, value = the voxel's field value,
         scales from the 3-NN distance (R/scene/gaussian_model.py:175-188);
         2^20 uniform samples per step, trilinear ground truth.
config2  129^3 lattice, spacing 1/128, interior vertices jittered U(+-0.3/128)
         (boundary vertices stay on the faces so the domain is [0,1]^3), each cube
         split into 6 Kuhn tetrahedra along its (0,0,0)-(1,1,1) diagonal;
         vertex scalar = 32-blob field; 1,000,000 Gaussians at random distinct
         vertices, identity rotations, weight 0.1, value = vertex scalar; 2^22
         uniform samples per step located in the tet mesh.
"""
import numpy as np

LOG_SCALE_MEAN = float(np.log(0.03))
LOG_SCALE_STD = 0.3


def random_gaussians(rng, P):
    means = rng.uniform(0.0, 1.0, (P, 3)).astype(np.float32)
    log_scales = rng.normal(LOG_SCALE_MEAN, LOG_SCALE_STD, (P, 3)).astype(np.float32)
    q = rng.normal(size=(P, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    rots = q.astype(np.float32)
    weights = rng.uniform(0.0, 1.0, (P, 1)).astype(np.float32)
    values = rng.uniform(0.0, 1.0, (P, 1)).astype(np.float32)
    return dict(means=means, log_scales=log_scales, rotations=rots, weights=weights, values=values)


def grid_points(n):
    g = np.linspace(0.0, 1.0, n, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float32)


def config0(seed=0):
    rng = np.random.default_rng(seed)
    g = random_gaussians(rng, 2048)
    a = rng.integers(1, 5, size=(3, 3))
    pts = grid_points(32)
    x = pts.astype(np.float64)
    target = sum(np.sin(2.0 * np.pi * (x @ a[k].astype(np.float64))) for k in range(3))
    g.update(samples=pts, target=target.astype(np.float32), freqs=a)
    return g


def _rot_np(q):
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y),
        2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x),
        2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y),
    ], axis=1).reshape(-1, 3, 3)


def blobs(rng, n):
    g = random_gaussians(rng, n)
    R = _rot_np(g["rotations"].astype(np.float64))
    s = np.exp(g["log_scales"].astype(np.float64))
    cov = R @ (s[:, :, None] ** 2 * np.transpose(R, (0, 2, 1)))
    inv = np.linalg.inv(cov)
    return dict(means=g["means"].astype(np.float64), inv=inv, w=g["weights"][:, 0].astype(np.float64))


def eval_blobs(b, pts, chunk=1 << 18):
    """sum_b w_b exp(-0.5 (x - mu_b)^T Sigma_b^-1 (x - mu_b)) in float64."""
    pts = np.asarray(pts, dtype=np.float64)
    out = np.zeros(pts.shape[0])
    for s in range(0, pts.shape[0], chunk):
        x = pts[s:s + chunk]
        acc = np.zeros(x.shape[0])
        for k in range(b["means"].shape[0]):
            d = x - b["means"][k]
            q = np.einsum("ni,ij,nj->n", d, b["inv"][k], d)
            acc += b["w"][k] * np.exp(-0.5 * q)
        out[s:s + chunk] = acc
    return out


def knn3_mean_sq(points):
    """Mean squared distance to the 3 nearest other points (scipy KD-tree)."""
    from scipy.spatial import cKDTree
    t = cKDTree(points.astype(np.float64))
    d, _ = t.query(points.astype(np.float64), k=4)
    return (d[:, 1:] ** 2).mean(axis=1)


def config1(seed=0, P=200_000, n=128, S=1 << 20):
    rng = np.random.default_rng(seed)
    b = blobs(rng, 64)
    grid = grid_points(n)
    field = eval_blobs(b, grid) + rng.normal(0.0, 0.01, grid.shape[0])
    vox = rng.choice(grid.shape[0], size=P, replace=False)
    means = grid[vox].astype(np.float32)
    d2 = np.maximum(knn3_mean_sq(means), 1e-7)
    log_s = np.log(np.sqrt(d2)).astype(np.float32)
    log_scales = np.repeat(log_s[:, None], 3, axis=1)
    rots = np.zeros((P, 4), np.float32)
    rots[:, 0] = 1.0
    weights = np.full((P, 1), 0.1, np.float32)
    values = field[vox].astype(np.float32)[:, None]
    samples = rng.uniform(0.0, 1.0, (S, 3)).astype(np.float32)
    return dict(means=means, log_scales=log_scales, rotations=rots, weights=weights, values=values,
                field=field.astype(np.float32).reshape(n, n, n), samples=samples, dims=(n, n, n))


def kuhn_tets(n):
    """Connectivity [6 n^3, 4] of the Kuhn split of an n^3-cube lattice whose
    (n+1)^3 vertices are numbered i + (n+1) (j + (n+1) k)."""
    m = n + 1
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (i + m * (j + m * k)).ravel().astype(np.int64)
    e = [1, m, m * m]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for p in perms:
        a = base
        b = a + e[p[0]]
        c = b + e[p[1]]
        d = base + e[0] + e[1] + e[2]
        tets.append(np.stack([a, b, c, d], axis=1))
    return np.stack(tets, axis=1).reshape(-1, 4)


def jittered_lattice(rng, n, jitter=0.3):
    m = n + 1
    g = np.arange(m, dtype=np.float64) / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(), Z.transpose(2, 1, 0).ravel()], 1)
    # vertex id = i + m (j + m k): x fastest
    jit = rng.uniform(-jitter / n, jitter / n, pts.shape)
    on_face = (pts <= 0.0) | (pts >= 1.0)
    jit[on_face] = 0.0
    return (pts + jit).astype(np.float32)


def config2(seed=0, n=128, P=1_000_000, S=1 << 22, n_blobs=32):
    rng = np.random.default_rng(seed)
    pts = jittered_lattice(rng, n)
    tets = kuhn_tets(n)
    b = blobs(rng, n_blobs)
    vals = eval_blobs(b, pts)
    vsel = rng.choice(pts.shape[0], size=min(P, pts.shape[0]), replace=False)
    means = pts[vsel]
    d2 = np.maximum(knn3_mean_sq(means), 1e-7)
    log_s = np.log(np.sqrt(d2)).astype(np.float32)
    log_scales = np.repeat(log_s[:, None], 3, axis=1)
    Pn = means.shape[0]
    rots = np.zeros((Pn, 4), np.float32)
    rots[:, 0] = 1.0
    weights = np.full((Pn, 1), 0.1, np.float32)
    values = vals[vsel].astype(np.float32)[:, None]
    samples = rng.uniform(0.0, 1.0, (S, 3)).astype(np.float32)
    return dict(means=means, log_scales=log_scales, rotations=rots, weights=weights, values=values,
                mesh_points=pts, tets=tets, mesh_values=vals, samples=samples)


# ---------------------------------------------------------------------------
# device-side builders used by bench.py (same definitions, torch on the GPU)
# ---------------------------------------------------------------------------
def _eval_blobs_torch(b, pts, device):
    import torch
    x = torch.as_tensor(pts, dtype=torch.float64, device=device)
    mu = torch.as_tensor(b["means"], device=device)
    inv = torch.as_tensor(b["inv"], device=device)
    w = torch.as_tensor(b["w"], device=device)
    out = torch.zeros(x.shape[0], dtype=torch.float64, device=device)
    for k in range(mu.shape[0]):
        d = x - mu[k]
        q = ((d @ inv[k]) * d).sum(1)
        out += w[k] * torch.exp(-0.5 * q)
    return out


def _init_gaussians(means, values, device):
    """R/scene/gaussian_model.py:175-188: 3-NN scales, identity rotations, w = 0.1."""
    import torch
    from simple_knn._C import distCUDA2
    m = torch.as_tensor(means, dtype=torch.float32, device=device).contiguous()
    d2 = torch.clamp_min(distCUDA2(m), 1e-7)
    log_s = torch.log(torch.sqrt(d2))[:, None].repeat(1, 3).contiguous()
    P = m.shape[0]
    rots = torch.zeros((P, 4), dtype=torch.float32, device=device)
    rots[:, 0] = 1.0
    weights = torch.full((P, 1), 0.1, dtype=torch.float32, device=device)
    vals = torch.as_tensor(values, dtype=torch.float32, device=device).reshape(P, 1).contiguous()
    return dict(means=m, log_scales=log_s, rotations=rots, weights=weights, values=vals, P=int(P))


def mesh_workload(device, seed=0, n=128, P=1_000_000, n_blobs=32):
    """config2 on the device: mesh (numpy), Gaussians (torch)."""
    rng = np.random.default_rng(seed)
    pts = jittered_lattice(rng, n)
    tets = kuhn_tets(n).astype(np.int32)
    b = blobs(rng, n_blobs)
    vals = _eval_blobs_torch(b, pts, device)
    vals = (vals - vals.min()) / (vals.max() - vals.min())   # R/scene/dataset_readers.py:10-17
    vsel = rng.choice(pts.shape[0], size=min(P, pts.shape[0]), replace=False)
    g = _init_gaussians(pts[vsel], vals.cpu().numpy()[vsel], device)
    g.update(mesh_points=pts, tets=tets, mesh_values=vals)
    return g


def grid_workload(device, seed=0, n=128, P=200_000):
    """config1 on the device; grid_fn(samples) is the trilinear ground truth."""
    import torch
    from gvc.sampling import grid_sample
    rng = np.random.default_rng(seed)
    b = blobs(rng, 64)
    grid = grid_points(n)
    field = _eval_blobs_torch(b, grid, device) + torch.as_tensor(rng.normal(0.0, 0.01, grid.shape[0]), device=device)
    field = field.float()
    vox = rng.choice(grid.shape[0], size=P, replace=False)
    g = _init_gaussians(grid[vox], field.cpu().numpy()[vox], device)
    # grid_points uses 'ij' order (x slowest); the probe wants x fastest
    f_xfast = field.reshape(n, n, n).permute(2, 1, 0).contiguous().reshape(-1)
    dims = np.array([n, n, n], np.int64)
    origin = np.zeros(3, np.float32)
    spacing = np.full(3, 1.0 / (n - 1), np.float32)

    def grid_fn(samples):
        return grid_sample(dims, origin, spacing, f_xfast, samples)

    g.update(grid_fn=grid_fn, field=field)
    return g


def batch_workload(device, seed=0, volumes=32, total_points=1 << 24, point_batches=2, n_blobs=8):
    """32 independent synthetic volumes of uneven size, fitted side by side.

    Per volume v (draw order: sides, Gaussian counts, then per volume blobs,
    Gaussian means, point batches): side_v = round(exp U[ln 32, ln 256]); P_v =
    round(exp U[ln 1e4, ln 2e6]); points per step proportional to side_v^3 with
    sum exactly ``total_points`` (largest remainders).  The field of a volume is
    a sum of ``n_blobs`` random blobs (the small-config distributions), scaled
    to [0, 1] by its maximum over the volume's points; Gaussians start at
    uniform random positions with the reference's initialisation (3-NN scales,
    identity rotations, weight 0.1, value = field at the mean).  Like the
    reference's precomputed point batches (R/train.py:240-289), ``point_batches``
    batches of points and their ground truth are drawn once; step k uses batch
    k mod point_batches.  Every volume lives in its own unit cube."""
    import torch
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(32.0), np.log(256.0), volumes))).astype(np.int64)
    P = np.rint(np.exp(rng.uniform(np.log(1e4), np.log(2e6), volumes))).astype(np.int64)
    vox = sides.astype(np.float64) ** 3
    share = total_points * vox / vox.sum()
    S = np.floor(share).astype(np.int64)
    rest = total_points - int(S.sum())
    S[np.argsort(-(share - S), kind="stable")[:rest]] += 1
    parts = []
    pts = [[] for _ in range(point_batches)]
    gts = [[] for _ in range(point_batches)]
    for v in range(volumes):
        b = blobs(rng, n_blobs)
        means = rng.uniform(0.0, 1.0, (int(P[v]), 3)).astype(np.float32)
        xs = [rng.uniform(0.0, 1.0, (int(S[v]), 3)).astype(np.float32) for _ in range(point_batches)]
        fx = [_eval_blobs_torch(b, x, device) for x in xs]
        fm = _eval_blobs_torch(b, means, device)
        top = max([float(f.max()) for f in fx if f.numel()] + [float(fm.max()), 1e-12])
        parts.append(_init_gaussians(means, (fm / top).float().cpu().numpy(), device))
        for k in range(point_batches):
            pts[k].append(torch.as_tensor(xs[k], device=device))
            gts[k].append((fx[k] / top).float())
    cat = {k: torch.cat([p[k] for p in parts]).contiguous()
           for k in ("means", "log_scales", "rotations", "weights", "values")}
    cat.update(sides=sides, gaussians=P, samples=S, P=int(P.sum()), S=int(S.sum()),
               point_batches=[torch.cat(x).contiguous() for x in pts],
               ground_truth=[torch.cat(g).contiguous() for g in gts])
    return cat


def mixed_cell_mesh(n=4, seed=0):
    """A VTK mesh of mixed 3D cells with planar faces: an n^3-cube lattice under a
    fixed affine map, each cube (uniformly at random) one hexahedron (12), two
    wedges (13), six pyramids around an added centre point (14) or six Kuhn
    tetrahedra (10).  Returns (points [N,3] float32, conn int64, types int64,
    offsets int64 [C] (start of each cell in conn))."""
    rng = np.random.default_rng(seed)
    g = np.arange(n + 1, dtype=np.float64) / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    A = np.array([[1.0, 0.15, -0.1], [0.05, 0.9, 0.2], [-0.1, 0.1, 1.1]])
    b = np.array([0.05, -0.02, 0.03])
    pts = [base @ A.T + b]

    def vid(i, j, k):
        return (i * (n + 1) + j) * (n + 1) + k

    conn, types = [], []
    extra = 0
    N0 = base.shape[0]
    kuhn = [[0, 1, 3, 7], [0, 1, 5, 7], [0, 2, 3, 7], [0, 2, 6, 7], [0, 4, 5, 7], [0, 4, 6, 7]]
    for i in range(n):
        for j in range(n):
            for k in range(n):
                # binary corner order c = dx + 2 dy + 4 dz, and VTK hexahedron order
                c = [vid(i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1)) for q in range(8)]
                hexa = [c[0], c[1], c[3], c[2], c[4], c[5], c[7], c[6]]
                kind = int(rng.integers(0, 4))
                if kind == 0:
                    conn.append(hexa); types.append(12)
                elif kind == 1:
                    conn.append([hexa[0], hexa[1], hexa[2], hexa[4], hexa[5], hexa[6]]); types.append(13)
                    conn.append([hexa[0], hexa[2], hexa[3], hexa[4], hexa[6], hexa[7]]); types.append(13)
                elif kind == 2:
                    centre = N0 + extra
                    extra += 1
                    pts.append((np.mean(base[c], axis=0) @ A.T + b)[None, :])
                    for face in ([0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4], [3, 2, 6, 7], [0, 3, 7, 4],
                                 [1, 2, 6, 5]):
                        conn.append([hexa[f] for f in face] + [centre]); types.append(14)
                else:
                    for t in kuhn:
                        conn.append([c[q] for q in t]); types.append(10)
    flat = np.concatenate([np.asarray(x, np.int64) for x in conn])
    lens = np.array([len(x) for x in conn], np.int64)
    offsets = np.concatenate([[0], np.cumsum(lens)[:-1]])
    return np.concatenate(pts).astype(np.float32), flat, np.asarray(types, np.int64), offsets


def precision_workload(device, seed=0, n=512, P=4_000_000, n_blobs=256, degenerate=0.02, stress=0.005):
    """The precision-study workload: an n^3 float32 field (a sum of ``n_blobs``
    random blobs, scaled to [0, 1], plus three sinusoid octaves of amplitude 0.1,
    0.05, 0.025 at 4, 8, 16 cycles per unit along random unit directions with
    random phases) and P Gaussians at uniform random positions initialised like
    the reference (3-NN scales, identity rotations, weight 0.1, value = field at
    the mean).  A fraction ``degenerate`` of them gets one log-scale set to
    ln 1e-4 (scale ratio about 35, below the FP64 threshold); a further
    ``stress`` fraction gets ln 1e-6 (ratio above 1000, i.e. cond(Sigma) > 1e6);
    both sets get random rotations.
    Returns the Gaussians (torch), the field (x fastest, for gvc_grid_sample) and
    the grid geometry; field_fn(samples) samples it trilinearly."""
    import torch
    from gvc.sampling import grid_sample
    rng = np.random.default_rng(seed)
    b = blobs(rng, n_blobs)
    dirs = rng.normal(size=(3, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    phases = rng.uniform(0, 2 * np.pi, 3)
    g = torch.arange(n, device=device, dtype=torch.float32) / (n - 1)
    field = torch.zeros((n, n, n), dtype=torch.float32, device=device)   # [z, y, x]
    mu = torch.as_tensor(b["means"], dtype=torch.float32, device=device)
    inv = torch.as_tensor(b["inv"], dtype=torch.float32, device=device)
    w = torch.as_tensor(b["w"], dtype=torch.float32, device=device)
    Z, Y, X = torch.meshgrid(g, g, g, indexing="ij")
    for k in range(n_blobs):
        dx, dy, dz = X - mu[k, 0], Y - mu[k, 1], Z - mu[k, 2]
        q = (inv[k, 0, 0] * dx * dx + inv[k, 1, 1] * dy * dy + inv[k, 2, 2] * dz * dz +
             2 * (inv[k, 0, 1] * dx * dy + inv[k, 0, 2] * dx * dz + inv[k, 1, 2] * dy * dz))
        field += w[k] * torch.exp(-0.5 * q)
    field = (field - field.min()) / (field.max() - field.min())
    for k, (amp, freq) in enumerate(((0.1, 4.0), (0.05, 8.0), (0.025, 16.0))):
        d = dirs[k]
        field += amp * torch.sin(2 * np.pi * freq * (float(d[0]) * X + float(d[1]) * Y + float(d[2]) * Z) +
                                 float(phases[k]))
    del X, Y, Z
    flat = field.reshape(-1).contiguous()
    dims = np.array([n, n, n], np.int64)
    origin = np.zeros(3, np.float32)
    spacing = np.full(3, 1.0 / (n - 1), np.float32)

    def field_fn(samples):
        return grid_sample(dims, origin, spacing, flat, samples)

    means = rng.uniform(0.0, 1.0, (P, 3)).astype(np.float32)
    m = torch.as_tensor(means, device=device)
    vals = field_fn(m).double().cpu().numpy()
    out = _init_gaussians(means, vals, device)
    nd, ns = int(round(degenerate * P)), int(round(stress * P))
    pick = rng.permutation(P)[:nd + ns]
    axis = torch.as_tensor(rng.integers(0, 3, nd + ns), device=device)
    rows = torch.as_tensor(pick, device=device)
    target = torch.full((nd + ns,), float(np.log(1e-4)), device=device)
    target[nd:] = float(np.log(1e-6))
    out["log_scales"][rows, axis] = target
    # rotated, as after training: a diagonal covariance inverts exactly in fp32
    q = rng.normal(size=(nd + ns, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    out["rotations"][rows] = torch.as_tensor(q, dtype=torch.float32, device=device)
    out.update(field=flat, field_fn=field_fn, degenerate=pick[:nd], stress=pick[nd:])
    return out
