// training.hip -- the per-step glue around the mixture, fused: parameter
// activations with their adjoint, and the fitting loss with its adjoint.
//
// Activations (R/scene/gaussian_model.py:31-48, :71-77, :114-132): scales =
// soft-capped exp(log_scales) -- s * (r_soft / r) with r = |s| + 1e-8 and
// r_soft = max_scale * tanh(r / max_scale) (max_scale <= 0: no cap) --
// rotations = q / max(|q|, 1e-12) (torch.nn.functional.normalize), weights and
// values = sigmoid.  One thread per Gaussian, forward and backward.
//
// Loss (R/train.py:389-420, R/utils/loss_utils.py:19-20):
//   recon = (gt != -1) & (cells != -1);  l1 = mean over recon of |cells - gt|,
//           or, when encode_surface, the mean over every sample (R/train.py:391-392)
//   fn    = (gt != -1) & (w > 0) & (w < t_fn),  t_fn = 0.0105 (surface) / 0.011
//   FN    = fn_reg * sum(clamp(0.011 - w[fn], 0)) / (#(vals > 0) + 1e-8)
//   FP    = fp_reg * sum(w[(gt == -1) & (w > 0)]) / (#(vals > 0) + 1e-8)   (surface only)
//   loss  = l1 + FN (+ FP when encode_surface)
// Sums are deterministic: FP64 per 256-thread block in a fixed tree order, then
// one block adds the block sums in order.  No memset, no atomics: the whole
// step can be captured in a graph.
#include <math.h>

#include "gvc_internal.h"

namespace gvc {

namespace {

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

__global__ void activate_fwd_kernel(int P, const float* __restrict__ ls, const float* __restrict__ q,
                                    const float* __restrict__ wl, const float* __restrict__ vl, float max_scale,
                                    float* __restrict__ scales, float* __restrict__ qn, float* __restrict__ w,
                                    float* __restrict__ v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float s0 = expf(ls[3 * i]), s1 = expf(ls[3 * i + 1]), s2 = expf(ls[3 * i + 2]);
    float f = 1.0f;
    if (max_scale > 0.0f) {
        const float r = sqrtf(s0 * s0 + s1 * s1 + s2 * s2) + 1e-8f;
        f = (max_scale * tanhf(r / max_scale)) / r;
    }
    scales[3 * i] = s0 * f;
    scales[3 * i + 1] = s1 * f;
    scales[3 * i + 2] = s2 * f;
    const float a = q[4 * i], b = q[4 * i + 1], c = q[4 * i + 2], d = q[4 * i + 3];
    const float n = fmaxf(sqrtf(a * a + b * b + c * c + d * d), 1e-12f);
    qn[4 * i] = a / n;
    qn[4 * i + 1] = b / n;
    qn[4 * i + 2] = c / n;
    qn[4 * i + 3] = d / n;
    w[i] = sigmoidf(wl[i]);
    v[i] = sigmoidf(vl[i]);
}

__global__ void activate_bwd_kernel(int P, const float* __restrict__ ls, const float* __restrict__ q,
                                    const float* __restrict__ wl, const float* __restrict__ vl, float max_scale,
                                    const float* __restrict__ gs, const float* __restrict__ gq,
                                    const float* __restrict__ gw, const float* __restrict__ gv,
                                    float* __restrict__ dls, float* __restrict__ dq, float* __restrict__ dwl,
                                    float* __restrict__ dvl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float s[3] = { expf(ls[3 * i]), expf(ls[3 * i + 1]), expf(ls[3 * i + 2]) };
    const float g[3] = { gs[3 * i], gs[3 * i + 1], gs[3 * i + 2] };
    float ds[3] = { g[0], g[1], g[2] };
    if (max_scale > 0.0f) {
        const float n = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        const float r = n + 1e-8f;
        const float t = tanhf(r / max_scale);
        const float f = max_scale * t / r;
        // d f / d r = (sech^2(r/M) r - M tanh(r/M)) / r^2
        const float dfdr = ((1.0f - t * t) * r - max_scale * t) / (r * r);
        const float gsum = g[0] * s[0] + g[1] * s[1] + g[2] * s[2];
        for (int k = 0; k < 3; ++k) ds[k] = g[k] * f + (n > 0.0f ? gsum * dfdr * s[k] / n : 0.0f);
    }
    for (int k = 0; k < 3; ++k) dls[3 * i + k] = ds[k] * s[k];
    const float a[4] = { q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3] };
    const float nq = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
    const float gqv[4] = { gq[4 * i], gq[4 * i + 1], gq[4 * i + 2], gq[4 * i + 3] };
    if (nq > 1e-12f) {
        float dot = 0.0f;
        for (int k = 0; k < 4; ++k) dot += a[k] * gqv[k];
        for (int k = 0; k < 4; ++k) dq[4 * i + k] = (gqv[k] - a[k] * dot / (nq * nq)) / nq;
    } else {
        for (int k = 0; k < 4; ++k) dq[4 * i + k] = gqv[k] / 1e-12f;
    }
    const float w = sigmoidf(wl[i]), v = sigmoidf(vl[i]);
    dwl[i] = gw[i] * w * (1.0f - w);
    dvl[i] = gv[i] * v * (1.0f - v);
}

// state: block partials [nb][6] (double: l1 sum, l1 count, fn sum, fn count,
// fp sum, fp count), then totals [6].
struct LossLayout {
    double* part;
    double* tot;
    LossLayout(Arena& a, int S)
    {
        const size_t nb = blocks_for((size_t)(S > 0 ? S : 1), 256);
        part = a.take<double>(6 * nb);
        tot = a.take<double>(8);
    }
};

struct LossTerms {
    bool recon, fn, fp;
    double l1, fnv, fpv;
};

// recon: the sample enters the L1 mean -- every sample when encode_surface
// (torch.abs(cells - gt).mean(), R/train.py:391-392), else the masked ones
// (R/train.py:389, :394).
__device__ __forceinline__ LossTerms loss_terms(float c, float w, float g, int surface)
{
    LossTerms t;
    t.recon = surface || ((g != -1.0f) && (c != -1.0f));
    t.l1 = t.recon ? fabs((double)(c - g)) : 0.0;
    const float tf = surface ? 0.0105f : 0.011f;
    t.fn = (g != -1.0f) && (w > 0.0f) && (w < tf);
    t.fnv = t.fn ? (double)fmaxf(0.011f - w, 0.0f) : 0.0;
    t.fp = surface && (g == -1.0f) && (w > 0.0f);
    t.fpv = t.fp ? (double)w : 0.0;
    return t;
}

__global__ __launch_bounds__(256) void loss_part_kernel(int S, const float* __restrict__ cells,
                                                        const float* __restrict__ weights,
                                                        const float* __restrict__ gt, int surface,
                                                        double* __restrict__ part)
{
    __shared__ double red[6][256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6] = { 0, 0, 0, 0, 0, 0 };
    if (i < S) {
        const LossTerms t = loss_terms(cells[i], weights[i], gt[i], surface);
        v[0] = t.l1;
        v[1] = t.recon ? 1.0 : 0.0;
        v[2] = t.fnv;
        v[3] = t.fnv > 0.0 ? 1.0 : 0.0;
        v[4] = t.fpv;
        v[5] = t.fpv > 0.0 ? 1.0 : 0.0;
    }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void loss_final_kernel(int nb, const double* __restrict__ part, float fn_reg,
                                                         float fp_reg, int surface, double* __restrict__ tot,
                                                         float* __restrict__ loss)
{
    __shared__ double red[6][256];
    double v[6] = { 0, 0, 0, 0, 0, 0 };
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int k = 0; k < 6; ++k) v[k] += part[6 * (size_t)b + k];
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; ++k) tot[k] = red[k][0];
        // an empty reconstruction mask gives NaN, like the mean of an empty tensor
        double L = red[0][0] / red[1][0];
        L += (double)fn_reg * red[2][0] / (red[3][0] + 1e-8);
        if (surface) L += (double)fp_reg * red[4][0] / (red[5][0] + 1e-8);
        *loss = (float)L;
    }
}

__global__ void loss_bwd_kernel(int S, const float* __restrict__ cells, const float* __restrict__ weights,
                                const float* __restrict__ gt, float fn_reg, float fp_reg, int surface,
                                const double* __restrict__ tot, const float* __restrict__ upstream,
                                float* __restrict__ dcells, float* __restrict__ dweights)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const float up = upstream ? *upstream : 1.0f;
    const float c = cells[i], w = weights[i], g = gt[i];
    const LossTerms t = loss_terms(c, w, g, surface);
    float dc = 0.0f, dw = 0.0f;
    if (t.recon) {
        const float d = c - g;
        const float sgn = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
        dc = up * sgn * (float)(1.0 / tot[1]);
    }
    if (t.fn && 0.011f - w >= 0.0f) dw -= up * (float)((double)fn_reg / (tot[3] + 1e-8));
    if (t.fp) dw += up * (float)((double)fp_reg / (tot[5] + 1e-8));
    dcells[i] = dc;
    dweights[i] = dw;
}

}  // namespace

int activate_forward(int P, const float* ls, const float* q, const float* wl, const float* vl, float max_scale,
                     float* scales, float* qn, float* w, float* v, hipStream_t st)
{
    if (P <= 0) return GVC_OK;
    activate_fwd_kernel<<<blocks_for(P, 256), 256, 0, st>>>(P, ls, q, wl, vl, max_scale, scales, qn, w, v);
    GVC_LAUNCH_CHECK("activate_fwd");
    return GVC_OK;
}

int activate_backward(int P, const float* ls, const float* q, const float* wl, const float* vl, float max_scale,
                      const float* gs, const float* gq, const float* gw, const float* gv, float* dls, float* dq,
                      float* dwl, float* dvl, hipStream_t st)
{
    if (P <= 0) return GVC_OK;
    activate_bwd_kernel<<<blocks_for(P, 256), 256, 0, st>>>(P, ls, q, wl, vl, max_scale, gs, gq, gw, gv, dls, dq,
                                                            dwl, dvl);
    GVC_LAUNCH_CHECK("activate_bwd");
    return GVC_OK;
}

size_t loss_state_bytes(int S)
{
    Arena a(nullptr, 0);
    LossLayout l(a, S);
    return a.off + 256;
}

int loss_forward(int S, const float* cells, const float* weights, const float* gt, float fn_reg, float fp_reg,
                 int surface, float* loss, void* state, size_t bytes, hipStream_t st)
{
    Arena a(state, bytes);
    LossLayout l(a, S);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "loss state buffer too small");
    const int nb = (int)blocks_for((size_t)(S > 0 ? S : 1), 256);
    loss_part_kernel<<<nb, 256, 0, st>>>(S, cells, weights, gt, surface, l.part);
    GVC_LAUNCH_CHECK("loss_part");
    loss_final_kernel<<<1, 256, 0, st>>>(nb, l.part, fn_reg, fp_reg, surface, l.tot, loss);
    GVC_LAUNCH_CHECK("loss_final");
    return GVC_OK;
}

int loss_backward(int S, const float* cells, const float* weights, const float* gt, float fn_reg, float fp_reg,
                  int surface, const float* upstream, const void* state, size_t bytes, float* dcells,
                  float* dweights, hipStream_t st)
{
    if (S <= 0) return GVC_OK;
    Arena a(const_cast<void*>(state), bytes);
    LossLayout l(a, S);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "loss state buffer too small");
    loss_bwd_kernel<<<blocks_for(S, 256), 256, 0, st>>>(S, cells, weights, gt, fn_reg, fp_reg, surface, l.tot,
                                                        upstream, dcells, dweights);
    GVC_LAUNCH_CHECK("loss_bwd");
    return GVC_OK;
}

}  // namespace gvc
