// optim.hip -- the optimiser and the model surgery of the training loop, over
// fixed-capacity parameter buffers.
//
// Adam (R/scene/gaussian_model.py:207-247: torch.optim.Adam, betas (0.9, 0.999),
// eps 1e-15, one group per parameter tensor).  One launch updates every tensor of
// the model; the update is torch's multi-tensor Adam step written per element:
//   m <- fma(1 - b1, g - m, m)           (lerp)
//   v <- v b2 + (1 - b2) g g
//   p <- p + step_size * m / (sqrt(v) / sqrt(1 - b2^t) + eps),
//   step_size = -lr / (1 - b1^t)
// with the scalars formed in double and rounded to float once, as torch forms
// them from Python floats.  The step count t and the learning rates live in a
// caller-owned device clock, advanced by device code, so a step captured in a
// graph keeps counting on replay.  The position learning rate follows the
// exponential schedule of R/utils/general_utils.py:35-69 (get_expon_lr_func),
// evaluated on the device from the clock's iteration (R/train.py:367).
//
// Surgery (R/scene/gaussian_model.py:396-612, :727-799; R/train.py:551-590):
//   - the error order: indices of the samples with gt != -1 by descending
//     |cells - gt| (ties: ascending index), the reference's torch.topk pick, with
//     the two counts its pick size needs (valid samples, Gaussians of weight
//     <= 0.005);
//   - densify_in_empty: rows [count, count + k) of every parameter take the
//     picked samples' positions, log-scales log(sqrt(max(knn3 * f, 1e-7))) on all
//     three axes (knn3 over the count + k positions; f = 0.1 there, 1 for the
//     initial cloud of create_from_pcd, :139-205), the identity rotation, the
//     weight logit of the initial weight, and the value logit of the sample's
//     ground truth; their Adam moments are zero (cat_tensors_to_optimizer);
//   - prune: rows whose weight sigmoid(logit) < min_weight (or a caller mask) are
//     removed from every parameter and every moment, the survivors keeping their
//     order (_prune_optimizer), through a staging copy in the workspace.
#include <math.h>

#include "gvc_internal.h"

namespace gvc {

namespace {

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

constexpr int kAdamThreads = 256;
constexpr int kAdamUnroll = 8;
constexpr int kAdamBlockElems = kAdamThreads * kAdamUnroll;

struct AdamArgs {
    float* p[GVC_ADAM_MAX_TENSORS];
    const float* g[GVC_ADAM_MAX_TENSORS];
    float* m[GVC_ADAM_MAX_TENSORS];
    float* v[GVC_ADAM_MAX_TENSORS];
    uint64_t numel[GVC_ADAM_MAX_TENSORS];
    uint32_t first_block[GVC_ADAM_MAX_TENSORS + 1];
    int lr_index[GVC_ADAM_MAX_TENSORS];
    int n;
};

// One thread: t += 1, then the per-group scalars of this step.
__global__ void adam_tick_kernel(gvc_train_clock* clock, double beta1, double beta2)
{
    if (threadIdx.x != 0) return;
    const int64_t t = clock->step + 1;
    clock->step = t;
    const double bc1 = 1.0 - pow(beta1, (double)t);
    const double bc2 = 1.0 - pow(beta2, (double)t);
    for (int k = 0; k < GVC_ADAM_MAX_TENSORS; ++k) clock->step_size[k] = (float)(-(clock->lr[k] / bc1));
    clock->bc2_sqrt = (float)sqrt(bc2);
}

__global__ __launch_bounds__(kAdamThreads) void adam_kernel(AdamArgs a, const gvc_train_clock* __restrict__ clock,
                                                            float w1, float b2, float c2, float eps)
{
    const uint32_t b = blockIdx.x;
    int t = 0;
    while (t + 1 < a.n && b >= a.first_block[t + 1]) ++t;
    const float ss = clock->step_size[a.lr_index[t]];
    const float bc2s = clock->bc2_sqrt;
    float* __restrict__ p = a.p[t];
    const float* __restrict__ g = a.g[t];
    float* __restrict__ m = a.m[t];
    float* __restrict__ v = a.v[t];
    const uint64_t n = a.numel[t];
    const uint64_t base = (uint64_t)(b - a.first_block[t]) * kAdamBlockElems + threadIdx.x;
    float gr[kAdamUnroll], mr[kAdamUnroll], vr[kAdamUnroll], pr[kAdamUnroll];
#pragma unroll
    for (int u = 0; u < kAdamUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kAdamThreads;
        const bool ok = i < n;
        gr[u] = ok ? g[i] : 0.0f;
        mr[u] = ok ? m[i] : 0.0f;
        vr[u] = ok ? v[i] : 0.0f;
        pr[u] = ok ? p[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kAdamUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kAdamThreads;
        if (i >= n) continue;
        const float gg = gr[u];
        const float mm = fmaf(w1, gg - mr[u], mr[u]);   // torch's lerp
        const float vv = vr[u] * b2 + c2 * gg * gg;
        const float den = sqrtf(vv) / bc2s + eps;
        m[i] = mm;
        v[i] = vv;
        p[i] = pr[u] + ss * (mm / den);
    }
}

// get_expon_lr_func(lr_init, lr_final, delay_steps, delay_mult, max_steps)(step)
__device__ double expon_lr(double step, const gvc_lr_schedule& s)
{
    if (step < 0.0 || (s.lr_init == 0.0 && s.lr_final == 0.0)) return 0.0;
    double delay = 1.0;
    if (s.delay_steps > 0) {
        double r = step / (double)s.delay_steps;
        r = r < 0.0 ? 0.0 : (r > 1.0 ? 1.0 : r);
        delay = s.delay_mult + (1.0 - s.delay_mult) * sin(0.5 * M_PI * r);
    }
    double t = step / (double)s.max_steps;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return delay * exp(log(s.lr_init) * (1.0 - t) + log(s.lr_final) * t);
}

__global__ void lr_advance_kernel(gvc_train_clock* clock, gvc_lr_schedule s, int lr_index, int64_t set_iteration)
{
    if (threadIdx.x != 0) return;
    const int64_t it = set_iteration >= 0 ? set_iteration : clock->iteration + 1;
    clock->iteration = it;
    clock->lr[lr_index] = expon_lr((double)it, s);
}

// ---- error order ---------------------------------------------------------------

// Ascending key = descending |cells - gt| over the valid samples (31 magnitude
// bits; NaN above +inf, as torch.topk ranks it); invalid samples (gt == -1) key
// 2^31, after every valid one.  Also counts the valid samples and the Gaussians
// with weight sigmoid(logit) <= low_weight: per block in LDS, one atomic per
// block and counter (integer sums: the counts do not depend on order).
constexpr int kKeyThreads = 256;
constexpr unsigned kKeyBlocks = 1024;

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

__global__ __launch_bounds__(kKeyThreads) void error_key_kernel(int S, const float* __restrict__ cells,
                                                                const float* __restrict__ gt,
                                                                uint32_t* __restrict__ keys, int count,
                                                                const float* __restrict__ weight_logits,
                                                                float low_weight, uint32_t* __restrict__ counts)
{
    __shared__ uint32_t part[2][kKeyThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t nvalid = 0, nlow = 0;
    const int n = S > count ? S : count;
    for (int i0 = blockIdx.x * kKeyThreads; i0 < n; i0 += gridDim.x * kKeyThreads) {
        const int i = i0 + threadIdx.x;
        bool valid = false, low = false;
        if (i < S) {
            const float g = gt[i];
            valid = g != -1.0f;
            const uint32_t mag = __float_as_uint(fabsf(cells[i] - g)) & 0x7fffffffu;
            keys[i] = valid ? 0x7fffffffu - mag : 0x80000000u;
        }
        if (weight_logits && i < count) low = sigmoidf(weight_logits[i]) <= low_weight;
        nvalid += (uint32_t)__popcll(__ballot(valid));
        nlow += (uint32_t)__popcll(__ballot(low));
    }
    if (lane == 0) {
        part[0][w] = nvalid;
        part[1][w] = nlow;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t t = 0;
        for (int k = 0; k < kKeyThreads / 64; ++k) t += part[threadIdx.x][k];
        if (t) atomicAdd(&counts[threadIdx.x], t);
    }
}

// ---- densify ---------------------------------------------------------------------

__global__ void gather_positions_kernel(int count, int k, const float* __restrict__ samples,
                                        const uint32_t* __restrict__ idx, float* __restrict__ xyz)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const size_t src = idx ? (size_t)idx[j] : (size_t)j;
    const size_t dst = (size_t)count + j;
    xyz[3 * dst + 0] = samples[3 * src + 0];
    xyz[3 * dst + 1] = samples[3 * src + 1];
    xyz[3 * dst + 2] = samples[3 * src + 2];
}

struct StoreArgs {
    float* prm[5];
    float* m[5];
    float* v[5];
};

// floats per row of the five parameters: xyz, weight, scaling, rotation, value
__host__ __device__ constexpr int width_of(int a) { return a == 0 || a == 2 ? 3 : (a == 3 ? 4 : 1); }

__global__ void init_rows_kernel(int count, int k, StoreArgs s, const float* __restrict__ gt,
                                 const uint32_t* __restrict__ idx, const float* __restrict__ d2, float dist_scale,
                                 float init_weight)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const size_t r = (size_t)count + j;
    const size_t src = idx ? (size_t)idx[j] : (size_t)j;
    const float ls = logf(sqrtf(fmaxf(d2[r] * dist_scale, 1e-7f)));
    s.prm[2][3 * r + 0] = ls;
    s.prm[2][3 * r + 1] = ls;
    s.prm[2][3 * r + 2] = ls;
    s.prm[3][4 * r + 0] = 1.0f;
    s.prm[3][4 * r + 1] = 0.0f;
    s.prm[3][4 * r + 2] = 0.0f;
    s.prm[3][4 * r + 3] = 0.0f;
    s.prm[1][r] = logf(init_weight / (1.0f - init_weight));
    const float g = gt[src];
    s.prm[4][r] = logf(g / (1.0f - g));
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        for (int c = 0; c < width_of(a); ++c) {
            if (s.m[a]) s.m[a][(size_t)width_of(a) * r + c] = 0.0f;
            if (s.v[a]) s.v[a][(size_t)width_of(a) * r + c] = 0.0f;
        }
    }
}

// ---- prune ------------------------------------------------------------------------

__global__ void keep_kernel(int count, const float* __restrict__ weight_logits, const uint8_t* __restrict__ mask,
                            float min_weight, uint32_t* __restrict__ keep)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    // prune iff weight < min_weight: a NaN weight is kept, as the reference's mask keeps it
    keep[i] = mask ? (mask[i] ? 0u : 1u) : ((sigmoidf(weight_logits[i]) < min_weight) ? 0u : 1u);
}

// The 15 arrays (5 parameters, 5 first moments, 5 second moments) in one grid:
// blockIdx.y picks the array.
struct CompactArgs {
    float* src[15];
    float* stage[15];
    int width[15];
};

__global__ void compact_kernel(int count, CompactArgs c, const uint32_t* __restrict__ keep,
                               const uint32_t* __restrict__ dest)
{
    const int a = blockIdx.y;
    const float* __restrict__ src = c.src[a];
    if (!src) return;
    float* __restrict__ stage = c.stage[a];
    const int w = c.width[a];
    const size_t n = (size_t)count * w;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const size_t row = e / (size_t)w;
        if (keep[row]) stage[(size_t)dest[row] * w + (e - row * w)] = src[e];
    }
}

__global__ void copy_back_kernel(int count, CompactArgs c, const uint32_t* __restrict__ dest,
                                 int* __restrict__ new_count)
{
    const int a = blockIdx.y;
    const uint32_t kept = dest[count];
    if (a == 0 && blockIdx.x == 0 && threadIdx.x == 0 && new_count) *new_count = (int)kept;
    float* __restrict__ dst = c.src[a];
    if (!dst) return;
    const float* __restrict__ stage = c.stage[a];
    const size_t n = (size_t)kept * c.width[a];
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = stage[e];
}

StoreArgs store_args(const gvc_store* s)
{
    StoreArgs a;
    float* prm[5] = { s->xyz, s->weight, s->scaling, s->rotation, s->value };
    for (int k = 0; k < 5; ++k) {
        a.prm[k] = prm[k];
        a.m[k] = s->exp_avg[k];
        a.v[k] = s->exp_avg_sq[k];
    }
    return a;
}

struct ErrorOrderLayout {
    uint32_t *keys, *keys_alt, *vals_alt, *scratch;
    ErrorOrderLayout(Arena& a, int S)
    {
        keys = a.take<uint32_t>((size_t)S);
        keys_alt = a.take<uint32_t>((size_t)S);
        vals_alt = a.take<uint32_t>((size_t)S);
        scratch = a.take<uint32_t>(radix_scratch_elems((size_t)S));
    }
};

struct PruneLayout {
    uint32_t *keep, *dest, *scan;
    float* stage[15];
    PruneLayout(Arena& a, int count)
    {
        keep = a.take<uint32_t>((size_t)count);
        dest = a.take<uint32_t>((size_t)count + 1);
        scan = a.take<uint32_t>(scan_scratch_elems((size_t)count));
        for (int k = 0; k < 15; ++k) stage[k] = a.take<float>((size_t)count * width_of(k % 5));
    }
};

}  // namespace

int adam_step(int n, const gvc_adam_tensor* tensors, gvc_train_clock* clock, double beta1, double beta2,
              double eps, hipStream_t st)
{
    AdamArgs a = {};
    uint32_t blocks = 0;
    int used = 0;
    for (int k = 0; k < n; ++k) {
        const gvc_adam_tensor& t = tensors[k];
        if (!t.grad || t.numel == 0) continue;   // a tensor without a gradient is skipped, as torch skips it
        if (!t.param || !t.exp_avg || !t.exp_avg_sq)
            return set_error(GVC_ERR_INVALID_ARGUMENT, "adam: NULL parameter or moment");
        if (t.lr_index < 0 || t.lr_index >= GVC_ADAM_MAX_TENSORS)
            return set_error(GVC_ERR_INVALID_ARGUMENT, "adam: lr_index out of range");
        const uint64_t nb = (t.numel + kAdamBlockElems - 1) / kAdamBlockElems;
        if (blocks + nb > 0x7fffffffull) return set_error(GVC_ERR_INVALID_ARGUMENT, "adam: too many elements");
        a.p[used] = t.param;
        a.g[used] = t.grad;
        a.m[used] = t.exp_avg;
        a.v[used] = t.exp_avg_sq;
        a.numel[used] = t.numel;
        a.lr_index[used] = t.lr_index;
        a.first_block[used] = blocks;
        blocks += (uint32_t)nb;
        ++used;
    }
    a.n = used;
    a.first_block[used] = blocks;
    // torch counts steps per parameter with a gradient: no gradient, no step
    if (used == 0) return GVC_OK;
    adam_tick_kernel<<<1, 64, 0, st>>>(clock, beta1, beta2);
    GVC_LAUNCH_CHECK("adam_tick");
    const float w1 = (float)(1.0 - beta1);
    const float c2 = (float)(1.0 - beta2);
    adam_kernel<<<blocks, kAdamThreads, 0, st>>>(a, clock, w1, (float)beta2, c2, (float)eps);
    GVC_LAUNCH_CHECK("adam");
    return GVC_OK;
}

int lr_advance(gvc_train_clock* clock, const gvc_lr_schedule* s, int lr_index, int64_t set_iteration,
               hipStream_t st)
{
    lr_advance_kernel<<<1, 64, 0, st>>>(clock, *s, lr_index, set_iteration);
    GVC_LAUNCH_CHECK("lr_advance");
    return GVC_OK;
}

size_t error_order_bytes(int S)
{
    Arena a(nullptr, 0);
    ErrorOrderLayout l(a, S);
    return a.off + 256;
}

int error_order(int S, const float* cells, const float* gt, int count, const float* weight_logits,
                float low_weight, uint32_t* order, uint32_t* counts, void* ws, size_t bytes, hipStream_t st)
{
    Arena a(ws, bytes);
    ErrorOrderLayout l(a, S);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "error order workspace too small");
    GVC_TRY(fill_u32(counts, 0u, 2, st));
    const int n = S > count ? S : count;
    if (n == 0) return GVC_OK;
    const unsigned nb = blocks_for((size_t)n, kKeyThreads) < kKeyBlocks ? blocks_for((size_t)n, kKeyThreads) : kKeyBlocks;
    error_key_kernel<<<nb, kKeyThreads, 0, st>>>(S, cells, gt, l.keys, count, weight_logits, low_weight, counts);
    GVC_LAUNCH_CHECK("error_key");
    if (S == 0) return GVC_OK;
    bool alt = false;
    GVC_TRY(radix_sort_u32(l.keys, order, l.keys_alt, l.vals_alt, (size_t)S, 32, l.scratch, &alt, st, nullptr,
                           true));
    if (alt) GVC_HIP(hipMemcpyAsync(order, l.vals_alt, (size_t)S * 4, hipMemcpyDeviceToDevice, st));
    return GVC_OK;
}

size_t knn3_bytes(int n);
int knn3(int n, const float* pts, float* out, void* ws, size_t bytes, hipStream_t st);

size_t densify_bytes(int total)
{
    Arena a(nullptr, 0);
    a.take<float>((size_t)total);
    a.take<char>(knn3_bytes(total));
    return a.off + 256;
}

int densify_in_empty(int count, int k, const gvc_store* s, const float* samples, const float* gt,
                     const uint32_t* idx, float dist_scale, float init_weight, void* ws, size_t bytes,
                     hipStream_t st)
{
    const int total = count + k;
    Arena a(ws, bytes);
    float* d2 = a.take<float>((size_t)total);
    const size_t kb = knn3_bytes(total);
    char* kws = a.take<char>(kb);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "densify workspace too small");
    if (k == 0) return GVC_OK;
    gather_positions_kernel<<<blocks_for(k, 256), 256, 0, st>>>(count, k, samples, idx, s->xyz);
    GVC_LAUNCH_CHECK("gather_positions");
    GVC_TRY(knn3(total, s->xyz, d2, kws, kb, st));
    init_rows_kernel<<<blocks_for(k, 256), 256, 0, st>>>(count, k, store_args(s), gt, idx, d2, dist_scale,
                                                         init_weight);
    GVC_LAUNCH_CHECK("init_rows");
    return GVC_OK;
}

size_t prune_bytes(int count)
{
    Arena a(nullptr, 0);
    PruneLayout l(a, count);
    return a.off + 256;
}

int prune(int count, const gvc_store* s, const uint8_t* mask, float min_weight, int* new_count, void* ws,
          size_t bytes, hipStream_t st)
{
    Arena a(ws, bytes);
    PruneLayout l(a, count);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "prune workspace too small");
    if (count == 0) return new_count ? fill_u32(reinterpret_cast<uint32_t*>(new_count), 0u, 1, st) : GVC_OK;
    keep_kernel<<<blocks_for(count, 256), 256, 0, st>>>(count, s->weight, mask, min_weight, l.keep);
    GVC_LAUNCH_CHECK("keep");
    GVC_TRY(exclusive_scan_u32(l.keep, l.dest, (size_t)count, true, l.scan, st));
    const StoreArgs sa = store_args(s);
    CompactArgs c;
    for (int k = 0; k < 15; ++k) {
        const int grp = k % 5;
        c.src[k] = k < 5 ? sa.prm[grp] : (k < 10 ? sa.m[grp] : sa.v[grp]);
        c.stage[k] = l.stage[k];
        c.width[k] = width_of(grp);
    }
    const unsigned gx = blocks_for((size_t)count * 4, 256) < 2048u ? blocks_for((size_t)count * 4, 256) : 2048u;
    compact_kernel<<<dim3(gx, 15), 256, 0, st>>>(count, c, l.keep, l.dest);
    GVC_LAUNCH_CHECK("compact");
    copy_back_kernel<<<dim3(gx, 15), 256, 0, st>>>(count, c, l.dest, new_count);
    GVC_LAUNCH_CHECK("copy_back");
    return GVC_OK;
}

}  // namespace gvc
