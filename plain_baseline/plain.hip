// plain.hip -- the "plain GPU sequence": a naive HIP restatement of the
// reference's default (sample-BVH) evaluation, used only as a benchmark
// baseline.  Reference: DGR/ext.cu:13-65 (sample boxes + BVH build, leaves of up
// to 257 points), CR/forward.cu:116-182 (one warp per Gaussian walks the sample
// BVH with the Gaussian's box, points of a leaf spread over the lanes, float
// atomics into the cells), CR/forward.cu:261-273 (separate normalise),
// CR/backward.cu:127-337 (one warp per Gaussian walks again and sums its own
// gradient).  Here a wavefront is 64 lanes; the sample BVH is a Morton-sorted
// implicit binary tree over 256-point leaves.  No compensation, no batching:
// each volume is its own launch sequence.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gvc_device.h"

using namespace gvc;

namespace {

constexpr int kLeafPoints = 256;
constexpr int kMaxDepth = 30;
constexpr float kLog2e = 1.4426950408889634f;

// exp(x) on the hardware base-2 exponential, as the fused kernels evaluate it
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * kLog2e); }

__global__ void plain_preprocess_kernel(int P, const float* __restrict__ means, const float* __restrict__ scales,
                                        const float* __restrict__ rots, const float* __restrict__ weights, float k,
                                        float* __restrict__ conics, float* __restrict__ boxes)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float mu[3] = { means[3 * i], means[3 * i + 1], means[3 * i + 2] };
    float sc[3] = { scales[3 * i], scales[3 * i + 1], scales[3 * i + 2] };
    float q[4] = { rots[4 * i], rots[4 * i + 1], rots[4 * i + 2], rots[4 * i + 3] };
    float cn[6], bx[6];
    preprocess_one(mu, sc, q, weights[i], k, false, cn, bx);
    for (int c = 0; c < 6; ++c) {
        conics[6 * i + c] = cn[c];
        boxes[6 * i + c] = bx[c];
    }
}

__global__ void plain_codes_kernel(int S, const float* __restrict__ samples, uint64_t* __restrict__ codes,
                                   uint32_t* __restrict__ idx)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    codes[s] = morton3d_unit(samples[3 * s], samples[3 * s + 1], samples[3 * s + 2], true);
    idx[s] = (uint32_t)s;
}

// points in Morton order, original index in w
__global__ void plain_gather_kernel(int S, const float* __restrict__ samples, const uint32_t* __restrict__ order,
                                    float4* __restrict__ pts)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint32_t j = order[s];
    pts[s] = make_float4(samples[3 * j], samples[3 * j + 1], samples[3 * j + 2], __uint_as_float(j));
}

// node n of the implicit tree: boxes lo[n], hi[n]; leaves are nodes NP-1 .. 2NP-2
__global__ void plain_leaf_box_kernel(int S, int NP, const float4* __restrict__ pts, float4* __restrict__ lo,
                                      float4* __restrict__ hi)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= NP) return;
    float4 l = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.f);
    float4 h = make_float4(-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), 0.f);
    const int b = j * kLeafPoints;
    const int e = b + kLeafPoints < S ? b + kLeafPoints : S;
    for (int s = b; s < e; ++s) {
        const float4 p = pts[s];
        l.x = fminf(l.x, p.x); l.y = fminf(l.y, p.y); l.z = fminf(l.z, p.z);
        h.x = fmaxf(h.x, p.x); h.y = fmaxf(h.y, p.y); h.z = fmaxf(h.z, p.z);
    }
    lo[NP - 1 + j] = l;
    hi[NP - 1 + j] = h;
}

__global__ void plain_level_kernel(int first, int count, float4* __restrict__ lo, float4* __restrict__ hi)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int n = first + t;
    const float4 al = lo[2 * n + 1], ah = hi[2 * n + 1], bl = lo[2 * n + 2], bh = hi[2 * n + 2];
    lo[n] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.f);
    hi[n] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.f);
}

__global__ void plain_zero_kernel(float* __restrict__ a, float* __restrict__ b, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] = 0.f; b[i] = 0.f; }
}

__device__ __forceinline__ bool plain_accept(const float* bx, const float* mu, const float* cn, float4 p, float* pw)
{
    if (!(bx[0] <= p.x && p.x <= bx[3] && bx[1] <= p.y && p.y <= bx[4] && bx[2] <= p.z && p.z <= bx[5]))
        return false;
    float r0, r1, r2;
    const float power =
        pair_power(cn[0], cn[1], cn[2], cn[3], cn[4], cn[5], p.x - mu[0], p.y - mu[1], p.z - mu[2], r0, r1, r2);
    if (!in_window(power)) return false;
    *pw = power;
    return true;
}

// Depth-first walk of the sample tree with Gaussian box bx; calls f(point) for
// every point of every leaf whose box meets bx, the leaf's points spread over
// the wave's lanes.  The walk is wave-uniform (one Gaussian per wave).
template <class F>
__device__ __forceinline__ void plain_walk(const float* bx, int NP, int S, const float4* __restrict__ lo,
                                           const float4* __restrict__ hi, const float4* __restrict__ pts, int lane,
                                           F&& f)
{
    int stack[kMaxDepth + 2];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int n = stack[--sp];
        const float4 l = lo[n], h = hi[n];
        if (!(l.x <= bx[3] && bx[0] <= h.x && l.y <= bx[4] && bx[1] <= h.y && l.z <= bx[5] && bx[2] <= h.z))
            continue;
        if (n >= NP - 1) {
            const int b = (n - (NP - 1)) * kLeafPoints;
            const int e = b + kLeafPoints < S ? b + kLeafPoints : S;
            for (int s = b + lane; s < e; s += 64) f(pts[s]);
        } else {
            stack[sp++] = 2 * n + 2;
            stack[sp++] = 2 * n + 1;
        }
    }
}

__global__ __launch_bounds__(256) void plain_forward_kernel(int P, const float* __restrict__ means,
                                                            const float* __restrict__ values,
                                                            const float* __restrict__ weights,
                                                            const float* __restrict__ conics,
                                                            const float* __restrict__ boxes, int NP, int S,
                                                            const float4* __restrict__ lo,
                                                            const float4* __restrict__ hi,
                                                            const float4* __restrict__ pts, float* __restrict__ U,
                                                            float* __restrict__ W,
                                                            unsigned long long* __restrict__ pairs)
{
    const int i = blockIdx.x * (blockDim.x / 64) + (int)(threadIdx.x / 64);
    const int lane = (int)(threadIdx.x & 63);
    if (i >= P) return;
    const float mu[3] = { means[3 * i], means[3 * i + 1], means[3 * i + 2] };
    float cn[6], bx[6];
    for (int c = 0; c < 6; ++c) { cn[c] = conics[6 * i + c]; bx[c] = boxes[6 * i + c]; }
    const float w = weights[i], v = values[i];
    unsigned count = 0;
    plain_walk(bx, NP, S, lo, hi, pts, lane, [&](float4 p) {
        float pw;
        if (!plain_accept(bx, mu, cn, p, &pw)) return;
        const float wgt = w * fast_exp(pw);
        const uint32_t s = __float_as_uint(p.w);
        atomicAdd(&W[s], wgt);
        atomicAdd(&U[s], wgt * v);
        ++count;
    });
    for (int d = 32; d >= 1; d >>= 1) count += __shfl_xor(count, d);
    if (lane == 0 && count) atomicAdd(pairs, (unsigned long long)count);
}

__global__ void plain_normalize_kernel(int S, float* __restrict__ U, float* __restrict__ W)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    if (W[s] <= (float)kWeightCutoff) {
        U[s] = -1.0f;
        W[s] = 0.0f;
    } else {
        U[s] = U[s] / W[s];
    }
}

__global__ __launch_bounds__(256) void plain_backward_kernel(
    int P, const float* __restrict__ means, const float* __restrict__ scales, const float* __restrict__ rots,
    const float* __restrict__ values, const float* __restrict__ weights, float kmod,
    const float* __restrict__ conics, const float* __restrict__ boxes, int NP, int S, const float4* __restrict__ lo,
    const float4* __restrict__ hi, const float4* __restrict__ pts, const float* __restrict__ V,
    const float* __restrict__ W, const float* __restrict__ gV, const float* __restrict__ gW,
    float* __restrict__ dmeans, float* __restrict__ dscales, float* __restrict__ drots, float* __restrict__ dvalues,
    float* __restrict__ dweights)
{
    const int i = blockIdx.x * (blockDim.x / 64) + (int)(threadIdx.x / 64);
    const int lane = (int)(threadIdx.x & 63);
    if (i >= P) return;
    const float mu[3] = { means[3 * i], means[3 * i + 1], means[3 * i + 2] };
    float cn[6], bx[6];
    for (int c = 0; c < 6; ++c) { cn[c] = conics[6 * i + c]; bx[c] = boxes[6 * i + c]; }
    const float w = weights[i], v = values[i];
    float g[11];
    for (int c = 0; c < 11; ++c) g[c] = 0.f;
    plain_walk(bx, NP, S, lo, hi, pts, lane, [&](float4 p) {
        float pw;
        if (!plain_accept(bx, mu, cn, p, &pw)) return;
        const uint32_t s = __float_as_uint(p.w);
        const float Ws = W[s];
        if (Ws <= (float)kWeightCutoff) return;
        const float dx = p.x - mu[0], dy = p.y - mu[1], dz = p.z - mu[2];
        const float e = fast_exp(pw), wgt = w * e, invW = 1.0f / Ws;
        const float t = gV[s] * (v - V[s]) * invW + gW[s];
        const float dq = -0.5f * wgt * t;
        g[0] += gV[s] * wgt * invW;
        g[1] += t * e;
        g[2] += -2.f * dq * (cn[0] * dx + cn[1] * dy + cn[2] * dz);
        g[3] += -2.f * dq * (cn[1] * dx + cn[3] * dy + cn[4] * dz);
        g[4] += -2.f * dq * (cn[2] * dx + cn[4] * dy + cn[5] * dz);
        g[5] += dq * dx * dx;
        g[6] += dq * 2.f * dx * dy;
        g[7] += dq * 2.f * dx * dz;
        g[8] += dq * dy * dy;
        g[9] += dq * 2.f * dy * dz;
        g[10] += dq * dz * dz;
    });
    for (int c = 0; c < 11; ++c)
        for (int d = 32; d >= 1; d >>= 1) g[c] += __shfl_xor(g[c], d);
    if (lane != 0) return;
    dvalues[i] = g[0];
    dweights[i] = g[1];
    dmeans[3 * i + 0] = g[2];
    dmeans[3 * i + 1] = g[3];
    dmeans[3 * i + 2] = g[4];
    const double g6[6] = { g[5], g[6], g[7], g[8], g[9], g[10] };
    double ds[3], dq4[4];
    covariance_chain(g6, scales + 3 * i, rots + 4 * i, kmod, ds, dq4);
    for (int c = 0; c < 3; ++c) dscales[3 * i + c] = (float)ds[c];
    for (int c = 0; c < 4; ++c) drots[4 * i + c] = (float)dq4[c];
}

inline unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

int depth_of(int S)
{
    const int leaves = S > 0 ? (S + kLeafPoints - 1) / kLeafPoints : 1;
    int d = 0;
    while ((1 << d) < leaves) ++d;
    return d;
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points (ctypes).  Sizes: conics/boxes 6P floats; codes S u64; idx S
// u32; pts S float4; lo/hi 2 NP - 1 float4 each with NP = plain_leaf_slots(S).
// The Morton sort between plain_codes and plain_gather is the caller's.
// ---------------------------------------------------------------------------
extern "C" {

int plain_leaf_slots(int S) { return 1 << depth_of(S); }

int plain_preprocess(int P, const float* means, const float* scales, const float* rots, const float* weights,
                     float k, float* conics, float* boxes, void* stream)
{
    if (P <= 0) return 0;
    plain_preprocess_kernel<<<blocks(P, 256), 256, 0, (hipStream_t)stream>>>(P, means, scales, rots, weights, k,
                                                                             conics, boxes);
    return (int)hipGetLastError();
}

int plain_codes(int S, const float* samples, uint64_t* codes, uint32_t* idx, void* stream)
{
    if (S <= 0) return 0;
    plain_codes_kernel<<<blocks(S, 256), 256, 0, (hipStream_t)stream>>>(S, samples, codes, idx);
    return (int)hipGetLastError();
}

int plain_build(int S, const float* samples, const uint32_t* order, float4* pts, float4* lo, float4* hi,
                void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int D = depth_of(S), NP = 1 << D;
    if (S > 0) plain_gather_kernel<<<blocks(S, 256), 256, 0, st>>>(S, samples, order, pts);
    plain_leaf_box_kernel<<<blocks(NP, 64), 64, 0, st>>>(S, NP, pts, lo, hi);
    for (int lev = D - 1; lev >= 0; --lev) {
        const int first = (1 << lev) - 1, count = 1 << lev;
        plain_level_kernel<<<blocks(count, 256), 256, 0, st>>>(first, count, lo, hi);
    }
    return (int)hipGetLastError();
}

int plain_forward(int P, const float* means, const float* values, const float* weights, const float* conics,
                  const float* boxes, int S, const float4* pts, const float4* lo, const float4* hi, float* out_cells,
                  float* out_weights, unsigned long long* pairs, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int NP = plain_leaf_slots(S);
    if (S > 0) plain_zero_kernel<<<blocks(S, 256), 256, 0, st>>>(out_cells, out_weights, S);
    if (P > 0 && S > 0)
        plain_forward_kernel<<<blocks(P, 4), 256, 0, st>>>(P, means, values, weights, conics, boxes, NP, S, lo, hi,
                                                           pts, out_cells, out_weights, pairs);
    if (S > 0) plain_normalize_kernel<<<blocks(S, 256), 256, 0, st>>>(S, out_cells, out_weights);
    return (int)hipGetLastError();
}

int plain_backward(int P, const float* means, const float* scales, const float* rots, const float* values,
                   const float* weights, float k, const float* conics, const float* boxes, int S, const float4* pts,
                   const float4* lo, const float4* hi, const float* out_cells, const float* out_weights,
                   const float* gcells, const float* gweights, float* dmeans, float* dscales, float* drots,
                   float* dvalues, float* dweights, void* stream)
{
    if (P <= 0) return 0;
    plain_backward_kernel<<<blocks(P, 4), 256, 0, (hipStream_t)stream>>>(
        P, means, scales, rots, values, weights, k, conics, boxes, plain_leaf_slots(S), S > 0 ? S : 0, lo, hi, pts,
        out_cells, out_weights, gcells, gweights, dmeans, dscales, drots, dvalues, dweights);
    return (int)hipGetLastError();
}

}  // extern "C"
