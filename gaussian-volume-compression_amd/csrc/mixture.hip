// mixture.hip -- the differentiable Gaussian-mixture evaluator on gfx950.
//
// Pipeline (one stream, no allocation, no host synchronisation):
//   sample set   : Morton keys of the samples (top 24 of 30 bits) -> stable radix sort -> float4
//                  points (x, y, z, original index) in Morton order; every 64
//                  consecutive points form one group = one wavefront.
//   forward      : top 24 bits of the means' 63-bit Morton codes (the reference's
//                  morton3d_unit) -> stable radix sort -> canonical preprocess in
//                  the caller's order written as 80-byte records at their BVH
//                  positions -> Karras radix tree over leaves of
//                  `leaf_size` consecutive records -> node boxes from a segment
//                  tree over the leaf boxes -> persistent wavefronts take groups
//                  from per-XCD queues; a group is cut into runs of consecutive
//                  lanes where it straddles a coarse Morton boundary, and each run
//                  walks the tree with its box (wide wave-uniform traversal,
//                  candidates in 64-wide chunks); each lane owns one point and
//                  accumulates w exp(power) and w exp(power) v with Kahan
//                  compensation.  The walk also counts, per group, the Gaussians
//                  with at least one accepted pair: those are the backward's
//                  records.
//   backward     : the same walk per group; each lane takes one candidate
//                  Gaussian of a chunk and loops over the run's points inside its
//                  box (LDS) accumulating the Gaussian's 11 partial sums -> one
//                  record per (Gaussian, group) at a position fixed by the
//                  forward's counts -> stable radix sort by Gaussian ->
//                  per-Gaussian FP64 sums by lane teams in BVH order (order
//                  fixed by the segment alone) -> covariance chain in FP64 ->
//                  gradients in the caller's Gaussian order.
// Reference path: CR/forward.cu:10-274, CR/backward.cu:9-660,
// CR/rasterizer_impl.cu:27-452, DGR/ext.cu:13-147 (CR/ = diff-gaussian-
// rasterization/cuda_rasterizer/, DGR/ = diff-gaussian-rasterization/).
#include <math.h>

#include <vector>
#include <type_traits>

#include "gvc_device.h"
#include "gvc_internal.h"

namespace gvc {

namespace {

constexpr float kLog2e = 1.4426950408889634f;

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float uniformf(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ int uniformi(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave-wide sum, every lane gets the result (all lanes active).
__device__ __forceinline__ uint32_t wave_sumu(uint32_t v)
{
    v += lane_xor<1>(v);
    v += lane_xor<2>(v);
    v += lane_xor<4>(v);
    v += lane_xor<8>(v);
    v += lane_xor<16>(v);
    return v + lane_xor<32>(v);
}

__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// ---------------------------------------------------------------------------
// batches: V independent volumes (Gaussians, samples, groups, leaves and
// internal nodes each a contiguous range).  The device table holds V, then
// five arrays of V+1 prefix offsets.  A single volume is the V = 1 table.
// ---------------------------------------------------------------------------
struct BatchView {
    int V;
    const int* P_off;  // Gaussians
    const int* S_off;  // samples (input order is volume-major)
    const int* G_off;  // 64-point groups
    const int* L_off;  // BVH leaves
    const int* N_off;  // internal BVH nodes
};

inline size_t batch_table_ints(int V) { return 4 + 5 * ((size_t)V + 1); }

inline BatchView batch_view(const void* table, int V)
{
    const int* t = reinterpret_cast<const int*>(table) + 4;
    const size_t n = (size_t)V + 1;
    return BatchView{ V, t, t + n, t + 2 * n, t + 3 * n, t + 4 * n };
}

// largest v with off[v] <= x (off ascending, off[0] <= x)
__device__ __forceinline__ int volume_of(const int* __restrict__ off, int V, int x)
{
    int lo = 0, hi = V - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// one volume of P Gaussians and S samples with leaves of L Gaussians
// The one-volume table of P Gaussians, S samples, leaf size L (one thread).
__device__ __forceinline__ void write_single_table(int* __restrict__ t, int P, int S, int L)
{
    const int NL = P > 0 ? (P + L - 1) / L : 0;
    const int v[5][2] = { { 0, P }, { 0, S }, { 0, (S + 63) / 64 }, { 0, NL }, { 0, NL > 1 ? NL - 1 : 0 } };
    t[0] = 1; t[1] = t[2] = t[3] = 0;
    for (int k = 0; k < 5; ++k) {
        t[4 + 2 * k] = v[k][0];
        t[4 + 2 * k + 1] = v[k][1];
    }
}

__global__ void single_table_kernel(int* __restrict__ t, int P, int S, int L)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) write_single_table(t, P, S, L);
}

// ---------------------------------------------------------------------------
// generic kernels
// ---------------------------------------------------------------------------
__global__ void iota_kernel(uint32_t* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

__global__ void morton63_kernel(const float* __restrict__ xyz, int n, int clamp, uint64_t* __restrict__ codes)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) codes[i] = morton3d_unit(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], clamp != 0);
}

// factors (optional, [P, 12]): the precise path's conic factor hi[6] lo[6], NaN
// rows for Gaussians off that path.
__global__ void preprocess_kernel(int P, const float* __restrict__ means, const float* __restrict__ scales,
                                  const float* __restrict__ rots, const float* __restrict__ weights, float k,
                                  float fp64_ratio, float factor_ratio, float* __restrict__ conics,
                                  float* __restrict__ boxes, float* __restrict__ factors)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float mu[3] = { means[3 * i], means[3 * i + 1], means[3 * i + 2] };
    float sc[3] = { scales[3 * i], scales[3 * i + 1], scales[3 * i + 2] };
    float q[4] = { rots[4 * i], rots[4 * i + 1], rots[4 * i + 2], rots[4 * i + 3] };
    float cn[6], bx[6], fh[6], fl[6];
    const bool precise = preprocess_one(mu, sc, q, weights[i], k, fp64_ratio, cn, bx, fh, fl, factor_ratio,
                                        factors != nullptr);
    if (conics)
        for (int c = 0; c < 6; ++c) conics[6 * i + c] = cn[c];
    if (boxes)
        for (int c = 0; c < 6; ++c) boxes[6 * i + c] = bx[c];
    if (factors)
        for (int c = 0; c < 6; ++c) {
            factors[12 * (size_t)i + c] = precise ? fh[c] : __builtin_nanf("");
            factors[12 * (size_t)i + 6 + c] = precise ? fl[c] : __builtin_nanf("");
        }
}

// ---------------------------------------------------------------------------
// sample set
// ---------------------------------------------------------------------------
struct SampleSetHeader {
    int S;
    int NG;
    int magic;
    int pad;
};
constexpr int kSampleMagic = 0x47564353;

__device__ __forceinline__ void write_sample_header(SampleSetHeader* h, int S, int NG)
{
    h->S = S;
    h->NG = NG;
    h->magic = kSampleMagic;
    h->pad = 0;
}

// 30-bit Morton key of a sample: coordinates clamped to the unit cube (NaN -> 0),
// 10 bits per axis.
// Bits of sample_code dropped from the sample set's sort key: the points are
// ordered on the top 24 bits, which leaves a 256^3 lattice of cells -- finer than
// the groups of 64 points of any set this build sorts -- in three 8-bit radix
// passes instead of four.
constexpr int kSampleKeyDrop = 6;
constexpr int kSampleKeyBits = 30 - kSampleKeyDrop;

__device__ __forceinline__ uint32_t sample_code(float x, float y, float z)
{
    const float v[3] = { x, y, z };
    uint32_t c = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float t = v[a];
        t = (t == t) ? fminf(fmaxf(t, 0.0f), 1.0f) : 0.0f;
        c |= spread10((uint32_t)(t * 1023.0f)) << a;
    }
    return c;
}

// With table != nullptr (a single-volume set), thread 0 also writes the set's
// one-volume table.
__global__ void sample_codes_kernel(const float* __restrict__ s, int S, uint32_t* __restrict__ keys,
                                    uint32_t* __restrict__ idx, int* __restrict__ table)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (table && i == 0) write_single_table(table, 0, S, 1);
    if (i >= S) return;
    // the sort key: the top 24 bits (8 per axis, 256^3 cells; ties keep the
    // caller's order)
    keys[i] = sample_code(s[3 * i], s[3 * i + 1], s[3 * i + 2]) >> kSampleKeyDrop;
    idx[i] = (uint32_t)i;
}

// Sorted points (volume-major, Morton order within a volume), each volume
// padded with NaN points (original index -1) to whole 64-point groups, so every
// group can be read with wave-uniform loads and belongs to one volume.  Thread 0
// also writes the set's header.
__global__ void sample_gather_kernel(const float* __restrict__ s, BatchView b, int S64,
                                     const uint32_t* __restrict__ perm, float4* __restrict__ pts,
                                     SampleSetHeader* __restrict__ header, int S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) write_sample_header(header, S, S64 / 64);
    if (i >= S64) return;
    const int v = b.V == 1 ? 0 : volume_of(b.G_off, b.V, i >> 6);
    const int r = i - 64 * b.G_off[v];
    if (r >= b.S_off[v + 1] - b.S_off[v]) {
        const float nan = __builtin_nanf("");
        pts[i] = make_float4(nan, nan, nan, __builtin_bit_cast(float, -1));
        return;
    }
    const uint32_t j = perm[b.S_off[v] + r];
    pts[i] = make_float4(s[3 * j], s[3 * j + 1], s[3 * j + 2], __builtin_bit_cast(float, (int)j));
}

// The header of a set without samples (sample_gather_kernel writes it otherwise).
__global__ void sample_header_kernel(SampleSetHeader* h, int S, int NG)
{
    write_sample_header(h, S, NG);
}

// ---------------------------------------------------------------------------
// Gaussian BVH
// ---------------------------------------------------------------------------
// The tree orders Gaussians by the top 24 bits of their 63-bit Morton codes
// (8 bits per axis, a 256^3 lattice, finer than the Gaussians of a volume): three
// 8-bit radix passes on 32-bit words instead of eight on 64-bit codes; equal keys
// keep the caller's order (stable sort) and the Karras build breaks ties by
// index.
constexpr int kGaussianCodeBits = 24;

// Thread 0 also clears the state's four error words and, for a state without a
// caller's table, writes its one-volume table (S samples, leaf size L).
__global__ void gaussian_codes_kernel(const float* __restrict__ means, int P, uint32_t* __restrict__ codes,
                                      uint32_t* __restrict__ err, int* __restrict__ table, int S, int L)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        for (int k = 0; k < 4; ++k) err[k] = 0u;
        if (table) write_single_table(table, P, S, L);
    }
    if (i >= P) return;
    codes[i] = (uint32_t)(morton3d_unit(means[3 * i], means[3 * i + 1], means[3 * i + 2], false) >>
                          (63 - kGaussianCodeBits));
}

// Canonical preprocess, in the caller's order (coalesced parameter reads and
// conic writes), written as the record of BVH index inv_perm[j].  With fac_lo
// (FP64 mode), a Gaussian on the precise path (scale ratio above factor_ratio:
// the factor of its FP64 conic above fp64_ratio, the float-float factor below)
// carries the hi parts of its conic's
// Cholesky factor in the conic slots and kPreciseBit in its id; the lo parts go
// to fac_lo[2j], [2j + 1], j = the caller's index.  Only those entries are
// written: a pair loop reads an entry only for an id with kPreciseBit, which
// this same launch sets.
__global__ __launch_bounds__(256) void record_kernel(int P, const uint32_t* __restrict__ inv_perm,
                                                     const float* __restrict__ means,
                                                     const float* __restrict__ scales, const float* __restrict__ rots,
                                                     const float* __restrict__ values,
                                                     const float* __restrict__ weights, float k, float fp64_ratio,
                                                     float factor_ratio, GRecord* __restrict__ rec,
                                                     float4* __restrict__ fac_lo, float* __restrict__ conics_out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P) return;
    const uint32_t i = inv_perm[j];
    float mu[3] = { means[3 * j], means[3 * j + 1], means[3 * j + 2] };
    float sc[3] = { scales[3 * j], scales[3 * j + 1], scales[3 * j + 2] };
    float q[4] = { rots[4 * j], rots[4 * j + 1], rots[4 * j + 2], rots[4 * j + 3] };
    const float w = weights[j];
    float cn[6], bx[6], fh[6], fl[6];
    const bool precise = preprocess_one(mu, sc, q, w, k, fp64_ratio, cn, bx, fh, fl, factor_ratio,
                                        fac_lo != nullptr);
    // element-wise selects (a pointer to either array would put both in scratch)
    GRecord r;
    r.mx = mu[0]; r.my = mu[1]; r.mz = mu[2]; r.w = w;
    r.c0 = precise ? fh[0] : cn[0]; r.c1 = precise ? fh[1] : cn[1]; r.c2 = precise ? fh[2] : cn[2];
    r.c3 = precise ? fh[3] : cn[3]; r.c4 = precise ? fh[4] : cn[4]; r.c5 = precise ? fh[5] : cn[5];
    r.v = values[j];
    r.id = (int)((uint32_t)j | (precise ? kPreciseBit : 0u));
    r.lox = bx[0]; r.loy = bx[1]; r.loz = bx[2];
    r.hix = bx[3]; r.hiy = bx[4]; r.hiz = bx[5];
    r.pad0 = 0.f; r.pad1 = 0.f;
    rec[i] = r;
    if (precise) {
        fac_lo[2 * (size_t)j] = make_float4(fl[0], fl[1], fl[2], fl[3]);
        fac_lo[2 * (size_t)j + 1] = make_float4(fl[4], fl[5], 0.f, 0.f);
    }
    if (conics_out)
        for (int c = 0; c < 6; ++c) conics_out[6 * j + c] = cn[c];
}

// Per volume, the walks' estimate of a Gaussian's reach when splitting groups:
// the mean box extent per axis over up to 1024 records taken at a fixed stride,
// summed in units of 2^-20 (integers: the same total in any order).  ext[4v..]
// = (sum x, sum y, sum z, count).
__device__ __forceinline__ void volume_extent(const GRecord* __restrict__ rec, const BatchView& bv, int v,
                                              unsigned long long* __restrict__ ext)
{
    __shared__ unsigned long long red[3][256];
    const long long g0 = bv.P_off[v], n = bv.P_off[v + 1] - g0;
    const long long m = n < 1024 ? n : 1024;
    unsigned long long sum[3] = { 0, 0, 0 };
    for (long long kk = threadIdx.x; kk < m; kk += 256) {
        const GRecord& r = rec[g0 + (kk * n) / m];
        const float e[3] = { r.hix - r.lox, r.hiy - r.loy, r.hiz - r.loz };
        for (int a = 0; a < 3; ++a)
            sum[a] += (e[a] >= 0.0f) ? (unsigned long long)(fminf(e[a], 1024.0f) * 1048576.0f) : 0ull;
    }
    for (int a = 0; a < 3; ++a) red[a][threadIdx.x] = sum[a];
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d)
            for (int a = 0; a < 3; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int a = 0; a < 3; ++a) ext[4 * (size_t)v + a] = red[a][0];
        ext[4 * (size_t)v + 3] = (unsigned long long)m;
    }
}

// First Gaussian of leaf j (leaves never span volumes).
__device__ __forceinline__ int leaf_first(const BatchView& b, int L, int j, int& end)
{
    const int v = b.V == 1 ? 0 : volume_of(b.L_off, b.V, j);
    const int g0 = b.P_off[v] + (j - b.L_off[v]) * L;
    end = min(g0 + L, b.P_off[v + 1]);
    return g0;
}

// Blocks [0, V): the extent sums of volume blockIdx.x (volume_extent); the
// others: the sorted code of every leaf's first Gaussian.  256 threads.
__global__ __launch_bounds__(256) void extent_leaf_codes_kernel(const GRecord* __restrict__ rec,
                                                                const uint32_t* __restrict__ sorted_codes,
                                                                BatchView b, int NL, int L,
                                                                unsigned long long* __restrict__ ext,
                                                                uint64_t* __restrict__ leaf_codes)
{
    if ((int)blockIdx.x < b.V) {
        volume_extent(rec, b, (int)blockIdx.x, ext);
        return;
    }
    const int j = ((int)blockIdx.x - b.V) * 256 + (int)threadIdx.x;
    if (j >= NL) return;
    int end;
    leaf_codes[j] = sorted_codes[leaf_first(b, L, j, end)];
}

__device__ __forceinline__ int karras_delta(const uint64_t* k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clzll(a ^ b);
}

// Karras (2012): one thread per internal node, within its volume's leaves;
// children and leaf ranges are stored as global indices.
__global__ void karras_kernel(const uint64_t* __restrict__ all_keys, BatchView bv, int NI, BNode* __restrict__ nodes)
{
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= NI) return;
    const int v = bv.V == 1 ? 0 : volume_of(bv.N_off, bv.V, gi);
    const int lb = bv.L_off[v], nb = bv.N_off[v];
    const int n = bv.L_off[v + 1] - lb;
    const uint64_t* keys = all_keys + lb;
    const int i = gi - nb;
    const int d = (karras_delta(keys, n, i, i + 1) - karras_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = karras_delta(keys, n, i, i - d);
    int lmax = 2;
    while (karras_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (karras_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = karras_delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) / 2;
        if (karras_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int left = (lo == gamma) ? ~(lb + gamma) : nb + gamma;
    const int right = (hi == gamma + 1) ? ~(lb + gamma + 1) : nb + gamma + 1;
    nodes[gi].left = left;
    nodes[gi].right = right;
    nodes[gi].pad0 = lb + lo;
    nodes[gi].pad1 = lb + hi;
}

// Node boxes.  A node of the radix tree covers the leaf range [first, last]; its
// box is the union of the leaves' boxes, read from a segment tree over the
// leaves: M = the power of two at or above NL, seg[M + j] = box of leaf j (NaN
// boxes for j >= NL), seg[i] = seg[2i] u seg[2i + 1]; a range is the union of the
// O(log) aligned nodes of its canonical decomposition.  Unions use fminf/fmaxf,
// which skip NaN boxes (a leaf whose Gaussians all have NaN boxes keeps a NaN
// box); min and max are exact, so any grouping of the unions gives the same box.
struct Box6 {
    float lo[3], hi[3];
};

__device__ __forceinline__ Box6 box_union(const Box6& a, const Box6& b)
{
    Box6 r;
    for (int c = 0; c < 3; ++c) {
        r.lo[c] = fminf(a.lo[c], b.lo[c]);
        r.hi[c] = fmaxf(a.hi[c], b.hi[c]);
    }
    return r;
}

// The NaN box is the identity of box_union (fminf / fmaxf skip NaN): it pads the
// segment tree and starts the range unions, so a range whose leaves all have NaN
// boxes keeps a NaN box.
__device__ __forceinline__ Box6 nan_box()
{
    Box6 b;
    for (int c = 0; c < 3; ++c) b.lo[c] = b.hi[c] = __builtin_nanf("");
    return b;
}

// Segment-tree levels built per block: a block of kSegBlock threads reduces
// kSegBlock consecutive nodes of one level through kSegLevels levels in LDS.
constexpr int kSegBlock = 256;
constexpr int kSegLevels = 8;   // log2(kSegBlock)

// One block: the nodes [base, base + kSegBlock) of a level (indices into seg[],
// level nodes base .. base + count), their parents, grandparents, ... up to
// kSegLevels levels (or the root) are formed in LDS and written.  With leaves
// (rec != nullptr) the level's boxes are first formed from the leaves' records.
__global__ __launch_bounds__(kSegBlock) void seg_level_kernel(const GRecord* __restrict__ rec, BatchView bv, int NL,
                                                               int L, Box6* __restrict__ seg, uint32_t base,
                                                               uint32_t count, int levels)
{
    __shared__ Box6 sb[kSegBlock];
    const int t = threadIdx.x;
    const uint32_t i = base + blockIdx.x * kSegBlock + t;   // node index at this level
    Box6 b = nan_box();
    if (i < base + count) {
        if (rec) {
            const uint32_t j = i - base;   // leaf j (base = M)
            if ((int)j < NL) {
                bool any = false;
                int g1;
                const int g0 = leaf_first(bv, L, (int)j, g1);
                for (int g = g0; g < g1; ++g) {
                    const GRecord& r = rec[g];
                    b.lo[0] = fminf(b.lo[0], r.lox); b.lo[1] = fminf(b.lo[1], r.loy); b.lo[2] = fminf(b.lo[2], r.loz);
                    b.hi[0] = fmaxf(b.hi[0], r.hix); b.hi[1] = fmaxf(b.hi[1], r.hiy); b.hi[2] = fmaxf(b.hi[2], r.hiz);
                    any |= (r.lox == r.lox);
                }
                if (!any)
                    for (int c = 0; c < 3; ++c) b.lo[c] = b.hi[c] = __builtin_nanf("");
            }
            seg[i] = b;
        } else {
            b = seg[i];
        }
    }
    sb[t] = b;
    uint32_t lvl_base = base;
    for (int k = 1; k <= levels; ++k) {
        __syncthreads();
        const int n = kSegBlock >> k;   // nodes of this block at level k
        Box6 u;
        if (t < n) u = box_union(sb[2 * t], sb[2 * t + 1]);
        __syncthreads();
        lvl_base >>= 1;
        if (t < n) {
            sb[t] = u;
            // the level holds count >> k nodes (fewer than a block's share when
            // the whole level fits in one block)
            const uint32_t q = blockIdx.x * (uint32_t)n + (uint32_t)t;
            if (q < (count >> k)) seg[lvl_base + q] = u;
        }
    }
}

// Union of leaves [a, b] (inclusive) by the canonical decomposition.
__device__ __forceinline__ Box6 range_box(const Box6* __restrict__ seg, uint32_t M, int a, int b)
{
    uint32_t l = M + (uint32_t)a, r = M + (uint32_t)b + 1u;
    Box6 x = nan_box(), y = nan_box();   // left part, right part
    while (l < r) {
        if (l & 1u) x = box_union(x, seg[l++]);
        if (r & 1u) y = box_union(seg[--r], y);
        l >>= 1;
        r >>= 1;
    }
    return box_union(x, y);
}

__global__ void node_box_kernel(const Box6* __restrict__ seg, uint32_t M, BatchView bv, int NI,
                                BNode* __restrict__ nodes, float* __restrict__ root_box)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NI) return;
    const int v = bv.V == 1 ? 0 : volume_of(bv.N_off, bv.V, i);
    BNode& n = nodes[i];
    const int lo = n.pad0, hi = n.pad1, left = n.left;
    // the left child covers [lo, gamma]; an internal left child has local index gamma
    const int gamma = left < 0 ? ~left : bv.L_off[v] + (left - bv.N_off[v]);
    const Box6 l = range_box(seg, M, lo, gamma);
    const Box6 r = range_box(seg, M, gamma + 1, hi);
    for (int c = 0; c < 3; ++c) {
        n.llo[c] = l.lo[c];
        n.lhi[c] = l.hi[c];
        n.rlo[c] = r.lo[c];
        n.rhi[c] = r.hi[c];
    }
    if (i == bv.N_off[v]) {
        const Box6 u = box_union(l, r);
        for (int c = 0; c < 3; ++c) {
            root_box[6 * v + c] = u.lo[c];
            root_box[6 * v + 3 + c] = u.hi[c];
        }
    }
}

// Root boxes of volumes with a single leaf (no internal node); empty volumes get
// an empty box.
// Thread 0 also clears the heavy-group count (*heavy) for group_cost_kernel.
__global__ void leaf_root_kernel(const Box6* __restrict__ leaves, BatchView bv, float* __restrict__ root_box,
                                 uint32_t* __restrict__ heavy)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v == 0) *heavy = 0u;
    if (v >= bv.V) return;
    const int nl = bv.L_off[v + 1] - bv.L_off[v];
    if (nl > 1) return;
    for (int c = 0; c < 3; ++c) {
        root_box[6 * v + c] = nl == 1 ? leaves[bv.L_off[v]].lo[c] : 1.0f;
        root_box[6 * v + 3 + c] = nl == 1 ? leaves[bv.L_off[v]].hi[c] : 0.0f;
    }
}

// The tree of a forward state, copied out (gvc_forward_tree): per BVH position
// the caller's index and the 32-bit Morton key (recomputed from the record's
// mean, as gaussian_codes_kernel formed it); per internal node its links, leaf
// range and both child boxes.
__global__ void tree_export_kernel(const GRecord* __restrict__ rec, int P, const BNode* __restrict__ nodes, int NI,
                                   uint32_t* __restrict__ order, uint32_t* __restrict__ keys,
                                   int32_t* __restrict__ links, float* __restrict__ boxes)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P) {
        const GRecord& r = rec[i];
        if (order) order[i] = (uint32_t)r.id & ~kPreciseBit;
        if (keys)
            keys[i] = (uint32_t)(morton3d_unit(r.mx, r.my, r.mz, false) >> (63 - kGaussianCodeBits));
    }
    if (i < NI) {
        const BNode& n = nodes[i];
        if (links) {
            links[4 * (size_t)i] = n.left;
            links[4 * (size_t)i + 1] = n.right;
            links[4 * (size_t)i + 2] = n.pad0;
            links[4 * (size_t)i + 3] = n.pad1;
        }
        if (boxes)
            for (int c = 0; c < 3; ++c) {
                boxes[12 * (size_t)i + c] = n.llo[c];
                boxes[12 * (size_t)i + 3 + c] = n.lhi[c];
                boxes[12 * (size_t)i + 6 + c] = n.rlo[c];
                boxes[12 * (size_t)i + 9 + c] = n.rhi[c];
            }
    }
}

// ---------------------------------------------------------------------------
// group walk: shared by the forward and the backward
//
// One wavefront owns one 64-point group.  Candidate Gaussians (those whose box
// overlaps the group's box) are enumerated in a fixed order:
//   * wide traversal: each step pops up to 64 tree nodes from a wave-private LDS
//     stack, one per lane (independent vector loads), tests both child boxes and
//     compacts the overlapping internal children back onto the stack and the
//     overlapping leaves into a leaf list (ballot + mbcnt, lane order);
//   * leaf expansion: the leaves' Gaussians are box-tested 64 at a time and
//     appended to a candidate ring; every 64 candidates are handed to `chunk`.
// If the wide stack has no room for a full step, one node is taken off it and
// its subtree is walked depth first with a second, depth-bounded stack.
// ---------------------------------------------------------------------------
struct GroupBox {
    float lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ bool box_overlaps(float lx, float ly, float lz, float hx, float hy, float hz,
                                             const GroupBox& b)
{
    return boxes_overlap(lx, ly, lz, hx, hy, hz, b.lx, b.ly, b.lz, b.hx, b.hy, b.hz);
}

constexpr int kWideStack = 256;
constexpr int kSeqStack = 128;
constexpr int kRing = 128;

struct WalkLds {
    int wstack[kWideStack];
    int sstack[kSeqStack];
    int leaves[128];
    int ring[kRing];
};

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ void lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Box of the points of the lanes with `in` set (wave-uniform result; NaN
// coordinates are skipped, as fminf / fmaxf skip them; no such point gives the
// empty box +inf .. -inf).
__device__ __forceinline__ GroupBox masked_box(float px, float py, float pz, bool in)
{
    auto lo = [&](float v) { return fkey_inv(wave_extreme_key<false>(in && v == v ? fkey(v) : kKeyPosInf)); };
    auto hi = [&](float v) { return fkey_inv(wave_extreme_key<true>(in && v == v ? fkey(v) : kKeyNegInf)); };
    GroupBox b;
    b.lx = lo(px); b.ly = lo(py); b.lz = lo(pz);
    b.hx = hi(px); b.hy = hi(py); b.hz = hi(pz);
    return b;
}

// Expected candidate count of a box, up to a constant factor: the volume of the
// box grown by the mean Gaussian box extent r.
__device__ __forceinline__ float box_cost(const GroupBox& b, const float r[3])
{
    return (b.hx - b.lx + r[0]) * (b.hy - b.ly + r[1]) * (b.hz - b.lz + r[2]);
}

// Runs: a group's valid lanes [0, nv) are cut into runs of consecutive
// lanes, each walked with its own box.  Consecutive Morton-sorted points can
// straddle a coarse Morton boundary; their common box is then much larger than
// the point clouds on either side.  A run [a, b) is cut where its first and last
// keys first differ (the keys are sorted, so the lanes with that bit set form a
// suffix) when the two halves' cost is below 3/4 of the whole's.  Returns the
// run starts as a lane bit mask.  Any partition gives the same sums; this one
// depends only on the group's points, the mean Gaussian extent and run_points
// (runs never exceed run_points lanes: 4, 8, 16, 32 or 64, a tuning parameter;
// at 64 a group has at most 8 runs).
__device__ __forceinline__ uint64_t split_runs(float px, float py, float pz, int nv, const float r[3], int run_points)
{
    const int lane = threadIdx.x & 63;
    if (nv <= 0) return 0ull;
    uint64_t starts = 1ull;
    for (int k = run_points; k < nv && k < 64; k += run_points) starts |= 1ull << k;
    if (nv < 16) return starts;
    const uint32_t code = lane < nv ? sample_code(px, py, pz) : 0xffffffffu;
    uint64_t pending = starts;
    int tests = 0;
    while (pending != 0ull && tests < 7) {
        const int a = __builtin_ctzll(pending);
        pending &= pending - 1;
        const uint64_t after = starts & ~((2ull << a) - 1ull);
        const int b = after ? __builtin_ctzll(after) : nv;
        if (b - a < 16) continue;
        const uint32_t ca = (uint32_t)__builtin_amdgcn_readlane((int)code, a);
        const uint32_t cb = (uint32_t)__builtin_amdgcn_readlane((int)code, b - 1);
        if (ca >= cb) continue;
        ++tests;
        const int hb = 31 - __builtin_clz(ca ^ cb);
        const bool inrun = lane >= a && lane < b;
        const uint64_t upper = __ballot(inrun && ((code >> hb) & 1u));
        const int sp = upper ? __builtin_ctzll(upper) : b;
        if (sp <= a || sp >= b) continue;
        const GroupBox L = masked_box(px, py, pz, inrun && lane < sp);
        const GroupBox R = masked_box(px, py, pz, inrun && lane >= sp);
        GroupBox B;
        B.lx = fminf(L.lx, R.lx); B.ly = fminf(L.ly, R.ly); B.lz = fminf(L.lz, R.lz);
        B.hx = fmaxf(L.hx, R.hx); B.hy = fmaxf(L.hy, R.hy); B.hz = fmaxf(L.hz, R.hz);
        if (box_cost(L, r) + box_cost(R, r) < 0.75f * box_cost(B, r)) {
            starts |= 1ull << sp;
            pending |= (1ull << a) | (1ull << sp);
        }
    }
    return starts;
}

// Dynamic group scheduling.  First the heavy groups (listed in the build by
// their expected cost, so that the longest walks start at once instead of
// trailing the pass): every wave takes the next one from one shared cursor.
// Then the rest: the groups are cut into kQueues contiguous slices, kQueues/8 per XCD
// (workgroups are dealt round-robin to the XCDs, so blockIdx & 7 names the
// wave's XCD); a wave takes the next group of its slice, skipping heavy ones,
// and once that slice is exhausted moves on to the following slices (its own
// XCD's first).  Each slice's counter is on its own cache line.  Returns
// NG when all are done.  The 8 queue words (queue[x * kLine]) and *heavy_cursor must
// be zero at launch.
// Device counters of the walks, each on its own 128-byte line: atomics on one
// line serialise in one L2 channel, and every group claim is one atomic.
constexpr int kLine = 32;                  // uint32 words per 128-byte line
constexpr int kQueues = 8;                 // group queues per pass, a multiple of the 8 XCDs
constexpr int kQFwd = 0;                   // forward group queues, a line each
constexpr int kQBwd = kQueues * kLine;     // backward group queues
constexpr int kQLog = 2 * kQueues * kLine; // candidate-log slot counter
constexpr int kQLane = kQLog + kLine;      // lane-walked groups
constexpr int kQHeavyF = kQLog + 2 * kLine;  // heavy-list cursor, forward
constexpr int kQHeavyB = kQLog + 3 * kLine;  // heavy-list cursor, backward
constexpr int kQWords = kQLog + 4 * kLine;

struct GroupSched {
    const uint32_t* heavy;   // [0] = count, [1..] = heavy groups
    uint32_t* cursor;
    const uint8_t* flag;     // 1 = heavy
    bool heavy_done;
};

__device__ __forceinline__ int next_group(uint32_t* __restrict__ queue, int NG, int& x, int& tries, GroupSched& sc)
{
    const int lane = threadIdx.x & 63;
    if (!sc.heavy_done) {
        const uint32_t nh = (uint32_t)uniformi((int)sc.heavy[0]);
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(sc.cursor, 1u);
        t = (uint32_t)__builtin_amdgcn_readlane((int)t, 0);
        if (t < nh) return uniformi((int)sc.heavy[1 + t]);
        sc.heavy_done = true;
    }
    while (tries < kQueues) {
        const uint32_t b = (uint32_t)(((uint64_t)NG * (uint32_t)x) / kQueues);
        const uint32_t e = (uint32_t)(((uint64_t)NG * (uint32_t)(x + 1)) / kQueues);
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&queue[x * kLine], 1u);
        t = (uint32_t)__builtin_amdgcn_readlane((int)t, 0);
        if (t < e - b) {
            const int g = (int)(b + t);
            if (uniformi((int)sc.flag[g]) != 0) continue;
            return g;
        }
        x = (x + 1) & (kQueues - 1);
        ++tries;
    }
    return NG;
}

// A wave's first queue: one of the slices of its XCD's eighth of the groups
// (blockIdx & 7 names the XCD).
__device__ __forceinline__ int home_queue()
{
    constexpr unsigned per = kQueues / 8;
    return (int)((blockIdx.x & 7u) * per + (blockIdx.x >> 3) % per);
}

__device__ __forceinline__ void mean_extent(const unsigned long long* __restrict__ ext, float r[3])
{
    const unsigned long long m = ext[3];
    for (int a = 0; a < 3; ++a)
        r[a] = uniformf((float)((double)ext[a] * (1.0 / 1048576.0) / (double)(m > 0 ? m : 1)));
}

// Candidate enumeration.  chunk(base, n) is called with the ring positions
// [base, base+n) (n <= 64, modulo kRing) holding Gaussian indices.  Written as
// one loop whose every pass does exactly one of: hand a chunk over, expand 64
// leaf Gaussians into the ring, or take one traversal step; so `chunk` has a
// single call site and is inlined once.
// One volume's tree: Gaussians [g_lo, g_hi), leaves [leaf_base, leaf_base + NL)
// of L Gaussians, root node `root` (NL > 1) and its box.
struct TreeView {
    const float* root_box;
    int g_lo, g_hi, leaf_base, NL, root;
};

__device__ __forceinline__ TreeView tree_of(const BatchView& b, const float* __restrict__ root_boxes, int v)
{
    TreeView t;
    t.root_box = root_boxes + 6 * v;
    t.g_lo = b.P_off[v];
    t.g_hi = b.P_off[v + 1];
    t.leaf_base = b.L_off[v];
    t.NL = b.L_off[v + 1] - b.L_off[v];
    t.root = b.N_off[v];
    return t;
}

template <class ChunkFn>
__device__ __forceinline__ bool enumerate_candidates(const GRecord* __restrict__ rec, const BNode* __restrict__ nodes,
                                                     const TreeView& tv, int L, const GroupBox& gb, WalkLds& s,
                                                     ChunkFn&& chunk)
{
    const float* __restrict__ root_box = tv.root_box;
    const int NL = tv.NL;
    const int lane = threadIdx.x & 63;
    uint32_t chead = 0, ctail = 0;
    bool ok = true;
    int nl = 0, eb = 0;             // leaves s.leaves[0..nl) under expansion; next Gaussian offset
    int sp = 0;                     // wide stack depth
    int dnode = -1, ssp = 0;        // depth-first walk of one subtree while dnode >= 0
    // t / L as umulhi(t, ceil(2^32 / L)): exact for t < 2^32 / L (here t < 128 L)
    const uint32_t l_recip = L > 1 ? (uint32_t)((0x100000000ull + (uint64_t)L - 1u) / (uint64_t)L) : 0u;
    if (NL <= 0) return true;
    if (NL == 1) {
        if (boxes_overlap_mask(root_box[0], root_box[1], root_box[2], root_box[3], root_box[4], root_box[5], gb.lx,
                               gb.ly, gb.lz, gb.hx, gb.hy, gb.hz) != 0ull) {
            if (lane == 0) s.leaves[0] = tv.leaf_base;
            nl = 1;
        }
    } else {
        if (lane == 0) s.wstack[0] = tv.root;
        sp = 1;
    }
    lds_fence();
    while (true) {
        // the walk state is wave-uniform: say so, so that it lives in scalar
        // registers and the loop branches on the scalar unit
        chead = (uint32_t)uniformi((int)chead);
        ctail = (uint32_t)uniformi((int)ctail);
        nl = uniformi(nl);
        eb = uniformi(eb);
        sp = uniformi(sp);
        dnode = uniformi(dnode);
        ssp = uniformi(ssp);
        const uint32_t pending = ctail - chead;
        const bool exp_done = eb >= nl * L;
        const bool trav_done = sp == 0 && dnode < 0;
        if (pending >= 64u || (pending > 0u && exp_done && trav_done)) {
            const int n = pending >= 64u ? 64 : (int)pending;
            chunk(chead, n);
            chead += (uint32_t)n;
            continue;
        }
        if (!exp_done) {
            // box-test the next 64 Gaussians of the pending leaves
            const int t = eb + lane;
            int g = tv.g_hi;
            if (t < nl * L) {
                const int q = L > 1 ? (int)__umulhi((uint32_t)t, l_recip) : t;
                g = tv.g_lo + (s.leaves[q] - tv.leaf_base) * L + (t - q * L);
            }
            const bool live = g < tv.g_hi;
            float4 b0 = make_float4(0.f, 0.f, 0.f, 0.f);
            float2 b1 = make_float2(0.f, 0.f);
            if (live) {
                b0 = *reinterpret_cast<const float4*>(&rec[g].lox);
                b1 = *reinterpret_cast<const float2*>(&rec[g].hiy);
            }
            const uint64_t bal = __ballot(live) & boxes_overlap_mask(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, gb.lx,
                                                                      gb.ly, gb.lz, gb.hx, gb.hy, gb.hz);
            if (__builtin_amdgcn_inverse_ballot_w64(bal)) s.ring[(ctail + lane_rank(bal)) & (kRing - 1)] = g;
            ctail += (uint32_t)__popcll(bal);
            eb += 64;
            lds_fence();
            continue;
        }
        if (trav_done) break;
        nl = 0;
        eb = 0;
        if (dnode >= 0) {
            // one node of the depth-first walk
            const BNode nd = nodes[dnode];
            // every lane tests the same node: the masks are all lanes or none
            const bool ovl = boxes_overlap_mask(nd.llo[0], nd.llo[1], nd.llo[2], nd.lhi[0], nd.lhi[1], nd.lhi[2],
                                                gb.lx, gb.ly, gb.lz, gb.hx, gb.hy, gb.hz) != 0ull;
            const bool ovr = boxes_overlap_mask(nd.rlo[0], nd.rlo[1], nd.rlo[2], nd.rhi[0], nd.rhi[1], nd.rhi[2],
                                                gb.lx, gb.ly, gb.lz, gb.hx, gb.hy, gb.hz) != 0ull;
            const int cl = uniformi(nd.left), cr = uniformi(nd.right);
            if (ovl && cl < 0) { if (lane == 0) s.leaves[nl] = ~cl; ++nl; }
            if (ovr && cr < 0) { if (lane == 0) s.leaves[nl] = ~cr; ++nl; }
            const bool dl = ovl && cl >= 0, dr = ovr && cr >= 0;
            if (dl) {
                if (dr) {
                    if (ssp >= kSeqStack) {
                        ok = false;
                        dnode = -1;
                        sp = 0;
                    } else {
                        if (lane == 0) s.sstack[ssp] = cr;
                        ++ssp;
                        dnode = cl;
                    }
                } else {
                    dnode = cl;
                }
            } else if (dr) {
                dnode = cr;
            } else {
                lds_fence();
                dnode = ssp > 0 ? uniformi(s.sstack[--ssp]) : -1;
            }
            lds_fence();
            continue;
        }
        int k = min(64, min(sp, kWideStack - sp));
        if (k == 0) {
            // no room for a wide step: walk the top subtree depth first
            dnode = uniformi(s.wstack[sp - 1]);
            sp -= 1;
            ssp = 0;
            continue;
        }
        const int base = sp - k;
        int node = -1;
        if (lane < k) node = s.wstack[base + lane];
        sp = base;
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, c = a, d = a;
        if (lane < k) {
            const float4* q = reinterpret_cast<const float4*>(&nodes[node]);
            a = q[0];
            b = q[1];
            c = q[2];
            d = q[3];
        }
        // a = llo.xyz lhi.x | b = lhi.yz rlo.xy | c = rlo.z rhi.xyz | d = left right . .
        const int cl = __builtin_bit_cast(int, d.x), cr = __builtin_bit_cast(int, d.y);
        const uint64_t act = __ballot(lane < k);
        const uint64_t ovl = act & boxes_overlap_mask(a.x, a.y, a.z, a.w, b.x, b.y, gb.lx, gb.ly, gb.lz, gb.hx,
                                                      gb.hy, gb.hz);
        const uint64_t ovr = act & boxes_overlap_mask(b.z, b.w, c.x, c.y, c.z, c.w, gb.lx, gb.ly, gb.lz, gb.hx,
                                                      gb.hy, gb.hz);
        const uint64_t inl = __ballot(cl >= 0), inr = __ballot(cr >= 0);   // internal children
        const uint64_t bpl = ovl & inl, bpr = ovr & inr, bll = ovl & ~inl, blr = ovr & ~inr;
        const int npl = __popcll(bpl), npr = __popcll(bpr), nll = __popcll(bll);
        lds_fence();
        if (__builtin_amdgcn_inverse_ballot_w64(bpl)) s.wstack[sp + lane_rank(bpl)] = cl;
        if (__builtin_amdgcn_inverse_ballot_w64(bpr)) s.wstack[sp + npl + lane_rank(bpr)] = cr;
        if (__builtin_amdgcn_inverse_ballot_w64(bll)) s.leaves[lane_rank(bll)] = ~cl;
        if (__builtin_amdgcn_inverse_ballot_w64(blr)) s.leaves[nll + lane_rank(blr)] = ~cr;
        sp += npl + npr;
        nl = nll + __popcll(blr);
        lds_fence();
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Lane mode.  Where a group's points are sparse next to its Gaussians, the
// shared walk of the group's box meets thousands of candidates for a few dozen
// pairs per point, one 64-candidate chunk after another on one wavefront.  Such
// a group is walked per lane instead: every lane runs a left-first depth-first
// walk with its own point as the query (so a point meets its Gaussians in
// ascending record order).  The lanes step in lockstep -- one node or one
// Gaussian each per step -- and visit() is called converged after every step.
// The forward and the backward take the same decision from the same inputs.
// ---------------------------------------------------------------------------
constexpr int kLaneStack = 64;
constexpr float kLaneModeCandidates = 8192.0f;

// Gaussians per unit volume of the volume's root box.
__device__ __forceinline__ float gaussian_density(const TreeView& tv)
{
    const float* __restrict__ rb = tv.root_box;
    const float vol = fmaxf(rb[3] - rb[0], 1e-12f) * fmaxf(rb[4] - rb[1], 1e-12f) * fmaxf(rb[5] - rb[2], 1e-12f);
    return (float)(tv.g_hi - tv.g_lo) / vol;
}

// policy 0: lane mode where the expected candidates of the shared walk exceed
// kLaneModeCandidates -- Gaussians per unit volume of the volume's root box times
// the run boxes (starts: the group's runs over lanes [0, nv)) grown by the mean
// Gaussian extent; 1: never; 2: always.
__device__ __forceinline__ bool lane_mode_group(const TreeView& tv, float px, float py, float pz, int nv,
                                                uint64_t starts, const float reach[3], int policy)
{
    if (policy != 0) return policy == 2;
    if (nv <= 0) return false;
    const int lane = threadIdx.x & 63;
    float cost = 0.f;
    uint64_t rs = starts;
    while (rs != 0ull) {
        const int a = __builtin_ctzll(rs);
        rs &= rs - 1;
        const int b = rs ? __builtin_ctzll(rs) : nv;
        cost += box_cost(masked_box(px, py, pz, lane >= a && lane < b), reach);
    }
    return uniformi(gaussian_density(tv) * cost > kLaneModeCandidates ? 1 : 0) != 0;
}

// Heavy groups: expected candidates (from four 16-point segment boxes: runs hold
// at least 16 points) above kHeavyCandidates.  One wave per group; heavy groups
// are flagged and listed in heavy[1 .. heavy[0]] (heavy[0] zero at launch).
constexpr float kHeavyCandidates = 4096.0f;

// Block 0 also clears the forward walk's queue words (its group queues and the
// words from kQLog on), which nothing reads before the walk.
__global__ __launch_bounds__(256) void group_cost_kernel(const float4* __restrict__ pts, int NG, BatchView bv,
                                                         const float* __restrict__ root_box,
                                                         const unsigned long long* __restrict__ ext_sum,
                                                         uint32_t* __restrict__ heavy, uint8_t* __restrict__ flag,
                                                         uint32_t* __restrict__ queues)
{
    if (blockIdx.x == 0) {
        for (int k = (int)threadIdx.x; k < kQueues * kLine; k += 256) queues[kQFwd + k] = 0u;
        for (int k = (int)threadIdx.x; k < kQWords - kQLog; k += 256) queues[kQLog + k] = 0u;
    }
    const int g = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (g >= NG) return;
    const int lane = threadIdx.x & 63;
    const int vol = bv.V == 1 ? 0 : uniformi(volume_of(bv.G_off, bv.V, g));
    const TreeView tv = tree_of(bv, root_box, vol);
    float reach[3];
    mean_extent(ext_sum + 4 * vol, reach);
    const float4 p = pts[(size_t)g * 64 + lane];
    const bool valid = __builtin_bit_cast(int, p.w) >= 0 && p.x == p.x && p.y == p.y && p.z == p.z;
    // min / max over each 16-lane segment: an inclusive scan of the coordinates'
    // integer keys within the row (DPP row shifts, one instruction per step;
    // lanes without a source keep their value), complete at the row's last lane
    auto seg = [&](float v, auto mx) {
        constexpr bool MAX = decltype(mx)::value;
        constexpr int id = MAX ? (int)0x80000000u : 0x7fffffff;
        int k = valid ? fkey(v) : (MAX ? kKeyNegInf : kKeyPosInf);
        auto step = [&](auto ctrl) {
            const int o = __builtin_amdgcn_update_dpp(id, k, decltype(ctrl)::value, 0xF, 0xF, false);
            k = MAX ? (o > k ? o : k) : (o < k ? o : k);
        };
        step(std::integral_constant<int, 0x111>{});
        step(std::integral_constant<int, 0x112>{});
        step(std::integral_constant<int, 0x114>{});
        step(std::integral_constant<int, 0x118>{});
        return fkey_inv(k);
    };
    const std::false_type kmin;
    const std::true_type kmax;
    const float lx = seg(p.x, kmin), ly = seg(p.y, kmin), lz = seg(p.z, kmin);
    const float hx = seg(p.x, kmax), hy = seg(p.y, kmax), hz = seg(p.z, kmax);
    float c = ((lane & 15) == 15 && hx >= lx) ? (hx - lx + reach[0]) * (hy - ly + reach[1]) * (hz - lz + reach[2]) : 0.f;
    c += lane_xor<16>(c);
    c += lane_xor<32>(c);
    const bool is_heavy = gaussian_density(tv) * c > kHeavyCandidates;
    if (lane == 15) {   // the segment sums end in lanes 15, 31, 47, 63
        flag[g] = is_heavy ? 1 : 0;
        if (is_heavy) heavy[1 + atomicAdd(heavy, 1u)] = (uint32_t)g;
    }
}

// visit(hit, gi, A = mean|w, B = conic c0..c3, C = c4 c5 v id, quad, r0, r1, r2,
// dx, dy, dz): hit = Gaussian gi's box holds the point and the power -0.5 quad is
// inside the window (the predicate of the shared walk).  Returns false if a lane's
// stack overflowed (that lane's walk is then incomplete).
template <bool PRECISE, class Visit>
__device__ __forceinline__ bool lane_walk(const GRecord* __restrict__ rec, const float4* __restrict__ fac_lo,
                                          const BNode* __restrict__ nodes, const TreeView& tv, int L, float px,
                                          float py, float pz, bool active, Visit&& visit)
{
    int stk[kLaneStack];
    int sp = 0, g = 0, gend = 0;
    bool ok = true;
    if (active && tv.NL > 0) stk[sp++] = tv.NL == 1 ? ~tv.leaf_base : tv.root;
    while (true) {
        const bool more = g < gend || sp > 0;
        if (__ballot(more) == 0ull) break;
        bool hit = false;
        int gi = -1;
        float4 A = make_float4(0.f, 0.f, 0.f, 0.f), B = A, C = A;
        float quad = 0.f, r0 = 0.f, r1 = 0.f, r2 = 0.f, dx = 0.f, dy = 0.f, dz = 0.f;
        if (g < gend) {
            gi = g++;
            const float4* q = reinterpret_cast<const float4*>(&rec[gi]);
            const float4 d = q[3], e = q[4];   // lo.xyz hi.x | hi.yz
            if (point_in_box(px, py, pz, d.x, d.y, d.z, d.w, e.x, e.y)) {
                A = q[0];
                B = q[1];
                C = q[2];
                quad = record_quad<PRECISE>(A, B, C, fac_lo, px, py, pz, dx, dy, dz, r0, r1, r2);
                hit = quad_in_window(quad);
            }
        } else if (sp > 0) {
            const int n = stk[--sp];
            if (n < 0) {
                g = tv.g_lo + (~n - tv.leaf_base) * L;
                gend = min(g + L, tv.g_hi);
            } else {
                const float4* q = reinterpret_cast<const float4*>(&nodes[n]);
                const float4 a = q[0], b = q[1], c = q[2], d = q[3];
                // a = llo.xyz lhi.x | b = lhi.yz rlo.xy | c = rlo.z rhi.xyz | d = left right . .
                const bool ovl = point_in_box(px, py, pz, a.x, a.y, a.z, a.w, b.x, b.y);
                const bool ovr = point_in_box(px, py, pz, b.z, b.w, c.x, c.y, c.z, c.w);
                if (sp + 2 > kLaneStack) {
                    ok = false;
                    sp = 0;
                } else {
                    if (ovr) stk[sp++] = __builtin_bit_cast(int, d.y);
                    if (ovl) stk[sp++] = __builtin_bit_cast(int, d.x);
                }
            }
        }
        visit(hit, gi, A, B, C, quad, r0, r1, r2, dx, dy, dz);
    }
    return ok;
}

// ---------------------------------------------------------------------------
// forward: lane = point.  Per chunk of <= 64 candidates: lane j holds candidate
// j's box and finds the run's points inside it by rank intervals (GroupAxes);
// a 64x64 bit transpose turns that into each point's candidate mask.  Every
// lane then walks its own set bits in candidate order (next candidate
// prefetched from LDS) and accumulates w e^power and w e^power v with Kahan
// compensation.
// ---------------------------------------------------------------------------
// Candidate log: the forward appends every chunk it evaluates to a slot of 832
// bytes {next slot, n, pad | Gaussian index per lane | candidate-major box-hit
// mask per candidate lane (the points inside its box: the backward's layout)};
// a group's slots form a linked list headed by group_head[g].
// Waves reserve slots 8 at a time from one counter.  The backward replays the
// log instead of walking the tree and box-testing again.
constexpr uint32_t kSlotBytes = 832;
constexpr uint32_t kSlotBlock = 8;
constexpr uint32_t kLogEnd = 0xffffffffu;
constexpr uint32_t kLogIncomplete = 0xfffffffeu;
constexpr uint32_t kLogLaneMode = 0xfffffffdu;   // the group was walked per lane: nothing logged

struct LogCursor {
    uint32_t next = 0, end = 0;     // reserved slots [next, end)
};

__device__ __forceinline__ uint32_t log_take(LogCursor& c, uint32_t* __restrict__ counter)
{
    if (c.next == c.end) {
        uint32_t b = 0;
        if ((threadIdx.x & 63) == 0) b = atomicAdd(counter, kSlotBlock);
        b = (uint32_t)__builtin_amdgcn_readlane((int)b, 0);
        c.next = b;
        c.end = b + kSlotBlock;
    }
    return c.next++;
}

// 64x64 bit-matrix transpose across the wavefront: lane r holds row r (bit c =
// element (r, c)); returns column `lane`.  Six block-swap stages: at stage j the
// lanes with bit j clear trade their bits with bit j set against the partner's
// (lane ^ j) bits with bit j clear.
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t a)
{
    const int lane = threadIdx.x & 63;
    uint32_t lo = (uint32_t)a, hi = (uint32_t)(a >> 32);
    {
        // lanes < 32 send their high word and keep the partner's low word
        const uint32_t send = (lane & 32) ? lo : hi;
        const uint32_t got = lane_xor<32>(send);
        if (lane & 32) lo = got;
        else hi = got;
    }
    auto stage = [&](auto jc, const uint32_t m) {
        constexpr int j = decltype(jc)::value;
        const uint32_t plo = lane_xor<j>(lo), phi = lane_xor<j>(hi);
        if (lane & j) {
            lo = (lo & ~m) | ((plo & ~m) >> j);
            hi = (hi & ~m) | ((phi & ~m) >> j);
        } else {
            lo = (lo & m) | ((plo & m) << j);
            hi = (hi & m) | ((phi & m) << j);
        }
    };
    stage(std::integral_constant<int, 16>{}, 0x0000ffffu);
    stage(std::integral_constant<int, 8>{}, 0x00ff00ffu);
    stage(std::integral_constant<int, 4>{}, 0x0f0f0f0fu);
    stage(std::integral_constant<int, 2>{}, 0x33333333u);
    stage(std::integral_constant<int, 1>{}, 0x55555555u);
    return ((uint64_t)hi << 32) | lo;
}

// Box-vs-group masks by rank intervals.  Per group, the 64 point coordinates are
// sorted along each axis (bitonic, key (value, lane); NaN and padding as +inf)
// and pref[a][k] = the lanes of the k smallest values.  A box [lo, hi] then holds,
// along axis a, exactly the points of rank r0 <= r < r1 with r0 = #(v < lo) and
// r1 = #(v <= hi) -- the same comparisons as lo <= v <= hi -- so its point mask
// is AND_a (pref[a][r1] & ~pref[a][r0]): two 7-step binary searches per axis
// instead of six compares per point.  Points with a NaN coordinate never pass a
// comparison and are removed by `finite`; a box with a NaN bound holds nothing.
struct GroupAxes {
    float vals[3][64];
    uint64_t pref[3][65];
};

// Bitonic sort of (v, id) over the wave's 64 lanes, ascending by v: the
// compare-exchange passes of merge size K, partner distance J = K/2 .. 1.  The
// lane that keeps the smaller of its pair takes the partner's entry if that is
// strictly smaller, the other lane if it is strictly larger, so equal values
// stay put on both sides (the pair stays a permutation).  Their order among
// themselves is immaterial: box_points only reads prefix masks at ranks
// #(v < x) and #(v <= x), which end between distinct values.
constexpr uint64_t keep_min_lanes(int K, int J)
{
    uint64_t m = 0;
    for (int l = 0; l < 64; ++l)
        if (((l & K) == 0) == ((l & J) == 0)) m |= 1ull << l;
    return m;
}

template <int K, int J>
__device__ __forceinline__ void bitonic_pass(float& v, int& id, int lane)
{
    const float ov = lane_xor<J>(v);
    const int oid = lane_xor<J>(id);
    constexpr uint64_t keep_min = keep_min_lanes(K, J);
    const uint64_t less = __builtin_amdgcn_ballot_w64(ov < v), more = __builtin_amdgcn_ballot_w64(ov > v);
    const bool take = __builtin_amdgcn_inverse_ballot_w64((keep_min & less) | (~keep_min & more));
    v = take ? ov : v;
    id = take ? oid : id;
    if constexpr (J > 1) bitonic_pass<K, J / 2>(v, id, lane);
}
template <int K>
__device__ __forceinline__ void bitonic_sort64(float& v, int& id, int lane)
{
    bitonic_pass<K, K / 2>(v, id, lane);
    if constexpr (K < 64) bitonic_sort64<2 * K>(v, id, lane);
}

// Returns the mask of lanes whose point has no NaN coordinate among `valid` lanes.
__device__ __forceinline__ uint64_t build_axes(float px, float py, float pz, bool valid, GroupAxes& ax)
{
    const int lane = threadIdx.x & 63;
    const bool finite = valid && px == px && py == py && pz == pz;
    const float inf = __builtin_inff();
    auto bit_or = [](uint32_t x, uint32_t y) { return x | y; };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float c = a == 0 ? px : (a == 1 ? py : pz);
        float v = finite ? c : inf;
        int id = lane;
        bitonic_sort64<2>(v, id, lane);
        ax.vals[a][lane] = v;
        // prefix masks: lane r holds the lanes of the r + 1 smallest values
        const uint32_t lo = wave_inclusive_scan(id < 32 ? 1u << id : 0u, bit_or);
        const uint32_t hi = wave_inclusive_scan(id >= 32 ? 1u << (id - 32) : 0u, bit_or);
        ax.pref[a][lane + 1] = ((uint64_t)hi << 32) | lo;
        if (lane == 0) ax.pref[a][0] = 0ull;
    }
    return __ballot(finite);
}

// #(vals < x) (strict) or #(vals <= x) over 64 ascending values in LDS.
// The search moves an LDS pointer (one add and one select per step; the probe's
// displacement is an immediate of the LDS read).  After the six halving steps at
// most 63 values have passed, so the last probe is always inside the array.
typedef __attribute__((address_space(3))) const float lds_cfloat;
template <bool STRICT>
__device__ __forceinline__ int rank64(const float* __restrict__ vals, float x)
{
    lds_cfloat* const v0 = (lds_cfloat*)vals;
    lds_cfloat* q = v0;   // q - v0 = values known to pass
#pragma unroll
    for (int st = 32; st >= 1; st >>= 1) {
        const float v = q[st - 1];
        if (STRICT ? (v < x) : (v <= x)) q += st;
    }
    const float v = q[0];
    if (STRICT ? (v < x) : (v <= x)) ++q;
    return (int)(q - v0);
}

// Points of the group inside the box (lox..hiz), before run and finite masks.
__device__ __forceinline__ uint64_t box_points(const GroupAxes& ax, float lox, float loy, float loz, float hix,
                                               float hiy, float hiz)
{
    const float lo[3] = { lox, loy, loz }, hi[3] = { hix, hiy, hiz };
    uint64_t m = ~0ull;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int r0 = rank64<true>(ax.vals[a], lo[a]);
        const int r1 = rank64<false>(ax.vals[a], hi[a]);
        m &= ax.pref[a][r1] & ~ax.pref[a][r0];
    }
    const bool ok = lox == lox && loy == loy && loz == loz && hix == hix && hiy == hiy && hiz == hiz;
    return ok ? m : 0ull;
}

struct FwdLds {
    WalkLds walk;
    float4 chunk[64][3];   // mean|w, conic c0..c3, c4 c5 v id
    GroupAxes axes;
};

template <bool KAHAN, bool PRECISE>
__global__ __launch_bounds__(1024) void forward_kernel(const GRecord* __restrict__ rec,
                                                      const float4* __restrict__ fac_lo,
                                                      const BNode* __restrict__ nodes,
                                                      const float* __restrict__ root_box, BatchView bv, int P, int L,
                                                      const float4* __restrict__ pts, int NG,
                                                      float* __restrict__ out_cells, float* __restrict__ out_weights,
                                                      float2* __restrict__ pvw,
                                                      uint32_t* __restrict__ group_records,
                                                      uint32_t* __restrict__ group_pairs, uint32_t* __restrict__ err,
                                                      uint32_t* __restrict__ queue,
                                                      const unsigned long long* __restrict__ ext_sum,
                                                      int log_on, unsigned char* __restrict__ log, uint32_t log_cap,
                                                      uint32_t* __restrict__ log_counter,
                                                      uint32_t* __restrict__ group_head, int run_points,
                                                      int lane_policy, const uint32_t* __restrict__ heavy,
                                                      uint32_t* __restrict__ heavy_cursor,
                                                      const uint8_t* __restrict__ group_flag)
{
    extern __shared__ float4 lds_dyn[];
    const int wid = uniformi(threadIdx.x >> 6), lane = threadIdx.x & 63;
    FwdLds& s = reinterpret_cast<FwdLds*>(lds_dyn)[wid];
    int x = home_queue(), tries = 0;
    GroupSched sched{ heavy, heavy_cursor, group_flag, false };
    bool ok = true;
    LogCursor cur;
    while (true) {
        const int g = next_group(queue, NG, x, tries, sched);
        if (g >= NG) break;
        const int vol = bv.V == 1 ? 0 : uniformi(volume_of(bv.G_off, bv.V, g));
        const TreeView tv = tree_of(bv, root_box, vol);
        float reach[3];
        mean_extent(ext_sum + 4 * vol, reach);
        const float4* __restrict__ gp = pts + (size_t)g * 64;
        const float4 pme = gp[lane];
        const float px = pme.x, py = pme.y, pz = pme.z;
        const int orig = __builtin_bit_cast(int, pme.w);
        const int nv = __popcll(__ballot(orig >= 0));
        const uint64_t starts = split_runs(px, py, pz, nv, reach, run_points);
        lds_fence();
        const uint64_t finite = build_axes(px, py, pz, orig >= 0, s.axes);
        lds_fence();

        float W = 0.f, cW = 0.f, U = 0.f, cU = 0.f;
        uint32_t npairs = 0, tp = 0, nrec = 0;   // accepted pairs (per lane, wave-uniform), records
        uint64_t runmask = 0;
        uint32_t head = kLogEnd, prev = kLogEnd;
        bool logged = log_on != 0;

        auto chunk = [&](uint32_t base, int n) {
            float lox = 1.f, loy = 1.f, loz = 1.f, hix = 0.f, hiy = 0.f, hiz = 0.f;  // empty box for idle lanes
            int gi = -1;
            if (lane < n) {
                gi = s.walk.ring[(base + lane) & (kRing - 1)];
                const float4* q = reinterpret_cast<const float4*>(&rec[gi]);
                const float4 a = q[0], b = q[1], c = q[2], d = q[3], e = q[4];
                s.chunk[lane][0] = a;
                s.chunk[lane][1] = b;
                s.chunk[lane][2] = c;
                lox = d.x; loy = d.y; loz = d.z; hix = d.w; hiy = e.x; hiz = e.y;
            }
            // lane j: the run's points inside candidate j's box; transposed, lane p:
            // the candidates whose box holds point p
            const uint64_t cm = lane < n ? box_points(s.axes, lox, loy, loz, hix, hiy, hiz) & runmask & finite : 0ull;
            uint64_t m = wave_transpose64(cm);
            lds_fence();
            if (logged) {
                const uint32_t slot = log_take(cur, log_counter);
                if (slot < log_cap) {
                    unsigned char* sb = log + (size_t)slot * kSlotBytes;
                    // streaming stores: the backward reads the log back long after
                    __builtin_nontemporal_store(gi, reinterpret_cast<int*>(sb + 64) + lane);
                    __builtin_nontemporal_store((unsigned long long)cm,
                                                reinterpret_cast<unsigned long long*>(sb + 320) + lane);
                    if (lane == 0) {
                        reinterpret_cast<uint32_t*>(sb)[0] = kLogEnd;
                        reinterpret_cast<uint32_t*>(sb)[1] = (uint32_t)n;
                        if (prev != kLogEnd) reinterpret_cast<uint32_t*>(log + (size_t)prev * kSlotBytes)[0] = slot;
                    }
                    if (prev == kLogEnd) head = slot;
                    prev = slot;
                } else {
                    logged = false;
                }
            }
            // the backward writes one record per candidate with a box hit in the run
            nrec += (uint32_t)__popcll(__ballot(cm != 0ull));
            // each point lane walks its candidates in ascending order, two per pass
            // (both loads issued before either pair is evaluated)
            const float4* __restrict__ cb = &s.chunk[0][0];
            // Branch-free: a rejected pair (or the missing second one) adds zero.
            auto pair = [&](const float4& A, const float4& B, const float4& C, bool live, float& wgt, float& wv) {
                float dx, dy, dz, r0, r1, r2;
                const float quad = record_quad<PRECISE>(A, B, C, fac_lo, px, py, pz, dx, dy, dz, r0, r1, r2);
                const bool in = live && quad_in_window(quad);
                const float e = quad_exp(quad);
                wgt = in ? A.w * e : 0.0f;
                wv = in ? wgt * C.z : 0.0f;
                npairs += in ? 1u : 0u;   // per lane: the candidate loop diverges
            };
            while (m != 0ull) {
                const uint32_t j0 = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const bool two = m != 0ull;
                const uint32_t j1 = two ? (uint32_t)__builtin_ctzll(m) : j0;
                m &= m - 1;  // stays 0 when there was no second bit
                const float4 A0 = cb[3 * j0], B0 = cb[3 * j0 + 1], C0 = cb[3 * j0 + 2];
                const float4 A1 = cb[3 * j1], B1 = cb[3 * j1 + 1], C1 = cb[3 * j1 + 2];
                float w0, v0, w1, v1;
                pair(A0, B0, C0, true, w0, v0);
                pair(A1, B1, C1, two, w1, v1);
                // the two terms are added first (one rounding of at most u of
                // their sum), then compensated into the running sums: the total
                // error stays within a few u of the sum
                const float wgt = w0 + w1, wv = v0 + v1;
                if (KAHAN) {
                    float y = wgt - cW;
                    float t = W + y;
                    cW = (t - W) - y;
                    W = t;
                    y = wv - cU;
                    t = U + y;
                    cU = (t - U) - y;
                    U = t;
                } else {
                    W += wgt;
                    U += wv;
                }
            }
            lds_fence();
        };
        const bool lane_mode = lane_mode_group(tv, px, py, pz, nv, starts, reach, lane_policy);
        if (lane_mode) {
            const bool active = orig >= 0 && px == px && py == py && pz == pz;
            ok = lane_walk<PRECISE>(rec, fac_lo, nodes, tv, L, px, py, pz, active,
                           [&](bool hit, int, const float4& A, const float4&, const float4& C, float quad, float,
                               float, float, float, float, float) {
                               tp += (uint32_t)__popcll(__ballot(hit));
                               if (!hit) return;
                               const float e = quad_exp(quad);
                               const float wgt = A.w * e;
                               const float wv = wgt * C.z;
                               if (KAHAN) {
                                   float y = wgt - cW;
                                   float t = W + y;
                                   cW = (t - W) - y;
                                   W = t;
                                   y = wv - cU;
                                   t = U + y;
                                   cU = (t - U) - y;
                                   U = t;
                               } else {
                                   W += wgt;
                                   U += wv;
                               }
                           }) && ok;
            nrec = tp;   // the backward writes one record per pair
            head = kLogLaneMode;
            if (lane == 0) atomicAdd(log_counter + (kQLane - kQLog), 1u);   // lane-walked groups
        } else {
            uint64_t rs = starts;
            while (rs != 0ull) {
                const int a = __builtin_ctzll(rs);
                rs &= rs - 1;
                const int b = rs ? __builtin_ctzll(rs) : nv;
                runmask = (b >= 64 ? ~0ull : ((1ull << b) - 1ull)) & ~((1ull << a) - 1ull);
                const GroupBox gb = masked_box(px, py, pz, lane >= a && lane < b);
                ok = enumerate_candidates(rec, nodes, tv, L, gb, s.walk, chunk) && ok;
            }
        }
        {
            float V = 0.0f, Wo = 0.0f;
            if (orig >= 0) {
                if ((double)W <= kWeightCutoff) {
                    V = -1.0f;
                } else {
                    V = U / W;
                    Wo = W;
                }
                // streaming stores: the loss reads these after the walk
                __builtin_nontemporal_store(V, out_cells + orig);
                __builtin_nontemporal_store(Wo, out_weights + orig);
            }
            // and in the sample set's order for the backward (coalesced)
            typedef float v2f __attribute__((ext_vector_type(2)));
            __builtin_nontemporal_store(v2f{ V, Wo }, reinterpret_cast<v2f*>(pvw) + (size_t)g * 64 + lane);
        }
        if (!lane_mode) tp = wave_sumu(npairs);
        if (lane == 0) {
            group_records[g] = nrec;
            group_pairs[g] = tp;
            if (log_on) {
                group_head[g] = logged ? head : kLogIncomplete;
                if (!logged) atomicOr(err, (uint32_t)GVC_ERRFLAG_CHUNK_OVERFLOW);
            }
        }
    }
    if (!ok && lane == 0) atomicOr(err, (uint32_t)GVC_ERRFLAG_STACK_OVERFLOW);
}

// ---------------------------------------------------------------------------
// Debug counts (gvc_forward_pair_counts): accepted pairs per sample and per
// Gaussian, the reference's count_intersections (CR/forward.cu:180, :257)
// printed by its debug mode (CR/rasterizer_impl.cu:175-216).  One wave per group
// replays the group's candidate log -- a bit transpose gives lane p its point's
// row of box hits, and it evaluates the predicate for each hit candidate -- or, for a lane-walked
// group, walks the tree per lane again.  Integer atomics only (the totals do not
// depend on their order).  Not on the hot path: the walks count nothing per
// Gaussian.
// ---------------------------------------------------------------------------
template <bool PRECISE>
__global__ __launch_bounds__(256) void pair_count_kernel(const GRecord* __restrict__ rec,
                                                         const float4* __restrict__ fac_lo,
                                                         const BNode* __restrict__ nodes,
                                                         const float* __restrict__ root_box, BatchView bv, int L,
                                                         const float4* __restrict__ pts, int NG,
                                                         const unsigned char* __restrict__ log,
                                                         const uint32_t* __restrict__ group_head,
                                                         uint32_t* __restrict__ sample_counts,
                                                         uint32_t* __restrict__ gaussian_counts,
                                                         uint32_t* __restrict__ err)
{
    __shared__ float4 cand[4][64][3];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = (int)blockIdx.x * 4 + wid;
    if (g >= NG) return;
    const uint32_t head = (uint32_t)uniformi((int)group_head[g]);
    const float4 pme = pts[(size_t)g * 64 + lane];
    const int orig = __builtin_bit_cast(int, pme.w);
    uint32_t count = 0;
    auto tally = [&](bool hit, const float4& C) {
        if (!hit) return;
        ++count;
        atomicAdd(&gaussian_counts[(uint32_t)__builtin_bit_cast(int, C.w) & ~kPreciseBit], 1u);
    };
    if (head == kLogIncomplete) {
        if (lane == 0) atomicOr(err, (uint32_t)GVC_ERRFLAG_CHUNK_OVERFLOW);
    } else if (head == kLogLaneMode) {
        const int vol = bv.V == 1 ? 0 : uniformi(volume_of(bv.G_off, bv.V, g));
        const TreeView tv = tree_of(bv, root_box, vol);
        const bool active = orig >= 0 && pme.x == pme.x && pme.y == pme.y && pme.z == pme.z;
        const bool ok = lane_walk<PRECISE>(rec, fac_lo, nodes, tv, L, pme.x, pme.y, pme.z, active,
                                           [&](bool hit, int, const float4&, const float4&, const float4& C, float,
                                               float, float, float, float, float, float) { tally(hit, C); });
        if (!ok) atomicOr(err, (uint32_t)GVC_ERRFLAG_STACK_OVERFLOW);
    } else {
        for (uint32_t slot = head; slot != kLogEnd;) {
            const unsigned char* sb = log + (size_t)slot * kSlotBytes;
            const uint32_t nxt = (uint32_t)uniformi((int)reinterpret_cast<const uint32_t*>(sb)[0]);
            const int n = uniformi((int)reinterpret_cast<const uint32_t*>(sb)[1]);
            if (lane < n) {
                const float4* q = reinterpret_cast<const float4*>(&rec[reinterpret_cast<const int*>(sb + 64)[lane]]);
                cand[wid][lane][0] = q[0];
                cand[wid][lane][1] = q[1];
                cand[wid][lane][2] = q[2];
            }
            const uint2 row = reinterpret_cast<const uint2*>(sb + 320)[lane];
            // the log keeps candidate-major rows; lane p: the candidates holding point p
            uint64_t m = wave_transpose64(((uint64_t)row.y << 32) | row.x);
            lds_fence();
            while (m != 0ull) {
                const int j = __builtin_ctzll(m);
                m &= m - 1;
                const float4 A = cand[wid][j][0], B = cand[wid][j][1], C = cand[wid][j][2];
                float dx, dy, dz, r0, r1, r2;
                const float quad = record_quad<PRECISE>(A, B, C, fac_lo, pme.x, pme.y, pme.z, dx, dy, dz, r0, r1, r2);
                tally(quad_in_window(quad), C);
            }
            lds_fence();
            slot = nxt;
        }
    }
    if (orig >= 0) sample_counts[orig] = count;
}

// ---------------------------------------------------------------------------
// backward: lane = candidate.  Per chunk each lane holds one candidate's record,
// builds the bitmask of the group's points inside its box (points by scalar
// loads), then walks those points (LDS, next point prefetched) accumulating its
// Gaussian's 11 partial sums; one record per candidate with an accepted pair, at
// the position fixed by the forward's per-group counts.
// ---------------------------------------------------------------------------
// One 64-byte line per record, so the reduction's gathers touch one line each.
struct alignas(64) PartialRecord {
    int gid;           // Gaussian index in BVH order
    float dv, dw;      // d/dvalue, d/dweight
    float dmx, dmy, dmz;
    float dc[6];       // d/dconic, packed, off-diagonals counted once each (factor frame on the precise path)
    float pad[4];
};
static_assert(sizeof(PartialRecord) == 64, "PartialRecord layout");

// A record and its sort key written with streaming stores: they are read back
// only after the walk (the sort and the sums), from HBM, and would otherwise
// take L2 lines from the records and nodes the walk reads again.
__device__ __forceinline__ void put_record(PartialRecord* __restrict__ out, uint32_t* __restrict__ keys, uint64_t pos,
                                           const PartialRecord& o)
{
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f* src = reinterpret_cast<const v4f*>(&o);
    v4f* dst = reinterpret_cast<v4f*>(out + pos);
#pragma unroll
    for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(src[k], dst + k);
    __builtin_nontemporal_store((uint32_t)o.gid, keys + pos);
}

struct BwdLdsBase {
    float4 ptsA[64];      // x, y, z, V
    float2 ptsB[64];      // gV / W, gW   (zero where W <= cutoff)
    float4 slot[64][3];   // one slice's 11 partial sums + accepted flag
    float4 cand[64][3];   // the chunk's candidates: mean|w, conic c0..c3, c4 c5 v .
    uint8_t owner[256];   // slice -> candidate lane
};
template <bool REPLAY>
struct BwdLds : BwdLdsBase {
};
template <>
struct BwdLds<false> : BwdLdsBase {
    WalkLds walk;
    GroupAxes axes;
};

// Inclusive prefix sum over the wavefront.
__device__ __forceinline__ uint32_t wave_inscan_u32(uint32_t v)
{
    return wave_inclusive_scan(v, [](uint32_t x, uint32_t y) { return x + y; });
}

// m with its k lowest set bits cleared (k < popcount(m)).
__device__ __forceinline__ uint64_t drop_low_bits(uint64_t m, uint32_t k)
{
    if (k == 0) return m;
    uint32_t w = (uint32_t)m;
    int base = 0;
    const uint32_t c0 = (uint32_t)__popc(w);
    if (k >= c0) {
        k -= c0;
        w = (uint32_t)(m >> 32);
        base = 32;
    }
#pragma unroll
    for (int st = 16; st >= 1; st >>= 1) {
        const uint32_t c = (uint32_t)__popc(w & ((1u << st) - 1u));
        if (k >= c) {
            k -= c;
            w >>= st;
            base += st;
        }
    }
    // bit `base` is the k-th set bit: keep it and everything above
    return m & ~((1ull << base) - 1ull);
}

// REPLAY = true: candidates and box hits come from the forward's candidate log.
// REPLAY = false: the group is walked again (runs, tree, box tests), as the
// forward did, so nothing but per-group counts is kept between the passes.
template <bool REPLAY, bool PRECISE>
__global__ __launch_bounds__(1024) void backward_kernel(
    const GRecord* __restrict__ rec, const float4* __restrict__ fac_lo, const BNode* __restrict__ nodes, const float* __restrict__ root_box, BatchView bv,
    int P, int L, const float4* __restrict__ pts, int NG, const float2* __restrict__ pvw,
    const float* __restrict__ gcells, const float* __restrict__ gweights,
    const uint32_t* __restrict__ group_offsets, uint64_t capacity, PartialRecord* __restrict__ out_rec,
    uint32_t* __restrict__ keys, uint32_t* __restrict__ err, uint32_t* __restrict__ queue,
    const unsigned long long* __restrict__ ext_sum, const unsigned char* __restrict__ log,
    const uint32_t* __restrict__ group_head, int run_points, int lane_policy, const uint32_t* __restrict__ heavy,
    uint32_t* __restrict__ heavy_cursor, const uint8_t* __restrict__ group_flag)
{
    extern __shared__ float4 lds_dyn[];
    const int wid = uniformi(threadIdx.x >> 6), lane = threadIdx.x & 63;
    BwdLds<REPLAY>& s = reinterpret_cast<BwdLds<REPLAY>*>(lds_dyn)[wid];
    const uint64_t lt = lanemask_lt();
    int x = home_queue(), tries = 0;
    GroupSched sched{ heavy, heavy_cursor, group_flag, false };
    bool ok = true;
    while (true) {
        const int g = next_group(queue, NG, x, tries, sched);
        if (g >= NG) break;
        uint32_t head = kLogEnd;
        if (REPLAY) {
            head = uniformi((int)group_head[g]);
            if (head == kLogEnd) continue;
            if (head == kLogIncomplete) {
                if (lane == 0) atomicOr(err, (uint32_t)GVC_ERRFLAG_CHUNK_OVERFLOW);
                continue;
            }
        }
        const float4 pme = pts[(size_t)g * 64 + lane];
        const int orig = __builtin_bit_cast(int, pme.w);
        const int vol = bv.V == 1 ? 0 : uniformi(volume_of(bv.G_off, bv.V, g));
        const TreeView tv = tree_of(bv, root_box, vol);
        float reach[3] = { 0.f, 0.f, 0.f };
        int nv = 0;
        uint64_t starts = 0ull;
        if (!REPLAY) {
            mean_extent(ext_sum + 4 * vol, reach);
            nv = __popcll(__ballot(orig >= 0));
            starts = split_runs(pme.x, pme.y, pme.z, nv, reach, run_points);
        }
        const bool lane_mode =
            REPLAY ? head == kLogLaneMode
                   : lane_mode_group(tv, pme.x, pme.y, pme.z, nv, starts, reach, lane_policy);
        float V = 0.f, gVi = 0.f, gWe = 0.f;
        if (orig >= 0) {
            // the four gathers in flight together (one round trip), then the test
            // the forward's value and weight from the state (coalesced), the
            // upstream gradients gathered, all in flight together
            const float2 vw = pvw[(size_t)g * 64 + lane];
            const float Wp = vw.y, Vp = vw.x, gV = gcells[orig], gW = gweights[orig];
            if ((double)Wp > kWeightCutoff) {
                V = Vp;
                gVi = gV / Wp;
                gWe = gW;
            }
        }
        lds_fence();
        s.ptsA[lane] = make_float4(pme.x, pme.y, pme.z, V);
        s.ptsB[lane] = make_float2(gVi, gWe);
        lds_fence();

        uint32_t base_out = group_offsets[g];
        const uint32_t limit = group_offsets[g + 1];

        // one chunk: n candidates, lane j's Gaussian gi and the points inside its box
        // (bit mask m, lanes j >= n empty); the candidates' records are in s.cand
        auto process = [&](int n, int gi, uint64_t m) {

            // Balanced pair loop: the candidates' point lists are cut into slices of
            // at most q pairs (q = ceil(H / 96) for H pairs in the chunk, so at most
            // H / q + 64 <= 160 slices: three rounds at most, two for a typical
            // chunk — measured faster than more, shorter slices), dealt to the lanes
            // 64 at a time; each slice's sums land in an LDS slot and the
            // candidate's lane adds its slices in order.
            const uint32_t h = (uint32_t)__popcll(m);
            const uint32_t H = uniformi((int)wave_sumu(h));
            float dv = 0.f, dw = 0.f, dmx = 0.f, dmy = 0.f, dmz = 0.f;
            float dc0 = 0.f, dc1 = 0.f, dc2 = 0.f, dc3 = 0.f, dc4 = 0.f, dc5 = 0.f;
            if (H > 0) {
                const uint32_t q = H > 96 ? (H + 95) / 96 : 1u;
                const uint32_t sj = (h + q - 1) / q;
                const uint32_t incl = wave_inscan_u32(sj);
                const uint32_t bj = incl - sj;
                const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
                for (uint32_t t = 0; t < sj; ++t) s.owner[bj + t] = (uint8_t)lane;
                lds_fence();
                for (uint32_t r0 = 0; r0 < total; r0 += 64) {
                    const uint32_t v = r0 + (uint32_t)lane;
                    const bool live = v < total;
                    const int j = live ? (int)s.owner[v] : lane;
                    const uint32_t tb = (uint32_t)__shfl((int)bj, j, 64);
                    const uint32_t mlo_j = (uint32_t)__shfl((int)(uint32_t)m, j, 64);
                    const uint32_t mhi_j = (uint32_t)__shfl((int)(uint32_t)(m >> 32), j, 64);
                    const float4 ca = s.cand[j][0], cb = s.cand[j][1], cc = s.cand[j][2];
                    const float cw = ca.w, cv = cc.z;
                    uint64_t mj = live ? drop_low_bits(((uint64_t)mhi_j << 32) | mlo_j, (v - tb) * q) : 0ull;
                    float sv = 0.f, sw = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
                    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f, s5 = 0.f;
                    for (uint32_t it = 0; it < q && mj != 0ull; ++it) {
                        const int p = __builtin_ctzll(mj);
                        mj &= mj - 1;
                        const float4 a = s.ptsA[p];
                        const float2 b = s.ptsB[p];
                        float dx, dy, dz, r0_, r1_, r2_;
                        const float quad = record_quad<PRECISE>(ca, cb, cc, fac_lo, a.x, a.y, a.z, dx, dy, dz,
                                                                r0_, r1_, r2_);
                        if (quad_in_window(quad)) {
                            const float e = quad_exp(quad);
                            const float wgt = cw * e;
                            sv = __builtin_fmaf(b.x, wgt, sv);
                            const float tt = __builtin_fmaf(b.x, cv - a.w, b.y);
                            sw = __builtin_fmaf(tt, e, sw);
                            const float st = wgt * tt;
                            sx = __builtin_fmaf(st, r0_, sx);
                            sy = __builtin_fmaf(st, r1_, sy);
                            sz = __builtin_fmaf(st, r2_, sz);
                            const float dq = -0.5f * st;
                            const float h0 = dq * dx, h1 = dq * dy, h2 = dq * dz;
                            const float h0x2 = h0 + h0, h1x2 = h1 + h1;
                            s0 = __builtin_fmaf(h0, dx, s0);
                            s1 = __builtin_fmaf(h0x2, dy, s1);
                            s2 = __builtin_fmaf(h0x2, dz, s2);
                            s3 = __builtin_fmaf(h1, dy, s3);
                            s4 = __builtin_fmaf(h1x2, dz, s4);
                            s5 = __builtin_fmaf(h2, dz, s5);
                        }
                    }
                    s.slot[lane][0] = make_float4(sv, sw, sx, sy);
                    s.slot[lane][1] = make_float4(sz, s0, s1, s2);
                    s.slot[lane][2] = make_float4(s3, s4, s5, 0.f);
                    lds_fence();
                    // this candidate's slices inside the round
                    const uint32_t lo = bj > r0 ? bj : r0;
                    const uint32_t hi = (bj + sj) < (r0 + 64) ? (bj + sj) : (r0 + 64);
                    for (uint32_t u = lo; u < hi; ++u) {
                        const float4 A = s.slot[u - r0][0], B = s.slot[u - r0][1], C = s.slot[u - r0][2];
                        dv += A.x; dw += A.y; dmx += A.z; dmy += A.w;
                        dmz += B.x; dc0 += B.y; dc1 += B.z; dc2 += B.w;
                        dc3 += C.x; dc4 += C.y; dc5 += C.z;
                    }
                    lds_fence();
                }
            }
            // one record per candidate with a box hit (the forward counted these)
            const bool has = h != 0u;
            const uint64_t bal = __ballot(has);
            if (has) {
                const uint64_t pos = (uint64_t)base_out + (uint32_t)__popcll(bal & lt);
                if (pos < limit && pos < capacity) {
                    PartialRecord o;
                    o.gid = gi;
                    o.dv = dv; o.dw = dw;
                    o.dmx = dmx; o.dmy = dmy; o.dmz = dmz;
                    o.dc[0] = dc0; o.dc[1] = dc1; o.dc[2] = dc2; o.dc[3] = dc3; o.dc[4] = dc4; o.dc[5] = dc5;
                    o.pad[0] = o.pad[1] = o.pad[2] = o.pad[3] = 0.f;
                    put_record(out_rec, keys, pos, o);
                } else {
                    atomicOr(err, (uint32_t)GVC_ERRFLAG_RECORD_OVERFLOW);
                }
            }
            base_out += (uint32_t)__popcll(bal);
        };
        if (lane_mode) {
            // one record per accepted pair, in step order then lane order
            const bool active = orig >= 0 && pme.x == pme.x && pme.y == pme.y && pme.z == pme.z;
            ok = lane_walk<PRECISE>(rec, fac_lo, nodes, tv, L, pme.x, pme.y, pme.z, active,
                           [&](bool hit, int gi, const float4& A, const float4&, const float4& C, float quad,
                               float r0_, float r1_, float r2_, float dx, float dy, float dz) {
                               const uint64_t bal = __ballot(hit);
                               if (hit) {
                                   const uint64_t pos = (uint64_t)base_out + (uint32_t)__popcll(bal & lt);
                                   if (pos < limit && pos < capacity) {
                                       const float e = quad_exp(quad);
                                       const float wgt = A.w * e;
                                       const float tt = __builtin_fmaf(gVi, C.z - V, gWe);
                                       const float st = wgt * tt;
                                       const float dq = -0.5f * st;
                                       const float h0 = dq * dx, h1 = dq * dy, h2 = dq * dz;
                                       const float h0x2 = h0 + h0, h1x2 = h1 + h1;
                                       PartialRecord o;
                                       o.gid = gi;
                                       o.dv = gVi * wgt;
                                       o.dw = tt * e;
                                       o.dmx = st * r0_; o.dmy = st * r1_; o.dmz = st * r2_;
                                       o.dc[0] = h0 * dx; o.dc[1] = h0x2 * dy; o.dc[2] = h0x2 * dz;
                                       o.dc[3] = h1 * dy; o.dc[4] = h1x2 * dz; o.dc[5] = h2 * dz;
                                       o.pad[0] = o.pad[1] = o.pad[2] = o.pad[3] = 0.f;
                                       put_record(out_rec, keys, pos, o);
                                   } else {
                                       atomicOr(err, (uint32_t)GVC_ERRFLAG_RECORD_OVERFLOW);
                                   }
                               }
                               base_out += (uint32_t)__popcll(bal);
                           }) && ok;
        } else if constexpr (REPLAY) {
            // replay the group's log; the next slot's header and rows are loaded
            // before the current chunk is processed
            const unsigned char* sb = log + (size_t)head * kSlotBytes;
            uint32_t nxt = uniformi((int)reinterpret_cast<const uint32_t*>(sb)[0]);
            int n = uniformi((int)reinterpret_cast<const uint32_t*>(sb)[1]);
            int gi = reinterpret_cast<const int*>(sb + 64)[lane];
            uint2 mr = reinterpret_cast<const uint2*>(sb + 320)[lane];
            while (true) {
                uint32_t nxt2 = kLogEnd;
                int n2 = 0, gi2 = -1;
                uint2 mr2 = make_uint2(0u, 0u);
                if (nxt != kLogEnd) {
                    const unsigned char* sb2 = log + (size_t)nxt * kSlotBytes;
                    nxt2 = uniformi((int)reinterpret_cast<const uint32_t*>(sb2)[0]);
                    n2 = uniformi((int)reinterpret_cast<const uint32_t*>(sb2)[1]);
                    gi2 = reinterpret_cast<const int*>(sb2 + 64)[lane];
                    mr2 = reinterpret_cast<const uint2*>(sb2 + 320)[lane];
                }
                const bool active = lane < n;
                const float4* q = reinterpret_cast<const float4*>(&rec[active ? gi : 0]);
                s.cand[lane][0] = q[0];
                s.cand[lane][1] = q[1];
                s.cand[lane][2] = q[2];
                // the logged row of lane j: the points inside candidate j's box
                const uint64_t tm = ((uint64_t)mr.y << 32) | mr.x;
                lds_fence();
                process(n, gi, active ? tm : 0ull);
                if (nxt == kLogEnd) break;
                nxt = nxt2;
                n = n2;
                gi = gi2;
                mr = mr2;
            }
        } else {
            // walk the group's runs again and box-test each chunk (lane = candidate)
            const uint64_t finite = build_axes(pme.x, pme.y, pme.z, orig >= 0, s.axes);
            lds_fence();
            uint64_t runmask = 0;
            auto chunk = [&](uint32_t base, int n) {
                const bool active = lane < n;
                const int gi = active ? s.walk.ring[(base + lane) & (kRing - 1)] : 0;
                const float4* q = reinterpret_cast<const float4*>(&rec[gi]);
                const float4 rd = q[3], re = q[4];
                s.cand[lane][0] = q[0];
                s.cand[lane][1] = q[1];
                s.cand[lane][2] = q[2];
                const uint64_t cm = active ? box_points(s.axes, rd.x, rd.y, rd.z, rd.w, re.x, re.y) & runmask & finite
                                           : 0ull;
                lds_fence();
                process(n, active ? gi : -1, cm);
            };
            uint64_t rs = starts;
            while (rs != 0ull) {
                const int a = __builtin_ctzll(rs);
                rs &= rs - 1;
                const int b = rs ? __builtin_ctzll(rs) : nv;
                runmask = (b >= 64 ? ~0ull : ((1ull << b) - 1ull)) & ~((1ull << a) - 1ull);
                const GroupBox gb = masked_box(pme.x, pme.y, pme.z, lane >= a && lane < b);
                ok = enumerate_candidates(rec, nodes, tv, L, gb, s.walk, chunk) && ok;
            }
        }
    }
    if (!ok && lane == 0) atomicOr(err, (uint32_t)GVC_ERRFLAG_STACK_OVERFLOW);
}

// Segment bounds of the sorted records: seg[2 g] = first record of Gaussian g,
// seg[2 g + 1] = one past its last (both zero for a Gaussian without records).
__global__ void segment_kernel(const uint32_t* __restrict__ keys, uint64_t n_max, const uint32_t* __restrict__ n_dev,
                               uint32_t P, uint32_t* __restrict__ seg)
{
    const uint64_t n = (uint64_t)*n_dev < n_max ? (uint64_t)*n_dev : n_max;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t key = keys[k];
    if (key >= P) return;
    if (k == 0 || keys[k - 1] != key) seg[2 * (size_t)key] = (uint32_t)k;
    if (k + 1 == n || keys[k + 1] != key) seg[2 * (size_t)key + 1] = (uint32_t)(k + 1);
}

// Gradient sums of one Gaussian, in BVH order: 11 doubles and a pad (96 bytes).
constexpr int kSumStride = 12;

// Segments longer than this are summed by a whole wave each (reduce_long_kernel):
// in sparse volumes Gaussians carry hundreds of records, and a few hundred waves
// walking them with 4-lane teams would trail the launch.
constexpr uint32_t kLongSegment = 128;

// FP64 sums of records [b, e): lane j of a team of TEAM lanes adds records b+j,
// b+j+TEAM, ... in order, then a fixed butterfly over the team leaves the same
// sums in every lane.  The order depends only on the segment and TEAM.
template <int TEAM>
__device__ __forceinline__ void team_sums(uint32_t b, uint32_t e, int j, const uint32_t* __restrict__ sorted_vals,
                                          const PartialRecord* __restrict__ parts, double acc[11])
{
    for (int c = 0; c < 11; ++c) acc[c] = 0.0;
    // A = gid dv dw dmx | B = dmy dmz dc0 dc1 | C = dc2 dc3 dc4 dc5
    auto add = [&](const float4& A, const float4& B, const float4& C) {
        acc[0] += A.y; acc[1] += A.z; acc[2] += A.w; acc[3] += B.x; acc[4] += B.y;
        acc[5] += B.z; acc[6] += B.w; acc[7] += C.x; acc[8] += C.y; acc[9] += C.z; acc[10] += C.w;
    };
    uint32_t k = b + (uint32_t)j;
    for (; k + TEAM < e; k += 2 * TEAM) {
        const float4* q0 = reinterpret_cast<const float4*>(&parts[sorted_vals[k]]);
        const float4* q1 = reinterpret_cast<const float4*>(&parts[sorted_vals[k + TEAM]]);
        const float4 A0 = q0[0], B0 = q0[1], C0 = q0[2];
        const float4 A1 = q1[0], B1 = q1[1], C1 = q1[2];
        add(A0, B0, C0);
        add(A1, B1, C1);
    }
    if (k < e) {
        const float4* q0 = reinterpret_cast<const float4*>(&parts[sorted_vals[k]]);
        add(q0[0], q0[1], q0[2]);
    }
#pragma unroll
    for (int c = 0; c < 11; ++c) {
#pragma unroll
        for (int o = 1; o < TEAM; o <<= 1) acc[c] += __shfl_xor(acc[c], o, 64);
    }
}

__device__ __forceinline__ void store_sums(double* __restrict__ sums, int gi, const double m[11])
{
    double2* o = reinterpret_cast<double2*>(sums + (size_t)kSumStride * gi);
    o[0] = make_double2(m[0], m[1]);
    o[1] = make_double2(m[2], m[3]);
    o[2] = make_double2(m[4], m[5]);
    o[3] = make_double2(m[6], m[7]);
    o[4] = make_double2(m[8], m[9]);
    o[5] = make_double2(m[10], 0.0);
}

// Per Gaussian, in BVH order (neighbouring Gaussians have segments of similar
// length): segments of at most kLongSegment records are summed by a team of TEAM
// lanes (team_sums); a wave covers 64 consecutive Gaussians in TEAM passes (lane
// L first loads Gaussian L's segment bounds; a pass whose teams have no records
// is skipped), hands lane L the sums of Gaussian L and writes them to sums[] in
// BVH order (zeros for a Gaussian without records).  Longer segments are listed
// for reduce_long_kernel.
template <int TEAM>
__global__ __launch_bounds__(256) void reduce_sums_kernel(int P, const uint2* __restrict__ seg,
                                                          const uint32_t* __restrict__ sorted_vals,
                                                          const PartialRecord* __restrict__ parts,
                                                          double* __restrict__ sums, uint32_t* __restrict__ long_list)
{
    static_assert(TEAM >= 1 && TEAM <= 64 && (TEAM & (TEAM - 1)) == 0, "team size");
    constexpr int kPer = 64 / TEAM;  // Gaussians per pass
    const int lane = threadIdx.x & 63;
    const int g0 = ((int)xcd_block(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6)) * 64;
    if (g0 >= P) return;
    const int team = lane / TEAM, j = lane % TEAM;
    const int gi = g0 + lane;
    const bool valid = gi < P;
    const uint2 be = valid ? seg[gi] : make_uint2(0u, 0u);
    uint32_t my_b = be.x, my_e = be.y;
    const bool is_long = my_e - my_b > kLongSegment;
    const uint64_t longs = __ballot(is_long);
    if (longs != 0ull) {
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(long_list, (uint32_t)__popcll(longs));
        slot = (uint32_t)__shfl((int)slot, 0, 64);
        if (is_long) {
            long_list[1 + slot + (uint32_t)__popcll(longs & lanemask_lt())] = (uint32_t)gi;
            my_e = my_b;  // summed by reduce_long_kernel
        }
    }
    double mine[11];
    for (int c = 0; c < 11; ++c) mine[c] = 0.0;
    for (int p = 0; p < TEAM; ++p) {
        const int src = p * kPer + team;
        const uint32_t b = (uint32_t)__shfl((int)my_b, src, 64);
        const uint32_t e = (uint32_t)__shfl((int)my_e, src, 64);
        if (__ballot(e > b) == 0ull) continue;
        double acc[11];
        team_sums<TEAM>(b, e, j, sorted_vals, parts, acc);
#pragma unroll
        for (int c = 0; c < 11; ++c) {
            const double x = __shfl(acc[c], (lane % kPer) * TEAM, 64);
            if (lane / kPer == p) mine[c] = x;
        }
    }
    if (!valid || is_long) return;
    store_sums(sums, gi, mine);
}

// One wave per listed Gaussian (long segments): team_sums with 64 lanes.  A
// fixed grid takes the list with a grid-stride loop over its device-side count.
__global__ __launch_bounds__(256) void reduce_long_kernel(const uint2* __restrict__ seg,
                                                          const uint32_t* __restrict__ sorted_vals,
                                                          const PartialRecord* __restrict__ parts,
                                                          const uint32_t* __restrict__ long_list,
                                                          double* __restrict__ sums)
{
    const int lane = threadIdx.x & 63;
    const uint32_t n = long_list[0];
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t w = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); w < n; w += waves) {
        const int gi = (int)long_list[1 + w];
        double acc[11];
        const uint2 be = seg[gi];
        team_sums<64>(be.x, be.y, lane, sorted_vals, parts, acc);
        if (lane == 0) store_sums(sums, gi, acc);
    }
}

// Per Gaussian, in the caller's order (coalesced parameter reads and gradient
// writes; one 96-byte read of the BVH-order sums): gradients of value, weight
// and mean, and the covariance chain (conic gradient -> Sigma -> M -> scales,
// rotations) in double from the float inputs.  A zero conic gradient (no
// records) gives zero scale and rotation gradients without the chain.
// fp64_ratio > 0 (FP64 mode): Gaussians on the precise path have their conic
// gradient summed in the factor frame (covariance_chain repeats the test).
__global__ __launch_bounds__(256) void chain_kernel(int P, const uint32_t* __restrict__ inv_perm,
                                                    const GRecord* __restrict__ rec, int precise_mode,
                                                    const double* __restrict__ sums, const float* __restrict__ scales,
                                                    const float* __restrict__ rots, float kmod,
                                                    float* __restrict__ dmeans, float* __restrict__ dscales,
                                                    float* __restrict__ drots, float* __restrict__ dvalues,
                                                    float* __restrict__ dweights)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= P) return;
    const uint32_t ri = inv_perm[id];
    const double2* s2 = reinterpret_cast<const double2*>(sums + (size_t)kSumStride * ri);
    double m[12];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const double2 v = s2[c];
        m[2 * c] = v.x;
        m[2 * c + 1] = v.y;
    }
    dvalues[id] = (float)m[0];
    dweights[id] = (float)m[1];
    dmeans[3 * id + 0] = (float)m[2];
    dmeans[3 * id + 1] = (float)m[3];
    dmeans[3 * id + 2] = (float)m[4];
    double ds[3] = { 0.0, 0.0, 0.0 }, dq[4] = { 0.0, 0.0, 0.0, 0.0 };
    const bool any = m[5] != 0.0 || m[6] != 0.0 || m[7] != 0.0 || m[8] != 0.0 || m[9] != 0.0 || m[10] != 0.0;
    if (any) {
        // the precise path's conic gradient is in the factor frame (record bit set
        // by the preprocess, in FP64 mode only: the record is not read otherwise)
        const bool factor_frame = precise_mode && ((uint32_t)rec[ri].id & kPreciseBit) != 0u;
        covariance_chain(m + 5, scales + 3 * id, rots + 4 * id, kmod, ds, dq, factor_frame);
    }
    for (int c = 0; c < 3; ++c) dscales[3 * id + c] = (float)ds[c];
    for (int c = 0; c < 4; ++c) drots[4 * id + c] = (float)dq[c];
}

// Team size of the gradient reduction (lanes per Gaussian).
constexpr int kReduceTeam = 4;

// The reduction's segment bounds (2 words per Gaussian, zero: no records) and
// long-segment count, cleared in one launch.
__global__ void clear_reduce_kernel(uint32_t* __restrict__ seg, size_t n, uint32_t* __restrict__ long_list)
{
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0) long_list[0] = 0u;
    for (size_t i = i0; i < n; i += (size_t)gridDim.x * blockDim.x) seg[i] = 0u;
}

__global__ void zero_floats_kernel(float* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Sum of the per-group pair counts: each block adds its share to *total (integer
// atomics: the same total in any order).
__device__ __forceinline__ void add_group_pairs(const uint32_t* __restrict__ group_pairs, int NG,
                                                unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long part[256];
    unsigned long long s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < NG; i += gridDim.x * blockDim.x) s += group_pairs[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part[0]) atomicAdd(total, part[0]);
}

// out[0] = records, out[1] = pairs (out[1] cleared by the caller)
__global__ __launch_bounds__(256) void stats_kernel(const uint32_t* __restrict__ group_pairs, int NG,
                                                    const uint32_t* __restrict__ offsets,
                                                    unsigned long long* __restrict__ out)
{
    add_group_pairs(group_pairs, NG, out + 1);
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = offsets[NG];
}

// accum[0] += pairs, accum[1] += records, accum[2] |= error flags
// (stream-ordered, capturable).
__global__ __launch_bounds__(256) void accumulate_kernel(const uint32_t* __restrict__ group_pairs, int NG,
                                                         const uint32_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ err,
                                                         unsigned long long* __restrict__ accum)
{
    add_group_pairs(group_pairs, NG, accum);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        accum[1] += offsets[NG];
        accum[2] |= *err;
    }
}

inline unsigned reduce_blocks(int NG)
{
    const long long b = ((long long)NG + 1023) / 1024;
    return (unsigned)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Leaves of the segment tree over the leaf boxes: the power of two at or above NL.
inline uint32_t seg_leaves(int NL)
{
    uint32_t m = 1;
    while ((int64_t)m < (int64_t)NL) m <<= 1;
    return m;
}

// Grid of a persistent walk: as many blocks as fit on the device at once (a
// multiple of 8, one share per XCD), never more than the groups need.
unsigned walk_blocks(const void* kernel, int threads, size_t lds, int NG)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, lds) != hipSuccess || per <= 0 ||
        cus <= 0) {
        (void)hipGetLastError();
        per = 1;
        cus = 256;
    }
    unsigned nb = (unsigned)per * (unsigned)cus;
    nb = nb >= 8 ? nb & ~7u : nb;
    const unsigned need = blocks_for((size_t)NG, (unsigned)(threads / 64));
    return nb < need ? nb : need;
}

// Walk kernels take their LDS (per-wave stacks, rings, staging) dynamically:
// block_threads / 64 waves per block.
int allow_lds(const void* kernel, size_t bytes)
{
    if (bytes > 160u * 1024u) return set_error(GVC_ERR_INVALID_ARGUMENT, "block_threads too large for the LDS");
    if (bytes > 64u * 1024u)
        GVC_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return GVC_OK;
}

int run_points_of(const gvc_params* prm) { return prm->run_points ? (int)prm->run_points : 64; }
int lane_policy_of(const gvc_params* prm)
{
    return (prm->flags & GVC_FLAG_LANE_ALWAYS) ? 2 : ((prm->flags & GVC_FLAG_LANE_NEVER) ? 1 : 0);
}

inline int key_bits_for(uint64_t maxv)
{
    int b = 0;
    while (b < 64 && (maxv >> b) != 0) ++b;
    return b;
}

// ---------------------------------------------------------------------------
// workspace layouts
// ---------------------------------------------------------------------------
// Host-side totals of a batch and its device table.
struct Batch {
    int V, P, S, NG, NL, NI, L;
    const void* table;
};

Batch single_batch(int P, int S, int L, const void* table)
{
    Batch b;
    b.V = 1; b.P = P; b.S = S; b.L = L;
    b.NG = (S + 63) / 64;
    b.NL = P > 0 ? (P + L - 1) / L : 0;
    b.NI = b.NL > 1 ? b.NL - 1 : 0;
    b.table = table;
    return b;
}

struct SampleSetLayout {
    SampleSetHeader* header;
    int* table;  // the set's own one-volume table
    float4* pts;
    uint32_t *keys, *idx, *keys_alt, *idx_alt, *scratch;
    // V: the volumes of a batch (the sort scratch, last, is sized for them)
    SampleSetLayout(Arena& a, int S, int NG, int V = 1)
    {
        header = a.take<SampleSetHeader>(1);
        table = a.take<int>(batch_table_ints(1));
        pts = a.take<float4>((size_t)NG * 64);
        keys = a.take<uint32_t>((size_t)S);
        idx = a.take<uint32_t>((size_t)S);
        keys_alt = a.take<uint32_t>((size_t)S);
        idx_alt = a.take<uint32_t>((size_t)S);
        scratch = a.take<uint32_t>(radix_scratch_elems((size_t)S, V));
    }
};

struct ForwardLayout {
    int* table;  // the state's own one-volume table
    GRecord* rec;
    float4* fac_lo;  // FP64 mode: lo parts of the precise-path factors, 2 per Gaussian (caller order)
    uint32_t *codes, *codes_alt;
    uint64_t* leaf_codes;
    uint32_t *idx, *idx_alt, *sort_scratch;
    uint32_t* inv_perm;  // caller's Gaussian index -> record (BVH) index
    BNode* nodes;
    Box6* seg;  // segment tree over the leaf boxes, 2 M entries (leaf j at M + j)
    float* root_box;  // 6 per volume
    uint32_t *group_records, *group_offsets, *group_pairs, *scan_scratch, *err;
    uint32_t* queues;  // walk counters, kQWords words (kQ* offsets)
    unsigned long long *stats, *ext_sum;
    uint32_t* group_head;
    uint32_t* heavy;        // [0] = count, [1..] = heavy groups (walked first)
    uint8_t* group_flag;    // 1 = heavy
    float2* pvw;            // per sorted point (group g, lane j at 64 g + j): the forward's (value, weight)
    unsigned char* log;
    uint32_t log_cap;
    int NL, NG;
    ForwardLayout(Arena& a, const Batch& b, unsigned flags, unsigned chunk_slots)
    {
        const int P = b.P;
        NL = b.NL;
        NG = b.NG;
        table = a.take<int>(batch_table_ints(1));
        rec = a.take<GRecord>((size_t)P);
        fac_lo = (flags & GVC_FLAG_FP64_COND) ? a.take<float4>(2 * (size_t)P) : nullptr;
        codes = a.take<uint32_t>((size_t)P);
        codes_alt = a.take<uint32_t>((size_t)P);
        idx = a.take<uint32_t>((size_t)P);
        idx_alt = a.take<uint32_t>((size_t)P);
        inv_perm = a.take<uint32_t>((size_t)P);
        sort_scratch = a.take<uint32_t>(radix_scratch_elems((size_t)P, b.V));
        leaf_codes = a.take<uint64_t>((size_t)NL);
        nodes = a.take<BNode>((size_t)(b.NI > 0 ? b.NI : 1));
        seg = a.take<Box6>(2 * (size_t)seg_leaves(NL));
        root_box = a.take<float>(6 * (size_t)b.V + 2);
        group_records = a.take<uint32_t>((size_t)NG + 1);
        group_offsets = a.take<uint32_t>((size_t)NG + 1);
        group_pairs = a.take<uint32_t>((size_t)NG + 1);
        scan_scratch = a.take<uint32_t>(scan_scratch_elems((size_t)NG) + 1);
        err = a.take<uint32_t>(4);
        queues = a.take<uint32_t>(kQWords);
        stats = a.take<unsigned long long>(4);
        ext_sum = a.take<unsigned long long>(4 * (size_t)b.V);
        group_head = a.take<uint32_t>((size_t)NG + 1);
        heavy = a.take<uint32_t>((size_t)NG + 1);
        group_flag = reinterpret_cast<uint8_t*>(a.take<uint32_t>(((size_t)NG + 3) / 4 + 1));
        pvw = a.take<float2>((size_t)NG * 64 + 1);
        const uint64_t cap = (flags & GVC_FLAG_CANDIDATE_LOG)
                                 ? (uint64_t)(chunk_slots ? chunk_slots : 16u) * (uint64_t)(NG > 0 ? NG : 1)
                                 : 1ull;
        log_cap = (uint32_t)(cap < 0xfff00000ull ? cap : 0xfff00000ull);
        log = reinterpret_cast<unsigned char*>(a.take<uint4>((size_t)log_cap * (kSlotBytes / 16)));
    }
};

struct BackwardLayout {
    PartialRecord* parts;
    uint32_t *keys, *vals, *keys_alt, *vals_alt, *scratch;
    uint2* seg;    // per Gaussian (BVH order): its records [x, y) in the sorted order
    double* sums;  // kSumStride per Gaussian, BVH order
    uint32_t* long_list;  // [0] = count, [1..] = Gaussians with long segments
    BackwardLayout(Arena& a, int P, uint64_t cap)
    {
        const size_t c = (size_t)(cap > 0 ? cap : 1);
        parts = a.take<PartialRecord>(c);
        keys = a.take<uint32_t>(c);
        vals = a.take<uint32_t>(c);
        keys_alt = a.take<uint32_t>(c);
        vals_alt = a.take<uint32_t>(c);
        scratch = a.take<uint32_t>(radix_scratch_elems(c));
        seg = a.take<uint2>((size_t)(P > 0 ? P : 1));
        sums = a.take<double>((size_t)kSumStride * (size_t)(P > 0 ? P : 1));
        long_list = a.take<uint32_t>((size_t)P + 1);
    }
};

int check_params(const gvc_params* prm)
{
    if (!prm) return set_error(GVC_ERR_INVALID_ARGUMENT, "params is NULL");
    if (prm->leaf_size < 1 || prm->leaf_size > 64)
        return set_error(GVC_ERR_INVALID_ARGUMENT, "leaf_size must be in [1, 64]");
    if (prm->block_threads < 64 || prm->block_threads > 1024 || (prm->block_threads % 64) != 0)
        return set_error(GVC_ERR_INVALID_ARGUMENT, "block_threads must be a multiple of 64 in [64, 1024]");
    if (prm->run_points != 0 && prm->run_points != 4 && prm->run_points != 8 && prm->run_points != 16 &&
        prm->run_points != 32 && prm->run_points != 64)
        return set_error(GVC_ERR_INVALID_ARGUMENT, "run_points must be 0, 4, 8, 16, 32 or 64");
    if ((prm->flags & GVC_FLAG_LANE_NEVER) && (prm->flags & GVC_FLAG_LANE_ALWAYS))
        return set_error(GVC_ERR_INVALID_ARGUMENT, "GVC_FLAG_LANE_NEVER and GVC_FLAG_LANE_ALWAYS are exclusive");
    if (!(prm->fp64_scale_ratio == 0.f || (prm->fp64_scale_ratio >= 1.f && prm->fp64_scale_ratio <= 1e18f)))
        return set_error(GVC_ERR_INVALID_ARGUMENT, "fp64_scale_ratio must be 0 (default) or in [1, 1e18]");
    if (!(prm->factor_scale_ratio == 0.f || (prm->factor_scale_ratio >= 1.f && prm->factor_scale_ratio <= 1e18f)))
        return set_error(GVC_ERR_INVALID_ARGUMENT, "factor_scale_ratio must be 0 (default) or in [1, 1e18]");
    return GVC_OK;
}

// The scale ratio above which a Gaussian takes the FP64 path (0 = the default).
float fp64_ratio_of(const gvc_params* prm) { return prm->fp64_scale_ratio > 0.f ? prm->fp64_scale_ratio : kCondRatio; }
// The scale ratio above which a Gaussian below the FP64 threshold takes the
// float-float factor (0 = the default).
float factor_ratio_of(const gvc_params* prm)
{
    return prm->factor_scale_ratio > 0.f ? prm->factor_scale_ratio : kFactorRatio;
}

// The table a call uses: the batch's, or for one volume the buffer's own.
const void* table_of(const Batch& b, const int* own) { return b.V == 1 && !b.table ? own : b.table; }

size_t sample_set_bytes_impl(int S, int NG, int V = 1)
{
    Arena a(nullptr, 0);
    SampleSetLayout l(a, S, NG, V);
    return a.off + 256;
}

int sample_set_build_impl(const Batch& bt, const float* samples, void* set, size_t bytes, hipStream_t st)
{
    const int S = bt.S, NG = bt.NG;
    Arena a(set, bytes);
    SampleSetLayout l(a, S, NG, bt.V);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "sample set buffer too small");
    if (!bt.table && S == 0) {
        single_table_kernel<<<1, 64, 0, st>>>(l.table, 0, S, 1);
        GVC_LAUNCH_CHECK("sample_table");
    }
    const BatchView bv = batch_view(table_of(bt, l.table), bt.V);
    if (S > 0) {
        // the codes kernel also writes the one-volume table of a set without one
        sample_codes_kernel<<<blocks_for(S, 256), 256, 0, st>>>(samples, S, l.keys, l.idx,
                                                               bt.table ? nullptr : l.table);
        GVC_LAUNCH_CHECK("sample_codes");
        // Morton order within each volume: a batch's samples are volume-major, and
        // each volume's range is sorted on its own (segmented sort)
        bool alt = false;
        GVC_TRY(radix_sort_u32(l.keys, l.idx, l.keys_alt, l.idx_alt, (size_t)S, kSampleKeyBits, l.scratch, &alt, st,
                               nullptr, false, bt.V > 1 ? bv.S_off : nullptr, bt.V));
        const uint32_t* perm = alt ? l.idx_alt : l.idx;
        sample_gather_kernel<<<blocks_for((size_t)NG * 64, 256), 256, 0, st>>>(samples, bv, NG * 64, perm, l.pts,
                                                                               l.header, S);
        GVC_LAUNCH_CHECK("sample_gather");
    } else {
        sample_header_kernel<<<1, 1, 0, st>>>(l.header, S, NG);
        GVC_LAUNCH_CHECK("sample_header");
    }
    return GVC_OK;
}

const float4* sample_set_points_impl(const void* sample_set, int S, int NG)
{
    Arena a(const_cast<void*>(sample_set), sample_set_bytes_impl(S, NG));
    SampleSetLayout l(a, S, NG);
    return l.pts;
}

size_t forward_state_bytes_impl(const Batch& b, const gvc_params* prm)
{
    if (check_params(prm) != GVC_OK) return 0;
    Arena a(nullptr, 0);
    ForwardLayout l(a, b, prm->flags, prm->chunk_slots);
    return a.off + 256;
}

int mixture_forward_impl(const Batch& bt, const float* means, const float* scales, const float* rots,
                         const float* values, const float* weights, float k, const void* sample_set,
                         const gvc_params* prm, float* out_cells, float* out_weights, float* conics, void* state,
                         size_t state_bytes, hipStream_t st)
{
    GVC_TRY(check_params(prm));
    const int P = bt.P, S = bt.S, L = prm->leaf_size;
    if (P < 0 || S < 0) return set_error(GVC_ERR_INVALID_ARGUMENT, "negative size");
    if (L != bt.L) return set_error(GVC_ERR_INVALID_ARGUMENT, "params.leaf_size differs from the batch's");
    Arena a(state, state_bytes);
    ForwardLayout l(a, bt, prm->flags, prm->chunk_slots);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "forward state buffer too small");
    const float4* pts = sample_set_points_impl(sample_set, S, bt.NG);
    const int NG = l.NG, NL = l.NL;
    const unsigned stages = prm->stages ? (unsigned)prm->stages : 7u;

    if (stages & GVC_STAGE_BUILD) {
        if (P == 0) {   // else gaussian_codes_kernel does both
            GVC_TRY(fill_u32(l.err, 0u, 4, st));
            if (!bt.table) {
                single_table_kernel<<<1, 64, 0, st>>>(l.table, P, S, L);
                GVC_LAUNCH_CHECK("table");
            }
        }
    }
    const BatchView bv = batch_view(table_of(bt, l.table), bt.V);
    if (P == 0) {
        // DGR/rasterize_points.cu:51: the reference skips the kernels, leaving zeros.
        if (S > 0) {
            zero_floats_kernel<<<blocks_for(S, 256), 256, 0, st>>>(out_cells, S);
            zero_floats_kernel<<<blocks_for(S, 256), 256, 0, st>>>(out_weights, S);
            GVC_LAUNCH_CHECK("zero");
        }
        GVC_TRY(fill_u32(l.group_records, 0u, (size_t)NG + 1, st));
        GVC_TRY(fill_u32(l.group_pairs, 0u, (size_t)NG + 1, st));
        GVC_TRY(exclusive_scan_u32(l.group_records, l.group_offsets, (size_t)NG, true, l.scan_scratch, st));
        return GVC_OK;
    }
    const float fp64_ratio = (prm->flags & GVC_FLAG_FP64_COND) ? fp64_ratio_of(prm) : 0.f;
    const float factor_ratio = (prm->flags & GVC_FLAG_FP64_COND) ? factor_ratio_of(prm) : 0.f;
    if (stages & GVC_STAGE_BUILD) {
        // 1. Morton order of the Gaussians (within each volume)
        gaussian_codes_kernel<<<blocks_for(P, 256), 256, 0, st>>>(means, P, l.codes, l.err,
                                                                  bt.table ? nullptr : l.table, S, L);
        GVC_LAUNCH_CHECK("gaussian_codes");
        // (a batch's Gaussians are volume-major: each volume's range is sorted on
        // its own)
        bool alt = false;
        // (the last pass also writes inv_perm: each caller index's BVH position)
        GVC_TRY(radix_sort_u32(l.codes, l.idx, l.codes_alt, l.idx_alt, (size_t)P, kGaussianCodeBits, l.sort_scratch,
                               &alt, st, nullptr, true, bt.V > 1 ? bv.P_off : nullptr, bt.V, l.inv_perm));
        const uint32_t* scodes = alt ? l.codes_alt : l.codes;
        // 2. records
        record_kernel<<<blocks_for(P, 256), 256, 0, st>>>(P, l.inv_perm, means, scales, rots, values, weights, k,
                                                          fp64_ratio, factor_ratio, l.rec, l.fac_lo, conics);
        GVC_LAUNCH_CHECK("record");
        // 3. trees, one per volume (with the volumes' extent sums, one launch)
        extent_leaf_codes_kernel<<<(unsigned)bt.V + blocks_for(NL, 256), 256, 0, st>>>(l.rec, scodes, bv, NL, L,
                                                                                      l.ext_sum, l.leaf_codes);
        GVC_LAUNCH_CHECK("extent_leaf_codes");
        if (bt.NI > 0) {
            karras_kernel<<<blocks_for(bt.NI, 256), 256, 0, st>>>(l.leaf_codes, bv, bt.NI, l.nodes);
            GVC_LAUNCH_CHECK("karras");
        }
        // 4. boxes: segment tree over the leaf boxes (kSegLevels levels per
        // launch), then both child boxes of every node by range queries
        const uint32_t M = seg_leaves(NL);
        if (NL > 0) {
            uint32_t lvl = M;   // first node of the level being reduced
            int total = 0;
            while ((1u << total) < M) ++total;   // levels above the leaves
            bool first = true;
            while (true) {
                const int lv = total < kSegLevels ? total : kSegLevels;
                seg_level_kernel<<<blocks_for(lvl, kSegBlock), kSegBlock, 0, st>>>(first ? l.rec : nullptr, bv, NL, L,
                                                                                  l.seg, lvl, lvl, lv);
                GVC_LAUNCH_CHECK("seg_level");
                first = false;
                total -= lv;
                lvl >>= lv;
                if (total <= 0) break;
            }
        }
        if (bt.NI > 0) {
            node_box_kernel<<<blocks_for(bt.NI, 256), 256, 0, st>>>(l.seg, M, bv, bt.NI, l.nodes, l.root_box);
            GVC_LAUNCH_CHECK("node_box");
        }
        leaf_root_kernel<<<blocks_for(bt.V, 64), 64, 0, st>>>(l.seg + M, bv, l.root_box, l.heavy);
        GVC_LAUNCH_CHECK("root_box");
        // 4b. heavy groups, walked first (the kernel also clears the walk's queue
        // words)
        if (NG > 0) {
            group_cost_kernel<<<blocks_for((size_t)NG, 4), 256, 0, st>>>(pts, NG, bv, l.root_box, l.ext_sum, l.heavy,
                                                                        l.group_flag, l.queues);
            GVC_LAUNCH_CHECK("group_cost");
        }
    }
    // 5. walk
    if ((stages & GVC_STAGE_WALK) && NG > 0) {
        const int T = prm->block_threads;
        if (!(stages & GVC_STAGE_BUILD)) {   // else group_cost_kernel cleared them
            GVC_TRY(fill_u32(l.queues + kQFwd, 0u, kQueues * kLine, st));
            GVC_TRY(fill_u32(l.queues + kQLog, 0u, kQWords - kQLog, st));
        }
        const int log_on = (prm->flags & GVC_FLAG_CANDIDATE_LOG) ? 1 : 0;
        const int rp = run_points_of(prm);
        const int lp = lane_policy_of(prm);
        const size_t lds = (size_t)(T / 64) * sizeof(FwdLds);
        auto launch = [&](auto kahan, auto precise) -> int {
            constexpr bool K = decltype(kahan)::value, PR = decltype(precise)::value;
            const void* kf = reinterpret_cast<const void*>(&forward_kernel<K, PR>);
            GVC_TRY(allow_lds(kf, lds));
            const unsigned nb = walk_blocks(kf, T, lds, NG);
            forward_kernel<K, PR><<<nb, T, lds, st>>>(l.rec, l.fac_lo, l.nodes, l.root_box, bv, P, L, pts, NG,
                                                      out_cells, out_weights, l.pvw, l.group_records, l.group_pairs,
                                                      l.err,
                                                      l.queues + kQFwd, l.ext_sum, log_on, l.log, l.log_cap,
                                                      l.queues + kQLog, l.group_head, rp, lp, l.heavy,
                                                      l.queues + kQHeavyF, l.group_flag);
            return GVC_OK;
        };
        using T1 = std::true_type;
        using F0 = std::false_type;
        const bool kahan = (prm->flags & GVC_FLAG_KAHAN) != 0, precise = l.fac_lo != nullptr;
        GVC_TRY(kahan ? (precise ? launch(T1{}, T1{}) : launch(T1{}, F0{}))
                      : (precise ? launch(F0{}, T1{}) : launch(F0{}, F0{})));
        GVC_LAUNCH_CHECK("forward");
    }
    // 6. record offsets for the backward
    if (stages & GVC_STAGE_COUNTS)
        GVC_TRY(exclusive_scan_u32(l.group_records, l.group_offsets, (size_t)NG, true, l.scan_scratch, st));
    return GVC_OK;
}

int forward_stats_impl(const Batch& bt, const void* state, const gvc_params* prm, uint64_t* n_records,
                       uint64_t* n_pairs, uint32_t* error_flags, uint64_t* chunk_slots_used, uint64_t* lane_groups,
                       hipStream_t st)
{
    GVC_TRY(check_params(prm));
    Arena a(const_cast<void*>(state), (size_t)-1);
    ForwardLayout l(a, bt, prm->flags, prm->chunk_slots);
    GVC_TRY(fill_u32(reinterpret_cast<uint32_t*>(l.stats), 0u, 4, st));
    stats_kernel<<<reduce_blocks(l.NG), 256, 0, st>>>(l.group_pairs, l.NG, l.group_offsets, l.stats);
    GVC_LAUNCH_CHECK("stats");
    unsigned long long h[2] = { 0, 0 };
    uint32_t e = 0, used[2] = { 0, 0 };
    GVC_HIP(hipMemcpyAsync(h, l.stats, sizeof(h), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipMemcpyAsync(&e, l.err, sizeof(e), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipMemcpyAsync(used, l.queues + kQLog, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipMemcpyAsync(used + 1, l.queues + kQLane, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipStreamSynchronize(st));
    if (n_records) *n_records = h[0];
    if (n_pairs) *n_pairs = h[1];
    if (error_flags) *error_flags = e;
    if (chunk_slots_used) *chunk_slots_used = used[0];
    if (lane_groups) *lane_groups = used[1];
    return GVC_OK;
}

int forward_pair_counts_impl(const Batch& bt, const void* state, const gvc_params* prm, const void* sample_set,
                             uint32_t* sample_counts, uint32_t* gaussian_counts, hipStream_t st)
{
    GVC_TRY(check_params(prm));
    if (!(prm->flags & GVC_FLAG_CANDIDATE_LOG))
        return set_error(GVC_ERR_UNSUPPORTED, "pair counts replay the candidate log: the forward needs GVC_FLAG_CANDIDATE_LOG");
    Arena a(const_cast<void*>(state), (size_t)-1);
    ForwardLayout l(a, bt, prm->flags, prm->chunk_slots);
    if (bt.S > 0) GVC_TRY(fill_u32(sample_counts, 0u, (size_t)bt.S, st));
    if (bt.P > 0) GVC_TRY(fill_u32(gaussian_counts, 0u, (size_t)bt.P, st));
    if (bt.P == 0 || bt.NG == 0) return GVC_OK;
    const float4* pts = sample_set_points_impl(sample_set, bt.S, bt.NG);
    const BatchView bv = batch_view(table_of(bt, l.table), bt.V);
    const unsigned nb = blocks_for((size_t)bt.NG, 4);
    if (l.fac_lo)
        pair_count_kernel<true><<<nb, 256, 0, st>>>(l.rec, l.fac_lo, l.nodes, l.root_box, bv, bt.L, pts, bt.NG, l.log,
                                                   l.group_head, sample_counts, gaussian_counts, l.err);
    else
        pair_count_kernel<false><<<nb, 256, 0, st>>>(l.rec, l.fac_lo, l.nodes, l.root_box, bv, bt.L, pts, bt.NG, l.log,
                                                    l.group_head, sample_counts, gaussian_counts, l.err);
    GVC_LAUNCH_CHECK("pair_count");
    return GVC_OK;
}

int forward_accumulate_impl(const Batch& bt, const void* state, const gvc_params* prm, unsigned long long* accum,
                            hipStream_t st)
{
    GVC_TRY(check_params(prm));
    Arena a(const_cast<void*>(state), (size_t)-1);
    ForwardLayout l(a, bt, prm->flags, prm->chunk_slots);
    accumulate_kernel<<<reduce_blocks(l.NG), 256, 0, st>>>(l.group_pairs, l.NG, l.group_offsets, l.err, accum);
    GVC_LAUNCH_CHECK("accumulate");
    return GVC_OK;
}

int mixture_backward_impl(const Batch& bt, const float* scales, const float* rots, float k, const void* sample_set,
                          const gvc_params* prm, const float* out_cells, const float* out_weights,
                          const float* gcells, const float* gweights, void* state, size_t state_bytes,
                          uint64_t capacity, void* ws, size_t ws_bytes, float* dmeans, float* dscales, float* drots,
                          float* dvalues, float* dweights, hipStream_t st)
{
    GVC_TRY(check_params(prm));
    const int P = bt.P, S = bt.S, L = prm->leaf_size;
    if (P <= 0) return GVC_OK;
    if (L != bt.L) return set_error(GVC_ERR_INVALID_ARGUMENT, "params.leaf_size differs from the batch's");
    if (capacity > 0xffffffffull) return set_error(GVC_ERR_INVALID_ARGUMENT, "record capacity above 2^32");
    Arena a(state, state_bytes);
    ForwardLayout l(a, bt, prm->flags, prm->chunk_slots);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "forward state buffer too small");
    Arena ba(ws, ws_bytes);
    BackwardLayout bl(ba, P, capacity);
    if (!ba.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "backward workspace too small");
    const float4* pts = sample_set_points_impl(sample_set, S, bt.NG);
    const BatchView bv = batch_view(table_of(bt, l.table), bt.V);
    const int NG = l.NG;
    const unsigned stages = prm->stages ? (unsigned)prm->stages : 3u;
    if ((stages & GVC_STAGE_WALK) && NG > 0 && capacity > 0) {
        const int T = prm->block_threads;
        GVC_TRY(fill_u32(l.queues + kQBwd, 0u, kQueues * kLine, st));
        GVC_TRY(fill_u32(l.queues + kQHeavyB, 0u, kLine, st));
        const int rp = run_points_of(prm);
        const int lp = lane_policy_of(prm);
        auto launch = [&](auto replay, auto precise) -> int {
            constexpr bool RE = decltype(replay)::value, PR = decltype(precise)::value;
            const void* kf = reinterpret_cast<const void*>(&backward_kernel<RE, PR>);
            const size_t lds = (size_t)(T / 64) * sizeof(BwdLds<RE>);
            GVC_TRY(allow_lds(kf, lds));
            const unsigned nb = walk_blocks(kf, T, lds, NG);
            backward_kernel<RE, PR><<<nb, T, lds, st>>>(l.rec, l.fac_lo, l.nodes, l.root_box, bv, P, L, pts, NG,
                                                        l.pvw, gcells, gweights, l.group_offsets,
                                                        capacity, bl.parts, bl.keys, l.err, l.queues + kQBwd,
                                                        l.ext_sum, l.log, l.group_head, rp, lp, l.heavy,
                                                        l.queues + kQHeavyB, l.group_flag);
            return GVC_OK;
        };
        using T1 = std::true_type;
        using F0 = std::false_type;
        const bool replay = (prm->flags & GVC_FLAG_CANDIDATE_LOG) != 0, precise = l.fac_lo != nullptr;
        GVC_TRY(replay ? (precise ? launch(T1{}, T1{}) : launch(T1{}, F0{}))
                       : (precise ? launch(F0{}, T1{}) : launch(F0{}, F0{})));
        GVC_LAUNCH_CHECK("backward");
    }
    if (!(stages & GVC_STAGE_REDUCE)) return GVC_OK;
    {
        const unsigned nb = blocks_for(2 * (size_t)P, 256);
        clear_reduce_kernel<<<nb < 4096u ? (nb > 0u ? nb : 1u) : 4096u, 256, 0, st>>>(
            reinterpret_cast<uint32_t*>(bl.seg), 2 * (size_t)P, bl.long_list);
        GVC_LAUNCH_CHECK("clear_reduce");
    }
    const uint32_t* skeys = bl.keys;
    const uint32_t* svals = bl.vals;
    if (NG > 0 && capacity > 0) {
        bool alt = false;
        // the records emitted: group_offsets[NG] (capped at the capacity)
        // values: the record index (generated by the first pass)
        GVC_TRY(radix_sort_u32(bl.keys, bl.vals, bl.keys_alt, bl.vals_alt, (size_t)capacity,
                               key_bits_for((uint64_t)P), bl.scratch, &alt, st, l.group_offsets + NG, true));
        skeys = alt ? bl.keys_alt : bl.keys;
        svals = alt ? bl.vals_alt : bl.vals;
        segment_kernel<<<blocks_for(capacity, 256), 256, 0, st>>>(skeys, capacity, l.group_offsets + NG, (uint32_t)P,
                                                                  reinterpret_cast<uint32_t*>(bl.seg));
        GVC_LAUNCH_CHECK("segment");
    }
    reduce_sums_kernel<kReduceTeam><<<blocks_for(P, 256), 256, 0, st>>>(P, bl.seg, svals, bl.parts,
                                                                        bl.sums, bl.long_list);
    GVC_LAUNCH_CHECK("reduce_sums");
    reduce_long_kernel<<<1024, 256, 0, st>>>(bl.seg, svals, bl.parts, bl.long_list, bl.sums);
    GVC_LAUNCH_CHECK("reduce_long");
    chain_kernel<<<blocks_for(P, 256), 256, 0, st>>>(P, l.inv_perm, l.rec, l.fac_lo != nullptr ? 1 : 0, bl.sums,
                                                     scales, rots, k, dmeans, dscales, drots, dvalues, dweights);
    GVC_LAUNCH_CHECK("chain");
    return GVC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// host entry points (called from capi.cpp)
// ---------------------------------------------------------------------------
int preprocess_launch(int P, const float* means, const float* scales, const float* rots, const float* weights,
                      float k, unsigned flags, float* conics, float* boxes, hipStream_t st)
{
    if (P <= 0) return GVC_OK;
    preprocess_kernel<<<blocks_for(P, 256), 256, 0, st>>>(P, means, scales, rots, weights, k,
                                                          (flags & GVC_FLAG_FP64_COND) ? kCondRatio : 0.f, 0.f,
                                                          conics, boxes, nullptr);
    GVC_LAUNCH_CHECK("preprocess");
    return GVC_OK;
}

int preprocess_factors_launch(int P, const float* means, const float* scales, const float* rots, const float* weights,
                              float k, const gvc_params* prm, float* conics, float* boxes, float* factors,
                              hipStream_t st)
{
    GVC_TRY(check_params(prm));
    if (P <= 0) return GVC_OK;
    const bool on = (prm->flags & GVC_FLAG_FP64_COND) != 0;
    preprocess_kernel<<<blocks_for(P, 256), 256, 0, st>>>(P, means, scales, rots, weights, k,
                                                          on ? fp64_ratio_of(prm) : 0.f,
                                                          on ? factor_ratio_of(prm) : 0.f, conics, boxes, factors);
    GVC_LAUNCH_CHECK("preprocess");
    return GVC_OK;
}

int morton_launch(int n, const float* xyz, int clamp, uint64_t* codes, hipStream_t st)
{
    if (n <= 0) return GVC_OK;
    morton63_kernel<<<blocks_for(n, 256), 256, 0, st>>>(xyz, n, clamp, codes);
    GVC_LAUNCH_CHECK("morton");
    return GVC_OK;
}

size_t sample_set_bytes(int S) { return sample_set_bytes_impl(S, (S + 63) / 64); }

int sample_set_build(int S, const float* samples, void* set, size_t bytes, hipStream_t st)
{
    return sample_set_build_impl(single_batch(0, S, 1, nullptr), samples, set, bytes, st);
}

const float4* sample_set_points(const void* sample_set, int S)
{
    return sample_set_points_impl(sample_set, S, (S + 63) / 64);
}

size_t forward_state_bytes(int P, int S, const gvc_params* prm)
{
    if (check_params(prm) != GVC_OK) return 0;
    return forward_state_bytes_impl(single_batch(P, S, prm->leaf_size, nullptr), prm);
}

int mixture_forward(int P, const float* means, const float* scales, const float* rots, const float* values,
                    const float* weights, float k, const void* sample_set, int S, const gvc_params* prm,
                    float* out_cells, float* out_weights, float* conics, void* state, size_t state_bytes,
                    hipStream_t st)
{
    GVC_TRY(check_params(prm));
    return mixture_forward_impl(single_batch(P, S, prm->leaf_size, nullptr), means, scales, rots, values, weights, k,
                                sample_set, prm, out_cells, out_weights, conics, state, state_bytes, st);
}

int forward_stats(const void* state, int P, int S, const gvc_params* prm, uint64_t* n_records, uint64_t* n_pairs,
                  uint32_t* error_flags, uint64_t* chunk_slots_used, uint64_t* lane_groups,
                       hipStream_t st)
{
    GVC_TRY(check_params(prm));
    return forward_stats_impl(single_batch(P, S, prm->leaf_size, nullptr), state, prm, n_records, n_pairs,
                              error_flags, chunk_slots_used, lane_groups, st);
}

int forward_accumulate(const void* state, int P, int S, const gvc_params* prm, unsigned long long* accum,
                       hipStream_t st)
{
    GVC_TRY(check_params(prm));
    return forward_accumulate_impl(single_batch(P, S, prm->leaf_size, nullptr), state, prm, accum, st);
}

int forward_pair_counts(const void* state, int P, int S, const gvc_params* prm, const void* sample_set,
                        uint32_t* sample_counts, uint32_t* gaussian_counts, hipStream_t st)
{
    GVC_TRY(check_params(prm));
    if (P < 0 || S < 0) return set_error(GVC_ERR_INVALID_ARGUMENT, "negative size");
    return forward_pair_counts_impl(single_batch(P, S, prm->leaf_size, nullptr), state, prm, sample_set,
                                    sample_counts, gaussian_counts, st);
}

int forward_tree(const void* state, int P, int S, const gvc_params* prm, uint32_t* order, uint32_t* keys,
                 int32_t* links, float* boxes, hipStream_t st)
{
    GVC_TRY(check_params(prm));
    if (P < 0 || S < 0) return set_error(GVC_ERR_INVALID_ARGUMENT, "negative size");
    const Batch bt = single_batch(P, S, prm->leaf_size, nullptr);
    Arena a(const_cast<void*>(state), forward_state_bytes_impl(bt, prm));
    ForwardLayout l(a, bt, prm->flags, prm->chunk_slots);
    const int n = P > bt.NI ? P : bt.NI;
    if (n > 0) {
        tree_export_kernel<<<blocks_for((size_t)n, 256), 256, 0, st>>>(l.rec, P, l.nodes, bt.NI, order, keys, links,
                                                                       boxes);
        GVC_LAUNCH_CHECK("tree_export");
    }
    return GVC_OK;
}

size_t backward_bytes(int P, uint64_t cap)
{
    Arena a(nullptr, 0);
    BackwardLayout l(a, P, cap);
    return a.off + 256;
}

int mixture_backward(int P, const float* means, const float* scales, const float* rots, const float* values,
                     const float* weights, float k, const void* sample_set, int S, const gvc_params* prm,
                     const float* out_cells, const float* out_weights, const float* gcells, const float* gweights,
                     void* state, size_t state_bytes, uint64_t capacity, void* ws, size_t ws_bytes, float* dmeans,
                     float* dscales, float* drots, float* dvalues, float* dweights, hipStream_t st)
{
    (void)means; (void)values; (void)weights;
    GVC_TRY(check_params(prm));
    return mixture_backward_impl(single_batch(P, S, prm->leaf_size, nullptr), scales, rots, k, sample_set, prm,
                                 out_cells, out_weights, gcells, gweights, state, state_bytes, capacity, ws, ws_bytes,
                                 dmeans, dscales, drots, dvalues, dweights, st);
}

// ---- batches ---------------------------------------------------------------
size_t batch_table_bytes(int V) { return V > 0 ? batch_table_ints(V) * sizeof(int) + 256 : 0; }

int batch_describe(int V, const int* P_counts, const int* S_counts, int L, void* table, size_t bytes,
                   gvc_batch* out, hipStream_t st)
{
    if (V < 1) return set_error(GVC_ERR_INVALID_ARGUMENT, "a batch needs at least one volume");
    if (L < 1 || L > 64) return set_error(GVC_ERR_INVALID_ARGUMENT, "leaf_size must be in [1, 64]");
    if (bytes < batch_table_ints(V) * sizeof(int)) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "batch table too small");
    std::vector<int> h(batch_table_ints(V), 0);
    h[0] = V;
    const size_t n = (size_t)V + 1;
    int* Po = h.data() + 4;
    int* So = Po + n;
    int* Go = So + n;
    int* Lo = Go + n;
    int* No = Lo + n;
    long long p = 0, s = 0, g = 0, l = 0, ni = 0;
    for (int v = 0; v < V; ++v) {
        if (P_counts[v] < 1 || S_counts[v] < 0)
            return set_error(GVC_ERR_INVALID_ARGUMENT, "every volume needs >= 1 Gaussian and >= 0 samples");
        Po[v] = (int)p; So[v] = (int)s; Go[v] = (int)g; Lo[v] = (int)l; No[v] = (int)ni;
        const long long nl = (P_counts[v] + L - 1) / L;
        p += P_counts[v]; s += S_counts[v]; g += (S_counts[v] + 63) / 64; l += nl; ni += nl > 1 ? nl - 1 : 0;
        if (p > 0x7fffffffll || s > 0x7fffffffll || 64 * g > 0x7fffffffll)
            return set_error(GVC_ERR_INVALID_ARGUMENT, "batch totals above 2^31");
    }
    Po[V] = (int)p; So[V] = (int)s; Go[V] = (int)g; Lo[V] = (int)l; No[V] = (int)ni;
    GVC_HIP(hipMemcpyAsync(table, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, st));
    GVC_HIP(hipStreamSynchronize(st));
    out->volumes = V;
    out->leaf_size = L;
    out->P_total = (int)p;
    out->S_total = (int)s;
    out->groups = (int)g;
    out->leaves = (int)l;
    out->internal_nodes = (int)ni;
    out->table = table;
    return GVC_OK;
}

static Batch from_c(const gvc_batch* b)
{
    Batch r;
    r.V = b->volumes; r.P = b->P_total; r.S = b->S_total; r.NG = b->groups; r.NL = b->leaves;
    r.NI = b->internal_nodes; r.L = b->leaf_size; r.table = b->table;
    return r;
}

size_t sample_set_bytes_batched(const gvc_batch* b)
{
    return sample_set_bytes_impl(b->S_total, b->groups, b->volumes);
}

int sample_set_build_batched(const gvc_batch* b, const float* samples, void* set, size_t bytes, hipStream_t st)
{
    return sample_set_build_impl(from_c(b), samples, set, bytes, st);
}

size_t forward_state_bytes_batched(const gvc_batch* b, const gvc_params* prm)
{
    return forward_state_bytes_impl(from_c(b), prm);
}

int mixture_forward_batched(const gvc_batch* b, const float* means, const float* scales, const float* rots,
                            const float* values, const float* weights, float k, const void* sample_set,
                            const gvc_params* prm, float* out_cells, float* out_weights, float* conics, void* state,
                            size_t state_bytes, hipStream_t st)
{
    return mixture_forward_impl(from_c(b), means, scales, rots, values, weights, k, sample_set, prm, out_cells,
                                out_weights, conics, state, state_bytes, st);
}

int forward_stats_batched(const gvc_batch* b, const void* state, const gvc_params* prm, uint64_t* n_records,
                          uint64_t* n_pairs, uint32_t* error_flags, uint64_t* chunk_slots_used, uint64_t* lane_groups,
                       hipStream_t st)
{
    return forward_stats_impl(from_c(b), state, prm, n_records, n_pairs, error_flags, chunk_slots_used, lane_groups,
                              st);
}

int forward_pair_counts_batched(const gvc_batch* b, const void* state, const gvc_params* prm, const void* sample_set,
                                uint32_t* sample_counts, uint32_t* gaussian_counts, hipStream_t st)
{
    if (prm->leaf_size != b->leaf_size)
        return set_error(GVC_ERR_INVALID_ARGUMENT, "params.leaf_size differs from the batch's");
    return forward_pair_counts_impl(from_c(b), state, prm, sample_set, sample_counts, gaussian_counts, st);
}

int forward_accumulate_batched(const gvc_batch* b, const void* state, const gvc_params* prm,
                               unsigned long long* accum, hipStream_t st)
{
    return forward_accumulate_impl(from_c(b), state, prm, accum, st);
}

int mixture_backward_batched(const gvc_batch* b, const float* scales, const float* rots, float k,
                             const void* sample_set, const gvc_params* prm, const float* out_cells,
                             const float* out_weights, const float* gcells, const float* gweights, void* state,
                             size_t state_bytes, uint64_t capacity, void* ws, size_t ws_bytes, float* dmeans,
                             float* dscales, float* drots, float* dvalues, float* dweights, hipStream_t st)
{
    return mixture_backward_impl(from_c(b), scales, rots, k, sample_set, prm, out_cells, out_weights, gcells, gweights,
                                 state, state_bytes, capacity, ws, ws_bytes, dmeans, dscales, drots, dvalues,
                                 dweights, st);
}

int sort_pairs_u64(int n, const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, int key_bits,
                   void* ws, size_t ws_bytes, hipStream_t st)
{
    if (n <= 0) return GVC_OK;
    Arena a(ws, ws_bytes);
    uint64_t* k2 = a.take<uint64_t>((size_t)n);
    uint32_t* v2 = a.take<uint32_t>((size_t)n);
    uint32_t* scr = a.take<uint32_t>(radix_scratch_elems((size_t)n));
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "sort workspace too small");
    GVC_HIP(hipMemcpyAsync(kout, kin, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    GVC_HIP(hipMemcpyAsync(vout, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    bool alt = false;
    GVC_TRY(radix_sort_u64(kout, vout, k2, v2, (size_t)n, key_bits, scr, &alt, st));
    if (alt) {
        GVC_HIP(hipMemcpyAsync(kout, k2, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
        GVC_HIP(hipMemcpyAsync(vout, v2, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    }
    return GVC_OK;
}

size_t sort_pairs_u64_bytes(int n)
{
    Arena a(nullptr, 0);
    a.take<uint64_t>((size_t)n);
    a.take<uint32_t>((size_t)n);
    a.take<uint32_t>(radix_scratch_elems((size_t)n));
    return a.off + 256;
}

}  // namespace gvc
