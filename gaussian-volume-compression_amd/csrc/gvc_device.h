// gvc_device.h -- canonical per-Gaussian and per-pair arithmetic on gfx950.
//
// Every function here rounds exactly like the CPU oracle (same operation order,
// explicit fmaf where a fused multiply-add is part of the definition; the whole
// library is built with -ffp-contract=off so nothing else is fused).  That is what
// makes pair membership, the box and the conic bit-identical between the GPU and
// the oracle.  Reference: CR/forward.cu:10-114 (preprocess), :160-171 (quadratic
// form and window), CR/config.h:9 (cutoff), with CR/ =
// diff-gaussian-rasterization/cuda_rasterizer/.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gvc {

constexpr double kWeightCutoff = 1e-2;  // CR/config.h:9
constexpr float kCondRatio = 1000.0f;   // cond(Sigma) = (s_max/s_min)^2 > 1e6

// Gaussian record, 80 bytes, stored in BVH (Morton) order.  One scalar
// s_load_dwordx16 + s_load_dwordx4 fetches it for a wave-uniform candidate.
struct alignas(16) GRecord {
    float mx, my, mz, w;       // mean, weight
    float c0, c1, c2, c3;      // conic xx xy xz yy
    float c4, c5, v;           // conic yz zz, value
    int id;                    // original Gaussian index
    float lox, loy, loz, hix;  // box
    float hiy, hiz;
    float pad0, pad1;
};
static_assert(sizeof(GRecord) == 80, "GRecord layout");

// Internal BVH node: boxes of both children plus child links (>= 0 internal node,
// < 0 leaf ~j).  64 bytes = one s_load_dwordx16.
struct alignas(16) BNode {
    float llo[3], lhi[3];
    float rlo[3], rhi[3];
    int left, right;
    int pad0, pad1;
};
static_assert(sizeof(BNode) == 64, "BNode layout");

// ---------------------------------------------------------------------------
// canonical log (double): x = 2^k f, f in [sqrt(1/2), sqrt(2)), log f = 2 atanh(s),
// s = (f-1)/(f+1), atanh by its odd Taylor series to s^21.  IEEE + - * / only.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double canon_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    uint64_t u = __builtin_bit_cast(uint64_t, x);
    if (u == 0x7ff0000000000000ull) return x;
    int k = 0;
    if ((u >> 52) == 0) {
        x = x * 18014398509481984.0;  // 2^54
        u = __builtin_bit_cast(uint64_t, x);
        k = -54;
    }
    k += (int)((u >> 52) & 0x7ff) - 1023;
    u = (u & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double f = __builtin_bit_cast(double, u);
    if (f > 1.4142135623730951) { f = f * 0.5; k += 1; }
    double s = (f - 1.0) / (f + 1.0);
    double s2 = s * s;
    double p = 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    double tail = s * s2 * p;
    double kd = (double)k;
    return (kd * 0.6931471803691238 + 2.0 * s) + (kd * 1.9082149292705877e-10 + 2.0 * tail);
}

// CR/forward.cu:79: m = sqrtf(-2 logf((0.1*cutoff)/w)); quotient in double, logf on
// its float rounding.  NaN for w < 1e-3 (the box then collapses onto the mean).
__device__ __forceinline__ float box_radius(float w)
{
    double q = (0.1 * kWeightCutoff) / (double)w;
    float qf = (float)q;
    float lf = (float)canon_log((double)qf);
    return __builtin_sqrtf(-2.0f * lf);
}

// CR/forward.cu:39-44, rows of the standard rotation matrix of unnormalised q.
__device__ __forceinline__ void rotation(float r, float x, float y, float z, float R[3][3])
{
    R[0][0] = 1.f - 2.f * (y * y + z * z);
    R[0][1] = 2.f * (x * y - r * z);
    R[0][2] = 2.f * (x * z + r * y);
    R[1][0] = 2.f * (x * y + r * z);
    R[1][1] = 1.f - 2.f * (x * x + z * z);
    R[1][2] = 2.f * (y * z - r * x);
    R[2][0] = 2.f * (x * z - r * y);
    R[2][1] = 2.f * (y * z + r * x);
    R[2][2] = 1.f - 2.f * (x * x + y * y);
}

__device__ __forceinline__ void rotation_d(float r_, float x_, float y_, float z_, double R[3][3])
{
    double r = r_, x = x_, y = y_, z = z_;
    R[0][0] = 1.0 - 2.0 * (y * y + z * z);
    R[0][1] = 2.0 * (x * y - r * z);
    R[0][2] = 2.0 * (x * z + r * y);
    R[1][0] = 2.0 * (x * y + r * z);
    R[1][1] = 1.0 - 2.0 * (x * x + z * z);
    R[1][2] = 2.0 * (y * z - r * x);
    R[2][0] = 2.0 * (x * z - r * y);
    R[2][1] = 2.0 * (y * z + r * x);
    R[2][2] = 1.0 - 2.0 * (x * x + y * y);
}

__device__ __forceinline__ bool ill_conditioned(float s0, float s1, float s2, float k)
{
    float a = fabsf(k * s0), b = fabsf(k * s1), c = fabsf(k * s2);
    float smax = fmaxf(fmaxf(a, b), c);
    float smin = fminf(fminf(a, b), c);
    return smax > kCondRatio * smin;
}

// CR/forward.cu:10-114 with the canonical operation order.  conic = packed
// inverse of Sigma + eps I (xx xy xz yy yz zz); box = lo.xyz hi.xyz.
__device__ __forceinline__ void preprocess_one(const float mu[3], const float sc[3], const float q[4],
                                               float w, float k, bool fp64_cond, float conic[6],
                                               float box[6])
{
    float R[3][3];
    rotation(q[0], q[1], q[2], q[3], R);
    float s[3] = { k * sc[0], k * sc[1], k * sc[2] };
    bool ill = fp64_cond && ill_conditioned(sc[0], sc[1], sc[2], k);
    if (!ill) {
        float M[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) M[a][c] = s[a] * R[c][a];
        float S00 = __builtin_fmaf(M[2][0], M[2][0], __builtin_fmaf(M[1][0], M[1][0], M[0][0] * M[0][0]));
        float S01 = __builtin_fmaf(M[2][0], M[2][1], __builtin_fmaf(M[1][0], M[1][1], M[0][0] * M[0][1]));
        float S02 = __builtin_fmaf(M[2][0], M[2][2], __builtin_fmaf(M[1][0], M[1][2], M[0][0] * M[0][2]));
        float S11 = __builtin_fmaf(M[2][1], M[2][1], __builtin_fmaf(M[1][1], M[1][1], M[0][1] * M[0][1]));
        float S12 = __builtin_fmaf(M[2][1], M[2][2], __builtin_fmaf(M[1][1], M[1][2], M[0][1] * M[0][2]));
        float S22 = __builtin_fmaf(M[2][2], M[2][2], __builtin_fmaf(M[1][2], M[1][2], M[0][2] * M[0][2]));
        float eps = (float)((double)fmaxf(fmaxf(fabsf(S00), fabsf(S11)), fabsf(S22)) * 1e-5);
        float a = S00 + eps, b = S01, c = S02, d = S11 + eps, e = S12, f = S22 + eps;
        float A00 = __builtin_fmaf(d, f, -(e * e));
        float A01 = __builtin_fmaf(c, e, -(b * f));
        float A02 = __builtin_fmaf(b, e, -(c * d));
        float A11 = __builtin_fmaf(a, f, -(c * c));
        float A12 = __builtin_fmaf(b, c, -(a * e));
        float A22 = __builtin_fmaf(a, d, -(b * b));
        float det = __builtin_fmaf(c, A02, __builtin_fmaf(b, A01, a * A00));
        float det_inv = (float)(1.0 / (double)det);
        conic[0] = A00 * det_inv;
        conic[1] = A01 * det_inv;
        conic[2] = A02 * det_inv;
        conic[3] = A11 * det_inv;
        conic[4] = A12 * det_inv;
        conic[5] = A22 * det_inv;
    } else {
        double Rd[3][3];
        rotation_d(q[0], q[1], q[2], q[3], Rd);
        double sd[3] = { (double)k * (double)sc[0], (double)k * (double)sc[1], (double)k * (double)sc[2] };
        double M[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) M[a][c] = sd[a] * Rd[c][a];
        double Sg[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) Sg[r][c] = M[0][r] * M[0][c] + M[1][r] * M[1][c] + M[2][r] * M[2][c];
        double mx = fabs(Sg[0][0]);
        if (fabs(Sg[1][1]) > mx) mx = fabs(Sg[1][1]);
        if (fabs(Sg[2][2]) > mx) mx = fabs(Sg[2][2]);
        double eps = mx * 1e-5;
        double a = Sg[0][0] + eps, b = Sg[0][1], c = Sg[0][2], d = Sg[1][1] + eps, e = Sg[1][2],
               f = Sg[2][2] + eps;
        double A00 = d * f - e * e, A01 = c * e - b * f, A02 = b * e - c * d;
        double A11 = a * f - c * c, A12 = b * c - a * e, A22 = a * d - b * b;
        double det = a * A00 + b * A01 + c * A02;
        conic[0] = (float)(A00 / det);
        conic[1] = (float)(A01 / det);
        conic[2] = (float)(A02 / det);
        conic[3] = (float)(A11 / det);
        conic[4] = (float)(A12 / det);
        conic[5] = (float)(A22 / det);
    }
    float m = box_radius(w);
    float sS[3] = { s[0] * m, s[1] * m, s[2] * m };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float p = mu[c];
        float t0 = fabsf(R[0][c] * sS[0]);
        float t1 = fabsf(R[1][c] * sS[1]);
        float t2 = fabsf(R[2][c] * sS[2]);
        float lo = ((p - t0) - t1) - t2;
        float hi = ((p + t0) + t1) + t2;
        box[c] = fminf(p, lo);
        box[3 + c] = fmaxf(p, hi);
    }
}

// Adjoint of the covariance chain in double: from the gradient of the packed
// conic (xx xy xz yy yz zz) to the gradients of the scales and the quaternion
// (r x y z).  The forward it differentiates: M = diag(k s) R^T, Sigma = M^T M,
// conic = (Sigma + eps I)^-1 with eps = 1e-5 max_i |Sigma_ii| (CR/forward.cu:26-75;
// the eps term's own derivative is kept on the largest diagonal entry).
__device__ __forceinline__ void covariance_chain(const double* __restrict__ g6, const float* __restrict__ sc,
                                                 const float* __restrict__ q, float kmod, double ds[3],
                                                 double dq[4])
{
    const float q4[4] = { q[0], q[1], q[2], q[3] };
    double R[3][3];
    rotation_d(q4[0], q4[1], q4[2], q4[3], R);
    const double km = kmod;
    const double s[3] = { km * sc[0], km * sc[1], km * sc[2] };
    double M[3][3];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) M[a][c] = s[a] * R[c][a];
    double Sg[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Sg[r][c] = M[0][r] * M[0][c] + M[1][r] * M[1][c] + M[2][r] * M[2][c];
    int jx = 0;
    double best = fabs(Sg[0][0]);
    if (fabs(Sg[1][1]) > best) { best = fabs(Sg[1][1]); jx = 1; }
    if (fabs(Sg[2][2]) > best) { best = fabs(Sg[2][2]); jx = 2; }
    const double eps = best * 1e-5;
    const double a_ = Sg[0][0] + eps, b_ = Sg[0][1], c_ = Sg[0][2], d_ = Sg[1][1] + eps, e_ = Sg[1][2],
                 f_ = Sg[2][2] + eps;
    const double A00 = d_ * f_ - e_ * e_, A01 = c_ * e_ - b_ * f_, A02 = b_ * e_ - c_ * d_;
    const double A11 = a_ * f_ - c_ * c_, A12 = b_ * c_ - a_ * e_, A22 = a_ * d_ - b_ * b_;
    const double det = a_ * A00 + b_ * A01 + c_ * A02;
    const double C[3][3] = { { A00 / det, A01 / det, A02 / det },
                             { A01 / det, A11 / det, A12 / det },
                             { A02 / det, A12 / det, A22 / det } };
    const double G[3][3] = { { g6[0], 0.5 * g6[1], 0.5 * g6[2] },
                             { 0.5 * g6[1], g6[3], 0.5 * g6[4] },
                             { 0.5 * g6[2], 0.5 * g6[4], g6[5] } };
    double CG[3][3], K[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) CG[r][c] = C[r][0] * G[0][c] + C[r][1] * G[1][c] + C[r][2] * G[2][c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) K[r][c] = -(CG[r][0] * C[0][c] + CG[r][1] * C[1][c] + CG[r][2] * C[2][c]);
    const double trH = K[0][0] + K[1][1] + K[2][2];
    const double sg = Sg[jx][jx] > 0 ? 1.0 : (Sg[jx][jx] < 0 ? -1.0 : 0.0);
    K[jx][jx] += trH * 1e-5 * sg;
    double dM[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) dM[r][c] = 2.0 * (M[r][0] * K[0][c] + M[r][1] * K[1][c] + M[r][2] * K[2][c]);
    for (int a = 0; a < 3; ++a) {
        const double v = dM[a][0] * R[0][a] + dM[a][1] * R[1][a] + dM[a][2] * R[2][a];
        ds[a] = km * v;
    }
    double Gq[3][3];
    for (int r = 0; r < 3; ++r)
        for (int a = 0; a < 3; ++a) Gq[r][a] = dM[a][r] * s[a];
    const double qr = q4[0], qx = q4[1], qy = q4[2], qz = q4[3];
    dq[0] = 2.0 * (qz * (Gq[1][0] - Gq[0][1]) + qy * (Gq[0][2] - Gq[2][0]) + qx * (Gq[2][1] - Gq[1][2]));
    dq[1] = 2.0 * (qy * (Gq[0][1] + Gq[1][0]) + qz * (Gq[0][2] + Gq[2][0]) + qr * (Gq[2][1] - Gq[1][2])) -
            4.0 * qx * (Gq[1][1] + Gq[2][2]);
    dq[2] = 2.0 * (qx * (Gq[0][1] + Gq[1][0]) + qr * (Gq[0][2] - Gq[2][0]) + qz * (Gq[1][2] + Gq[2][1])) -
            4.0 * qy * (Gq[0][0] + Gq[2][2]);
    dq[3] = 2.0 * (qr * (Gq[1][0] - Gq[0][1]) + qx * (Gq[0][2] + Gq[2][0]) + qy * (Gq[1][2] + Gq[2][1])) -
            4.0 * qz * (Gq[0][0] + Gq[1][1]);
}

// Canonical quadratic form and window (CR/forward.cu:160-166 in a fixed order).
// Returns the fp32 power; the pair is accepted iff in_window(power).
__device__ __forceinline__ float pair_power(float c0, float c1, float c2, float c3, float c4, float c5,
                                            float dx, float dy, float dz, float& r0, float& r1, float& r2)
{
    r0 = __builtin_fmaf(c2, dz, __builtin_fmaf(c1, dy, c0 * dx));
    r1 = __builtin_fmaf(c4, dz, __builtin_fmaf(c3, dy, c1 * dx));
    r2 = __builtin_fmaf(c5, dz, __builtin_fmaf(c4, dy, c2 * dx));
    float quad = __builtin_fmaf(dz, r2, __builtin_fmaf(dy, r1, dx * r0));
    return -0.5f * quad;
}

__device__ __forceinline__ bool in_window(float power) { return !(power < -14.0f || power > 0.0f); }

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).  For a
// grid over spatially sorted data, map block b to an index such that every XCD
// works through one contiguous 1/8 of the data, so neighbours share its L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    const unsigned x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ __forceinline__ bool boxes_overlap(float alx, float aly, float alz, float ahx, float ahy,
                                              float ahz, float blx, float bly, float blz, float bhx,
                                              float bhy, float bhz)
{
    return alx <= bhx && blx <= ahx && aly <= bhy && bly <= ahy && alz <= bhz && blz <= ahz;
}

// Morton bit spread, R/gaussian_renderer/__init__.py:49-56.
__device__ __forceinline__ uint64_t part1by2(int64_t n)
{
    uint64_t v = (uint64_t)n & 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// float -> int64 truncation as torch's .long() does on x86 and CUDA: NaN and
// out-of-range values give INT64_MIN.
__device__ __forceinline__ int64_t f32_to_i64_trunc(float v)
{
    if (v != v || v >= 9.2233720368547758e18f || v < -9.2233720368547758e18f) return INT64_MIN;
    return (int64_t)v;
}

__device__ __forceinline__ uint64_t morton3d_unit(float x, float y, float z, bool clamp)
{
    const float scale = 2097151.0f;
    float v[3] = { x, y, z };
    uint64_t c = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float t = v[a];
        if (clamp && t == t) t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        c |= part1by2(f32_to_i64_trunc(t * scale)) << a;
    }
    return c;
}

}  // namespace gvc
