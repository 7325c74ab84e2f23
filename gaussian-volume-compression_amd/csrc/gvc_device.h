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
#include <type_traits>

namespace gvc {

constexpr double kWeightCutoff = 1e-2;  // CR/config.h:9
constexpr float kCondRatio = 1000.0f;   // default FP64 threshold: cond(Sigma) = (s_max/s_min)^2 > 1e6
constexpr float kFactorRatio = 10.0f;   // default float-float factor threshold (FP64 mode): cond(Sigma) > 100
constexpr uint32_t kPreciseBit = 0x80000000u;  // GRecord::id bit: the Gaussian is on the precise path

// Gaussian record, 80 bytes, stored in BVH (Morton) order.  One scalar
// s_load_dwordx16 + s_load_dwordx4 fetches it for a wave-uniform candidate.
struct alignas(16) GRecord {
    float mx, my, mz, w;       // mean, weight
    float c0, c1, c2, c3;      // conic xx xy xz yy
    float c4, c5, v;           // conic yz zz, value
    int id;                    // original Gaussian index | kPreciseBit on the precise path
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

__device__ __forceinline__ bool ill_conditioned(float s0, float s1, float s2, float k, float ratio)
{
    float a = fabsf(k * s0), b = fabsf(k * s1), c = fabsf(k * s2);
    float smax = fmaxf(fmaxf(a, b), c);
    float smin = fminf(fminf(a, b), c);
    return smax > ratio * smin;
}

// Upper Cholesky factor U of an FP64 conic (packed xx xy xz yy yz zz), conic =
// U^T U, as hi + lo float pairs (u00 u01 u02 u11 u12 u22).  False when the conic
// is not numerically positive definite (the Gaussian then keeps its conic).
__device__ __forceinline__ bool conic_factor_d(const double C[6], double u[6])
{
    if (!(C[0] > 0.0)) return false;
    u[0] = __builtin_sqrt(C[0]);
    u[1] = C[1] / u[0];
    u[2] = C[2] / u[0];
    const double a1 = C[3] - u[1] * u[1];
    if (!(a1 > 0.0)) return false;
    u[3] = __builtin_sqrt(a1);
    u[4] = (C[4] - u[1] * u[2]) / u[3];
    const double a2 = (C[5] - u[2] * u[2]) - u[4] * u[4];
    if (!(a2 > 0.0)) return false;
    u[5] = __builtin_sqrt(a2);
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (!(fabsf((float)u[k]) <= 3.4e38f)) return false;
    return true;
}

__device__ __forceinline__ bool conic_factor(const double C[6], float hi[6], float lo[6])
{
    double u[6];
    if (!conic_factor_d(C, u)) return false;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        hi[k] = (float)u[k];
        lo[k] = (float)(u[k] - (double)hi[k]);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Float-float arithmetic: a value is h + l with |l| <= ulp(h) / 2, every step in
// fp32 (Knuth's TwoSum, fma residual products, Dekker's normalisation), about
// 44 bits of precision.  The compensated-fp32 factor below runs the FP64 path's
// operation sequence in it, so Gaussians between the factor threshold and the
// FP64 threshold get an accurate conic factor without FP64 arithmetic.
// ---------------------------------------------------------------------------
struct FF {
    float h, l;
};

__device__ __forceinline__ FF ff_quick(float a, float b)  // |a| >= |b| or a == 0
{
    const float s = a + b;
    return FF{ s, b - (s - a) };
}
__device__ __forceinline__ FF ff_two_sum(float a, float b)
{
    const float s = a + b;
    const float bb = s - a;
    return FF{ s, (a - (s - bb)) + (b - bb) };
}
__device__ __forceinline__ FF ff_two_prod(float a, float b)
{
    const float p = a * b;
    return FF{ p, __builtin_fmaf(a, b, -p) };
}
__device__ __forceinline__ FF ff_neg(FF x) { return FF{ -x.h, -x.l }; }
__device__ __forceinline__ FF ff_add(FF x, FF y)
{
    FF s = ff_two_sum(x.h, y.h);
    const FF t = ff_two_sum(x.l, y.l);
    s.l += t.h;
    s = ff_quick(s.h, s.l);
    s.l += t.l;
    return ff_quick(s.h, s.l);
}
__device__ __forceinline__ FF ff_sub(FF x, FF y) { return ff_add(x, ff_neg(y)); }
__device__ __forceinline__ FF ff_mul(FF x, FF y)
{
    FF p = ff_two_prod(x.h, y.h);
    p.l = __builtin_fmaf(x.h, y.l, __builtin_fmaf(x.l, y.h, p.l));
    return ff_quick(p.h, p.l);
}
__device__ __forceinline__ FF ff_mul_f(FF x, float y)
{
    FF p = ff_two_prod(x.h, y);
    p.l = __builtin_fmaf(x.l, y, p.l);
    return ff_quick(p.h, p.l);
}
// x / y: three quotient digits, each from the remainder's leading float.
__device__ __forceinline__ FF ff_div(FF x, FF y)
{
    const float q1 = x.h / y.h;
    FF r = ff_sub(x, ff_mul_f(y, q1));
    const float q2 = r.h / y.h;
    r = ff_sub(r, ff_mul_f(y, q2));
    const float q3 = r.h / y.h;
    return ff_add(ff_quick(q1, q2), FF{ q3, 0.0f });
}
// sqrt(x) for x > 0: the float root and one Newton correction from the remainder.
__device__ __forceinline__ FF ff_sqrt(FF x)
{
    const float s = __builtin_sqrtf(x.h);
    const FF r = ff_sub(x, ff_two_prod(s, s));
    return ff_quick(s, r.h / (s + s));
}
__device__ __forceinline__ bool ff_less(FF a, FF b) { return a.h < b.h || (a.h == b.h && a.l < b.l); }

// 1e-5 as a float-float constant (hi = the float nearest 1e-5, lo = the rest).
constexpr float kEpsHi = (float)1e-5;
constexpr float kEpsLo = (float)(1e-5 - (double)(float)1e-5);

// The upper Cholesky factor of the conic (Sigma + eps I)^-1 of scales sc (times
// k) and quaternion q, computed in float-float arithmetic along the FP64 path's
// operation sequence (rotation, M = S R^T, Sigma = M^T M, eps = 1e-5 max diag,
// adjugate, determinant, quotients, Cholesky); entries u00 u01 u02 u11 u12 u22
// as hi + lo floats.  False when the conic is not positive definite in this
// arithmetic or a factor entry is not finite (the Gaussian keeps its fp32 conic).
__device__ __forceinline__ bool conic_factor_ff(const float sc[3], const float q[4], float k, float hi[6],
                                                float lo[6])
{
    const float r = q[0], x = q[1], y = q[2], z = q[3];
    const FF one{ 1.0f, 0.0f };
    auto two = [](FF v) { return FF{ 2.0f * v.h, 2.0f * v.l }; };
    FF R[3][3];
    R[0][0] = ff_sub(one, two(ff_add(ff_two_prod(y, y), ff_two_prod(z, z))));
    R[0][1] = two(ff_sub(ff_two_prod(x, y), ff_two_prod(r, z)));
    R[0][2] = two(ff_add(ff_two_prod(x, z), ff_two_prod(r, y)));
    R[1][0] = two(ff_add(ff_two_prod(x, y), ff_two_prod(r, z)));
    R[1][1] = ff_sub(one, two(ff_add(ff_two_prod(x, x), ff_two_prod(z, z))));
    R[1][2] = two(ff_sub(ff_two_prod(y, z), ff_two_prod(r, x)));
    R[2][0] = two(ff_sub(ff_two_prod(x, z), ff_two_prod(r, y)));
    R[2][1] = two(ff_add(ff_two_prod(y, z), ff_two_prod(r, x)));
    R[2][2] = ff_sub(one, two(ff_add(ff_two_prod(x, x), ff_two_prod(y, y))));
    FF M[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const FF sa = ff_two_prod(k, sc[a]);
#pragma unroll
        for (int c = 0; c < 3; ++c) M[a][c] = ff_mul(sa, R[c][a]);
    }
    auto sig = [&](int i, int j) {
        return ff_add(ff_add(ff_mul(M[0][i], M[0][j]), ff_mul(M[1][i], M[1][j])), ff_mul(M[2][i], M[2][j]));
    };
    const FF S00 = sig(0, 0), S01 = sig(0, 1), S02 = sig(0, 2), S11 = sig(1, 1), S12 = sig(1, 2), S22 = sig(2, 2);
    FF mx = S00;   // diagonal entries are sums of squares: no absolute value needed
    if (ff_less(mx, S11)) mx = S11;
    if (ff_less(mx, S22)) mx = S22;
    const FF eps = ff_mul(mx, FF{ kEpsHi, kEpsLo });
    const FF a = ff_add(S00, eps), b = S01, c = S02, d = ff_add(S11, eps), e = S12, f = ff_add(S22, eps);
    const FF A00 = ff_sub(ff_mul(d, f), ff_mul(e, e));
    const FF A01 = ff_sub(ff_mul(c, e), ff_mul(b, f));
    const FF A02 = ff_sub(ff_mul(b, e), ff_mul(c, d));
    const FF A11 = ff_sub(ff_mul(a, f), ff_mul(c, c));
    const FF A12 = ff_sub(ff_mul(b, c), ff_mul(a, e));
    const FF A22 = ff_sub(ff_mul(a, d), ff_mul(b, b));
    const FF det = ff_add(ff_add(ff_mul(a, A00), ff_mul(b, A01)), ff_mul(c, A02));
    if (!(det.h > 0.0f)) return false;
    const FF C0 = ff_div(A00, det), C1 = ff_div(A01, det), C2 = ff_div(A02, det);
    const FF C3 = ff_div(A11, det), C4 = ff_div(A12, det), C5 = ff_div(A22, det);
    if (!(C0.h > 0.0f)) return false;
    FF u[6];
    u[0] = ff_sqrt(C0);
    u[1] = ff_div(C1, u[0]);
    u[2] = ff_div(C2, u[0]);
    const FF a1 = ff_sub(C3, ff_mul(u[1], u[1]));
    if (!(a1.h > 0.0f)) return false;
    u[3] = ff_sqrt(a1);
    u[4] = ff_div(ff_sub(C4, ff_mul(u[1], u[2])), u[3]);
    const FF a2 = ff_sub(ff_sub(C5, ff_mul(u[2], u[2])), ff_mul(u[4], u[4]));
    if (!(a2.h > 0.0f)) return false;
    u[5] = ff_sqrt(a2);
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        if (!(fabsf(u[t].h) <= 3.4e38f && fabsf(u[t].l) <= 3.4e38f)) return false;
        hi[t] = u[t].h;
        lo[t] = u[t].l;
    }
    return true;
}

// CR/forward.cu:10-114 with the canonical operation order.  conic = packed
// inverse of Sigma + eps I (xx xy xz yy yz zz); box = lo.xyz hi.xyz.
// fp64_ratio > 0: Gaussians whose scale ratio exceeds it (cond(Sigma) above its
// square) take the FP64 covariance inverse; with want_factor, such a Gaussian
// also gets the Cholesky factor of that FP64 conic in fac_hi/fac_lo and returns
// true (the precise path: its pairs are evaluated from the factor, precise_power).
// factor_ratio > 0 (with want_factor): a Gaussian whose scale ratio exceeds it
// but not fp64_ratio keeps its fp32 conic and gets the factor in float-float
// arithmetic (conic_factor_ff), and returns true when that succeeds: the precise
// path without FP64.
__device__ __forceinline__ bool preprocess_one(const float mu[3], const float sc[3], const float q[4],
                                               float w, float k, float fp64_ratio, float conic[6],
                                               float box[6], float* fac_hi = nullptr, float* fac_lo = nullptr,
                                               float factor_ratio = 0.f, bool want_factor = false)
{
    float R[3][3];
    rotation(q[0], q[1], q[2], q[3], R);
    float s[3] = { k * sc[0], k * sc[1], k * sc[2] };
    bool ill = fp64_ratio > 0.f && ill_conditioned(sc[0], sc[1], sc[2], k, fp64_ratio);
    bool precise = false;
    if (!ill) {
        if (want_factor && factor_ratio > 0.f && ill_conditioned(sc[0], sc[1], sc[2], k, factor_ratio))
            precise = conic_factor_ff(sc, q, k, fac_hi, fac_lo);
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
        const double cd[6] = { A00 / det, A01 / det, A02 / det, A11 / det, A12 / det, A22 / det };
#pragma unroll
        for (int t = 0; t < 6; ++t) conic[t] = (float)cd[t];
        if (want_factor) precise = conic_factor(cd, fac_hi, fac_lo);
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
    return precise;
}

// Adjoint of the covariance chain in double: from the gradient of the packed
// conic (xx xy xz yy yz zz) to the gradients of the scales and the quaternion
// (r x y z).  The forward it differentiates: M = diag(k s) R^T, Sigma = M^T M,
// conic = (Sigma + eps I)^-1 with eps = 1e-5 max_i |Sigma_ii| (CR/forward.cu:26-75;
// the eps term's own derivative is kept on the largest diagonal entry).
// factor_frame: the Gaussian is on the precise path (its record carries the
// precise bit), so g6 is with respect to the factor-frame outer products (sum of
// dquad y y^T, y = U d, conic = U^T U), and dL/dSigma = -C G C is formed as
// -U^T G_y U with U the Cholesky factor of the FP64 conic (exact algebra with
// G = U^-1 G_y U^-T; no cancellation against the conic's large entries).
__device__ __forceinline__ void covariance_chain(const double* __restrict__ g6, const float* __restrict__ sc,
                                                 const float* __restrict__ q, float kmod, double ds[3],
                                                 double dq[4], bool factor_frame = false)
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
    const double c6[6] = { C[0][0], C[0][1], C[0][2], C[1][1], C[1][2], C[2][2] };
    double u6[6];
    if (factor_frame && conic_factor_d(c6, u6)) {
        const double U[3][3] = { { u6[0], u6[1], u6[2] }, { 0.0, u6[3], u6[4] }, { 0.0, 0.0, u6[5] } };
        for (int r = 0; r < 3; ++r)   // CG = U^T G_y
            for (int c = 0; c < 3; ++c) CG[r][c] = U[0][r] * G[0][c] + U[1][r] * G[1][c] + U[2][r] * G[2][c];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) K[r][c] = -(CG[r][0] * U[0][c] + CG[r][1] * U[1][c] + CG[r][2] * U[2][c]);
    } else {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) CG[r][c] = C[r][0] * G[0][c] + C[r][1] * G[1][c] + C[r][2] * G[2][c];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) K[r][c] = -(CG[r][0] * C[0][c] + CG[r][1] * C[1][c] + CG[r][2] * C[2][c]);
    }
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
// pair_quad returns the fp32 quadratic form; the power is -0.5f * quad and the
// pair is accepted iff in_window(power).
__device__ __forceinline__ float pair_quad(float c0, float c1, float c2, float c3, float c4, float c5,
                                           float dx, float dy, float dz, float& r0, float& r1, float& r2)
{
    r0 = __builtin_fmaf(c2, dz, __builtin_fmaf(c1, dy, c0 * dx));
    r1 = __builtin_fmaf(c4, dz, __builtin_fmaf(c3, dy, c1 * dx));
    r2 = __builtin_fmaf(c5, dz, __builtin_fmaf(c4, dy, c2 * dx));
    return __builtin_fmaf(dz, r2, __builtin_fmaf(dy, r1, dx * r0));
}

__device__ __forceinline__ float pair_power(float c0, float c1, float c2, float c3, float c4, float c5,
                                            float dx, float dy, float dz, float& r0, float& r1, float& r2)
{
    return -0.5f * pair_quad(c0, c1, c2, c3, c4, c5, dx, dy, dz, r0, r1, r2);
}

__device__ __forceinline__ bool in_window(float power) { return !(power < -14.0f || power > 0.0f); }

// The same window decided on quad: in_window(-0.5f * q) == quad_in_window(q) for
// every float q.  -0.5f * q is exact except for an odd multiple of the smallest
// subnormal, which rounds half to even; so power < -14 iff q > 28, and power > 0
// iff q < -2^-149 (q = -2^-149 gives power +0); NaN passes both.
__device__ __forceinline__ bool quad_in_window(float q) { return !(q > 28.0f || q < -0x1p-149f); }

// e^power = 2^(power log2 e) with power = -0.5f * q: q * (-0.5 log2 e) rounds the
// same product once (the halving is exact), so the weight equals the one formed
// from the power, bit for bit (subnormal powers give 1 either way).
constexpr float kLog2eF = 1.4426950408889634f;
constexpr float kNegHalfLog2e = -0.5f * kLog2eF;
__device__ __forceinline__ float quad_exp(float q) { return __builtin_amdgcn_exp2f(q * kNegHalfLog2e); }

// Knuth's TwoSum: s + e == a + b exactly (round to nearest, no contraction).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// (s, c) += (uh + ul)(dh + dl): uh dh exactly (product and its fma residual, the
// product added by TwoSum), the cross terms uh dl + ul dh rounded once, ul dl dropped.
__device__ __forceinline__ void ff_madd(float& s, float& c, float uh, float ul, float dh, float dl)
{
    const float p = uh * dh;
    const float pe = __builtin_fmaf(uh, dh, -p);
    float t, e;
    two_sum(s, p, t, e);
    s = t;
    c = c + (e + (pe + __builtin_fmaf(uh, dl, ul * dh)));
}

// Power of a pair on the precise path.  The conic is U^T U with U upper
// triangular (u00 u01 u02 u11 u12 u22 = h*, each entry h + l); d = x - mu exactly
// as dh + dl; y = U d in float-float arithmetic, each component rounded once;
// power = -0.5 |y|^2.  The fp32 conic form d^T C d cancels against entries up to
// cond times larger than its value; here the cancellation is carried exactly.
// y feeds precise_conic_d.  Window and use as for pair_power.
__device__ __forceinline__ float precise_power(float px, float py, float pz, float mx, float my, float mz,
                                               float h0, float h1, float h2, float h3, float h4, float h5,
                                               const float4& l0123, const float4& l45, float& y0, float& y1,
                                               float& y2)
{
    float xh, xl, yh, yl, zh, zl;
    two_sum(px, -mx, xh, xl);
    two_sum(py, -my, yh, yl);
    two_sum(pz, -mz, zh, zl);
    float s = 0.f, c = 0.f;
    ff_madd(s, c, h0, l0123.x, xh, xl);
    ff_madd(s, c, h1, l0123.y, yh, yl);
    ff_madd(s, c, h2, l0123.z, zh, zl);
    y0 = s + c;
    s = 0.f; c = 0.f;
    ff_madd(s, c, h3, l0123.w, yh, yl);
    ff_madd(s, c, h4, l45.x, zh, zl);
    y1 = s + c;
    s = 0.f; c = 0.f;
    ff_madd(s, c, h5, l45.y, zh, zl);
    y2 = s + c;
    const float quad = __builtin_fmaf(y2, y2, __builtin_fmaf(y1, y1, y0 * y0));
    return -0.5f * quad;
}
// precise_power's quadratic form (-0.5f * precise_quad == precise_power).
__device__ __forceinline__ float precise_quad(float px, float py, float pz, float mx, float my, float mz, float h0,
                                              float h1, float h2, float h3, float h4, float h5, const float4& l0123,
                                              const float4& l45, float& y0, float& y1, float& y2)
{
    (void)precise_power(px, py, pz, mx, my, mz, h0, h1, h2, h3, h4, h5, l0123, l45, y0, y1, y2);
    return __builtin_fmaf(y2, y2, __builtin_fmaf(y1, y1, y0 * y0));
}

// conic d = U^T y on the precise path (the mean's gradient term of pair_power's r).
__device__ __forceinline__ void precise_conic_d(float h0, float h1, float h2, float h3, float h4, float h5,
                                                float y0, float y1, float y2, float& r0, float& r1, float& r2)
{
    r0 = h0 * y0;
    r1 = __builtin_fmaf(h3, y1, h1 * y0);
    r2 = __builtin_fmaf(h5, y2, __builtin_fmaf(h4, y1, h2 * y0));
}

// Pair evaluation of a record's candidate (A = mean|w, B = c0..c3, C = c4 c5 v id):
// the canonical fp32 form, or with PRECISE and the id's precise bit set, the
// factor form (hi parts in the conic slots, lo parts in fac_lo[2 j], [2 j + 1] for
// the Gaussian's caller index j, the id without kPreciseBit).
// r = conic d, and (dx, dy, dz) = the vector whose outer products the adjoint sums
// for the conic's gradient: d = x - mu (fp32) on the conic form, y = U d on the
// precise path (the chain then takes that gradient in the factor frame).
template <bool PRECISE>
__device__ __forceinline__ float record_power(const float4& A, const float4& B, const float4& C,
                                              const float4* __restrict__ fac_lo, float px, float py,
                                              float pz, float& dx, float& dy, float& dz, float& r0, float& r1,
                                              float& r2)
{
    dx = px - A.x;
    dy = py - A.y;
    dz = pz - A.z;
    if constexpr (PRECISE) {
        const int id = __builtin_bit_cast(int, C.w);
        if (id < 0) {
            const size_t j = (size_t)(uint32_t)(id & 0x7fffffff);
            const float4 l0 = fac_lo[2 * j], l1 = fac_lo[2 * j + 1];
            float y0, y1, y2;
            const float pw = precise_power(px, py, pz, A.x, A.y, A.z, B.x, B.y, B.z, B.w, C.x, C.y, l0, l1, y0,
                                           y1, y2);
            precise_conic_d(B.x, B.y, B.z, B.w, C.x, C.y, y0, y1, y2, r0, r1, r2);
            dx = y0;
            dy = y1;
            dz = y2;
            return pw;
        }
    }
    return pair_power(B.x, B.y, B.z, B.w, C.x, C.y, dx, dy, dz, r0, r1, r2);
}

// record_power's quadratic form: -0.5f * record_quad == record_power.
template <bool PRECISE>
__device__ __forceinline__ float record_quad(const float4& A, const float4& B, const float4& C,
                                             const float4* __restrict__ fac_lo, float px, float py, float pz,
                                             float& dx, float& dy, float& dz, float& r0, float& r1, float& r2)
{
    dx = px - A.x;
    dy = py - A.y;
    dz = pz - A.z;
    if constexpr (PRECISE) {
        const int id = __builtin_bit_cast(int, C.w);
        if (id < 0) {
            const size_t j = (size_t)(uint32_t)(id & 0x7fffffff);
            const float4 l0 = fac_lo[2 * j], l1 = fac_lo[2 * j + 1];
            float y0, y1, y2;
            const float q = precise_quad(px, py, pz, A.x, A.y, A.z, B.x, B.y, B.z, B.w, C.x, C.y, l0, l1, y0, y1,
                                         y2);
            precise_conic_d(B.x, B.y, B.z, B.w, C.x, C.y, y0, y1, y2, r0, r1, r2);
            dx = y0;
            dy = y1;
            dz = y2;
            return q;
        }
    }
    return pair_quad(B.x, B.y, B.z, B.w, C.x, C.y, dx, dy, dz, r0, r1, r2);
}

// ---------------------------------------------------------------------------
// Cross-lane exchanges on the VALU (no LDS round trip), for wave64 code whose 64
// lanes are all active: lane_xor<J>(v) is the value of the lane whose index
// differs in bit J.  J = 1, 2, 4, 8 use DPP within rows of 16 lanes (quad
// permutations; the half-row mirror followed by a quad reversal for 4; a row
// rotation by 8); J = 16, 32 use gfx950's v_permlane16_swap / v_permlane32_swap.
// ---------------------------------------------------------------------------
// dpp_mov: lanes outside ROWS, and lanes without a source when ZERO, get 0.  A
// full-row move that cannot lack a source (a permutation) leaves the old value
// undefined, so no register has to be cleared first.
template <int CTRL, int ROWS = 0xF, bool ZERO = false>
__device__ __forceinline__ int dpp_mov(int v)
{
    if constexpr (ROWS == 0xF && !ZERO)
        return __builtin_amdgcn_mov_dpp(v, CTRL, ROWS, 0xF, false);
    else
        return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, ZERO);
}

template <int J>
__device__ __forceinline__ int lane_xor(int v)
{
    if constexpr (J == 1) {
        return dpp_mov<0xB1>(v);                 // quad_perm [1, 0, 3, 2]
    } else if constexpr (J == 2) {
        return dpp_mov<0x4E>(v);                 // quad_perm [2, 3, 0, 1]
    } else if constexpr (J == 4) {
        return dpp_mov<0x1B>(dpp_mov<0x141>(v)); // half-row mirror (i ^ 7), then quad_perm [3, 2, 1, 0] (i ^ 3)
    } else if constexpr (J == 8) {
        return dpp_mov<0x128>(v);                // row_ror:8
    } else if constexpr (J == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (threadIdx.x & 16u) ? (int)r[0] : (int)r[1];
    } else {
        static_assert(J == 32, "lane_xor: J must be a power of two up to 32");
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (threadIdx.x & 32u) ? (int)r[0] : (int)r[1];
    }
}
template <int J>
__device__ __forceinline__ float lane_xor(float v)
{
    return __builtin_bit_cast(float, lane_xor<J>(__builtin_bit_cast(int, v)));
}
template <int J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v)
{
    return (uint32_t)lane_xor<J>((int)v);
}

// The value of lane (lane - D) of the same wave, 0 for lanes below D (D = 1, 2,
// 4, 8 shift within rows of 16; callers combine rows with row_carry below).
template <int D>
__device__ __forceinline__ uint32_t row_shr(uint32_t v)
{
    return (uint32_t)dpp_mov<0x110 + D, 0xF, true>((int)v);
}

// Inclusive scan over the wave's 64 lanes in lane order with an associative,
// commutative OP whose identity is 0 (add, or): Hillis-Steele within each row of
// 16 lanes (DPP row shifts), then the last lane of row 0 broadcast into row 1
// and of row 2 into row 3, then the last lane of row 1 into rows 2 and 3.
template <class OP>
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, OP op)
{
    v = op(v, row_shr<1>(v));
    v = op(v, row_shr<2>(v));
    v = op(v, row_shr<4>(v));
    v = op(v, row_shr<8>(v));
    v = op(v, (uint32_t)dpp_mov<0x142, 0xA>((int)v));  // row_bcast:15 into rows 1, 3
    v = op(v, (uint32_t)dpp_mov<0x143, 0xC>((int)v));  // row_bcast:31 into rows 2, 3
    return v;
}

// Order-preserving integer key of a float (an involution): non-negative floats
// keep their bits, negative ones flip all but the sign bit, so signed integer
// order is float order (-0 just below +0).  NaN has no meaningful key: callers
// exclude it.
__device__ __forceinline__ int fkey(float x)
{
    const int b = __builtin_bit_cast(int, x);
    return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float fkey_inv(int k) { return __builtin_bit_cast(float, k ^ ((k >> 31) & 0x7fffffff)); }
constexpr int kKeyPosInf = 0x7f800000;        // fkey(+inf)
constexpr int kKeyNegInf = (int)0x807fffffu;  // fkey(-inf)

// Min / max of integer keys over the wave (all lanes active), returned
// wave-uniform: an inclusive scan within each row of 16 lanes (DPP row shifts; a
// lane without a source keeps its value), then row_bcast:15 into rows 1 and 3 and
// row_bcast:31 into rows 2 and 3, so lane 63 holds the whole wave's result.  The
// old value of each DPP move is the operation's identity, so the move folds into
// the min / max: one VALU instruction per step.
template <bool MAX>
__device__ __forceinline__ int wave_extreme_key(int k)
{
    constexpr int id = MAX ? (int)0x80000000u : 0x7fffffff;
    auto step = [&](auto ctrl, auto rows) {
        const int o = __builtin_amdgcn_update_dpp(id, k, decltype(ctrl)::value, decltype(rows)::value, 0xF, false);
        k = MAX ? (o > k ? o : k) : (o < k ? o : k);
    };
    step(std::integral_constant<int, 0x111>{}, std::integral_constant<int, 0xF>{});  // row_shr:1
    step(std::integral_constant<int, 0x112>{}, std::integral_constant<int, 0xF>{});  // row_shr:2
    step(std::integral_constant<int, 0x114>{}, std::integral_constant<int, 0xF>{});  // row_shr:4
    step(std::integral_constant<int, 0x118>{}, std::integral_constant<int, 0xF>{});  // row_shr:8
    step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{});  // row_bcast:15
    step(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xC>{});  // row_bcast:31
    return __builtin_amdgcn_readlane(k, 63);
}

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).  For a
// grid over spatially sorted data, map block b to an index such that every XCD
// works through one contiguous 1/8 of the data, so neighbours share its L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    const unsigned x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Six ordered comparisons (false on NaN), each a vector compare into a scalar
// lane mask, and-ed on the scalar unit; the lane's own bit of the result is the
// answer.  Same value as the && chain for every active lane, without the
// compiler's bit-packing of the chain on the vector unit.
__device__ __forceinline__ bool all6(bool c0, bool c1, bool c2, bool c3, bool c4, bool c5)
{
    const uint64_t m = __builtin_amdgcn_ballot_w64(c0) & __builtin_amdgcn_ballot_w64(c1) &
                       __builtin_amdgcn_ballot_w64(c2) & __builtin_amdgcn_ballot_w64(c3) &
                       __builtin_amdgcn_ballot_w64(c4) & __builtin_amdgcn_ballot_w64(c5);
    return __builtin_amdgcn_inverse_ballot_w64(m);
}

__device__ __forceinline__ bool boxes_overlap(float alx, float aly, float alz, float ahx, float ahy,
                                              float ahz, float blx, float bly, float blz, float bhx,
                                              float bhy, float bhz)
{
    return all6(alx <= bhx, blx <= ahx, aly <= bhy, bly <= ahy, alz <= bhz, blz <= ahz);
}

// Wave mask of the active lanes whose box [al, ah] overlaps [bl, bh] (the same
// comparisons as boxes_overlap), for callers that go on with lane masks.
__device__ __forceinline__ uint64_t boxes_overlap_mask(float alx, float aly, float alz, float ahx, float ahy,
                                                       float ahz, float blx, float bly, float blz, float bhx,
                                                       float bhy, float bhz)
{
    return __builtin_amdgcn_ballot_w64(alx <= bhx) & __builtin_amdgcn_ballot_w64(blx <= ahx) &
           __builtin_amdgcn_ballot_w64(aly <= bhy) & __builtin_amdgcn_ballot_w64(bly <= ahy) &
           __builtin_amdgcn_ballot_w64(alz <= bhz) & __builtin_amdgcn_ballot_w64(blz <= ahz);
}

// Set bits of m in the lanes below this one (mbcnt: two vector instructions).
__device__ __forceinline__ uint32_t lane_rank(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// lo <= p <= hi on every axis (closed box), in the same form.
__device__ __forceinline__ bool point_in_box(float px, float py, float pz, float lx, float ly, float lz, float hx,
                                             float hy, float hz)
{
    return all6(lx <= px, px <= hx, ly <= py, py <= hy, lz <= pz, pz <= hz);
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
