/*
 * gvc_oracle.c -- CPU restatement of the reference's Gaussian-mixture hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for orc_cpu_plain_*, as the timed CPU baseline).  The product path never
 * calls it.
 *
 * Citations use R/ = the reference repository root, CR/ =
 * R/diff-gaussian-rasterization/cuda_rasterizer/.
 *
 * What is restated (and how each function is pinned by tests/test_oracle_*.py):
 *   orc_log                 canonical double log (range reduction + atanh series);
 *                           pinned against math.log (<= 1 ulp) on 2e5 inputs.
 *   orc_box_radius          m(w) = sqrtf(-2 logf(1e-3/w))       CR/forward.cu:79
 *                           pinned against the closed form in float64.
 *   orc_preprocess          conic + AABB per Gaussian           CR/forward.cu:10-114
 *                           pinned: identity quaternion -> axis box mu +- m s and
 *                           conic = 1/(s^2+eps); conic*cov = I; 90-degree
 *                           rotations permute the axes; random rotations against
 *                           a float64 numpy inverse.
 *   orc_morton3d_unit       63-bit Morton code        R/gaussian_renderer/__init__.py:45-58
 *                           (clamp variant: R/train.py:291-315); pinned by known
 *                           answers (0 -> 0, 1 -> 2^63-1, fp32 truncation, wrap).
 *   orc_forward             brute force, canonical predicate, FP64 contributions
 *                           CR/forward.cu:185-259 (+ :261-274 normalise); pinned by
 *                           single-Gaussian closed form, convexity, permutation
 *                           invariance, and a float64 torch restatement.
 *   orc_backward            FP64 per-pair adjoint CR/backward.cu:257-298 and the
 *                           covariance chain derived from the forward formula (the
 *                           reference's own chain CR/backward.cu:68-124 is not
 *                           followed: it does not match autograd); pinned against
 *                           torch float64 autograd and central differences.
 *   orc_karras_build        Karras (2012) binary radix tree over sorted codes;
 *                           pinned against a recursive top-down split on small n.
 *   orc_cpu_plain_fwdbwd    plain fp32 brute force (CPU baseline, OpenMP).
 *
 * Canonical fp32 operation order (shared with the HIP kernels; gcc builds this
 * file with -ffp-contract=off so no multiply-add is fused unless written as
 * fmaf, which glibc rounds correctly):
 *   rotation        reference expression order, plain float ops
 *   Sigma_ij        fmaf(M2i,M2j, fmaf(M1i,M1j, M0i*M0j)),  M_ai = (k s_a) Rstd[i][a]
 *   epsilon         (float)((double)max|Sigma_ii| * 1e-5)
 *   adjugate        fmaf(p,q,-(r*s)) per 2x2 minor;  det = fmaf(c,A02,fmaf(b,A01,a*A00))
 *   det_inv         (float)(1.0/(double)det)
 *   box             lo = fminf(mu, ((mu-|t0|)-|t1|)-|t2|), t_k = Rstd[k][c]*(k s_k m)
 *   quad            r_i = fmaf(C_i2,dz,fmaf(C_i1,dy,C_i0*dx)); q = fmaf(dz,r2,fmaf(dy,r1,dx*r0))
 *   window          power = -0.5f*q; accept iff !(power < -14 || power > 0)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_WEIGHT_CUTOFF 1e-2 /* CR/config.h:9 */
#define ORC_COND_RATIO 1000.0f /* cond(Sigma) = (s_max/s_min)^2 > 1e6 */
#define ORC_FACTOR_RATIO 10.0f /* float-float factor below the FP64 threshold: cond > 100 */

/* ------------------------------------------------------------------------- */
/* canonical log: x = 2^k * f, f in [sqrt(1/2), sqrt(2));  log f = 2 atanh(s),  */
/* s = (f-1)/(f+1), |s| <= 0.1716;  atanh by its odd Taylor series (11 terms).  */
/* Only IEEE + - * / are used, so every conforming double unit reproduces it.   */
/* ------------------------------------------------------------------------- */
static const double ORC_LN2_HI = 0.6931471803691238;     /* ln2 to 32 bits   */
static const double ORC_LN2_LO = 1.9082149292705877e-10; /* ln2 - ORC_LN2_HI */

double orc_log(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    if (x != x) return x;                       /* NaN */
    if (x < 0.0) return (x - x) / (x - x);      /* NaN */
    if (x == 0.0) return -1.0 / (x * x);        /* -inf (x*x = +0) */
    if (u == 0x7ff0000000000000ull) return x;   /* +inf */
    int k = 0;
    if ((u >> 52) == 0) {                       /* subnormal: scale by 2^54 */
        x = x * 18014398509481984.0;
        memcpy(&u, &x, 8);
        k = -54;
    }
    k += (int)((u >> 52) & 0x7ff) - 1023;
    u = (u & 0x000fffffffffffffull) | 0x3ff0000000000000ull; /* f in [1,2) */
    double f;
    memcpy(&f, &u, 8);
    if (f > 1.4142135623730951) { f = f * 0.5; k += 1; }
    double s = (f - 1.0) / (f + 1.0);
    double s2 = s * s;
    /* 2*atanh(s) = 2 s (1 + s2/3 + s2^2/5 + ... + s2^10/21), Horner from the top */
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
    double tail = s * s2 * p;                   /* s^3/3 + s^5/5 + ... */
    double kd = (double)k;
    return (kd * ORC_LN2_HI + 2.0 * s) + (kd * ORC_LN2_LO + 2.0 * tail);
}

/* CR/forward.cu:79 -- m = sqrtf(-2*logf((0.1*WEIGHT_CUTOFF)/w)); the quotient is a
 * double, rounded to float for logf.  logf is the canonical log rounded to float. */
float orc_box_radius(float w)
{
    double q = (0.1 * ORC_WEIGHT_CUTOFF) / (double)w;
    float qf = (float)q;
    float lf = (float)orc_log((double)qf);
    return sqrtf(-2.0f * lf);
}

/* CR/forward.cu:39-44: the standard rotation matrix of q = (r,x,y,z), written row by
 * row (glm stores these rows as its columns).  q is not normalised (:33). */
static void orc_rotation(const float q[4], float R[3][3])
{
    float r = q[0], x = q[1], y = q[2], z = q[3];
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

static void orc_rotation_d(const float q[4], double R[3][3])
{
    double r = q[0], x = q[1], y = q[2], z = q[3];
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

/* True when the scale ratio s_max/s_min exceeds ratio (cond(Sigma) = ratio^2 on the
 * unregularised Sigma with unit q; the default ratio 1000 is cond > 1e6). */
int orc_ill_conditioned(const float s[3], float scale_modifier, float ratio)
{
    float a = fabsf(scale_modifier * s[0]);
    float b = fabsf(scale_modifier * s[1]);
    float c = fabsf(scale_modifier * s[2]);
    float smax = fmaxf(fmaxf(a, b), c);
    float smin = fminf(fminf(a, b), c);
    return smax > ratio * smin;
}

/*
 * Upper Cholesky factor U of an FP64 conic (packed xx xy xz yy yz zz), conic =
 * U^T U, entries u00 u01 u02 u11 u12 u22 as hi + lo floats.  Returns 0 when the
 * conic is not numerically positive definite (the Gaussian keeps its conic).
 */
int orc_conic_factor(const double C[6], float hi[6], float lo[6])
{
    if (!(C[0] > 0.0)) return 0;
    const double u00 = sqrt(C[0]);
    const double u01 = C[1] / u00, u02 = C[2] / u00;
    const double a1 = C[3] - u01 * u01;
    if (!(a1 > 0.0)) return 0;
    const double u11 = sqrt(a1);
    const double u12 = (C[4] - u01 * u02) / u11;
    const double a2 = (C[5] - u02 * u02) - u12 * u12;
    if (!(a2 > 0.0)) return 0;
    const double u[6] = { u00, u01, u02, u11, u12, sqrt(a2) };
    for (int k = 0; k < 6; ++k) {
        hi[k] = (float)u[k];
        if (!(fabsf(hi[k]) <= 3.4e38f)) return 0;
        lo[k] = (float)(u[k] - (double)hi[k]);
    }
    return 1;
}

/* ------------------------------------------------------------------------- */
/* Float-float arithmetic (value = h + l, fp32 operations only): TwoSum, fma     */
/* residual products, Dekker's normalisation; three-digit quotient, one Newton   */
/* step for the root.  orc_conic_factor_ff runs the FP64 path's formulas in it   */
/* (the same operation sequence as the HIP kernels' compensated-fp32 factor).    */
/* ------------------------------------------------------------------------- */
typedef struct { float h, l; } orc_ff;

static inline orc_ff ff_quick(float a, float b)
{
    orc_ff r;
    r.h = a + b;
    r.l = b - (r.h - a);
    return r;
}
static inline orc_ff ff_two_sum(float a, float b)
{
    orc_ff r;
    r.h = a + b;
    float bb = r.h - a;
    r.l = (a - (r.h - bb)) + (b - bb);
    return r;
}
static inline orc_ff ff_two_prod(float a, float b)
{
    orc_ff r;
    r.h = a * b;
    r.l = fmaf(a, b, -r.h);
    return r;
}
static inline orc_ff ff_neg(orc_ff x) { orc_ff r = { -x.h, -x.l }; return r; }
static inline orc_ff ff_add(orc_ff x, orc_ff y)
{
    orc_ff s = ff_two_sum(x.h, y.h);
    orc_ff t = ff_two_sum(x.l, y.l);
    s.l += t.h;
    s = ff_quick(s.h, s.l);
    s.l += t.l;
    return ff_quick(s.h, s.l);
}
static inline orc_ff ff_sub(orc_ff x, orc_ff y) { return ff_add(x, ff_neg(y)); }
static inline orc_ff ff_mul(orc_ff x, orc_ff y)
{
    orc_ff p = ff_two_prod(x.h, y.h);
    p.l = fmaf(x.h, y.l, fmaf(x.l, y.h, p.l));
    return ff_quick(p.h, p.l);
}
static inline orc_ff ff_mul_f(orc_ff x, float y)
{
    orc_ff p = ff_two_prod(x.h, y);
    p.l = fmaf(x.l, y, p.l);
    return ff_quick(p.h, p.l);
}
static inline orc_ff ff_div(orc_ff x, orc_ff y)
{
    float q1 = x.h / y.h;
    orc_ff r = ff_sub(x, ff_mul_f(y, q1));
    float q2 = r.h / y.h;
    r = ff_sub(r, ff_mul_f(y, q2));
    float q3 = r.h / y.h;
    orc_ff q3f = { q3, 0.0f };
    return ff_add(ff_quick(q1, q2), q3f);
}
static inline orc_ff ff_sqrt(orc_ff x)
{
    float s = sqrtf(x.h);
    orc_ff r = ff_sub(x, ff_two_prod(s, s));
    return ff_quick(s, r.h / (s + s));
}
static inline int ff_less(orc_ff a, orc_ff b) { return a.h < b.h || (a.h == b.h && a.l < b.l); }
static inline orc_ff ff_twice(orc_ff v) { orc_ff r = { 2.0f * v.h, 2.0f * v.l }; return r; }

/* Upper Cholesky factor of (Sigma + eps I)^-1 in float-float arithmetic, along
 * the FP64 path: rotation of q, M = S R^T (S = k s), Sigma = M^T M, eps = 1e-5 max
 * diag, adjugate, determinant, quotients, Cholesky.  hi/lo = the entries' float
 * parts.  Returns 0 when not positive definite or not finite.  Pinned in
 * tests/test_oracle.py: agrees with the FP64 factor (orc_conic_factor of
 * orc_conic_fp64) to ~1e-12 relative and U^T U reproduces the FP64 conic. */
int orc_conic_factor_ff(const float sc[3], const float q[4], float k, float hi[6], float lo[6])
{
    const float r = q[0], x = q[1], y = q[2], z = q[3];
    const orc_ff one = { 1.0f, 0.0f };
    orc_ff R[3][3];
    R[0][0] = ff_sub(one, ff_twice(ff_add(ff_two_prod(y, y), ff_two_prod(z, z))));
    R[0][1] = ff_twice(ff_sub(ff_two_prod(x, y), ff_two_prod(r, z)));
    R[0][2] = ff_twice(ff_add(ff_two_prod(x, z), ff_two_prod(r, y)));
    R[1][0] = ff_twice(ff_add(ff_two_prod(x, y), ff_two_prod(r, z)));
    R[1][1] = ff_sub(one, ff_twice(ff_add(ff_two_prod(x, x), ff_two_prod(z, z))));
    R[1][2] = ff_twice(ff_sub(ff_two_prod(y, z), ff_two_prod(r, x)));
    R[2][0] = ff_twice(ff_sub(ff_two_prod(x, z), ff_two_prod(r, y)));
    R[2][1] = ff_twice(ff_add(ff_two_prod(y, z), ff_two_prod(r, x)));
    R[2][2] = ff_sub(one, ff_twice(ff_add(ff_two_prod(x, x), ff_two_prod(y, y))));
    orc_ff M[3][3];
    for (int a = 0; a < 3; ++a) {
        orc_ff sa = ff_two_prod(k, sc[a]);
        for (int c = 0; c < 3; ++c) M[a][c] = ff_mul(sa, R[c][a]);
    }
    orc_ff S[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j)
            S[i][j] = ff_add(ff_add(ff_mul(M[0][i], M[0][j]), ff_mul(M[1][i], M[1][j])), ff_mul(M[2][i], M[2][j]));
    orc_ff mx = S[0][0];
    if (ff_less(mx, S[1][1])) mx = S[1][1];
    if (ff_less(mx, S[2][2])) mx = S[2][2];
    const float eps_hi = (float)1e-5;
    const orc_ff eps_c = { eps_hi, (float)(1e-5 - (double)eps_hi) };
    orc_ff eps = ff_mul(mx, eps_c);
    orc_ff a = ff_add(S[0][0], eps), b = S[0][1], c = S[0][2], d = ff_add(S[1][1], eps), e = S[1][2],
           f = ff_add(S[2][2], eps);
    orc_ff A00 = ff_sub(ff_mul(d, f), ff_mul(e, e));
    orc_ff A01 = ff_sub(ff_mul(c, e), ff_mul(b, f));
    orc_ff A02 = ff_sub(ff_mul(b, e), ff_mul(c, d));
    orc_ff A11 = ff_sub(ff_mul(a, f), ff_mul(c, c));
    orc_ff A12 = ff_sub(ff_mul(b, c), ff_mul(a, e));
    orc_ff A22 = ff_sub(ff_mul(a, d), ff_mul(b, b));
    orc_ff det = ff_add(ff_add(ff_mul(a, A00), ff_mul(b, A01)), ff_mul(c, A02));
    if (!(det.h > 0.0f)) return 0;
    orc_ff C0 = ff_div(A00, det), C1 = ff_div(A01, det), C2 = ff_div(A02, det);
    orc_ff C3 = ff_div(A11, det), C4 = ff_div(A12, det), C5 = ff_div(A22, det);
    if (!(C0.h > 0.0f)) return 0;
    orc_ff u[6];
    u[0] = ff_sqrt(C0);
    u[1] = ff_div(C1, u[0]);
    u[2] = ff_div(C2, u[0]);
    orc_ff a1 = ff_sub(C3, ff_mul(u[1], u[1]));
    if (!(a1.h > 0.0f)) return 0;
    u[3] = ff_sqrt(a1);
    u[4] = ff_div(ff_sub(C4, ff_mul(u[1], u[2])), u[3]);
    orc_ff a2 = ff_sub(ff_sub(C5, ff_mul(u[2], u[2])), ff_mul(u[4], u[4]));
    if (!(a2.h > 0.0f)) return 0;
    u[5] = ff_sqrt(a2);
    for (int t = 0; t < 6; ++t) {
        if (!(fabsf(u[t].h) <= 3.4e38f && fabsf(u[t].l) <= 3.4e38f)) return 0;
        hi[t] = u[t].h;
        lo[t] = u[t].l;
    }
    return 1;
}

/* Knuth's TwoSum: s + e == a + b exactly. */
static inline void orc_two_sum(float a, float b, float *s, float *e)
{
    float t = a + b;
    float bb = t - a;
    *e = (a - (t - bb)) + (b - bb);
    *s = t;
}

/* (s, c) += (uh + ul)(dh + dl): uh dh exactly (product, fma residual, TwoSum),
 * cross terms uh dl + ul dh rounded once, ul dl dropped. */
static inline void orc_ff_madd(float *s, float *c, float uh, float ul, float dh, float dl)
{
    float p = uh * dh;
    float pe = fmaf(uh, dh, -p);
    float t, e;
    orc_two_sum(*s, p, &t, &e);
    *s = t;
    *c = *c + (e + (pe + fmaf(uh, dl, ul * dh)));
}

/* Power of a pair on the precise path: d = x - mu exactly as dh + dl (TwoSum),
 * y = U d in float-float arithmetic, each component rounded once to float,
 * power = -0.5 |y|^2 with quad = fma(y2, y2, fma(y1, y1, y0 y0)).  fac = hi[6] lo[6]. */
float orc_precise_power(const float *fac, const float *mu, const float *x)
{
    const float *h = fac, *l = fac + 6;
    float xh, xl, yh, yl, zh, zl;
    orc_two_sum(x[0], -mu[0], &xh, &xl);
    orc_two_sum(x[1], -mu[1], &yh, &yl);
    orc_two_sum(x[2], -mu[2], &zh, &zl);
    float s = 0.f, c = 0.f;
    orc_ff_madd(&s, &c, h[0], l[0], xh, xl);
    orc_ff_madd(&s, &c, h[1], l[1], yh, yl);
    orc_ff_madd(&s, &c, h[2], l[2], zh, zl);
    float y0 = s + c;
    s = 0.f; c = 0.f;
    orc_ff_madd(&s, &c, h[3], l[3], yh, yl);
    orc_ff_madd(&s, &c, h[4], l[4], zh, zl);
    float y1 = s + c;
    s = 0.f; c = 0.f;
    orc_ff_madd(&s, &c, h[5], l[5], zh, zl);
    float y2 = s + c;
    float quad = fmaf(y2, y2, fmaf(y1, y1, y0 * y0));
    return -0.5f * quad;
}

/*
 * CR/forward.cu:10-114 (preprocessCUDA).  Outputs per Gaussian:
 *   conic[6] = packed inverse of (Sigma + eps I): xx xy xz yy yz zz
 *   box[6]   = lo.xyz, hi.xyz
 * fp64_cond != 0: Gaussians whose scale ratio exceeds fp64_ratio (0 = 1000, i.e.
 * cond(Sigma) > 1e6) get their covariance, epsilon, determinant and adjugate in
 * double (plain ops), conic = (float)(adj/det).  factor_out (P x 12, may be NULL):
 * for those, the Cholesky factor of the double conic as hi[6] lo[6]
 * (orc_conic_factor) -- the precise path; NaN in every other row.  Gaussians
 * whose scale ratio exceeds factor_ratio (0 = 10) but not the FP64 threshold keep
 * the fp32 conic and get the float-float factor (orc_conic_factor_ff) instead.
 */
void orc_preprocess(int P, const float *means, const float *scales, const float *rots,
                    const float *weights, float scale_modifier, int fp64_cond, float fp64_ratio,
                    float factor_ratio, float *conic_out, float *box_out, int32_t *ill_out, float *factor_out)
{
    const float ratio = fp64_ratio > 0.f ? fp64_ratio : ORC_COND_RATIO;
    const float fratio = factor_ratio > 0.f ? factor_ratio : ORC_FACTOR_RATIO;
    for (int i = 0; i < P; ++i) {
        const float *q = rots + 4 * i;
        float s[3] = { scale_modifier * scales[3 * i + 0], scale_modifier * scales[3 * i + 1],
                       scale_modifier * scales[3 * i + 2] };
        float R[3][3];
        orc_rotation(q, R);
        int ill = fp64_cond ? orc_ill_conditioned(scales + 3 * i, scale_modifier, ratio) : 0;
        if (ill_out) ill_out[i] = ill;
        float *cn = conic_out + 6 * i;
        if (factor_out)
            for (int k = 0; k < 12; ++k) factor_out[12 * i + k] = NAN;
        if (!ill) {
            float M[3][3]; /* M[a][i] = s_a * Rstd[i][a] */
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) M[a][c] = s[a] * R[c][a];
            float S00 = fmaf(M[2][0], M[2][0], fmaf(M[1][0], M[1][0], M[0][0] * M[0][0]));
            float S01 = fmaf(M[2][0], M[2][1], fmaf(M[1][0], M[1][1], M[0][0] * M[0][1]));
            float S02 = fmaf(M[2][0], M[2][2], fmaf(M[1][0], M[1][2], M[0][0] * M[0][2]));
            float S11 = fmaf(M[2][1], M[2][1], fmaf(M[1][1], M[1][1], M[0][1] * M[0][1]));
            float S12 = fmaf(M[2][1], M[2][2], fmaf(M[1][1], M[1][2], M[0][1] * M[0][2]));
            float S22 = fmaf(M[2][2], M[2][2], fmaf(M[1][2], M[1][2], M[0][2] * M[0][2]));
            float eps = (float)((double)fmaxf(fmaxf(fabsf(S00), fabsf(S11)), fabsf(S22)) * 1e-5);
            float a = S00 + eps, b = S01, c = S02, d = S11 + eps, e = S12, f = S22 + eps;
            float A00 = fmaf(d, f, -(e * e));
            float A01 = fmaf(c, e, -(b * f));
            float A02 = fmaf(b, e, -(c * d));
            float A11 = fmaf(a, f, -(c * c));
            float A12 = fmaf(b, c, -(a * e));
            float A22 = fmaf(a, d, -(b * b));
            float det = fmaf(c, A02, fmaf(b, A01, a * A00));
            float det_inv = (float)(1.0 / (double)det);
            cn[0] = A00 * det_inv;
            cn[1] = A01 * det_inv;
            cn[2] = A02 * det_inv;
            cn[3] = A11 * det_inv;
            cn[4] = A12 * det_inv;
            cn[5] = A22 * det_inv;
            if (fp64_cond && factor_out && orc_ill_conditioned(scales + 3 * i, scale_modifier, fratio)) {
                float hi[6], lo[6];
                if (orc_conic_factor_ff(scales + 3 * i, q, scale_modifier, hi, lo))
                    for (int k = 0; k < 6; ++k) {
                        factor_out[12 * i + k] = hi[k];
                        factor_out[12 * i + 6 + k] = lo[k];
                    }
            }
        } else {
            double Rd[3][3];
            orc_rotation_d(q, Rd);
            double sd[3] = { (double)scale_modifier * (double)scales[3 * i + 0],
                             (double)scale_modifier * (double)scales[3 * i + 1],
                             (double)scale_modifier * (double)scales[3 * i + 2] };
            double M[3][3];
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) M[a][c] = sd[a] * Rd[c][a];
            double Sg[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    Sg[r][c] = M[0][r] * M[0][c] + M[1][r] * M[1][c] + M[2][r] * M[2][c];
            double mx = fabs(Sg[0][0]);
            if (fabs(Sg[1][1]) > mx) mx = fabs(Sg[1][1]);
            if (fabs(Sg[2][2]) > mx) mx = fabs(Sg[2][2]);
            double eps = mx * 1e-5;
            double a = Sg[0][0] + eps, b = Sg[0][1], c = Sg[0][2], d = Sg[1][1] + eps,
                   e = Sg[1][2], f = Sg[2][2] + eps;
            double A00 = d * f - e * e, A01 = c * e - b * f, A02 = b * e - c * d;
            double A11 = a * f - c * c, A12 = b * c - a * e, A22 = a * d - b * b;
            double det = a * A00 + b * A01 + c * A02;
            const double cd[6] = { A00 / det, A01 / det, A02 / det, A11 / det, A12 / det, A22 / det };
            for (int k = 0; k < 6; ++k) cn[k] = (float)cd[k];
            if (factor_out) {
                float hi[6], lo[6];
                if (orc_conic_factor(cd, hi, lo))
                    for (int k = 0; k < 6; ++k) {
                        factor_out[12 * i + k] = hi[k];
                        factor_out[12 * i + 6 + k] = lo[k];
                    }
            }
        }
        /* AABB: CR/forward.cu:79-113 */
        float m = orc_box_radius(weights[i]);
        float sS[3] = { s[0] * m, s[1] * m, s[2] * m };
        float *bx = box_out + 6 * i;
        for (int c = 0; c < 3; ++c) {
            float p = means[3 * i + c];
            float t0 = fabsf(R[0][c] * sS[0]);
            float t1 = fabsf(R[1][c] * sS[1]);
            float t2 = fabsf(R[2][c] * sS[2]);
            float lo = ((p - t0) - t1) - t2;
            float hi = ((p + t0) + t1) + t2;
            bx[c] = fminf(p, lo);
            bx[3 + c] = fmaxf(p, hi);
        }
    }
}

/* Canonical per-pair predicate.  Returns 1 if (Gaussian, point) contributes and
 * writes the fp32 power.  fac (12 floats or NULL): a Gaussian on the precise path
 * (fac[0] not NaN) takes its power from orc_precise_power. */
static inline int orc_pair_accept(const float *box, const float *mu, const float *cn, const float *fac,
                                  const float *x, float *power_out)
{
    if (!(box[0] <= x[0] && x[0] <= box[3] && box[1] <= x[1] && x[1] <= box[4] &&
          box[2] <= x[2] && x[2] <= box[5]))
        return 0;
    if (fac && fac[0] == fac[0]) {
        float pw = orc_precise_power(fac, mu, x);
        if (pw < -14.0f || pw > 0.0f) return 0;
        *power_out = pw;
        return 1;
    }
    float dx = x[0] - mu[0], dy = x[1] - mu[1], dz = x[2] - mu[2];
    float r0 = fmaf(cn[2], dz, fmaf(cn[1], dy, cn[0] * dx));
    float r1 = fmaf(cn[4], dz, fmaf(cn[3], dy, cn[1] * dx));
    float r2 = fmaf(cn[5], dz, fmaf(cn[4], dy, cn[2] * dx));
    float quad = fmaf(dz, r2, fmaf(dy, r1, dx * r0));
    float power = -0.5f * quad;
    if (power < -14.0f || power > 0.0f) return 0;
    *power_out = power;
    return 1;
}

int orc_pair_accept_ext(const float *box, const float *mu, const float *cn, const float *fac, const float *x,
                        float *power_out)
{
    return orc_pair_accept(box, mu, cn, fac, x, power_out);
}

/* Dense membership mask (small problems only): mask[i*S + s] = 1 iff the canonical
 * predicate accepts (Gaussian i, point s); power[i*S + s] = its fp32 power. */
void orc_pair_mask(int P, int S, const float *means, const float *conic, const float *box,
                   const float *factor, const float *samples, uint8_t *mask, float *power)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; ++i)
        for (int s = 0; s < S; ++s) {
            float pw = 0.f;
            int a = orc_pair_accept(box + 6 * i, means + 3 * i, conic + 6 * i, factor ? factor + 12 * i : NULL,
                                    samples + 3 * s, &pw);
            mask[(size_t)i * S + s] = (uint8_t)a;
            if (power) power[(size_t)i * S + s] = a ? pw : 0.f;
        }
}

/*
 * Brute-force forward, CR/forward.cu:185-259 + :261-274 with the canonical
 * predicate.  Contributions in double: d = x - mu and quad in double from the
 * fp32 conic (conic_d may be NULL) or from an FP64 conic (conic_d given).
 * Outputs V (cells) and W (weights) in double; sentinel (-1, 0) when W <= 1e-2.
 * count_out[s] = number of accepted pairs at point s (may be NULL).
 */
/* FP64 contribution of an accepted pair to (W, U). */
static inline void orc_pair_add(int i, const float *x, const float *means, const float *values,
                                const float *weights, const float *conic, const double *conic_d,
                                double *W, double *U)
{
    const float *cf = conic + 6 * i;
    double c[6];
    for (int k = 0; k < 6; ++k) c[k] = conic_d ? conic_d[6 * i + k] : (double)cf[k];
    double dx = (double)x[0] - means[3 * i + 0];
    double dy = (double)x[1] - means[3 * i + 1];
    double dz = (double)x[2] - means[3 * i + 2];
    double quad = dx * (c[0] * dx + c[1] * dy + c[2] * dz) +
                  dy * (c[1] * dx + c[3] * dy + c[4] * dz) +
                  dz * (c[2] * dx + c[4] * dy + c[5] * dz);
    double wgt = (double)weights[i] * exp(-0.5 * quad);
    *W += wgt;
    *U += wgt * (double)values[i];
}

static inline void orc_finish(double W, double U, int cnt, int s, double *V_out, double *W_out,
                              int32_t *count_out)
{
    if (W <= ORC_WEIGHT_CUTOFF) {
        V_out[s] = -1.0;
        W_out[s] = 0.0;
    } else {
        V_out[s] = U / W;
        W_out[s] = W;
    }
    if (count_out) count_out[s] = cnt;
}

void orc_forward(int P, int S, const float *means, const float *values, const float *weights,
                 const float *conic, const float *box, const double *conic_d, const float *factor,
                 const float *samples, double *V_out, double *W_out, int32_t *count_out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int s = 0; s < S; ++s) {
        const float *x = samples + 3 * s;
        double W = 0.0, U = 0.0;
        int cnt = 0;
        for (int i = 0; i < P; ++i) {
            float pw;
            if (!orc_pair_accept(box + 6 * i, means + 3 * i, conic + 6 * i, factor ? factor + 12 * i : NULL, x,
                                 &pw))
                continue;
            orc_pair_add(i, x, means, values, weights, conic, conic_d, &W, &U);
            ++cnt;
        }
        orc_finish(W, U, cnt, s, V_out, W_out, count_out);
    }
}

/* The same forward with the Gaussians binned in a uniform grid of G^3 cells over
 * the finite part of their boxes (a box reaching past the grid is clamped to it;
 * a box with a NaN bound holds nothing and is not binned).  Each cell lists its
 * Gaussians in ascending index, so every sample visits its accepted pairs in the
 * order of orc_forward and the FP64 sums are identical.  Returns 0, or -1 when
 * the bins would exceed max_entries. */
int orc_forward_binned(int P, int S, const float *means, const float *values, const float *weights,
                       const float *conic, const float *box, const double *conic_d, const float *factor,
                       const float *samples, int G, long long max_entries, double *V_out,
                       double *W_out, int32_t *count_out)
{
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int i = 0; i < P; ++i)
        for (int a = 0; a < 3; ++a) {
            double l = box[6 * i + a], h = box[6 * i + 3 + a];
            if (isfinite(l) && l < lo[a]) lo[a] = l;
            if (isfinite(h) && h > hi[a]) hi[a] = h;
        }
    for (int a = 0; a < 3; ++a)
        if (!(lo[a] <= hi[a])) { lo[a] = 0.0; hi[a] = 1.0; }
    if (G < 1) G = 1;
    const double w[3] = { (hi[0] - lo[0]) / G + 1e-300, (hi[1] - lo[1]) / G + 1e-300, (hi[2] - lo[2]) / G + 1e-300 };
#define ORC_CELL(v, a) ((int)fmin(fmax(floor(((v) - lo[a]) / w[a]), 0.0), (double)(G - 1)))
    const size_t nc = (size_t)G * G * G;
    long long *start = (long long *)calloc(nc + 1, sizeof(long long));
    if (!start) return -1;
    for (int pass = 0; pass < 2; ++pass) {
        int *list = NULL;
        long long *fill = NULL;
        if (pass == 1) {
            for (size_t c = 0; c < nc; ++c) start[c + 1] += start[c];
            if (start[nc] > max_entries) { free(start); return -1; }
            list = (int *)malloc(sizeof(int) * (size_t)(start[nc] > 0 ? start[nc] : 1));
            fill = (long long *)malloc(sizeof(long long) * nc);
            if (!list || !fill) { free(list); free(fill); free(start); return -1; }
            memcpy(fill, start, sizeof(long long) * nc);
        }
        for (int i = 0; i < P; ++i) {
            const float *b = box + 6 * i;
            if (b[0] != b[0] || b[1] != b[1] || b[2] != b[2] || b[3] != b[3] || b[4] != b[4] || b[5] != b[5])
                continue;
            const int x0 = ORC_CELL(b[0], 0), x1 = ORC_CELL(b[3], 0);
            const int y0 = ORC_CELL(b[1], 1), y1 = ORC_CELL(b[4], 1);
            const int z0 = ORC_CELL(b[2], 2), z1 = ORC_CELL(b[5], 2);
            for (int z = z0; z <= z1; ++z)
                for (int y = y0; y <= y1; ++y)
                    for (int x = x0; x <= x1; ++x) {
                        const size_t c = ((size_t)z * G + y) * G + x;
                        if (pass == 0) ++start[c + 1];
                        else list[fill[c]++] = i;
                    }
        }
        if (pass == 1) {
#pragma omp parallel for schedule(dynamic, 64)
            for (int s = 0; s < S; ++s) {
                const float *x = samples + 3 * s;
                double Wv = 0.0, U = 0.0;
                int cnt = 0;
                if (x[0] == x[0] && x[1] == x[1] && x[2] == x[2]) {
                    const size_t c = ((size_t)ORC_CELL(x[2], 2) * G + ORC_CELL(x[1], 1)) * G + ORC_CELL(x[0], 0);
                    for (long long k = start[c]; k < start[c + 1]; ++k) {
                        const int i = list[k];
                        float pw;
                        if (!orc_pair_accept(box + 6 * i, means + 3 * i, conic + 6 * i,
                                             factor ? factor + 12 * i : NULL, x, &pw))
                            continue;
                        orc_pair_add(i, x, means, values, weights, conic, conic_d, &Wv, &U);
                        ++cnt;
                    }
                }
                orc_finish(Wv, U, cnt, s, V_out, W_out, count_out);
            }
            free(list);
            free(fill);
        }
    }
#undef ORC_CELL
    free(start);
    return 0;
}

/*
 * Per-Gaussian covariance chain: gradient of the packed conic (as used in the
 * quadratic form with the off-diagonals counted twice) -> scales, rotations.
 * Derived from the forward formula (Sigma = Rstd S^2 Rstd^T, eps regularisation,
 * C = (Sigma + eps I)^-1):
 *   G  = sym(gc) with off-diagonals halved;   H = -C G C
 *   K  = H + tr(H) 1e-5 sign(Sigma_jj) E_jj   (j = argmax |Sigma_ii|, first on ties)
 *   dM = 2 M K,  M = S Rstd^T
 *   ds_a = k (dM Rstd)_aa;   dRstd = dM^T S;   dq from dRstd (chain through R(q)).
 * Computed entirely in double from the float inputs.
 */
void orc_cov_chain(const float *scale3, const float *q4, float scale_modifier, const double *C6,
                   const double *gc6, double *dscale3, double *dq4)
{
    double R[3][3];
    orc_rotation_d(q4, R);
    double k = scale_modifier;
    double s[3] = { k * scale3[0], k * scale3[1], k * scale3[2] };
    double M[3][3];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) M[a][c] = s[a] * R[c][a];
    double Sg[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Sg[r][c] = M[0][r] * M[0][c] + M[1][r] * M[1][c] + M[2][r] * M[2][c];
    double C[3][3] = { { C6[0], C6[1], C6[2] }, { C6[1], C6[3], C6[4] }, { C6[2], C6[4], C6[5] } };
    double G[3][3] = { { gc6[0], 0.5 * gc6[1], 0.5 * gc6[2] },
                       { 0.5 * gc6[1], gc6[3], 0.5 * gc6[4] },
                       { 0.5 * gc6[2], 0.5 * gc6[4], gc6[5] } };
    double CG[3][3], H[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) CG[r][c] = C[r][0] * G[0][c] + C[r][1] * G[1][c] + C[r][2] * G[2][c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) H[r][c] = -(CG[r][0] * C[0][c] + CG[r][1] * C[1][c] + CG[r][2] * C[2][c]);
    /* epsilon = 1e-5 * max|Sigma_ii| (first index wins ties, as max(max(a,b),c) in the
     * reference picks a over b and the larger over c) */
    int j = 0;
    double best = fabs(Sg[0][0]);
    if (fabs(Sg[1][1]) > best) { best = fabs(Sg[1][1]); j = 1; }
    if (fabs(Sg[2][2]) > best) { best = fabs(Sg[2][2]); j = 2; }
    double trH = H[0][0] + H[1][1] + H[2][2];
    double sg = Sg[j][j] > 0 ? 1.0 : (Sg[j][j] < 0 ? -1.0 : 0.0);
    double K[3][3];
    memcpy(K, H, sizeof(K));
    K[j][j] += trH * 1e-5 * sg;
    double dM[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) dM[r][c] = 2.0 * (M[r][0] * K[0][c] + M[r][1] * K[1][c] + M[r][2] * K[2][c]);
    for (int a = 0; a < 3; ++a) {
        double acc = dM[a][0] * R[0][a] + dM[a][1] * R[1][a] + dM[a][2] * R[2][a];
        dscale3[a] = k * acc;
    }
    double Gq[3][3]; /* dL/dRstd[i][a] = dM[a][i] * s_a */
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 3; ++a) Gq[i][a] = dM[a][i] * s[a];
    double r = q4[0], x = q4[1], y = q4[2], z = q4[3];
    dq4[0] = 2.0 * (z * (Gq[1][0] - Gq[0][1]) + y * (Gq[0][2] - Gq[2][0]) + x * (Gq[2][1] - Gq[1][2]));
    dq4[1] = 2.0 * (y * (Gq[0][1] + Gq[1][0]) + z * (Gq[0][2] + Gq[2][0]) + r * (Gq[2][1] - Gq[1][2])) -
             4.0 * x * (Gq[1][1] + Gq[2][2]);
    dq4[2] = 2.0 * (x * (Gq[0][1] + Gq[1][0]) + r * (Gq[0][2] - Gq[2][0]) + z * (Gq[1][2] + Gq[2][1])) -
             4.0 * y * (Gq[0][0] + Gq[2][2]);
    dq4[3] = 2.0 * (r * (Gq[1][0] - Gq[0][1]) + x * (Gq[0][2] + Gq[2][0]) + y * (Gq[1][2] + Gq[2][1])) -
             4.0 * z * (Gq[0][0] + Gq[1][1]);
}

/* FP64 conic of the forward formula (plain double ops) -- used by the chain. */
void orc_conic_fp64(const float *scale3, const float *q4, float scale_modifier, double *C6)
{
    double R[3][3];
    orc_rotation_d(q4, R);
    double k = scale_modifier;
    double s[3] = { k * scale3[0], k * scale3[1], k * scale3[2] };
    double M[3][3];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) M[a][c] = s[a] * R[c][a];
    double Sg[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Sg[r][c] = M[0][r] * M[0][c] + M[1][r] * M[1][c] + M[2][r] * M[2][c];
    double mx = fabs(Sg[0][0]);
    if (fabs(Sg[1][1]) > mx) mx = fabs(Sg[1][1]);
    if (fabs(Sg[2][2]) > mx) mx = fabs(Sg[2][2]);
    double eps = mx * 1e-5;
    double a = Sg[0][0] + eps, b = Sg[0][1], c = Sg[0][2], d = Sg[1][1] + eps, e = Sg[1][2], f = Sg[2][2] + eps;
    double A00 = d * f - e * e, A01 = c * e - b * f, A02 = b * e - c * d;
    double A11 = a * f - c * c, A12 = b * c - a * e, A22 = a * d - b * b;
    double det = a * A00 + b * A01 + c * A02;
    C6[0] = A00 / det; C6[1] = A01 / det; C6[2] = A02 / det;
    C6[3] = A11 / det; C6[4] = A12 / det; C6[5] = A22 / det;
}

void orc_conic_fp64_n(int P, const float *scales, const float *rots, float scale_modifier, double *out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; ++i) orc_conic_fp64(scales + 3 * i, rots + 4 * i, scale_modifier, out + 6 * i);
}

/*
 * Brute-force backward in double.  Pair membership: canonical fp32 predicate
 * (frozen).  Contributions: FP64 conic of the forward formula.  Per pair
 * (CR/backward.cu:257-298), skipping points with W <= 1e-2:
 *   dv += gV wgt / W;  a = gV (v - V) / W;  t = a + gW;  dw += t e;
 *   dquad = -wgt t / 2;  dmu += -2 dquad (C d);  dC += dquad (dx^2, 2dxdy, 2dxdz, dy^2, 2dydz, dz^2)
 * then the covariance chain.  V, W are the FP64 forward outputs.
 * Outputs (double): dmeans[3P], dscales[3P], drots[4P], dvalues[P], dweights[P],
 * and the per-Gaussian packed conic gradient dconic[6P] (may be NULL).
 */
void orc_backward(int P, int S, const float *means, const float *scales, const float *rots,
                  const float *values, const float *weights, float scale_modifier,
                  const float *conic, const float *box, const float *factor, const float *samples,
                  const double *V, const double *W, const double *gV, const double *gW, double *dmeans,
                  double *dscales, double *drots, double *dvalues, double *dweights,
                  double *dconic_out)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < P; ++i) {
        double C[6];
        orc_conic_fp64(scales + 3 * i, rots + 4 * i, scale_modifier, C);
        double dv = 0, dw = 0, dmu[3] = { 0, 0, 0 }, dc[6] = { 0, 0, 0, 0, 0, 0 };
        for (int s = 0; s < S; ++s) {
            const float *x = samples + 3 * s;
            float pw;
            if (!orc_pair_accept(box + 6 * i, means + 3 * i, conic + 6 * i, factor ? factor + 12 * i : NULL, x,
                                 &pw))
                continue;
            if (W[s] <= ORC_WEIGHT_CUTOFF) continue;
            double dx = (double)x[0] - means[3 * i + 0];
            double dy = (double)x[1] - means[3 * i + 1];
            double dz = (double)x[2] - means[3 * i + 2];
            double r0 = C[0] * dx + C[1] * dy + C[2] * dz;
            double r1 = C[1] * dx + C[3] * dy + C[4] * dz;
            double r2 = C[2] * dx + C[4] * dy + C[5] * dz;
            double quad = dx * r0 + dy * r1 + dz * r2;
            double e = exp(-0.5 * quad);
            double wgt = (double)weights[i] * e;
            double invW = 1.0 / W[s];
            dv += gV[s] * wgt * invW;
            double a = gV[s] * ((double)values[i] - V[s]) * invW;
            double t = a + gW[s];
            dw += t * e;
            double dquad = -0.5 * wgt * t;
            dmu[0] += -2.0 * dquad * r0;
            dmu[1] += -2.0 * dquad * r1;
            dmu[2] += -2.0 * dquad * r2;
            dc[0] += dquad * dx * dx;
            dc[1] += dquad * 2.0 * dx * dy;
            dc[2] += dquad * 2.0 * dx * dz;
            dc[3] += dquad * dy * dy;
            dc[4] += dquad * 2.0 * dy * dz;
            dc[5] += dquad * dz * dz;
        }
        dvalues[i] = dv;
        dweights[i] = dw;
        for (int c = 0; c < 3; ++c) dmeans[3 * i + c] = dmu[c];
        if (dconic_out)
            for (int c = 0; c < 6; ++c) dconic_out[6 * i + c] = dc[c];
        orc_cov_chain(scales + 3 * i, rots + 4 * i, scale_modifier, C, dc, dscales + 3 * i, drots + 4 * i);
    }
}

/* ------------------------------------------------------------------------- */
/* Morton codes: R/gaussian_renderer/__init__.py:45-58 (no clamp) and          */
/* R/train.py:291-315 (clamp to [0,1] first).  q = (int64)(x * (2^21-1)) with */
/* the product in float32 and truncation toward zero; NaN and out-of-range     */
/* conversions give INT64_MIN (the x86/CUDA conversion result).                */
/* ------------------------------------------------------------------------- */
static inline uint64_t orc_part1by2(int64_t n)
{
    uint64_t v = (uint64_t)n & 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

static inline int64_t orc_f32_to_i64_trunc(float v)
{
    if (v != v || v >= 9.2233720368547758e18f || v < -9.2233720368547758e18f) return INT64_MIN;
    return (int64_t)v;
}

void orc_morton3d_unit(int n, const float *xyz, int clamp, uint64_t *codes)
{
    const float scale = 2097151.0f; /* (1 << 21) - 1 */
    for (int i = 0; i < n; ++i) {
        uint64_t c = 0;
        for (int a = 0; a < 3; ++a) {
            float v = xyz[3 * i + a];
            if (clamp && v == v) v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            int64_t qv = orc_f32_to_i64_trunc(v * scale);
            c |= orc_part1by2(qv) << a;
        }
        codes[i] = c;
    }
}

/* ------------------------------------------------------------------------- */
/* Karras (2012) binary radix tree over n sorted 64-bit keys.  Equal keys are   */
/* ordered by index: delta(i,j) = clz(k_i ^ k_j), or 64 + clz32(i ^ j) when     */
/* equal.  Internal node i in [0, n-2]; child encoding: >= 0 internal node,     */
/* < 0 leaf ~j.  Also returns the [first,last] leaf range of every internal.   */
/* ------------------------------------------------------------------------- */
static inline int orc_clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
static inline int orc_clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

static inline int orc_delta(const uint64_t *k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    if (k[i] == k[j]) return 64 + orc_clz32((uint32_t)i ^ (uint32_t)j);
    return orc_clz64(k[i] ^ k[j]);
}

void orc_karras_build(int n, const uint64_t *keys, int32_t *left, int32_t *right,
                      int32_t *first, int32_t *last)
{
    for (int i = 0; i < n - 1; ++i) {
        int d = (orc_delta(keys, n, i, i + 1) - orc_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = orc_delta(keys, n, i, i - d);
        int lmax = 2;
        while (orc_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (orc_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = orc_delta(keys, n, i, j);
        int s = 0;
        int t = l;
        do {
            t = (t + 1) / 2;
            if (orc_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        int gamma = i + s * d + (d < 0 ? d : 0);
        int lo = i < j ? i : j, hi = i < j ? j : i;
        left[i] = (lo == gamma) ? ~gamma : gamma;
        right[i] = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
        first[i] = lo;
        last[i] = hi;
    }
}

/* ------------------------------------------------------------------------- */
/* Plain fp32 CPU baseline: the reference formula on CPU, brute force with the  */
/* box predicate, no compensation: the reference formula run on the CPU.       */
/* Evaluates points [0, S) of `samples` against all P Gaussians; forward then   */
/* backward with upstream dL/dV = gV, dL/dW = gW.  Per-Gaussian gradients are   */
/* accumulated per thread and summed (fp32 packed-conic gradient, no chain).    */
/* Returns the number of accepted pairs.                                        */
/* ------------------------------------------------------------------------- */
long long orc_cpu_plain_fwdbwd(int P, int S, const float *means, const float *values,
                               const float *weights, const float *conic, const float *box,
                               const float *samples, const float *gV, const float *gW,
                               float *V_out, float *W_out, float *grad11)
{
    long long pairs = 0;
#pragma omp parallel reduction(+ : pairs)
    {
        float *acc = (float *)calloc((size_t)P * 11, sizeof(float));
#pragma omp for schedule(dynamic, 32)
        for (int s = 0; s < S; ++s) {
            const float *x = samples + 3 * s;
            float W = 0.f, U = 0.f;
            for (int i = 0; i < P; ++i) {
                float pw;
                if (!orc_pair_accept(box + 6 * i, means + 3 * i, conic + 6 * i, NULL, x, &pw)) continue;
                float wgt = weights[i] * expf(pw);
                W += wgt;
                U += wgt * values[i];
                ++pairs;
            }
            float V;
            if (W <= ORC_WEIGHT_CUTOFF) { V = -1.f; W = 0.f; } else { V = U / W; }
            V_out[s] = V;
            W_out[s] = W;
            if (W <= ORC_WEIGHT_CUTOFF) continue;
            float invW = 1.f / W;
            for (int i = 0; i < P; ++i) {
                float pw;
                if (!orc_pair_accept(box + 6 * i, means + 3 * i, conic + 6 * i, NULL, x, &pw)) continue;
                const float *cn = conic + 6 * i;
                float dx = x[0] - means[3 * i], dy = x[1] - means[3 * i + 1], dz = x[2] - means[3 * i + 2];
                float e = expf(pw), wgt = weights[i] * e;
                float a = gV[s] * (values[i] - V) * invW;
                float t = a + gW[s];
                float dq = -0.5f * wgt * t;
                float *g = acc + 11 * (size_t)i;
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
            }
        }
#pragma omp critical
        {
            if (grad11)
                for (size_t k = 0; k < (size_t)P * 11; ++k) grad11[k] += acc[k];
        }
        free(acc);
    }
    return pairs;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
