// exact.h -- exact sign of small polynomial predicates on float32 coordinates.
//
// Method (Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast Robust
// Geometric Predicates", 1997): a double-precision evaluation with a forward error
// bound decides almost every case; otherwise the predicate is expanded into a sum
// of products of input coordinates, every product is made exact with TwoProduct
// (fma), and the sign of the exact sum is read off a nonoverlapping expansion
// built with TwoSum (grow-expansion with zero elimination).
//
// Products of two float32 values are exact in double; products of three are
// exact as a (hi, lo) pair as long as no partial product underflows (inputs with
// |x| >= 2^-100 or x == 0).
#pragma once
#include <hip/hip_runtime.h>

namespace gvc {

__host__ __device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    double bv = s - a;
    double av = s - bv;
    e = (a - av) + (b - bv);
}

__host__ __device__ __forceinline__ void two_prod(double a, double b, double& p, double& e)
{
    p = a * b;
    e = __builtin_fma(a, b, -p);
}

// Sign of the exact sum of n doubles.
template <int N>
__host__ __device__ inline int exact_sum_sign(const double (&t)[N], int n)
{
    double h[N];
    int m = 0;
    for (int k = 0; k < n; ++k) {
        double q = t[k];
        int mm = 0;
        for (int j = 0; j < m; ++j) {
            double s, e;
            two_sum(q, h[j], s, e);
            q = s;
            if (e != 0.0) h[mm++] = e;
        }
        if (q != 0.0) h[mm++] = q;
        m = mm;
    }
    // the largest-magnitude component (last) carries the sign
    if (m == 0) return 0;
    return h[m - 1] > 0.0 ? 1 : -1;
}

// orient3d(a, b, c, d) = det[a-d; b-d; c-d] (the signed volume x 6).
__host__ __device__ inline int orient3d_sign(const float* a, const float* b, const float* c, const float* d,
                                             double* value_out)
{
    const double adx = (double)a[0] - d[0], ady = (double)a[1] - d[1], adz = (double)a[2] - d[2];
    const double bdx = (double)b[0] - d[0], bdy = (double)b[1] - d[1], bdz = (double)b[2] - d[2];
    const double cdx = (double)c[0] - d[0], cdy = (double)c[1] - d[1], cdz = (double)c[2] - d[2];
    const double bc = bdx * cdy - bdy * cdx, ca = cdx * ady - cdy * adx, ab = adx * bdy - ady * bdx;
    const double det = adz * bc + bdz * ca + cdz * ab;
    if (value_out) *value_out = det;
    const double perm = (fabs(bdx * cdy) + fabs(bdy * cdx)) * fabs(adz) + (fabs(cdx * ady) + fabs(cdy * adx)) * fabs(bdz) +
                        (fabs(adx * bdy) + fabs(ady * bdx)) * fabs(cdz);
    // Shewchuk's o3derrboundA = (7 + 56 eps) eps, eps = 2^-53
    const double bound = 7.7715611723761082e-16 * perm;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    // exact: the 4x4 determinant det[[a,1],[b,1],[c,1],[d,1]] expanded into the 24
    // signed triple products x_i y_j z_k over ordered triples of distinct points.
    const float* P[4] = { a, b, c, d };
    double terms[48];
    int n = 0;
    // det[a-d, b-d, c-d] = det4 of rows (p, 1): sum over permutations (i,j,k,l) of
    // sign * x_i y_j z_k (the 1-column picks l).  sign from the permutation parity.
    const int perms[24][4] = { { 0, 1, 2, 3 }, { 0, 1, 3, 2 }, { 0, 2, 1, 3 }, { 0, 2, 3, 1 }, { 0, 3, 1, 2 },
                               { 0, 3, 2, 1 }, { 1, 0, 2, 3 }, { 1, 0, 3, 2 }, { 1, 2, 0, 3 }, { 1, 2, 3, 0 },
                               { 1, 3, 0, 2 }, { 1, 3, 2, 0 }, { 2, 0, 1, 3 }, { 2, 0, 3, 1 }, { 2, 1, 0, 3 },
                               { 2, 1, 3, 0 }, { 2, 3, 0, 1 }, { 2, 3, 1, 0 }, { 3, 0, 1, 2 }, { 3, 0, 2, 1 },
                               { 3, 1, 0, 2 }, { 3, 1, 2, 0 }, { 3, 2, 0, 1 }, { 3, 2, 1, 0 } };
    for (int p = 0; p < 24; ++p) {
        const int i = perms[p][0], j = perms[p][1], k = perms[p][2], l = perms[p][3];
        // parity by counting inversions
        int inv = 0;
        const int q[4] = { i, j, k, l };
        for (int u = 0; u < 4; ++u)
            for (int v = u + 1; v < 4; ++v) inv += q[u] > q[v];
        const double xy = (double)P[i][0] * (double)P[j][1];  // exact
        double hi, lo;
        two_prod(xy, (double)P[k][2], hi, lo);
        const double s = (inv & 1) ? -1.0 : 1.0;
        terms[n++] = s * hi;
        terms[n++] = s * lo;
    }
    // det4 of the rows (a,1), (b,1), (c,1), (d,1) equals det[a-d; b-d; c-d]
    // (subtract the last row, expand along the last column).
    return exact_sum_sign(terms, n);
}

}  // namespace gvc
