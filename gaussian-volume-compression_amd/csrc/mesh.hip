// mesh.hip -- the mesh sampler: exact point location in tetrahedral meshes with
// barycentric interpolation, and trilinear probing of uniform grids.
//
// Replaces GS/sample_mesh.cu (gpu_sampleu / gpu_sample over Viskores Probe,
// :155-167, :380-390; GS/ = gpu_sampling/).  Differences by design: device
// pointers in and out (no host staging), int32 cell ids are returned, nothing is
// printed, and location is exact:
//   point p is in the closed tet T = (v0,v1,v2,v3) iff, for every face, orient3d
//   of the face and p is zero or has the sign of orient3d(T) (degenerate tets
//   hold nothing); among several containing tets the lowest index wins.  As
//   Probe's cells are closed (parametric coordinates within [0, 1] up to a
//   tolerance), a point on the mesh boundary is inside -- the reference's own
//   evaluation grid clips samples onto the boundary planes -- and a point on a
//   face, edge or vertex shared by several tets takes the lowest of them.
//
// Locator: a uniform grid over the mesh's bounding box (about one cell per tet),
// each cell split 3 x 3 x 3 into sub-cells.  A cell lists every tet that
// may meet it as an 8-byte entry: the tet id and a 27-bit mask of the sub-cells
// the tet may meet -- a sub-cell (expanded by the rounding of the point-to-cell
// mapping) is dropped when it misses the tet's bounding box or lies strictly
// outside one of the tet's face planes (double, with a rounding allowance).  So
// most candidates are rejected from the entry alone, and a cell that the tet's
// box reaches but the tet does not gets no entry.  Each cell's list starts on a
// 128-byte line and is padded with empty entries (mask 0) to whole lines of 16
// entries, so a sample reads its cell's list one whole line per load.  The
// survivors are tested with a per-tet barycentric matrix M (computed in double,
// stored as float) -- a 48-byte filter record {M, v3}; the vertex ids sit in a
// separate array, read only by the exact predicate and vertex interpolation:
// b = M (p - v3), evaluated in double; a decision is taken only when every |b_i|
// exceeds a margin e_i = 1e-6 sum_j |M_ij d_j| + 2e-9 |M|_1 |d|_1 with
// d = p - v3 (and e_3 = e_0 + e_1 + e_2 + 1e-15 for b_3 = 1 - b_0 - b_1 - b_2):
// the first term covers the float rounding of each M_ij (2^-24, a 16x margin),
// the second the error of the double inverse, |dM_ij| <= 6 u cond_F(T) |M|_F
// with u = 2^-53 and cond_F(T) <= 1e6 (a 300x margin).  Otherwise -- and
// always for tets with cond_F(T) > 1e6 -- the exact predicate decides.  A cell's
// entries are sorted by tet index, so the scan stops at the first tet that
// contains the point (lowest index wins) or at the first undecided one (the
// exact pass redoes it).
#include <math.h>
#include <string.h>

#include "exact.h"
#include "gvc_device.h"
#include "gvc_internal.h"

namespace gvc {

struct alignas(16) TetRec {
    float m[9];   // rows of inverse([v0-v3 | v1-v3 | v2-v3]); m[0] = NaN: exact test only
    float v3[3];
};
static_assert(sizeof(TetRec) == 48, "TetRec layout");

// Entries per 128-byte line of a cell's list; an empty entry has tet id
// kNoTet and mask 0 (never a candidate).
constexpr int kLineEntries = 16;
constexpr uint32_t kNoTet = 0xffffffffu;

struct alignas(16) TetGridHeader {
    float lo[3];
    float cell[3];  // cell size per axis
    int dims[3];
    int n_tets;
    int n_points;
    float hi[3];    // bounding box of the mesh points: no tet reaches outside it
    unsigned long long n_entries;
    unsigned long long capacity;
};

namespace {

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Inverse of [v0-v3 | v1-v3 | v2-v3] in double from the float vertices (rows give
// b_0..b_2 = inv (p - v3)); ok = finite, non-zero determinant and
// cond_F = |T|_F |inv|_F <= 1e6.
__device__ void tet_inverse(const float* __restrict__ pts, const int v[4], double inv[9], bool& ok)
{
    double V[4][3];
    for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) V[k][a] = pts[3 * v[k] + a];
    // T columns: v0-v3, v1-v3, v2-v3
    double T[3][3];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) T[a][c] = V[c][a] - V[3][a];
    const double A00 = T[1][1] * T[2][2] - T[1][2] * T[2][1];
    const double A01 = T[0][2] * T[2][1] - T[0][1] * T[2][2];
    const double A02 = T[0][1] * T[1][2] - T[0][2] * T[1][1];
    const double A10 = T[1][2] * T[2][0] - T[1][0] * T[2][2];
    const double A11 = T[0][0] * T[2][2] - T[0][2] * T[2][0];
    const double A12 = T[0][2] * T[1][0] - T[0][0] * T[1][2];
    const double A20 = T[1][0] * T[2][1] - T[1][1] * T[2][0];
    const double A21 = T[0][1] * T[2][0] - T[0][0] * T[2][1];
    const double A22 = T[0][0] * T[1][1] - T[0][1] * T[1][0];
    const double det = T[0][0] * A00 + T[0][1] * A10 + T[0][2] * A20;
    inv[0] = A00 / det; inv[1] = A01 / det; inv[2] = A02 / det;
    inv[3] = A10 / det; inv[4] = A11 / det; inv[5] = A12 / det;
    inv[6] = A20 / det; inv[7] = A21 / det; inv[8] = A22 / det;
    double nT = 0.0, nI = 0.0;
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) nT += T[a][c] * T[a][c];
    for (int k = 0; k < 9; ++k) nI += inv[k] * inv[k];
    ok = det != 0.0 && isfinite(det) && sqrt(nT * nI) <= 1e6;
}

// Per-tet barycentric matrix (float copy of tet_inverse) and bounding box.
__global__ void tet_records_kernel(const float* __restrict__ pts, const int* __restrict__ tets, int C,
                                   TetRec* __restrict__ recs, int4* __restrict__ verts, float* __restrict__ boxes)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C) return;
    TetRec r;
    const int v[4] = { tets[4 * t], tets[4 * t + 1], tets[4 * t + 2], tets[4 * t + 3] };
    for (int a = 0; a < 3; ++a) {
        r.v3[a] = pts[3 * v[3] + a];
        float lo = pts[3 * v[0] + a], hi = lo;
        for (int k = 1; k < 4; ++k) {
            float x = pts[3 * v[k] + a];
            lo = fminf(lo, x);
            hi = fmaxf(hi, x);
        }
        boxes[6 * t + a] = lo;
        boxes[6 * t + 3 + a] = hi;
    }
    double inv[9];
    bool ok;
    tet_inverse(pts, v, inv, ok);
    for (int k = 0; k < 9; ++k) r.m[k] = ok ? (float)inv[k] : __builtin_nanf("");
    recs[t] = r;
    verts[t] = make_int4(v[0], v[1], v[2], v[3]);
}

// Bound values (gvc_tet_locator_bind_values): per tet the affine interpolant
// value(p) = w3 + g . (p - v3), g_j = sum_k inv_kj (w_k - w3), in double.  Tets
// whose matrix is not kept (cond_F > 1e6, degenerate) store the four vertex
// values instead and are interpolated from orientation volumes.
__global__ void tet_coef_kernel(const float* __restrict__ pts, const TetRec* __restrict__ recs,
                                const int4* __restrict__ verts, const double* __restrict__ values, int C,
                                double4* __restrict__ coef)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C) return;
    const float m0 = recs[t].m[0];
    const int4 vi = verts[t];
    const int v[4] = { vi.x, vi.y, vi.z, vi.w };
    const double w0 = values[v[0]], w1 = values[v[1]], w2 = values[v[2]], w3 = values[v[3]];
    if (m0 != m0) {
        coef[t] = make_double4(w0, w1, w2, w3);
        return;
    }
    double inv[9];
    bool ok;
    tet_inverse(pts, v, inv, ok);
    const double d0 = w0 - w3, d1 = w1 - w3, d2 = w2 - w3;
    coef[t] = make_double4(w3, inv[0] * d0 + inv[3] * d1 + inv[6] * d2, inv[1] * d0 + inv[4] * d1 + inv[7] * d2,
                           inv[2] * d0 + inv[5] * d1 + inv[8] * d2);
}

__device__ __forceinline__ int cell_coord(float x, float lo, float cell, int dim)
{
    float t = (x - lo) / cell;
    int c = (int)floorf(t);
    return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

// Sub-cells: each grid cell is split 3 x 3 x 3; an entry carries a 27-bit mask
// of the sub-cells its tet may meet (bit s0 + 3 s1 + 9 s2).  A point's sub-cell
// comes from its grid coordinate u = (x - lo) / cell (float, as for the cell):
// s = floor(3 (u - c)) clamped to [0, 2].  Float rounding moves u by at most
// 2^-23 dims + 2^-24 (the subtraction, the division, the scaling by 3), so every
// sub-cell is tested on the build side expanded by kSubMargin (dims + 2) cells
// (an 8x margin).
constexpr int kSub = 3;
constexpr double kSubMargin = 0x1p-20;

__device__ __forceinline__ int sub_index(float u, int c)
{
    const int s = (int)floorf((u - (float)c) * (float)kSub);
    return s < 0 ? 0 : (s >= kSub ? kSub - 1 : s);
}

// Inward face planes of a tet from its float vertices, in double: every point x
// of the closed tet has n_f . x + o_f >= 0 for f = 0..3 (face f is opposite
// vertex f).  A face is used to prune only if the side of the opposite vertex is
// certain: |n_f . (v_f - a_f)| above a bound 1e3 x larger than the rounding of
// the cross product and dot product (~ 8 u sum |n_i| |d_i| with u = 2^-53);
// otherwise (nearly degenerate tets) the tet is pruned by its bounding box only.
// tol_f covers the rounding of n_f and of the evaluation at any point within
// `reach` of the face's first vertex (again with a margin of ~1e3).
struct FacePlanes {
    double n[4][3];
    double o[4];
    double tol[4];
    bool use;
};

__device__ void face_planes(const float* __restrict__ pts, const int v[4], double reach, FacePlanes& F)
{
    double V[4][3];
    for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) V[k][a] = (double)pts[3 * v[k] + a];
    F.use = true;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int i0 = f == 0 ? 1 : 0, i1 = f <= 1 ? 2 : 1, i2 = f <= 2 ? 3 : 2;
        double u[3], w[3], d[3];
        for (int a = 0; a < 3; ++a) {
            u[a] = V[i1][a] - V[i0][a];
            w[a] = V[i2][a] - V[i0][a];
            d[a] = V[f][a] - V[i0][a];
        }
        double n[3] = { u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0] };
        const double m[3] = { fabs(u[1] * w[2]) + fabs(u[2] * w[1]), fabs(u[2] * w[0]) + fabs(u[0] * w[2]),
                              fabs(u[0] * w[1]) + fabs(u[1] * w[0]) };
        const double side = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
        const double bound = 1e-12 * (m[0] * fabs(d[0]) + m[1] * fabs(d[1]) + m[2] * fabs(d[2]));
        if (!(fabs(side) > bound)) F.use = false;
        const double sg = side < 0.0 ? -1.0 : 1.0;
        for (int a = 0; a < 3; ++a) F.n[f][a] = sg * n[a];
        F.o[f] = -(F.n[f][0] * V[i0][0] + F.n[f][1] * V[i0][1] + F.n[f][2] * V[i0][2]);
        F.tol[f] = 1e-12 * (m[0] + m[1] + m[2]) * (reach + fabs(V[i0][0]) + fabs(V[i0][1]) + fabs(V[i0][2]));
    }
}

// Sub-cell mask of a tet in grid cell ci: bit set unless the expanded sub-cell
// misses the tet's bounding box or lies strictly outside one of its face planes
// (a conservative test: a set bit may still miss the tet).
__device__ __forceinline__ uint32_t cell_mask(const TetGridHeader* __restrict__ h, const int ci[3],
                                              const float* __restrict__ b, const FacePlanes& F)
{
    // per axis: sub-cell centres, half-width (with the margin) and the
    // sub-cells that meet the tet's bounding box
    double ctr[3][kSub], half[3];
    uint32_t axis_ok[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // margin: the float rounding of the point's grid coordinate, plus the
        // double rounding of the sub-cell bounds themselves
        const double cs = (double)h->cell[a], lo = (double)h->lo[a];
        const double mg = cs * kSubMargin * (double)(h->dims[a] + 2) +
                          1e-15 * (fabs(lo) + fabs((double)h->hi[a]) + cs * h->dims[a]);
        half[a] = cs * (0.5 / kSub) + mg;
        axis_ok[a] = 0u;
#pragma unroll
        for (int s = 0; s < kSub; ++s) {
            const double xlo = lo + cs * (ci[a] + (double)s / kSub) - mg;
            const double xhi = lo + cs * (ci[a] + (double)(s + 1) / kSub) + mg;
            ctr[a][s] = lo + cs * (ci[a] + (s + 0.5) / kSub);
            if (xhi >= (double)b[a] && xlo <= (double)b[3 + a]) axis_ok[a] |= 1u << s;
        }
    }
    uint32_t mask = 0u;
#pragma unroll
    for (int s2 = 0; s2 < kSub; ++s2)
#pragma unroll
        for (int s1 = 0; s1 < kSub; ++s1)
#pragma unroll
            for (int s0 = 0; s0 < kSub; ++s0)
                if ((axis_ok[0] >> s0) & (axis_ok[1] >> s1) & (axis_ok[2] >> s2) & 1u)
                    mask |= 1u << (s0 + kSub * (s1 + kSub * s2));
    if (!F.use || mask == 0u) return mask;
    // face by face: the sub-cells whose highest point over the face's plane is
    // not below -tol (n . ctr + |n| . half + o, separable per axis)
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        double tx[kSub], ty[kSub], tz[kSub];
#pragma unroll
        for (int s = 0; s < kSub; ++s) {
            tx[s] = F.n[f][0] * ctr[0][s] + fabs(F.n[f][0]) * half[0];
            ty[s] = F.n[f][1] * ctr[1][s] + fabs(F.n[f][1]) * half[1];
            tz[s] = F.n[f][2] * ctr[2][s] + fabs(F.n[f][2]) * half[2] + F.o[f];
        }
        uint32_t keep = 0u;
#pragma unroll
        for (int s2 = 0; s2 < kSub; ++s2)
#pragma unroll
            for (int s1 = 0; s1 < kSub; ++s1)
#pragma unroll
                for (int s0 = 0; s0 < kSub; ++s0)
                    if ((tx[s0] + ty[s1]) + tz[s2] >= -F.tol[f]) keep |= 1u << (s0 + kSub * (s1 + kSub * s2));
        mask &= keep;
    }
    return mask;
}

// Face planes of tet t, with the reach of every point tested against them: the
// tet's bounding-box diagonal plus two cells per axis.
__device__ __forceinline__ void tet_planes_for(const float* __restrict__ pts, const int* __restrict__ tets, int t,
                                               const float* __restrict__ b, const TetGridHeader* __restrict__ h,
                                               FacePlanes& F)
{
    const int v[4] = { tets[4 * t], tets[4 * t + 1], tets[4 * t + 2], tets[4 * t + 3] };
    double reach = 0.0;
    for (int a = 0; a < 3; ++a) reach += ((double)b[3 + a] - b[a]) + 2.0 * (double)h->cell[a];
    face_planes(pts, v, reach, F);
}

// 64-bit total of the per-cell entry counts (the scan that turns them into
// offsets is 32-bit, so the build refuses totals near 2^32; below that bound the
// scans' index steps cannot wrap).
__global__ __launch_bounds__(256) void count_total_kernel(const uint32_t* __restrict__ counts, size_t n,
                                                          unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long part[256];
    unsigned long long v = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) v += counts[i];
    part[threadIdx.x] = v;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part[0]) atomicAdd(total, part[0]);
}

// The build kernels keep a tet's four face planes and the sub-cell geometry in
// registers (about 240 VGPRs, no scratch): launched with kBuildThreads-thread
// blocks.
constexpr int kBuildThreads = 64;

__global__ __launch_bounds__(kBuildThreads) void tet_count_kernel(const float* __restrict__ pts, const int* __restrict__ tets,
                                 const float* __restrict__ boxes, int C, const TetGridHeader* __restrict__ h,
                                 uint32_t* __restrict__ counts)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C) return;
    const float* b = boxes + 6 * t;
    if (!(b[0] == b[0] && b[1] == b[1] && b[2] == b[2])) return;
    FacePlanes F;
    tet_planes_for(pts, tets, t, b, h, F);
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        c0[a] = cell_coord(b[a], h->lo[a], h->cell[a], h->dims[a]);
        c1[a] = cell_coord(b[3 + a], h->lo[a], h->cell[a], h->dims[a]);
    }
    for (int k = c0[2]; k <= c1[2]; ++k)
        for (int j = c0[1]; j <= c1[1]; ++j)
            for (int i = c0[0]; i <= c1[0]; ++i) {
                const int ci[3] = { i, j, k };
                if (cell_mask(h, ci, b, F) == 0u) continue;
                atomicAdd(&counts[(size_t)i + (size_t)h->dims[0] * ((size_t)j + (size_t)h->dims[1] * k)], 1u);
            }
}

// Each cell's entry count rounded up to whole lines of kLineEntries.
__global__ __launch_bounds__(256) void pad_counts_kernel(uint32_t* __restrict__ counts, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) counts[i] = (counts[i] + (kLineEntries - 1)) & ~(uint32_t)(kLineEntries - 1);
}

// Every entry slot empty: tet id kNoTet, mask 0.
__global__ __launch_bounds__(256) void empty_entries_kernel(uint2* __restrict__ entries, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        entries[i] = make_uint2(kNoTet, 0u);
}

// entry = (tet id, sub-cell mask of the tet in the entry's cell); cells whose
// mask is empty get no entry (the same test as tet_count_kernel).  A cell's
// entries fill the front of its padded list (cursor = the cell's entry count).
__global__ __launch_bounds__(kBuildThreads) void tet_fill_kernel(const float* __restrict__ pts, const int* __restrict__ tets,
                                const float* __restrict__ boxes, int C, const TetGridHeader* __restrict__ h,
                                const uint32_t* __restrict__ starts, uint32_t* __restrict__ cursor,
                                uint2* __restrict__ entries, unsigned long long capacity)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C) return;
    const float* b = boxes + 6 * t;
    if (!(b[0] == b[0] && b[1] == b[1] && b[2] == b[2])) return;
    FacePlanes F;
    tet_planes_for(pts, tets, t, b, h, F);
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        c0[a] = cell_coord(b[a], h->lo[a], h->cell[a], h->dims[a]);
        c1[a] = cell_coord(b[3 + a], h->lo[a], h->cell[a], h->dims[a]);
    }
    for (int k = c0[2]; k <= c1[2]; ++k)
        for (int j = c0[1]; j <= c1[1]; ++j)
            for (int i = c0[0]; i <= c1[0]; ++i) {
                const int ci[3] = { i, j, k };
                const uint32_t q = cell_mask(h, ci, b, F);
                if (q == 0u) continue;
                const size_t g = (size_t)i + (size_t)h->dims[0] * ((size_t)j + (size_t)h->dims[1] * k);
                const unsigned long long pos = (unsigned long long)starts[g] + atomicAdd(&cursor[g], 1u);
                if (pos < capacity) entries[pos] = make_uint2((uint32_t)t, q);
            }
}

// Sorts each cell's entries (the first counts[g] of its list; the empty padding
// stays behind them) by tet index, so the locate kernel can stop at the first
// containing tet.  One wavefront per cell: up to kSortCap entries are staged
// in LDS and each lane places its entries at their rank (count of smaller tet
// indices; a tet appears at most once per cell); longer cells fall back to an
// insertion sort by one lane.
constexpr int kSortCap = 512;

__global__ __launch_bounds__(256) void tet_sort_cells_kernel(const uint32_t* __restrict__ starts,
                                                             const uint32_t* __restrict__ counts, size_t n_grid,
                                                             uint2* __restrict__ entries)
{
    __shared__ uint2 buf[4][kSortCap];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t g = (size_t)blockIdx.x * 4 + wid;
    if (g >= n_grid) return;
    const uint32_t b = starts[g], n = counts[g], e = b + n;
    if (n <= 1) return;
    if (n <= (uint32_t)kSortCap) {
        for (uint32_t i = lane; i < n; i += 64) buf[wid][i] = entries[b + i];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = lane; i < n; i += 64) {
            const uint32_t key = buf[wid][i].x;
            uint32_t rank = 0;
            for (uint32_t j = 0; j < n; ++j) rank += buf[wid][j].x < key ? 1u : 0u;
            entries[b + rank] = buf[wid][i];
        }
        return;
    }
    if (lane != 0) return;
    for (uint32_t i = b + 1; i < e; ++i) {
        const uint2 v = entries[i];
        uint32_t j = i;
        while (j > b && entries[j - 1].x > v.x) {
            entries[j] = entries[j - 1];
            --j;
        }
        entries[j] = v;
    }
}

// Exact containment of p in the closed tet with vertex ids v.
__device__ __noinline__ bool tet_contains_exact(const float* __restrict__ pts, const int4 v, const float* p)
{
    const float* V[4] = { pts + 3 * v.x, pts + 3 * v.y, pts + 3 * v.z, pts + 3 * v.w };
    const int sT = orient3d_sign(V[0], V[1], V[2], V[3], nullptr);
    if (sT == 0) return false;
    for (int i = 0; i < 4; ++i) {
        // face opposite i, remaining vertices in increasing order; o_i = s_i orient(F, p)
        const float* F[3];
        int m = 0;
        for (int k = 0; k < 4; ++k)
            if (k != i) F[m++] = V[k];
        const int si = ((3 - i) & 1) ? -1 : 1;
        const int sg = orient3d_sign(F[0], F[1], F[2], p, nullptr);
        if (sg != 0 && si * sg != sT) return false;
    }
    return true;
}

__device__ __forceinline__ double orient3d_value(const float* a, const float* b, const float* c, const float* d)
{
    const double adx = (double)a[0] - d[0], ady = (double)a[1] - d[1], adz = (double)a[2] - d[2];
    const double bdx = (double)b[0] - d[0], bdy = (double)b[1] - d[1], bdz = (double)b[2] - d[2];
    const double cdx = (double)c[0] - d[0], cdy = (double)c[1] - d[1], cdz = (double)c[2] - d[2];
    return adz * (bdx * cdy - bdy * cdx) + bdz * (cdx * ady - cdy * adx) + cdz * (adx * bdy - ady * bdx);
}

// Barycentric coordinates of p in the tet with vertex ids v from double
// orientation volumes.
__device__ void tet_bary(const float* __restrict__ pts, const int4 v, const float* p, double b[4])
{
    const float* V[4] = { pts + 3 * v.x, pts + 3 * v.y, pts + 3 * v.z, pts + 3 * v.w };
    const double vol = orient3d_value(V[0], V[1], V[2], V[3]);
    b[0] = -orient3d_value(V[1], V[2], V[3], p) / vol;
    b[1] = orient3d_value(V[0], V[2], V[3], p) / vol;
    b[2] = -orient3d_value(V[0], V[1], V[3], p) / vol;
    b[3] = orient3d_value(V[0], V[1], V[2], p) / vol;
}

// 1 inside, 0 outside, -1 undecided (exact test needed).
__device__ __forceinline__ int tet_filter(const TetRec& r, const float* p)
{
    if (r.m[0] != r.m[0]) return -1;
    const double dx = (double)p[0] - r.v3[0], dy = (double)p[1] - r.v3[1], dz = (double)p[2] - r.v3[2];
    float l1m = 0.0f;
    for (int k = 0; k < 9; ++k) l1m += fabsf(r.m[k]);
    const double eabs = 2e-9 * (double)l1m * (fabs(dx) + fabs(dy) + fabs(dz));
    double b[4], e[4];
    for (int i = 0; i < 3; ++i) {
        const double t0 = (double)r.m[3 * i] * dx, t1 = (double)r.m[3 * i + 1] * dy, t2 = (double)r.m[3 * i + 2] * dz;
        b[i] = t0 + t1 + t2;
        e[i] = 1e-6 * (fabs(t0) + fabs(t1) + fabs(t2)) + eabs;
    }
    b[3] = 1.0 - b[0] - b[1] - b[2];
    e[3] = e[0] + e[1] + e[2] + 1e-15;
    bool all_in = true;
    for (int i = 0; i < 4; ++i) {
        if (b[i] < -e[i]) return 0;
        all_in = all_in && (b[i] > e[i]);
    }
    return all_in ? 1 : -1;
}

// A sample the filter pass leaves undecided holds cell_ids[s] = -2 - i, i = its
// position in that pass's order (decided samples hold >= -1).
__device__ __forceinline__ int mark_undecided(int i) { return -2 - i; }

constexpr int kLocateThreads = 256;

// EXACT = false (filter pass): threads follow the sample set's Morton order when
// one is given (sorted != nullptr; neighbouring threads share grid cells and
// tets; results go to the original index carried in .w), else the caller's
// order; a sample whose first undecided candidate comes before any decided one
// is marked (mark_undecided, no result written).  EXACT = true: threads follow
// the caller's order (coalesced reads of the marks); only marked samples are
// redone, undecided candidates by the exact predicate.
// A lane reads its cell's list one 128-byte line (16 entries) per load, keeps
// the entries whose sub-cell mask holds the point's sub-cell (in tet order, in
// LDS) and tests those with the filter, the next candidate's record in flight,
// stopping at the first decided containment; so a wave runs the filter once per
// candidate of its busiest lane.  Values: coef != nullptr (bound values)
// evaluates the tet's affine interpolant (one 32-byte read); otherwise
// barycentric coordinates from orientation volumes weight values[] at the four
// vertices.
template <bool EXACT>
__global__ __launch_bounds__(kLocateThreads) void tet_locate_kernel(
    const float* __restrict__ pts, const TetRec* __restrict__ recs, const int4* __restrict__ verts,
    const TetGridHeader* __restrict__ h, const uint32_t* __restrict__ starts, const uint2* __restrict__ entries,
    const double* __restrict__ values, const double4* __restrict__ coef, int S, const float* __restrict__ samples,
    const float4* __restrict__ sorted, int* __restrict__ cell_ids, double* __restrict__ out_values,
    float* __restrict__ bary_out)
{
    __shared__ int cand[kLineEntries][kLocateThreads];
    const int tid = threadIdx.x;
    const int i0 = (int)xcd_block(blockIdx.x, gridDim.x) * kLocateThreads + tid;
    if (i0 >= S) return;
    int s = i0, i = i0;
    if (EXACT) {
        const int mark = cell_ids[s];
        if (mark > -2) return;
        i = -2 - mark;
    }
    float p[3];
    if (sorted) {
        const float4 q = sorted[i];
        p[0] = q.x; p[1] = q.y; p[2] = q.z;
        if (!EXACT) {
            s = __builtin_bit_cast(int, q.w);
            if (s < 0) return;
        }
    } else {
        p[0] = samples[3 * i]; p[1] = samples[3 * i + 1]; p[2] = samples[3 * i + 2];
    }
    int best = -1;
    TetRec br;   // best's record
    bool undecided = false;
    bool inside_grid = true;
    int c[3], sb = 0;
    for (int a = 0, w = 1; a < 3; ++a, w *= kSub) {
        if (!(p[a] >= h->lo[a] && p[a] <= h->hi[a])) inside_grid = false;
        const float t = (p[a] - h->lo[a]) / h->cell[a];
        int ci = inside_grid ? (int)floorf(t) : 0;
        c[a] = ci < 0 ? 0 : (ci >= h->dims[a] ? h->dims[a] - 1 : ci);
        sb += w * sub_index(t, c[a]);
    }
    const uint32_t sbit = 1u << sb;
    if (inside_grid) {
        // entries of a cell are sorted by tet index: the first tet that contains
        // p is the answer (lowest index wins), unless an earlier one is undecided
        const size_t g = (size_t)c[0] + (size_t)h->dims[0] * ((size_t)c[1] + (size_t)h->dims[1] * c[2]);
        uint32_t k = starts[g];
        const uint32_t e = starts[g + 1];
        bool done = false;
        while (!done && k < e) {
            const uint4* __restrict__ line = reinterpret_cast<const uint4*>(entries + k);
            uint4 en[kLineEntries / 2];
#pragma unroll
            for (int u = 0; u < kLineEntries / 2; ++u) en[u] = line[u];
            k += kLineEntries;
            int n = 0;
#pragma unroll
            for (int u = 0; u < kLineEntries / 2; ++u) {
                if (en[u].y & sbit) cand[n++][tid] = (int)en[u].x;
                if (en[u].w & sbit) cand[n++][tid] = (int)en[u].z;
            }
            TetRec rn;
            if (n > 0) rn = recs[cand[0][tid]];
            for (int j = 0; j < n; ++j) {
                const int t = cand[j][tid];
                const TetRec r = rn;
                if (j + 1 < n) rn = recs[cand[j + 1][tid]];
                const int f = tet_filter(r, p);
                if (f == 1) {
                    best = t;
                    br = r;
                    done = true;
                    break;
                }
                if (f == -1) {
                    if (EXACT) {
                        if (tet_contains_exact(pts, verts[t], p)) {
                            best = t;
                            br = r;
                            done = true;
                            break;
                        }
                    } else {
                        undecided = true;
                        done = true;
                        break;
                    }
                }
            }
        }
    }
    if (!EXACT && undecided) {
        cell_ids[s] = mark_undecided(i);
        return;
    }
    cell_ids[s] = best;
    double v = -1.0;
    double bb[4] = { 0, 0, 0, 0 };
    if (best >= 0) {
        // the matrix is kept (finite m) for every tet the filter accepts
        const bool affine = coef && br.m[0] == br.m[0];
        double4 cf = make_double4(0.0, 0.0, 0.0, 0.0);
        if (coef && out_values) cf = coef[best];
        int4 bv = make_int4(0, 0, 0, 0);
        if (bary_out || (out_values && !affine && (coef || values))) {
            bv = verts[best];
            tet_bary(pts, bv, p, bb);
        }
        if (out_values) {
            if (affine)
                v = cf.x + cf.y * ((double)p[0] - br.v3[0]) + cf.z * ((double)p[1] - br.v3[1]) +
                    cf.w * ((double)p[2] - br.v3[2]);
            else if (coef)
                v = bb[0] * cf.x + bb[1] * cf.y + bb[2] * cf.z + bb[3] * cf.w;
            else if (values)
                v = bb[0] * values[bv.x] + bb[1] * values[bv.y] + bb[2] * values[bv.z] + bb[3] * values[bv.w];
        }
    }
    if (out_values) out_values[s] = v;
    if (bary_out)
        for (int k = 0; k < 4; ++k) bary_out[4 * s + k] = (float)bb[k];
}

// Layout: everything the locate kernel reads sits at offsets fixed by (C, cap);
// the grid arrays come last, sized for the largest grid grid_dims can choose.
struct TetLayout {
    TetGridHeader* header;
    unsigned* bbox;
    TetRec* recs;
    int4* verts;     // vertex ids per tet
    uint2* entries;
    float* boxes;
    double4* coef;   // bound values' per-tet interpolants (NaN until bound)
    uint32_t *starts, *cursor, *scan_scratch;
    TetLayout(Arena& a, int C, size_t n_grid_max, unsigned long long cap)
    {
        header = a.take<TetGridHeader>(1);
        bbox = a.take<unsigned>(8);
        recs = a.take<TetRec>((size_t)(C > 0 ? C : 1));
        verts = a.take<int4>((size_t)(C > 0 ? C : 1));
        entries = a.take<uint2>((size_t)(cap > 0 ? cap : 1));
        starts = a.take<uint32_t>(n_grid_max + 1);
        cursor = a.take<uint32_t>(n_grid_max);
        scan_scratch = a.take<uint32_t>(scan_scratch_elems(n_grid_max) + 1);
        boxes = a.take<float>((size_t)(C > 0 ? C : 1) * 6);
        coef = a.take<double4>((size_t)(C > 0 ? C : 1));
    }
};

// Grid resolution: about one grid cell per kTetsPerCell tets (at most 2^26), cells
// as cubic as the bounding box allows; the cell count never exceeds max_grid_cells.
constexpr double kTetsPerCell = 1.0;
double grid_target(int C)
{
    double target = C > kTetsPerCell ? C / kTetsPerCell : 1.0;
    return target > 67108864.0 ? 67108864.0 : target;
}
size_t max_grid_cells(int C) { return (size_t)(2.0 * grid_target(C)) + 8; }

void grid_dims(int C, const float lo[3], const float hi[3], int dims[3], float cell[3])
{
    double ext[3];
    double vol = 1.0;
    int nz = 0;
    for (int a = 0; a < 3; ++a) {
        ext[a] = (double)hi[a] - lo[a];
        if (ext[a] > 0) { vol *= ext[a]; ++nz; }
    }
    const double target = grid_target(C);
    double h = nz > 0 ? pow(vol / target, 1.0 / nz) : 1.0;
    for (int iter = 0; iter < 64; ++iter) {
        size_t total = 1;
        for (int a = 0; a < 3; ++a) {
            if (ext[a] > 0 && h > 0) {
                double d = ceil(ext[a] / h);
                dims[a] = (int)(d < 1 ? 1 : (d > 65536 ? 65536 : d));
            } else {
                dims[a] = 1;
            }
            total *= (size_t)dims[a];
        }
        if (total <= max_grid_cells(C)) break;
        h *= 1.26;
    }
    for (int a = 0; a < 3; ++a) {
        if (ext[a] > 0) {
            cell[a] = (float)(ext[a] / dims[a]);
            // guard against a cell size that rounds below ext/dims
            while ((double)cell[a] * dims[a] < ext[a]) cell[a] = nextafterf(cell[a], INFINITY);
        } else {
            cell[a] = 1.0f;
        }
    }
}


// ---------------------------------------------------------------------------
// Non-tetrahedral cells of a mixed mesh (gpu_sampleu over VTK hexahedra,
// wedges and pyramids).  The cell holding a sample is found by the tetrahedral
// locator on the cells' tetrahedral splits; its value is the cell's own
// interpolant: the VTK shape functions N_i at the parametric coordinates of the
// sample (the isoparametric map inverted by Newton's method in double, from
// the cell centre), value = sum N_i v_i.  Shape functions, parametric (r, s, t):
//   hexahedron (12), points 0-3 bottom, 4-7 top:
//     (1-r)(1-s)(1-t), r(1-s)(1-t), rs(1-t), (1-r)s(1-t), (1-r)(1-s)t, r(1-s)t, rst, (1-r)st
//   wedge (13), points 0-2 one triangle, 3-5 the other:
//     (1-r-s)(1-t), r(1-t), s(1-t), (1-r-s)t, rt, st
//   pyramid (14), points 0-3 base, 4 apex:
//     (1-r)(1-s)(1-t), r(1-s)(1-t), rs(1-t), (1-r)s(1-t), t
// ---------------------------------------------------------------------------
__device__ __forceinline__ int shape_functions(int type, double r, double s, double t, double N[8], double dN[8][3])
{
    const double rm = 1.0 - r, sm = 1.0 - s, tm = 1.0 - t;
    if (type == 12) {
        const double a[8] = { rm * sm, r * sm, r * s, rm * s, rm * sm, r * sm, r * s, rm * s };
        const double dr[8] = { -sm, sm, s, -s, -sm, sm, s, -s };
        const double ds[8] = { -rm, -r, r, rm, -rm, -r, r, rm };
        for (int i = 0; i < 8; ++i) {
            const double w = i < 4 ? tm : t, dw = i < 4 ? -1.0 : 1.0;
            N[i] = a[i] * w;
            dN[i][0] = dr[i] * w;
            dN[i][1] = ds[i] * w;
            dN[i][2] = a[i] * dw;
        }
        return 8;
    }
    if (type == 13) {
        const double a[3] = { 1.0 - r - s, r, s };
        const double dr[3] = { -1.0, 1.0, 0.0 };
        const double ds[3] = { -1.0, 0.0, 1.0 };
        for (int i = 0; i < 6; ++i) {
            const int k = i % 3;
            const double w = i < 3 ? tm : t, dw = i < 3 ? -1.0 : 1.0;
            N[i] = a[k] * w;
            dN[i][0] = dr[k] * w;
            dN[i][1] = ds[k] * w;
            dN[i][2] = a[k] * dw;
        }
        return 6;
    }
    // pyramid
    const double a[4] = { rm * sm, r * sm, r * s, rm * s };
    const double dr[4] = { -sm, sm, s, -s };
    const double ds[4] = { -rm, -r, r, rm };
    for (int i = 0; i < 4; ++i) {
        N[i] = a[i] * tm;
        dN[i][0] = dr[i] * tm;
        dN[i][1] = ds[i] * tm;
        dN[i][2] = -a[i];
    }
    N[4] = t;
    dN[4][0] = 0.0;
    dN[4][1] = 0.0;
    dN[4][2] = 1.0;
    return 5;
}

__global__ void cell_interp_kernel(int S, const float* __restrict__ samples, const int* __restrict__ cell,
                                   const uint8_t* __restrict__ types, const int* __restrict__ offsets,
                                   const int* __restrict__ conn, const float* __restrict__ pts,
                                   const double* __restrict__ values, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const int c = cell[i];
    if (c < 0) return;
    const int type = types[c];
    if (type != 12 && type != 13 && type != 14) return;
    const int* ids = conn + offsets[c];
    const int nv = type == 12 ? 8 : (type == 13 ? 6 : 5);
    double X[8][3];
    for (int k = 0; k < nv; ++k)
        for (int a = 0; a < 3; ++a) X[k][a] = (double)pts[3 * ids[k] + a];
    const double p[3] = { (double)samples[3 * i], (double)samples[3 * i + 1], (double)samples[3 * i + 2] };
    double xi[3] = { 0.5, 0.5, 0.5 };
    if (type == 13) xi[0] = xi[1] = 1.0 / 3.0;
    if (type == 14) xi[2] = 0.2;
    double N[8], dN[8][3];
    for (int it = 0; it < 32; ++it) {
        shape_functions(type, xi[0], xi[1], xi[2], N, dN);
        double F[3] = { -p[0], -p[1], -p[2] }, J[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
        for (int k = 0; k < nv; ++k)
            for (int a = 0; a < 3; ++a) {
                F[a] += N[k] * X[k][a];
                for (int b = 0; b < 3; ++b) J[a][b] += dN[k][b] * X[k][a];
            }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                     c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        if (det == 0.0 || det != det) break;
        const double inv = 1.0 / det;
        const double d0 = (F[0] * c00 + F[1] * (J[0][2] * J[2][1] - J[0][1] * J[2][2]) +
                           F[2] * (J[0][1] * J[1][2] - J[0][2] * J[1][1])) * inv;
        const double d1 = (F[0] * c01 + F[1] * (J[0][0] * J[2][2] - J[0][2] * J[2][0]) +
                           F[2] * (J[0][2] * J[1][0] - J[0][0] * J[1][2])) * inv;
        const double d2 = (F[0] * c02 + F[1] * (J[0][1] * J[2][0] - J[0][0] * J[2][1]) +
                           F[2] * (J[0][0] * J[1][1] - J[0][1] * J[1][0])) * inv;
        xi[0] -= d0;
        xi[1] -= d1;
        xi[2] -= d2;
        if (fabs(d0) + fabs(d1) + fabs(d2) < 1e-14) break;
    }
    shape_functions(type, xi[0], xi[1], xi[2], N, dN);
    double v = 0.0;
    for (int k = 0; k < nv; ++k) v += N[k] * values[ids[k]];
    out[i] = v;
}
}  // namespace

size_t tet_locator_bytes(int n_points, int C, unsigned long long cap)
{
    (void)n_points;
    Arena a(nullptr, 0);
    TetLayout l(a, C, max_grid_cells(C), cap);
    return a.off + 256;
}

int tet_locator_build(int n_points, const float* pts, int C, const int* tets, unsigned long long cap, void* loc,
                      size_t bytes, unsigned long long* needed, hipStream_t st)
{
    if (needed) *needed = 0;
    if (n_points <= 0 || C <= 0) return set_error(GVC_ERR_INVALID_ARGUMENT, "empty mesh");
    Arena a(loc, bytes);
    TetLayout l(a, C, max_grid_cells(C), cap);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "tet locator buffer too small");
    GVC_TRY(bbox_launch(pts, n_points, l.bbox, st));
    unsigned hb[6];
    GVC_HIP(hipMemcpyAsync(hb, l.bbox, sizeof(hb), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipStreamSynchronize(st));
    TetGridHeader H;
    memset(&H, 0, sizeof(H));
    float lo[3], hi[3];
    bbox_from_keys(hb, lo, hi);
    for (int a2 = 0; a2 < 3; ++a2) {
        if (!(lo[a2] <= hi[a2])) return set_error(GVC_ERR_INVALID_ARGUMENT, "mesh points have no finite bounding box");
        H.lo[a2] = lo[a2];
        H.hi[a2] = hi[a2];
    }
    grid_dims(C, lo, hi, H.dims, H.cell);
    H.n_tets = C;
    H.n_points = n_points;
    H.capacity = cap;
    const size_t n_grid = (size_t)H.dims[0] * H.dims[1] * H.dims[2];
    GVC_HIP(hipMemcpyAsync(l.header, &H, sizeof(H), hipMemcpyHostToDevice, st));
    tet_records_kernel<<<blocks_for(C, 256), 256, 0, st>>>(pts, tets, C, l.recs, l.verts, l.boxes);
    GVC_LAUNCH_CHECK("tet_records");
    GVC_TRY(fill_u32(reinterpret_cast<uint32_t*>(l.coef), 0xffffffffu, (size_t)C * 8, st));   // NaN: unbound
    GVC_TRY(fill_u32(l.starts, 0u, n_grid + 1, st));
    tet_count_kernel<<<blocks_for(C, kBuildThreads), kBuildThreads, 0, st>>>(pts, tets, l.boxes, C, l.header,
                                                                              l.starts);
    GVC_LAUNCH_CHECK("tet_count");
    pad_counts_kernel<<<blocks_for(n_grid, 256), 256, 0, st>>>(l.starts, n_grid);
    GVC_LAUNCH_CHECK("pad_counts");
    // the bounding-box words are free again: 64-bit total of the padded lists
    unsigned long long* total64 = reinterpret_cast<unsigned long long*>(l.bbox);
    GVC_TRY(fill_u32(l.bbox, 0u, 2, st));
    count_total_kernel<<<256, 256, 0, st>>>(l.starts, n_grid, total64);
    GVC_LAUNCH_CHECK("count_total");
    unsigned long long total_all = 0;
    GVC_HIP(hipMemcpyAsync(&total_all, total64, sizeof(total_all), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipStreamSynchronize(st));
    if (total_all > 0xffffffffull - 64ull)
        return set_error(GVC_ERR_INVALID_ARGUMENT, "mesh too large for the tet locator (more than 2^32 - 64 grid entries)");
    GVC_TRY(exclusive_scan_u32(l.starts, l.starts, n_grid, true, l.scan_scratch, st));
    const uint32_t total = (uint32_t)total_all;
    if (needed) *needed = total;
    if (total > cap) return set_error(GVC_ERR_CAPACITY, "tet locator entry capacity too small");
    if (total > 0) {
        const unsigned nb_empty = blocks_for(total, 256);
        empty_entries_kernel<<<nb_empty < 8192u ? nb_empty : 8192u, 256, 0, st>>>(l.entries, total);
        GVC_LAUNCH_CHECK("empty_entries");
    }
    GVC_TRY(fill_u32(l.cursor, 0u, n_grid, st));
    tet_fill_kernel<<<blocks_for(C, kBuildThreads), kBuildThreads, 0, st>>>(pts, tets, l.boxes, C, l.header,
                                                                             l.starts, l.cursor, l.entries, cap);
    GVC_LAUNCH_CHECK("tet_fill");
    tet_sort_cells_kernel<<<blocks_for(n_grid, 4), 256, 0, st>>>(l.starts, l.cursor, n_grid, l.entries);
    GVC_LAUNCH_CHECK("tet_sort_cells");
    H.n_entries = total;
    GVC_HIP(hipMemcpyAsync(l.header, &H, sizeof(H), hipMemcpyHostToDevice, st));
    GVC_HIP(hipStreamSynchronize(st));
    return GVC_OK;
}

int cell_interp(int S, const float* samples, const int* cell_ids, const uint8_t* types, const int* offsets,
                const int* conn, const float* pts, const double* values, double* out, hipStream_t st)
{
    if (S <= 0) return GVC_OK;
    cell_interp_kernel<<<blocks_for(S, 256), 256, 0, st>>>(S, samples, cell_ids, types, offsets, conn, pts, values,
                                                           out);
    GVC_LAUNCH_CHECK("cell_interp");
    return GVC_OK;
}

int tet_locator_bind_values(int n_points, const float* pts, int C, void* loc, size_t bytes, unsigned long long cap,
                            const double* values, hipStream_t st)
{
    (void)n_points;
    if (C <= 0) return GVC_OK;
    Arena a(loc, bytes);
    TetLayout l(a, C, max_grid_cells(C), cap);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "tet locator buffer too small");
    tet_coef_kernel<<<blocks_for(C, 256), 256, 0, st>>>(pts, l.recs, l.verts, values, C, l.coef);
    GVC_LAUNCH_CHECK("tet_coef");
    return GVC_OK;
}

int tet_locate_interp(int n_points, const float* pts, int C, const void* loc, size_t bytes, unsigned long long cap,
                      const double* values, bool bound, int S, const float* samples, const void* sample_set,
                      int* cell_ids, double* out_values, float* bary, hipStream_t st)
{
    (void)n_points;
    if (S <= 0) return GVC_OK;
    Arena a(const_cast<void*>(loc), bytes);
    TetLayout l(a, C, max_grid_cells(C), cap);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "tet locator buffer too small");
    const float4* sorted = sample_set ? sample_set_points(sample_set, S) : nullptr;
    if ((size_t)S > (size_t)INT32_MAX - 2) return set_error(GVC_ERR_INVALID_ARGUMENT, "too many samples");
    const double4* coef = bound ? l.coef : nullptr;
    if (bound) values = nullptr;
    const unsigned nb = blocks_for((size_t)S, kLocateThreads);
    tet_locate_kernel<false><<<nb, kLocateThreads, 0, st>>>(pts, l.recs, l.verts, l.header, l.starts, l.entries, values, coef,
                                                            S, samples, sorted, cell_ids, out_values, bary);
    GVC_LAUNCH_CHECK("tet_locate");
    tet_locate_kernel<true><<<nb, kLocateThreads, 0, st>>>(pts, l.recs, l.verts, l.header, l.starts, l.entries, values, coef,
                                                           S, samples, sorted, cell_ids, out_values, bary);
    GVC_LAUNCH_CHECK("tet_locate_exact");
    return GVC_OK;
}

// ---------------------------------------------------------------------------
// uniform grid probe (gpu_sample): trilinear in the hexahedral cells of an
// nx x ny x nz point grid, value -1 (cast to the value type) outside.
// t_a = (p_a - origin_a) / spacing_a in float; valid iff 0 <= t_a <= n_a - 1;
// cell i_a = min(floor(t_a), n_a - 2); weights and the sum in double; integer
// types truncate toward zero.
// ---------------------------------------------------------------------------
// -1 for floating types; the two's-complement pattern of -1 (all ones) for the
// unsigned types, which is what the x86 conversion of -1.0 leaves in the low bits.
template <class T>
__device__ __forceinline__ T invalid_value()
{
    return (T)(-1.0);
}
template <>
__device__ __forceinline__ uint16_t invalid_value<uint16_t>()
{
    return (uint16_t)0xffffu;
}
template <>
__device__ __forceinline__ uint8_t invalid_value<uint8_t>()
{
    return (uint8_t)0xffu;
}

template <class T>
__global__ void grid_sample_kernel(int nx, int ny, int nz, float ox, float oy, float oz, float sx, float sy, float sz,
                                   const T* __restrict__ vals, int S, const float* __restrict__ samples,
                                   T* __restrict__ out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int n[3] = { nx, ny, nz };
    const float o[3] = { ox, oy, oz }, sp[3] = { sx, sy, sz };
    int i[3];
    double f[3];
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        const float t = (samples[3 * s + a] - o[a]) / sp[a];
        if (!(t >= 0.0f && t <= (float)(n[a] - 1))) ok = false;
        int c = ok ? (int)floorf(t) : 0;
        if (c > n[a] - 2) c = n[a] - 2;
        if (c < 0) c = 0;
        i[a] = c;
        f[a] = ok ? (double)(t - (float)c) : 0.0;
        if (n[a] == 1) f[a] = 0.0;
    }
    if (!ok) {
        out[s] = invalid_value<T>();
        return;
    }
    double acc = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
                if (w == 0.0) continue;
                const int xi = min(i[0] + dx, nx - 1), yi = min(i[1] + dy, ny - 1), zi = min(i[2] + dz, nz - 1);
                acc += w * (double)vals[(size_t)xi + (size_t)nx * ((size_t)yi + (size_t)ny * zi)];
            }
    out[s] = (T)acc;
}

int grid_sample(const long long dims[3], const float origin[3], const float spacing[3], int dtype, const void* values,
                int S, const float* samples, void* out, hipStream_t st)
{
    if (S <= 0) return GVC_OK;
    for (int a = 0; a < 3; ++a)
        if (dims[a] < 1 || dims[a] > (1ll << 30)) return set_error(GVC_ERR_INVALID_ARGUMENT, "grid dims out of range");
    const int nx = (int)dims[0], ny = (int)dims[1], nz = (int)dims[2];
    const unsigned nb = blocks_for(S, 256);
#define GVC_GRID(T)                                                                                                   \
    grid_sample_kernel<T><<<nb, 256, 0, st>>>(nx, ny, nz, origin[0], origin[1], origin[2], spacing[0], spacing[1],    \
                                              spacing[2], (const T*)values, S, samples, (T*)out)
    switch (dtype) {
    case 0: GVC_GRID(double); break;
    case 1: GVC_GRID(float); break;
    case 2: GVC_GRID(uint16_t); break;
    case 3: GVC_GRID(uint8_t); break;
    default: return set_error(GVC_ERR_INVALID_ARGUMENT, "unsupported value dtype");
    }
#undef GVC_GRID
    GVC_LAUNCH_CHECK("grid_sample");
    return GVC_OK;
}

}  // namespace gvc
