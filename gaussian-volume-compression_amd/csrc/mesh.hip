// mesh.hip -- the mesh sampler: exact point location in tetrahedral meshes with
// barycentric interpolation, and trilinear probing of uniform grids.
//
// Replaces GS/sample_mesh.cu (gpu_sampleu / gpu_sample over Viskores Probe,
// :155-167, :380-390; GS/ = gpu_sampling/).  Differences by design: device
// pointers in and out (no host staging), int32 cell ids are returned, nothing is
// printed, and location is exact:
//   point p is in tet T = (v0,v1,v2,v3) iff, for every face, orient3d of the face
//   and p has the sign of orient3d(T), with simulation of simplicity resolving
//   exact zeros: p is perturbed to p + (e, e^2, e^3), so a point on a shared face,
//   edge or vertex belongs to exactly one tet of a conforming mesh.  Among several
//   containing tets (overlapping meshes) the lowest index wins.
// Candidates come from a uniform grid over the mesh's bounding box: every tet is
// listed in each grid cell its bounding box overlaps (CSR; order inside a cell is
// irrelevant because the lowest containing index is taken).
#include <math.h>
#include <string.h>

#include "exact.h"
#include "gvc_internal.h"

namespace gvc {

struct alignas(16) TetRec {
    int v[4];
    float lo[3], hi[3];
    float pad[2];
};
static_assert(sizeof(TetRec) == 48, "TetRec layout");

struct alignas(16) TetGridHeader {
    float lo[3];
    float cell[3];  // cell size per axis
    int dims[3];
    int n_tets;
    int n_points;
    int pad[3];
    unsigned long long n_entries;
    unsigned long long capacity;
};

namespace {

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__global__ void tet_records_kernel(const float* __restrict__ pts, const int* __restrict__ tets, int C,
                                   TetRec* __restrict__ recs)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C) return;
    TetRec r;
    for (int k = 0; k < 4; ++k) r.v[k] = tets[4 * t + k];
    for (int a = 0; a < 3; ++a) {
        float lo = pts[3 * r.v[0] + a], hi = lo;
        for (int k = 1; k < 4; ++k) {
            float x = pts[3 * r.v[k] + a];
            lo = fminf(lo, x);
            hi = fmaxf(hi, x);
        }
        r.lo[a] = lo;
        r.hi[a] = hi;
    }
    r.pad[0] = r.pad[1] = 0.f;
    recs[t] = r;
}

__device__ __forceinline__ int cell_coord(float x, float lo, float cell, int dim)
{
    float t = (x - lo) / cell;
    int c = (int)floorf(t);
    return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

__global__ void tet_count_kernel(const TetRec* __restrict__ recs, int C, const TetGridHeader* __restrict__ h,
                                 uint32_t* __restrict__ counts)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C) return;
    const TetRec r = recs[t];
    if (!(r.lo[0] == r.lo[0] && r.lo[1] == r.lo[1] && r.lo[2] == r.lo[2])) return;
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        c0[a] = cell_coord(r.lo[a], h->lo[a], h->cell[a], h->dims[a]);
        c1[a] = cell_coord(r.hi[a], h->lo[a], h->cell[a], h->dims[a]);
    }
    for (int k = c0[2]; k <= c1[2]; ++k)
        for (int j = c0[1]; j <= c1[1]; ++j)
            for (int i = c0[0]; i <= c1[0]; ++i)
                atomicAdd(&counts[(size_t)i + (size_t)h->dims[0] * ((size_t)j + (size_t)h->dims[1] * k)], 1u);
}

__global__ void tet_fill_kernel(const TetRec* __restrict__ recs, int C, const TetGridHeader* __restrict__ h,
                                const uint32_t* __restrict__ starts, uint32_t* __restrict__ cursor,
                                int* __restrict__ entries, unsigned long long capacity)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C) return;
    const TetRec r = recs[t];
    if (!(r.lo[0] == r.lo[0] && r.lo[1] == r.lo[1] && r.lo[2] == r.lo[2])) return;
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        c0[a] = cell_coord(r.lo[a], h->lo[a], h->cell[a], h->dims[a]);
        c1[a] = cell_coord(r.hi[a], h->lo[a], h->cell[a], h->dims[a]);
    }
    for (int k = c0[2]; k <= c1[2]; ++k)
        for (int j = c0[1]; j <= c1[1]; ++j)
            for (int i = c0[0]; i <= c1[0]; ++i) {
                const size_t g = (size_t)i + (size_t)h->dims[0] * ((size_t)j + (size_t)h->dims[1] * k);
                const unsigned long long pos = (unsigned long long)starts[g] + atomicAdd(&cursor[g], 1u);
                if (pos < capacity) entries[pos] = t;
            }
}

// Containment of p in tet r with the simulation-of-simplicity tie rule; on
// success writes the four barycentric coordinates.
__device__ bool tet_contains(const float* __restrict__ pts, const TetRec& r, const float p[3], double bary[4])
{
    const float* V[4] = { pts + 3 * r.v[0], pts + 3 * r.v[1], pts + 3 * r.v[2], pts + 3 * r.v[3] };
    double ovol;
    const int sT = orient3d_sign(V[0], V[1], V[2], V[3], &ovol);
    if (sT == 0) return false;
    double o[4];
    for (int i = 0; i < 4; ++i) {
        // face opposite i, remaining vertices in increasing order; o_i = s_i orient(F, p)
        const float* F[3];
        int m = 0;
        for (int k = 0; k < 4; ++k)
            if (k != i) F[m++] = V[k];
        const int si = ((3 - i) & 1) ? -1 : 1;
        double val;
        int sg = orient3d_sign(F[0], F[1], F[2], p, &val);
        if (sg == 0) {
            // d/dp orient(F, p) = -(F1 - F0) x (F2 - F0); p + (e, e^2, e^3)
            for (int a = 0; a < 3 && sg == 0; ++a) sg = -normal_sign(F[0], F[1], F[2], a);
        }
        if (si * sg != sT) return false;
        o[i] = si * val;
    }
    for (int i = 0; i < 4; ++i) bary[i] = o[i] / ovol;
    return true;
}

// sorted != nullptr: the samples are taken from a sample set (Morton order, so
// neighbouring threads share grid cells and tets) and results go to the original
// index carried in .w.
__global__ void tet_locate_kernel(const float* __restrict__ pts, const TetRec* __restrict__ recs,
                                  const TetGridHeader* __restrict__ h, const uint32_t* __restrict__ starts,
                                  const int* __restrict__ entries, const double* __restrict__ values, int S,
                                  const float* __restrict__ samples, const float4* __restrict__ sorted,
                                  int* __restrict__ cell_ids, double* __restrict__ out_values,
                                  float* __restrict__ bary_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    int s = i;
    float p[3];
    if (sorted) {
        const float4 q = sorted[i];
        p[0] = q.x; p[1] = q.y; p[2] = q.z;
        s = __builtin_bit_cast(int, q.w);
        if (s < 0) return;
    } else {
        p[0] = samples[3 * s]; p[1] = samples[3 * s + 1]; p[2] = samples[3 * s + 2];
    }
    int best = -1;
    double bb[4] = { 0, 0, 0, 0 };
    bool inside_grid = true;
    int c[3];
    for (int a = 0; a < 3; ++a) {
        const float t = (p[a] - h->lo[a]) / h->cell[a];
        if (!(t >= 0.0f) || !(t <= (float)h->dims[a])) inside_grid = false;
        int ci = (int)floorf(t);
        c[a] = ci < 0 ? 0 : (ci >= h->dims[a] ? h->dims[a] - 1 : ci);
    }
    if (inside_grid) {
        const size_t g = (size_t)c[0] + (size_t)h->dims[0] * ((size_t)c[1] + (size_t)h->dims[1] * c[2]);
        const uint32_t b = starts[g], e = starts[g + 1];
        for (uint32_t k = b; k < e; ++k) {
            const int t = entries[k];
            if (best >= 0 && t > best) continue;
            const TetRec r = recs[t];
            if (p[0] < r.lo[0] || p[0] > r.hi[0] || p[1] < r.lo[1] || p[1] > r.hi[1] || p[2] < r.lo[2] ||
                p[2] > r.hi[2])
                continue;
            double bc[4];
            if (tet_contains(pts, r, p, bc)) {
                best = t;
                for (int i = 0; i < 4; ++i) bb[i] = bc[i];
            }
        }
    }
    cell_ids[s] = best;
    if (out_values) {
        double v = -1.0;
        if (best >= 0 && values) {
            const TetRec r = recs[best];
            v = bb[0] * values[r.v[0]] + bb[1] * values[r.v[1]] + bb[2] * values[r.v[2]] + bb[3] * values[r.v[3]];
        }
        out_values[s] = v;
    }
    if (bary_out)
        for (int i = 0; i < 4; ++i) bary_out[4 * s + i] = (float)bb[i];
}

// Layout: everything the locate kernel reads sits at offsets fixed by (C, cap);
// the grid arrays come last, sized for the largest grid grid_dims can choose.
struct TetLayout {
    TetGridHeader* header;
    unsigned* bbox;
    TetRec* recs;
    int* entries;
    uint32_t *starts, *cursor, *scan_scratch;
    TetLayout(Arena& a, int C, size_t n_grid_max, unsigned long long cap)
    {
        header = a.take<TetGridHeader>(1);
        bbox = a.take<unsigned>(8);
        recs = a.take<TetRec>((size_t)(C > 0 ? C : 1));
        entries = a.take<int>((size_t)(cap > 0 ? cap : 1));
        starts = a.take<uint32_t>(n_grid_max + 1);
        cursor = a.take<uint32_t>(n_grid_max);
        scan_scratch = a.take<uint32_t>(scan_scratch_elems(n_grid_max) + 1);
    }
};

// Grid resolution: about one grid cell per two tets (at most 2^26), cells as
// cubic as the bounding box allows; the cell count never exceeds max_grid_cells.
double grid_target(int C)
{
    double target = C > 2 ? C / 2.0 : 1.0;
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
    TetLayout& l0 = l;
    GVC_TRY(bbox_launch(pts, n_points, l0.bbox, st));
    unsigned hb[6];
    GVC_HIP(hipMemcpyAsync(hb, l0.bbox, sizeof(hb), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipStreamSynchronize(st));
    TetGridHeader H;
    memset(&H, 0, sizeof(H));
    float lo[3], hi[3];
    bbox_from_keys(hb, lo, hi);
    for (int a = 0; a < 3; ++a) {
        if (!(lo[a] <= hi[a])) return set_error(GVC_ERR_INVALID_ARGUMENT, "mesh points have no finite bounding box");
        H.lo[a] = lo[a];
    }
    grid_dims(C, lo, hi, H.dims, H.cell);
    H.n_tets = C;
    H.n_points = n_points;
    H.capacity = cap;
    const size_t n_grid = (size_t)H.dims[0] * H.dims[1] * H.dims[2];
    GVC_HIP(hipMemcpyAsync(l.header, &H, sizeof(H), hipMemcpyHostToDevice, st));
    tet_records_kernel<<<blocks_for(C, 256), 256, 0, st>>>(pts, tets, C, l.recs);
    GVC_LAUNCH_CHECK("tet_records");
    GVC_TRY(fill_u32(l.starts, 0u, n_grid + 1, st));
    tet_count_kernel<<<blocks_for(C, 256), 256, 0, st>>>(l.recs, C, l.header, l.starts);
    GVC_LAUNCH_CHECK("tet_count");
    GVC_TRY(exclusive_scan_u32(l.starts, l.starts, n_grid, true, l.scan_scratch, st));
    uint32_t total = 0;
    GVC_HIP(hipMemcpyAsync(&total, l.starts + n_grid, sizeof(total), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipStreamSynchronize(st));
    if (needed) *needed = total;
    if (total > cap) return set_error(GVC_ERR_CAPACITY, "tet locator entry capacity too small");
    GVC_TRY(fill_u32(l.cursor, 0u, n_grid, st));
    tet_fill_kernel<<<blocks_for(C, 256), 256, 0, st>>>(l.recs, C, l.header, l.starts, l.cursor, l.entries, cap);
    GVC_LAUNCH_CHECK("tet_fill");
    H.n_entries = total;
    GVC_HIP(hipMemcpyAsync(l.header, &H, sizeof(H), hipMemcpyHostToDevice, st));
    GVC_HIP(hipStreamSynchronize(st));
    return GVC_OK;
}

int tet_locate_interp(int n_points, const float* pts, int C, const void* loc, size_t bytes, unsigned long long cap,
                      const double* values, int S, const float* samples, const void* sample_set, int* cell_ids,
                      double* out_values, float* bary, hipStream_t st)
{
    (void)n_points;
    if (S <= 0) return GVC_OK;
    Arena a(const_cast<void*>(loc), bytes);
    TetLayout l(a, C, max_grid_cells(C), cap);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "tet locator buffer too small");
    const float4* sorted = sample_set ? sample_set_points(sample_set, S) : nullptr;
    tet_locate_kernel<<<blocks_for(S, 128), 128, 0, st>>>(pts, l.recs, l.header, l.starts, l.entries, values, S,
                                                          samples, sorted, cell_ids, out_values, bary);
    GVC_LAUNCH_CHECK("tet_locate");
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
