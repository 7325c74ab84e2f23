// knn.hip -- distCUDA2: per point, the mean squared distance to its 3 nearest
// other points (the simple-knn operator used at R/scene/gaussian_model.py:175-178
// and :730-733; simple-knn itself is not in the reference tree).
//
// Exact search on a uniform grid (about two points per cell): every point scans
// Chebyshev rings of cells around its own until the third-best squared distance
// is below the squared distance to the nearest unscanned cell (with a 1e-3
// relative margin so float rounding of distances cannot end the search early).
// Distances use the canonical float order dx*dx + dy*dy + dz*dz (no fma); the
// result is ((d1 + d2) + d3) / 3 over the three smallest, ascending.  Points with
// fewer than three others get +inf for the missing ones (as an unfilled heap).
#include <math.h>

#include "gvc_internal.h"

namespace gvc {

namespace {

struct alignas(16) KnnHeader {
    float lo[3];
    float cell[3];
    int dims[3];
    int n;
};

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

double knn_target(int n)
{
    double t = n > 2 ? n / 2.0 : 1.0;
    return t > 67108864.0 ? 67108864.0 : t;
}
size_t knn_max_cells(int n) { return (size_t)(2.0 * knn_target(n)) + 8; }

void knn_grid(int n, const float lo[3], const float hi[3], int dims[3], float cell[3])
{
    double ext[3], vol = 1.0;
    int nz = 0;
    for (int a = 0; a < 3; ++a) {
        ext[a] = (double)hi[a] - lo[a];
        if (ext[a] > 0) { vol *= ext[a]; ++nz; }
    }
    double h = nz > 0 ? pow(vol / knn_target(n), 1.0 / nz) : 1.0;
    for (int iter = 0; iter < 64; ++iter) {
        size_t total = 1;
        for (int a = 0; a < 3; ++a) {
            dims[a] = (ext[a] > 0 && h > 0) ? (int)fmin(fmax(ceil(ext[a] / h), 1.0), 65536.0) : 1;
            total *= (size_t)dims[a];
        }
        if (total <= knn_max_cells(n)) break;
        h *= 1.26;
    }
    for (int a = 0; a < 3; ++a) {
        if (ext[a] > 0) {
            cell[a] = (float)(ext[a] / dims[a]);
            while ((double)cell[a] * dims[a] < ext[a]) cell[a] = nextafterf(cell[a], INFINITY);
        } else {
            cell[a] = 1.0f;
        }
    }
}

__device__ __forceinline__ int kcell(float x, float lo, float cell, int dim)
{
    int c = (int)floorf((x - lo) / cell);
    return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

__device__ __forceinline__ bool finite3(float x, float y, float z)
{
    return isfinite(x) && isfinite(y) && isfinite(z);
}

__global__ void knn_count_kernel(const float* __restrict__ p, int n, const KnnHeader* __restrict__ h,
                                 uint32_t* __restrict__ counts)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
    if (!finite3(x, y, z)) return;
    const int cx = kcell(x, h->lo[0], h->cell[0], h->dims[0]);
    const int cy = kcell(y, h->lo[1], h->cell[1], h->dims[1]);
    const int cz = kcell(z, h->lo[2], h->cell[2], h->dims[2]);
    atomicAdd(&counts[(size_t)cx + (size_t)h->dims[0] * ((size_t)cy + (size_t)h->dims[1] * cz)], 1u);
}

__global__ void knn_fill_kernel(const float* __restrict__ p, int n, const KnnHeader* __restrict__ h,
                                const uint32_t* __restrict__ starts, uint32_t* __restrict__ cursor,
                                int* __restrict__ order)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
    if (!finite3(x, y, z)) return;
    const int cx = kcell(x, h->lo[0], h->cell[0], h->dims[0]);
    const int cy = kcell(y, h->lo[1], h->cell[1], h->dims[1]);
    const int cz = kcell(z, h->lo[2], h->cell[2], h->dims[2]);
    const size_t g = (size_t)cx + (size_t)h->dims[0] * ((size_t)cy + (size_t)h->dims[1] * cz);
    order[starts[g] + atomicAdd(&cursor[g], 1u)] = i;
}

__device__ __forceinline__ void insert3(float d, float& b0, float& b1, float& b2)
{
    if (d < b2) {
        if (d < b1) {
            b2 = b1;
            if (d < b0) { b1 = b0; b0 = d; } else { b1 = d; }
        } else {
            b2 = d;
        }
    }
}

__global__ void knn_query_kernel(const float* __restrict__ p, int n, const KnnHeader* __restrict__ h,
                                 const uint32_t* __restrict__ starts, const int* __restrict__ order,
                                 float* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
    if (!finite3(x, y, z)) {
        out[i] = __builtin_nanf("");
        return;
    }
    const int D[3] = { h->dims[0], h->dims[1], h->dims[2] };
    const float P[3] = { x, y, z };
    int c[3];
    for (int a = 0; a < 3; ++a) c[a] = kcell(P[a], h->lo[a], h->cell[a], D[a]);
    const int rmax = max(D[0], max(D[1], D[2]));
    float b0 = __builtin_inff(), b1 = b0, b2 = b0;
    auto scan_cell = [&](int ii, int j, int k) {
        const size_t g = (size_t)ii + (size_t)D[0] * ((size_t)j + (size_t)D[1] * k);
        for (uint32_t q = starts[g]; q < starts[g + 1]; ++q) {
            const int o = order[q];
            if (o == i) continue;
            const float dx = x - p[3 * o], dy = y - p[3 * o + 1], dz = z - p[3 * o + 2];
            const float d = (dx * dx + dy * dy) + dz * dz;
            insert3(d, b0, b1, b2);
        }
    };
    for (int r = 0; r <= rmax; ++r) {
        // cells at Chebyshev distance exactly r from c, clipped to the grid
        const int k0 = max(c[2] - r, 0), k1 = min(c[2] + r, D[2] - 1);
        const int j0 = max(c[1] - r, 0), j1 = min(c[1] + r, D[1] - 1);
        const int i0 = max(c[0] - r, 0), i1 = min(c[0] + r, D[0] - 1);
        for (int k = k0; k <= k1; ++k)
            for (int j = j0; j <= j1; ++j) {
                const bool shell = (k == c[2] - r) || (k == c[2] + r) || (j == c[1] - r) || (j == c[1] + r);
                if (shell) {
                    for (int ii = i0; ii <= i1; ++ii) scan_cell(ii, j, k);
                } else {
                    if (c[0] - r >= 0) scan_cell(c[0] - r, j, k);
                    if (r > 0 && c[0] + r <= D[0] - 1) scan_cell(c[0] + r, j, k);
                }
            }
        // lower bound on the distance to any point in rings > r
        double gap = INFINITY;
        for (int a = 0; a < 3; ++a) {
            if (c[a] - r > 0) gap = fmin(gap, (double)P[a] - ((double)h->lo[a] + (double)(c[a] - r) * h->cell[a]));
            if (c[a] + r < D[a] - 1)
                gap = fmin(gap, ((double)h->lo[a] + (double)(c[a] + r + 1) * h->cell[a]) - (double)P[a]);
        }
        if (gap == INFINITY) break;  // every cell scanned
        if (gap > 0 && (double)b2 <= 0.999 * gap * gap) break;
    }
    out[i] = ((b0 + b1) + b2) / 3.0f;
}

struct KnnLayout {
    KnnHeader* header;
    unsigned* bbox;
    uint32_t *starts, *cursor, *scan_scratch;
    int* order;
    KnnLayout(Arena& a, int n, size_t n_grid)
    {
        header = a.take<KnnHeader>(1);
        bbox = a.take<unsigned>(8);
        starts = a.take<uint32_t>(n_grid + 1);
        cursor = a.take<uint32_t>(n_grid);
        scan_scratch = a.take<uint32_t>(scan_scratch_elems(n_grid) + 1);
        order = a.take<int>((size_t)(n > 0 ? n : 1));
    }
};

}  // namespace

size_t knn3_bytes(int n)
{
    Arena a(nullptr, 0);
    KnnLayout l(a, n, knn_max_cells(n));
    return a.off + 256;
}

int knn3(int n, const float* pts, float* out, void* ws, size_t bytes, hipStream_t st)
{
    if (n <= 0) return GVC_OK;
    Arena a0(ws, bytes);
    KnnLayout l0(a0, n, 1);
    GVC_TRY(bbox_launch(pts, n, l0.bbox, st));
    unsigned hb[6];
    GVC_HIP(hipMemcpyAsync(hb, l0.bbox, sizeof(hb), hipMemcpyDeviceToHost, st));
    GVC_HIP(hipStreamSynchronize(st));
    KnnHeader H;
    float lo[3], hi[3];
    bbox_from_keys(hb, lo, hi);
    for (int a = 0; a < 3; ++a) {
        if (!(lo[a] <= hi[a])) { lo[a] = 0.f; hi[a] = 0.f; }  // no finite point
        H.lo[a] = lo[a];
    }
    knn_grid(n, lo, hi, H.dims, H.cell);
    H.n = n;
    const size_t n_grid = (size_t)H.dims[0] * H.dims[1] * H.dims[2];
    Arena a(ws, bytes);
    KnnLayout l(a, n, n_grid);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "knn workspace too small");
    GVC_HIP(hipMemcpyAsync(l.header, &H, sizeof(H), hipMemcpyHostToDevice, st));
    GVC_TRY(fill_u32(l.starts, 0u, n_grid + 1, st));
    knn_count_kernel<<<blocks_for(n, 256), 256, 0, st>>>(pts, n, l.header, l.starts);
    GVC_LAUNCH_CHECK("knn_count");
    GVC_TRY(exclusive_scan_u32(l.starts, l.starts, n_grid, true, l.scan_scratch, st));
    GVC_TRY(fill_u32(l.cursor, 0u, n_grid, st));
    knn_fill_kernel<<<blocks_for(n, 256), 256, 0, st>>>(pts, n, l.header, l.starts, l.cursor, l.order);
    GVC_LAUNCH_CHECK("knn_fill");
    knn_query_kernel<<<blocks_for(n, 128), 128, 0, st>>>(pts, n, l.header, l.starts, l.order, out);
    GVC_LAUNCH_CHECK("knn_query");
    // the host header copy above must outlive its transfer
    GVC_HIP(hipStreamSynchronize(st));
    return GVC_OK;
}

}  // namespace gvc
