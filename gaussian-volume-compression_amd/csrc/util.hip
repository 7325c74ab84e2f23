// util.hip -- fill and device-wide exclusive scan (reduce-then-scan, 2048 items
// per 256-thread block, recursive over block sums).  Stream-ordered, no
// allocation, no memset nodes (safe inside a captured hipGraph).
#include <string.h>

#include "gvc_internal.h"

namespace gvc {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Exclusive scan of the 256 per-thread values of a block; returns the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total)
{
    __shared__ uint32_t wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        uint32_t s = wsum[w];
        if (w < wid) base += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + x - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_kernel(const uint32_t* in, uint32_t* out, size_t n,
                                                                 uint32_t* block_sums, uint32_t* total_out)
{
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) {
        if (block_sums) block_sums[blockIdx.x] = tot;
        if (total_out) *total_out = tot;
    }
}

// Adds each tile's offset; with total_out, block 0 also copies the scanned total
// (block_offsets[gridDim.x]) there.
__global__ __launch_bounds__(kScanThreads) void scan_add_kernel(uint32_t* out, size_t n, const uint32_t* block_offsets,
                                                                uint32_t* total_out)
{
    const size_t base = (size_t)blockIdx.x * kScanTile;
    const uint32_t add = block_offsets[blockIdx.x];
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
        if (base + k < n) out[base + k] += add;
    if (total_out && blockIdx.x == 0 && threadIdx.x == 0) *total_out = block_offsets[gridDim.x];
}

// Second step of a two-level scan when the tile sums are few (at most
// kPrefixSums): each block sums the tile sums before its own tile (integer sums:
// the same in any order), adds that to its tile, and the last block also writes
// the total.  Replaces the scan of the tile sums and its launch.
constexpr size_t kPrefixSums = 4096;

__global__ __launch_bounds__(kScanThreads) void scan_add_prefix_kernel(uint32_t* out, size_t n,
                                                                       const uint32_t* block_sums,
                                                                       uint32_t* total_out)
{
    __shared__ uint32_t red[kScanThreads / 64];
    const uint32_t b = blockIdx.x;
    uint32_t part = 0;
    for (uint32_t k = threadIdx.x; k < b; k += kScanThreads) part += block_sums[k];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) add += red[w];
    const size_t base = (size_t)b * kScanTile;
    if (add)
        for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
            if (base + k < n) out[base + k] += add;
    if (total_out && b == gridDim.x - 1 && threadIdx.x == 0) *total_out = add + block_sums[b];
}

__device__ __forceinline__ unsigned ord_key(float f)
{
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void bbox_init_kernel(unsigned* b)
{
    if (threadIdx.x < 3) b[threadIdx.x] = 0xffffffffu;
    else if (threadIdx.x < 6) b[threadIdx.x] = 0u;
}

__global__ void bbox_kernel(const float* __restrict__ pts, int n, unsigned* b)
{
    __shared__ unsigned s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int a = 0; a < 3; ++a) {
            float v = pts[3 * i + a];
            if (v != v) continue;
            unsigned k = ord_key(v);
            atomicMin(&s[a], k);
            atomicMax(&s[3 + a], k);
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&b[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&b[threadIdx.x], s[threadIdx.x]);
}

}  // namespace

int bbox_launch(const float* pts, int n, unsigned* keys, hipStream_t st)
{
    bbox_init_kernel<<<1, 64, 0, st>>>(keys);
    GVC_LAUNCH_CHECK("bbox_init");
    if (n > 0) {
        unsigned nb = (unsigned)((n + 255) / 256);
        if (nb > 1024) nb = 1024;
        bbox_kernel<<<nb, 256, 0, st>>>(pts, n, keys);
        GVC_LAUNCH_CHECK("bbox");
    }
    return GVC_OK;
}

void bbox_from_keys(const unsigned keys[6], float lo[3], float hi[3])
{
    for (int a = 0; a < 3; ++a) {
        unsigned kl = keys[a], kh = keys[3 + a];
        unsigned ul = (kl & 0x80000000u) ? (kl & 0x7fffffffu) : ~kl;
        unsigned uh = (kh & 0x80000000u) ? (kh & 0x7fffffffu) : ~kh;
        memcpy(&lo[a], &ul, 4);
        memcpy(&hi[a], &uh, 4);
    }
}

int fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st)
{
    if (n == 0) return GVC_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    fill_u32_kernel<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(p, value, n);
    GVC_LAUNCH_CHECK("fill_u32");
    return GVC_OK;
}

size_t scan_scratch_elems(size_t n)
{
    if (n <= (size_t)kScanTile) return 0;
    size_t nb = (n + kScanTile - 1) / kScanTile;
    return nb + 1 + scan_scratch_elems(nb);
}

int exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, bool with_total, uint32_t* scratch,
                       hipStream_t st)
{
    if (n == 0) {
        if (with_total) return fill_u32(out, 0u, 1, st);
        return GVC_OK;
    }
    if (n <= (size_t)kScanTile) {
        scan_tile_kernel<<<1, kScanThreads, 0, st>>>(in, out, n, nullptr, with_total ? out + n : nullptr);
        GVC_LAUNCH_CHECK("scan_tile");
        return GVC_OK;
    }
    size_t nb = (n + kScanTile - 1) / kScanTile;
    uint32_t* sums = scratch;  // nb + 1 entries
    scan_tile_kernel<<<dim3((unsigned)nb), kScanThreads, 0, st>>>(in, out, n, sums, nullptr);
    GVC_LAUNCH_CHECK("scan_tile");
    if (nb <= kPrefixSums) {
        scan_add_prefix_kernel<<<dim3((unsigned)nb), kScanThreads, 0, st>>>(out, n, sums, with_total ? out + n : nullptr);
        GVC_LAUNCH_CHECK("scan_add_prefix");
        return GVC_OK;
    }
    GVC_TRY(exclusive_scan_u32(sums, sums, nb, true, scratch + nb + 1, st));
    scan_add_kernel<<<dim3((unsigned)nb), kScanThreads, 0, st>>>(out, n, sums, with_total ? out + n : nullptr);
    GVC_LAUNCH_CHECK("scan_add");
    return GVC_OK;
}

}  // namespace gvc
