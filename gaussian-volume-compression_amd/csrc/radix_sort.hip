// radix_sort.hip -- stable LSD radix sort of (key, uint32 value) pairs, 8-bit
// digits, 2048-item tiles per 256-thread block.  Per pass:
//   1. per-tile digit histogram (LDS integer atomics: order-independent counts)
//   2. exclusive scan of the [digit][tile] histogram -> global digit offsets
//   3. per-tile stable ranking (wave ballot peer masks + cross-wave prefix in LDS,
//      items taken in striped rounds so round-major order is input order), an LDS
//      reorder into tile-sorted order, then a coalesced write-out.
// Stability makes every pass, and therefore the sort, deterministic.  An optional
// device-side count (n_dev) sorts only the first min(n, *n_dev) items: the grid is
// sized for n, tiles past the count stay idle and contribute empty histograms.  Replaces the
// sorts of R/gaussian_renderer/__init__.py:77 and R/train.py:311 (torch.argsort)
// and the CUB usage of CR/rasterizer_impl.cu; nothing is allocated inside.
#include "gvc_internal.h"

namespace gvc {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kWaves = kThreads / 64;
constexpr int kBins = 256;

__device__ __forceinline__ size_t count_of(size_t n, const uint32_t* n_dev)
{
    if (!n_dev) return n;
    const size_t m = *n_dev;
    return m < n ? m : n;
}

template <class K>
__global__ __launch_bounds__(kThreads) void radix_hist_kernel(const K* keys, size_t n_max, const uint32_t* n_dev,
                                                              int shift, uint32_t* hist, uint32_t n_tiles)
{
    const size_t n = count_of(n_max, n_dev);
    __shared__ uint32_t h[kBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        size_t i = base + (size_t)r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(kThreads) void radix_scatter_kernel(const K* keys, const uint32_t* vals, K* keys_out,
                                                                 uint32_t* vals_out, size_t n_max,
                                                                 const uint32_t* n_dev, int shift,
                                                                 const uint32_t* offsets, uint32_t n_tiles)
{
    const size_t n = count_of(n_max, n_dev);
    if ((size_t)blockIdx.x * kTile >= n) return;
    __shared__ uint32_t cnt[kWaves][kBins];
    __shared__ uint32_t run[kBins];
    __shared__ uint32_t start[kBins];
    __shared__ K skeys[kTile];
    __shared__ uint32_t svals[kTile];
    __shared__ uint32_t wtot[kWaves];

    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const size_t base = (size_t)blockIdx.x * kTile;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    run[t] = 0;
    K k_reg[kItems];
    uint32_t v_reg[kItems];
    uint32_t d_reg[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        size_t i = base + (size_t)r * kThreads + t;
        bool ok = i < n;
        k_reg[r] = ok ? keys[i] : K(0);
        v_reg[r] = ok ? vals[i] : 0u;
        d_reg[r] = ok ? ((uint32_t)(k_reg[r] >> shift) & 0xffu) : 256u;
    }
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) cnt[w][t] = 0;
        __syncthreads();
        const uint32_t d = d_reg[r];
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt_mask);
        if (d < 256u && below == 0) cnt[wid][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t acc = run[t];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                uint32_t c = cnt[w][t];
                cnt[w][t] = acc;
                acc += c;
            }
            run[t] = acc;
        }
        __syncthreads();
        rank[r] = (d < 256u) ? cnt[wid][d] + below : 0u;
        __syncthreads();
    }
    // tile-local exclusive scan of the digit counts
    {
        uint32_t x = run[t];
        uint32_t inc = x;
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
            uint32_t y = __shfl_up(inc, dd, 64);
            if (lane >= dd) inc += y;
        }
        if (lane == 63) wtot[wid] = inc;
        __syncthreads();
        uint32_t b = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
            if (w < wid) b += wtot[w];
        start[t] = b + inc - x;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const uint32_t d = d_reg[r];
        if (d < 256u) {
            const uint32_t pos = start[d] + rank[r];
            skeys[pos] = k_reg[r];
            svals[pos] = v_reg[r];
        }
    }
    __syncthreads();
    const size_t tile_n = (n - base) < (size_t)kTile ? (n - base) : (size_t)kTile;
    for (uint32_t k = t; k < tile_n; k += kThreads) {
        const K key = skeys[k];
        const uint32_t d = (uint32_t)(key >> shift) & 0xffu;
        const size_t g = (size_t)offsets[(size_t)d * n_tiles + blockIdx.x] + (k - start[d]);
        keys_out[g] = key;
        vals_out[g] = svals[k];
    }
}

template <class K>
int radix_sort_impl(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, size_t n, const uint32_t* n_dev,
                    int key_bits, uint32_t* scratch, bool* result_in_alt, hipStream_t st)
{
    *result_in_alt = false;
    if (n <= 1 || key_bits <= 0) return GVC_OK;
    if (n > 0xffffffffull) return set_error(GVC_ERR_INVALID_ARGUMENT, "radix sort: more than 2^32 items");
    const uint32_t n_tiles = (uint32_t)((n + kTile - 1) / kTile);
    uint32_t* hist = scratch;
    uint32_t* scan_scratch = scratch + (size_t)kBins * n_tiles;
    K* src_k = keys;
    uint32_t* src_v = vals;
    K* dst_k = keys_alt;
    uint32_t* dst_v = vals_alt;
    bool in_alt = false;
    for (int shift = 0; shift < key_bits; shift += 8) {
        radix_hist_kernel<K><<<n_tiles, kThreads, 0, st>>>(src_k, n, n_dev, shift, hist, n_tiles);
        GVC_LAUNCH_CHECK("radix_hist");
        GVC_TRY(exclusive_scan_u32(hist, hist, (size_t)kBins * n_tiles, false, scan_scratch, st));
        radix_scatter_kernel<K><<<n_tiles, kThreads, 0, st>>>(src_k, src_v, dst_k, dst_v, n, n_dev, shift, hist,
                                                              n_tiles);
        GVC_LAUNCH_CHECK("radix_scatter");
        K* tk = src_k; src_k = dst_k; dst_k = tk;
        uint32_t* tv = src_v; src_v = dst_v; dst_v = tv;
        in_alt = !in_alt;
    }
    *result_in_alt = in_alt;
    return GVC_OK;
}

}  // namespace

size_t radix_scratch_elems(size_t n)
{
    const size_t n_tiles = (n + kTile - 1) / kTile;
    return (size_t)kBins * n_tiles + scan_scratch_elems((size_t)kBins * n_tiles) + 1;
}

int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n, int key_bits,
                   uint32_t* scratch, bool* result_in_alt, hipStream_t st, const uint32_t* n_dev)
{
    return radix_sort_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, n_dev, key_bits, scratch, result_in_alt,
                                     st);
}

int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, size_t n, int key_bits,
                   uint32_t* scratch, bool* result_in_alt, hipStream_t st)
{
    return radix_sort_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, nullptr, key_bits, scratch, result_in_alt,
                                     st);
}

}  // namespace gvc
