// radix_sort.hip -- stable LSD radix sort of (key, uint32 value) pairs.
// Digits of at most 8 bits, as few passes as the key width needs (20-bit keys:
// three 7-bit passes); 8192-item tiles per 512-thread block.  Per pass:
//   1. per-tile digit histogram (per-wave LDS counters, integer atomics: the
//      counts do not depend on order)
//   2. exclusive scan of the [digit][tile] histogram -> global digit offsets
//   3. per-tile stable ranking: each wave owns 1024 consecutive items and ranks
//      them 64 at a time (ballot peer masks; the lowest peer lane updates the
//      wave's private digit counter), a cross-wave prefix per digit, an LDS
//      reorder into tile-sorted order, then a write-out in digit runs.
// Stability makes every pass, and therefore the sort, deterministic.  An optional
// device-side count (n_dev) sorts only the first min(n, *n_dev) items: the grid is
// sized for n, tiles past the count stay idle and contribute empty histograms.
// Optional identity values: the first pass takes item i's value to be i (the
// caller's value array is then only written).
// Segmented mode (seg_off: nseg + 1 item offsets on the device): the items are
// already grouped into contiguous segments (the volumes of a batch) and every
// segment is sorted on its own, in place of a further pass on a segment key.
// Tiles never span segments (a segment of n_s items takes ceil(n_s / 8192)
// tiles); the histogram of each pass is laid out segment-major -- column
// (d, t) = (digits) * tau_s + d * T_s + (t - tau_s) for tile t of segment s, whose
// T_s tiles start at tile tau_s -- so one exclusive scan gives every tile its
// output offsets, segments included.  Replaces the sorts of
// R/gaussian_renderer/__init__.py:77 and R/train.py:311 (torch.argsort) and the
// CUB usage of CR/rasterizer_impl.cu; nothing is allocated inside.
#include "gvc_internal.h"

namespace gvc {

namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 16;                  // per thread (rounds per wave)
constexpr int kTile = kThreads * kItems;    // 8192
constexpr int kWaveItems = 64 * kItems;     // consecutive items a wave ranks
constexpr int kMaxBins = 256;

__device__ __forceinline__ size_t count_of(size_t n, const uint32_t* n_dev)
{
    if (!n_dev) return n;
    const size_t m = *n_dev;
    return m < n ? m : n;
}

// A tile: items [begin, end), histogram column of digit d = col + d * stride.
struct TileSpan {
    size_t begin, end;
    size_t col, stride;
};

// tiles (segmented mode): per tile {begin, end, tau_s, T_s}; empty tiles past the
// last segment's have begin == end and tau_s = the tile, T_s = 1.
__device__ __forceinline__ TileSpan tile_span(const uint4* __restrict__ tiles, size_t n, uint32_t n_tiles,
                                              uint32_t digits)
{
    TileSpan sp;
    if (tiles) {
        const uint4 q = tiles[blockIdx.x];
        sp.begin = q.x;
        sp.end = q.y;
        sp.col = (size_t)digits * q.z + (blockIdx.x - q.z);
        sp.stride = q.w;
    } else {
        sp.begin = (size_t)blockIdx.x * kTile;
        sp.end = sp.begin + kTile < n ? sp.begin + kTile : n;
        sp.col = blockIdx.x;
        sp.stride = n_tiles;
    }
    return sp;
}

// First tile of every segment (segmented mode): tau[s] = sum of ceil(n_s' / kTile)
// over the segments before s, tau[nseg] = all tiles.  One thread (a batch has few
// segments).
__global__ void seg_tiles_kernel(const int* __restrict__ seg_off, int nseg, uint32_t* __restrict__ tau)
{
    uint32_t t = 0;
    for (int s = 0; s < nseg; ++s) {
        tau[s] = t;
        const uint32_t ns = (uint32_t)(seg_off[s + 1] - seg_off[s]);
        t += (ns + kTile - 1) / kTile;
    }
    tau[nseg] = t;
}

__global__ void tile_table_kernel(const int* __restrict__ seg_off, int nseg, const uint32_t* __restrict__ tau,
                                  uint32_t max_tiles, uint4* __restrict__ tiles)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= max_tiles) return;
    if (t >= tau[nseg]) {
        tiles[t] = make_uint4(0u, 0u, t, 1u);
        return;
    }
    int lo = 0, hi = nseg - 1;   // the segment s with tau[s] <= t < tau[s + 1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (tau[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t b = (uint32_t)seg_off[lo] + (t - tau[lo]) * (uint32_t)kTile;
    const uint32_t e = b + (uint32_t)kTile < (uint32_t)seg_off[lo + 1] ? b + (uint32_t)kTile : (uint32_t)seg_off[lo + 1];
    tiles[t] = make_uint4(b, e, tau[lo], tau[lo + 1] - tau[lo]);
}

template <class K>
__global__ __launch_bounds__(kThreads) void radix_hist_kernel(const K* __restrict__ keys, size_t n_max,
                                                              const uint32_t* __restrict__ n_dev, int shift,
                                                              uint32_t mask, uint32_t* __restrict__ hist,
                                                              uint32_t n_tiles, const uint4* __restrict__ tiles)
{
    const size_t n = count_of(n_max, n_dev);
    const TileSpan sp = tile_span(tiles, n, n_tiles, mask + 1u);
    __shared__ uint32_t h[kWaves][kMaxBins];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = t; i < kWaves * kMaxBins; i += kThreads) (&h[0][0])[i] = 0u;
    __syncthreads();
    const size_t base = sp.begin + (size_t)w * kWaveItems + lane;
    // all of a thread's keys in flight at once, then the counts
    K kv[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const size_t i = base + (size_t)r * 64;
        kv[r] = i < sp.end ? keys[i] : K(0);
    }
#pragma unroll
    for (int r = 0; r < kItems; ++r)
        if (base + (size_t)r * 64 < sp.end) atomicAdd(&h[w][(uint32_t)(kv[r] >> shift) & mask], 1u);
    __syncthreads();
    for (uint32_t d = t; d <= mask; d += kThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int v = 0; v < kWaves; ++v) s += h[v][d];
        hist[(size_t)d * sp.stride + sp.col] = s;
    }
}

// Dynamic LDS: kTile keys, kTile values.
template <class K>
constexpr size_t scatter_lds() { return (size_t)kTile * (sizeof(K) + sizeof(uint32_t)); }

template <class K>
__global__ __launch_bounds__(kThreads) void radix_scatter_kernel(const K* __restrict__ keys,
                                                                 const uint32_t* __restrict__ vals,
                                                                 K* __restrict__ keys_out,
                                                                 uint32_t* __restrict__ vals_out, size_t n_max,
                                                                 const uint32_t* __restrict__ n_dev, int shift,
                                                                 uint32_t mask, const uint32_t* __restrict__ offsets,
                                                                 uint32_t n_tiles, const uint4* __restrict__ tiles,
                                                                 uint32_t* __restrict__ inv_out)
{
    const TileSpan sp = tile_span(tiles, count_of(n_max, n_dev), n_tiles, mask + 1u);
    const size_t base = sp.begin, n = sp.end;   // this tile: items [base, n)
    if (base >= n) return;
    extern __shared__ __align__(16) unsigned char dyn[];
    K* skeys = reinterpret_cast<K*>(dyn);
    uint32_t* svals = reinterpret_cast<uint32_t*>(dyn + (size_t)kTile * sizeof(K));
    __shared__ uint32_t cnt[kWaves][kMaxBins];
    __shared__ uint32_t start[kMaxBins];
    __shared__ uint32_t gofs[kMaxBins];
    __shared__ uint32_t wsum[kWaves];

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int i = t; i < kWaves * kMaxBins; i += kThreads) (&cnt[0][0])[i] = 0u;
    const size_t wbase = base + (size_t)w * kWaveItems + lane;

    K kr[kItems];
    uint32_t vr[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const size_t i = wbase + (size_t)r * 64;
        const bool ok = i < n;
        kr[r] = ok ? keys[i] : K(0);
        vr[r] = ok ? (vals ? vals[i] : (uint32_t)i) : 0u;
    }
    __syncthreads();
    // wave-private stable ranks: round r holds the wave's items r*64 .. r*64+63
    uint32_t rk[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const bool ok = wbase + (size_t)r * 64 < n;
        const uint32_t d = ok ? ((uint32_t)(kr[r] >> shift) & mask) : (uint32_t)kMaxBins;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int leader = __builtin_ctzll(peers);
        uint32_t old = 0u;
        if (ok && lane == leader) {
            old = cnt[w][d];
            cnt[w][d] = old + (uint32_t)__popcll(peers);
        }
        old = __shfl(old, leader, 64);
        rk[r] = old + (uint32_t)__popcll(peers & lt_mask);
    }
    __syncthreads();
    // per digit: exclusive prefix over the waves, then over the digits
    {
        const uint32_t d = (uint32_t)t;
        uint32_t tot = 0u;
        if (d < (uint32_t)kMaxBins) {
#pragma unroll
            for (int v = 0; v < kWaves; ++v) {
                const uint32_t c = cnt[v][d];
                cnt[v][d] = tot;
                tot += c;
            }
        }
        uint32_t inc = tot;
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
            const uint32_t y = __shfl_up(inc, dd, 64);
            if (lane >= dd) inc += y;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (d < (uint32_t)kMaxBins) {
            uint32_t b = 0u;
            for (int v = 0; v < w; ++v) b += wsum[v];
            const uint32_t s0 = b + inc - tot;
            start[d] = s0;
            gofs[d] = (d <= mask) ? offsets[(size_t)d * sp.stride + sp.col] - s0 : 0u;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        if (wbase + (size_t)r * 64 < n) {
            const uint32_t d = (uint32_t)(kr[r] >> shift) & mask;
            const uint32_t pos = start[d] + cnt[w][d] + rk[r];
            skeys[pos] = kr[r];
            svals[pos] = vr[r];
        }
    }
    __syncthreads();
    const uint32_t tile_n = (uint32_t)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    for (uint32_t k = t; k < tile_n; k += kThreads) {
        const K key = skeys[k];
        const uint32_t d = (uint32_t)(key >> shift) & mask;
        const size_t g = (size_t)(gofs[d] + k);
        keys_out[g] = key;
        vals_out[g] = svals[k];
        if (inv_out) inv_out[svals[k]] = (uint32_t)g;   // last pass: the inverse permutation
    }
}

template <class K>
int allow_scatter_lds()
{
    static bool done = false;
    if (!done) {
        GVC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&radix_scatter_kernel<K>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)scatter_lds<K>()));
        done = true;
    }
    return GVC_OK;
}

// Tiles of a sort of n items in nseg segments (nseg <= 1: not segmented).
inline size_t max_tiles(size_t n, int nseg)
{
    return (n + kTile - 1) / kTile + (nseg > 1 ? (size_t)nseg : 0);
}

template <class K>
int radix_sort_impl(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, size_t n, const uint32_t* n_dev,
                    int key_bits, uint32_t* scratch, bool* result_in_alt, hipStream_t st, bool vals_identity,
                    const int* seg_off = nullptr, int nseg = 1, uint32_t* inv_out = nullptr)
{
    *result_in_alt = false;
    if (n <= 1 || key_bits <= 0) {
        // nothing to reorder; identity values still have to be written
        if ((vals_identity || inv_out) && n > 1)
            return set_error(GVC_ERR_INVALID_ARGUMENT, "radix sort: identity values need a key to sort on");
        if (vals_identity && n == 1) GVC_TRY(fill_u32(vals, 0u, 1, st));
        if (inv_out && n == 1) GVC_TRY(fill_u32(inv_out, 0u, 1, st));
        return GVC_OK;
    }
    if (n > 0xffffffffull) return set_error(GVC_ERR_INVALID_ARGUMENT, "radix sort: more than 2^32 items");
    if (key_bits > (int)(8 * sizeof(K))) key_bits = (int)(8 * sizeof(K));
    GVC_TRY(allow_scatter_lds<K>());
    const bool segmented = seg_off && nseg > 1;
    if (segmented && n_dev) return set_error(GVC_ERR_INVALID_ARGUMENT, "radix sort: segments with a device count");
    const uint32_t n_tiles = (uint32_t)max_tiles(n, segmented ? nseg : 1);
    uint32_t* hist = scratch;
    uint32_t* scan_scratch = scratch + (size_t)kMaxBins * n_tiles;
    uint4* tiles = nullptr;
    if (segmented) {
        uint32_t* tau = scan_scratch + ((scan_scratch_elems((size_t)kMaxBins * n_tiles) + 3) & ~(size_t)3);
        tiles = reinterpret_cast<uint4*>(tau + (((size_t)nseg + 1 + 3) & ~(size_t)3));
        seg_tiles_kernel<<<1, 1, 0, st>>>(seg_off, nseg, tau);
        GVC_LAUNCH_CHECK("seg_tiles");
        tile_table_kernel<<<(n_tiles + 255) / 256, 256, 0, st>>>(seg_off, nseg, tau, n_tiles, tiles);
        GVC_LAUNCH_CHECK("tile_table");
    }
    const int passes = (key_bits + 7) / 8;
    const int width = (key_bits + passes - 1) / passes;
    K* src_k = keys;
    const uint32_t* src_v = vals_identity ? nullptr : vals;
    K* dst_k = keys_alt;
    uint32_t* dst_v = vals_alt;
    uint32_t* other_v = vals;
    bool in_alt = false;
    for (int shift = 0; shift < key_bits; shift += width) {
        const int bits = (key_bits - shift) < width ? (key_bits - shift) : width;
        const uint32_t mask = (1u << bits) - 1u;
        radix_hist_kernel<K><<<n_tiles, kThreads, 0, st>>>(src_k, n, n_dev, shift, mask, hist, n_tiles, tiles);
        GVC_LAUNCH_CHECK("radix_hist");
        GVC_TRY(exclusive_scan_u32(hist, hist, (size_t)(mask + 1) * n_tiles, false, scan_scratch, st));
        const bool last = shift + width >= key_bits;
        radix_scatter_kernel<K><<<n_tiles, kThreads, scatter_lds<K>(), st>>>(src_k, src_v, dst_k, dst_v, n, n_dev,
                                                                            shift, mask, hist, n_tiles, tiles,
                                                                            last ? inv_out : nullptr);
        GVC_LAUNCH_CHECK("radix_scatter");
        K* tk = src_k; src_k = dst_k; dst_k = tk;
        uint32_t* nv = other_v;
        other_v = dst_v;
        src_v = dst_v;
        dst_v = nv;
        in_alt = !in_alt;
    }
    *result_in_alt = in_alt;
    return GVC_OK;
}

}  // namespace

size_t radix_scratch_elems(size_t n, int segments)
{
    const size_t n_tiles = max_tiles(n, segments);
    const size_t scan = scan_scratch_elems((size_t)kMaxBins * n_tiles);
    size_t e = (size_t)kMaxBins * n_tiles + scan + 1;
    if (segments > 1)   // 16-byte aligned tau and tile table after the scan scratch
        e = (size_t)kMaxBins * n_tiles + ((scan + 3) & ~(size_t)3) + (((size_t)segments + 1 + 3) & ~(size_t)3) +
            4 * n_tiles + 4;
    return e;
}

int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n, int key_bits,
                   uint32_t* scratch, bool* result_in_alt, hipStream_t st, const uint32_t* n_dev, bool vals_identity,
                   const int* seg_off, int segments, uint32_t* inv_out)
{
    return radix_sort_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, n_dev, key_bits, scratch, result_in_alt,
                                     st, vals_identity, seg_off, segments, inv_out);
}

int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, size_t n, int key_bits,
                   uint32_t* scratch, bool* result_in_alt, hipStream_t st)
{
    return radix_sort_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, nullptr, key_bits, scratch, result_in_alt,
                                     st, false);
}

}  // namespace gvc
