// metrics.hip -- the sums behind the volume evaluation of R/volume_metrics.py:318-369
// in one pass over (cells, gt).
//
// The reference evaluates a fitted model on a grid (or on mesh points) and
// prints L1, L2, five PSNR variants and false-negative / false-positive counts,
// each a separate torch reduction.  Here one kernel reads every (cell, gt) pair
// once and forms all the sums in FP64 (a fixed tree per 256-thread block, then
// one block adds the block sums in order: deterministic, no atomics, no memset),
// written to out[GVC_METRIC_SUMS]:
//   0  sum |c - g|                       1  sum (c - g)^2
//   2  sum (c - g)^2 over (g != -1 & c != -1)      3  count of that mask
//   4  sum g^2                           5  sum (c0 - g)^2,  c0 = (c == -1 ? 0 : c)
//   6  sum (c0 - g)^2 over g != -1       7  count g != -1
//   8  count g == -1                     9  count c == -1 & g != -1   (false negatives)
//   10 count c != -1 & g == -1   (false positives)
#include <mutex>

#include "gvc_internal.h"

namespace gvc {

namespace {

constexpr int kMetricThreads = 256;
constexpr int kSums = 11;

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ void metric_terms(float c, float g, double v[kSums])
{
    const double d = (double)c - (double)g;
    const bool gv = g != -1.0f, cv = c != -1.0f;
    const double c0 = cv ? (double)c : 0.0;
    const double d0 = c0 - (double)g;
    v[0] = fabs(d);
    v[1] = d * d;
    v[2] = (gv && cv) ? d * d : 0.0;
    v[3] = (gv && cv) ? 1.0 : 0.0;
    v[4] = (double)g * (double)g;
    v[5] = d0 * d0;
    v[6] = gv ? d0 * d0 : 0.0;
    v[7] = gv ? 1.0 : 0.0;
    v[8] = gv ? 0.0 : 1.0;
    v[9] = (!cv && gv) ? 1.0 : 0.0;
    v[10] = (cv && !gv) ? 1.0 : 0.0;
}

template <bool FINAL>
__global__ __launch_bounds__(kMetricThreads) void metrics_kernel(int n, const float* __restrict__ cells,
                                                                 const float* __restrict__ gt,
                                                                 const double* __restrict__ part_in,
                                                                 double* __restrict__ out)
{
    __shared__ double red[kSums][kMetricThreads];
    double v[kSums];
#pragma unroll
    for (int k = 0; k < kSums; ++k) v[k] = 0.0;
    if (FINAL) {
        for (int b = threadIdx.x; b < n; b += kMetricThreads)
            for (int k = 0; k < kSums; ++k) v[k] += part_in[(size_t)kSums * b + k];
    } else {
        const int i = blockIdx.x * kMetricThreads + threadIdx.x;
        if (i < n) metric_terms(cells[i], gt[i], v);
    }
    for (int k = 0; k < kSums; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int d = kMetricThreads / 2; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d)
            for (int k = 0; k < kSums; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x < kSums) out[(size_t)kSums * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Scratch of the allocation-free entry point (gvc_volume_metrics_noalloc): the
// block partials of up to kStaticBlocks blocks, i.e. samples up to
// kStaticBlocks * kMetricThreads = 2^24, in a file-scope device array (5.8 MB
// per GPU, held until the process ends).  Host code never passes `scratch`
// itself to a kernel: its device address is looked up per device once.
constexpr unsigned kStaticBlocks = 65536;
constexpr int kStaticMaxSamples = (int)(kStaticBlocks * kMetricThreads);
static __device__ __align__(256) double scratch[(size_t)kSums * kStaticBlocks];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (hipGetSymbolAddress with that device
// current, once per device).
int scratch_of_current_device(double** out)
{
    int dev = 0;
    GVC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(GVC_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        GVC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = scratch_ptr[dev];
    return GVC_OK;
}

}  // namespace

int volume_metrics_noalloc(int S, const float* cells, const float* gt, double* out, hipStream_t st)
{
    if (S > kStaticMaxSamples)
        return set_error(GVC_ERR_CAPACITY, "gvc_volume_metrics_noalloc: more than 2^24 samples (its static scratch "
                                           "holds 65536 block partials); use gvc_volume_metrics with a workspace");
    double* part = nullptr;
    GVC_TRY(scratch_of_current_device(&part));
    const unsigned nb = blocks_for((size_t)(S > 0 ? S : 1), kMetricThreads);
    // every partial the final pass reads is written by this call's first pass:
    // nothing has to start at zero
    metrics_kernel<false><<<nb, kMetricThreads, 0, st>>>(S, cells, gt, nullptr, part);
    GVC_LAUNCH_CHECK("metrics");
    metrics_kernel<true><<<1, kMetricThreads, 0, st>>>((int)nb, nullptr, nullptr, part, out);
    GVC_LAUNCH_CHECK("metrics_final");
    return GVC_OK;
}

size_t volume_metrics_bytes(int S)
{
    Arena a(nullptr, 0);
    a.take<double>((size_t)kSums * blocks_for((size_t)(S > 0 ? S : 1), kMetricThreads));
    return a.off + 256;
}

int volume_metrics(int S, const float* cells, const float* gt, double* out, void* ws, size_t bytes, hipStream_t st)
{
    Arena a(ws, bytes);
    const unsigned nb = blocks_for((size_t)(S > 0 ? S : 1), kMetricThreads);
    double* part = a.take<double>((size_t)kSums * nb);
    if (!a.fits()) return set_error(GVC_ERR_WORKSPACE_TOO_SMALL, "volume metrics workspace too small");
    metrics_kernel<false><<<nb, kMetricThreads, 0, st>>>(S, cells, gt, nullptr, part);
    GVC_LAUNCH_CHECK("metrics");
    metrics_kernel<true><<<1, kMetricThreads, 0, st>>>((int)nb, nullptr, nullptr, part, out);
    GVC_LAUNCH_CHECK("metrics_final");
    return GVC_OK;
}

}  // namespace gvc
