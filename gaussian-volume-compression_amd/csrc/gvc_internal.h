// gvc_internal.h -- host-side plumbing shared by the HIP sources of libgvc.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/gvc.h"

namespace gvc {

// Record a HIP error text for gvc_last_error() and map it to a status.
int set_error(int status, const char* what);
int hip_status(hipError_t e, const char* what);

#define GVC_HIP(call)                                              \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return ::gvc::hip_status(_e, #call); \
    } while (0)

#define GVC_TRY(call)                 \
    do {                              \
        int _s = (call);              \
        if (_s != GVC_OK) return _s;  \
    } while (0)

#define GVC_LAUNCH_CHECK(name) GVC_HIP(hipGetLastError())

// Bump allocator over a caller-owned workspace.  With base == nullptr it only
// measures (size queries run the same layout code as the real call).
struct Arena {
    char* base;
    size_t cap;
    size_t off = 0;
    Arena(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
    template <class T>
    T* take(size_t n)
    {
        off = (off + 255) & ~size_t(255);
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += n * sizeof(T);
        return p;
    }
    bool fits() const { return base == nullptr || off <= cap; }
};

// ---- util.hip --------------------------------------------------------------
int fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st);
// Exclusive scan of n uint32 into out (in-place allowed); out[n] = total when
// with_total (out must then hold n+1).  Scratch from scan_scratch_elems(n).
size_t scan_scratch_elems(size_t n);
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, bool with_total, uint32_t* scratch,
                       hipStream_t st);

// Bounding box of n float3 points (NaN coordinates skipped) as order-preserving
// uint keys: keys[0..2] = min, keys[3..5] = max (deterministic integer atomics).
int bbox_launch(const float* pts, int n, unsigned* keys, hipStream_t st);
void bbox_from_keys(const unsigned keys[6], float lo[3], float hi[3]);

// ---- mixture.hip ------------------------------------------------------------
// Sorted points (x, y, z, original index as int bits) of a sample set, padded with
// NaN points (index -1) to whole 64-point groups.
const float4* sample_set_points(const void* sample_set, int S);

// ---- radix_sort.hip --------------------------------------------------------
// Stable LSD radix sort of (key, value) pairs on the low key_bits bits.
// Scratch elements (uint32) needed for n items in `segments` segments.  n_dev
// (optional, device memory): sort only the first min(n, *n_dev) items.
// vals_identity: vals is not read; item i's value is i.  seg_off (optional,
// device memory, segments + 1 item offsets): the items form contiguous segments,
// each sorted on its own (the order of the segments is kept).  inv_out
// (optional, n entries; values a permutation of 0..n-1): inv_out[value] = its
// sorted position, written by the last pass.
size_t radix_scratch_elems(size_t n, int segments = 1);
int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n,
                   int key_bits, uint32_t* scratch, bool* result_in_alt, hipStream_t st,
                   const uint32_t* n_dev = nullptr, bool vals_identity = false, const int* seg_off = nullptr,
                   int segments = 1, uint32_t* inv_out = nullptr);
int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, size_t n,
                   int key_bits, uint32_t* scratch, bool* result_in_alt, hipStream_t st);

}  // namespace gvc
