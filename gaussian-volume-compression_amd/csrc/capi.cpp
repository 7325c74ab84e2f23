// capi.cpp -- extern "C" entry points of libgvc (declared in include/gvc.h).
// Thin: argument checks, status mapping, then the stream-ordered launchers.
#include <stdio.h>
#include <string.h>

#include "gvc_internal.h"

namespace gvc {

static thread_local char g_last_error[512] = "";

int set_error(int status, const char* what)
{
    snprintf(g_last_error, sizeof(g_last_error), "%s", what ? what : "");
    return status;
}

int hip_status(hipError_t e, const char* what)
{
    snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what ? what : "hip", hipGetErrorString(e));
    return GVC_ERR_HIP;
}

// mixture.hip
int preprocess_launch(int P, const float* means, const float* scales, const float* rots, const float* weights,
                      float k, unsigned flags, float* conics, float* boxes, hipStream_t st);
int preprocess_factors_launch(int P, const float* means, const float* scales, const float* rots, const float* weights,
                              float k, const gvc_params* prm, float* conics, float* boxes, float* factors,
                              hipStream_t st);
int morton_launch(int n, const float* xyz, int clamp, uint64_t* codes, hipStream_t st);
size_t sample_set_bytes(int S);
int sample_set_build(int S, const float* samples, void* set, size_t bytes, hipStream_t st);
size_t forward_state_bytes(int P, int S, const gvc_params* prm);
int mixture_forward(int P, const float* means, const float* scales, const float* rots, const float* values,
                    const float* weights, float k, const void* sample_set, int S, const gvc_params* prm,
                    float* out_cells, float* out_weights, float* conics, void* state, size_t state_bytes,
                    hipStream_t st);
int forward_stats(const void* state, int P, int S, const gvc_params* prm, uint64_t* n_records, uint64_t* n_pairs,
                  uint32_t* error_flags, uint64_t* chunk_slots_used, uint64_t* lane_groups,
                  hipStream_t st);
size_t backward_bytes(int P, uint64_t cap);
int forward_tree(const void* state, int P, int S, const gvc_params* prm, uint32_t* order, uint32_t* keys,
                 int32_t* links, float* boxes, hipStream_t st);
int forward_pair_counts(const void* state, int P, int S, const gvc_params* prm, const void* sample_set,
                        uint32_t* sample_counts, uint32_t* gaussian_counts, hipStream_t st);
int forward_accumulate(const void* state, int P, int S, const gvc_params* prm, unsigned long long* accum,
                       hipStream_t st);
int mixture_backward(int P, const float* means, const float* scales, const float* rots, const float* values,
                     const float* weights, float k, const void* sample_set, int S, const gvc_params* prm,
                     const float* out_cells, const float* out_weights, const float* gcells, const float* gweights,
                     void* state, size_t state_bytes, uint64_t capacity, void* ws, size_t ws_bytes, float* dmeans,
                     float* dscales, float* drots, float* dvalues, float* dweights, hipStream_t st);
int sort_pairs_u64(int n, const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, int key_bits,
                   void* ws, size_t ws_bytes, hipStream_t st);
size_t sort_pairs_u64_bytes(int n);

// batches (mixture.hip)
size_t batch_table_bytes(int V);
int batch_describe(int V, const int* P_counts, const int* S_counts, int L, void* table, size_t bytes, gvc_batch* out,
                   hipStream_t st);
size_t sample_set_bytes_batched(const gvc_batch* b);
int sample_set_build_batched(const gvc_batch* b, const float* samples, void* set, size_t bytes, hipStream_t st);
size_t forward_state_bytes_batched(const gvc_batch* b, const gvc_params* prm);
int mixture_forward_batched(const gvc_batch* b, const float* means, const float* scales, const float* rots,
                            const float* values, const float* weights, float k, const void* sample_set,
                            const gvc_params* prm, float* out_cells, float* out_weights, float* conics, void* state,
                            size_t state_bytes, hipStream_t st);
int forward_stats_batched(const gvc_batch* b, const void* state, const gvc_params* prm, uint64_t* n_records,
                          uint64_t* n_pairs, uint32_t* error_flags, uint64_t* chunk_slots_used, uint64_t* lane_groups,
                          hipStream_t st);
int forward_accumulate_batched(const gvc_batch* b, const void* state, const gvc_params* prm,
                               unsigned long long* accum, hipStream_t st);
int forward_pair_counts_batched(const gvc_batch* b, const void* state, const gvc_params* prm, const void* sample_set,
                                uint32_t* sample_counts, uint32_t* gaussian_counts, hipStream_t st);
int mixture_backward_batched(const gvc_batch* b, const float* scales, const float* rots, float k,
                             const void* sample_set, const gvc_params* prm, const float* out_cells,
                             const float* out_weights, const float* gcells, const float* gweights, void* state,
                             size_t state_bytes, uint64_t capacity, void* ws, size_t ws_bytes, float* dmeans,
                             float* dscales, float* drots, float* dvalues, float* dweights, hipStream_t st);

// training.hip
int activate_forward(int P, const float* ls, const float* q, const float* wl, const float* vl, float max_scale,
                     float* scales, float* qn, float* w, float* v, hipStream_t st);
int activate_backward(int P, const float* ls, const float* q, const float* wl, const float* vl, float max_scale,
                      const float* gs, const float* gq, const float* gw, const float* gv, float* dls, float* dq,
                      float* dwl, float* dvl, hipStream_t st);
size_t loss_state_bytes(int S);
int loss_forward(int S, const float* cells, const float* weights, const float* gt, float fn_reg, float fp_reg,
                 int surface, float* loss, void* state, size_t bytes, hipStream_t st);
int loss_backward(int S, const float* cells, const float* weights, const float* gt, float fn_reg, float fp_reg,
                  int surface, const float* upstream, const void* state, size_t bytes, float* dcells,
                  float* dweights, hipStream_t st);

// mesh.hip, knn.hip
size_t tet_locator_bytes(int n_points, int C, unsigned long long cap);
int tet_locator_build(int n_points, const float* pts, int C, const int* tets, unsigned long long cap, void* loc,
                      size_t bytes, unsigned long long* needed, hipStream_t st);
int tet_locator_bind_values(int n_points, const float* pts, int C, void* loc, size_t bytes, unsigned long long cap,
                            const double* values, hipStream_t st);
int tet_locate_interp(int n_points, const float* pts, int C, const void* loc, size_t bytes, unsigned long long cap,
                      const double* values, bool bound, int S, const float* samples, const void* sample_set,
                      int* cell_ids, double* out_values, float* bary, hipStream_t st);
int grid_sample(const long long dims[3], const float origin[3], const float spacing[3], int dtype, const void* values,
                int S, const float* samples, void* out, hipStream_t st);
int cell_interp(int S, const float* samples, const int* cell_ids, const uint8_t* types, const int* offsets,
                const int* conn, const float* pts, const double* values, double* out, hipStream_t st);
size_t knn3_bytes(int n);
int knn3(int n, const float* pts, float* out, void* ws, size_t bytes, hipStream_t st);

// optim.hip
int adam_step(int n, const gvc_adam_tensor* tensors, gvc_train_clock* clock, double beta1, double beta2,
              double eps, hipStream_t st);
int lr_advance(gvc_train_clock* clock, const gvc_lr_schedule* s, int lr_index, int64_t set_iteration,
               hipStream_t st);
size_t error_order_bytes(int S);
int error_order(int S, const float* cells, const float* gt, int count, const float* weight_logits,
                float low_weight, uint32_t* order, uint32_t* counts, void* ws, size_t bytes, hipStream_t st);
size_t densify_bytes(int total);
int densify_in_empty(int count, int k, const gvc_store* s, const float* samples, const float* gt,
                     const uint32_t* idx, float dist_scale, float init_weight, void* ws, size_t bytes,
                     hipStream_t st);
size_t prune_bytes(int count);
// metrics.hip
size_t volume_metrics_bytes(int S);
int volume_metrics(int S, const float* cells, const float* gt, double* out, void* ws, size_t bytes, hipStream_t st);
int volume_metrics_noalloc(int S, const float* cells, const float* gt, double* out, hipStream_t st);
int prune(int count, const gvc_store* s, const uint8_t* mask, float min_weight, int* new_count, void* ws,
          size_t bytes, hipStream_t st);

}  // namespace gvc

using namespace gvc;

#define GVC_REQUIRE(cond, msg)                                        \
    do {                                                              \
        if (!(cond)) return set_error(GVC_ERR_INVALID_ARGUMENT, msg); \
    } while (0)

extern "C" {

int gvc_version(void) { return 2; }

const char* gvc_status_string(int status)
{
    switch (status) {
    case GVC_OK: return "ok";
    case GVC_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GVC_ERR_WORKSPACE_TOO_SMALL: return "workspace too small";
    case GVC_ERR_HIP: return "HIP runtime error";
    case GVC_ERR_UNSUPPORTED: return "unsupported";
    case GVC_ERR_CAPACITY: return "capacity exceeded";
    default: return "unknown status";
    }
}

const char* gvc_last_error(void) { return g_last_error; }

int gvc_preprocess(int P, const float* means3D, const float* scales, const float* rotations, const float* weights,
                   float scale_modifier, unsigned flags, float* conics, float* boxes, void* stream)
{
    GVC_REQUIRE(P >= 0, "P < 0");
    if (P == 0) return GVC_OK;
    GVC_REQUIRE(means3D && scales && rotations && weights, "NULL input");
    return preprocess_launch(P, means3D, scales, rotations, weights, scale_modifier, flags, conics, boxes,
                             (hipStream_t)stream);
}

int gvc_preprocess_factors(int P, const float* means3D, const float* scales, const float* rotations,
                           const float* weights, float scale_modifier, const gvc_params* params, float* conics,
                           float* boxes, float* factors, void* stream)
{
    GVC_REQUIRE(P >= 0, "P < 0");
    GVC_REQUIRE(params, "NULL params");
    if (P == 0) return GVC_OK;
    GVC_REQUIRE(means3D && scales && rotations && weights, "NULL input");
    return preprocess_factors_launch(P, means3D, scales, rotations, weights, scale_modifier, params, conics, boxes,
                                     factors, (hipStream_t)stream);
}

int gvc_morton3d_unit(int n, const float* xyz, int clamp, uint64_t* codes, void* stream)
{
    GVC_REQUIRE(n >= 0, "n < 0");
    if (n == 0) return GVC_OK;
    GVC_REQUIRE(xyz && codes, "NULL pointer");
    return morton_launch(n, xyz, clamp, codes, (hipStream_t)stream);
}

size_t gvc_sort_pairs_u64_bytes(int n) { return n > 0 ? sort_pairs_u64_bytes(n) : 256; }

int gvc_sort_pairs_u64(int n, const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                       uint32_t* vals_out, int key_bits, void* workspace, size_t workspace_bytes, void* stream)
{
    GVC_REQUIRE(n >= 0, "n < 0");
    GVC_REQUIRE(key_bits >= 0 && key_bits <= 64, "key_bits must be in [0, 64]");
    if (n == 0) return GVC_OK;
    GVC_REQUIRE(keys_in && vals_in && keys_out && vals_out && workspace, "NULL pointer");
    return sort_pairs_u64(n, keys_in, vals_in, keys_out, vals_out, key_bits, workspace, workspace_bytes,
                          (hipStream_t)stream);
}

size_t gvc_sample_set_bytes(int S) { return S >= 0 ? sample_set_bytes(S) : 0; }

int gvc_sample_set_build(int S, const float* samples, void* sample_set, size_t sample_set_bytes_, void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    GVC_REQUIRE(sample_set, "NULL sample set");
    GVC_REQUIRE(S == 0 || samples, "NULL samples");
    return sample_set_build(S, samples, sample_set, sample_set_bytes_, (hipStream_t)stream);
}

size_t gvc_forward_state_bytes(int P, int S, const gvc_params* params)
{
    if (P < 0 || S < 0 || !params) return 0;
    return forward_state_bytes(P, S, params);
}

int gvc_mixture_forward(int P, const float* means3D, const float* scales, const float* rotations,
                        const float* values, const float* weights, float scale_modifier, const void* sample_set, int S,
                        const gvc_params* params, float* out_cells, float* out_weights, float* conics, void* state,
                        size_t state_bytes, void* stream)
{
    GVC_REQUIRE(P >= 0 && S >= 0, "negative size");
    GVC_REQUIRE(sample_set && state, "NULL sample set or state");
    GVC_REQUIRE(S == 0 || (out_cells && out_weights), "NULL outputs");
    GVC_REQUIRE(P == 0 || (means3D && scales && rotations && values && weights), "NULL Gaussian input");
    return mixture_forward(P, means3D, scales, rotations, values, weights, scale_modifier, sample_set, S, params,
                           out_cells, out_weights, conics, state, state_bytes, (hipStream_t)stream);
}

int gvc_forward_stats(const void* state, int P, int S, const gvc_params* params, uint64_t* n_records,
                      uint64_t* n_pairs, uint32_t* error_flags, uint64_t* chunk_slots_used, uint64_t* lane_groups,
                      void* stream)
{
    GVC_REQUIRE(state, "NULL state");
    return forward_stats(state, P, S, params, n_records, n_pairs, error_flags, chunk_slots_used,
                         lane_groups, (hipStream_t)stream);
}

int gvc_forward_accumulate(const void* state, int P, int S, const gvc_params* params, uint64_t* accum, void* stream)
{
    GVC_REQUIRE(state && accum, "NULL pointer");
    return forward_accumulate(state, P, S, params, reinterpret_cast<unsigned long long*>(accum), (hipStream_t)stream);
}

int gvc_forward_tree(const void* state, int P, int S, const gvc_params* params, uint32_t* order, uint32_t* keys,
                     int32_t* nodes, float* boxes, void* stream)
{
    GVC_REQUIRE(state && params, "NULL pointer");
    return forward_tree(state, P, S, params, order, keys, nodes, boxes, (hipStream_t)stream);
}

int gvc_forward_pair_counts(const void* state, int P, int S, const gvc_params* params, const void* sample_set,
                            uint32_t* sample_counts, uint32_t* gaussian_counts, void* stream)
{
    GVC_REQUIRE(state && params && sample_set && (S == 0 || sample_counts) && (P == 0 || gaussian_counts),
                "NULL pointer");
    return forward_pair_counts(state, P, S, params, sample_set, sample_counts, gaussian_counts, (hipStream_t)stream);
}

size_t gvc_backward_bytes(int P, uint64_t record_capacity)
{
    return P >= 0 ? backward_bytes(P, record_capacity) : 0;
}

int gvc_mixture_backward(int P, const float* means3D, const float* scales, const float* rotations,
                         const float* values, const float* weights, float scale_modifier, const void* sample_set, int S,
                         const gvc_params* params, const float* out_cells, const float* out_weights,
                         const float* dL_dout_cells, const float* dL_dout_cell_weights, void* state, size_t state_bytes,
                         uint64_t record_capacity, void* workspace, size_t workspace_bytes, float* dL_dmeans3D,
                         float* dL_dscales, float* dL_drotations, float* dL_dvalues, float* dL_dweights, void* stream)
{
    GVC_REQUIRE(P >= 0 && S >= 0, "negative size");
    if (P == 0) return GVC_OK;
    GVC_REQUIRE(sample_set && state && workspace, "NULL sample set, state or workspace");
    GVC_REQUIRE(scales && rotations, "NULL scales or rotations");
    GVC_REQUIRE(S == 0 || (out_cells && out_weights && dL_dout_cells && dL_dout_cell_weights), "NULL per-sample input");
    GVC_REQUIRE(dL_dmeans3D && dL_dscales && dL_drotations && dL_dvalues && dL_dweights, "NULL gradient output");
    return mixture_backward(P, means3D, scales, rotations, values, weights, scale_modifier, sample_set, S, params,
                            out_cells, out_weights, dL_dout_cells, dL_dout_cell_weights, state, state_bytes, record_capacity,
                            workspace, workspace_bytes, dL_dmeans3D, dL_dscales, dL_drotations, dL_dvalues,
                            dL_dweights, (hipStream_t)stream);
}

size_t gvc_tet_locator_bytes(int n_points, int n_tets, uint64_t entry_capacity)
{
    if (n_points < 0 || n_tets < 0) return 0;
    return tet_locator_bytes(n_points, n_tets, entry_capacity);
}

int gvc_tet_locator_build(int n_points, const float* points, int n_tets, const int32_t* tets, uint64_t entry_capacity,
                          void* locator, size_t locator_bytes, uint64_t* entries_needed, void* stream)
{
    GVC_REQUIRE(n_points > 0 && n_tets > 0, "empty mesh");
    GVC_REQUIRE(points && tets && locator, "NULL pointer");
    GVC_REQUIRE(entry_capacity <= 0xffffffffull, "entry capacity above 2^32");
    unsigned long long need = 0;
    int st = tet_locator_build(n_points, points, n_tets, tets, entry_capacity, locator, locator_bytes, &need,
                               (hipStream_t)stream);
    if (entries_needed) *entries_needed = need;
    return st;
}

int gvc_tet_locate_interp(int n_points, const float* points, int n_tets, const void* locator, size_t locator_bytes,
                          uint64_t entry_capacity, const double* point_values, int S, const float* samples,
                          int32_t* cell_ids, double* values_out, float* bary_out, void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    if (S == 0) return GVC_OK;
    GVC_REQUIRE(points && locator && samples && cell_ids, "NULL pointer");
    GVC_REQUIRE(!values_out || point_values, "values_out needs point_values");
    return tet_locate_interp(n_points, points, n_tets, locator, locator_bytes, entry_capacity, point_values, false,
                             S, samples, nullptr, cell_ids, values_out, bary_out, (hipStream_t)stream);
}

int gvc_tet_locate_interp_set(int n_points, const float* points, int n_tets, const void* locator,
                              size_t locator_bytes, uint64_t entry_capacity, const double* point_values,
                              const void* sample_set, int S, int32_t* cell_ids, double* values_out, float* bary_out,
                              void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    if (S == 0) return GVC_OK;
    GVC_REQUIRE(points && locator && sample_set && cell_ids, "NULL pointer");
    GVC_REQUIRE(!values_out || point_values, "values_out needs point_values");
    return tet_locate_interp(n_points, points, n_tets, locator, locator_bytes, entry_capacity, point_values, false,
                             S, nullptr, sample_set, cell_ids, values_out, bary_out, (hipStream_t)stream);
}

int gvc_tet_locator_bind_values(int n_points, const float* points, int n_tets, void* locator, size_t locator_bytes,
                                uint64_t entry_capacity, const double* point_values, void* stream)
{
    GVC_REQUIRE(n_points > 0 && n_tets > 0, "empty mesh");
    GVC_REQUIRE(points && locator && point_values, "NULL pointer");
    return tet_locator_bind_values(n_points, points, n_tets, locator, locator_bytes, entry_capacity, point_values,
                                   (hipStream_t)stream);
}

int gvc_tet_locate_bound(int n_points, const float* points, int n_tets, const void* locator, size_t locator_bytes,
                         uint64_t entry_capacity, int S, const float* samples, const void* sample_set,
                         int32_t* cell_ids, double* values_out, float* bary_out, void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    if (S == 0) return GVC_OK;
    GVC_REQUIRE(points && locator && cell_ids && (samples || sample_set), "NULL pointer");
    return tet_locate_interp(n_points, points, n_tets, locator, locator_bytes, entry_capacity, nullptr, true, S,
                             sample_set ? nullptr : samples, sample_set, cell_ids, values_out, bary_out,
                             (hipStream_t)stream);
}

int gvc_cell_interp(int S, const float* samples, const int32_t* cell_ids, const uint8_t* cell_types,
                    const int32_t* cell_offsets, const int32_t* connectivity, const float* points,
                    const double* point_values, double* values_out, void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    if (S == 0) return GVC_OK;
    GVC_REQUIRE(samples && cell_ids && cell_types && cell_offsets && connectivity && points && point_values &&
                    values_out,
                "NULL pointer");
    return cell_interp(S, samples, cell_ids, cell_types, cell_offsets, connectivity, points, point_values, values_out,
                       (hipStream_t)stream);
}

int gvc_grid_sample(const int64_t* dims, const float* origin, const float* spacing, int value_dtype,
                    const void* values, int S, const float* samples, void* out, void* stream)
{
    GVC_REQUIRE(dims && origin && spacing, "NULL grid description");
    GVC_REQUIRE(S >= 0, "S < 0");
    if (S == 0) return GVC_OK;
    GVC_REQUIRE(values && samples && out, "NULL pointer");
    const long long d[3] = { (long long)dims[0], (long long)dims[1], (long long)dims[2] };
    return grid_sample(d, origin, spacing, value_dtype, values, S, samples, out, (hipStream_t)stream);
}

size_t gvc_knn3_bytes(int n) { return n >= 0 ? knn3_bytes(n) : 0; }

int gvc_knn3(int n, const float* points, float* mean_sq_dist, void* workspace, size_t workspace_bytes, void* stream)
{
    GVC_REQUIRE(n >= 0, "n < 0");
    if (n == 0) return GVC_OK;
    GVC_REQUIRE(points && mean_sq_dist && workspace, "NULL pointer");
    return knn3(n, points, mean_sq_dist, workspace, workspace_bytes, (hipStream_t)stream);
}

size_t gvc_batch_table_bytes(int volumes) { return batch_table_bytes(volumes); }

int gvc_batch_describe(int volumes, const int* gaussians_per_volume, const int* samples_per_volume, int leaf_size,
                       void* table, size_t table_bytes, gvc_batch* out, void* stream)
{
    GVC_REQUIRE(gaussians_per_volume && samples_per_volume && table && out, "NULL pointer");
    return batch_describe(volumes, gaussians_per_volume, samples_per_volume, leaf_size, table, table_bytes, out,
                          (hipStream_t)stream);
}

#define GVC_REQUIRE_BATCH(b)                                                                                  \
    GVC_REQUIRE((b) && (b)->table && (b)->volumes >= 1, "batch not described (gvc_batch_describe)")

size_t gvc_sample_set_bytes_batched(const gvc_batch* batch)
{
    return (batch && batch->volumes >= 1) ? sample_set_bytes_batched(batch) : 0;
}

int gvc_sample_set_build_batched(const gvc_batch* batch, const float* samples, void* sample_set, size_t bytes,
                                 void* stream)
{
    GVC_REQUIRE_BATCH(batch);
    GVC_REQUIRE(sample_set && (batch->S_total == 0 || samples), "NULL pointer");
    return sample_set_build_batched(batch, samples, sample_set, bytes, (hipStream_t)stream);
}

size_t gvc_forward_state_bytes_batched(const gvc_batch* batch, const gvc_params* params)
{
    return (batch && batch->volumes >= 1 && params) ? forward_state_bytes_batched(batch, params) : 0;
}

int gvc_mixture_forward_batched(const gvc_batch* batch, const float* means3D, const float* scales,
                                const float* rotations, const float* values, const float* weights,
                                float scale_modifier, const void* sample_set, const gvc_params* params,
                                float* out_cells, float* out_weights, float* conics, void* state, size_t state_bytes,
                                void* stream)
{
    GVC_REQUIRE_BATCH(batch);
    GVC_REQUIRE(means3D && scales && rotations && values && weights && sample_set && state && out_cells &&
                    out_weights,
                "NULL pointer");
    return mixture_forward_batched(batch, means3D, scales, rotations, values, weights, scale_modifier, sample_set,
                                   params, out_cells, out_weights, conics, state, state_bytes, (hipStream_t)stream);
}

int gvc_forward_stats_batched(const gvc_batch* batch, const void* state, const gvc_params* params,
                              uint64_t* n_records, uint64_t* n_pairs, uint32_t* error_flags,
                              uint64_t* chunk_slots_used, uint64_t* lane_groups, void* stream)
{
    GVC_REQUIRE_BATCH(batch);
    GVC_REQUIRE(state, "NULL state");
    return forward_stats_batched(batch, state, params, n_records, n_pairs, error_flags, chunk_slots_used,
                                 lane_groups, (hipStream_t)stream);
}

int gvc_forward_pair_counts_batched(const gvc_batch* batch, const void* state, const gvc_params* params,
                                    const void* sample_set, uint32_t* sample_counts, uint32_t* gaussian_counts,
                                    void* stream)
{
    GVC_REQUIRE_BATCH(batch);
    GVC_REQUIRE(state && params && sample_set && (batch->S_total == 0 || sample_counts) &&
                    (batch->P_total == 0 || gaussian_counts),
                "NULL pointer");
    return forward_pair_counts_batched(batch, state, params, sample_set, sample_counts, gaussian_counts,
                                       (hipStream_t)stream);
}

int gvc_forward_accumulate_batched(const gvc_batch* batch, const void* state, const gvc_params* params,
                                   uint64_t* accum, void* stream)
{
    GVC_REQUIRE_BATCH(batch);
    GVC_REQUIRE(state && accum, "NULL pointer");
    return forward_accumulate_batched(batch, state, params, reinterpret_cast<unsigned long long*>(accum),
                                      (hipStream_t)stream);
}

int gvc_mixture_backward_batched(const gvc_batch* batch, const float* scales, const float* rotations,
                                 float scale_modifier, const void* sample_set, const gvc_params* params,
                                 const float* out_cells, const float* out_weights, const float* dL_dout_cells,
                                 const float* dL_dout_cell_weights, void* state, size_t state_bytes,
                                 uint64_t record_capacity, void* workspace, size_t workspace_bytes,
                                 float* dL_dmeans3D, float* dL_dscales, float* dL_drotations, float* dL_dvalues,
                                 float* dL_dweights, void* stream)
{
    GVC_REQUIRE_BATCH(batch);
    GVC_REQUIRE(scales && rotations && sample_set && out_cells && out_weights && dL_dout_cells &&
                    dL_dout_cell_weights && state && workspace && dL_dmeans3D && dL_dscales && dL_drotations &&
                    dL_dvalues && dL_dweights,
                "NULL pointer");
    return mixture_backward_batched(batch, scales, rotations, scale_modifier, sample_set, params, out_cells,
                                    out_weights, dL_dout_cells, dL_dout_cell_weights, state, state_bytes,
                                    record_capacity, workspace, workspace_bytes, dL_dmeans3D, dL_dscales,
                                    dL_drotations, dL_dvalues, dL_dweights, (hipStream_t)stream);
}

int gvc_activate_forward(int P, const float* log_scales, const float* rotations, const float* weight_logits,
                         const float* value_logits, float max_scale, float* scales, float* rotations_n,
                         float* weights, float* values, void* stream)
{
    GVC_REQUIRE(P >= 0, "P < 0");
    if (P == 0) return GVC_OK;
    GVC_REQUIRE(log_scales && rotations && weight_logits && value_logits && scales && rotations_n && weights && values,
                "NULL pointer");
    return activate_forward(P, log_scales, rotations, weight_logits, value_logits, max_scale, scales, rotations_n,
                            weights, values, (hipStream_t)stream);
}

int gvc_activate_backward(int P, const float* log_scales, const float* rotations, const float* weight_logits,
                          const float* value_logits, float max_scale, const float* dL_dscales,
                          const float* dL_drotations_n, const float* dL_dweights, const float* dL_dvalues,
                          float* dL_dlog_scales, float* dL_drotations, float* dL_dweight_logits,
                          float* dL_dvalue_logits, void* stream)
{
    GVC_REQUIRE(P >= 0, "P < 0");
    if (P == 0) return GVC_OK;
    GVC_REQUIRE(log_scales && rotations && weight_logits && value_logits && dL_dscales && dL_drotations_n &&
                    dL_dweights && dL_dvalues && dL_dlog_scales && dL_drotations && dL_dweight_logits &&
                    dL_dvalue_logits,
                "NULL pointer");
    return activate_backward(P, log_scales, rotations, weight_logits, value_logits, max_scale, dL_dscales,
                             dL_drotations_n, dL_dweights, dL_dvalues, dL_dlog_scales, dL_drotations,
                             dL_dweight_logits, dL_dvalue_logits, (hipStream_t)stream);
}

size_t gvc_loss_state_bytes(int S) { return S >= 0 ? loss_state_bytes(S) : 0; }

int gvc_loss_forward(int S, const float* cells, const float* weights, const float* gt, float fn_reg, float fp_reg,
                     int encode_surface, float* loss, void* state, size_t state_bytes, void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    GVC_REQUIRE(loss && state && (S == 0 || (cells && weights && gt)), "NULL pointer");
    return loss_forward(S, cells, weights, gt, fn_reg, fp_reg, encode_surface, loss, state, state_bytes,
                        (hipStream_t)stream);
}

int gvc_loss_backward(int S, const float* cells, const float* weights, const float* gt, float fn_reg,
                      float fp_reg, int encode_surface, const float* dL_dloss, const void* state,
                      size_t state_bytes, float* dL_dcells, float* dL_dweights, void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    if (S == 0) return GVC_OK;
    GVC_REQUIRE(cells && weights && gt && state && dL_dcells && dL_dweights, "NULL pointer");
    return loss_backward(S, cells, weights, gt, fn_reg, fp_reg, encode_surface, dL_dloss, state, state_bytes,
                         dL_dcells, dL_dweights, (hipStream_t)stream);
}

int gvc_lr_advance(gvc_train_clock* clock, const gvc_lr_schedule* schedule, int lr_index, int64_t set_iteration,
                   void* stream)
{
    GVC_REQUIRE(clock && schedule, "NULL pointer");
    GVC_REQUIRE(lr_index >= 0 && lr_index < GVC_ADAM_MAX_TENSORS, "lr_index out of range");
    GVC_REQUIRE(schedule->max_steps > 0, "max_steps <= 0");
    return lr_advance(clock, schedule, lr_index, set_iteration, (hipStream_t)stream);
}

int gvc_adam_step(int n_tensors, const gvc_adam_tensor* tensors, gvc_train_clock* clock, double beta1,
                  double beta2, double eps, void* stream)
{
    GVC_REQUIRE(n_tensors >= 0 && n_tensors <= GVC_ADAM_MAX_TENSORS, "n_tensors out of range");
    if (n_tensors == 0) return GVC_OK;
    GVC_REQUIRE(tensors && clock, "NULL pointer");
    GVC_REQUIRE(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0, "betas outside [0, 1)");
    return adam_step(n_tensors, tensors, clock, beta1, beta2, eps, (hipStream_t)stream);
}

size_t gvc_error_order_bytes(int S) { return S >= 0 ? error_order_bytes(S) : 0; }

int gvc_error_order(int S, const float* cells, const float* gt, int count, const float* weight_logits,
                    float low_weight, uint32_t* order, uint32_t* counts, void* workspace, size_t workspace_bytes,
                    void* stream)
{
    GVC_REQUIRE(S >= 0 && count >= 0, "S < 0 or count < 0");
    GVC_REQUIRE(counts && workspace && (S == 0 || (cells && gt && order)), "NULL pointer");
    GVC_REQUIRE(count == 0 || weight_logits, "NULL weight_logits");
    return error_order(S, cells, gt, count, weight_logits, low_weight, order, counts, workspace, workspace_bytes,
                       (hipStream_t)stream);
}

static bool store_ok(const gvc_store* s)
{
    if (!s || !s->xyz || !s->weight || !s->scaling || !s->rotation || !s->value) return false;
    // moments all present or all absent
    int have = 0;
    for (int k = 0; k < 5; ++k) have += (s->exp_avg[k] != nullptr) + (s->exp_avg_sq[k] != nullptr);
    return have == 0 || have == 10;
}

size_t gvc_densify_bytes(int total) { return total >= 0 ? densify_bytes(total) : 0; }

int gvc_densify_in_empty(int count, int k, const gvc_store* store, const float* samples, const float* gt,
                         const uint32_t* idx, float dist_scale, float init_weight, void* workspace,
                         size_t workspace_bytes, void* stream)
{
    GVC_REQUIRE(count >= 0 && k >= 0, "count < 0 or k < 0");
    GVC_REQUIRE(store_ok(store), "incomplete gvc_store");
    GVC_REQUIRE((long long)count + k <= store->capacity, "count + k exceeds the store capacity");
    if (k == 0) return GVC_OK;
    GVC_REQUIRE(samples && gt && workspace, "NULL pointer");
    return densify_in_empty(count, k, store, samples, gt, idx, dist_scale, init_weight, workspace, workspace_bytes,
                            (hipStream_t)stream);
}

size_t gvc_prune_bytes(int count) { return count >= 0 ? prune_bytes(count) : 0; }

int gvc_prune(int count, const gvc_store* store, const uint8_t* prune_mask, float min_weight, int* new_count,
              void* workspace, size_t workspace_bytes, void* stream)
{
    GVC_REQUIRE(count >= 0, "count < 0");
    GVC_REQUIRE(store_ok(store), "incomplete gvc_store");
    GVC_REQUIRE(count <= store->capacity, "count exceeds the store capacity");
    GVC_REQUIRE(new_count && workspace, "NULL pointer");
    return prune(count, store, prune_mask, min_weight, new_count, workspace, workspace_bytes, (hipStream_t)stream);
}

size_t gvc_volume_metrics_bytes(int S) { return S >= 0 ? volume_metrics_bytes(S) : 0; }

int gvc_volume_metrics(int S, const float* cells, const float* gt, double* out, void* workspace,
                       size_t workspace_bytes, void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    GVC_REQUIRE(out && workspace && (S == 0 || (cells && gt)), "NULL pointer");
    return volume_metrics(S, cells, gt, out, workspace, workspace_bytes, (hipStream_t)stream);
}

int gvc_volume_metrics_noalloc(int S, const float* cells, const float* gt, double* out, void* stream)
{
    GVC_REQUIRE(S >= 0, "S < 0");
    GVC_REQUIRE(out && (S == 0 || (cells && gt)), "NULL pointer");
    return volume_metrics_noalloc(S, cells, gt, out, (hipStream_t)stream);
}

}  // extern "C"
