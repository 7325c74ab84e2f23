/*
 * gvc.h -- C ABI of libgvc, the MI355X (gfx950) implementation of the
 * differentiable Gaussian-mixture evaluator, its mesh sampler and its 3-NN
 * initialiser.
 *
 * Conventions for every entry point:
 *   - plain device pointers and sizes; no torch or HIP types in the signatures;
 *     `stream` is a hipStream_t passed as void* (NULL = the null stream);
 *   - nothing allocates device memory and nothing synchronises with the host,
 *     except the two *_stats / *_count readers that say so;  scratch memory is a
 *     caller-owned workspace whose size the matching *_bytes query returns;
 *   - state lives in caller-owned buffers (a "sample set", a "forward state"),
 *     never in library statics, so calls on different buffers are independent;
 *   - return value: GVC_OK or a gvc_status code; gvc_last_error() gives the text
 *     of the last failure on the calling thread.
 *
 * Citations: R/ = the reference repository root, DGR/ = R/diff-gaussian-
 * rasterization/, CR/ = DGR/cuda_rasterizer/, GS/ = R/gpu_sampling/.
 */
#ifndef GVC_H_
#define GVC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gvc_status {
    GVC_OK = 0,
    GVC_ERR_INVALID_ARGUMENT = 1,
    GVC_ERR_WORKSPACE_TOO_SMALL = 2,
    GVC_ERR_HIP = 3,
    GVC_ERR_UNSUPPORTED = 4,
    GVC_ERR_CAPACITY = 5
};

/* gvc_params.flags */
#define GVC_FLAG_KAHAN 1u      /* compensated per-point sums in the forward       */
#define GVC_FLAG_FP64_COND 2u  /* FP64 covariance inverse where cond(Sigma) > 1e6
                                  (gvc_params.fp64_scale_ratio squared); in the
                                  mixture calls those Gaussians' pairs are then
                                  evaluated from the Cholesky factor of that FP64
                                  conic in compensated fp32 (the precise path).
                                  Gaussians above gvc_params.factor_scale_ratio
                                  (cond > 100 by default) but below the FP64
                                  threshold take the precise path too, with the
                                  factor computed in float-float arithmetic (no
                                  FP64)                                          */
#define GVC_FLAG_CANDIDATE_LOG 4u  /* forward keeps a candidate log that the backward
                                     * replays; clear: the backward walks the tree again */
/* How a 64-point group is walked.  Default: the group shares one walk of its
 * box, except where its points are sparse next to its Gaussians (expected
 * candidates of the shared walk above 8192), where every lane walks the tree for
 * its own point.  These two flags force one way for every group (not both). */
#define GVC_FLAG_LANE_NEVER 8u
#define GVC_FLAG_LANE_ALWAYS 16u

/* gvc_params.stages: 0 runs every stage of a call; otherwise only the selected
 * ones, in order, so a caller can bracket each stage with events. */
#define GVC_STAGE_BUILD 1u   /* forward: Morton sort, records, tree, node boxes     */
#define GVC_STAGE_WALK 2u    /* forward: walk; backward: partial records            */
#define GVC_STAGE_COUNTS 4u  /* forward: scan of the per-group record counts        */
#define GVC_STAGE_REDUCE 1u  /* backward: sort, segments, FP64 sums, chain          */

/* Launch/tuning parameters of the mixture kernels (chosen by the autotuner). */
typedef struct gvc_params {
    int leaf_size;        /* Gaussians per BVH leaf: 1..64                        */
    int block_threads;    /* threads per block of the walk kernels: 64..1024      */
    unsigned flags;       /* GVC_FLAG_*                                           */
    unsigned stages;      /* GVC_STAGE_* mask, 0 = all                            */
    unsigned chunk_slots; /* candidate-log slots per 64-point group, 0 = 16       */
    unsigned run_points;  /* most points walked together: 4, 8, 16, 32, 64 (0 = 64) */
    float fp64_scale_ratio; /* GVC_FLAG_FP64_COND threshold on s_max / s_min:
                               0 = 1000 (cond(Sigma) > 1e6), else in [1, 1e18]   */
    float factor_scale_ratio; /* GVC_FLAG_FP64_COND: s_max / s_min above which a
                                 Gaussian below the FP64 threshold takes the
                                 float-float factor: 0 = 10, else in [1, 1e18]   */
} gvc_params;

/* Error flags reported by gvc_forward_stats (bitwise OR). */
#define GVC_ERRFLAG_STACK_OVERFLOW 1u
#define GVC_ERRFLAG_RECORD_OVERFLOW 2u
#define GVC_ERRFLAG_CHUNK_OVERFLOW 4u  /* candidate log full: rerun the forward with more chunk_slots */

int gvc_version(void);
const char *gvc_status_string(int status);
const char *gvc_last_error(void);

/* Per-Gaussian conic and axis-aligned box, the canonical fp32 preprocess.
 * Replaces preprocessCUDA, CR/forward.cu:10-114 (launched :285).
 * conics: [P,6] packed xx xy xz yy yz zz (may be NULL); boxes: [P,6] lo.xyz hi.xyz
 * (may be NULL).  flags: GVC_FLAG_FP64_COND. */
int gvc_preprocess(int P, const float *means3D, const float *scales, const float *rotations,
                   const float *weights, float scale_modifier, unsigned flags, float *conics,
                   float *boxes, void *stream);

/* gvc_preprocess with the precise path's factors: params->flags (GVC_FLAG_FP64_COND)
 * and its two ratios select the paths as in the mixture calls; factors [P,12] (may
 * be NULL) = u00 u01 u02 u11 u12 u22 hi parts then lo parts of the conic's Cholesky
 * factor for Gaussians on the precise path, NaN elsewhere.  For inspection and
 * tests of the precise path (CR/forward.cu:10-114 has no counterpart of it). */
int gvc_preprocess_factors(int P, const float *means3D, const float *scales, const float *rotations,
                           const float *weights, float scale_modifier, const gvc_params *params, float *conics,
                           float *boxes, float *factors, void *stream);

/* 63-bit Morton codes of [n,3] points, the reference's morton3d_unit,
 * R/gaussian_renderer/__init__.py:45-58 (clamp = 0) and the clamped variant of
 * R/train.py:300-309 (clamp = 1). */
int gvc_morton3d_unit(int n, const float *xyz, int clamp, uint64_t *codes, void *stream);

/* Stable sort of (u64 key, u32 value) pairs on the low key_bits bits; the sorted
 * pairs are written to keys_out/vals_out.  Replaces torch.argsort in
 * R/gaussian_renderer/__init__.py:77 and R/train.py:311 (those are unstable; this
 * one orders equal keys by input position). */
size_t gvc_sort_pairs_u64_bytes(int n);
int gvc_sort_pairs_u64(int n, const uint64_t *keys_in, const uint32_t *vals_in, uint64_t *keys_out,
                       uint32_t *vals_out, int key_bits, void *workspace, size_t workspace_bytes,
                       void *stream);

/* Sample set: the points the mixture is evaluated at, spatially sorted into
 * 64-point groups.  Replaces BuildBVH/buildBoxes, DGR/ext.cu:13-65, whose static
 * sample BVH becomes this caller-owned buffer. */
size_t gvc_sample_set_bytes(int S);
int gvc_sample_set_build(int S, const float *samples, void *sample_set, size_t sample_set_bytes,
                         void *stream);

/* Forward: (cells, weights) of the normalised mixture at the S samples.
 * Replaces _C.rasterize_gaussians = RasterizeGaussiansCUDA, DGR/ext.cu:67-105,
 * DGR/rasterize_points.cu:17-73, CR/rasterizer_impl.cu:27-239 (kernels K3, K4,
 * K5/K7, K6: CR/forward.cu:10-274).  out_cells/out_weights: [S]; sentinel (-1, 0)
 * where the total weight is <= 1e-2; P == 0 leaves zeros (as the reference).
 * conics: optional [P,6] output (the reference's third forward result).
 * `state` (gvc_forward_state_bytes) keeps the Gaussian records and BVH and the
 * per-group record counts for the backward, and with GVC_FLAG_CANDIDATE_LOG the
 * candidate log: every 64-candidate chunk a group's walk visited, with its
 * 64x64 box-hit bit mask, in params->chunk_slots x groups slots of 832 bytes.
 * It must not be reused until the backward ran. */
size_t gvc_forward_state_bytes(int P, int S, const gvc_params *params);
int gvc_mixture_forward(int P, const float *means3D, const float *scales, const float *rotations,
                        const float *values, const float *weights, float scale_modifier,
                        const void *sample_set, int S, const gvc_params *params, float *out_cells,
                        float *out_weights, float *conics, void *state, size_t state_bytes,
                        void *stream);

/* BLOCKING: copies the forward's counters to the host (one small D2H copy and a
 * stream synchronisation).  n_records = (Gaussian, 64-point group) partials the
 * backward will emit; n_pairs = accepted (Gaussian, point) pairs; error_flags =
 * GVC_ERRFLAG_* raised so far on this state; chunk_slots_used = candidate-log
 * slots the forward asked for (a chunk_slots of ceil(used / groups) suffices
 * for the same inputs; a lower bound when the log overflowed); lane_groups =
 * groups walked per lane.  Any pointer may be NULL. */
int gvc_forward_stats(const void *state, int P, int S, const gvc_params *params, uint64_t *n_records,
                      uint64_t *n_pairs, uint32_t *error_flags, uint64_t *chunk_slots_used,
                      uint64_t *lane_groups, void *stream);

/* Stream-ordered, non-blocking counterpart of gvc_forward_stats for timed or
 * graph-captured loops: accum (3 x uint64 on the device) gets accum[0] += pairs,
 * accum[1] += records, accum[2] |= error flags of this state so far (call it after
 * the backward to include the backward's flags).  No counterpart in the reference:
 * it replaces the reference's blocking per-call pair-count read-back
 * (CR/rasterizer_impl.cu:163-166) for callers that must not synchronise. */
int gvc_forward_accumulate(const void *state, int P, int S, const gvc_params *params, uint64_t *accum,
                           void *stream);

/* The Gaussian BVH a forward built into `state` (single volume), copied out for
 * inspection: the counterpart of the reference's static Gaussian BVH
 * (DGR/ext.cu:8-10, :39-42), which cuBQL's radixBuilder builds over the Gaussian
 * boxes (CR/rasterizer_impl.cu:78-83).  Here the tree is a Karras (2012) radix tree
 * over leaves of params->leaf_size consecutive Gaussians in Morton order.
 *   order [P] u32: the caller's index of the Gaussian at each BVH position (stable
 *         sort of the keys, equal keys in caller order); leaf j holds positions
 *         [j L, min((j + 1) L, P));
 *   keys  [P] u32: the sorted keys, the top 24 bits of morton3d_unit of the means;
 *   nodes [NL - 1][4] i32 (NL = ceil(P / L) leaves; none when NL < 2): left child,
 *         right child (>= 0 internal node, < 0 leaf ~j), first and last leaf under
 *         the node; node 0 is the root.  Ties between equal leaf keys are broken by
 *         leaf index;
 *   boxes [NL - 1][12] f32: the left child's box lo.xyz hi.xyz, then the right
 *         child's: the union (fminf/fmaxf, NaN boxes skipped) of the boxes of the
 *         Gaussians under it.
 * Any output may be NULL.  Stream-ordered, no host synchronisation. */
int gvc_forward_tree(const void *state, int P, int S, const gvc_params *params, uint32_t *order, uint32_t *keys,
                     int32_t *nodes, float *boxes, void *stream);

/* Debug counts of a forward (state from a forward with GVC_FLAG_CANDIDATE_LOG,
 * with the same sample set): sample_counts [S] u32 = accepted pairs per sample
 * (original sample order), gaussian_counts [P] u32 = accepted pairs per Gaussian
 * (caller order).  The reference's count_intersections (CR/forward.cu:180, :257),
 * which its debug mode reduces to max/mean (CR/rasterizer_impl.cu:175-216).  Replays
 * the candidate log (or re-walks lane-walked groups); stream-ordered;
 * GVC_ERR_UNSUPPORTED without the candidate log. */
int gvc_forward_pair_counts(const void *state, int P, int S, const gvc_params *params, const void *sample_set,
                            uint32_t *sample_counts, uint32_t *gaussian_counts, void *stream);

/* Backward: gradients for means3D [P,3], scales [P,3], rotations [P,4], values
 * [P,1], weights [P,1] from dL/dout_cells, dL/dout_cell_weights [S].  Replaces
 * _C.rasterize_gaussians_backward, DGR/ext.cu:107-147, DGR/rasterize_points.cu:75-137,
 * CR/rasterizer_impl.cu:243-452 (kernels K9-K11, CR/backward.cu:9-660); the
 * covariance chain is derived from the forward formula instead of following
 * CR/backward.cu:68-124.  Reduction is deterministic: per-(Gaussian, group)
 * partials are sorted by Gaussian (stable) and each Gaussian's are summed in FP64
 * in an order fixed by its own sorted partials; no float atomics.
 * record_capacity >= n_records of the forward (gvc_forward_stats) or the call
 * flags GVC_ERRFLAG_RECORD_OVERFLOW and the gradients are incomplete.
 * out_cells/out_weights are the outputs of the forward that filled `state`: the
 * walk reads the same values from the copy that forward kept in the state, in
 * the sample set's order. */
size_t gvc_backward_bytes(int P, uint64_t record_capacity);
int gvc_mixture_backward(int P, const float *means3D, const float *scales, const float *rotations,
                         const float *values, const float *weights, float scale_modifier,
                         const void *sample_set, int S, const gvc_params *params,
                         const float *out_cells, const float *out_weights, const float *dL_dout_cells,
                         const float *dL_dout_cell_weights, void *state, size_t state_bytes,
                         uint64_t record_capacity, void *workspace, size_t workspace_bytes,
                         float *dL_dmeans3D, float *dL_dscales, float *dL_drotations,
                         float *dL_dvalues, float *dL_dweights, void *stream);

/* ---- mesh sampler ------------------------------------------------------------
 * Tetrahedral meshes: exact point location in closed tets (exact orient3d signs; a
 * point on a face, edge or vertex is inside, on the mesh boundary too, as Probe's
 * closed cells hold it; where several tets hold a point the lowest index wins) and
 * barycentric interpolation of a per-point field.  Replaces sample_meshu over
 * Viskores Probe, GS/sample_mesh.cu:242-414 (cells of VTK type 10), and returns the
 * int32 containing-cell id (-1 outside) besides the value (-1.0 outside, :382).
 * points [n_points,3] f32, tets [n_tets,4] i32 (device).  The locator buffer
 * (gvc_tet_locator_bytes) holds a uniform grid of candidate lists with room for
 * entry_capacity (tet, grid cell) entry slots; each cell's list is padded to a
 * multiple of 16 slots (one 128-byte line each).
 * gvc_tet_locator_build is BLOCKING (two small D2H reads); GVC_ERR_CAPACITY with
 * *entries_needed set means: rebuild with a capacity >= *entries_needed. */
size_t gvc_tet_locator_bytes(int n_points, int n_tets, uint64_t entry_capacity);
int gvc_tet_locator_build(int n_points, const float *points, int n_tets, const int32_t *tets,
                          uint64_t entry_capacity, void *locator, size_t locator_bytes,
                          uint64_t *entries_needed, void *stream);
/* cell_ids [S] i32 (required); values_out [S] f64 (optional, needs point_values
 * [n_points] f64); bary_out [S,4] f32 (optional). */
int gvc_tet_locate_interp(int n_points, const float *points, int n_tets, const void *locator,
                          size_t locator_bytes, uint64_t entry_capacity, const double *point_values,
                          int S, const float *samples, int32_t *cell_ids, double *values_out,
                          float *bary_out, void *stream);

/* Same, for the points of a sample set (gvc_sample_set_build): they are processed
 * in the set's Morton order (cache-friendly) and results land at the original
 * sample index.  Lets a training step sort its samples once for both the
 * sampler and the mixture. */
int gvc_tet_locate_interp_set(int n_points, const float *points, int n_tets, const void *locator,
                              size_t locator_bytes, uint64_t entry_capacity, const double *point_values,
                              const void *sample_set, int S, int32_t *cell_ids, double *values_out,
                              float *bary_out, void *stream);

/* Bound values: gvc_tet_locator_bind_values stores, per tet, the affine
 * interpolant of point_values [n_points] f64 (value(p) = w3 + g . (p - v3) in
 * double; for tets whose barycentric matrix is ill-conditioned, cond_F > 1e6,
 * the four vertex values) in the locator; stream-ordered, copies nothing back.
 * gvc_tet_locate_bound then locates like gvc_tet_locate_interp (samples [S,3] in
 * caller order) or gvc_tet_locate_interp_set (sample_set != NULL, samples
 * ignored) and evaluates the bound interpolant: one 32-byte read per located
 * sample instead of four vertices and four values.  Values agree with
 * gvc_tet_locate_interp's to the rounding of double (1e-9 relative covers tets
 * up to cond_F = 1e6).  Before the first bind, values_out is NaN where a tet
 * holds the sample.  Rebind after the values change.  No counterpart in the
 * reference, whose sampler takes the values with every call
 * (GS/sample_mesh.cu:380-390); for callers that sample one field many times. */
int gvc_tet_locator_bind_values(int n_points, const float *points, int n_tets, void *locator,
                                size_t locator_bytes, uint64_t entry_capacity, const double *point_values,
                                void *stream);
int gvc_tet_locate_bound(int n_points, const float *points, int n_tets, const void *locator,
                         size_t locator_bytes, uint64_t entry_capacity, int S, const float *samples,
                         const void *sample_set, int32_t *cell_ids, double *values_out, float *bary_out,
                         void *stream);

/* Mixed meshes (the rest of sample_meshu's cell table, GS/sample_mesh.cu:300-336,
 * interpolated like Viskores Probe): for each sample whose cell (cell_ids[i], e.g.
 * the parent of the tetrahedron gvc_tet_locate_interp found in the cells'
 * tetrahedral splits) is a VTK hexahedron (12), wedge (13) or pyramid (14),
 * values_out[i] = sum_k N_k(r,s,t) point_values[conn[cell_offsets[c] + k]] with
 * the VTK shape functions at the sample's parametric coordinates (Newton's
 * method in double).  Other samples are left untouched.  cell_types uint8 [C],
 * cell_offsets int32 [C], connectivity int32. */
int gvc_cell_interp(int S, const float *samples, const int32_t *cell_ids, const uint8_t *cell_types,
                    const int32_t *cell_offsets, const int32_t *connectivity, const float *points,
                    const double *point_values, double *values_out, void *stream);

/* Uniform point grid (dims nx,ny,nz; origin; spacing -- HOST arrays), trilinear
 * interpolation, -1 outside cast to the value type (all-ones bits for the
 * unsigned types).  value_dtype: 0 f64, 1 f32, 2 u16, 3 u8; out has that type.
 * Replaces sample_mesh over Viskores Probe, GS/sample_mesh.cu:26-240. */
int gvc_grid_sample(const int64_t *dims, const float *origin, const float *spacing, int value_dtype,
                    const void *values, int S, const float *samples, void *out, void *stream);

/* ---- 3-NN initialiser ----------------------------------------------------------
 * distCUDA2 of simple-knn (call sites R/scene/gaussian_model.py:175-178, :730-733):
 * per point the mean squared distance to its 3 nearest other points.  Exact.
 * BLOCKING (one small D2H read of the bounding box). */
size_t gvc_knn3_bytes(int n);
int gvc_knn3(int n, const float *points, float *mean_sq_dist, void *workspace, size_t workspace_bytes,
             void *stream);

/* ---- batches of independent volumes ---------------------------------------------
 * V volumes of uneven size fitted in one persistent launch per pass: each volume
 * has its own Gaussians, samples, Morton order and BVH; groups never mix volumes.
 * Inputs are concatenated volume-major (volume 0's Gaussians, then volume 1's,
 * ...; the same for samples).  No counterpart in the reference, which fits one
 * volume per process (R/train.py); with V = 1 results equal the single-volume
 * calls bit for bit.  gvc_batch_describe copies the offset table to the device
 * (BLOCKING; call once per batch shape, outside captured regions). */
typedef struct gvc_batch {
    int volumes;
    int leaf_size;       /* must equal params->leaf_size of the calls below */
    int P_total, S_total;
    int groups;          /* sum of ceil(S_v / 64)                           */
    int leaves;          /* sum of ceil(P_v / leaf_size)                    */
    int internal_nodes;  /* sum of max(leaves_v - 1, 0)                     */
    const void *table;   /* device offset table                             */
} gvc_batch;

size_t gvc_batch_table_bytes(int volumes);
int gvc_batch_describe(int volumes, const int *gaussians_per_volume, const int *samples_per_volume, int leaf_size,
                       void *table, size_t table_bytes, gvc_batch *out, void *stream);
size_t gvc_sample_set_bytes_batched(const gvc_batch *batch);
int gvc_sample_set_build_batched(const gvc_batch *batch, const float *samples, void *sample_set, size_t bytes,
                                 void *stream);
size_t gvc_forward_state_bytes_batched(const gvc_batch *batch, const gvc_params *params);
int gvc_mixture_forward_batched(const gvc_batch *batch, const float *means3D, const float *scales,
                                const float *rotations, const float *values, const float *weights,
                                float scale_modifier, const void *sample_set, const gvc_params *params,
                                float *out_cells, float *out_weights, float *conics, void *state,
                                size_t state_bytes, void *stream);
int gvc_forward_stats_batched(const gvc_batch *batch, const void *state, const gvc_params *params,
                              uint64_t *n_records, uint64_t *n_pairs, uint32_t *error_flags,
                              uint64_t *chunk_slots_used, uint64_t *lane_groups, void *stream);
int gvc_forward_accumulate_batched(const gvc_batch *batch, const void *state, const gvc_params *params,
                                   uint64_t *accum, void *stream);
/* gvc_forward_pair_counts over a batch: counts at the concatenated (volume-major)
 * sample and Gaussian positions. */
int gvc_forward_pair_counts_batched(const gvc_batch *batch, const void *state, const gvc_params *params,
                                    const void *sample_set, uint32_t *sample_counts, uint32_t *gaussian_counts,
                                    void *stream);
int gvc_mixture_backward_batched(const gvc_batch *batch, const float *scales, const float *rotations,
                                 float scale_modifier, const void *sample_set, const gvc_params *params,
                                 const float *out_cells, const float *out_weights, const float *dL_dout_cells,
                                 const float *dL_dout_cell_weights, void *state, size_t state_bytes,
                                 uint64_t record_capacity, void *workspace, size_t workspace_bytes,
                                 float *dL_dmeans3D, float *dL_dscales, float *dL_drotations, float *dL_dvalues,
                                 float *dL_dweights, void *stream);

/* ---- training-step glue ----------------------------------------------------------
 * Activations, R/scene/gaussian_model.py:31-48, :71-77, :114-132 (get_scaling with
 * _apply_cap, get_rotation, get_weight, get_values): scales [P,3] = s * r_soft / r
 * with s = exp(log_scales), r = |s| + 1e-8, r_soft = max_scale * tanh(r / max_scale)
 * (max_scale <= 0 disables the cap; the reference uses 0.02); rotations [P,4] =
 * q / max(|q|, 1e-12); weights, values [P] = sigmoid of the logits.  The backward
 * maps output gradients to gradients of the raw parameters. */
int gvc_activate_forward(int P, const float *log_scales, const float *rotations, const float *weight_logits,
                         const float *value_logits, float max_scale, float *scales, float *rotations_n,
                         float *weights, float *values, void *stream);
int gvc_activate_backward(int P, const float *log_scales, const float *rotations, const float *weight_logits,
                          const float *value_logits, float max_scale, const float *dL_dscales,
                          const float *dL_drotations_n, const float *dL_dweights, const float *dL_dvalues,
                          float *dL_dlog_scales, float *dL_drotations, float *dL_dweight_logits,
                          float *dL_dvalue_logits, void *stream);

/* Fitting loss of R/train.py:389-420 (l1_loss: R/utils/loss_utils.py:19-20):
 * L1 over (gt != -1) & (cells != -1) as a mean (over every sample when
 * encode_surface, R/train.py:391-392), plus fn_reg x the mean positive
 * clamp(0.011 - w) over (gt != -1) & (0 < w < t) with t = 0.011 (0.0105 when
 * encode_surface), plus, when encode_surface, fp_reg x the mean positive w over
 * (gt == -1) & (w > 0).  loss: one float on the device.  Sums in FP64 in a fixed
 * order (deterministic); `state` (gvc_loss_state_bytes) keeps them for the
 * backward, which writes dL/dcells and dL/dweights scaled by *dL_dloss (device
 * float, NULL = 1). */
size_t gvc_loss_state_bytes(int S);
int gvc_loss_forward(int S, const float *cells, const float *weights, const float *gt, float fn_reg, float fp_reg,
                     int encode_surface, float *loss, void *state, size_t state_bytes, void *stream);
int gvc_loss_backward(int S, const float *cells, const float *weights, const float *gt, float fn_reg,
                      float fp_reg, int encode_surface, const float *dL_dloss, const void *state,
                      size_t state_bytes, float *dL_dcells, float *dL_dweights, void *stream);

/* ---- optimiser and model surgery ---------------------------------------------------
 * The training loop's optimiser and densify/prune (R/scene/gaussian_model.py:207-256,
 * :396-612, :727-799; R/train.py:367, :551-597), over parameter buffers of a fixed
 * capacity: nothing is reallocated, rows [0, count) are the live Gaussians. */
#define GVC_ADAM_MAX_TENSORS 8

/* Device clock of a training loop (caller-owned device memory, zero-initialised,
 * lr[] set by the caller).  `iteration` is advanced by gvc_lr_advance, `step` (the
 * Adam step count) by gvc_adam_step; both on the device, so a captured step keeps
 * counting on replay. */
typedef struct gvc_train_clock {
    int64_t iteration;
    int64_t step;
    double lr[GVC_ADAM_MAX_TENSORS];           /* learning rate per lr_index */
    float step_size[GVC_ADAM_MAX_TENSORS];     /* written by gvc_adam_step */
    float bc2_sqrt;                            /* written by gvc_adam_step */
    float reserved[7];
} gvc_train_clock;

/* get_expon_lr_func (R/utils/general_utils.py:35-69). */
typedef struct gvc_lr_schedule {
    double lr_init, lr_final, delay_mult;
    int64_t delay_steps, max_steps;
} gvc_lr_schedule;

/* update_learning_rate (R/scene/gaussian_model.py:249-255): clock.iteration =
 * set_iteration (>= 0) or clock.iteration + 1 (set_iteration < 0), then
 * clock.lr[lr_index] = schedule(iteration).  Stream-ordered, no host sync. */
int gvc_lr_advance(gvc_train_clock *clock, const gvc_lr_schedule *schedule, int lr_index, int64_t set_iteration,
                   void *stream);

/* One tensor of an Adam step: numel floats of parameter, gradient and both moments.
 * grad == NULL skips the tensor (torch skips a parameter without a gradient). */
typedef struct gvc_adam_tensor {
    float *param;
    const float *grad;
    float *exp_avg;
    float *exp_avg_sq;
    uint64_t numel;
    int lr_index;
} gvc_adam_tensor;

/* optimizer.step() of torch.optim.Adam(betas, eps) as the reference configures it
 * (R/scene/gaussian_model.py:240, stepped at R/train.py:593-595), for up to
 * GVC_ADAM_MAX_TENSORS tensors in one launch: clock.step += 1, then per element
 * m = lerp(m, g, 1 - beta1), v = v*beta2 + (1 - beta2)*g*g,
 * p += -lr/(1 - beta1^t) * m / (sqrt(v)/sqrt(1 - beta2^t) + eps).  No step is
 * counted when every grad is NULL.  beta1, beta2, eps are doubles, as torch holds
 * them: 1 - beta and the bias corrections are formed in double, then rounded to
 * float.  Stream-ordered, no host sync. */
int gvc_adam_step(int n_tensors, const gvc_adam_tensor *tensors, gvc_train_clock *clock, double beta1,
                  double beta2, double eps, void *stream);

/* The densification pick of R/train.py:562-578 (torch.topk of |cells - gt| over
 * gt != -1, k from the Gaussian count, cap_max and the low-weight count):
 * order[0..S) = the indices of the valid samples by descending error, ties in
 * ascending index, then the invalid ones; counts (device uint32 [2]) = the number
 * of valid samples and the number of the `count` Gaussians whose weight
 * sigmoid(weight_logits[i]) <= low_weight (the reference uses 0.005).
 * Stream-ordered. */
size_t gvc_error_order_bytes(int S);
int gvc_error_order(int S, const float *cells, const float *gt, int count, const float *weight_logits,
                    float low_weight, uint32_t *order, uint32_t *counts, void *workspace, size_t workspace_bytes,
                    void *stream);

/* The trainable model: the five raw parameter arrays of capacity rows (means
 * [C,3], weight logits [C], log-scales [C,3], quaternions [C,4], value logits [C]:
 * the reference optimiser's groups xyz, weight, scaling, rotation, value, in that
 * order) and their Adam moments in the same order and shapes (NULL: no moments yet,
 * as before the first step). */
typedef struct gvc_store {
    int capacity;
    float *xyz, *weight, *scaling, *rotation, *value;
    float *exp_avg[5];
    float *exp_avg_sq[5];
} gvc_store;

/* densify_in_empty (R/scene/gaussian_model.py:727-763) and the initial cloud of
 * create_from_pcd (:139-205): rows [count, count + k) take positions
 * samples[idx[j]] (idx NULL: samples[j]), log-scales log(sqrt(max(d * dist_scale,
 * 1e-7))) on all three axes with d = the mean squared distance to the 3 nearest
 * of the count + k positions (distCUDA2), rotation (1, 0, 0, 0), weight logit
 * log(w0 / (1 - w0)) for w0 = init_weight, value logit log(g / (1 - g)) for
 * g = gt[idx[j]], and zero moments.  The reference uses dist_scale 0.1 and w0 0.01
 * when densifying, 1 and 0.1 for the initial cloud.  count + k <= capacity; every
 * idx[j] must index `samples` and `gt` (not checked on the device).
 * BLOCKING (the 3-NN reads back a bounding box). */
size_t gvc_densify_bytes(int total);
int gvc_densify_in_empty(int count, int k, const gvc_store *store, const float *samples, const float *gt,
                         const uint32_t *idx, float dist_scale, float init_weight, void *workspace,
                         size_t workspace_bytes, void *stream);

/* prune_points / the prune of densify_and_prune (R/scene/gaussian_model.py:524-554,
 * :786): remove the rows with prune_mask[i] != 0, or (prune_mask NULL) with
 * sigmoid(weight logit) < min_weight (a NaN weight is kept), from every parameter
 * and moment; the survivors keep their order.  *new_count (device int) = the rows
 * kept.  Stream-ordered (reading new_count is the caller's sync). */
size_t gvc_prune_bytes(int count);
int gvc_prune(int count, const gvc_store *store, const uint8_t *prune_mask, float min_weight, int *new_count,
              void *workspace, size_t workspace_bytes, void *stream);

/* ---- evaluation ----------------------------------------------------------------
 * The sums behind the volume evaluation of R/volume_metrics.py:318-369 (L1, L2, the
 * five PSNR variants, false negatives and positives), in one pass over S (cells,
 * gt) pairs, FP64, deterministic: out (device double [GVC_METRIC_SUMS]) =
 *   [0] sum |c-g|  [1] sum (c-g)^2  [2] sum (c-g)^2 over g != -1 & c != -1  [3] its count
 *   [4] sum g^2  [5] sum (c0-g)^2 with c0 = (c == -1 ? 0 : c)  [6] that over g != -1
 *   [7] count g != -1  [8] count g == -1  [9] count c == -1 & g != -1  [10] count c != -1 & g == -1
 * Stream-ordered. */
#define GVC_METRIC_SUMS 11
size_t gvc_volume_metrics_bytes(int S);
int gvc_volume_metrics(int S, const float *cells, const float *gt, double *out, void *workspace,
                       size_t workspace_bytes, void *stream);

/* gvc_volume_metrics without a workspace and without allocating: no counterpart
 * in the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight evaluation loops, harnesses that forbid hipMalloc).
 * Its block partials live in a static device array of the library (5.8 MB per
 * GPU, kept until the process ends), so S is limited to 2^24 samples
 * (GVC_ERR_CAPACITY beyond).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (different streams, host threads) race and must be ordered
 * by the caller.  Same results as gvc_volume_metrics, bit for bit. */
int gvc_volume_metrics_noalloc(int S, const float *cells, const float *gt, double *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GVC_H_ */
