/*
 * splinegrids_hip.h -- C-ABI of the MI355X (gfx950) grid-evaluation library.
 *
 * This is the drop-in boundary for the hot path of SplineGrids.jl.  The
 * reference reaches its kernels through KernelAbstractions' backend dispatch
 * (backend = get_backend(eval); kernel(backend)(...; ndrange); synchronize):
 * src/spline_grid.jl:216-228, src/adjoint.jl:70-81,
 * src/spline_dimension.jl:236-240, src/utils.jl:20-28.  A Julia package
 * extension replaces those launches with ccall-s into the functions below
 * (the Julia binding is given with the integration notes).
 *
 * Conventions
 *   - Plain C types only; every function returns SG_OK (0) or an SG_ERR_*
 *     code and never aborts.  sg_last_error() gives the calling thread's
 *     message for the last failure.
 *   - Device pointers refer to contiguous COLUMN-MAJOR arrays, exactly as
 *     Julia stores them: control_points[n_1,...,n_Nin, Nout],
 *     weights[n_1,...,n_Nin], eval[m_1,...,m_Nin, Nout].  The output channel
 *     is the slowest axis.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Calls are stream-ordered; the reference synchronizes after every launch,
 *     which the caller reproduces with sg_stream_synchronize.
 *   - A plan is not re-entrant: concurrent calls on one plan must be
 *     serialized by the caller (they share the plan's adjoint workspace).
 */
#ifndef SPLINEGRIDS_HIP_H
#define SPLINEGRIDS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SG_OK 0
#define SG_ERR_INVALID 1     /* bad argument or descriptor */
#define SG_ERR_HIP 2         /* a HIP runtime call failed */
#define SG_ERR_UNSUPPORTED 3 /* valid input outside what this build handles */
#define SG_ERR_NO_MEMORY 4   /* device allocation failed (plan set-up only) */

/* value types of the caller's arrays (Tv in the reference) */
#define SG_F32 0
#define SG_F64 1

/* arithmetic tiers for the tensor-product contraction.  Basis tables are
 * always computed in FP64. */
#define SG_TIER_NATIVE 0    /* FP32 arrays: FP32 FMA; FP64 arrays: FP64 FMA */
#define SG_TIER_SPLIT_F16 1 /* FP32 arrays: hi+lo FP16 split on MFMA, FP32 acc */
#define SG_TIER_BF16 2      /* FP32 arrays: BF16 operands on MFMA, FP32 acc */

#define SG_MAX_NIN 8
#define SG_MAX_FIELDS 4

typedef struct sg_plan sg_plan;

/* One input dimension.  Mirrors SplineDimension (src/spline_dimension.jl:
 * 25-66): degree, max_derivative_order, knot_vector.knots_all and
 * sample_points.  Host pointers, FP64 values (Float32 values promote
 * exactly).  n_basis must equal n_knots - degree - 1 (src/utils.jl:31-33). */
typedef struct {
  int32_t degree;
  int32_t n_basis;
  int32_t n_samples;
  int32_t max_derivative_order;
  const double* knots_all;     /* n_basis + degree + 1 values, sorted */
  const double* sample_points; /* n_samples values */
} sg_dim_desc;

/* A grid: SplineGrid / NURBSGrid (src/spline_grid.jl:19-58, :100-108). */
typedef struct {
  int32_t n_in;     /* 1 .. SG_MAX_NIN */
  int32_t n_out;    /* >= 1 */
  int32_t is_nurbs; /* weights are passed to sg_evaluate iff nonzero */
  int32_t device;   /* HIP device ordinal the plan lives on */
  const sg_dim_desc* dims;
} sg_grid_desc;

/* Thread-local message of the last failed call on this thread. */
const char* sg_last_error(void);

int sg_version(int32_t* major, int32_t* minor, int32_t* patch);

/* set_sample_indices! + evaluate!(::SplineDimension) on caller arrays.
 * Replaces src/utils.jl:19-29 (set_sample_indices_kernel,
 * src/util_kernels.jl:22-49) and src/spline_dimension.jl:231-242
 * (spline_dimension_kernel, :160-215).  Device pointers:
 *   knots_all[n_knots], sample_points[n_samples]    (dtype)
 *   sample_indices[n_samples]                       (int32, 1-based, out)
 *   eval[n_samples, degree+1, max_derivative_order+1] (dtype, out)
 * Arithmetic is FP64 whatever dtype is; results are rounded to dtype. */
int sg_dimension_evaluate(const void* knots_all, int32_t n_knots,
                          const void* sample_points, int32_t n_samples,
                          int32_t degree, int32_t max_derivative_order,
                          int32_t* sample_indices, void* eval, int32_t dtype,
                          void* stream);

/* Build a plan: uploads knots and samples, computes every dimension's FP64
 * span table and basis table (with all derivative orders up to D_d) on the
 * GPU, derives the FP32 / split-FP16 / BF16 operand tables, picks the kernels
 * and tile shapes, and allocates the adjoint workspace.  All allocation
 * happens here; evaluate calls allocate nothing. */
int sg_plan_create(const sg_grid_desc* desc, void* stream, sg_plan** plan);
int sg_plan_destroy(sg_plan* plan);

/* Copy dimension `dim`'s FP64 basis table ([m, p+1, D+1], column-major) and
 * its 1-based sample indices back to host memory (either may be NULL). */
int sg_plan_basis(const sg_plan* plan, int32_t dim, double* table,
                  int32_t* sample_indices);

/* Bytes of device memory the plan holds (tables + workspace). */
int sg_plan_device_bytes(const sg_plan* plan, int64_t* bytes);

/* One line of text naming the kernels and tile shapes the plan chose (for
 * logs and tests; no counterpart in the reference), truncated to capacity-1
 * characters and NUL-terminated. */
int sg_plan_describe(const sg_plan* plan, char* text, int32_t capacity);

/* evaluate!(grid; derivative_order, control_points, eval): replaces the
 * spline_eval_kernel launch of src/spline_grid.jl:200-230 (kernel :119-183).
 * eval is overwritten.  weights must be non-NULL iff the plan is NURBS, and
 * NURBS take only zero derivative orders (src/validation.jl:81-88). */
int sg_evaluate(sg_plan* plan, const int32_t* derivative_order,
                const void* control_points, const void* weights, void* eval,
                int32_t dtype, int32_t tier, void* stream);

/* Several derivative fields of one B-spline grid in one pass over the
 * control points (n_fields <= SG_MAX_FIELDS; derivative_orders is
 * [n_fields][n_in] row-major).  Same result as n_fields sg_evaluate calls. */
int sg_evaluate_fields(sg_plan* plan, int32_t n_fields,
                       const int32_t* derivative_orders,
                       const void* control_points, void* const* evals,
                       int32_t dtype, int32_t tier, void* stream);

/* Derivative fields of a NURBS plan's rational map R = sum N w P / sum N w.
 * No counterpart in the reference, which rejects NURBS derivatives
 * (src/validation.jl:81-88; sg_evaluate keeps that behaviour): the quotient
 * rule over the homogeneous B-spline fields of (w P, w), combined in FP64.
 * derivative_orders is [n_fields][n_in] (n_fields <= SG_MAX_FIELDS, each
 * order <= the dimension's max_derivative_order, at most 16 orders below
 * the requested ones); evals[f] has the shape of eval.  workspace: device
 * memory of sg_rational_workspace_bytes bytes (homogeneous control points
 * and fields); the call allocates nothing. */
int sg_rational_workspace_bytes(const sg_plan* plan, int32_t n_fields,
                                const int32_t* derivative_orders, int32_t dtype,
                                int64_t* bytes);
int sg_evaluate_rational_fields(sg_plan* plan, int32_t n_fields,
                                const int32_t* derivative_orders,
                                const void* control_points, const void* weights,
                                void* const* evals, void* workspace,
                                int64_t workspace_bytes, int32_t dtype, void* stream);

/* The adjoint of a NURBS plan's evaluate! (derivative order 0) in its
 * control points, the weights fixed: control_points = w_k [B^T (eval /
 * w(x))]_k.  No counterpart in the reference, whose evaluate_adjoint! rejects
 * NURBS (src/adjoint.jl:57; sg_evaluate_adjoint keeps that behaviour).
 * accumulate as in sg_evaluate_adjoint.  workspace: device memory of
 * sg_rational_adjoint_workspace_bytes bytes; the call allocates nothing. */
int sg_rational_adjoint_workspace_bytes(const sg_plan* plan, int32_t dtype, int64_t* bytes);
int sg_evaluate_rational_adjoint(sg_plan* plan, const void* weights, const void* eval,
                                 void* control_points, void* workspace,
                                 int64_t workspace_bytes, int32_t dtype, int32_t accumulate,
                                 void* stream);

/* evaluate_adjoint!(grid; derivative_order, control_points, eval): replaces
 * src/adjoint.jl:52-83 (kernel :1-40).  Computes M^T eval into
 * control_points without atomics; the result is deterministic.
 * accumulate = 0 overwrites control_points (the reference zeroes it first,
 * src/adjoint.jl:61); accumulate = 1 adds to it (reverse-mode AD).
 * B-spline plans only (src/adjoint.jl:57).  The split-FP16 and BF16 tiers
 * run on the matrix-core adjoint (FP32 plans whose band windows fit 16 x 16
 * control points per tile); other plans return SG_ERR_UNSUPPORTED for them. */
int sg_evaluate_adjoint(sg_plan* plan, const int32_t* derivative_order,
                        const void* eval, void* control_points, int32_t dtype,
                        int32_t tier, int32_t accumulate, void* stream);

/* synchronize(backend) after a launch (e.g. src/spline_grid.jl:228). */
int sg_stream_synchronize(void* stream);

/* ---- batches: many independent grids in one persistent launch ----------
 * No counterpart in the reference, which evaluates one grid per evaluate!
 * call; the batch is for callers that hold many small grids (the per-grid
 * loop of evaluate! calls is what it replaces).  Every grid is evaluated
 * with derivative order 0 on FP32 arrays. */
typedef struct sg_batch sg_batch;

/* Build per-grid FP64 tables and one list of tiles over all grids, sorted
 * largest first; allocates all device memory of the batch. */
int sg_batch_create(const sg_grid_desc* grids, int32_t n_grids, void* stream,
                    sg_batch** batch);
int sg_batch_destroy(sg_batch* batch);

/* Device pointers of every grid's arrays (weights[i] NULL for B-spline
 * grids, or weights itself NULL); kept until set again. */
int sg_batch_set_arrays(sg_batch* batch, const void* const* control_points,
                        const void* const* weights, void* const* evals);

/* Enqueue the evaluation of every grid: a small kernel that resets the
 * work-queue counters, then the persistent kernel.  Only kernel launches, so
 * the call can be captured into a hipGraph and replayed.  tier: native, or
 * split-FP16 / BF16 when every grid of the batch takes the MFMA tiers
 * (SG_ERR_UNSUPPORTED otherwise); each grid then gets the result of its own
 * sg_evaluate with that tier. */
int sg_batch_evaluate(sg_batch* batch, int32_t tier, void* stream);

/* The loop that a batch replaces, for comparison: one sg_evaluate-style
 * launch per grid (native tier) on `stream`. */
int sg_batch_evaluate_each(sg_batch* batch, void* stream);

/* Time candidate tilings (workgroups per CU {2,3,4,6} x rows per wave
 * {8,16,32}; each grid's rows are halved until its tiles fit the LDS share of
 * one workgroup) on the arrays set last and keep the fastest; `reps` timed
 * launches each.  A graph captured from sg_batch_evaluate keeps the tiling of
 * its capture, so once a capture has been recorded the batch refuses to
 * retune (SG_ERR_INVALID): autotune first, then capture. */
int sg_batch_autotune(sg_batch* batch, int32_t reps, void* stream);

/* Totals of the batch: tiles, sample points, grids evaluated by the
 * persistent kernel, and the tuned rows per wave / workgroups per CU. */
int sg_batch_info(const sg_batch* batch, int64_t* n_tiles, int64_t* n_samples,
                  int32_t* n_persistent, int32_t* rows_per_wave,
                  int32_t* blocks_per_cu);

/* One line of text with the batch's tiling, LDS per workgroup and resident
 * workgroups per CU (for logs and tests), truncated to capacity-1 characters
 * and NUL-terminated. */
int sg_batch_describe(const sg_batch* batch, char* text, int32_t capacity);

#ifdef __cplusplus
}
#endif
#endif
