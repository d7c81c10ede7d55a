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
#define SG_U8 2 /* flags (0/1) of the THB bookkeeping, where a call says so */

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

/* sg_dimension_evaluate without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight loops that move sample points every iteration, harnesses
 * that forbid hipMalloc).  Same arguments, kernels and results; the scratch
 * is a fixed device array of 2^20 doubles (8 MiB) held by every GPU the
 * library runs on until the process ends, so n_samples * (degree + 1) *
 * (max_derivative_order + 1) must be at most 2^20 (FP64) or 2^19 (FP32), or
 * the call fails with SG_ERR_INVALID before launching anything.  No
 * synchronisation: stream-ordered.  Every call on one GPU shares the array:
 * calls that overlap on a GPU (other streams or threads) race and must be
 * serialised by the caller. */
int sg_dimension_evaluate_noalloc(const void* knots_all, int32_t n_knots,
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

/* ---- refinement matrices and locally refined (THB) control points ------
 *
 * A refinement matrix as the reference's RefinementMatrix stores it
 * (src/refinement_matrix.jl:1-60): m x n, consecutive non-zeros in every row
 * and column; 1-based device arrays row_pointer[m] and column_start[m] (int32)
 * and the non-zeros nzval[nnz] (value type given by the call's dtype); row i
 * holds nzval[row_pointer[i] .. row_pointer[i+1]-1] (row_pointer[m+1] =
 * nnz + 1) at columns column_start[i], column_start[i] + 1, ...  The struct
 * itself is host memory. */
typedef struct {
  int32_t m, n, nnz;
  const int32_t* row_pointer;
  const int32_t* column_start;
  const void* nzval;
} sg_refmat;

/* The (n_basis + 1) x n_basis refinement matrix of inserting knot_new into
 * knots_all (device, n_basis + degree + 1 values of dtype), knot_span_index
 * 1-based with knots_all[k] <= knot_new < knots_all[k + 1]: replaces
 * build_refinement_matrix_kernel and RefinementMatrix(spline_dimension,
 * knot_span_index, knot_new) (src/refinement.jl:3-84).  Outputs (device):
 * row_pointer, column_start [n_basis + 1] and nzval [n_basis + degree + 1].
 * alpha is computed in FP64 and rounded to dtype. */
int sg_refinement_insert_knot(const void* knots_all, int32_t n_basis, int32_t degree,
                              int32_t knot_span_index, double knot_new, int32_t* row_pointer,
                              int32_t* column_start, void* nzval, int32_t dtype, void* stream);

/* C = A B (Base.:*, src/refinement_matrix.jl:273-327) in two calls: the
 * non-zeros per row of C and their first column (device [A.m] each); the
 * caller forms row_pointer (1 + exclusive prefix sum) and allocates nzval,
 * then the values (C->nzval is written).  A and B must be valid
 * (sg_refinement_validate). */
int sg_refinement_product_rows(const sg_refmat* A, const sg_refmat* B, int32_t* n_nonzero,
                               int32_t* column_start, void* stream);
int sg_refinement_product_values(const sg_refmat* A, const sg_refmat* B, const sg_refmat* C,
                                 int32_t dtype, void* stream);

/* validate_refinement_matrix_kernel (src/refinement_matrix.jl:134-181): one
 * flag per row (device uint8 [m]), 1 where the row is valid. */
int sg_refinement_validate(const sg_refmat* A, uint8_t* valid_rows, void* stream);

/* mult!(Y, As, B, dims_refinement) (src/refinement_matrix.jl:421-445): Y and B
 * are column-major device arrays of nd dimensions (the last the output
 * channel, never refined); dims[k] (1-based) is the dimension mats[k]
 * refines (shape_y = mats[k].m, shape_b = mats[k].n there; other dimensions
 * equal).  Y[I] = sum_J prod_d R_d[I_d, J_d] B[J], overwritten.  FP32/FP64. */
int sg_refinement_mult(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                       int32_t n_mats, const sg_refmat* mats, const int32_t* dims, const void* B,
                       void* Y, int32_t dtype, void* stream);

/* mult_adjoint!(B, As, Y, dims_refinement) (src/adjoint.jl:127-152) without
 * atomics: B = M^T Y (overwritten, or added to with accumulate = 1), each
 * element of B gathering the rows that hold it in a fixed order.  dtype
 * SG_U8 is the Flag form of deactivate_overwritten_control_points!
 * (src/control_points.jl:624-680): B[J] = 1 where some I with Y[I] != 0
 * holds J (structure only; the nzval are not read). */
int sg_refinement_mult_adjoint(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                               int32_t n_mats, const sg_refmat* mats, const int32_t* dims,
                               const void* Y, void* B, int32_t dtype, int32_t accumulate,
                               void* stream);

/* local_refinement_kernel (src/control_points.jl:296-311): for i < n_idx,
 * control_points[indices[i, :]..., c] = values[i, c].  control_points has
 * shape (shape[0..nin-1]..., nout); indices [n_idx, nin] int32 1-based and
 * values [n_idx, nout], all column-major device arrays; F32, F64 or U8.
 * A row of indices outside the grid is skipped. */
int sg_local_refinement_overwrite(int32_t nin, int32_t nout, const int32_t* shape, int64_t n_idx,
                                  const int32_t* indices, const void* values,
                                  void* control_points, int32_t dtype, void* stream);

/* values[i, c] = control_points[indices[i, :]..., c]; with clear = 1 that
 * entry is then zeroed (local_refinement_adjoint_kernel, src/adjoint.jl:
 * 154-170), with clear = 0 it is kept (the read of refinement_values_new_
 * kernel, src/control_points.jl:427-456).  A row of indices outside the grid
 * reads 0. */
int sg_local_refinement_gather(int32_t nin, int32_t nout, const int32_t* shape, int64_t n_idx,
                               const int32_t* indices, void* values, void* control_points,
                               int32_t clear, int32_t dtype, void* stream);

#ifdef __cplusplus
}
#endif
#endif
