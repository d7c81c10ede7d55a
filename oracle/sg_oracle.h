/*
 * sg_oracle.h -- CPU restatement of the SplineGrids.jl grid-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (and the CPU baseline that bench.py times).  Nothing in the
 * product (splinegrids.jl_amd/) links, loads or calls it.
 *
 * Conventions follow the reference exactly:
 *   - arrays are column-major (first index fastest), the output channel is the
 *     slowest axis of control points and eval;
 *   - sample_indices are 1-based spans, as produced by
 *     src/util_kernels.jl:22-49;
 *   - basis tables are laid out [m, p+1, D+1] (src/spline_dimension.jl:135-141).
 *
 * Every function is restated from the reference file:line given beside it.
 * Pinning (what checks each function): see the header of oracle/oracle.py.
 */
#ifndef SG_ORACLE_H
#define SG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGO_MAX_NIN 8

/* src/knot_vector.jl:28-37 + src/util_kernels.jl:1-20: repeat values[i]
 * mults[i] times.  Returns the expanded length. */
int32_t sgo_expand_knots(const double* values, const int32_t* mults, int32_t k,
                         double* knots_all);

/* src/util_kernels.jl:22-49: i = #{knots_all[k] <= t} (linear scan, stops at
 * the first knot > t), clamped to [p+1, n_knots-p-1]; 1-based. */
void sgo_sample_indices(const double* t, int32_t m, const double* knots_all,
                        int32_t n_knots, int32_t p, int32_t* idx1);

/* src/spline_dimension.jl:160-215: Cox-de Boor triangle plus derivative
 * recursion, same operation order, in double.  eval is [m, p+1, D+1]. */
void sgo_basis(const double* knots_all, const double* t, const int32_t* idx1,
               int32_t m, int32_t p, int32_t max_deriv, double* eval);

/* src/spline_grid.jl:119-183: direct per-sample sum over (p+1)^Nin control
 * points, NURBS weights fused, eval overwritten.  tables[d] is the [m_d, p_d+1,
 * D_d+1] table of dimension d, idx1[d] its 1-based spans.  W may be NULL.
 * Only the slab j_last in [last_begin, last_end) of the slowest sample
 * dimension is evaluated and written (eval is then sized for that slab);
 * pass 0, m[nin-1] for the whole grid.  nthreads > 1 uses OpenMP. */
void sgo_eval_direct_f64(int32_t nin, int32_t nout, const int32_t* n,
                         const int32_t* m, const int32_t* p, const int32_t* D,
                         const double* const* tables,
                         const int32_t* const* idx1, const int32_t* deriv,
                         const double* P, const double* W, double* eval,
                         int32_t last_begin, int32_t last_end,
                         int32_t nthreads);

/* Same algorithm with Tv = Float32 (the reference's default value type):
 * float tables, control points, weights and arithmetic. */
void sgo_eval_direct_f32(int32_t nin, int32_t nout, const int32_t* n,
                         const int32_t* m, const int32_t* p, const int32_t* D,
                         const float* const* tables,
                         const int32_t* const* idx1, const int32_t* deriv,
                         const float* P, const float* W, float* eval,
                         int32_t last_begin, int32_t last_end,
                         int32_t nthreads);

/* The direct sum of src/spline_grid.jl:148-182 at a list of samples given by
 * their column-major linear sample index; out is [npts, nout] column-major. */
void sgo_eval_points_f64(int32_t nin, int32_t nout, const int32_t* n,
                         const int32_t* m, const int32_t* p, const int32_t* D,
                         const double* const* tables,
                         const int32_t* const* idx1, const int32_t* deriv,
                         const double* P, const double* W,
                         const int64_t* lin, int64_t npts, double* out,
                         int32_t nthreads);

/* src/adjoint.jl:1-40 and :52-83: Pbar is zeroed (:61), then every sample
 * scatters prod(B) * ebar into the (p+1)^Nin control points it touches.
 * Sequential, so the summation order is fixed. */
void sgo_adjoint_direct_f64(int32_t nin, int32_t nout, const int32_t* n,
                            const int32_t* m, const int32_t* p,
                            const int32_t* D, const double* const* tables,
                            const int32_t* const* idx1, const int32_t* deriv,
                            const double* ebar, double* Pbar);

/* The same scatter in Float32 with OpenMP atomics, as the reference's
 * threaded CPU backend runs it (Atomix.@atomic, src/adjoint.jl:35-37). */
void sgo_adjoint_direct_f32(int32_t nin, int32_t nout, const int32_t* n,
                            const int32_t* m, const int32_t* p,
                            const int32_t* D, const float* const* tables,
                            const int32_t* const* idx1, const int32_t* deriv,
                            const float* ebar, float* Pbar, int32_t nthreads);

/* ---- refinement matrices and THB control points ----
 *
 * A refinement matrix (src/refinement_matrix.jl:1-60) as the reference
 * stores it, 1-based: row i (1..m) holds nzval[row_pointer[i] ..
 * row_pointer[i+1] - 1] (row_pointer[m+1] = nnz + 1) at columns
 * column_start[i], column_start[i] + 1, ... */
typedef struct {
  int32_t m, n, nnz;
  const int32_t* row_pointer;
  const int32_t* column_start;
  const double* nzval;
} sgo_refmat;

/* mult! (src/refinement_matrix.jl:421-445, kernel :365-403): Y and B have
 * nd dimensions (the last one the output channel); refmat_of[d] is 0 when
 * dimension d is not refined (then shape_y[d] == shape_b[d]), else the
 * 1-based index into mats of its refinement matrix.  For every Y index I,
 * Y[I] = sum over the column ranges J of prod_d R_d[I_d, J_d] * B[J], the
 * ranges walked first dimension fastest and the product taken B[J] first,
 * then dimension 1, 2, ... */
void sgo_refmat_mult(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                     const sgo_refmat* mats, const int32_t* refmat_of, const double* B,
                     double* Y);

/* mult_adjoint! (src/adjoint.jl:127-152, kernel :85-125): B zeroed, then
 * every Y index scatters prod_d R_d[I_d, J_d] * Y[I] into B[J]. */
void sgo_refmat_mult_adjoint(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                             const sgo_refmat* mats, const int32_t* refmat_of,
                             const double* Y, double* B);

/* The Flag form of the same scatter (deactivate_overwritten_control_points!,
 * src/control_points.jl:624-680 with Flag, src/utils.jl:236-247): B[J] is
 * set to 1 when some I with Y[I] != 0 holds J in its column ranges (the
 * values of the matrices do not matter); B is zeroed first. */
void sgo_refmat_mult_adjoint_flags(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                                   const sgo_refmat* mats, const int32_t* refmat_of,
                                   const uint8_t* Y, uint8_t* B);

/* local_refinement_kernel (src/control_points.jl:296-311): for i < n_idx,
 * cp[indices[i, :]..., c] = values[i, c]; indices [n_idx, nin] and values
 * [n_idx, nout] column-major, indices 1-based; cp has shape (shape..., nout). */
void sgo_local_refinement_overwrite(int32_t nin, int32_t nout, const int32_t* shape,
                                    int64_t n_idx, const int32_t* indices,
                                    const double* values, double* cp);

/* local_refinement_adjoint_kernel (src/adjoint.jl:154-170): values[i, c] =
 * cp[indices[i, :]..., c], then that entry of cp is zeroed. */
void sgo_local_refinement_gather(int32_t nin, int32_t nout, const int32_t* shape,
                                 int64_t n_idx, const int32_t* indices, double* values,
                                 double* cp);

int32_t sgo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
