/*
 * sg_oracle.c -- CPU restatement of the SplineGrids.jl grid-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
 * timed by bench.py.  The product never links or calls this file.
 *
 * Built with -ffp-contract=off so that a*b+c is two roundings, as in the
 * reference's Julia kernels (Julia does not contract without @fastmath or
 * muladd).  Each function cites the reference file:line it restates.
 */
#include "sg_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int32_t sgo_max_threads(void) {
#ifdef _OPENMP
  return (int32_t)omp_get_max_threads();
#else
  return 1;
#endif
}

/* src/util_kernels.jl:1-20 (launched from src/knot_vector.jl:31-36). */
int32_t sgo_expand_knots(const double* values, const int32_t* mults, int32_t k,
                         double* knots_all) {
  int32_t pos = 0;
  for (int32_t i = 0; i < k; ++i)
    for (int32_t r = 0; r < mults[i]; ++r) knots_all[pos++] = values[i];
  return pos;
}

/* src/util_kernels.jl:22-49. */
void sgo_sample_indices(const double* t, int32_t m, const double* knots_all,
                        int32_t n_knots, int32_t p, int32_t* idx1) {
  for (int32_t l = 0; l < m; ++l) {
    const double s = t[l];
    int32_t count = 0;
    for (int32_t k = 0; k < n_knots; ++k) {
      if (s < knots_all[k]) break; /* :35-41, ties go right */
      ++count;
    }
    const int32_t lo = p + 1, hi = n_knots - p - 1; /* :43-47 */
    if (count < lo) count = lo;
    if (count > hi) count = hi;
    idx1[l] = count;
  }
}

/* src/spline_dimension.jl:160-215.  Same loop nest and operation order;
 * 1-based reference indices are shifted to 0-based here.  eval/eval_prev are
 * per-sample scratch of shape [p+1, D+1]. */
void sgo_basis(const double* knots_all, const double* t, const int32_t* idx1,
               int32_t m, int32_t p, int32_t max_deriv, double* eval) {
  const int32_t w = p + 1, nd = max_deriv + 1;
  double* cur = (double*)calloc((size_t)w * nd, sizeof(double));
  double* prev = (double*)calloc((size_t)w * nd, sizeof(double));
#define CUR(r, o) cur[(r) + w * (o)]
#define PREV(r, o) prev[(r) + w * (o)]
  for (int32_t l = 0; l < m; ++l) {
    const double tt = t[l];
    const int32_t i = idx1[l]; /* 1-based span */
    memset(prev, 0, sizeof(double) * (size_t)w * nd); /* :174-176 */
    memset(cur, 0, sizeof(double) * (size_t)w * nd);
    CUR(0, 0) = 1.0; /* :179 */
    PREV(0, 0) = 1.0; /* :180 */
    for (int32_t k = 1; k <= p; ++k) { /* :183 */
      memset(cur, 0, sizeof(double) * (size_t)w * nd); /* :185-187 */
      for (int32_t kk = 1; kk <= k; ++kk) { /* :189 */
        /* knots_all[i + kk - k], knots_all[i + kk] are 1-based (:190-191) */
        const double t_min = knots_all[i + kk - k - 1];
        const double t_max = knots_all[i + kk - 1];
        const double dt = t_max - t_min;
        double b_prev = PREV(kk - 1, 0);
        const double frac = b_prev / dt;
        CUR(kk - 1, 0) += frac * (t_max - tt); /* :196 */
        CUR(kk, 0) = frac * (tt - t_min);      /* :197 */
        for (int32_t d = 1; d <= max_deriv + k - p; ++d) { /* :200 */
          b_prev = PREV(kk - 1, d - 1);
          const double contrib = b_prev * k / dt; /* :202 */
          CUR(kk - 1, d) -= contrib;              /* :203 */
          CUR(kk, d) = contrib;                   /* :204 */
        }
      }
      if (k != p) memcpy(prev, cur, sizeof(double) * (size_t)w * nd); /* :209-213 */
    }
    for (int32_t o = 0; o < nd; ++o)
      for (int32_t r = 0; r < w; ++r)
        eval[l + (size_t)m * (r + (size_t)w * o)] = CUR(r, o);
  }
#undef CUR
#undef PREV
  free(cur);
  free(prev);
}

/* Shared geometry of one grid: strides and per-dimension offsets. */
typedef struct {
  int32_t nin, nout;
  int64_t n_cp, n_samples;
  int64_t cp_stride[SGO_MAX_NIN], s_stride[SGO_MAX_NIN];
  int32_t width[SGO_MAX_NIN];
  int64_t n_terms;
} geom_t;

static void geom_init(geom_t* g, int32_t nin, int32_t nout, const int32_t* n,
                      const int32_t* m, const int32_t* p) {
  g->nin = nin;
  g->nout = nout;
  g->n_cp = 1;
  g->n_samples = 1;
  g->n_terms = 1;
  for (int32_t d = 0; d < nin; ++d) {
    g->cp_stride[d] = g->n_cp;
    g->s_stride[d] = g->n_samples;
    g->n_cp *= n[d];
    g->n_samples *= m[d];
    g->width[d] = p[d] + 1;
    g->n_terms *= p[d] + 1;
  }
}

/* src/spline_grid.jl:129-183, templated on the value type T.  One sample:
 * J (multi-index) -> out[c * ostride] for c < nout. */
#define DEFINE_EVAL_ONE(T, SUF)                                                \
  static void eval_one_##SUF(const geom_t* g, const int32_t* m,               \
                             const int32_t* p, const T* const* tables,        \
                             const int32_t* const* idx1,                      \
                             const int32_t* deriv, const T* P, const T* W,    \
                             const int32_t* J, T* out, int64_t ostride) {     \
    const int32_t nin = g->nin, nout = g->nout;                               \
    int64_t base[SGO_MAX_NIN];                                                \
    const T* tab[SGO_MAX_NIN];                                                \
    for (int32_t d = 0; d < nin; ++d) {                                       \
      base[d] = (int64_t)idx1[d][J[d]] - p[d] - 1; /* :145-146, 0-based */    \
      tab[d] = tables[d] + J[d] + (int64_t)m[d] * ((int64_t)(p[d] + 1) *      \
                                                   deriv[d]);                 \
    }                                                                         \
    T acc[64];                                                                \
    for (int32_t c = 0; c < nout; ++c) acc[c] = (T)0; /* :138-140 */          \
    T denom = (T)0;                                                           \
    int32_t I[SGO_MAX_NIN];                                                   \
    for (int32_t d = 0; d < nin; ++d) I[d] = 0;                               \
    for (int64_t term = 0; term < g->n_terms; ++term) { /* :148 */            \
      int64_t cp = 0;                                                         \
      T prod = (T)1; /* :153 */                                               \
      for (int32_t d = 0; d < nin; ++d) {                                     \
        cp += (base[d] + I[d]) * g->cp_stride[d];                             \
        prod *= tab[d][(int64_t)m[d] * I[d]]; /* :155-158 */                  \
      }                                                                       \
      if (W) { /* :161-165 */                                                 \
        prod = prod * W[cp];                                                  \
        denom += prod;                                                        \
      }                                                                       \
      for (int32_t c = 0; c < nout; ++c) /* :171-174 */                       \
        acc[c] += prod * P[cp + g->n_cp * c];                                 \
      for (int32_t d = 0; d < nin; ++d) { /* column-major odometer */         \
        if (++I[d] < g->width[d]) break;                                      \
        I[d] = 0;                                                             \
      }                                                                       \
    }                                                                         \
    if (W)                                                                    \
      for (int32_t c = 0; c < nout; ++c) acc[c] /= denom; /* :178-182 */      \
    for (int32_t c = 0; c < nout; ++c) out[c * ostride] = acc[c];             \
  }

DEFINE_EVAL_ONE(double, f64)
DEFINE_EVAL_ONE(float, f32)

#define DEFINE_EVAL_DIRECT(T, SUF)                                             \
  void sgo_eval_direct_##SUF(                                                 \
      int32_t nin, int32_t nout, const int32_t* n, const int32_t* m,          \
      const int32_t* p, const int32_t* D, const T* const* tables,             \
      const int32_t* const* idx1, const int32_t* deriv, const T* P,           \
      const T* W, T* eval, int32_t last_begin, int32_t last_end,              \
      int32_t nthreads) {                                                     \
    (void)D;                                                                  \
    geom_t g;                                                                 \
    geom_init(&g, nin, nout, n, m, p);                                        \
    const int64_t slab = g.s_stride[nin - 1]; /* samples per last index */    \
    const int64_t n_out = slab * (int64_t)(last_end - last_begin);            \
    if (nthreads < 1) nthreads = 1;                                           \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")        \
    for (int64_t q = 0; q < n_out; ++q) {                                     \
      int32_t J[SGO_MAX_NIN];                                                 \
      int64_t r = q;                                                          \
      for (int32_t d = 0; d < nin - 1; ++d) {                                 \
        J[d] = (int32_t)(r % m[d]);                                           \
        r /= m[d];                                                            \
      }                                                                       \
      J[nin - 1] = (int32_t)r + last_begin;                                   \
      eval_one_##SUF(&g, m, p, tables, idx1, deriv, P, W, J, eval + q,        \
                     n_out);                                                  \
    }                                                                         \
  }

DEFINE_EVAL_DIRECT(double, f64)
DEFINE_EVAL_DIRECT(float, f32)

void sgo_eval_points_f64(int32_t nin, int32_t nout, const int32_t* n,
                         const int32_t* m, const int32_t* p, const int32_t* D,
                         const double* const* tables,
                         const int32_t* const* idx1, const int32_t* deriv,
                         const double* P, const double* W,
                         const int64_t* lin, int64_t npts, double* out,
                         int32_t nthreads) {
  (void)D;
  geom_t g;
  geom_init(&g, nin, nout, n, m, p);
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t q = 0; q < npts; ++q) {
    int32_t J[SGO_MAX_NIN];
    int64_t r = lin[q];
    for (int32_t d = 0; d < nin; ++d) {
      J[d] = (int32_t)(r % m[d]);
      r /= m[d];
    }
    eval_one_f64(&g, m, p, tables, idx1, deriv, P, W, J, out + q, npts);
  }
}

/* src/adjoint.jl:1-40 (kernel) and :52-83 (zeroing at :61). */
#define DEFINE_ADJOINT(T, SUF, ATOMIC)                                         \
  static void adjoint_##SUF(int32_t nin, int32_t nout, const int32_t* n,      \
                            const int32_t* m, const int32_t* p,               \
                            const T* const* tables,                           \
                            const int32_t* const* idx1, const int32_t* deriv, \
                            const T* ebar, T* Pbar, int32_t nthreads) {       \
    geom_t g;                                                                 \
    geom_init(&g, nin, nout, n, m, p);                                        \
    memset(Pbar, 0, sizeof(T) * (size_t)(g.n_cp * nout));                     \
    if (nthreads < 1) nthreads = 1;                                           \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")        \
    for (int64_t q = 0; q < g.n_samples; ++q) {                               \
      int32_t J[SGO_MAX_NIN], I[SGO_MAX_NIN];                                 \
      int64_t base[SGO_MAX_NIN];                                              \
      const T* tab[SGO_MAX_NIN];                                              \
      int64_t r = q;                                                          \
      for (int32_t d = 0; d < nin; ++d) {                                     \
        J[d] = (int32_t)(r % m[d]);                                           \
        r /= m[d];                                                            \
        base[d] = (int64_t)idx1[d][J[d]] - p[d] - 1; /* :18-19 */             \
        tab[d] = tables[d] + J[d] +                                           \
                 (int64_t)m[d] * ((int64_t)(p[d] + 1) * deriv[d]);            \
        I[d] = 0;                                                             \
      }                                                                       \
      for (int64_t term = 0; term < g.n_terms; ++term) { /* :21 */            \
        int64_t cp = 0;                                                       \
        T prod = (T)1;                                                        \
        for (int32_t d = 0; d < nin; ++d) {                                   \
          cp += (base[d] + I[d]) * g.cp_stride[d];                            \
          prod *= tab[d][(int64_t)m[d] * I[d]]; /* :28-31 */                  \
        }                                                                     \
        for (int32_t c = 0; c < nout; ++c) { /* :34-38 */                     \
          const T v = prod * ebar[q + g.n_samples * c];                       \
          ATOMIC                                                              \
          Pbar[cp + g.n_cp * c] += v;                                         \
        }                                                                     \
        for (int32_t d = 0; d < nin; ++d) {                                   \
          if (++I[d] < g.width[d]) break;                                     \
          I[d] = 0;                                                           \
        }                                                                     \
      }                                                                       \
    }                                                                         \
  }

DEFINE_ADJOINT(double, f64, )
DEFINE_ADJOINT(float, f32, _Pragma("omp atomic"))

void sgo_adjoint_direct_f64(int32_t nin, int32_t nout, const int32_t* n,
                            const int32_t* m, const int32_t* p,
                            const int32_t* D, const double* const* tables,
                            const int32_t* const* idx1, const int32_t* deriv,
                            const double* ebar, double* Pbar) {
  (void)D;
  /* one thread: a fixed summation order, so the checker is deterministic */
  adjoint_f64(nin, nout, n, m, p, tables, idx1, deriv, ebar, Pbar, 1);
}

void sgo_adjoint_direct_f32(int32_t nin, int32_t nout, const int32_t* n,
                            const int32_t* m, const int32_t* p,
                            const int32_t* D, const float* const* tables,
                            const int32_t* const* idx1, const int32_t* deriv,
                            const float* ebar, float* Pbar, int32_t nthreads) {
  (void)D;
  adjoint_f32(nin, nout, n, m, p, tables, idx1, deriv, ebar, Pbar, nthreads);
}

/* ---- refinement matrices and THB control points ---- */

/* 1-based column range of row i (1-based) of a refinement matrix:
 * src/refinement_matrix.jl:103-131 (get_column_end, get_column_range). */
static void refmat_row(const sgo_refmat* A, int32_t i, int32_t* cs, int32_t* ce) {
  const int32_t rp_next = (i == A->m) ? A->nnz + 1 : A->row_pointer[i];
  const int32_t rp = A->row_pointer[i - 1];
  *cs = A->column_start[i - 1];
  *ce = *cs + (rp_next - rp) - 1;
}

typedef struct {
  int32_t nd;
  int64_t stride_y[SGO_MAX_NIN + 1], stride_b[SGO_MAX_NIN + 1];
  int64_t size_y;
} refmat_geom;

static void refmat_geom_init(refmat_geom* g, int32_t nd, const int32_t* shape_y,
                             const int32_t* shape_b) {
  g->nd = nd;
  int64_t sy = 1, sb = 1;
  for (int32_t d = 0; d < nd; ++d) {
    g->stride_y[d] = sy;
    g->stride_b[d] = sb;
    sy *= shape_y[d];
    sb *= shape_b[d];
  }
  g->size_y = sy;
}

/* For Y index I (0-based, from linear q): per dimension the first column
 * (1-based), the number of columns and the first nzval position (1-based,
 * 0 for an unrefined dimension): src/utils.jl:204-234 (get_row_extends). */
static void refmat_extends(const refmat_geom* g, const int32_t* shape_y, const sgo_refmat* mats,
                           const int32_t* refmat_of, int64_t q, int32_t* I, int32_t* col0,
                           int32_t* ncol, int32_t* rp) {
  int64_t r = q;
  for (int32_t d = 0; d < g->nd; ++d) {
    I[d] = (int32_t)(r % shape_y[d]);
    r /= shape_y[d];
    if (refmat_of[d] == 0) {
      col0[d] = I[d] + 1;
      ncol[d] = 1;
      rp[d] = 0;
    } else {
      const sgo_refmat* A = &mats[refmat_of[d] - 1];
      int32_t cs, ce;
      refmat_row(A, I[d] + 1, &cs, &ce);
      col0[d] = cs;
      ncol[d] = ce - cs + 1;
      rp[d] = A->row_pointer[I[d]];
    }
  }
}

#define REFMAT_WALK(BODY)                                                      \
  for (int64_t q = 0; q < g.size_y; ++q) {                                     \
    int32_t I[SGO_MAX_NIN + 1], col0[SGO_MAX_NIN + 1], ncol[SGO_MAX_NIN + 1],  \
        rp[SGO_MAX_NIN + 1], Jb[SGO_MAX_NIN + 1];                              \
    refmat_extends(&g, shape_y, mats, refmat_of, q, I, col0, ncol, rp);        \
    int64_t nterms = 1;                                                        \
    for (int32_t d = 0; d < nd; ++d) {                                         \
      nterms *= ncol[d];                                                       \
      Jb[d] = 0;                                                               \
    }                                                                          \
    for (int64_t t = 0; t < nterms; ++t) {                                     \
      int64_t jb = 0;                                                          \
      for (int32_t d = 0; d < nd; ++d) jb += (int64_t)(col0[d] - 1 + Jb[d]) * g.stride_b[d]; \
      BODY                                                                     \
      for (int32_t d = 0; d < nd; ++d) { /* column-major odometer */           \
        if (++Jb[d] < ncol[d]) break;                                          \
        Jb[d] = 0;                                                             \
      }                                                                        \
    }                                                                          \
  }

void sgo_refmat_mult(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                     const sgo_refmat* mats, const int32_t* refmat_of, const double* B,
                     double* Y) {
  refmat_geom g;
  refmat_geom_init(&g, nd, shape_y, shape_b);
  for (int64_t q = 0; q < g.size_y; ++q) Y[q] = 0.0; /* :386 out = zero */
  REFMAT_WALK({
    double contrib = B[jb]; /* :391 */
    for (int32_t d = 0; d < nd; ++d)
      if (refmat_of[d] != 0) /* :392-398 */
        contrib *= mats[refmat_of[d] - 1].nzval[rp[d] + Jb[d] - 1];
    Y[q] += contrib; /* :399-402 */
  })
}

void sgo_refmat_mult_adjoint(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                             const sgo_refmat* mats, const int32_t* refmat_of,
                             const double* Y, double* B) {
  refmat_geom g;
  refmat_geom_init(&g, nd, shape_y, shape_b);
  int64_t size_b = 1;
  for (int32_t d = 0; d < nd; ++d) size_b *= shape_b[d];
  for (int64_t k = 0; k < size_b; ++k) B[k] = 0.0; /* src/adjoint.jl:135 */
  REFMAT_WALK({
    double contrib = Y[q]; /* :109 */
    for (int32_t d = 0; d < nd; ++d)
      if (refmat_of[d] != 0) /* :110-116 */
        contrib *= mats[refmat_of[d] - 1].nzval[rp[d] + Jb[d] - 1];
    B[jb] += contrib; /* :122 (atomic in the reference) */
  })
}

void sgo_refmat_mult_adjoint_flags(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                                   const sgo_refmat* mats, const int32_t* refmat_of,
                                   const uint8_t* Y, uint8_t* B) {
  refmat_geom g;
  refmat_geom_init(&g, nd, shape_y, shape_b);
  int64_t size_b = 1;
  for (int32_t d = 0; d < nd; ++d) size_b *= shape_b[d];
  for (int64_t k = 0; k < size_b; ++k) B[k] = 0;
  REFMAT_WALK({
    if (Y[q]) B[jb] = 1; /* src/adjoint.jl:117-120: a true Flag is written */
  })
}

void sgo_local_refinement_overwrite(int32_t nin, int32_t nout, const int32_t* shape,
                                    int64_t n_idx, const int32_t* indices,
                                    const double* values, double* cp) {
  int64_t ncp = 1;
  for (int32_t d = 0; d < nin; ++d) ncp *= shape[d];
  for (int64_t i = 0; i < n_idx; ++i) {
    int64_t lin = 0, stride = 1;
    for (int32_t d = 0; d < nin; ++d) {
      lin += (int64_t)(indices[i + n_idx * d] - 1) * stride; /* :305 */
      stride *= shape[d];
    }
    for (int32_t c = 0; c < nout; ++c) cp[lin + ncp * c] = values[i + n_idx * c]; /* :307-310 */
  }
}

void sgo_local_refinement_gather(int32_t nin, int32_t nout, const int32_t* shape,
                                 int64_t n_idx, const int32_t* indices, double* values,
                                 double* cp) {
  int64_t ncp = 1;
  for (int32_t d = 0; d < nin; ++d) ncp *= shape[d];
  for (int64_t i = 0; i < n_idx; ++i) {
    int64_t lin = 0, stride = 1;
    for (int32_t d = 0; d < nin; ++d) {
      lin += (int64_t)(indices[i + n_idx * d] - 1) * stride; /* :164 */
      stride *= shape[d];
    }
    for (int32_t c = 0; c < nout; ++c) { /* :166-169 */
      values[i + n_idx * c] = cp[lin + ncp * c];
      cp[lin + ncp * c] = 0.0;
    }
  }
}
