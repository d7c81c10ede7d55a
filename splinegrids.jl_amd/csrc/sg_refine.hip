// sg_refine.hip -- refinement matrices and locally refined (THB) control
// points on the GPU.
//
// A refinement matrix R (m x n) has consecutive non-zeros in every row and
// column, stored as the reference's RefinementMatrix
// (src/refinement_matrix.jl:1-60): 1-based row_pointer[m], column_start[m],
// and the non-zeros of every row in nzval.  Its rows' first and last columns
// are non-decreasing, so R^T is banded the same way: the rows holding
// column j are one contiguous range, found by two binary searches.
//
// Kernels (one thread per output element, no atomics, fixed summation
// order):
//   insert_knot_rows   the (n+1) x n matrix of one knot insertion
//                      (src/refinement.jl:3-84)
//   product_rows /     C = A B, row extents then values
//   product_values     (src/refinement_matrix.jl:184-327)
//   validate_rows      validate_refinement_matrix_kernel (:134-181)
//   refmat_mult        mult! (:365-445): Y[I] = sum_J prod_d R_d[I_d, J_d] B[J]
//   refmat_mult_adj    mult_adjoint! (src/adjoint.jl:85-152) as a gather:
//                      B[J] = sum over the rows I holding J of
//                      prod_d R_d[I_d, J_d] Y[I], rows in increasing linear
//                      index (the order in which a sequential scatter adds
//                      them); on flags (uint8) the structural OR of
//                      deactivate_overwritten_control_points!
//   lr_overwrite /     local_refinement_kernel (src/control_points.jl:296-311)
//   lr_gather          and local_refinement_adjoint_kernel (src/adjoint.jl:154-170)
//
// The products are taken in the reference's order (the array value first,
// then dimension 1, 2, ...) with FP contraction off, so FP64 results equal a
// sequential evaluation of the reference's loops.
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int kMaxRefDims = SG_MAX_NIN + 1;  // input dimensions and the channel

template <typename T>
struct DevRefmat {
  int m, n, nnz;
  const int32_t* row_pointer;  // 1-based
  const int32_t* column_start;
  const T* nzval;
};

// 1-based first and last column of 1-based row i (get_column_range,
// src/refinement_matrix.jl:103-131).
template <typename T>
__device__ inline void row_range(const DevRefmat<T>& A, int i, int& cs, int& ce) {
  const int rp_next = i == A.m ? A.nnz + 1 : A.row_pointer[i];
  cs = A.column_start[i - 1];
  ce = cs + (rp_next - A.row_pointer[i - 1]) - 1;
}

template <typename T>
struct MultArgs {
  DevRefmat<T> mats[kMaxRefDims];
  int refmat_of[kMaxRefDims];  // 0: dimension not refined, else 1 + index into mats
  int shape_y[kMaxRefDims], shape_b[kMaxRefDims];
  long long stride_y[kMaxRefDims], stride_b[kMaxRefDims];
  long long size_y, size_b;
  int nd;
};

// ---- single knot insertion: rows of the (n+1) x n matrix ----

template <typename T, typename TK>
__global__ __launch_bounds__(256) void insert_knot_rows(const TK* __restrict__ knots, int n,
                                                        int degree, int span, double knot_new,
                                                        int32_t* __restrict__ row_pointer,
                                                        int32_t* __restrict__ column_start,
                                                        T* __restrict__ nzval) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;  // 1-based row
  if (i > n + 1) return;
  if (i <= span - degree) {  // left of the new knot: unchanged
    row_pointer[i - 1] = i;
    column_start[i - 1] = i;
    nzval[i - 1] = T(1);
  } else if (i <= span) {    // a combination of two old basis functions
    const double ki = static_cast<double>(knots[i - 1]);
    const double alpha = (knot_new - ki) / (static_cast<double>(knots[i + degree - 1]) - ki);
    const int p0 = 2 * i - span + degree - 1;
    row_pointer[i - 1] = p0;
    column_start[i - 1] = i - 1;
    nzval[p0 - 1] = static_cast<T>(1.0 - alpha);
    nzval[p0] = static_cast<T>(alpha);
  } else {                   // right of it: shifted by one
    row_pointer[i - 1] = i + degree;
    column_start[i - 1] = i - 1;
    nzval[i + degree - 1] = T(1);
  }
}

// ---- products ----

// Row i of C = A B covers the union of B's rows A's row i selects: with
// both matrices valid (starts and ends non-decreasing, no gaps) it runs from
// the first column of B's row cs_A(i) to the last of B's row ce_A(i).
__global__ __launch_bounds__(256) void product_rows(DevRefmat<float> A, DevRefmat<float> B,
                                                    int32_t* __restrict__ n_nonzero,
                                                    int32_t* __restrict__ column_start) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (i > A.m) return;
  int acs, ace, b0, b1, c0, c1;
  row_range(A, i, acs, ace);
  row_range(B, acs, b0, b1);
  row_range(B, ace, c0, c1);
  n_nonzero[i - 1] = c1 - b0 + 1;
  column_start[i - 1] = b0;
}

template <typename T>
__global__ __launch_bounds__(256) void product_values(DevRefmat<T> A, DevRefmat<T> B,
                                                      DevRefmat<T> C, T* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (i > A.m) return;
  int acs, ace, ccs, cce;
  row_range(A, i, acs, ace);
  row_range(C, i, ccs, cce);
  const int pc0 = C.row_pointer[i - 1];
  for (int j = ccs; j <= cce; ++j) out[pc0 - 1 + (j - ccs)] = T(0);
  int pa = A.row_pointer[i - 1];
  for (int k = acs; k <= ace; ++k, ++pa) {
    int bcs, bce;
    row_range(B, k, bcs, bce);
    for (int j = max(ccs, bcs); j <= min(cce, bce); ++j)
      out[pc0 - 1 + (j - ccs)] += A.nzval[pa - 1] * B.nzval[B.row_pointer[k - 1] + j - bcs - 1];
  }
}

__global__ __launch_bounds__(256) void validate_rows(DevRefmat<float> A,
                                                     uint8_t* __restrict__ valid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (i > A.m) return;
  int cs, ce;
  row_range(A, i, cs, ce);
  bool ok = ce >= cs;
  if (ok) ok = cs >= 1 && ce <= A.n;
  if (i == 1) ok = cs == 1 && A.row_pointer[0] == 1;
  if (ok && i > 1) {
    int csp, cep;
    row_range(A, i - 1, csp, cep);
    ok = csp <= cs && cs <= cep + 1 && ce >= cep;
  }
  valid[i - 1] = ok ? 1 : 0;
}

// ---- mult! ----

template <typename T>
__global__ __launch_bounds__(256) void refmat_mult(MultArgs<T> a, const T* __restrict__ B,
                                                   T* __restrict__ Y) {
#pragma clang fp contract(off)
  const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.size_y) return;
  int col0[kMaxRefDims], ncol[kMaxRefDims], rp[kMaxRefDims], jb[kMaxRefDims];
  long long r = q;
  long long nterms = 1;
  for (int d = 0; d < a.nd; ++d) {
    const int I = static_cast<int>(r % a.shape_y[d]);
    r /= a.shape_y[d];
    if (a.refmat_of[d] == 0) {
      col0[d] = I + 1;
      ncol[d] = 1;
      rp[d] = 0;
    } else {
      const DevRefmat<T>& A = a.mats[a.refmat_of[d] - 1];
      int cs, ce;
      row_range(A, I + 1, cs, ce);
      col0[d] = cs;
      ncol[d] = ce - cs + 1;
      rp[d] = A.row_pointer[I];
    }
    jb[d] = 0;
    nterms *= ncol[d];
  }
  T out = T(0);
  for (long long t = 0; t < nterms; ++t) {
    long long off = 0;
    for (int d = 0; d < a.nd; ++d) off += static_cast<long long>(col0[d] - 1 + jb[d]) * a.stride_b[d];
    T contrib = B[off];
    for (int d = 0; d < a.nd; ++d)
      if (a.refmat_of[d] != 0) contrib *= a.mats[a.refmat_of[d] - 1].nzval[rp[d] + jb[d] - 1];
    out += contrib;
    for (int d = 0; d < a.nd; ++d) {  // column-major odometer
      if (++jb[d] < ncol[d]) break;
      jb[d] = 0;
    }
  }
  Y[q] = out;
}

// ---- mult_adjoint! as a gather ----

// The 1-based rows of A holding column j: [lo, hi] (empty when lo > hi).
// Row ends are non-decreasing, so lo is the first row ending at or after j;
// row starts are non-decreasing, so hi is the last row starting at or
// before j.
template <typename T>
__device__ inline void rows_of_column(const DevRefmat<T>& A, int j, int& lo, int& hi) {
  int a = 1, b = A.m + 1;
  while (a < b) {
    const int mid = (a + b) >> 1;
    int cs, ce;
    row_range(A, mid, cs, ce);
    if (ce >= j) b = mid; else a = mid + 1;
  }
  lo = a;
  a = 0;
  b = A.m;  // the last row with start <= j, 0 if none
  while (a < b) {
    const int mid = (a + b + 1) >> 1;
    if (A.column_start[mid - 1] <= j) a = mid; else b = mid - 1;
  }
  hi = a;
}

template <typename T, bool FLAGS>
__global__ __launch_bounds__(256) void refmat_mult_adj(MultArgs<T> a, const void* __restrict__ Yv,
                                                       void* __restrict__ Bv, int accumulate) {
#pragma clang fp contract(off)
  const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.size_b) return;
  int J[kMaxRefDims], ilo[kMaxRefDims], nrow[kMaxRefDims], ib[kMaxRefDims];
  long long r = q;
  long long nterms = 1;
  for (int d = 0; d < a.nd; ++d) {
    J[d] = static_cast<int>(r % a.shape_b[d]) + 1;
    r /= a.shape_b[d];
    if (a.refmat_of[d] == 0) {
      ilo[d] = J[d];
      nrow[d] = 1;
    } else {
      int lo, hi;
      rows_of_column(a.mats[a.refmat_of[d] - 1], J[d], lo, hi);
      ilo[d] = lo;
      nrow[d] = max(0, hi - lo + 1);
    }
    ib[d] = 0;
    nterms *= nrow[d];
  }
  if constexpr (FLAGS) {
    const uint8_t* Y = static_cast<const uint8_t*>(Yv);
    uint8_t* B = static_cast<uint8_t*>(Bv);
    uint8_t any = accumulate ? B[q] : 0;
    for (long long t = 0; t < nterms && !any; ++t) {
      long long off = 0;
      for (int d = 0; d < a.nd; ++d) off += static_cast<long long>(ilo[d] - 1 + ib[d]) * a.stride_y[d];
      if (Y[off]) any = 1;
      for (int d = 0; d < a.nd; ++d) {
        if (++ib[d] < nrow[d]) break;
        ib[d] = 0;
      }
    }
    B[q] = any;
  } else {
    const T* Y = static_cast<const T*>(Yv);
    T* B = static_cast<T*>(Bv);
    T sum = accumulate ? B[q] : T(0);
    for (long long t = 0; t < nterms; ++t) {
      long long off = 0;
      for (int d = 0; d < a.nd; ++d) off += static_cast<long long>(ilo[d] - 1 + ib[d]) * a.stride_y[d];
      T contrib = Y[off];
      for (int d = 0; d < a.nd; ++d)
        if (a.refmat_of[d] != 0) {
          const DevRefmat<T>& A = a.mats[a.refmat_of[d] - 1];
          const int i = ilo[d] + ib[d];
          contrib *= A.nzval[A.row_pointer[i - 1] + J[d] - A.column_start[i - 1] - 1];
        }
      sum += contrib;
      for (int d = 0; d < a.nd; ++d) {
        if (++ib[d] < nrow[d]) break;
        ib[d] = 0;
      }
    }
    B[q] = sum;
  }
}

// ---- local refinement: overwrite and gather ----

struct LrArgs {
  int nin, nout;
  long long n_idx, ncp;
  long long stride[SG_MAX_NIN];
  int shape[SG_MAX_NIN];
};

// Linear offset of control point i's 1-based indices, or -1 when one of them
// lies outside the grid (that entry is then skipped, never accessed).
__device__ inline long long lr_offset(const LrArgs& a, const int32_t* __restrict__ idx,
                                      long long i) {
  long long lin = 0;
  for (int d = 0; d < a.nin; ++d) {
    const int k = idx[i + a.n_idx * d];
    if (k < 1 || k > a.shape[d]) return -1;
    lin += static_cast<long long>(k - 1) * a.stride[d];
  }
  return lin;
}

template <typename T>
__global__ __launch_bounds__(256) void lr_overwrite(LrArgs a, const int32_t* __restrict__ idx,
                                                    const T* __restrict__ values,
                                                    T* __restrict__ cp) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_idx) return;
  const long long lin = lr_offset(a, idx, i);
  if (lin < 0) return;
  for (int c = 0; c < a.nout; ++c) cp[lin + a.ncp * c] = values[i + a.n_idx * c];
}

template <typename T>
__global__ __launch_bounds__(256) void lr_gather(LrArgs a, const int32_t* __restrict__ idx,
                                                 T* __restrict__ values, T* __restrict__ cp,
                                                 int clear) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_idx) return;
  const long long lin = lr_offset(a, idx, i);
  for (int c = 0; c < a.nout; ++c) {
    if (lin < 0) {
      values[i + a.n_idx * c] = T(0);
      continue;
    }
    values[i + a.n_idx * c] = cp[lin + a.ncp * c];
    if (clear) cp[lin + a.ncp * c] = T(0);
  }
}

inline unsigned blocks_of(long long n) { return static_cast<unsigned>((n + 255) / 256); }

template <typename T>
DevRefmat<T> dev_refmat(const sg_refmat& r) {
  DevRefmat<T> d;
  d.m = r.m;
  d.n = r.n;
  d.nnz = r.nnz;
  d.row_pointer = r.row_pointer;
  d.column_start = r.column_start;
  d.nzval = static_cast<const T*>(r.nzval);
  return d;
}

int check_refmat(const sg_refmat& r, const char* what) {
  if (r.m < 1 || r.n < 1 || r.nnz < 1)
    return fail(SG_ERR_INVALID, std::string(what) + ": empty refinement matrix");
  if (!r.row_pointer || !r.column_start || !r.nzval)
    return fail(SG_ERR_INVALID, std::string(what) + ": null device pointer");
  return SG_OK;
}

template <typename T>
int fill_mult_args(MultArgs<T>& a, int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                   int32_t n_mats, const sg_refmat* mats, const int32_t* dims) {
  if (nd < 1 || nd > kMaxRefDims) return fail(SG_ERR_INVALID, "nd must be in 1..9");
  if (n_mats < 0 || n_mats > nd - 1 || (n_mats > 0 && (!mats || !dims)))
    return fail(SG_ERR_INVALID, "bad refinement matrix list");
  if (!shape_y || !shape_b) return fail(SG_ERR_INVALID, "null shape");
  a = MultArgs<T>();
  a.nd = nd;
  for (int d = 0; d < nd; ++d) a.refmat_of[d] = 0;
  for (int k = 0; k < n_mats; ++k) {
    const int d = dims[k] - 1;  // dimensions are 1-based, as dims_refinement
    if (d < 0 || d >= nd - 1) return fail(SG_ERR_INVALID, "refinement dimension out of range");
    if (a.refmat_of[d] != 0) return fail(SG_ERR_INVALID, "a dimension is refined twice");
    int rc = check_refmat(mats[k], "mult");
    if (rc != SG_OK) return rc;
    if (mats[k].m != shape_y[d] || mats[k].n != shape_b[d])
      return fail(SG_ERR_INVALID, "refinement matrix size does not match the arrays");
    a.refmat_of[d] = k + 1;
    a.mats[k] = dev_refmat<T>(mats[k]);
  }
  long long sy = 1, sb = 1;
  for (int d = 0; d < nd; ++d) {
    if (shape_y[d] < 0 || shape_b[d] < 0) return fail(SG_ERR_INVALID, "negative size");
    if (a.refmat_of[d] == 0 && shape_y[d] != shape_b[d])
      return fail(SG_ERR_INVALID, "unrefined dimensions must have equal sizes");
    a.shape_y[d] = shape_y[d];
    a.shape_b[d] = shape_b[d];
    a.stride_y[d] = sy;
    a.stride_b[d] = sb;
    sy *= shape_y[d];
    sb *= shape_b[d];
  }
  a.size_y = sy;
  a.size_b = sb;
  return SG_OK;
}

int fill_lr_args(LrArgs& a, int32_t nin, int32_t nout, const int32_t* shape, int64_t n_idx) {
  if (nin < 1 || nin > SG_MAX_NIN || nout < 1 || !shape || n_idx < 0)
    return fail(SG_ERR_INVALID, "bad local refinement arguments");
  a.nin = nin;
  a.nout = nout;
  a.n_idx = n_idx;
  long long s = 1;
  for (int d = 0; d < nin; ++d) {
    if (shape[d] < 0) return fail(SG_ERR_INVALID, "negative size");
    a.stride[d] = s;
    a.shape[d] = shape[d];
    s *= shape[d];
  }
  a.ncp = s;
  return SG_OK;
}

}  // namespace

}  // namespace sg

using namespace sg;

extern "C" {

int sg_refinement_insert_knot(const void* knots_all, int32_t n_basis, int32_t degree,
                              int32_t knot_span_index, double knot_new, int32_t* row_pointer,
                              int32_t* column_start, void* nzval, int32_t dtype, void* stream) {
  if (degree < 0 || n_basis < degree + 1) return fail(SG_ERR_INVALID, "bad n_basis or degree");
  if (knot_span_index < degree + 1 || knot_span_index > n_basis)
    return fail(SG_ERR_INVALID, "knot_span_index must be in degree+1 .. n_basis");
  if (!knots_all || !row_pointer || !column_start || !nzval)
    return fail(SG_ERR_INVALID, "null device pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned blocks = blocks_of(n_basis + 1);
  if (dtype == SG_F32)
    hipLaunchKernelGGL((insert_knot_rows<float, float>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const float*>(knots_all), n_basis, degree, knot_span_index,
                       knot_new, row_pointer, column_start, static_cast<float*>(nzval));
  else if (dtype == SG_F64)
    hipLaunchKernelGGL((insert_knot_rows<double, double>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const double*>(knots_all), n_basis, degree, knot_span_index,
                       knot_new, row_pointer, column_start, static_cast<double*>(nzval));
  else
    return fail(SG_ERR_INVALID, "dtype must be SG_F32 or SG_F64");
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SG_OK : hip_fail(e, "knot insertion launch");
}

int sg_refinement_product_rows(const sg_refmat* A, const sg_refmat* B, int32_t* n_nonzero,
                               int32_t* column_start, void* stream) {
  if (!A || !B || !n_nonzero || !column_start) return fail(SG_ERR_INVALID, "NULL argument");
  int rc = check_refmat(*A, "product");
  if (rc == SG_OK) rc = check_refmat(*B, "product");
  if (rc != SG_OK) return rc;
  if (A->n != B->m) return fail(SG_ERR_INVALID, "Inner dimensions must match");
  hipLaunchKernelGGL(product_rows, dim3(blocks_of(A->m)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), dev_refmat<float>(*A), dev_refmat<float>(*B),
                     n_nonzero, column_start);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SG_OK : hip_fail(e, "product rows launch");
}

int sg_refinement_product_values(const sg_refmat* A, const sg_refmat* B, const sg_refmat* C,
                                 int32_t dtype, void* stream) {
  if (!A || !B || !C) return fail(SG_ERR_INVALID, "NULL argument");
  int rc = check_refmat(*A, "product");
  if (rc == SG_OK) rc = check_refmat(*B, "product");
  if (rc == SG_OK) rc = check_refmat(*C, "product");
  if (rc != SG_OK) return rc;
  if (A->n != B->m || C->m != A->m || C->n != B->n)
    return fail(SG_ERR_INVALID, "Inner dimensions must match");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == SG_F32)
    hipLaunchKernelGGL(product_values<float>, dim3(blocks_of(A->m)), dim3(256), 0, s,
                       dev_refmat<float>(*A), dev_refmat<float>(*B), dev_refmat<float>(*C),
                       const_cast<float*>(static_cast<const float*>(C->nzval)));
  else if (dtype == SG_F64)
    hipLaunchKernelGGL(product_values<double>, dim3(blocks_of(A->m)), dim3(256), 0, s,
                       dev_refmat<double>(*A), dev_refmat<double>(*B), dev_refmat<double>(*C),
                       const_cast<double*>(static_cast<const double*>(C->nzval)));
  else
    return fail(SG_ERR_INVALID, "dtype must be SG_F32 or SG_F64");
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SG_OK : hip_fail(e, "product values launch");
}

int sg_refinement_validate(const sg_refmat* A, uint8_t* valid_rows, void* stream) {
  if (!A || !valid_rows) return fail(SG_ERR_INVALID, "NULL argument");
  if (A->m < 1) return SG_OK;
  if (!A->row_pointer || !A->column_start) return fail(SG_ERR_INVALID, "null device pointer");
  hipLaunchKernelGGL(validate_rows, dim3(blocks_of(A->m)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), dev_refmat<float>(*A), valid_rows);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SG_OK : hip_fail(e, "validate launch");
}

int sg_refinement_mult(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                       int32_t n_mats, const sg_refmat* mats, const int32_t* dims, const void* B,
                       void* Y, int32_t dtype, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == SG_F32 || dtype == SG_F64) {
    long long size_y = 0;
    if (dtype == SG_F32) {
      MultArgs<float> a;
      int rc = fill_mult_args(a, nd, shape_y, shape_b, n_mats, mats, dims);
      if (rc != SG_OK) return rc;
      size_y = a.size_y;
      if (size_y > 0 && (!B || !Y)) return fail(SG_ERR_INVALID, "null device pointer");
      if (size_y > 0)
        hipLaunchKernelGGL(refmat_mult<float>, dim3(blocks_of(size_y)), dim3(256), 0, s, a,
                           static_cast<const float*>(B), static_cast<float*>(Y));
    } else {
      MultArgs<double> a;
      int rc = fill_mult_args(a, nd, shape_y, shape_b, n_mats, mats, dims);
      if (rc != SG_OK) return rc;
      size_y = a.size_y;
      if (size_y > 0 && (!B || !Y)) return fail(SG_ERR_INVALID, "null device pointer");
      if (size_y > 0)
        hipLaunchKernelGGL(refmat_mult<double>, dim3(blocks_of(size_y)), dim3(256), 0, s, a,
                           static_cast<const double*>(B), static_cast<double*>(Y));
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? SG_OK : hip_fail(e, "mult launch");
  }
  return fail(SG_ERR_INVALID, "dtype must be SG_F32 or SG_F64");
}

int sg_refinement_mult_adjoint(int32_t nd, const int32_t* shape_y, const int32_t* shape_b,
                               int32_t n_mats, const sg_refmat* mats, const int32_t* dims,
                               const void* Y, void* B, int32_t dtype, int32_t accumulate,
                               void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == SG_F32 || dtype == SG_U8) {
    MultArgs<float> a;
    int rc = fill_mult_args(a, nd, shape_y, shape_b, n_mats, mats, dims);
    if (rc != SG_OK) return rc;
    if (a.size_b > 0 && (!B || (a.size_y > 0 && !Y)))
      return fail(SG_ERR_INVALID, "null device pointer");
    if (a.size_b > 0) {
      if (dtype == SG_F32)
        hipLaunchKernelGGL((refmat_mult_adj<float, false>), dim3(blocks_of(a.size_b)), dim3(256),
                           0, s, a, Y, B, accumulate);
      else
        hipLaunchKernelGGL((refmat_mult_adj<float, true>), dim3(blocks_of(a.size_b)), dim3(256),
                           0, s, a, Y, B, accumulate);
    }
  } else if (dtype == SG_F64) {
    MultArgs<double> a;
    int rc = fill_mult_args(a, nd, shape_y, shape_b, n_mats, mats, dims);
    if (rc != SG_OK) return rc;
    if (a.size_b > 0 && (!B || (a.size_y > 0 && !Y)))
      return fail(SG_ERR_INVALID, "null device pointer");
    if (a.size_b > 0)
      hipLaunchKernelGGL((refmat_mult_adj<double, false>), dim3(blocks_of(a.size_b)), dim3(256),
                         0, s, a, Y, B, accumulate);
  } else {
    return fail(SG_ERR_INVALID, "dtype must be SG_F32, SG_F64 or SG_U8");
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SG_OK : hip_fail(e, "mult_adjoint launch");
}

int sg_local_refinement_overwrite(int32_t nin, int32_t nout, const int32_t* shape, int64_t n_idx,
                                  const int32_t* indices, const void* values,
                                  void* control_points, int32_t dtype, void* stream) {
  LrArgs a;
  int rc = fill_lr_args(a, nin, nout, shape, n_idx);
  if (rc != SG_OK) return rc;
  if (n_idx == 0) return SG_OK;
  if (!indices || !values || !control_points) return fail(SG_ERR_INVALID, "null device pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned blocks = blocks_of(n_idx);
  if (dtype == SG_F32)
    hipLaunchKernelGGL(lr_overwrite<float>, dim3(blocks), dim3(256), 0, s, a, indices,
                       static_cast<const float*>(values), static_cast<float*>(control_points));
  else if (dtype == SG_F64)
    hipLaunchKernelGGL(lr_overwrite<double>, dim3(blocks), dim3(256), 0, s, a, indices,
                       static_cast<const double*>(values), static_cast<double*>(control_points));
  else if (dtype == SG_U8)
    hipLaunchKernelGGL(lr_overwrite<uint8_t>, dim3(blocks), dim3(256), 0, s, a, indices,
                       static_cast<const uint8_t*>(values), static_cast<uint8_t*>(control_points));
  else
    return fail(SG_ERR_INVALID, "dtype must be SG_F32, SG_F64 or SG_U8");
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SG_OK : hip_fail(e, "local refinement launch");
}

int sg_local_refinement_gather(int32_t nin, int32_t nout, const int32_t* shape, int64_t n_idx,
                               const int32_t* indices, void* values, void* control_points,
                               int32_t clear, int32_t dtype, void* stream) {
  LrArgs a;
  int rc = fill_lr_args(a, nin, nout, shape, n_idx);
  if (rc != SG_OK) return rc;
  if (n_idx == 0) return SG_OK;
  if (!indices || !values || !control_points) return fail(SG_ERR_INVALID, "null device pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned blocks = blocks_of(n_idx);
  if (dtype == SG_F32)
    hipLaunchKernelGGL(lr_gather<float>, dim3(blocks), dim3(256), 0, s, a, indices,
                       static_cast<float*>(values), static_cast<float*>(control_points), clear);
  else if (dtype == SG_F64)
    hipLaunchKernelGGL(lr_gather<double>, dim3(blocks), dim3(256), 0, s, a, indices,
                       static_cast<double*>(values), static_cast<double*>(control_points), clear);
  else if (dtype == SG_U8)
    hipLaunchKernelGGL(lr_gather<uint8_t>, dim3(blocks), dim3(256), 0, s, a, indices,
                       static_cast<uint8_t*>(values), static_cast<uint8_t*>(control_points), clear);
  else
    return fail(SG_ERR_INVALID, "dtype must be SG_F32, SG_F64 or SG_U8");
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SG_OK : hip_fail(e, "local refinement launch");
}

}  // extern "C"
