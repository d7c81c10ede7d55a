// sg_basis.hip -- per-dimension FP64 basis tables on the GPU.
//
// One thread per sample point.  The span search replaces the linear scan of
// set_sample_indices_kernel (src/util_kernels.jl:22-49) by a binary search
// over the sorted knots: the first knot > t is the same index either way
// (ties go right, NaN runs to the end), clamped to [p+1, n_knots-p-1].
// The Cox-de Boor triangle and the derivative recursion follow
// spline_dimension_kernel (src/spline_dimension.jl:160-215) operation by
// operation, with FP contraction off, so the FP64 table equals a
// sequential FP64 evaluation of the same recursion bit for bit.
#include "sg_internal.hpp"

namespace sg {

namespace {

template <typename TIn>
__global__ __launch_bounds__(256) void basis_kernel(
    const TIn* __restrict__ knots, int n_knots, const TIn* __restrict__ t,
    int m, int p, int D, double* __restrict__ cur, double* __restrict__ prev,
    int32_t* __restrict__ span0, int32_t* __restrict__ span1) {
#pragma clang fp contract(off)
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= m) return;
  const double tt = static_cast<double>(t[l]);

  // upper bound: number of knots <= t
  int lo = 0, hi = n_knots;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (tt < static_cast<double>(knots[mid]))
      hi = mid;
    else
      lo = mid + 1;
  }
  int i = lo;  // 1-based span of the reference
  if (i < p + 1) i = p + 1;
  if (i > n_knots - p - 1) i = n_knots - p - 1;
  if (span0) span0[l] = i - 1;
  if (span1) span1[l] = i;

  const int w = p + 1;
  const size_t ms = static_cast<size_t>(m);
#define CUR(r, o) cur[static_cast<size_t>((r) + w * (o)) * ms + l]
#define PREV(r, o) prev[static_cast<size_t>((r) + w * (o)) * ms + l]
  for (int o = 0; o <= D; ++o)
    for (int r = 0; r < w; ++r) PREV(r, o) = 0.0;  // :174-176
  CUR(0, 0) = 1.0;                                  // :179
  PREV(0, 0) = 1.0;                                 // :180
  for (int k = 1; k <= p; ++k) {                    // :183
    for (int o = 0; o <= D; ++o)
      for (int r = 0; r < w; ++r) CUR(r, o) = 0.0;  // :185-187
    for (int kk = 1; kk <= k; ++kk) {               // :189
      const double t_min = static_cast<double>(knots[i + kk - k - 1]);
      const double t_max = static_cast<double>(knots[i + kk - 1]);
      const double dt = t_max - t_min;
      double b_prev = PREV(kk - 1, 0);
      const double frac = b_prev / dt;
      CUR(kk - 1, 0) += frac * (t_max - tt);        // :196
      CUR(kk, 0) = frac * (tt - t_min);             // :197
      for (int d = 1; d <= D + k - p; ++d) {        // :200
        b_prev = PREV(kk - 1, d - 1);
        const double contrib = b_prev * static_cast<double>(k) / dt;
        CUR(kk - 1, d) -= contrib;                  // :203
        CUR(kk, d) = contrib;                       // :204
      }
    }
    if (k != p)
      for (int o = 0; o <= D; ++o)
        for (int r = 0; r < w; ++r) PREV(r, o) = CUR(r, o);  // :209-213
  }
#undef CUR
#undef PREV
}

// Reference layout [m, p+1, D+1] -> kernel layout [D+1][m][p+1].
__global__ __launch_bounds__(256) void kernel_tables(
    const double* __restrict__ table64, int m, int p, int D,
    double* __restrict__ ktab64, float* __restrict__ ktab32) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * (D + 1)) return;
  const int j = idx % m, o = idx / m;
  const int w = p + 1;
  for (int a = 0; a < w; ++a) {
    const double v = table64[j + static_cast<size_t>(m) * (a + w * o)];
    const size_t dst = (static_cast<size_t>(o) * m + j) * w + a;
    ktab64[dst] = v;
    ktab32[dst] = static_cast<float>(v);
  }
}

__global__ __launch_bounds__(256) void cast_f64_f32(const double* __restrict__ in,
                                                    size_t count,
                                                    float* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) out[i] = static_cast<float>(in[i]);
}

__global__ __launch_bounds__(256) void copy_f64(const double* __restrict__ in,
                                                size_t count,
                                                double* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) out[i] = in[i];
}

inline unsigned blocks_for(size_t n) {
  return static_cast<unsigned>((n + 255) / 256);
}

}  // namespace

template <typename TIn>
hipError_t launch_basis(const double* /*knots64*/, const TIn* knots_in,
                        int n_knots, const TIn* t_in, int m, int p, int D,
                        double* table64, double* scratch, int32_t* span0,
                        int32_t* span1, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(basis_kernel<TIn>, dim3(blocks_for(m)), dim3(256), 0, s,
                     knots_in, n_knots, t_in, m, p, D, table64, scratch, span0,
                     span1);
  return hipGetLastError();
}

template hipError_t launch_basis<double>(const double*, const double*, int,
                                         const double*, int, int, int, double*,
                                         double*, int32_t*, int32_t*,
                                         hipStream_t);
template hipError_t launch_basis<float>(const double*, const float*, int,
                                        const float*, int, int, int, double*,
                                        double*, int32_t*, int32_t*,
                                        hipStream_t);

hipError_t launch_kernel_tables(const double* table64, int m, int p, int D,
                                double* ktab64, float* ktab32, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(kernel_tables, dim3(blocks_for(static_cast<size_t>(m) * (D + 1))),
                     dim3(256), 0, s, table64, m, p, D, ktab64, ktab32);
  return hipGetLastError();
}

hipError_t launch_cast_table(const double* table64, size_t count, float* out,
                             hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(cast_f64_f32, dim3(blocks_for(count)), dim3(256), 0, s,
                     table64, count, out);
  return hipGetLastError();
}

hipError_t launch_copy_f64(const double* in, size_t count, double* out,
                           hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(copy_f64, dim3(blocks_for(count)), dim3(256), 0, s, in,
                     count, out);
  return hipGetLastError();
}

}  // namespace sg
