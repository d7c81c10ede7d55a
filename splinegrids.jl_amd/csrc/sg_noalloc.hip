// sg_noalloc.hip -- sg_dimension_evaluate_noalloc: the basis tables of one
// dimension without any device allocation.
//
// sg_dimension_evaluate allocates its Cox-de Boor scratch (and, for FP32
// output, the FP64 table before rounding) with hipMalloc on every call and
// synchronises to free it.  Callers that cannot allocate -- code under HIP
// graph capture, tight loops that move sample points every iteration,
// harnesses that forbid hipMalloc -- take this variant: the same kernels,
// with the scratch in a file-scope device array of kScratchDoubles doubles
// (8 MiB, held by every GPU the library runs on until the process ends), and
// no synchronisation (stream-ordered like every evaluate call).
//
// The array's address is looked up once per device with
// hipGetSymbolAddress, on the first call made with that device current; the
// kernels only ever receive that address.  Every call on a GPU shares the
// array, so calls that overlap on one GPU (different streams, or threads)
// race and must be serialised by the caller.  Nothing in the array needs to
// start at zero: the basis kernel clears each sample's triangle before use.
#include <mutex>

#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr size_t kScratchDoubles = size_t(1) << 20;  // 8 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once).
int scratch_address(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(SG_ERR_UNSUPPORTED, "device ordinal above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(SG_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SG_OK;
}

}  // namespace

}  // namespace sg

using namespace sg;

extern "C" int sg_dimension_evaluate_noalloc(const void* knots_all, int32_t n_knots,
                                             const void* sample_points, int32_t n_samples,
                                             int32_t degree, int32_t max_derivative_order,
                                             int32_t* sample_indices, void* eval,
                                             int32_t dtype, void* stream) {
  if (degree < 0) return fail(SG_ERR_INVALID, "degree must be >= 0");
  if (max_derivative_order < 0 || max_derivative_order > degree)
    return fail(SG_ERR_INVALID,
                "The max_degree must be positive and derivatives order higher "
                "than `degree` are all 0.");
  if (n_knots < 2 * degree + 2)
    return fail(SG_ERR_INVALID, "n_knots must be at least 2 * degree + 2 (n_basis > degree)");
  if (n_samples < 0) return fail(SG_ERR_INVALID, "n_samples must be >= 0");
  if (dtype != SG_F32 && dtype != SG_F64)
    return fail(SG_ERR_INVALID, "dtype must be SG_F32 or SG_F64");
  if (n_samples == 0) return SG_OK;
  if (!knots_all || !sample_points || !sample_indices || !eval)
    return fail(SG_ERR_INVALID, "null device pointer");
  // FP64 output: the scratch holds the previous-degree triangle; FP32 output:
  // also the FP64 table before its rounding to FP32
  const size_t count =
      static_cast<size_t>(n_samples) * (degree + 1) * (max_derivative_order + 1);
  const size_t need = (dtype == SG_F32 ? 2 : 1) * count;
  if (need > kScratchDoubles)
    return fail(SG_ERR_INVALID,
                "sg_dimension_evaluate_noalloc: n_samples * (degree + 1) * "
                "(max_derivative_order + 1) = " + std::to_string(count) +
                " exceeds the fixed scratch (" +
                std::to_string(kScratchDoubles / (dtype == SG_F32 ? 2 : 1)) +
                " for this value type); use sg_dimension_evaluate");
  double* ptr = nullptr;
  const int rc = scratch_address(&ptr);
  if (rc != SG_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  if (dtype == SG_F32) {
    double* table = ptr;
    double* prev = ptr + count;
    e = launch_basis<float>(nullptr, static_cast<const float*>(knots_all), n_knots,
                            static_cast<const float*>(sample_points), n_samples, degree,
                            max_derivative_order, table, prev, nullptr, sample_indices, s);
    if (e == hipSuccess) e = launch_cast_table(table, count, static_cast<float*>(eval), s);
  } else {
    e = launch_basis<double>(nullptr, static_cast<const double*>(knots_all), n_knots,
                             static_cast<const double*>(sample_points), n_samples, degree,
                             max_derivative_order, static_cast<double*>(eval), ptr, nullptr,
                             sample_indices, s);
  }
  return e == hipSuccess ? SG_OK : hip_fail(e, "basis kernel launch");
}
