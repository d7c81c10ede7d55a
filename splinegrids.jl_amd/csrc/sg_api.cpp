// sg_api.cpp -- the C-ABI of libsplinegrids_hip.so (include/splinegrids_hip.h).
//
// Argument checks here return SG_ERR_INVALID with a message; the
// reference's own Julia validation (src/validation.jl) stays on the caller's
// side of the boundary so its messages and log lines are unchanged.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sstream>

#include "sg_plan_host.hpp"

namespace sg {

namespace {
thread_local std::string g_error;

struct DeviceGuard {
  int prev = -1;
  hipError_t err = hipSuccess;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) err = hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

template <typename T>
hipError_t dalloc(Plan& pl, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, count * sizeof(T));
  if (e != hipSuccess) return e;
  pl.allocations.push_back(v);
  pl.device_bytes += count * sizeof(T);
  *p = static_cast<T*>(v);
  return hipSuccess;
}

void free_plan(Plan* pl) {
  if (!pl) return;
  for (void* v : pl->allocations) (void)hipFree(v);
  delete pl;
}

int checked(hipError_t e, const char* what) {
  return e == hipSuccess ? SG_OK : hip_fail(e, what);
}

}  // namespace

void set_error(const std::string& msg) { g_error = msg; }

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  std::ostringstream os;
  os << what << ": " << hipGetErrorName(e) << " (" << hipGetErrorString(e) << ")";
  g_error = os.str();
  return e == hipErrorOutOfMemory ? SG_ERR_NO_MEMORY : SG_ERR_HIP;
}

}  // namespace sg

using namespace sg;

struct sg_plan {
  Plan p;
};

extern "C" {

const char* sg_last_error(void) { return g_error.c_str(); }

int sg_version(int32_t* major, int32_t* minor, int32_t* patch) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  if (patch) *patch = 0;
  return SG_OK;
}

int sg_stream_synchronize(void* stream) {
  return checked(hipStreamSynchronize(static_cast<hipStream_t>(stream)),
                 "hipStreamSynchronize");
}

int sg_dimension_evaluate(const void* knots_all, int32_t n_knots,
                          const void* sample_points, int32_t n_samples,
                          int32_t degree, int32_t max_derivative_order,
                          int32_t* sample_indices, void* eval, int32_t dtype,
                          void* stream) {
  if (degree < 0) return fail(SG_ERR_INVALID, "degree must be >= 0");
  if (max_derivative_order < 0 || max_derivative_order > degree)
    return fail(SG_ERR_INVALID,
                "The max_degree must be positive and derivatives order higher "
                "than `degree` are all 0.");
  if (n_knots < 2 * degree + 2)
    return fail(SG_ERR_INVALID, "n_knots must be at least 2 * degree + 2 (n_basis > degree)");
  if (n_samples < 0) return fail(SG_ERR_INVALID, "n_samples must be >= 0");
  if (dtype != SG_F32 && dtype != SG_F64) return fail(SG_ERR_INVALID, "dtype must be SG_F32 or SG_F64");
  if (n_samples == 0) return SG_OK;
  if (!knots_all || !sample_points || !sample_indices || !eval)
    return fail(SG_ERR_INVALID, "null device pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t count = static_cast<size_t>(n_samples) * (degree + 1) * (max_derivative_order + 1);
  double* scratch = nullptr;
  double* table = nullptr;
  hipError_t e = hipMalloc(&scratch, count * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  if (dtype == SG_F32) {
    e = hipMalloc(&table, count * sizeof(double));
    if (e != hipSuccess) {
      (void)hipFree(scratch);
      return hip_fail(e, "hipMalloc");
    }
    e = launch_basis<float>(nullptr, static_cast<const float*>(knots_all), n_knots,
                            static_cast<const float*>(sample_points), n_samples, degree,
                            max_derivative_order, table, scratch, nullptr, sample_indices, s);
    if (e == hipSuccess) e = launch_cast_table(table, count, static_cast<float*>(eval), s);
  } else {
    e = launch_basis<double>(nullptr, static_cast<const double*>(knots_all), n_knots,
                             static_cast<const double*>(sample_points), n_samples, degree,
                             max_derivative_order, static_cast<double*>(eval), scratch,
                             nullptr, sample_indices, s);
  }
  hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(scratch);
  if (table) (void)hipFree(table);
  if (e != hipSuccess) return hip_fail(e, "basis kernel launch");
  return checked(e2, "basis kernel");
}

}  // extern "C"

namespace sg {

int validate_desc(const sg_grid_desc* desc) {
  if (!desc || !desc->dims) return fail(SG_ERR_INVALID, "grid descriptor is NULL");
  if (desc->n_in < 1 || desc->n_in > SG_MAX_NIN)
    return fail(SG_ERR_INVALID, "n_in must be in 1..8");
  if (desc->n_out < 1) return fail(SG_ERR_INVALID, "n_out must be >= 1");
  for (int d = 0; d < desc->n_in; ++d) {
    const sg_dim_desc& dd = desc->dims[d];
    std::ostringstream os;
    os << "dimension " << d + 1 << ": ";
    if (dd.degree < 0) return fail(SG_ERR_INVALID, os.str() + "degree must be >= 0");
    if (dd.n_basis < dd.degree + 1)
      return fail(SG_ERR_INVALID, os.str() + "n_basis must be >= degree + 1");
    if (dd.n_samples < 0) return fail(SG_ERR_INVALID, os.str() + "n_samples must be >= 0");
    if (dd.max_derivative_order < 0 || dd.max_derivative_order > dd.degree)
      return fail(SG_ERR_INVALID, os.str() +
                  "The max_degree must be positive and derivatives order higher than "
                  "`degree` are all 0.");
    if (!dd.knots_all) return fail(SG_ERR_INVALID, os.str() + "knots_all is NULL");
    if (dd.n_samples > 0 && !dd.sample_points)
      return fail(SG_ERR_INVALID, os.str() + "sample_points is NULL");
    const int nk = dd.n_basis + dd.degree + 1;
    for (int k = 1; k < nk; ++k)
      if (!(dd.knots_all[k - 1] <= dd.knots_all[k]))
        return fail(SG_ERR_INVALID, os.str() + "knots_all must be sorted");
  }
  return SG_OK;
}

// Uploads the gather lists of tiling tl (partial windows w1 x w2 x w3max per
// channel): the lists adj_gather walks.
static hipError_t upload_gather_lists(Plan& pl, Tiling& tl, int w1, int w2) {
  GatherLists g = gather_lists(pl, tl, w1, w2);
  std::vector<int32_t>* ptr = g.ptr;
  std::vector<long long>* off = g.off;
  tl.gchan = g.gchan;
  tl.gfast = g.gfast;
  hipError_t e = hipSuccess;
  for (int d = 0; d < 3 && e == hipSuccess; ++d) {
    e = dalloc(pl, &tl.gptr[d], ptr[d].size());
    if (e == hipSuccess) e = dalloc(pl, &tl.goff[d], off[d].size());
    if (e == hipSuccess)
      e = hipMemcpy(tl.gptr[d], ptr[d].data(), ptr[d].size() * sizeof(int32_t),
                    hipMemcpyHostToDevice);
    if (e == hipSuccess && !off[d].empty())
      e = hipMemcpy(tl.goff[d], off[d].data(), off[d].size() * sizeof(long long),
                    hipMemcpyHostToDevice);
  }
  return e;
}

// Build a plan for a validated descriptor on the current device.  On success
// *result owns every allocation; on failure everything is freed.
int build_plan(const sg_grid_desc* desc, hipStream_t s, bool want_adjoint, Plan** result) {
  *result = nullptr;
  std::unique_ptr<Plan, void (*)(Plan*)> pl(new Plan(), free_plan);
  pl->nin = desc->n_in;
  pl->nout = desc->n_out;
  pl->nurbs = desc->is_nurbs ? 1 : 0;
  pl->device = desc->device;
  pl->dims.resize(desc->n_in);
  std::vector<double*> scratch(desc->n_in, nullptr);
  int rc = SG_OK;
  for (int d = 0; d < desc->n_in && rc == SG_OK; ++d) {
    const sg_dim_desc& dd = desc->dims[d];
    DimDev& D = pl->dims[d];
    D.p = dd.degree;
    D.n = dd.n_basis;
    D.m = dd.n_samples;
    D.D = dd.max_derivative_order;
    const int nk = D.n + D.p + 1;
    const size_t tcount = static_cast<size_t>(D.m) * (D.p + 1) * (D.D + 1);
    hipError_t e = dalloc(*pl, &D.knots, nk);
    if (e == hipSuccess) e = dalloc(*pl, &D.t, D.m);
    if (e == hipSuccess) e = dalloc(*pl, &D.span, D.m);
    if (e == hipSuccess) e = dalloc(*pl, &D.table64, tcount);
    if (e == hipSuccess) e = dalloc(*pl, &D.ktab64, tcount);
    if (e == hipSuccess) e = dalloc(*pl, &D.ktab32, tcount);
    if (e == hipSuccess) e = dalloc(*pl, &scratch[d], tcount);
    if (e == hipSuccess)
      e = hipMemcpyAsync(D.knots, dd.knots_all, nk * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && D.m > 0)
      e = hipMemcpyAsync(D.t, dd.sample_points, D.m * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = launch_basis<double>(D.knots, D.knots, nk, D.t, D.m, D.p, D.D, D.table64,
                               scratch[d], D.span, nullptr, s);
    if (e == hipSuccess) e = launch_kernel_tables(D.table64, D.m, D.p, D.D, D.ktab64, D.ktab32, s);
    if (e != hipSuccess) rc = hip_fail(e, "plan set-up");
  }
  if (rc != SG_OK) return rc;
  // spans back to the host for the tiling and the transposed band lists
  for (int d = 0; d < desc->n_in && rc == SG_OK; ++d) {
    DimDev& D = pl->dims[d];
    D.h_span.resize(D.m);
    if (D.m > 0) {
      hipError_t e = hipMemcpyAsync(D.h_span.data(), D.span, D.m * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) rc = hip_fail(e, "span copy");
    }
  }
  if (rc != SG_OK) return rc;
  if ((rc = checked(hipStreamSynchronize(s), "plan set-up")) != SG_OK) return rc;

  for (int d = 0; d < desc->n_in && rc == SG_OK; ++d) {
    DimDev& D = pl->dims[d];
    D.monotone = std::is_sorted(D.h_span.begin(), D.h_span.end());
    // power-of-two scale of each derivative order's table (MFMA tiers)
    const size_t per_order = static_cast<size_t>(D.m) * (D.p + 1);
    std::vector<double> tab(per_order * (D.D + 1));
    if (!tab.empty()) {
      hipError_t e = hipMemcpy(tab.data(), D.table64, tab.size() * sizeof(double),
                               hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_fail(e, "table copy");
    }
    D.tab_exp.assign(D.D + 1, 0);
    for (int o = 0; o <= D.D; ++o) {
      double mx = 0.0;
      for (size_t i = 0; i < per_order; ++i) mx = std::max(mx, std::fabs(tab[o * per_order + i]));
      int e = 0;
      if (mx > 0.0) std::frexp(mx, &e);
      D.tab_exp[o] = e;
    }
    // transposed band lists: samples whose band holds k, in sample order
    std::vector<int32_t> ptr, js;
    band_lists(D, ptr, js);
    hipError_t e = dalloc(*pl, &D.csr_ptr, ptr.size());
    if (e == hipSuccess) e = dalloc(*pl, &D.csr_js, js.size());
    if (e == hipSuccess)
      e = hipMemcpy(D.csr_ptr, ptr.data(), ptr.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && !js.empty())
      e = hipMemcpy(D.csr_js, js.data(), js.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "band list upload");
  }
  if (rc != SG_OK) return rc;

  plan_tilings(*pl, want_adjoint, adjoint_mma_shape_ok(*pl), adjoint_mma_resident_waves);
  const bool ok = pl->stream_ok;
  if (ok && (pl->adj.n_tiles1 > 0 || pl->adjm.n_tiles1 > 0)) {
    hipError_t e = hipSuccess;
    if (pl->adj.n_tiles1 > 0) e = upload_gather_lists(*pl, pl->adj, pl->adj.w1max, pl->adj.w2max);
    if (e == hipSuccess && pl->adjm.n_tiles1 > 0)
      e = upload_gather_lists(*pl, pl->adjm, kAdjMmaWin, kAdjMmaWin);
    if (e == hipSuccess && pl->adjm.n_tiles1 > 0 && pl->nin == 2) {
      const float one = 1.f;
      const int32_t zero = 0;
      e = dalloc(*pl, &pl->unit_tab, 1);
      if (e == hipSuccess) e = dalloc(*pl, &pl->unit_span, 1);
      if (e == hipSuccess) e = hipMemcpy(pl->unit_tab, &one, sizeof(one), hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(pl->unit_span, &zero, sizeof(zero), hipMemcpyHostToDevice);
    }
    pl->work_bytes = std::max(adjoint_work_bytes(*pl), adjoint_mma_work_bytes(*pl));
    unsigned char* w = nullptr;
    if (e == hipSuccess) e = dalloc(*pl, &w, pl->work_bytes);
    pl->work = w;
    if (e != hipSuccess) return hip_fail(e, "tiling upload");
  }
  // the basis scratch is only needed during set-up
  for (double* sc : scratch) {
    auto it = std::find(pl->allocations.begin(), pl->allocations.end(), static_cast<void*>(sc));
    if (sc && it != pl->allocations.end()) {
      (void)hipFree(sc);
      pl->allocations.erase(it);
    }
  }
  size_t bytes = 0;
  for (int d = 0; d < pl->nin; ++d) {
    const DimDev& D = pl->dims[d];
    bytes += static_cast<size_t>(D.m) * (D.p + 1) * (D.D + 1) * sizeof(double);
  }
  pl->device_bytes -= bytes;
  *result = pl.release();
  return SG_OK;
}

void destroy_plan(Plan* pl) { free_plan(pl); }

}  // namespace sg

extern "C" {

int sg_plan_create(const sg_grid_desc* desc, void* stream, sg_plan** out) {
  if (!out) return fail(SG_ERR_INVALID, "plan output pointer is NULL");
  *out = nullptr;
  int rc = validate_desc(desc);
  if (rc != SG_OK) return rc;
  DeviceGuard guard(desc->device);
  if (guard.err != hipSuccess) return hip_fail(guard.err, "hipSetDevice");
  Plan* p = nullptr;
  rc = build_plan(desc, static_cast<hipStream_t>(stream), true, &p);
  if (rc != SG_OK) return rc;
  sg_plan* h = new sg_plan();
  h->p = std::move(*p);
  p->allocations.clear();
  delete p;
  *out = h;
  return SG_OK;
}

int sg_plan_destroy(sg_plan* plan) {
  if (!plan) return SG_OK;
  DeviceGuard guard(plan->p.device);
  for (void* v : plan->p.allocations) (void)hipFree(v);
  delete plan;
  return SG_OK;
}

int sg_plan_basis(const sg_plan* plan, int32_t dim, double* table, int32_t* sample_indices) {
  if (!plan) return fail(SG_ERR_INVALID, "plan is NULL");
  if (dim < 0 || dim >= plan->p.nin) return fail(SG_ERR_INVALID, "dimension out of range");
  const DimDev& D = plan->p.dims[dim];
  DeviceGuard guard(plan->p.device);
  if (table && D.m > 0) {
    const size_t count = static_cast<size_t>(D.m) * (D.p + 1) * (D.D + 1);
    hipError_t e = hipMemcpy(table, D.table64, count * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "table copy");
  }
  if (sample_indices)
    for (int j = 0; j < D.m; ++j) sample_indices[j] = D.h_span[j] + 1;
  return SG_OK;
}

int sg_plan_device_bytes(const sg_plan* plan, int64_t* bytes) {
  if (!plan || !bytes) return fail(SG_ERR_INVALID, "NULL argument");
  *bytes = static_cast<int64_t>(plan->p.device_bytes);
  return SG_OK;
}

int sg_plan_describe(const sg_plan* plan, char* text, int32_t capacity) {
  if (!plan || !text || capacity < 1) return fail(SG_ERR_INVALID, "NULL argument");
  const Plan& pl = plan->p;
  std::ostringstream os;
  auto tiling = [&](const char* name, const Tiling& t) {
    os << name << "=";
    if (t.n_tiles1 == 0) {
      os << "none ";
      return;
    }
    os << t.tile_width() << "x" << t.chunk_rows() << "x" << t.j3_per_block << ",tiles="
       << t.n_tiles1 << "x" << t.n_chunks2 << "x" << t.n_z << ",win=" << t.w1max << "x"
       << t.w2max << "x" << t.w3max << " ";
  };
  os << "forward=" << (pl.stream_ok ? "stream" : "direct") << " ";
  tiling("fwd", pl.fwd);
  tiling("mma", pl.mma);
  const char* slow = (pl.stream_ok && pl.adj.n_tiles1 > 0) ? "stream" : "csr";
  os << "adjoint=" << (pl.adjm.n_tiles1 > 0 ? "mma_f32+" : "") << slow << " ";
  tiling("adj", pl.adj);
  tiling("adjm", pl.adjm);
  os << "work_bytes=" << pl.work_bytes;
  const std::string s = os.str();
  const size_t n = std::min(s.size(), static_cast<size_t>(capacity - 1));
  std::memcpy(text, s.data(), n);
  text[n] = '\0';
  return SG_OK;
}

static int check_orders(const Plan& pl, int nf, const int32_t* orders) {
  if (!orders) return fail(SG_ERR_INVALID, "derivative_order is NULL");
  for (int f = 0; f < nf; ++f)
    for (int d = 0; d < pl.nin; ++d) {
      const int o = orders[f * pl.nin + d];
      if (o < 0 || o > pl.dims[d].D) {
        std::ostringstream os;
        os << "The maximum derivative order available for spline dimension " << d + 1
           << " is " << pl.dims[d].D << ", got " << o << ".";
        return fail(SG_ERR_INVALID, os.str());
      }
      if (pl.nurbs && o != 0)
        return fail(SG_ERR_INVALID, "Computing derivatives of NURBS is currently not supported.");
    }
  return SG_OK;
}

static long long n_samples(const Plan& pl) {
  long long n = 1;
  for (const DimDev& d : pl.dims) n *= d.m;
  return n;
}

static long long n_control(const Plan& pl) {
  long long n = 1;
  for (const DimDev& d : pl.dims) n *= d.n;
  return n;
}

static int check_common(const sg_plan* plan, int32_t dtype, int32_t tier) {
  if (!plan) return fail(SG_ERR_INVALID, "plan is NULL");
  if (dtype != SG_F32 && dtype != SG_F64) return fail(SG_ERR_INVALID, "dtype must be SG_F32 or SG_F64");
  if (tier != SG_TIER_NATIVE && tier != SG_TIER_SPLIT_F16 && tier != SG_TIER_BF16)
    return fail(SG_ERR_INVALID, "unknown tier");
  if (tier != SG_TIER_NATIVE) {
    if (dtype != SG_F32)
      return fail(SG_ERR_UNSUPPORTED, "the split-FP16 and BF16 tiers take FP32 arrays");
    if (!mma_supported(plan->p))
      return fail(SG_ERR_UNSUPPORTED,
                  "the split-FP16 and BF16 tiers need Nin 2-3, monotone sample points and "
                  "degrees <= 5 in dimensions 1-2");
  }
  return SG_OK;
}

static hipError_t forward(const Plan& pl, int tier, int nf, const int32_t* orders,
                          const void* P, const void* W, void* const* outs, int dtype,
                          hipStream_t s) {
  if (tier == SG_TIER_NATIVE) return launch_forward(pl, nf, orders, P, W, outs, dtype, s);
  if (n_samples(pl) == 0) return hipSuccess;
  return launch_forward_mma(pl, tier, nf, orders, static_cast<const float*>(P),
                            static_cast<const float*>(W), reinterpret_cast<float* const*>(outs),
                            s);
}

int sg_evaluate_fields(sg_plan* plan, int32_t n_fields, const int32_t* derivative_orders,
                       const void* control_points, void* const* evals, int32_t dtype,
                       int32_t tier, void* stream);

int sg_evaluate(sg_plan* plan, const int32_t* derivative_order, const void* control_points,
                const void* weights, void* eval, int32_t dtype, int32_t tier, void* stream) {
  int rc = check_common(plan, dtype, tier);
  if (rc != SG_OK) return rc;
  const Plan& pl = plan->p;
  if ((rc = check_orders(pl, 1, derivative_order)) != SG_OK) return rc;
  if (pl.nurbs && !weights) return fail(SG_ERR_INVALID, "NURBS plan needs weights");
  if (!pl.nurbs && weights) return fail(SG_ERR_INVALID, "B-spline plan takes no weights");
  if (n_samples(pl) == 0) return SG_OK;  // empty arrays may have NULL data
  if (!control_points || !eval) return fail(SG_ERR_INVALID, "null device pointer");
  DeviceGuard guard(pl.device);
  if (guard.err != hipSuccess) return hip_fail(guard.err, "hipSetDevice");
  void* outs[1] = {eval};
  return checked(forward(pl, tier, 1, derivative_order, control_points, weights, outs, dtype,
                         static_cast<hipStream_t>(stream)),
                 "forward launch");
}

int sg_evaluate_fields(sg_plan* plan, int32_t n_fields, const int32_t* derivative_orders,
                       const void* control_points, void* const* evals, int32_t dtype,
                       int32_t tier, void* stream) {
  int rc = check_common(plan, dtype, tier);
  if (rc != SG_OK) return rc;
  const Plan& pl = plan->p;
  if (pl.nurbs) return fail(SG_ERR_INVALID, "sg_evaluate_fields takes B-spline plans");
  if (n_fields < 1 || n_fields > SG_MAX_FIELDS)
    return fail(SG_ERR_INVALID, "n_fields must be in 1..4");
  if ((rc = check_orders(pl, n_fields, derivative_orders)) != SG_OK) return rc;
  if (n_samples(pl) == 0) return SG_OK;
  if (!control_points || !evals) return fail(SG_ERR_INVALID, "null device pointer");
  for (int f = 0; f < n_fields; ++f)
    if (!evals[f]) return fail(SG_ERR_INVALID, "null eval pointer");
  DeviceGuard guard(pl.device);
  if (guard.err != hipSuccess) return hip_fail(guard.err, "hipSetDevice");
  return checked(forward(pl, tier, n_fields, derivative_orders, control_points, nullptr, evals,
                         dtype, static_cast<hipStream_t>(stream)),
                 "forward launch");
}

// Orders of rational derivative fields: 0 <= o <= D per dimension.
static int check_rational(const sg_plan* plan, int32_t n_fields, const int32_t* orders,
                          int32_t dtype) {
  if (!plan) return fail(SG_ERR_INVALID, "plan is NULL");
  if (dtype != SG_F32 && dtype != SG_F64) return fail(SG_ERR_INVALID, "dtype must be SG_F32 or SG_F64");
  const Plan& pl = plan->p;
  if (!pl.nurbs) return fail(SG_ERR_INVALID, "rational derivatives need a NURBS plan");
  if (n_fields < 1 || n_fields > SG_MAX_FIELDS)
    return fail(SG_ERR_INVALID, "n_fields must be 1.." + std::to_string(SG_MAX_FIELDS));
  if (!orders) return fail(SG_ERR_INVALID, "derivative_orders is NULL");
  for (int f = 0; f < n_fields; ++f)
    for (int d = 0; d < pl.nin; ++d) {
      const int o = orders[f * pl.nin + d];
      if (o < 0 || o > pl.dims[d].D) {
        std::ostringstream os;
        os << "The maximum derivative order available for spline dimension " << d + 1
           << " is " << pl.dims[d].D << ", got " << o << ".";
        return fail(SG_ERR_INVALID, os.str());
      }
    }
  if (!rational_closure_ok(pl, n_fields, orders))
    return fail(SG_ERR_UNSUPPORTED, "too many derivative orders below the requested ones "
                                    "(at most 16)");
  return SG_OK;
}

int sg_rational_workspace_bytes(const sg_plan* plan, int32_t n_fields,
                                const int32_t* derivative_orders, int32_t dtype,
                                int64_t* bytes) {
  int rc = check_rational(plan, n_fields, derivative_orders, dtype);
  if (rc != SG_OK) return rc;
  if (!bytes) return fail(SG_ERR_INVALID, "bytes is NULL");
  *bytes = static_cast<int64_t>(rational_workspace_bytes(
      plan->p, n_fields, derivative_orders, dtype == SG_F32 ? sizeof(float) : sizeof(double)));
  return SG_OK;
}

int sg_evaluate_rational_fields(sg_plan* plan, int32_t n_fields, const int32_t* derivative_orders,
                                const void* control_points, const void* weights,
                                void* const* evals, void* workspace, int64_t workspace_bytes,
                                int32_t dtype, void* stream) {
  int rc = check_rational(plan, n_fields, derivative_orders, dtype);
  if (rc != SG_OK) return rc;
  const Plan& pl = plan->p;
  if (n_samples(pl) == 0) return SG_OK;
  if (!control_points || !weights || !evals || !workspace)
    return fail(SG_ERR_INVALID, "null device pointer");
  for (int f = 0; f < n_fields; ++f)
    if (!evals[f]) return fail(SG_ERR_INVALID, "null device pointer");
  const size_t need = rational_workspace_bytes(pl, n_fields, derivative_orders,
                                               dtype == SG_F32 ? sizeof(float) : sizeof(double));
  if (workspace_bytes < 0 || static_cast<size_t>(workspace_bytes) < need)
    return fail(SG_ERR_INVALID, "workspace too small: " + std::to_string(need) + " bytes needed");
  DeviceGuard guard(pl.device);
  if (guard.err != hipSuccess) return hip_fail(guard.err, "hipSetDevice");
  return checked(launch_rational(pl, n_fields, derivative_orders, control_points, weights, evals,
                                 workspace, dtype, static_cast<hipStream_t>(stream)),
                 "rational launch");
}

int sg_rational_adjoint_workspace_bytes(const sg_plan* plan, int32_t dtype, int64_t* bytes) {
  if (!plan) return fail(SG_ERR_INVALID, "plan is NULL");
  if (dtype != SG_F32 && dtype != SG_F64) return fail(SG_ERR_INVALID, "dtype must be SG_F32 or SG_F64");
  if (!plan->p.nurbs) return fail(SG_ERR_INVALID, "the rational adjoint needs a NURBS plan");
  if (!bytes) return fail(SG_ERR_INVALID, "bytes is NULL");
  *bytes = static_cast<int64_t>(
      rational_adjoint_workspace_bytes(plan->p, dtype == SG_F32 ? sizeof(float) : sizeof(double)));
  return SG_OK;
}

int sg_evaluate_rational_adjoint(sg_plan* plan, const void* weights, const void* eval,
                                 void* control_points, void* workspace, int64_t workspace_bytes,
                                 int32_t dtype, int32_t accumulate, void* stream) {
  int64_t need = 0;
  int rc = sg_rational_adjoint_workspace_bytes(plan, dtype, &need);
  if (rc != SG_OK) return rc;
  const Plan& pl = plan->p;
  if (n_control(pl) == 0) return SG_OK;
  if (!weights || !control_points || !workspace || (!eval && n_samples(pl) > 0))
    return fail(SG_ERR_INVALID, "null device pointer");
  if (workspace_bytes < need)
    return fail(SG_ERR_INVALID, "workspace too small: " + std::to_string(need) + " bytes needed");
  DeviceGuard guard(pl.device);
  if (guard.err != hipSuccess) return hip_fail(guard.err, "hipSetDevice");
  return checked(launch_rational_adjoint(pl, weights, eval, control_points, workspace, dtype,
                                         accumulate ? 1 : 0, static_cast<hipStream_t>(stream)),
                 "rational adjoint launch");
}

int sg_evaluate_adjoint(sg_plan* plan, const int32_t* derivative_order, const void* eval,
                        void* control_points, int32_t dtype, int32_t tier, int32_t accumulate,
                        void* stream) {
  int rc = check_common(plan, dtype, tier);
  if (rc != SG_OK) return rc;
  const Plan& pl = plan->p;
  if (pl.nurbs) return fail(SG_ERR_INVALID, "Adjoint evaluation not supported for NURBS.");
  if (tier != SG_TIER_NATIVE && pl.adjm.n_tiles1 == 0)
    return fail(SG_ERR_UNSUPPORTED,
                "the split-FP16 and BF16 adjoint tiers run on the matrix-core adjoint, which "
                "needs band windows of at most 16 control points per 64 samples of dimension 1 "
                "and per 16..64 samples of dimension 2 (and SG_ADJ_MMA not 0)");
  if ((rc = check_orders(pl, 1, derivative_order)) != SG_OK) return rc;
  if (n_control(pl) == 0) return SG_OK;
  if (!control_points || (!eval && n_samples(pl) > 0))
    return fail(SG_ERR_INVALID, "null device pointer");
  DeviceGuard guard(pl.device);
  if (guard.err != hipSuccess) return hip_fail(guard.err, "hipSetDevice");
  if (tier != SG_TIER_NATIVE && n_samples(pl) > 0)
    return checked(launch_adjoint_mma(pl, tier, derivative_order,
                                      static_cast<const float*>(eval),
                                      static_cast<float*>(control_points), accumulate ? 1 : 0,
                                      static_cast<hipStream_t>(stream)),
                   "adjoint launch");
  return checked(launch_adjoint(pl, derivative_order, eval, control_points, dtype,
                                accumulate ? 1 : 0, static_cast<hipStream_t>(stream)),
                 "adjoint launch");
}

}  // extern "C"
