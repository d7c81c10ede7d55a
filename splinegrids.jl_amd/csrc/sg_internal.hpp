// sg_internal.hpp -- plan structures and kernel launchers shared by the
// translation units of libsplinegrids_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/splinegrids_hip.h"

namespace sg {

// One wavefront of lanes covers this many consecutive samples of dimension 1
// (the fastest axis of eval), so every store of a row writes 256 contiguous
// bytes in FP32.
constexpr int kLanes = 64;
constexpr int kWaves = 4;               // waves per workgroup
constexpr int kThreads = kLanes * kWaves;
constexpr int kMaxStreamDegree = 5;     // streaming kernels: p1, p2 <= 5
constexpr int kMaxLds = 64 * 1024;      // LDS bytes one workgroup may claim

struct DimDev {
  int p = 0, n = 0, m = 0, D = 0;
  double* knots = nullptr;    // [n + p + 1]
  double* t = nullptr;        // [m]
  int32_t* span = nullptr;    // [m] 0-based span s: band is k = s-p .. s
  double* table64 = nullptr;  // [m, p+1, D+1] reference layout
  double* ktab64 = nullptr;   // [D+1][m][p+1]: the p+1 values of a sample
  float* ktab32 = nullptr;    //   are contiguous (kernel layout)
  std::vector<int32_t> h_span;
  bool monotone = true;       // spans non-decreasing in the sample index
  // For every basis function k: first and last sample whose band holds k
  // (last < first if none); only meaningful when monotone.
  int32_t* supp = nullptr;    // [n][2]
  // Transposed band lists for the general adjoint: samples j whose band
  // holds k, with their slot a = k - (s_j - p).
  int32_t* csr_ptr = nullptr; // [n + 1]
  int32_t* csr_js = nullptr;  // [m * (p+1)]: j * (p + 1) + a
};

struct Plan {
  int nin = 0, nout = 0, nurbs = 0, device = 0;
  std::vector<DimDev> dims;
  bool stream_ok = false;     // streaming (sum-factorised) kernels apply
  int rows_per_wave = 32;     // rows of dimension 2 per wave
  int n_tiles1 = 0, n_chunks2 = 0;
  int w1max = 0, w2max = 0;   // largest band windows of a tile / chunk
  int w2wave_max = 0;         // largest window of one wave's rows
  // adjoint: tile / chunk ranges holding each k1 / k2
  int32_t* tile_of_k1 = nullptr;   // [n1][2]
  int32_t* chunk_of_k2 = nullptr;  // [n2][2]
  void* work = nullptr;            // adjoint partial slabs
  size_t work_bytes = 0;
  size_t device_bytes = 0;
  std::vector<void*> allocations;
};

// ---- error reporting (sg_api.cpp) ----
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// ---- sg_basis.hip ----
// Span search + Cox-de Boor with derivatives (FP64).  Writes table64 in the
// reference layout; scratch has the same size.  span0 gets 0-based spans,
// span1 (optional) 1-based ones.
template <typename TIn>
hipError_t launch_basis(const double* knots64, const TIn* knots_in, int n_knots,
                        const TIn* t_in, int m, int p, int D, double* table64,
                        double* scratch, int32_t* span0, int32_t* span1,
                        hipStream_t s);
hipError_t launch_kernel_tables(const double* table64, int m, int p, int D,
                                double* ktab64, float* ktab32, hipStream_t s);
hipError_t launch_cast_table(const double* table64, size_t count, float* out,
                             hipStream_t s);
hipError_t launch_copy_f64(const double* in, size_t count, double* out,
                           hipStream_t s);

// ---- sg_forward.hip ----
// Fields: derivative orders [nf][nin]; one output pointer per field.
hipError_t launch_forward(const Plan& plan, int nf, const int32_t* orders,
                          const void* P, const void* W, void* const* outs,
                          int dtype, hipStream_t s);

// ---- sg_adjoint.hip ----
hipError_t launch_adjoint(const Plan& plan, const int32_t* orders,
                          const void* ebar, void* Pbar, int dtype,
                          int accumulate, hipStream_t s);
size_t adjoint_work_bytes(const Plan& plan);

}  // namespace sg
