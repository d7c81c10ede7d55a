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
constexpr int kAdjMmaWin = 16;         // matrix-core adjoint: window edge (one MFMA tile)

// Workgroup copy of n >= 1 elements of NF sources into LDS:
// store(f, i, load(f, i)) for f < NF, i < n.  Each thread issues all NF * U
// loads of a pass before its first LDS write, so staging a tile costs one
// memory latency per pass instead of one per element (a plain loop waits for
// each load before its write).  Loads past n re-read element n - 1.
template <int NF, int U, typename Load, typename Store>
__device__ inline void stage_to_lds(int n, Load load, Store store) {
  for (int i0 = threadIdx.x; i0 < n; i0 += U * kThreads) {
    decltype(load(0, 0)) v[NF][U];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int u = 0; u < U; ++u) v[f][u] = load(f, min(i0 + u * kThreads, n - 1));
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + u * kThreads < n) store(f, i0 + u * kThreads, v[f][u]);
  }
}

// A pointer into global memory, typed so.  Pointers read from memory (not
// kernel arguments) are otherwise generic, and generic (FLAT) accesses count
// on the LDS counter too, so every LDS wait would also wait for them.
template <typename T>
using Global = __attribute__((address_space(1))) T;
template <typename T>
__device__ inline Global<T>* global_ptr(T* p) {
  return (Global<T>*)p;
}

// A pointer into memory that no kernel writes while it runs, typed as the
// constant address space: loads at wave-uniform addresses through it become
// scalar loads (s_load), which the compiler cannot prove safe for a plain
// pointer read from memory.
template <typename T>
using Constant = __attribute__((address_space(4))) T;
template <typename T>
__device__ inline Constant<T>* constant_ptr(T* p) {
  return (Constant<T>*)p;
}

// 1/x for the NURBS quotient: FP32 takes the hardware reciprocal (at most
// 1 ulp from the correctly rounded value; 1/0 = inf, so 0/0 stays NaN), FP64
// one IEEE division.  Each sample then multiplies its channels by it.
__device__ inline float nurbs_recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double nurbs_recip(double x) { return 1.0 / x; }

// floor(n / d) for n * d < 2^32 from r = udiv_rcp(d): with r = ceil(2^32 / d)
// = (2^32 + e) / d (0 <= e < d), n r / 2^32 = n / d + n e / (d 2^32), and
// n e / 2^32 < 1 / d keeps the floor exact.
__host__ __device__ inline unsigned udiv_rcp(unsigned d) { return d > 1 ? 0xFFFFFFFFu / d + 1 : 0; }
__device__ inline int udiv_fast(int n, int d, unsigned r) {
  return d > 1 ? static_cast<int>(__umulhi(static_cast<unsigned>(n), r)) : n;
}

// A value that every lane of the wave holds (read from one LDS address),
// moved to a scalar register: branches on it stay scalar branches and FMAs
// take it as a scalar operand.
__device__ inline int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float wave_uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ inline double wave_uniform(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(b)));
  const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | lo);
}

struct DimDev {
  int p = 0, n = 0, m = 0, D = 0;
  double* knots = nullptr;    // [n + p + 1]
  double* t = nullptr;        // [m]
  int32_t* span = nullptr;    // [m] 0-based span s: band is k = s-p .. s
  double* table64 = nullptr;  // [m, p+1, D+1] reference layout
  double* ktab64 = nullptr;   // [D+1][m][p+1]: the p+1 values of a sample
  float* ktab32 = nullptr;    //   are contiguous (kernel layout)
  std::vector<int32_t> h_span;
  std::vector<int> tab_exp;   // per order: max|B| * 2^-e in [0.5, 1)
  bool monotone = true;       // spans non-decreasing in the sample index
  // Transposed band lists for the general adjoint: samples j whose band
  // holds k, with their slot a = k - (s_j - p).
  int32_t* csr_ptr = nullptr; // [n + 1]
  int32_t* csr_js = nullptr;  // [m * (p+1)]: j * (p + 1) + a
};

// Tiling of the sample grid for the streaming kernels: tiles of kLanes
// samples along dimension 1, chunks of kWaves * rows_per_wave rows along
// dimension 2, one workgroup per (tile, chunk, sample of dimension 3).
struct Tiling {
  int rows_per_wave = 32;
  int wx = 1;                 // waves side by side along dimension 1 (1, 2, 4)
  int n_tiles1 = 0, n_chunks2 = 0;
  int w1max = 0, w2max = 0;   // largest band windows of a tile / chunk
  int w2wave_max = 0;         // largest window of one wave's rows
  int xcd_map = 0;            // workgroup numbering, see xcd_logical_block
  // adjoint only: samples of dimension 3 per workgroup and their k3 windows
  int j3_per_block = 1, n_z = 1, w3max = 1;
  // adjoint gather: for dimension d and control point k, the windows that
  // hold k, [gptr[d][k], gptr[d][k + 1]) in goff[d], each as the part of the
  // element offset in the partial buffer that dimension d contributes (the
  // offsets of the three dimensions and the channel add up)
  int32_t* gptr[3] = {nullptr, nullptr, nullptr};  // [n_d + 1]
  long long* goff[3] = {nullptr, nullptr, nullptr};
  long long gchan = 0;  // element offset between the channels of a window
  int gfast = 0;        // window counts small enough for udiv_fast in the gather
  // explicit tile width / chunk height (0: derived from wx, rows_per_wave)
  int tile_w = 0, chunk_h = 0;
  int tile_width() const { return tile_w ? tile_w : kLanes * wx; }
  int chunk_rows() const { return chunk_h ? chunk_h : rows_per_wave * (kWaves / wx); }
};

struct Plan {
  int nin = 0, nout = 0, nurbs = 0, device = 0;
  std::vector<DimDev> dims;
  bool stream_ok = false;     // streaming (sum-factorised) kernels apply
  Tiling fwd, adj, mma;
  Tiling adjm;                // FP32 matrix-core adjoint (n_tiles1 == 0: not used)
  float* unit_tab = nullptr;  // {1.0f} and {0}: dimension 3 of a 2-D grid for
  int32_t* unit_span = nullptr;  //   the matrix-core adjoint
  int nt_stores = 0;          // non-temporal stores of eval
  void* work = nullptr;       // adjoint partial slabs (either adjoint path)
  size_t work_bytes = 0;
  size_t device_bytes = 0;
  std::vector<void*> allocations;
};

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  When rows of eval are not 128-byte aligned, neighbouring
// tiles of a row share the cache line at their edge; numbering the 1-D grid
// so that such tiles run on one XCD lets that line be fetched (or merged)
// once.  Modes (each a bijection of [0, gridDim.x)):
//   0  logical = physical (consecutive tiles on consecutive XCDs)
//   1  every XCD takes one contiguous range of logical blocks
//   2  groups of G consecutive logical blocks (one row of tiles) go to one
//      XCD, groups dealt round-robin over the XCDs
constexpr int kXcds = 8;
__device__ inline int xcd_logical_block(int mode, int G) {
  const int n = static_cast<int>(gridDim.x);
  const int p = static_cast<int>(blockIdx.x);
  const int xcd = p % kXcds, local = p / kXcds;
  if (mode == 1) {
    const int q = n / kXcds, r = n % kXcds;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
  }
  if (mode == 2 && G > 0) {
    const int ng = n / G;
    const int nmain = (ng / kXcds) * kXcds * G;
    if (p >= nmain) return p;
    return (xcd + kXcds * (local / G)) * G + local % G;
  }
  return p;
}

// ---- plans (sg_api.cpp) ----
int validate_desc(const sg_grid_desc* desc);
int build_plan(const sg_grid_desc* desc, hipStream_t s, bool want_adjoint, Plan** result);
void destroy_plan(Plan* pl);

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
size_t fwd_stream_lds(const Plan& plan, int nc, size_t elem);
// Fields: derivative orders [nf][nin]; one output pointer per field.
hipError_t launch_forward(const Plan& plan, int nf, const int32_t* orders,
                          const void* P, const void* W, void* const* outs,
                          int dtype, hipStream_t s);

// ---- sg_mfma.hip ----
bool mma_supported(const Plan& plan);
hipError_t launch_forward_mma(const Plan& plan, int tier, int nf, const int32_t* orders,
                              const float* P, const float* W, float* const* outs,
                              hipStream_t s);
// Batched MFMA tiers: argument records (one per grid channel block), their
// occupancy and the persistent launch over the work queues (sg_queue.hpp).
size_t mma_args_size();
void mma_fill_args(const Plan& plan, int c0, int nc, const float* P, const float* W, float* out,
                   void* record);
int batch_mma_resident(int tier, size_t lds);
hipError_t launch_batch_mma(int tier, const void* args, const void* tiles, int* counters,
                            int blocks, size_t lds, hipStream_t s);

// ---- sg_rational.hip ----
// Derivative fields of a NURBS plan's rational map (quotient rule over the
// homogeneous B-spline fields); the workspace holds the homogeneous control
// points and fields.
int rational_closure_ok(const Plan& plan, int nf, const int32_t* orders);
size_t rational_workspace_bytes(const Plan& plan, int nf, const int32_t* orders, size_t elem);
hipError_t launch_rational(const Plan& plan, int nf, const int32_t* orders, const void* P,
                           const void* W, void* const* outs, void* work, int dtype,
                           hipStream_t s);
// The adjoint of a NURBS plan's evaluate! in its control points (weights
// fixed): Pbar_k = w_k [B^T (ebar / w(x))]_k.
size_t rational_adjoint_workspace_bytes(const Plan& plan, size_t elem);
hipError_t launch_rational_adjoint(const Plan& plan, const void* W, const void* ebar, void* Pbar,
                                   void* work, int dtype, int accumulate, hipStream_t s);

// ---- sg_adjoint.hip ----
hipError_t launch_adjoint(const Plan& plan, const int32_t* orders,
                          const void* ebar, void* Pbar, int dtype,
                          int accumulate, hipStream_t s);
size_t adjoint_work_bytes(const Plan& plan);
// Sum the partial windows [unit][w3max][nout][w2][w1] of tiling tl into
// Pbar (overwrite or accumulate), one thread per control point.
hipError_t launch_adjoint_gather(const Plan& plan, const Tiling& tl,
                                 const void* part, void* Pbar, int dtype, int accumulate,
                                 hipStream_t s);

// ---- sg_adjoint_mma.hip ----
bool adjoint_mma_shape_ok(const Plan& plan);
int adjoint_mma_resident_waves(const Plan& plan);  // on the current device
size_t adjoint_mma_work_bytes(const Plan& plan);
// scale (native tier only): when not null, ebar is multiplied sample by
// sample by the field scale[j3][j2][j1], shared by the channels, as it is
// loaded (the NURBS adjoint's 1 / w(x))
hipError_t launch_adjoint_mma(const Plan& plan, int tier, const int32_t* orders, const float* ebar,
                              float* Pbar, int accumulate, hipStream_t s,
                              const float* scale = nullptr);

}  // namespace sg
