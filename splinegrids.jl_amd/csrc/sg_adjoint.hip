// sg_adjoint.hip -- evaluate_adjoint!(grid): Pbar = M^T ebar, no atomics.
//
// Replaces spline_eval_adjoint_kernel (src/adjoint.jl:1-40), which lets
// every sample scatter into its (p+1)^Nin control points with
// Atomix.@atomic (:35-37).  Here every output is owned by exactly one
// thread, sums run in a fixed order, and results are bit-reproducible.
//
// Streaming path (Nin = 2, 3; monotone spans; p1, p2 <= 5), two kernels:
//   adj_stream  one workgroup per (64-sample tile of dim 1, chunk of rows of
//               dim 2, sample j3).  Each lane (sample j1) walks its wave's
//               rows j2 and accumulates U[k2] += B2[j2, b] * ebar[j1, j2, j3]
//               in a sliding window of p2+1 registers; an entry that leaves
//               the window is final and goes to the wave's LDS slab.  The
//               workgroup then contracts the slabs along dimension 1 (each
//               (k1, k2) output sums the lanes whose band holds k1) into a
//               partial slab of the tile's band window, written to the plan's
//               workspace.
//   adj_gather  one thread per control point (k1, k2, k3, c): sums the
//               partial slabs of the tiles and chunks whose windows hold
//               (k1, k2), weighted by B3[j3, k3 - s3 + p3] over the samples j3
//               whose band holds k3 (the transposed band list of dimension 3).
// General path (any Nin, any degree, any sample order): adj_csr, one thread
// per control point, walking each dimension's transposed band list.
#include "sg_internal.hpp"

namespace sg {

namespace {

template <typename T>
struct AdjStreamArgs {
  const T* E;
  T* part;
  const T* tab1;  // [m1][p1+1] at the requested order
  const T* tab2;  // [m2][P2+1]
  const int32_t* span1;
  const int32_t* span2;
  int m1, m2, p1;
  int c0, nout;
  int rows_per_wave;
  int w1max, w2max, w2wave;
  int n_tiles1, n_chunks2;
  long long m12, nsamp;
};

template <typename T, int P2, int NC>
__global__ __launch_bounds__(kThreads) void adj_stream(AdjStreamArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int PM = kMaxStreamDegree;
  T* S = reinterpret_cast<T*>(smem);                        // [4][w2wave][NC][64]
  T* Bl = S + static_cast<size_t>(kWaves) * a.w2wave * NC * kLanes;  // [64][PM+1]
  int* Sl = reinterpret_cast<int*>(Bl + kLanes * (PM + 1));  // [64]
  int* wlo = Sl + kLanes;                                    // [4]
  int* whi = wlo + kWaves;                                   // [4]
  int* llo = whi + kWaves;                                   // [w1max]
  int* lhi = llo + a.w1max;                                  // [w1max]

  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const int j1_0 = blockIdx.x * kLanes;
  const int chunk = a.rows_per_wave * kWaves;
  const int r0 = blockIdx.y * chunk;
  const int r1 = min(r0 + chunk, a.m2);
  const int j3 = blockIdx.z;
  const int j1_last = min(j1_0 + kLanes - 1, a.m1 - 1);
  const int p1 = a.p1;
  const int lo1 = a.span1[j1_0] - p1;
  const int W1 = a.span1[j1_last] - lo1 + 1;
  const int lo2 = a.span2[r0] - P2;
  const int W2 = a.span2[r1 - 1] - lo2 + 1;

  const bool active = (j1_0 + lane) < a.m1;
  const int j1 = active ? j1_0 + lane : a.m1 - 1;
  const int s1 = a.span1[j1];

  const int slab_elems = kWaves * a.w2wave * NC * kLanes;
  for (int i = threadIdx.x; i < slab_elems; i += kThreads) S[i] = T(0);
  if (wave == 0) {
    Sl[lane] = s1;
#pragma unroll
    for (int t = 0; t <= PM; ++t)
      Bl[lane * (PM + 1) + t] =
          (active && t <= p1) ? a.tab1[static_cast<size_t>(j1) * (p1 + 1) + min(t, p1)] : T(0);
  }
  __syncthreads();

  // ---- stage A: stream the wave's rows, sliding window over k2 ----
  const int rb = r0 + wave * a.rows_per_wave;
  const int re = min(rb + a.rows_per_wave, r1);
  if (rb < re) {
    T* Sw = S + static_cast<size_t>(wave) * a.w2wave * NC * kLanes;
    int cur = a.span2[rb];
    const int w0 = cur - P2;
    T U[P2 + 1][NC];
#pragma unroll
    for (int q = 0; q <= P2; ++q)
#pragma unroll
      for (int c = 0; c < NC; ++c) U[q][c] = T(0);
    const long long ebase = static_cast<long long>(j1) + a.m12 * j3 + a.nsamp * a.c0;
    for (int j2 = rb; j2 < re; ++j2) {
      const int s2 = a.span2[j2];
      while (cur < s2) {
        const int row = cur - P2 - w0;
#pragma unroll
        for (int c = 0; c < NC; ++c) Sw[(row * NC + c) * kLanes + lane] = U[0][c];
#pragma unroll
        for (int q = 0; q < P2; ++q)
#pragma unroll
          for (int c = 0; c < NC; ++c) U[q][c] = U[q + 1][c];
#pragma unroll
        for (int c = 0; c < NC; ++c) U[P2][c] = T(0);
        ++cur;
      }
      const T* b2p = a.tab2 + static_cast<size_t>(j2) * (P2 + 1);
      T bb[P2 + 1];
#pragma unroll
      for (int q = 0; q <= P2; ++q) bb[q] = b2p[q];
      const long long eoff = ebase + static_cast<long long>(a.m1) * j2;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const T e = active ? a.E[eoff + a.nsamp * c] : T(0);
#pragma unroll
        for (int q = 0; q <= P2; ++q) U[q][c] += bb[q] * e;
      }
    }
#pragma unroll
    for (int q = 0; q <= P2; ++q) {
      const int row = cur - P2 + q - w0;
#pragma unroll
      for (int c = 0; c < NC; ++c) Sw[(row * NC + c) * kLanes + lane] = U[q][c];
    }
    if (lane == 0) {
      wlo[wave] = w0;
      whi[wave] = cur;
    }
  } else if (lane == 0) {
    wlo[wave] = 1;
    whi[wave] = 0;  // empty window
  }
  // lanes whose band holds k1: s1 in [k1, k1 + p1] (spans monotone in lanes)
  for (int q1 = threadIdx.x; q1 < W1; q1 += kThreads) {
    const int k1 = lo1 + q1;
    int lo = 0, hi = kLanes;  // first lane with Sl >= k1
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (a.span1[min(j1_0 + mid, a.m1 - 1)] < k1) lo = mid + 1; else hi = mid;
    }
    int lo2_ = lo, hi2_ = kLanes;  // first lane with Sl > k1 + p1
    while (lo2_ < hi2_) {
      const int mid = (lo2_ + hi2_) >> 1;
      if (a.span1[min(j1_0 + mid, a.m1 - 1)] <= k1 + p1) lo2_ = mid + 1; else hi2_ = mid;
    }
    llo[q1] = lo;
    lhi[q1] = lo2_;  // exclusive
  }
  __syncthreads();

  // ---- stage B: contract along dimension 1 into the partial slab ----
  const size_t unit = (static_cast<size_t>(j3) * a.n_chunks2 + blockIdx.y) * a.n_tiles1 +
                      blockIdx.x;
  const size_t plane = static_cast<size_t>(a.w1max) * a.w2max;
  T* out = a.part + (unit * a.nout + a.c0) * plane;
  const int nouts = W1 * W2 * NC;
  for (int idx = threadIdx.x; idx < nouts; idx += kThreads) {
    const int q1 = idx % W1;
    const int rest = idx / W1;
    const int q2 = rest % W2;
    const int c = rest / W2;
    const int k1 = lo1 + q1, k2 = lo2 + q2;
    T sum = T(0);
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      if (k2 < wlo[w] || k2 > whi[w]) continue;
      const T* Sw = S + (static_cast<size_t>(w) * a.w2wave + (k2 - wlo[w])) * NC * kLanes +
                    c * kLanes;
      for (int l = llo[q1]; l < lhi[q1]; ++l)
        sum += Bl[l * (PM + 1) + (k1 - Sl[l] + p1)] * Sw[l];
    }
    out[c * plane + static_cast<size_t>(q2) * a.w1max + q1] = sum;
  }
}

template <typename T>
struct AdjGatherArgs {
  const T* part;
  T* out;
  const int32_t* tile_of_k1;
  const int32_t* chunk_of_k2;
  const int32_t* csr3_ptr;  // dimension 3 transposed band list (3-D)
  const int32_t* csr3_js;
  const T* tab3;
  const int32_t* span1;
  const int32_t* span2;
  int n1, n2, nin;
  int p1, p2, p3;
  int nout, w1max, w2max, n_tiles1, n_chunks2, rows_chunk;
  int accumulate;
  long long ncp;
};

template <typename T>
__global__ __launch_bounds__(256) void adj_gather(AdjGatherArgs<T> a) {
  const long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= a.ncp * a.nout) return;
  const int c = static_cast<int>(idx / a.ncp);
  long long r = idx - static_cast<long long>(c) * a.ncp;
  const int k1 = static_cast<int>(r % a.n1);
  r /= a.n1;
  const int k2 = static_cast<int>(r % a.n2);
  const int k3 = static_cast<int>(r / a.n2);
  const int ta = a.tile_of_k1[2 * k1], tb = a.tile_of_k1[2 * k1 + 1];
  const int ca = a.chunk_of_k2[2 * k2], cb = a.chunk_of_k2[2 * k2 + 1];
  const size_t plane = static_cast<size_t>(a.w1max) * a.w2max;
  int e0 = 0, e1 = 1;
  if (a.nin == 3) {
    e0 = a.csr3_ptr[k3];
    e1 = a.csr3_ptr[k3 + 1];
  }
  T sum = T(0);
  for (int e = e0; e < e1; ++e) {
    int j3 = 0;
    T b3 = T(1);
    if (a.nin == 3) {
      const int js = a.csr3_js[e];
      j3 = js / (a.p3 + 1);
      b3 = a.tab3[js];
    }
    T inner = T(0);
    for (int by = ca; by <= cb; ++by) {
      const int lo2 = a.span2[by * a.rows_chunk] - a.p2;
      for (int bx = ta; bx <= tb; ++bx) {
        const int lo1 = a.span1[bx * kLanes] - a.p1;
        const size_t unit = (static_cast<size_t>(j3) * a.n_chunks2 + by) * a.n_tiles1 + bx;
        inner += a.part[(unit * a.nout + c) * plane +
                        static_cast<size_t>(k2 - lo2) * a.w1max + (k1 - lo1)];
      }
    }
    sum += b3 * inner;
  }
  a.out[idx] = a.accumulate ? a.out[idx] + sum : sum;
}

template <typename T>
struct AdjCsrArgs {
  const T* E;
  T* out;
  const T* tab[SG_MAX_NIN];
  const int32_t* ptr[SG_MAX_NIN];
  const int32_t* js[SG_MAX_NIN];
  int nin, nout;
  int n[SG_MAX_NIN];
  int w[SG_MAX_NIN];           // p_d + 1
  long long sst[SG_MAX_NIN];   // sample strides
  long long ncp, nsamp;
  int accumulate;
};

template <typename T>
__global__ __launch_bounds__(256) void adj_csr(AdjCsrArgs<T> a) {
  const long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= a.ncp * a.nout) return;
  const int c = static_cast<int>(idx / a.ncp);
  long long r = idx - static_cast<long long>(c) * a.ncp;
  int lo[SG_MAX_NIN], hi[SG_MAX_NIN], e[SG_MAX_NIN];
  bool empty = false;
#pragma unroll
  for (int d = 0; d < SG_MAX_NIN; ++d) {
    if (d < a.nin) {
      const int k = static_cast<int>(r % a.n[d]);
      r /= a.n[d];
      lo[d] = a.ptr[d][k];
      hi[d] = a.ptr[d][k + 1];
      e[d] = lo[d];
      empty |= lo[d] >= hi[d];
    }
  }
  T sum = T(0);
  if (!empty) {
    const T* E = a.E + a.nsamp * c;
    while (true) {
      T prod = T(1);
      long long lin = 0;
#pragma unroll
      for (int d = 0; d < SG_MAX_NIN; ++d)
        if (d < a.nin) {
          const int js = a.js[d][e[d]];
          prod *= a.tab[d][js];
          lin += static_cast<long long>(js / a.w[d]) * a.sst[d];
        }
      sum += prod * E[lin];
      bool carry = true;
#pragma unroll
      for (int d = 0; d < SG_MAX_NIN; ++d) {
        if (carry && d < a.nin) {
          if (++e[d] < hi[d])
            carry = false;
          else
            e[d] = lo[d];
        }
      }
      if (carry) break;
    }
  }
  a.out[idx] = a.accumulate ? a.out[idx] + sum : sum;
}

template <typename T>
const T* ktab(const DimDev& d, int order);
template <>
const float* ktab<float>(const DimDev& d, int order) {
  return d.ktab32 + static_cast<size_t>(order) * d.m * (d.p + 1);
}
template <>
const double* ktab<double>(const DimDev& d, int order) {
  return d.ktab64 + static_cast<size_t>(order) * d.m * (d.p + 1);
}

template <typename T>
using AdjKernel = void (*)(AdjStreamArgs<T>);

template <typename T, int P2>
AdjKernel<T> pick_nc(int nc) {
  switch (nc) {
    case 1: return adj_stream<T, P2, 1>;
    case 2: return adj_stream<T, P2, 2>;
    case 3: return adj_stream<T, P2, 3>;
    case 4: return adj_stream<T, P2, 4>;
  }
  return nullptr;
}

template <typename T>
AdjKernel<T> pick_adj(int p2, int nc) {
  switch (p2) {
    case 0: return pick_nc<T, 0>(nc);
    case 1: return pick_nc<T, 1>(nc);
    case 2: return pick_nc<T, 2>(nc);
    case 3: return pick_nc<T, 3>(nc);
    case 4: return pick_nc<T, 4>(nc);
    case 5: return pick_nc<T, 5>(nc);
  }
  return nullptr;
}

template <typename T>
size_t stream_lds(const Plan& pl, int nc) {
  return static_cast<size_t>(kWaves) * pl.w2wave_max * nc * kLanes * sizeof(T) +
         kLanes * (kMaxStreamDegree + 1) * sizeof(T) +
         (kLanes + 2 * kWaves + 2 * pl.w1max) * sizeof(int);
}

template <typename T>
int max_nc(const Plan& pl) {
  for (int nc = 4; nc >= 1; --nc)
    if (stream_lds<T>(pl, nc) <= static_cast<size_t>(kMaxLds)) return nc;
  return 0;
}

template <typename T>
hipError_t adjoint_stream(const Plan& pl, const int32_t* orders, const T* E,
                          T* Pbar, int accumulate, hipStream_t s) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  const int nin = pl.nin;
  AdjStreamArgs<T> a{};
  a.E = E;
  a.part = static_cast<T*>(pl.work);
  a.tab1 = ktab<T>(d1, orders[0]);
  a.tab2 = ktab<T>(d2, orders[1]);
  a.span1 = d1.span;
  a.span2 = d2.span;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.p1 = d1.p;
  a.nout = pl.nout;
  a.rows_per_wave = pl.rows_per_wave;
  a.w1max = pl.w1max;
  a.w2max = pl.w2max;
  a.w2wave = pl.w2wave_max;
  a.n_tiles1 = pl.n_tiles1;
  a.n_chunks2 = pl.n_chunks2;
  a.m12 = static_cast<long long>(d1.m) * d2.m;
  a.nsamp = a.m12 * (nin == 3 ? pl.dims[2].m : 1);
  const unsigned gz = nin == 3 ? static_cast<unsigned>(pl.dims[2].m) : 1u;
  const dim3 grid(pl.n_tiles1, pl.n_chunks2, gz);
  const int ncmax = max_nc<T>(pl);
  if (ncmax == 0) return hipErrorInvalidValue;
  for (int c0 = 0; c0 < pl.nout;) {
    const int nc = std::min(ncmax, pl.nout - c0);
    AdjKernel<T> k = pick_adj<T>(d2.p, nc);
    if (!k) return hipErrorInvalidValue;
    a.c0 = c0;
    hipLaunchKernelGGL(k, grid, dim3(kThreads), stream_lds<T>(pl, nc), s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    c0 += nc;
  }

  AdjGatherArgs<T> g{};
  g.part = static_cast<const T*>(pl.work);
  g.out = Pbar;
  g.tile_of_k1 = pl.tile_of_k1;
  g.chunk_of_k2 = pl.chunk_of_k2;
  g.span1 = d1.span;
  g.span2 = d2.span;
  g.n1 = d1.n;
  g.n2 = d2.n;
  g.nin = nin;
  g.p1 = d1.p;
  g.p2 = d2.p;
  if (nin == 3) {
    const DimDev& d3 = pl.dims[2];
    g.csr3_ptr = d3.csr_ptr;
    g.csr3_js = d3.csr_js;
    g.tab3 = ktab<T>(d3, orders[2]);
    g.p3 = d3.p;
  }
  g.nout = pl.nout;
  g.w1max = pl.w1max;
  g.w2max = pl.w2max;
  g.n_tiles1 = pl.n_tiles1;
  g.n_chunks2 = pl.n_chunks2;
  g.rows_chunk = pl.rows_per_wave * kWaves;
  g.accumulate = accumulate;
  g.ncp = 1;
  for (int d = 0; d < nin; ++d) g.ncp *= pl.dims[d].n;
  const long long total = g.ncp * pl.nout;
  hipLaunchKernelGGL(adj_gather<T>, dim3(static_cast<unsigned>((total + 255) / 256)),
                     dim3(256), 0, s, g);
  return hipGetLastError();
}

template <typename T>
hipError_t adjoint_csr(const Plan& pl, const int32_t* orders, const T* E,
                       T* Pbar, int accumulate, hipStream_t s) {
  AdjCsrArgs<T> a{};
  a.E = E;
  a.out = Pbar;
  a.nin = pl.nin;
  a.nout = pl.nout;
  a.accumulate = accumulate;
  a.ncp = 1;
  a.nsamp = 1;
  for (int d = 0; d < pl.nin; ++d) {
    const DimDev& dd = pl.dims[d];
    a.tab[d] = ktab<T>(dd, orders[d]);
    a.ptr[d] = dd.csr_ptr;
    a.js[d] = dd.csr_js;
    a.n[d] = dd.n;
    a.w[d] = dd.p + 1;
    a.sst[d] = a.nsamp;
    a.ncp *= dd.n;
    a.nsamp *= dd.m;
  }
  const long long total = a.ncp * pl.nout;
  hipLaunchKernelGGL(adj_csr<T>, dim3(static_cast<unsigned>((total + 255) / 256)),
                     dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

size_t adjoint_work_bytes(const Plan& pl) {
  if (!pl.stream_ok) return 0;
  const size_t units = static_cast<size_t>(pl.n_tiles1) * pl.n_chunks2 *
                       (pl.nin == 3 ? pl.dims[2].m : 1);
  return units * pl.nout * pl.w1max * pl.w2max * sizeof(double);
}

hipError_t launch_adjoint(const Plan& pl, const int32_t* orders,
                          const void* ebar, void* Pbar, int dtype,
                          int accumulate, hipStream_t s) {
  long long ncp = 1, nsamp = 1;
  for (const DimDev& d : pl.dims) {
    ncp *= d.n;
    nsamp *= d.m;
  }
  if (ncp == 0) return hipSuccess;
  const bool stream = pl.stream_ok && nsamp > 0 &&
                      (dtype == SG_F32 ? max_nc<float>(pl) : max_nc<double>(pl)) > 0;
  if (dtype == SG_F32) {
    auto e = static_cast<const float*>(ebar);
    auto o = static_cast<float*>(Pbar);
    return stream ? adjoint_stream<float>(pl, orders, e, o, accumulate, s)
                  : adjoint_csr<float>(pl, orders, e, o, accumulate, s);
  }
  auto e = static_cast<const double*>(ebar);
  auto o = static_cast<double*>(Pbar);
  return stream ? adjoint_stream<double>(pl, orders, e, o, accumulate, s)
                : adjoint_csr<double>(pl, orders, e, o, accumulate, s);
}

}  // namespace sg
