// sg_rational.hip -- derivative fields of a NURBS grid's rational map
// R = A / w, A = sum N w P, w = sum N w, and its adjoint in the control
// points.  The reference rejects both (src/validation.jl:81-88,
// src/adjoint.jl:57).
//
// Adjoint: with the weights fixed R is linear in P, R = B (w P) / w(x), so
// Pbar_k = w_k [B^T (ebar / w(x))]_k.  FP32 grids on the matrix-core adjoint:
// 1 / w(x) in one streaming pass (the weight field in registers, as the
// B-spline forward computes it), then the B-spline adjoint multiplying ebar
// by it as it loads ebar, then a product.  Otherwise ebar / w(x) is written
// by that pass and the B-spline adjoint reads it.
//
// Two passes.  (1) The homogeneous control points (w P_c, w) go through the
// B-spline forward (launch_forward on a view of the plan without weights and
// with Nout + 1 channels) for every order b below a requested one, giving
// A^(b) and w^(b).  (2) One thread per (sample, channel) applies Leibniz's
// rule to A = w R in FP64 (The NURBS Book, eq. 4.8 / Alg. A4.4):
//   R^(g) = (A^(g) - sum_{0 < b <= g} C(g, b) w^(b) R^(g-b)) / w,
// C(g, b) = prod_d binom(g_d, b_d), over the orders g of the closure in
// increasing total order, and writes the requested fields.
//
// Value and first derivatives (every requested order of total order <= 1)
// take one fused pass instead (rat_grad_stream): the forward's tiles and
// stage 1 over the homogeneous window (w P_c, w), a second window contracted
// with B3' when d/dx3 is asked for, and in each row the sums A, w and their
// first derivatives in registers, then R = A / w and
// dR/dx_d = (A_d - w_d R) / w before the stores; nothing but the requested
// fields goes through HBM.
#include <algorithm>
#include <array>
#include <vector>

#include "sg_tiles.hpp"

namespace sg {

namespace {

constexpr int kMaxClosure = 16;  // orders below the requested ones
constexpr int kMaxTerms = 96;    // Leibniz terms over the closure

template <typename T>
struct RatArgs {
  const T* F;  // [closure][nout + 1][nsamp]: A^(b) per channel, then w^(b)
  T* out[SG_MAX_FIELDS];
  int out_idx[SG_MAX_FIELDS];  // closure index of each requested field
  int nfields, nclosure, nterms, nout;
  long long nsamp;
  // term t: R[tg] -= coef * w[tb] * R[tr], terms grouped by tg in order
  unsigned char tg[kMaxTerms], tb[kMaxTerms], tr[kMaxTerms];
  float coef[kMaxTerms];
};

// ---- fused value and first derivatives ----

template <typename T>
struct RatGradArgs {
  BandSrc<T> src;    // control points and weights, dimension 3 at order 0
  const T* tab3d;    // dimension 3 at order 1 (3-D grids asking for d/dx3)
  const T* tab1[2];  // dimension 1 at orders 0 and 1 [m1][p1+1]
  const int32_t* span1;
  int m1, m2, p1, c0, rows_per_wave, wx, ntx, nty, xcd_map;
  long long nsamp;
  int store[4];  // value, d/dx1, d/dx2, d/dx3: written or not
};

// One tile (the forward's tiles: kLanes * wx samples j1, a chunk of rows j2,
// one sample j3) for channels [c0, c0 + NC): stage 1 stages the homogeneous
// window G0 (B3 contraction) and, with D3, G1 (B3' contraction); each lane
// walks its rows with the register windows V0 = b1 . G0, V0d = b1' . G0 and
// V1 = b1 . G1 over k2, and a row's sums are A = B2 . V0, A_1 = B2 . V0d,
// A_2 = B2' . V0, A_3 = B2 . V1 (the weight as channel NC of each).
template <typename T, int NIN, int P2, int NC, bool D3>
__global__ __launch_bounds__(kThreads) void rat_grad_stream(
    RatGradArgs<T> a, const int32_t* __restrict__ span2, const T* __restrict__ tb2,
    const T* __restrict__ tb2d, T* __restrict__ o_val, T* __restrict__ o_d1,
    T* __restrict__ o_d2, T* __restrict__ o_d3) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* G = reinterpret_cast<T*>(smem);
  constexpr int NCH = NC + 1;
  constexpr int PM = kMaxStreamDegree;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const int wy = kWaves / a.wx;
  const FwdTile tile = fwd_tile(a.xcd_map, a.ntx, a.nty, kLanes * a.wx, a.rows_per_wave * wy,
                                a.m1, a.m2, NIN);
  const int j1_0 = tile.j1_0, r0 = tile.r0, r1 = tile.r1, j3 = tile.j3;
  const int lo1 = a.span1[j1_0] - a.p1;
  const int W1 = a.span1[tile.j1_last] - lo1 + 1;
  const int lo2 = span2[r0] - P2;
  const int W2 = span2[r1 - 1] - lo2 + 1;
  const int cs = W1 * W2;

  stage1_band_window<T, NIN, NC, true>(G, a.src, j3, lo1, W1, lo2, W2);
  if constexpr (D3) {
    BandSrc<T> src = a.src;
    src.tab3 = a.tab3d;
    stage1_band_window<T, NIN, NC, true>(G + NCH * cs, src, j3, lo1, W1, lo2, W2);
  }
  __syncthreads();

  const int wave_x = wave % a.wx, wave_y = wave / a.wx;
  const int rb = r0 + wave_y * a.rows_per_wave;
  const int re = min(rb + a.rows_per_wave, r1);
  const int j1 = j1_0 + kLanes * wave_x + lane;
  // no barrier follows: lanes past the grid, and waves without rows, leave
  if (rb >= re || j1 >= a.m1) return;
  const int p1 = a.p1;
  const int q1b = a.span1[j1] - p1 - lo1;
  T b1[2][PM + 1];  // zero past p1 (the clamped index re-reads the band's end)
#pragma unroll
  for (int o = 0; o < 2; ++o)
#pragma unroll
    for (int t = 0; t <= PM; ++t) {
      const T v = a.tab1[o][static_cast<size_t>(j1) * (p1 + 1) + min(t, p1)];
      b1[o][t] = t <= p1 ? v : T(0);
    }

  constexpr int NV = D3 ? 3 : 2;  // windows: V0, V0d (, V1)
  T V[P2 + 1][NV][NCH];
  auto column = [&](int k2, T (&v)[NV][NCH]) {
    const T* g = G + (k2 - lo2) * W1 + q1b;
    T x[D3 ? 2 : 1][NCH][PM + 1];  // every read of the column before the sums
#pragma unroll
    for (int h = 0; h < (D3 ? 2 : 1); ++h)
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
        for (int t = 0; t <= PM; ++t) x[h][ch][t] = g[(h * NCH + ch) * cs + min(t, p1)];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
      for (int t = 0; t <= PM; ++t) {
        s0 += b1[0][t] * x[0][ch][t];
        s1 += b1[1][t] * x[0][ch][t];
        if constexpr (D3) s2 += b1[0][t] * x[D3 ? 1 : 0][ch][t];
      }
      v[0][ch] = s0;
      v[1][ch] = s1;
      if constexpr (D3) v[NV - 1][ch] = s2;
    }
  };
  int cur = span2[rb];
#pragma unroll
  for (int q = 0; q <= P2; ++q) column(cur - P2 + q, V[q]);

  long long o = static_cast<long long>(j1) + static_cast<long long>(a.m1) * a.m2 * j3 +
                a.nsamp * a.c0 + static_cast<long long>(a.m1) * rb;
  for (int j2 = rb; j2 < re; ++j2, o += a.m1) {
    const int s2 = span2[j2];
    if (s2 != cur) {
      if (s2 - cur > P2) {
        cur = s2;
#pragma unroll
        for (int q = 0; q <= P2; ++q) column(cur - P2 + q, V[q]);
      } else {
        while (cur < s2) {
#pragma unroll
          for (int q = 0; q < P2; ++q)
#pragma unroll
            for (int h = 0; h < NV; ++h)
#pragma unroll
              for (int ch = 0; ch < NCH; ++ch) V[q][h][ch] = V[q + 1][h][ch];
          ++cur;
          column(cur, V[P2]);
        }
      }
    }
    const T* bb0 = tb2 + static_cast<size_t>(j2) * (P2 + 1);
    const T* bb1 = tb2d + static_cast<size_t>(j2) * (P2 + 1);
    T A[NCH], A1[NCH], A2[NCH], A3[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
#pragma unroll
      for (int q = 0; q <= P2; ++q) {
        s0 += bb0[q] * V[q][0][ch];
        s1 += bb0[q] * V[q][1][ch];
        s2 += bb1[q] * V[q][0][ch];
        if constexpr (D3) s3 += bb0[q] * V[q][NV - 1][ch];
      }
      A[ch] = s0;
      A1[ch] = s1;
      A2[ch] = s2;
      A3[ch] = s3;
    }
    const T inv = T(1) / A[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const T R = A[c] * inv;
      const long long i = o + a.nsamp * c;
      if (a.store[0]) o_val[i] = R;
      if (a.store[1]) o_d1[i] = (A1[c] - A1[NC] * R) * inv;
      if (a.store[2]) o_d2[i] = (A2[c] - A2[NC] * R) * inv;
      if constexpr (D3)
        if (a.store[3]) o_d3[i] = (A3[c] - A3[NC] * R) * inv;
    }
  }
}

template <typename T>
using RatGradKernel = void (*)(RatGradArgs<T>, const int32_t*, const T*, const T*, T*, T*, T*,
                               T*);

template <typename T, int NIN, int P2, bool D3>
RatGradKernel<T> pick_rat_grad_nc(int nc) {
  switch (nc) {
    case 1: return rat_grad_stream<T, NIN, P2, 1, D3>;
    case 2: return rat_grad_stream<T, NIN, P2, 2, D3>;
    case 3: return rat_grad_stream<T, NIN, P2, 3, D3>;
  }
  return nullptr;
}

template <typename T, int NIN, bool D3>
RatGradKernel<T> pick_rat_grad_p2(int p2, int nc) {
  switch (p2) {
    case 0: return pick_rat_grad_nc<T, NIN, 0, D3>(nc);
    case 1: return pick_rat_grad_nc<T, NIN, 1, D3>(nc);
    case 2: return pick_rat_grad_nc<T, NIN, 2, D3>(nc);
    case 3: return pick_rat_grad_nc<T, NIN, 3, D3>(nc);
    case 4: return pick_rat_grad_nc<T, NIN, 4, D3>(nc);
    case 5: return pick_rat_grad_nc<T, NIN, 5, D3>(nc);
  }
  return nullptr;
}

template <typename T>
const T* order_table(const DimDev& d, int order);
template <>
const float* order_table<float>(const DimDev& d, int order) {
  return d.ktab32 + static_cast<size_t>(order) * d.m * (d.p + 1);
}
template <>
const double* order_table<double>(const DimDev& d, int order) {
  return d.ktab64 + static_cast<size_t>(order) * d.m * (d.p + 1);
}

constexpr int kRatGradMaxNC = 3;  // channels per launch (the weight comes on top)

// The fused pass applies when every field has total order <= 1, no field
// repeats, the grid takes the streaming kernels and the windows fit; returns
// the channels per launch, or 0.
int rat_grad_channels(const Plan& pl, int nf, const int32_t* orders, size_t elem) {
  if (!pl.stream_ok || (pl.nin != 2 && pl.nin != 3)) return 0;
  int seen[4] = {0, 0, 0, 0};
  bool d3 = false;
  for (int f = 0; f < nf; ++f) {
    int total = 0, kind = 0;
    for (int d = 0; d < pl.nin; ++d) {
      const int o = orders[f * pl.nin + d];
      total += o;
      if (o == 1) kind = d + 1;
    }
    if (total > 1 || seen[kind]++) return 0;
    d3 = d3 || kind == 3;
  }
  const Tiling& t = pl.fwd;
  for (int nc = std::min(kRatGradMaxNC, pl.nout); nc >= 1; --nc) {
    const size_t lds = static_cast<size_t>(d3 ? 2 : 1) * (nc + 1) * t.w1max * t.w2max * elem;
    if (lds <= static_cast<size_t>(kMaxLds)) return nc;
  }
  return 0;
}

template <typename T>
hipError_t rational_first_order(const Plan& pl, int nf, const int32_t* orders, const T* P,
                                const T* W, T* const* outs, int nc_max, hipStream_t s) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  RatGradArgs<T> a{};
  T* out[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int f = 0; f < nf; ++f) {
    int kind = 0;
    for (int d = 0; d < pl.nin; ++d)
      if (orders[f * pl.nin + d] == 1) kind = d + 1;
    out[kind] = outs[f];
    a.store[kind] = 1;
  }
  a.src.P = P;
  a.src.W = W;
  a.src.n1 = d1.n;
  a.src.n2 = d2.n;
  a.src.ncp = static_cast<long long>(d1.n) * d2.n;
  a.span1 = d1.span;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.p1 = d1.p;
  a.rows_per_wave = pl.fwd.rows_per_wave;
  a.wx = pl.fwd.wx;
  a.ntx = pl.fwd.n_tiles1;
  a.nty = pl.fwd.n_chunks2;
  a.xcd_map = pl.fwd.xcd_map;
  a.nsamp = static_cast<long long>(d1.m) * d2.m;
  // order-1 tables only where a derivative is asked for (D may be 0 elsewhere)
  a.tab1[0] = order_table<T>(d1, 0);
  a.tab1[1] = order_table<T>(d1, a.store[1] ? 1 : 0);
  const T* tb2 = order_table<T>(d2, 0);
  const T* tb2d = order_table<T>(d2, a.store[2] ? 1 : 0);
  unsigned gz = 1;
  const bool d3 = a.store[3] != 0;
  if (pl.nin == 3) {
    const DimDev& D3 = pl.dims[2];
    a.src.span3 = D3.span;
    a.src.p3 = D3.p;
    a.src.tab3 = order_table<T>(D3, 0);
    a.tab3d = order_table<T>(D3, d3 ? 1 : 0);
    a.src.ncp *= D3.n;
    a.nsamp *= D3.m;
    gz = static_cast<unsigned>(D3.m);
  }
  const dim3 grid(pl.fwd.n_tiles1 * pl.fwd.n_chunks2 * gz);
  for (int c0 = 0; c0 < pl.nout; c0 += nc_max) {
    const int nc = std::min(nc_max, pl.nout - c0);
    RatGradKernel<T> k = pl.nin == 2 ? pick_rat_grad_p2<T, 2, false>(d2.p, nc)
                         : d3        ? pick_rat_grad_p2<T, 3, true>(d2.p, nc)
                                     : pick_rat_grad_p2<T, 3, false>(d2.p, nc);
    if (!k) return hipErrorInvalidValue;
    a.c0 = c0;
    a.src.c0 = c0;
    const size_t lds = static_cast<size_t>(d3 ? 2 : 1) * (nc + 1) * pl.fwd.w1max *
                       pl.fwd.w2max * sizeof(T);
    hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a, d2.span, tb2, tb2d, out[0], out[1],
                       out[2], out[3]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
__global__ void homogenize(const T* P, const T* W, T* H, long long ncp, int nout) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= ncp) return;
  const T w = W[i];
  for (int c = 0; c < nout; ++c) H[i + ncp * c] = w * P[i + ncp * c];
  H[i + ncp * nout] = w;
}

// One thread per sample, all channels in turn (the weight fields are read
// once).  A thread's closure values sit in a column of LDS, [2][K][256]
// doubles (w, then R), indexed at run time without scratch memory.
constexpr int kCombineThreads = 256;

template <typename T, int K>
__global__ __launch_bounds__(kCombineThreads) void rational_combine(RatArgs<T> a) {
  extern __shared__ double sh[];
  constexpr int S = kCombineThreads;  // stride between a thread's entries
  double* w = sh + threadIdx.x;
  double* R = sh + static_cast<size_t>(K) * S + threadIdx.x;
  const long long fstride = a.nsamp * (a.nout + 1);
  for (long long j = static_cast<long long>(blockIdx.x) * S + threadIdx.x; j < a.nsamp;
       j += static_cast<long long>(gridDim.x) * S) {
    // the closure size is compile-time: each group of K loads is issued at once
    T v[K];
#pragma unroll
    for (int g = 0; g < K; ++g) v[g] = a.F[g * fstride + a.nout * a.nsamp + j];
#pragma unroll
    for (int g = 0; g < K; ++g) w[g * S] = static_cast<double>(v[g]);
    for (int c = 0; c < a.nout; ++c) {
#pragma unroll
      for (int g = 0; g < K; ++g) v[g] = a.F[g * fstride + c * a.nsamp + j];
#pragma unroll
      for (int g = 0; g < K; ++g) R[g * S] = static_cast<double>(v[g]);
      int t = 0;
#pragma unroll
      for (int g = 0; g < K; ++g) {
        double acc = R[g * S];
        for (; t < a.nterms && a.tg[t] == g; ++t)
          acc -= static_cast<double>(a.coef[t]) * w[a.tb[t] * S] * R[a.tr[t] * S];
        R[g * S] = acc / w[0];
      }
      for (int f = 0; f < a.nfields; ++f)
        a.out[f][c * a.nsamp + j] = static_cast<T>(R[a.out_idx[f] * S]);
    }
  }
}

template <typename T, int K>
void* combine_kernel_k(int k) {
  if constexpr (K > kMaxClosure) {
    return nullptr;
  } else {
    return k == K ? reinterpret_cast<void*>(rational_combine<T, K>) : combine_kernel_k<T, K + 1>(k);
  }
}

// ebar / w(x) for every channel in one pass: the forward's tiles over the
// weights as a one-channel B-spline (stage 1 window of w, the register
// window over k2), and each row's w(x) divides the row of ebar of every
// channel as it is read; with E null, the pass writes 1 / w(x) instead.
template <typename T>
struct DivArgs {
  BandSrc<T> src;  // P = the weights, one channel
  const T* tab1;
  const int32_t* span1;
  const T* E;
  T* out;
  int m1, m2, p1, nout, rows_per_wave, wx, ntx, nty, xcd_map;
  long long nsamp;
};

template <typename T, int NIN, int P2>
__global__ __launch_bounds__(kThreads) void divide_by_weight_field(
    DivArgs<T> a, const int32_t* __restrict__ span2, const T* __restrict__ tb2) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* G = reinterpret_cast<T*>(smem);
  constexpr int PM = kMaxStreamDegree;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const int wy = kWaves / a.wx;
  const FwdTile tile = fwd_tile(a.xcd_map, a.ntx, a.nty, kLanes * a.wx, a.rows_per_wave * wy,
                                a.m1, a.m2, NIN);
  const int j1_0 = tile.j1_0, r0 = tile.r0, r1 = tile.r1, j3 = tile.j3;
  const int lo1 = a.span1[j1_0] - a.p1;
  const int W1 = a.span1[tile.j1_last] - lo1 + 1;
  const int lo2 = span2[r0] - P2;
  const int W2 = span2[r1 - 1] - lo2 + 1;
  stage1_band_window<T, NIN, 1, false>(G, a.src, j3, lo1, W1, lo2, W2);
  __syncthreads();
  const int wave_x = wave % a.wx, wave_y = wave / a.wx;
  const int rb = r0 + wave_y * a.rows_per_wave;
  const int re = min(rb + a.rows_per_wave, r1);
  const int j1 = j1_0 + kLanes * wave_x + lane;
  if (rb >= re || j1 >= a.m1) return;  // no barrier follows
  const int p1 = a.p1;
  const int q1b = a.span1[j1] - p1 - lo1;
  T b1[PM + 1];
#pragma unroll
  for (int t = 0; t <= PM; ++t) {
    const T v = a.tab1[static_cast<size_t>(j1) * (p1 + 1) + min(t, p1)];
    b1[t] = t <= p1 ? v : T(0);
  }
  T V[P2 + 1];
  auto column = [&](int k2) {
    const T* g = G + (k2 - lo2) * W1 + q1b;
    T s = T(0);
#pragma unroll
    for (int t = 0; t <= PM; ++t) s += b1[t] * g[min(t, p1)];
    return s;
  };
  int cur = span2[rb];
#pragma unroll
  for (int q = 0; q <= P2; ++q) V[q] = column(cur - P2 + q);
  long long o = static_cast<long long>(j1) + static_cast<long long>(a.m1) * a.m2 * j3 +
                static_cast<long long>(a.m1) * rb;
  for (int j2 = rb; j2 < re; ++j2, o += a.m1) {
    const int s2 = span2[j2];
    if (s2 != cur) {
      if (s2 - cur > P2) {
        cur = s2;
#pragma unroll
        for (int q = 0; q <= P2; ++q) V[q] = column(cur - P2 + q);
      } else {
        while (cur < s2) {
#pragma unroll
          for (int q = 0; q < P2; ++q) V[q] = V[q + 1];
          ++cur;
          V[P2] = column(cur);
        }
      }
    }
    const T* bb = tb2 + static_cast<size_t>(j2) * (P2 + 1);
    T w = T(0);
#pragma unroll
    for (int q = 0; q <= P2; ++q) w += bb[q] * V[q];
    const T inv = T(1) / w;
    if (a.E == nullptr)
      a.out[o] = inv;
    else
      for (int c = 0; c < a.nout; ++c) a.out[o + a.nsamp * c] = a.E[o + a.nsamp * c] * inv;
  }
}

template <typename T, int NIN>
auto pick_divide(int p2) -> void (*)(DivArgs<T>, const int32_t*, const T*) {
  switch (p2) {
    case 0: return divide_by_weight_field<T, NIN, 0>;
    case 1: return divide_by_weight_field<T, NIN, 1>;
    case 2: return divide_by_weight_field<T, NIN, 2>;
    case 3: return divide_by_weight_field<T, NIN, 3>;
    case 4: return divide_by_weight_field<T, NIN, 4>;
    case 5: return divide_by_weight_field<T, NIN, 5>;
  }
  return nullptr;
}

// ebar / w(x), channel by channel (the weight field w(x) in D): the general
// case, after the weight field went through the B-spline forward.
template <typename T>
__global__ void divide_by_field(const T* E, const T* D, T* out, long long nsamp, int nout) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nsamp) return;
  const double inv = 1.0 / static_cast<double>(D[i]);
  for (int c = 0; c < nout; ++c)
    out[i + nsamp * c] = static_cast<T>(static_cast<double>(E[i + nsamp * c]) * inv);
}

// The one-pass division (or, with E null, 1 / w(x) alone): streaming grids
// (2-D, 3-D) whose weight window fits; false otherwise.
template <typename T>
bool divide_streaming(const Plan& pl, const T* W, const T* E, T* out, hipStream_t s,
                      hipError_t& e) {
  if (!pl.stream_ok || (pl.nin != 2 && pl.nin != 3)) return false;
  const size_t lds = static_cast<size_t>(pl.fwd.w1max) * pl.fwd.w2max * sizeof(T);
  if (lds > static_cast<size_t>(kMaxLds)) return false;
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  DivArgs<T> a{};
  a.src.P = W;
  a.src.n1 = d1.n;
  a.src.n2 = d2.n;
  a.src.ncp = static_cast<long long>(d1.n) * d2.n;
  a.tab1 = order_table<T>(d1, 0);
  a.span1 = d1.span;
  a.E = E;
  a.out = out;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.p1 = d1.p;
  a.nout = pl.nout;
  a.rows_per_wave = pl.fwd.rows_per_wave;
  a.wx = pl.fwd.wx;
  a.ntx = pl.fwd.n_tiles1;
  a.nty = pl.fwd.n_chunks2;
  a.xcd_map = pl.fwd.xcd_map;
  a.nsamp = static_cast<long long>(d1.m) * d2.m;
  unsigned gz = 1;
  if (pl.nin == 3) {
    const DimDev& d3 = pl.dims[2];
    a.src.span3 = d3.span;
    a.src.p3 = d3.p;
    a.src.tab3 = order_table<T>(d3, 0);
    a.src.ncp *= d3.n;
    a.nsamp *= d3.m;
    gz = static_cast<unsigned>(d3.m);
  }
  auto k = pl.nin == 2 ? pick_divide<T, 2>(d2.p) : pick_divide<T, 3>(d2.p);
  if (!k) return false;
  hipLaunchKernelGGL(k, dim3(pl.fwd.n_tiles1 * pl.fwd.n_chunks2 * gz), dim3(kThreads), lds, s, a,
                     d2.span, order_table<T>(d2, 0));
  e = hipGetLastError();
  return true;
}

// Pbar = (accumulate ? Pbar : 0) + w * Q, channel by channel.
template <typename T>
__global__ void weight_products(const T* Q, const T* W, T* Pbar, long long ncp, int nout,
                                int accumulate) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= ncp) return;
  const T w = W[i];
  for (int c = 0; c < nout; ++c) {
    const T v = w * Q[i + ncp * c];
    Pbar[i + ncp * c] = accumulate ? Pbar[i + ncp * c] + v : v;
  }
}

template <typename T>
hipError_t rational_adjoint(const Plan& pl, const T* W, const T* E, T* Pbar, void* work,
                            int accumulate, hipStream_t s) {
  long long ncp = 1, nsamp = 1;
  for (const DimDev& d : pl.dims) {
    ncp *= d.n;
    nsamp *= d.m;
  }
  // workspace: w(x) [nsamp], ebar / w(x) [nsamp][nout], B^T (ebar / w) [ncp][nout]
  T* D = static_cast<T*>(work);
  T* Et = D + nsamp;
  T* Q = Et + nsamp * pl.nout;
  const int dtype = sizeof(T) == 4 ? SG_F32 : SG_F64;
  const int32_t zero[SG_MAX_NIN] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipError_t e = hipSuccess;
  Plan bp = pl;  // a view: the B-spline adjoint of the same dimensions
  bp.nurbs = 0;
  if constexpr (sizeof(T) == 4) {
    // 1 / w(x) into D, then the matrix-core adjoint scales ebar as it loads
    if (nsamp > 0 && pl.adjm.n_tiles1 > 0 && divide_streaming<T>(pl, W, nullptr, D, s, e)) {
      if (e != hipSuccess) return e;
      e = launch_adjoint_mma(bp, SG_TIER_NATIVE, zero, E, Q, 0, s, D);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(weight_products<T>, dim3(static_cast<unsigned>((ncp + 255) / 256)),
                         dim3(256), 0, s, Q, W, Pbar, ncp, pl.nout, accumulate);
      return hipGetLastError();
    }
  }
  if (nsamp > 0 && !divide_streaming(pl, W, E, Et, s, e)) {
    Plan wp = pl;  // a view: the weights as a one-channel B-spline
    wp.nurbs = 0;
    wp.nout = 1;
    void* dout[1] = {D};
    e = launch_forward(wp, 1, zero, W, nullptr, dout, dtype, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(divide_by_field<T>, dim3(static_cast<unsigned>((nsamp + 255) / 256)),
                       dim3(256), 0, s, E, D, Et, nsamp, pl.nout);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return e;
  e = launch_adjoint(bp, zero, Et, Q, dtype, 0, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(weight_products<T>, dim3(static_cast<unsigned>((ncp + 255) / 256)),
                     dim3(256), 0, s, Q, W, Pbar, ncp, pl.nout, accumulate);
  return hipGetLastError();
}

using Order = std::array<int, SG_MAX_NIN>;

// The closure of the requested orders (every b <= a requested a), ordered by
// total order then lexicographically, so that b - c precedes b.
std::vector<Order> closure_of(const Plan& pl, int nf, const int32_t* orders) {
  std::vector<Order> cl;
  for (int f = 0; f < nf; ++f) {
    Order a{};
    for (int d = 0; d < pl.nin; ++d) a[d] = orders[f * pl.nin + d];
    Order b{};
    for (;;) {  // all b <= a
      if (std::find(cl.begin(), cl.end(), b) == cl.end()) cl.push_back(b);
      int d = 0;
      while (d < pl.nin && b[d] == a[d]) b[d++] = 0;
      if (d == pl.nin) break;
      ++b[d];
    }
  }
  std::sort(cl.begin(), cl.end(), [](const Order& x, const Order& y) {
    int sx = 0, sy = 0;
    for (int d = 0; d < SG_MAX_NIN; ++d) sx += x[d], sy += y[d];
    return sx != sy ? sx < sy : x < y;
  });
  return cl;
}

long long binom(int n, int k) {
  long long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

template <typename T>
hipError_t rational(const Plan& pl, int nf, const int32_t* orders, const T* P, const T* W,
                    T* const* outs, void* work, hipStream_t s) {
  if (const int nc = rat_grad_channels(pl, nf, orders, sizeof(T)))
    return rational_first_order(pl, nf, orders, P, W, outs, nc, s);
  const std::vector<Order> cl = closure_of(pl, nf, orders);
  const int K = static_cast<int>(cl.size());
  long long ncp = 1, nsamp = 1;
  for (const DimDev& d : pl.dims) {
    ncp *= d.n;
    nsamp *= d.m;
  }
  auto index = [&](const Order& o) {
    return static_cast<int>(std::find(cl.begin(), cl.end(), o) - cl.begin());
  };
  RatArgs<T> a{};
  a.nclosure = K;
  a.nout = pl.nout;
  a.nsamp = nsamp;
  a.nfields = nf;
  for (int g = 0; g < K; ++g)
    for (int b = 1; b <= g; ++b) {
      bool below = true;
      long long coef = 1;
      Order r{};
      for (int d = 0; d < pl.nin; ++d) {
        below = below && cl[b][d] <= cl[g][d];
        if (!below) break;
        coef *= binom(cl[g][d], cl[b][d]);
        r[d] = cl[g][d] - cl[b][d];
      }
      if (!below) continue;
      a.tg[a.nterms] = static_cast<unsigned char>(g);
      a.tb[a.nterms] = static_cast<unsigned char>(b);
      a.tr[a.nterms] = static_cast<unsigned char>(index(r));
      a.coef[a.nterms] = static_cast<float>(coef);
      ++a.nterms;
    }
  for (int f = 0; f < nf; ++f) {
    Order o{};
    for (int d = 0; d < pl.nin; ++d) o[d] = orders[f * pl.nin + d];
    a.out_idx[f] = index(o);
    a.out[f] = outs[f];
  }
  // pass 1: homogeneous control points, then A^(b), w^(b) for the closure
  T* H = static_cast<T*>(work);
  T* F = H + ncp * (pl.nout + 1);
  hipLaunchKernelGGL(homogenize<T>, dim3(static_cast<unsigned>((ncp + 255) / 256)), dim3(256),
                     0, s, P, W, H, ncp, pl.nout);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  Plan hp = pl;  // a view: the same tables, no weights, Nout + 1 channels
  hp.nurbs = 0;
  hp.nout = pl.nout + 1;
  std::vector<int32_t> co(static_cast<size_t>(K) * pl.nin);
  std::vector<void*> fo(K);
  for (int g = 0; g < K; ++g) {
    for (int d = 0; d < pl.nin; ++d) co[g * pl.nin + d] = cl[g][d];
    fo[g] = F + static_cast<long long>(g) * nsamp * (pl.nout + 1);
  }
  e = launch_forward(hp, K, co.data(), H, nullptr, fo.data(), sizeof(T) == 4 ? SG_F32 : SG_F64, s);
  if (e != hipSuccess) return e;
  // pass 2: the quotient rule
  a.F = F;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const long long blocks = std::min<long long>((nsamp + kCombineThreads - 1) / kCombineThreads,
                                               static_cast<long long>(cus) * 8);
  const size_t lds = 2 * static_cast<size_t>(K) * kCombineThreads * sizeof(double);
  void* k = combine_kernel_k<T, 1>(K);
  if (!k) return hipErrorInvalidValue;
  void* args[] = {&a};
  return hipLaunchKernel(k, dim3(static_cast<unsigned>(std::max(1LL, blocks))),
                         dim3(kCombineThreads), args, lds, s);
}

}  // namespace

int rational_closure_ok(const Plan& pl, int nf, const int32_t* orders) {
  const std::vector<Order> cl = closure_of(pl, nf, orders);
  if (static_cast<int>(cl.size()) > kMaxClosure) return 0;
  int terms = 0;
  for (size_t g = 0; g < cl.size(); ++g)
    for (size_t b = 1; b <= g; ++b) {
      bool below = true;
      for (int d = 0; d < pl.nin; ++d) below = below && cl[b][d] <= cl[g][d];
      terms += below ? 1 : 0;
    }
  return terms <= kMaxTerms ? 1 : 0;
}

size_t rational_workspace_bytes(const Plan& pl, int nf, const int32_t* orders, size_t elem) {
  long long ncp = 1, nsamp = 1;
  for (const DimDev& d : pl.dims) {
    ncp *= d.n;
    nsamp *= d.m;
  }
  const size_t K = closure_of(pl, nf, orders).size();
  return elem * static_cast<size_t>(pl.nout + 1) *
         (static_cast<size_t>(ncp) + K * static_cast<size_t>(nsamp));
}

hipError_t launch_rational(const Plan& pl, int nf, const int32_t* orders, const void* P,
                           const void* W, void* const* outs, void* work, int dtype,
                           hipStream_t s) {
  if (dtype == SG_F32)
    return rational<float>(pl, nf, orders, static_cast<const float*>(P),
                           static_cast<const float*>(W), reinterpret_cast<float* const*>(outs),
                           work, s);
  return rational<double>(pl, nf, orders, static_cast<const double*>(P),
                          static_cast<const double*>(W), reinterpret_cast<double* const*>(outs),
                          work, s);
}

size_t rational_adjoint_workspace_bytes(const Plan& pl, size_t elem) {
  long long ncp = 1, nsamp = 1;
  for (const DimDev& d : pl.dims) {
    ncp *= d.n;
    nsamp *= d.m;
  }
  return elem * (static_cast<size_t>(nsamp) * (1 + pl.nout) + static_cast<size_t>(ncp) * pl.nout);
}

hipError_t launch_rational_adjoint(const Plan& pl, const void* W, const void* ebar, void* Pbar,
                                   void* work, int dtype, int accumulate, hipStream_t s) {
  if (dtype == SG_F32)
    return rational_adjoint<float>(pl, static_cast<const float*>(W),
                                   static_cast<const float*>(ebar), static_cast<float*>(Pbar),
                                   work, accumulate, s);
  return rational_adjoint<double>(pl, static_cast<const double*>(W),
                                  static_cast<const double*>(ebar), static_cast<double*>(Pbar),
                                  work, accumulate, s);
}

}  // namespace sg
