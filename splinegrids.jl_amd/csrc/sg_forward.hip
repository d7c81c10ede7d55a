// sg_forward.hip -- evaluate!(grid): the tensor-product contraction.
//
// Replaces spline_eval_kernel (src/spline_grid.jl:119-183), which sums all
// (p+1)^Nin terms per sample and read-modify-writes eval in global memory for
// every term.  Here the contraction is sum-factorised into banded mode
// products and every eval element is written exactly once:
//
//   stage 1 (workgroup, LDS): the control-point band window of the tile,
//            contracted along dimension 3 for this workgroup's sample j3:
//            G[k2][k1] = sum_d B3[j3, d] * (w*)P[k1, k2, s3-p3+d]
//   stage 2 (per lane = one sample j1): walk the tile's rows j2 in order.
//            Spans are non-decreasing in j2, so the p2+1 values
//            V[k2] = sum_a B1[j1, a] * G[k2][s1-p1+a] needed by row j2 form a
//            sliding register window that advances only when the span does.
//            eval[j1, j2, j3] = sum_b B2[j2, b] * V[s2-p2+b]
//
// Lanes run along dimension 1 (the fastest axis of eval), so each store
// instruction of a wave writes 64 contiguous elements.  Row-uniform data
// (spans and B2 of row j2) is read with scalar loads.  NURBS carry the
// weight as one more channel; each sample takes one reciprocal of it.
//
// A direct kernel (one thread per sample, all (p+1)^Nin terms, terms in the
// reference's column-major order) covers every case the streaming kernel
// does not: Nin = 1 or > 3, degree > 5 in dimensions 1-2, or sample points
// whose spans are not monotone.
#include "sg_tiles.hpp"

namespace sg {

namespace {

template <typename T>
struct StreamArgs {
  BandSrc<T> src;                // control points of the tile (stage 1)
  const T* tab1[SG_MAX_FIELDS];  // dimension 1 at each field's order [m1][p1+1]
  const int32_t* span1;
  int m1, m2;
  int p1;
  int c0;                        // first channel handled by this launch
  int rows_per_wave;
  int wx;                        // waves side by side along dimension 1
  int ntx, nty;                  // tiles along dimension 1, chunks along 2
  int xcd_map;                   // workgroup numbering mode
  int nt;                        // non-temporal stores of eval
  long long nsamp;               // elements per channel of eval
};

// Per-row data of the row loop: dimension 2's spans, the B2 table of every
// field, and the output of every field.  They are kernel parameters of their
// own, declared __restrict__, so that the compiler may read the row-uniform
// spans and B2 rows with scalar loads (the scalar cache serves the whole
// wave, no vector registers or lane broadcasts are spent on them, and scalar
// loads do not share the vector-memory counter with the stores of eval).
template <typename T>
struct RowPtrs {
  const int32_t* span2;
  const T* tab2[SG_MAX_FIELDS];
  T* out[SG_MAX_FIELDS];
};

template <typename T, int NIN, int P2, int NF, int NC, bool NURBS>
__global__ __launch_bounds__(kThreads) void fwd_stream(
    StreamArgs<T> a, const int32_t* __restrict__ span2, const T* __restrict__ tb0,
    const T* __restrict__ tb1, const T* __restrict__ tb2, const T* __restrict__ tb3,
    T* __restrict__ ou0, T* __restrict__ ou1, T* __restrict__ ou2, T* __restrict__ ou3) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* G = reinterpret_cast<T*>(smem);
  constexpr int NCH = NC + (NURBS ? 1 : 0);
  constexpr int PM = kMaxStreamDegree;
  const T* const tab2[4] = {tb0, tb1, tb2, tb3};
  T* const outs[4] = {ou0, ou1, ou2, ou3};

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
  const int cs = W1 * W2;  // LDS elements per channel

  // ---- stage 1 ----
  stage1_band_window<T, NIN, NC, NURBS>(G, a.src, j3, lo1, W1, lo2, W2);
  __syncthreads();

  // ---- stage 2 ----
  const int wave_x = wave % a.wx, wave_y = wave / a.wx;
  const int rb = r0 + wave_y * a.rows_per_wave;
  const int re = min(rb + a.rows_per_wave, r1);
  const int j1 = j1_0 + kLanes * wave_x + lane;
  // no barrier follows: lanes past the grid, and waves without rows, leave
  if (rb >= re || j1 >= a.m1) return;
  const int p1 = a.p1;
  const int q1b = a.span1[j1] - p1 - lo1;

  // b1 is zero past p1, so the window sums below take all PM+1 terms without
  // a branch (the clamped index re-reads the last element of the band)
  T b1[NF][PM + 1];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int t = 0; t <= PM; ++t) {
      const T v = a.tab1[f][static_cast<size_t>(j1) * (p1 + 1) + min(t, p1)];
      b1[f][t] = t <= p1 ? v : T(0);
    }

  T V[P2 + 1][NF][NCH];
  // V[q] holds k2 = cur - P2 + q
  auto column = [&](int k2, T (&v)[NF][NCH]) {
    const T* g = G + (k2 - lo2) * W1 + q1b;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        T s = T(0);
#pragma unroll
        for (int t = 0; t <= PM; ++t) s += b1[f][t] * g[ch * cs + min(t, p1)];
        v[f][ch] = s;
      }
  };

  int cur = span2[rb];
#pragma unroll
  for (int q = 0; q <= P2; ++q) column(cur - P2 + q, V[q]);

  // Span and B2 rows of each row come from scalar loads issued one row
  // ahead.  Rows go in pairs with two register sets, A and B, so that no
  // prefetched value is copied, and the scalar pointers advance by one row
  // (they stay on the last row at the end instead of being recomputed from a
  // clamped index).
  const int32_t* sp = span2 + rb;
  const T* tp[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) tp[f] = tab2[f] + static_cast<size_t>(rb) * (P2 + 1);
  auto fetch = [&](int& s2, T (&bb)[NF][P2 + 1]) {
    s2 = *sp;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int q = 0; q <= P2; ++q) bb[f][q] = tp[f][q];
  };
  auto advance = [&](bool more) {
    if (more) {
      ++sp;
#pragma unroll
      for (int f = 0; f < NF; ++f) tp[f] += P2 + 1;
    }
  };

  long long orow = static_cast<long long>(j1) + static_cast<long long>(a.m1) * a.m2 * j3 +
                   a.nsamp * a.c0 + static_cast<long long>(a.m1) * rb;
  // one row: bring the window to span s2, then every field's outputs
  auto row = [&](int s2, const T (&bb)[NF][P2 + 1]) {
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
            for (int f = 0; f < NF; ++f)
#pragma unroll
              for (int ch = 0; ch < NCH; ++ch) V[q][f][ch] = V[q + 1][f][ch];
          ++cur;
          column(cur, V[P2]);
        }
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      T num[NCH];
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        T s = T(0);
#pragma unroll
        for (int q = 0; q <= P2; ++q) s += bb[f][q] * V[q][f][ch];
        num[ch] = s;
      }
      const T inv = NURBS ? nurbs_recip(num[NCH - 1]) : T(1);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const T v = NURBS ? num[c] * inv : num[c];
        T* dst = outs[f] + orow + a.nsamp * c;
        if (a.nt)
          __builtin_nontemporal_store(v, dst);
        else
          *dst = v;
      }
    }
    orow += a.m1;
  };

  int sA, sB;
  T bA[NF][P2 + 1], bB[NF][P2 + 1];
  fetch(sA, bA);  // row rb
  int j2 = rb;
  for (; j2 + 1 < re; j2 += 2) {
    advance(true);
    fetch(sB, bB);  // row j2 + 1
    row(sA, bA);
    advance(j2 + 2 < re);
    fetch(sA, bA);  // row j2 + 2 (row j2 + 1 again after the last pair)
    row(sB, bB);
  }
  if (j2 < re) row(sA, bA);
}

// ---- direct kernel (general case) ----

template <typename T>
struct DirectArgs {
  const T* P;
  const T* W;
  T* out;
  const T* tab[SG_MAX_NIN];  // [m_d][p_d+1] at the requested order
  const int32_t* span[SG_MAX_NIN];
  int nin, nout;
  int m[SG_MAX_NIN];
  int p[SG_MAX_NIN];
  long long cps[SG_MAX_NIN];  // control-point strides
  long long ncp, nsamp, nterms;
};

template <typename T, bool NURBS>
__global__ __launch_bounds__(256) void fwd_direct(DirectArgs<T> a) {
  const long long q = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (q >= a.nsamp) return;
  Global<const T>* tb[SG_MAX_NIN];
  int I[SG_MAX_NIN];
  long long base = 0, r = q;
#pragma unroll
  for (int d = 0; d < SG_MAX_NIN; ++d) {
    if (d < a.nin) {
      const int j = static_cast<int>(r % a.m[d]);
      r /= a.m[d];
      base += static_cast<long long>(global_ptr(a.span[d])[j] - a.p[d]) * a.cps[d];
      tb[d] = global_ptr(a.tab[d]) + static_cast<size_t>(j) * (a.p[d] + 1);
    }
  }
  for (int c0 = 0; c0 < a.nout; c0 += 4) {
    T acc[4] = {T(0), T(0), T(0), T(0)};
    T den = T(0);
#pragma unroll
    for (int d = 0; d < SG_MAX_NIN; ++d) I[d] = 0;
    for (long long term = 0; term < a.nterms; ++term) {
      T prod = T(1);
      long long cp = base;
#pragma unroll
      for (int d = 0; d < SG_MAX_NIN; ++d)
        if (d < a.nin) {
          prod *= tb[d][I[d]];
          cp += I[d] * a.cps[d];
        }
      if constexpr (NURBS) {
        prod = prod * a.W[cp];
        den += prod;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c0 + c < a.nout) acc[c] += prod * a.P[cp + a.ncp * (c0 + c)];
      // column-major odometer over the band offsets
      bool carry = true;
#pragma unroll
      for (int d = 0; d < SG_MAX_NIN; ++d) {
        if (carry && d < a.nin) {
          if (++I[d] <= a.p[d])
            carry = false;
          else
            I[d] = 0;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c0 + c < a.nout)
        a.out[q + a.nsamp * (c0 + c)] = NURBS ? acc[c] / den : acc[c];
  }
}

// ---- kernel selection ----

template <typename T>
using StreamKernel = void (*)(StreamArgs<T>, const int32_t*, const T*, const T*, const T*,
                              const T*, T*, T*, T*, T*);

template <typename T, int NIN, int P2>
StreamKernel<T> pick_variant(int nf, int nc, bool nurbs) {
  if (nurbs) {
    switch (nc) {
      case 1: return fwd_stream<T, NIN, P2, 1, 1, true>;
      case 2: return fwd_stream<T, NIN, P2, 1, 2, true>;
      case 3: return fwd_stream<T, NIN, P2, 1, 3, true>;
      case 4: return fwd_stream<T, NIN, P2, 1, 4, true>;
    }
    return nullptr;
  }
  if (nf == 1) {
    switch (nc) {
      case 1: return fwd_stream<T, NIN, P2, 1, 1, false>;
      case 2: return fwd_stream<T, NIN, P2, 1, 2, false>;
      case 3: return fwd_stream<T, NIN, P2, 1, 3, false>;
      case 4: return fwd_stream<T, NIN, P2, 1, 4, false>;
    }
    return nullptr;
  }
  if (nc != 1) return nullptr;
  if (nf == 2) return fwd_stream<T, NIN, P2, 2, 1, false>;
  if (nf == 3) return fwd_stream<T, NIN, P2, 3, 1, false>;
  if (nf == 4) return fwd_stream<T, NIN, P2, 4, 1, false>;
  return nullptr;
}

template <typename T, int NIN>
StreamKernel<T> pick_stream(int p2, int nf, int nc, bool nurbs) {
  switch (p2) {
    case 0: return pick_variant<T, NIN, 0>(nf, nc, nurbs);
    case 1: return pick_variant<T, NIN, 1>(nf, nc, nurbs);
    case 2: return pick_variant<T, NIN, 2>(nf, nc, nurbs);
    case 3: return pick_variant<T, NIN, 3>(nf, nc, nurbs);
    case 4: return pick_variant<T, NIN, 4>(nf, nc, nurbs);
    case 5: return pick_variant<T, NIN, 5>(nf, nc, nurbs);
  }
  return nullptr;
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

}  // namespace

// LDS of fwd_stream: the band window G of nc (+1 weight) channels.
size_t fwd_stream_lds(const Plan& pl, int nc, size_t elem) {
  const Tiling& t = pl.fwd;
  return static_cast<size_t>(nc + (pl.nurbs ? 1 : 0)) * t.w1max * t.w2max * elem;
}

namespace {

template <typename T>
hipError_t forward_stream(const Plan& pl, int nf, const int32_t* orders,
                          const T* P, const T* W, T* const* outs,
                          hipStream_t s) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  const bool nurbs = pl.nurbs != 0;
  const int nin = pl.nin;
  StreamArgs<T> a{};
  a.src.P = P;
  a.src.W = W;
  a.src.n1 = d1.n;
  a.src.n2 = d2.n;
  a.span1 = d1.span;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.p1 = d1.p;
  a.rows_per_wave = pl.fwd.rows_per_wave;
  a.wx = pl.fwd.wx;
  a.ntx = pl.fwd.n_tiles1;
  a.nty = pl.fwd.n_chunks2;
  a.xcd_map = pl.fwd.xcd_map;
  a.nt = pl.nt_stores;
  a.src.ncp = 1;
  a.nsamp = 1;
  for (int d = 0; d < nin; ++d) {
    a.src.ncp *= pl.dims[d].n;
    a.nsamp *= pl.dims[d].m;
  }
  unsigned gz = 1;
  if (nin == 3) {
    const DimDev& d3 = pl.dims[2];
    a.src.span3 = d3.span;
    a.src.p3 = d3.p;
    gz = static_cast<unsigned>(d3.m);
  }
  const dim3 grid(pl.fwd.n_tiles1 * pl.fwd.n_chunks2 * gz);

  // Fields that share the dimension-3 order share stage 1; launch each group
  // in blocks of up to 4 fields (B-splines) and each channel block.
  std::vector<int> done(nf, 0);
  for (int f0 = 0; f0 < nf; ++f0) {
    if (done[f0]) continue;
    std::vector<int> group;
    for (int f = f0; f < nf && static_cast<int>(group.size()) < SG_MAX_FIELDS; ++f) {
      if (done[f]) continue;
      if (nin == 3 && orders[f * nin + 2] != orders[f0 * nin + 2]) continue;
      group.push_back(f);
      done[f] = 1;
      if (nurbs) break;
    }
    const int gnf = static_cast<int>(group.size());
    const T* tab2[SG_MAX_FIELDS] = {nullptr, nullptr, nullptr, nullptr};
    T* out[SG_MAX_FIELDS] = {nullptr, nullptr, nullptr, nullptr};
    for (int g = 0; g < gnf; ++g) {
      const int f = group[g];
      a.tab1[g] = ktab<T>(d1, orders[f * nin + 0]);
      tab2[g] = ktab<T>(d2, orders[f * nin + 1]);
      out[g] = outs[f];
    }
    if (nin == 3) a.src.tab3 = ktab<T>(pl.dims[2], orders[group[0] * nin + 2]);
    int c0 = 0;
    while (c0 < pl.nout) {
      const int left = pl.nout - c0;
      const int nc = (gnf > 1) ? 1 : (left >= 4 ? 4 : left);
      StreamKernel<T> k = (nin == 2) ? pick_stream<T, 2>(d2.p, gnf, nc, nurbs)
                                     : pick_stream<T, 3>(d2.p, gnf, nc, nurbs);
      if (!k) return hipErrorInvalidValue;
      a.c0 = c0;
      a.src.c0 = c0;
      const size_t lds = fwd_stream_lds(pl, nc, sizeof(T));
      hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a, d2.span, tab2[0], tab2[1],
                         tab2[2], tab2[3], out[0], out[1], out[2], out[3]);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      c0 += nc;
    }
  }
  return hipSuccess;
}

template <typename T>
hipError_t forward_direct(const Plan& pl, int nf, const int32_t* orders,
                          const T* P, const T* W, T* const* outs,
                          hipStream_t s) {
  DirectArgs<T> a{};
  a.P = P;
  a.W = W;
  a.nin = pl.nin;
  a.nout = pl.nout;
  a.ncp = 1;
  a.nsamp = 1;
  a.nterms = 1;
  for (int d = 0; d < pl.nin; ++d) {
    const DimDev& dd = pl.dims[d];
    a.span[d] = dd.span;
    a.m[d] = dd.m;
    a.p[d] = dd.p;
    a.cps[d] = a.ncp;
    a.ncp *= dd.n;
    a.nsamp *= dd.m;
    a.nterms *= dd.p + 1;
  }
  const unsigned blocks = static_cast<unsigned>((a.nsamp + 255) / 256);
  for (int f = 0; f < nf; ++f) {
    for (int d = 0; d < pl.nin; ++d) a.tab[d] = ktab<T>(pl.dims[d], orders[f * pl.nin + d]);
    a.out = outs[f];
    if (pl.nurbs)
      hipLaunchKernelGGL((fwd_direct<T, true>), dim3(blocks), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((fwd_direct<T, false>), dim3(blocks), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

hipError_t launch_forward(const Plan& pl, int nf, const int32_t* orders,
                          const void* P, const void* W, void* const* outs,
                          int dtype, hipStream_t s) {
  long long nsamp = 1;
  for (const DimDev& d : pl.dims) nsamp *= d.m;
  if (nsamp == 0) return hipSuccess;
  if (dtype == SG_F32) {
    auto o = reinterpret_cast<float* const*>(outs);
    auto p = static_cast<const float*>(P);
    auto w = static_cast<const float*>(W);
    return pl.stream_ok ? forward_stream<float>(pl, nf, orders, p, w, o, s)
                        : forward_direct<float>(pl, nf, orders, p, w, o, s);
  }
  auto o = reinterpret_cast<double* const*>(outs);
  auto p = static_cast<const double*>(P);
  auto w = static_cast<const double*>(W);
  return pl.stream_ok ? forward_stream<double>(pl, nf, orders, p, w, o, s)
                      : forward_direct<double>(pl, nf, orders, p, w, o, s);
}

}  // namespace sg
