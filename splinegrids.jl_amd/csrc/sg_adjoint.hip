// sg_adjoint.hip -- evaluate_adjoint!(grid): Pbar = M^T ebar, no atomics.
//
// Replaces spline_eval_adjoint_kernel (src/adjoint.jl:1-40), which lets
// every sample scatter into its (p+1)^Nin control points with
// Atomix.@atomic (:35-37).  Here every output is owned by exactly one
// thread, sums run in a fixed order, and results are bit-reproducible.
//
// Streaming path (Nin = 2, 3; monotone spans; p1, p2 <= 5), two kernels:
//   adj_stream  one workgroup per (64-sample tile of dimension 1, chunk of
//               rows of dimension 2, range of samples j3 of dimension 3).
//               For each j3 of its range:
//     A  each lane (sample j1) walks its wave's rows j2 and accumulates
//        U[k2] += B2[j2, b] * ebar[j1, j2, j3] in a sliding window of p2+1
//        registers; an entry that leaves the window is final and goes to
//        the wave's LDS slab.  Row loads are double-buffered in batches of 4.
//     B  the workgroup contracts the slabs along dimension 1: output
//        (k1, k2) sums the lanes whose band holds k1 (B1 transposed per k1
//        in LDS, 16 bytes per LDS read).
//     C  the (k1, k2) slab is added, weighted by B3[j3, a], into a ring of
//        p3+1 k3-slabs in LDS; a k3 that leaves the band of the next j3 is
//        final for this workgroup and is written to its partial window.
//   adj_gather  one thread per control point (k1, k2, k3, c): sums the
//               partial windows of the few workgroups whose windows hold
//               (k1, k2, k3).
// General path (any Nin, any degree, any sample order): adj_csr, one thread
// per control point, walking each dimension's transposed band list.
#include "sg_plan_host.hpp"

namespace sg {

namespace {

template <typename T>
struct AdjStreamArgs {
  const T* E;
  T* part;
  const T* tab1;  // [m1][p1+1] at the requested order
  const T* tab2;  // [m2][P2+1]
  const T* tab3;  // [m3][p3+1] (3-D) or null (2-D)
  const int32_t* span1;
  const int32_t* span2;
  const int32_t* span3;
  int m1, m2, m3, p1, p3;
  int c0, nout;
  int rows_per_wave, j3_per_block;
  int ntx, nty, xcd_map;
  int w1max, w2max, w2wave, w3max;
  long long m12, nsamp;
};

// sum_{l in [l0, l1)} a[l] * b[l] for 16-byte aligned rows, 16 bytes at a time
__device__ inline float dot_lanes(const float* a, const float* b, int l0, int l1) {
  float s = 0.f;
  for (int l = l0; l < l1; l += 4) {
    const float4 x = *reinterpret_cast<const float4*>(a + l);
    const float4 y = *reinterpret_cast<const float4*>(b + l);
    s += x.x * y.x;
    s += x.y * y.y;
    s += x.z * y.z;
    s += x.w * y.w;
  }
  return s;
}
__device__ inline double dot_lanes(const double* a, const double* b, int l0, int l1) {
  double s = 0.0;
  for (int l = l0; l < l1; l += 2) {
    const double2 x = *reinterpret_cast<const double2*>(a + l);
    const double2 y = *reinterpret_cast<const double2*>(b + l);
    s += x.x * y.x;
    s += x.y * y.y;
  }
  return s;
}

constexpr int kRowBatch = 4;  // rows per load batch; two batches are in flight

// LDS rows of 64 lanes are padded by 16 bytes so that rows read by
// neighbouring threads start in different banks.
template <typename T>
constexpr int lane_stride() { return kLanes + 16 / static_cast<int>(sizeof(T)); }

template <typename T, int P2, int NC>
__global__ __launch_bounds__(kThreads) void adj_stream(AdjStreamArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  // S   [4][w2wave][NC][LS]  finished k2 entries of each wave (one j3)
  // BlT [w1max][LS]          B1 of lane l at k1 = lo1 + q1 (0 off the band)
  // R3  [p3+1][w1max*w2max*NC] ring of k3-slabs
  constexpr int LS = lane_stride<T>();
  T* S = reinterpret_cast<T*>(smem);
  T* BlT = S + static_cast<size_t>(kWaves) * a.w2wave * NC * LS;
  T* R3 = BlT + static_cast<size_t>(a.w1max) * LS;
  const int slab = a.w1max * a.w2max * NC;
  // RT [chunk][P2+1] B2 of the chunk's rows; RS [chunk] their spans (LDS
  // reads keep vector-memory loads of tables out of the row loop)
  T* RT = R3 + static_cast<size_t>(a.p3 + 1) * slab;
  const int chunk_cap = a.rows_per_wave * kWaves;
  int* RS = reinterpret_cast<int*>(RT + static_cast<size_t>(chunk_cap) * (P2 + 1));
  int* wlo = RS + chunk_cap;
  int* whi = wlo + kWaves;
  int* llo = whi + kWaves;   // [w1max]
  int* lhi = llo + a.w1max;  // [w1max]

  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const int L = xcd_logical_block(a.xcd_map, a.ntx);
  const int bx = L % a.ntx, by = (L / a.ntx) % a.nty, bz = L / (a.ntx * a.nty);
  const int j1_0 = bx * kLanes;
  const int chunk = a.rows_per_wave * kWaves;
  const int r0 = by * chunk;
  const int r1 = min(r0 + chunk, a.m2);
  const int j3a = bz * a.j3_per_block;
  const int j3b = min(j3a + a.j3_per_block, a.m3);
  const int j1_last = min(j1_0 + kLanes - 1, a.m1 - 1);
  const int p1 = a.p1, p3 = a.p3;
  const int lo1 = a.span1[j1_0] - p1;
  const int W1 = a.span1[j1_last] - lo1 + 1;
  const int lo2 = a.span2[r0] - P2;
  const int W2 = a.span2[r1 - 1] - lo2 + 1;
  const int lo3 = a.tab3 ? a.span3[j3a] - p3 : 0;
  const int nouts = W1 * W2 * NC;

  const bool active = (j1_0 + lane) < a.m1;
  const int j1 = active ? j1_0 + lane : a.m1 - 1;

  for (int i = threadIdx.x; i < W1 * kLanes; i += kThreads) {
    const int q1 = i / kLanes, l = i % kLanes;
    const int jl = j1_0 + l;
    T v = T(0);
    if (jl < a.m1) {
      const int off = lo1 + q1 - (a.span1[jl] - p1);  // slot of k1 in lane l's band
      if (off >= 0 && off <= p1) v = a.tab1[static_cast<size_t>(jl) * (p1 + 1) + off];
    }
    BlT[q1 * LS + l] = v;
  }
  // lanes whose band holds k1: s1 in [k1, k1 + p1] (spans monotone in lanes)
  for (int q1 = threadIdx.x; q1 < W1; q1 += kThreads) {
    const int k1 = lo1 + q1;
    int lo = 0, hi = kLanes;  // first lane with s1 >= k1
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (a.span1[min(j1_0 + mid, a.m1 - 1)] < k1) lo = mid + 1; else hi = mid;
    }
    int lo_b = lo, hi_b = kLanes;  // first lane with s1 > k1 + p1
    while (lo_b < hi_b) {
      const int mid = (lo_b + hi_b) >> 1;
      if (a.span1[min(j1_0 + mid, a.m1 - 1)] <= k1 + p1) lo_b = mid + 1; else hi_b = mid;
    }
    llo[q1] = lo & ~3;  // whole 16-byte groups; BlT is 0 off the band
    lhi[q1] = (lo_b + 3) & ~3;
  }
  for (int i = threadIdx.x; i < (p3 + 1) * slab; i += kThreads) R3[i] = T(0);
  {
    const size_t rbase = static_cast<size_t>(r0) * (P2 + 1);
    stage_to_lds<1, 1>(r1 - r0, [&](int, int i) { return a.span2[r0 + i]; },
                       [&](int, int i, int v) { RS[i] = v; });
    stage_to_lds<1, 4>((r1 - r0) * (P2 + 1), [&](int, int i) { return a.tab2[rbase + i]; },
                       [&](int, int i, T v) { RT[i] = v; });
  }

  // partial window of this workgroup: [w3max][nout][w2max][w1max]
  const size_t plane = static_cast<size_t>(a.w1max) * a.w2max;
  T* part = a.part + static_cast<size_t>(L) * a.w3max * a.nout * plane;
  // k3's ring slab is final: write it and clear it (each thread touches only
  // the ring entries it accumulates, so no barrier is needed around this)
  auto flush = [&](int k3) {
    T* ring = R3 + static_cast<size_t>((k3 - lo3) % (p3 + 1)) * slab;
    T* dst = part + static_cast<size_t>(k3 - lo3) * a.nout * plane;
    for (int idx = threadIdx.x; idx < nouts; idx += kThreads) {
      const int q1 = idx % W1, rest = idx / W1, q2 = rest % W2, c = rest / W2;
      dst[(a.c0 + c) * plane + static_cast<size_t>(q2) * a.w1max + q1] = ring[idx];
      ring[idx] = T(0);
    }
  };

  const int rb = r0 + wave * a.rows_per_wave;
  const int re = min(rb + a.rows_per_wave, r1);
  T* Sw = S + static_cast<size_t>(wave) * a.w2wave * NC * LS;
  int cur3 = a.tab3 ? a.span3[j3a] : 0;
  for (int j3 = j3a; j3 < j3b; ++j3) {
    const int s3 = a.tab3 ? a.span3[j3] : 0;
    while (cur3 < s3) {
      flush(cur3 - p3);
      ++cur3;
    }
    __syncthreads();  // BlT/llo ready (first j3); S free again (later j3)

    // ---- A: stream the wave's rows, sliding window over k2 ----
    if (rb < re) {
      int cur = wave_uniform(RS[rb - r0]);
      const int w0 = cur - P2;
      T U[P2 + 1][NC];
#pragma unroll
      for (int q = 0; q <= P2; ++q)
#pragma unroll
        for (int c = 0; c < NC; ++c) U[q][c] = T(0);
      const long long ebase = static_cast<long long>(j1) + a.m12 * j3 + a.nsamp * a.c0;
      auto load = [&](T (&e)[kRowBatch][NC], int jb) {
#pragma unroll
        for (int r = 0; r < kRowBatch; ++r) {
          const bool ok = active && (jb + r < re);
          const long long eoff = ebase + static_cast<long long>(a.m1) * min(jb + r, re - 1);
#pragma unroll
          for (int c = 0; c < NC; ++c) e[r][c] = ok ? a.E[eoff + a.nsamp * c] : T(0);
        }
      };
      auto consume = [&](const T (&e)[kRowBatch][NC], int jb) {
#pragma unroll
        for (int r = 0; r < kRowBatch; ++r) {
          const int j2 = jb + r;
          if (j2 >= re) break;
          const int s2 = wave_uniform(RS[j2 - r0]);
          while (cur < s2) {
            const int row = cur - P2 - w0;
#pragma unroll
            for (int c = 0; c < NC; ++c) Sw[(row * NC + c) * LS + lane] = U[0][c];
#pragma unroll
            for (int q = 0; q < P2; ++q)
#pragma unroll
              for (int c = 0; c < NC; ++c) U[q][c] = U[q + 1][c];
#pragma unroll
            for (int c = 0; c < NC; ++c) U[P2][c] = T(0);
            ++cur;
          }
          const T* b2p = RT + static_cast<size_t>(j2 - r0) * (P2 + 1);
          T bb[P2 + 1];
#pragma unroll
          for (int q = 0; q <= P2; ++q) bb[q] = wave_uniform(b2p[q]);
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int q = 0; q <= P2; ++q) U[q][c] += bb[q] * e[r][c];
        }
      };
      T ea[kRowBatch][NC], eb[kRowBatch][NC];
      load(ea, rb);
      for (int jb = rb; jb < re; jb += 2 * kRowBatch) {
        if (jb + kRowBatch < re) load(eb, jb + kRowBatch);
        consume(ea, jb);
        if (jb + 2 * kRowBatch < re) load(ea, jb + 2 * kRowBatch);
        if (jb + kRowBatch < re) consume(eb, jb + kRowBatch);
      }
#pragma unroll
      for (int q = 0; q <= P2; ++q) {
        const int row = cur - P2 + q - w0;
#pragma unroll
        for (int c = 0; c < NC; ++c) Sw[(row * NC + c) * LS + lane] = U[q][c];
      }
      if (lane == 0) {
        wlo[wave] = w0;
        whi[wave] = cur;
      }
    } else if (lane == 0) {
      wlo[wave] = 1;
      whi[wave] = 0;  // empty window
    }
    __syncthreads();

    // ---- B + C: contract along dimension 1, add into the k3 ring ----
    for (int idx = threadIdx.x; idx < nouts; idx += kThreads) {
      const int q1 = idx % W1;
      const int rest = idx / W1;
      const int q2 = rest % W2;
      const int c = rest / W2;
      const int k2 = lo2 + q2;
      const T* bt = BlT + q1 * LS;
      const int l0 = llo[q1], l1 = lhi[q1];
      T sum = T(0);
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        if (k2 < wlo[w] || k2 > whi[w]) continue;
        const T* sw = S + ((static_cast<size_t>(w) * a.w2wave + (k2 - wlo[w])) * NC + c) * LS;
        sum += dot_lanes(bt, sw, l0, l1);
      }
      if (a.tab3) {
        const T* b3 = a.tab3 + static_cast<size_t>(j3) * (p3 + 1);
        for (int t = 0; t <= p3; ++t) {
          const int k3 = s3 - p3 + t;
          R3[static_cast<size_t>((k3 - lo3) % (p3 + 1)) * slab + idx] += b3[t] * sum;
        }
      } else {
        R3[idx] += sum;
      }
    }
  }
  for (int k3 = cur3 - p3; k3 <= cur3; ++k3) flush(k3);
}

template <typename T>
struct AdjGatherArgs {
  const T* part;
  T* out;
  // per dimension d and control point k: the windows that hold k, as
  // offsets [ptr[k], ptr[k+1]) whose sum over the dimensions (plus the
  // channel's) is the element offset in part
  const int32_t* ptr1;
  const long long* off1;
  const int32_t* ptr2;
  const long long* off2;
  const int32_t* ptr3;
  const long long* off3;
  int n1, n2, nout;
  int accumulate;
  int fast;  // udiv_fast is exact for every window number of this plan
  long long ncp, chan;
};

// Lanes per control point in adj_gather: lane `sub` of a group sums the
// windows sub, sub + kGatherLanes, ... (kBatch loads in flight), then the
// group adds its partial sums with a fixed butterfly, so every result is
// the same sum in the same order on every run.
constexpr int kGatherLanes = 4;

template <typename T>
__global__ __launch_bounds__(256) void adj_gather(AdjGatherArgs<T> a) {
  const long long gidx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  const long long idx = gidx / kGatherLanes;
  const int sub = static_cast<int>(gidx % kGatherLanes);
  const bool live = idx < a.ncp * a.nout;
  const long long cidx = live ? idx : 0;  // dead lanes join the butterfly
  const int c = static_cast<int>(cidx / a.ncp);
  long long r = cidx - static_cast<long long>(c) * a.ncp;
  const int k1 = static_cast<int>(r % a.n1);
  r /= a.n1;
  const int k2 = static_cast<int>(r % a.n2);
  const int k3 = static_cast<int>(r / a.n2);
  const int e1a = a.ptr1[k1], n1e = a.ptr1[k1 + 1] - e1a;
  const int e2a = a.ptr2[k2], n2e = a.ptr2[k2 + 1] - e2a;
  const int e3a = a.ptr3[k3], n3e = a.ptr3[k3 + 1] - e3a;
  // the n1e * n2e * n3e windows that hold (k1, k2, k3), dimension 1 fastest:
  // window f is (i1, i2, i3) with f = i1 + n1e (i2 + n2e i3), split with
  // udiv_fast where the plan's counts keep it exact (a.fast, uniform)
  const int n12 = n1e * n2e;
  const unsigned r1 = udiv_rcp(static_cast<unsigned>(n1e)), r12 = udiv_rcp(static_cast<unsigned>(n12));
  const long long base = static_cast<long long>(c) * a.chan;
  constexpr int kBatch = 4;
  const int total = live ? n12 * n3e : 0;
  T sum = T(0);
  for (int f0 = sub; f0 < total; f0 += kBatch * kGatherLanes) {
    T v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int f = min(f0 + u * kGatherLanes, total - 1);
      const int i3 = a.fast ? udiv_fast(f, n12, r12) : f / n12, f12 = f - i3 * n12;
      const int i2 = a.fast ? udiv_fast(f12, n1e, r1) : f12 / n1e, i1 = f12 - i2 * n1e;
      v[u] = a.part[base + a.off1[e1a + i1] + a.off2[e2a + i2] + a.off3[e3a + i3]];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (f0 + u * kGatherLanes < total) sum += v[u];
  }
#pragma unroll
  for (int m = kGatherLanes / 2; m >= 1; m /= 2) sum += __shfl_xor(sum, m, kGatherLanes);
  if (live && sub == 0) a.out[idx] = a.accumulate ? a.out[idx] + sum : sum;
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

}  // namespace


namespace {

template <typename T>
int max_nc(const Plan& pl) {
  for (int nc = 4; nc >= 1; --nc)
    if (adjoint_stream_lds(pl, nc, sizeof(T)) <= static_cast<size_t>(kMaxLds)) return nc;
  return 0;
}

template <typename T>
hipError_t adjoint_stream(const Plan& pl, const int32_t* orders, const T* E,
                          T* Pbar, int accumulate, hipStream_t s) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  const int nin = pl.nin;
  const Tiling& tl = pl.adj;
  AdjStreamArgs<T> a{};
  a.E = E;
  a.part = static_cast<T*>(pl.work);
  a.tab1 = ktab<T>(d1, orders[0]);
  a.tab2 = ktab<T>(d2, orders[1]);
  a.span1 = d1.span;
  a.span2 = d2.span;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.m3 = 1;
  a.p1 = d1.p;
  a.p3 = 0;
  if (nin == 3) {
    const DimDev& d3 = pl.dims[2];
    a.tab3 = ktab<T>(d3, orders[2]);
    a.span3 = d3.span;
    a.m3 = d3.m;
    a.p3 = d3.p;
  }
  a.nout = pl.nout;
  a.rows_per_wave = tl.rows_per_wave;
  a.j3_per_block = tl.j3_per_block;
  a.ntx = tl.n_tiles1;
  a.nty = tl.n_chunks2;
  a.xcd_map = tl.xcd_map;
  a.w1max = tl.w1max;
  a.w2max = tl.w2max;
  a.w2wave = tl.w2wave_max;
  a.w3max = tl.w3max;
  a.m12 = static_cast<long long>(d1.m) * d2.m;
  a.nsamp = a.m12 * a.m3;
  const dim3 grid(tl.n_tiles1 * tl.n_chunks2 * tl.n_z);
  const int ncmax = max_nc<T>(pl);
  if (ncmax == 0) return hipErrorInvalidValue;
  for (int c0 = 0; c0 < pl.nout;) {
    const int nc = std::min(ncmax, pl.nout - c0);
    AdjKernel<T> k = pick_adj<T>(d2.p, nc);
    if (!k) return hipErrorInvalidValue;
    a.c0 = c0;
    hipLaunchKernelGGL(k, grid, dim3(kThreads), adjoint_stream_lds(pl, nc, sizeof(T)), s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    c0 += nc;
  }

  return launch_adjoint_gather(pl, tl, pl.work, Pbar, sizeof(T) == 4 ? SG_F32 : SG_F64,
                               accumulate, s);
}

template <typename T>
hipError_t gather(const Plan& pl, const Tiling& tl, const T* part, T* Pbar, int accumulate,
                  hipStream_t s) {
  AdjGatherArgs<T> g{};
  g.part = part;
  g.out = Pbar;
  g.ptr1 = tl.gptr[0];
  g.off1 = tl.goff[0];
  g.ptr2 = tl.gptr[1];
  g.off2 = tl.goff[1];
  g.ptr3 = tl.gptr[2];
  g.off3 = tl.goff[2];
  g.n1 = pl.dims[0].n;
  g.n2 = pl.dims[1].n;
  g.nout = pl.nout;
  g.chan = tl.gchan;
  g.fast = tl.gfast;
  g.accumulate = accumulate;
  g.ncp = 1;
  for (int d = 0; d < pl.nin; ++d) g.ncp *= pl.dims[d].n;
  const long long total = g.ncp * pl.nout * kGatherLanes;
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

hipError_t launch_adjoint_gather(const Plan& pl, const Tiling& tl, const void* part, void* Pbar,
                                 int dtype, int accumulate, hipStream_t s) {
  if (dtype == SG_F32)
    return gather<float>(pl, tl, static_cast<const float*>(part), static_cast<float*>(Pbar),
                         accumulate, s);
  return gather<double>(pl, tl, static_cast<const double*>(part), static_cast<double*>(Pbar),
                        accumulate, s);
}

size_t adjoint_work_bytes(const Plan& pl) {
  const Tiling& t = pl.adj;
  if (t.n_tiles1 == 0) return 0;
  const size_t units = static_cast<size_t>(t.n_tiles1) * t.n_chunks2 * t.n_z;
  return units * t.w3max * pl.nout * t.w1max * t.w2max * sizeof(double);
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
  const bool stream = pl.stream_ok && pl.adj.n_tiles1 > 0 && nsamp > 0 &&
                      (dtype == SG_F32 ? max_nc<float>(pl) : max_nc<double>(pl)) > 0;
  if (dtype == SG_F32) {
    auto e = static_cast<const float*>(ebar);
    auto o = static_cast<float*>(Pbar);
    if (pl.adjm.n_tiles1 > 0 && nsamp > 0)
      return launch_adjoint_mma(pl, SG_TIER_NATIVE, orders, e, o, accumulate, s);
    return stream ? adjoint_stream<float>(pl, orders, e, o, accumulate, s)
                  : adjoint_csr<float>(pl, orders, e, o, accumulate, s);
  }
  auto e = static_cast<const double*>(ebar);
  auto o = static_cast<double*>(Pbar);
  return stream ? adjoint_stream<double>(pl, orders, e, o, accumulate, s)
                : adjoint_csr<double>(pl, orders, e, o, accumulate, s);
}

}  // namespace sg
