// sg_rational.hip -- derivative fields of a NURBS grid's rational map
// R = A / w, A = sum N w P, w = sum N w, and its adjoint in the control
// points.  The reference rejects both (src/validation.jl:81-88,
// src/adjoint.jl:57).
//
// Adjoint: with the weights fixed R is linear in P, R = B (w P) / w(x), so
// Pbar_k = w_k [B^T (ebar / w(x))]_k: the weight field by the B-spline
// forward, a division, the B-spline adjoint, a product.
//
// Two passes.  (1) The homogeneous control points (w P_c, w) go through the
// B-spline forward (launch_forward on a view of the plan without weights and
// with Nout + 1 channels) for every order b below a requested one, giving
// A^(b) and w^(b).  (2) One thread per (sample, channel) applies Leibniz's
// rule to A = w R in FP64 (The NURBS Book, eq. 4.8 / Alg. A4.4):
//   R^(g) = (A^(g) - sum_{0 < b <= g} C(g, b) w^(b) R^(g-b)) / w,
// C(g, b) = prod_d binom(g_d, b_d), over the orders g of the closure in
// increasing total order, and writes the requested fields.
#include <algorithm>
#include <array>
#include <vector>

#include "sg_internal.hpp"

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

// ebar / w(x), channel by channel (the weight field w(x) in D).
template <typename T>
__global__ void divide_by_field(const T* E, const T* D, T* out, long long nsamp, int nout) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nsamp) return;
  const double inv = 1.0 / static_cast<double>(D[i]);
  for (int c = 0; c < nout; ++c)
    out[i + nsamp * c] = static_cast<T>(static_cast<double>(E[i + nsamp * c]) * inv);
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
  Plan wp = pl;  // a view: the weights as a one-channel B-spline
  wp.nurbs = 0;
  wp.nout = 1;
  void* dout[1] = {D};
  hipError_t e = launch_forward(wp, 1, zero, W, nullptr, dout, dtype, s);
  if (e != hipSuccess) return e;
  if (nsamp > 0) {
    hipLaunchKernelGGL(divide_by_field<T>, dim3(static_cast<unsigned>((nsamp + 255) / 256)),
                       dim3(256), 0, s, E, D, Et, nsamp, pl.nout);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  Plan bp = pl;  // a view: the B-spline adjoint of the same dimensions
  bp.nurbs = 0;
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
