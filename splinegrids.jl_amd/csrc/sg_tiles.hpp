// sg_tiles.hpp -- pieces shared by the forward kernels: the control-point
// source of a tile and stage 1 (the tile's band window staged in LDS).
#pragma once

#include "sg_internal.hpp"

namespace sg {

// Where a tile's control points come from.
template <typename T>
struct BandSrc {
  const T* P;            // control points [n1, n2, (n3,) Nout]
  const T* W;            // weights (NURBS) or null
  const T* tab3;         // dimension 3 (3-D) at the requested order [m3][p3+1]
  const int32_t* span3;
  int n1, n2, p3;
  int c0;                // first channel of the launch
  long long ncp;         // elements per channel of P
};

// Stage 1: the band window of control points of a tile, contracted along
// dimension 3 for the tile's samples j3_0 .. j3_0 + nplanes - 1 (PL at most):
//   G[w][c][q2][q1] = sum_d B3[j3_0 + w, d] * (w *) P[lo1+q1, lo2+q2, s3-p3+d, c0+c]
//   G[w][NC][q2][q1] = sum_d B3[j3_0 + w, d] * w[...]                    (NURBS)
// for q1 < W1, q2 < W2 (2-D: one plane, no dimension-3 sum).  Each
// control-point plane k3 in the union of the planes' bands is loaded once
// and serves every plane whose band holds it; each plane sums its terms in
// increasing k3 (the order of d), as a single-plane contraction would.  All
// threads of the block take part; the caller synchronises before reading G.
template <typename T, int NIN, int NC, bool NURBS, int PL>
__device__ inline void stage1_planes(T* G, const BandSrc<T>& b, int j3_0, int nplanes, int lo1,
                                     int W1, int lo2, int W2) {
  constexpr int NCH = NC + (NURBS ? 1 : 0);
  const int cs = W1 * W2;
  int s3[PL], lo3 = 0, hi3 = 0;
  if constexpr (NIN == 3) {
#pragma unroll
    for (int w = 0; w < PL; ++w) s3[w] = b.span3[j3_0 + min(w, nplanes - 1)];
    lo3 = s3[0] - b.p3;
    hi3 = s3[PL - 1];
  }
  for (int idx = threadIdx.x; idx < cs; idx += blockDim.x) {
    const int q2 = idx / W1, q1 = idx - q2 * W1;
    const long long cpi =
        static_cast<long long>(lo1 + q1) + static_cast<long long>(b.n1) * (lo2 + q2);
    T acc[PL][NCH];
#pragma unroll
    for (int w = 0; w < PL; ++w)
#pragma unroll
      for (int c = 0; c < NCH; ++c) acc[w][c] = T(0);
    if constexpr (NIN == 2) {
      if constexpr (NURBS) {
        const T wt = b.W[cpi];
        acc[0][NC] = wt;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[0][c] = wt * b.P[cpi + b.ncp * (b.c0 + c)];
      } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[0][c] = b.P[cpi + b.ncp * (b.c0 + c)];
      }
    } else {
      const long long plane = static_cast<long long>(b.n1) * b.n2;
      for (int k3 = lo3; k3 <= hi3; ++k3) {
        const long long off = cpi + plane * k3;
        T v[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = b.P[off + b.ncp * (b.c0 + c)];
        const T wk = NURBS ? b.W[off] : T(1);
#pragma unroll
        for (int w = 0; w < PL; ++w) {
          const int d = k3 - (s3[w] - b.p3);
          if (w < nplanes && d >= 0 && d <= b.p3) {
            const T bd = b.tab3[static_cast<size_t>(j3_0 + w) * (b.p3 + 1) + d];
            if constexpr (NURBS) {
              const T bw = bd * wk;
              acc[w][NC] += bw;
#pragma unroll
              for (int c = 0; c < NC; ++c) acc[w][c] += bw * v[c];
            } else {
#pragma unroll
              for (int c = 0; c < NC; ++c) acc[w][c] += bd * v[c];
            }
          }
        }
      }
    }
#pragma unroll
    for (int w = 0; w < PL; ++w)
      if (w < nplanes)  // the LDS holds nplanes planes
#pragma unroll
        for (int c = 0; c < NCH; ++c) G[(w * NCH + c) * cs + idx] = acc[w][c];
  }
}

// Tile of a forward workgroup: tile_width samples of dimension 1 from j1_0,
// rows [r0, r1) of dimension 2, sample j3 of dimension 3.
struct FwdTile {
  int j1_0, j1_last, r0, r1, j3;
};

__device__ inline FwdTile fwd_tile(int xcd_map, int ntx, int nty, int tile_width,
                                   int rows_per_chunk, int m1, int m2, int nin) {
  const int L = xcd_logical_block(xcd_map, ntx);
  const int bx = L % ntx, by = (L / ntx) % nty;
  FwdTile t;
  t.j1_0 = bx * tile_width;
  t.j1_last = min(t.j1_0 + tile_width - 1, m1 - 1);
  t.r0 = by * rows_per_chunk;
  t.r1 = min(t.r0 + rows_per_chunk, m2);
  t.j3 = nin == 3 ? L / (ntx * nty) : 0;
  return t;
}

}  // namespace sg
