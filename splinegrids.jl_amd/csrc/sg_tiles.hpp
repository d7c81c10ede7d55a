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
// dimension 3 for sample j3:
//   G[c][q2][q1] = sum_d B3[j3, d] * (w *) P[lo1+q1, lo2+q2, s3-p3+d, c0+c]
//   G[NC][q2][q1] = sum_d B3[j3, d] * w[...]                    (NURBS)
// for q1 < W1, q2 < W2 (2-D: no dimension-3 sum).  All threads of the block
// take part; the caller synchronises before reading G.
template <typename T, int NIN, int NC, bool NURBS>
__device__ inline void stage1_band_window(T* G, const BandSrc<T>& b, int j3, int lo1,
                                          int W1, int lo2, int W2) {
  const int cs = W1 * W2;
  for (int idx = threadIdx.x; idx < cs; idx += blockDim.x) {
    const int q2 = idx / W1, q1 = idx - q2 * W1;
    const long long cpi =
        static_cast<long long>(lo1 + q1) + static_cast<long long>(b.n1) * (lo2 + q2);
    T acc[NC];
    T accw = T(0);
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
    if constexpr (NIN == 2) {
      if constexpr (NURBS) {
        const T w = b.W[cpi];
        accw = w;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = w * b.P[cpi + b.ncp * (b.c0 + c)];
      } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = b.P[cpi + b.ncp * (b.c0 + c)];
      }
    } else {
      const int s3 = b.span3[j3];
      const T* b3 = b.tab3 + static_cast<size_t>(j3) * (b.p3 + 1);
      const long long plane = static_cast<long long>(b.n1) * b.n2;
      for (int d = 0; d <= b.p3; ++d) {
        const long long off = cpi + plane * (s3 - b.p3 + d);
        const T bd = b3[d];
        if constexpr (NURBS) {
          const T bw = bd * b.W[off];
          accw += bw;
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] += bw * b.P[off + b.ncp * (b.c0 + c)];
        } else {
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] += bd * b.P[off + b.ncp * (b.c0 + c)];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) G[c * cs + idx] = acc[c];
    if constexpr (NURBS) G[NC * cs + idx] = accw;
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
