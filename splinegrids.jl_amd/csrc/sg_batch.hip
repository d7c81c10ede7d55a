// sg_batch.hip -- many independent grids in one persistent launch.
//
// A batch holds one plan (FP64 tables, spans) per grid and a single list of
// tiles over all grids -- the same tiles as fwd_stream: 64 samples of
// dimension 1 x a chunk of rows of dimension 2 x one sample of dimension 3
// -- in one queue per XCD.  Grids are assigned whole to the eight queues,
// largest first onto the least loaded one, so each grid's control points are
// fetched into one XCD's L2 rather than into all eight.  A persistent kernel
// with a few workgroups per CU takes tiles from the queue of its own XCD
// (workgroups are dispatched to XCDs round-robin, so workgroup b runs on XCD
// b mod 8) through an atomic counter in global memory, and once that queue
// is empty from the other queues in turn, so uneven grids still balance
// across the chip.  The counters are reset by a small kernel before each
// launch.  Degrees, channel counts, NURBS weights and sizes are per tile at
// run time; the p2+1 register window is selected by a switch over p2.
#include <algorithm>
#include <numeric>
#include <cstring>
#include <sstream>
#include <vector>

#include "sg_plan_host.hpp"
#include "sg_tiles.hpp"

namespace sg {

struct BGrid {
  const float* P;
  const float* W;
  float* out;
  const float* tab1;
  const float* tab2;
  const int32_t* span1;
  const int32_t* span2;
  const float* lists;  // 3-D grids: plane list of each tile row bz, list_stride floats apart
  int nin, nout, nurbs;
  int n1, n2, p1, p2, p3;
  int m1, m2, m3;
  int rows_per_wave;
  int list_stride;
  int vec;  // samples j1 per lane (batch_vec)
  int pad;
  long long ncp, nsamp;
};


namespace {



// A wave walks at most kRowSegs * kLanes rows of a tile: the chunk of a 3-D
// tile has at most kWaves * kMaxBatchRows rows.
constexpr int kMaxBatchRows = 32;
constexpr int kRowSegs = 2;
static_assert(kWaves * kMaxBatchRows <= kRowSegs * kLanes, "row tables too small");

__device__ __attribute__((always_inline)) inline float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Everything a channel-block pass of a tile needs, set up once per tile.
struct TileCtx {
  Global<const float>* P;
  Global<const float>* W;
  Global<float>* out;
  float* smem;
  // 3-D tiles: the plane list (plane_list_host), read with scalar loads;
  // null for 2-D tiles
  Constant<const float>* list;
  int nk;             // ... and its length
  int n1, nch, cs, W1, lo1, lo2, rb, re, j3_0, j3, planes, q1b, m1, m2;
  int pz;  // 3-D tiles: this lane's plane of the tile (its window slab)
  unsigned cs_rcp, w1_rcp;  // udiv_rcp of cs and W1
  long long ncp, nsamp, plane_cp;
  bool active;
  int j1;
  float b1[kMaxStreamDegree + 1];
  // tiles of two samples per lane: the lane's second sample j1b = j1 + 1
  bool activeb;
  int j1b, q1bb;
  float b1b[kMaxStreamDegree + 1];
  // the wave's rows rb .. re - 1 (at most kRowSegs * kLanes of them): lane l
  // holds the span and B2 row of row rb + l in segment 0 and of row
  // rb + kLanes + l in segment 1; the row loop reads them with readlane, so
  // a row costs no memory access before its window sums
  int rspan[kRowSegs];
  float rtab[kRowSegs][kMaxStreamDegree + 1];
};

// Window indices are below kMaxLds / 4 = 2^14, so udiv_fast(n, d) is exact
// for them (n * d < 2^28).


// Stage 1 of a 2-D tile: the window of channels [c0, c0 + NC) (times the
// weight, and the weight last, for NURBS) copied into LDS, U elements per
// thread in flight.
template <int NC, bool NURBS>
__device__ __attribute__((always_inline)) void stage1_2d(const TileCtx& x, int c0) {
  constexpr int NCH = NC + (NURBS ? 1 : 0);
  constexpr int U = NCH <= 2 ? 4 : 2;
  const int cs = x.cs;
  for (int e0 = threadIdx.x; e0 < cs; e0 += U * kThreads) {
    float v[U][NCH];
    int e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      e[u] = min(e0 + u * kThreads, cs - 1);
      const int q2 = udiv_fast(e[u], x.W1, x.w1_rcp), q1 = e[u] - q2 * x.W1;
      const long long cpi = static_cast<long long>(x.lo1 + q1) +
                            static_cast<long long>(x.n1) * (x.lo2 + q2);
      const float w = NURBS ? x.W[cpi] : 1.f;
#pragma unroll
      for (int c = 0; c < NC; ++c) v[u][c] = x.P[cpi + x.ncp * (c0 + c)];
      if constexpr (NURBS) {
#pragma unroll
        for (int c = 0; c < NC; ++c) v[u][c] *= w;
        v[u][NC] = w;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * kThreads < cs) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) x.smem[c * cs + e[u]] = v[u][c];
      }
  }
}

// Stage 1 of a 3-D tile: each window element of every plane,
// G[pz][k2][k1] = sum over the plane list of w[pz] * (W *) P[k1, k2, k3],
// two list entries (and two elements for up to two channels) in flight.
template <int NC, bool NURBS>
__device__ __attribute__((always_inline)) void stage1_3d(const TileCtx& x, int c0) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  constexpr int NCH = NC + (NURBS ? 1 : 0);
  constexpr int U = NCH <= 2 ? 2 : 1;
  const int cs = x.cs, nk = x.nk;
  Constant<const float>* list = x.list;
  for (int e0 = threadIdx.x; e0 < cs; e0 += U * kThreads) {
    long long cpi[U];
    int e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      e[u] = min(e0 + u * kThreads, cs - 1);
      const int q2 = udiv_fast(e[u], x.W1, x.w1_rcp), q1 = e[u] - q2 * x.W1;
      cpi[u] = static_cast<long long>(x.lo1 + q1) + static_cast<long long>(x.n1) * (x.lo2 + q2);
    }
    float acc[U][kWaves][NCH];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int pz = 0; pz < kWaves; ++pz)
#pragma unroll
        for (int c = 0; c < NCH; ++c) acc[u][pz][c] = 0.f;
    for (int i = 0; i < nk; i += 2) {
      float v[2][U][NCH];
      f4 wv[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        Constant<const float>* ent = list + (i + h) * kEntry;
        wv[h] = *reinterpret_cast<Constant<const f4>*>(ent);
        const long long plane = x.plane_cp * reinterpret_cast<Constant<const int>*>(ent)[kEntryK];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const long long off = cpi[u] + plane;
          const float w = NURBS ? x.W[off] : 1.f;
#pragma unroll
          for (int c = 0; c < NC; ++c) v[h][u][c] = x.P[off + x.ncp * (c0 + c)];
          if constexpr (NURBS) {
#pragma unroll
            for (int c = 0; c < NC; ++c) v[h][u][c] *= w;
            v[h][u][NC] = w;
          }
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int pz = 0; pz < kWaves; ++pz)
#pragma unroll
            for (int c = 0; c < NCH; ++c) acc[u][pz][c] += wv[h][pz] * v[h][u][c];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * kThreads < cs) {
#pragma unroll
        for (int pz = 0; pz < kWaves; ++pz)
          if (pz < x.planes) {
            float* Gp = x.smem + static_cast<size_t>(pz) * x.nch * cs;
#pragma unroll
            for (int c = 0; c < NCH; ++c) Gp[c * cs + e[u]] = acc[u][pz][c];
          }
      }
  }
}

// One channel block [c0, c0 + NC) of a tile: stage 1 (band windows of every
// plane in LDS), then each wave's rows.  P1, P2, NC and NURBS are
// compile-time, so a window column is exactly P1 + 1 LDS reads and FMAs per
// channel, and the row loop does exactly NC (+ weight) window sums and stores.
template <int P1, int P2, int NC, bool NURBS, int VEC>
__device__ __attribute__((always_inline)) void tile_pass(const TileCtx& x, int c0) {
  typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
  constexpr int NCH = NC + (NURBS ? 1 : 0);
  const int cs = x.cs;
  // ---- stage 1: band windows, contracted along dimension 3 ----
  if (x.list == nullptr)
    stage1_2d<NC, NURBS>(x, c0);
  else
    stage1_3d<NC, NURBS>(x, c0);
  __syncthreads();

  // ---- stage 2: rows of this wave, sliding window over k2 ----
  // The loop is uniform over the wave (the row tables are read with
  // readlane, which needs every lane's register); lanes outside the tile
  // compute on clamped in-window indices and do not store.
  if (x.rb < x.re) {
    // 3-D tiles: each lane reads its plane's slab; 2-D tiles: the one plane
    const float* G = x.smem + (x.list != nullptr ? static_cast<size_t>(x.pz) * x.nch * cs : 0);
    // V[q][h][ch]: window column k2 = cur - P2 + q of the lane's sample h
    float V[P2 + 1][VEC][NCH];
    // all (P1 + 1) * NCH * VEC window reads of a column in flight before the sums
    auto column = [&](int k2, float (&v)[VEC][NCH]) {
      float t[VEC][NCH][P1 + 1];
#pragma unroll
      for (int h = 0; h < VEC; ++h) {
        const float* gr = G + (k2 - x.lo2) * x.W1 + (h == 0 ? x.q1b : x.q1bb);
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
          for (int q = 0; q <= P1; ++q) t[h][ch][q] = gr[ch * cs + q];
      }
#pragma unroll
      for (int h = 0; h < VEC; ++h)
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
          float s = 0.f;
#pragma unroll
          for (int q = 0; q <= P1; ++q) s += (h == 0 ? x.b1[q] : x.b1b[q]) * t[h][ch][q];
          v[h][ch] = s;
        }
    };
    int cur = __builtin_amdgcn_readlane(x.rspan[0], 0);
#pragma unroll
    for (int q = 0; q <= P2; ++q) column(cur - P2 + q, V[q]);
    long long o = static_cast<long long>(x.j1) + static_cast<long long>(x.m1) * x.m2 * x.j3 +
                  x.nsamp * c0 + static_cast<long long>(x.m1) * x.rb;
#pragma unroll 1
    for (int sgm = 0; sgm < kRowSegs; ++sgm) {
      const int nr = min(kLanes, x.re - x.rb - sgm * kLanes);  // rows of the segment
      if (nr <= 0) break;
      const int rs = sgm == 0 ? x.rspan[0] : x.rspan[1];
      float rt[P2 + 1];
#pragma unroll
      for (int q = 0; q <= P2; ++q) rt[q] = sgm == 0 ? x.rtab[0][q] : x.rtab[1][q];
      for (int l = 0; l < nr; ++l) {
        const int s2 = __builtin_amdgcn_readlane(rs, l);
        float bb[P2 + 1];
#pragma unroll
        for (int q = 0; q <= P2; ++q) bb[q] = readlane_f(rt[q], l);
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
                for (int h = 0; h < VEC; ++h)
#pragma unroll
                  for (int ch = 0; ch < NCH; ++ch) V[q][h][ch] = V[q + 1][h][ch];
              ++cur;
              column(cur, V[P2]);
            }
          }
        }
        float val[VEC][NC];
#pragma unroll
        for (int h = 0; h < VEC; ++h) {
          float num[NCH];
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch) {
            float s = 0.f;
#pragma unroll
            for (int q = 0; q <= P2; ++q) s += bb[q] * V[q][h][ch];
            num[ch] = s;
          }
          const float inv = NURBS ? nurbs_recip(num[NC]) : 1.f;
#pragma unroll
          for (int c = 0; c < NC; ++c) val[h][c] = NURBS ? num[c] * inv : num[c];
        }
        if constexpr (VEC == 2) {
          // both samples: one 8-byte store per channel (4-byte aligned, as
          // global stores allow); the tile's last odd sample alone: 4 bytes
          if (x.activeb) {
#pragma unroll
            for (int c = 0; c < NC; ++c)
              *reinterpret_cast<Global<f2u>*>(x.out + o + x.nsamp * c) = f2u{val[0][c], val[1][c]};
          } else if (x.active) {
#pragma unroll
            for (int c = 0; c < NC; ++c) x.out[o + x.nsamp * c] = val[0][c];
          }
        } else if (x.active) {
#pragma unroll
          for (int c = 0; c < NC; ++c) x.out[o + x.nsamp * c] = val[0][c];
        }
        o += x.m1;
      }
    }
  }
  __syncthreads();  // G is rewritten by the next channel block / tile
}

template <int P1, int P2, bool NURBS, int VEC>
__device__ __attribute__((always_inline)) void tile_pass_nc(const TileCtx& x, int c0, int nc) {
  switch (nc) {
    case 1: tile_pass<P1, P2, 1, NURBS, VEC>(x, c0); break;
    case 2: tile_pass<P1, P2, 2, NURBS, VEC>(x, c0); break;
    case 3: tile_pass<P1, P2, 3, NURBS, VEC>(x, c0); break;
    default:
      // two samples per lane: blocks of at most kVec2Slabs slabs
      if constexpr (VEC == 1 || !NURBS) tile_pass<P1, P2, 4, NURBS, VEC>(x, c0);
      break;
  }
}

// VEC samples j1 per lane: tiles of kLanes * VEC samples (VEC = 2: grids whose
// rows are longer than one wave; a tile stores its part of a row in 8-byte
// pairs)
template <int P1, int P2, int VEC>
__device__ __attribute__((always_inline)) void batch_tile(const BGrid& g, const BTile& t,
                                                          float* smem, int lane, int wave) {
  // the grid's arrays as global pointers (global loads and stores)
  const auto span1 = global_ptr(g.span1), span2 = global_ptr(g.span2);
  const auto tab1 = global_ptr(g.tab1);
  const int chunk_cap = g.rows_per_wave * kWaves;
  constexpr int tw = kLanes * VEC;  // samples j1 per tile
  const int j1_0 = t.bx * tw;
  const int j1_last = min(j1_0 + tw - 1, g.m1 - 1);
  const int r0 = t.by * chunk_cap, r1 = min(r0 + chunk_cap, g.m2);
  constexpr int p1 = P1;
  TileCtx x;
  x.lo1 = span1[j1_0] - p1;
  x.W1 = span1[j1_last] - x.lo1 + 1;
  x.lo2 = span2[r0] - P2;
  const int W2 = span2[r1 - 1] - x.lo2 + 1;
  x.cs = x.W1 * W2;
  x.cs_rcp = udiv_rcp(static_cast<unsigned>(x.cs));
  x.w1_rcp = udiv_rcp(static_cast<unsigned>(x.W1));
  const bool nurbs = g.nurbs != 0;
  x.nch = min(kBC, g.nout) + (nurbs ? 1 : 0);
  // 3-D grids: the tile holds kWaves consecutive samples j3 (from
  // kWaves * t.bz), one per wave, and every wave walks all rows of the chunk;
  // a tile 32 or 16 samples wide (one sample per lane) packs 2 or 4 planes
  // into the lane groups of a wave, which then walks a half or a quarter of
  // the rows (the same rows in every group, so spans and B2 rows stay uniform
  // over the wave);
  // 2-D grids: one plane, the waves split the chunk's rows
  const bool zpack = g.nin == 3;
  x.j3_0 = zpack ? t.bz * kWaves : 0;
  x.planes = zpack ? min(kWaves, g.m3 - x.j3_0) : 1;
  // LDS: [planes][nch][cs] window slabs (weight last; space for kWaves planes
  // is reserved for 3-D grids, as the host's LDS size assumes)
  x.smem = smem;
  x.P = global_ptr(g.P);
  x.W = global_ptr(g.W);
  x.out = global_ptr(g.out);
  x.n1 = g.n1;
  const int width = min(tw, g.m1 - j1_0);  // samples j1 of the tile
  const int groups = zpack && VEC == 1 ? (width <= 16 ? 4 : width <= 32 ? 2 : 1) : 1;
  const int group_lanes = kLanes / groups;
  const int grp = lane / group_lanes, li = lane - grp * group_lanes;
  const int slot0 = wave * groups;  // the wave's first plane slot
  x.pz = zpack ? (slot0 + grp) % kWaves : 0;
  if (zpack) {
    const int part = slot0 / kWaves;  // the wave's share of the rows
    const int len = (r1 - r0 + groups - 1) / groups;
    const bool any = slot0 % kWaves < x.planes;
    x.rb = any ? min(r0 + part * len, r1) : r1;
    x.re = any ? min(r0 + (part + 1) * len, r1) : r1;
  } else {
    x.rb = r0 + wave * g.rows_per_wave;
    x.re = min(x.rb + g.rows_per_wave, r1);
  }
  x.rb = wave_uniform(x.rb);  // scalar row bounds for the row loops
  x.re = wave_uniform(x.re);
  x.j3 = x.j3_0 + x.pz;
  // the row tables, loaded before stage 1 so that their latency hides under
  // it; rows past re - 1 repeat row re - 1 (re >= 1, so every read is a row
  // of the grid, also for a wave with no rows)
  {
    const auto tab2 = global_ptr(g.tab2);
#pragma unroll
    for (int sgm = 0; sgm < kRowSegs; ++sgm) {
      const int r = min(x.rb + sgm * kLanes + lane, x.re - 1);
      x.rspan[sgm] = span2[r];
#pragma unroll
      for (int q = 0; q <= P2; ++q) x.rtab[sgm][q] = tab2[static_cast<size_t>(r) * (P2 + 1) + q];
    }
  }
  x.m1 = g.m1;
  x.m2 = g.m2;
  x.ncp = g.ncp;
  x.nsamp = g.nsamp;
  x.plane_cp = static_cast<long long>(g.n1) * g.n2;
  // lanes outside the tile compute on the tile's last sample (in the window)
  // and do not store
  const bool plane_ok = !zpack || x.pz < x.planes;
  x.active = VEC * li < width && plane_ok;
  x.j1 = x.active ? j1_0 + VEC * li : j1_last;
  x.q1b = span1[x.j1] - p1 - x.lo1;
  x.activeb = VEC == 2 && VEC * li + 1 < width && plane_ok;
  x.j1b = x.activeb ? x.j1 + 1 : x.j1;
  x.q1bb = VEC == 2 ? span1[x.j1b] - p1 - x.lo1 : x.q1b;
  x.list = zpack ? constant_ptr(g.lists) + static_cast<size_t>(t.bz) * g.list_stride : nullptr;
  x.nk = zpack ? reinterpret_cast<Constant<const int>*>(x.list)[kEntryN] : 0;
#pragma unroll
  for (int q = 0; q <= P1; ++q) x.b1[q] = tab1[static_cast<size_t>(x.j1) * (P1 + 1) + q];
  if constexpr (VEC == 2) {
#pragma unroll
    for (int q = 0; q <= P1; ++q) x.b1b[q] = tab1[static_cast<size_t>(x.j1b) * (P1 + 1) + q];
  }
  // two samples per lane: at most kVec2Slabs slabs per block (registers)
  const int bc = VEC == 2 ? kVec2Slabs - (nurbs ? 1 : 0) : kBC;
  for (int c0 = 0; c0 < g.nout; c0 += bc) {
    const int nc = min(bc, g.nout - c0);
    if (nurbs)
      tile_pass_nc<P1, P2, true, VEC>(x, c0, nc);
    else
      tile_pass_nc<P1, P2, false, VEC>(x, c0, nc);
  }
}

template <int P1, int P2>
__device__ __attribute__((always_inline)) void batch_tile_v(const BGrid& g, const BTile& t,
                                                            float* smem, int lane, int wave) {
  // the host gives two samples per lane only to degrees <= kVec2MaxDegree
  if constexpr (P1 <= kVec2MaxDegree && P2 <= kVec2MaxDegree) {
    if (g.vec == 2) {
      batch_tile<P1, P2, 2>(g, t, smem, lane, wave);
      return;
    }
  }
  batch_tile<P1, P2, 1>(g, t, smem, lane, wave);
}

template <int P1>
__device__ __attribute__((always_inline)) void batch_tile_p2(const BGrid& g, const BTile& t,
                                                             float* smem, int lane, int wave) {
  switch (g.p2) {
    case 0: batch_tile_v<P1, 0>(g, t, smem, lane, wave); break;
    case 1: batch_tile_v<P1, 1>(g, t, smem, lane, wave); break;
    case 2: batch_tile_v<P1, 2>(g, t, smem, lane, wave); break;
    case 3: batch_tile_v<P1, 3>(g, t, smem, lane, wave); break;
    case 4: batch_tile_v<P1, 4>(g, t, smem, lane, wave); break;
    default: batch_tile_v<P1, 5>(g, t, smem, lane, wave); break;
  }
}

// At most 128 VGPRs (four waves per SIMD): the row loop keeps a window
// column's LDS reads in flight together and the row tables in registers,
// which outweighs the residency of six workgroups per CU.
__global__ __launch_bounds__(kThreads, 4) void batch_persistent(const BGrid* __restrict__ grids,
                                                             const BTile* __restrict__ tiles,
                                                             int* counters) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* smem = reinterpret_cast<float*>(smem_raw);
  __shared__ int s_tile[2];
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const int home = static_cast<int>(blockIdx.x) % kXcds;
  int visited = 0;  // thread 0: queues found empty so far
  int slot = 0;     // s_tile[slot]: the tile to take next
  if (threadIdx.x == 0) s_tile[0] = pop_tile(counters, home, visited);
  for (;;) {
    __syncthreads();
    // every thread read the same value: the exit is uniform, and with the
    // index in a scalar register the tile and grid records are scalar loads
    // (branches on them stay scalar branches)
    const int ti = wave_uniform(s_tile[slot]);
    if (ti < 0) break;
    // thread 0 claims the following tile now and publishes it once this one
    // is done, so the queue's atomic round trip overlaps this tile's work
    int next = -1;
    if (threadIdx.x == 0) next = pop_tile(counters, home, visited);
    const BTile t = tiles[ti];
    const BGrid g = grids[wave_uniform(t.grid)];
    switch (g.p1) {
      case 0: batch_tile_p2<0>(g, t, smem, lane, wave); break;
      case 1: batch_tile_p2<1>(g, t, smem, lane, wave); break;
      case 2: batch_tile_p2<2>(g, t, smem, lane, wave); break;
      case 3: batch_tile_p2<3>(g, t, smem, lane, wave); break;
      case 4: batch_tile_p2<4>(g, t, smem, lane, wave); break;
      default: batch_tile_p2<5>(g, t, smem, lane, wave); break;
    }
    // every thread read s_tile[slot] before the barriers of the tile; the
    // barrier at the top of the loop publishes the other slot
    if (threadIdx.x == 0) s_tile[slot ^ 1] = next;
    slot ^= 1;
  }
}

}  // namespace

}  // namespace sg

using namespace sg;

struct sg_batch {
  std::vector<Plan*> plans;       // one per grid
  std::vector<int> candidates;    // grids the persistent kernel can take
  std::vector<int> persistent;    // ... and takes under the current tiling
  std::vector<int> fixed_separate;  // grids that always need their own launch
  std::vector<int> separate;      // grids evaluated with their own launches
  std::vector<BGrid> hgrids;
  BGrid* dgrids = nullptr;
  BTile* dtiles = nullptr;
  float* dlists = nullptr;  // plane lists of the 3-D grids' tile rows
  int* counters = nullptr;  // kQueueInts: queue counters, then queue starts
  // MFMA tiers: one argument record per (grid, channel block) and their tiles
  // in queues of their own; mma_ok when every grid takes the tiers
  bool mma_ok = false;
  std::vector<unsigned char> hmargs;
  std::vector<int> mrec_grid, mrec_c0, mrec_nc;
  void* dmargs = nullptr;
  BTile* dtiles_m = nullptr;
  int* counters_m = nullptr;
  int n_tiles_m = 0;
  size_t lds_m = 0;
  std::vector<int> separate_m;  // grids whose windows exceed the tier batch's LDS cap
  int n_tiles = 0;
  size_t lds = 0;
  int rows_per_wave = 16;
  int blocks_per_cu = 3;
  int n_cu = 0;
  int device = 0;
  long long n_samples = 0;
  std::vector<const void*> P, W;
  std::vector<void*> out;
  // set once sg_batch_evaluate has been recorded into a stream capture: a
  // captured launch keeps the tile buffer, grid size and LDS of the tiling
  // at capture time, so the tiling may no longer change (autotune refuses)
  bool captured = false;
};

namespace {

void free_batch(sg_batch* b) {
  if (!b) return;
  for (Plan* p : b->plans) destroy_plan(p);
  if (b->dgrids) (void)hipFree(b->dgrids);
  if (b->dtiles) (void)hipFree(b->dtiles);
  if (b->dlists) (void)hipFree(b->dlists);
  if (b->counters) (void)hipFree(b->counters);
  if (b->dmargs) (void)hipFree(b->dmargs);
  if (b->dtiles_m) (void)hipFree(b->dtiles_m);
  if (b->counters_m) (void)hipFree(b->counters_m);
  delete b;
}



// The plane lists of every 3-D grid's tile rows (kWaves samples j3 each),
// built on the host from the FP32 order-0 table the kernels use (copied back
// once), in one device buffer.
hipError_t upload_plane_lists(sg_batch* b, hipStream_t s) {
  std::vector<float> lists;
  std::vector<size_t> at(b->plans.size(), 0);
  std::vector<float> tab;
  hipError_t e = hipStreamSynchronize(s);  // the plans' tables are built on s
  for (size_t i = 0; i < b->plans.size() && e == hipSuccess; ++i) {
    const Plan& pl = *b->plans[i];
    if (pl.nin != 3) continue;
    const DimDev& d3 = pl.dims[2];
    tab.resize(static_cast<size_t>(d3.m) * (d3.p + 1));
    e = hipMemcpy(tab.data(), d3.ktab32, tab.size() * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) break;
    const int stride = plane_list_floats(d3.p), nz = (d3.m + kWaves - 1) / kWaves;
    at[i] = lists.size();
    lists.resize(lists.size() + static_cast<size_t>(nz) * stride);
    for (int z = 0; z < nz; ++z)
      plane_list_host(d3.h_span.data(), tab.data(), d3.p, z * kWaves,
                      std::min(kWaves, d3.m - z * kWaves), lists.data() + at[i] + size_t(z) * stride);
    b->hgrids[i].list_stride = stride;
  }
  if (e != hipSuccess || lists.empty()) return e;
  e = hipMalloc(&b->dlists, lists.size() * sizeof(float));
  if (e == hipSuccess)
    e = hipMemcpy(b->dlists, lists.data(), lists.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  for (size_t i = 0; i < b->plans.size(); ++i)
    if (b->plans[i]->nin == 3) b->hgrids[i].lists = b->dlists + at[i];
  return hipSuccess;
}

// MFMA tiers: argument records and tiles of every grid (the grids' arrays
// are filled into the records by sg_batch_set_arrays).  Leaves mma_ok false
// when a grid cannot take the tiers.  A workgroup reserves at most the LDS
// that lets kMmaBatchPerCu of them share a CU; grids whose windows need more
// are evaluated with their own tier launches after the persistent one.
constexpr int kMmaBatchPerCu = 4;

int build_mma_tiles(sg_batch* b) {
  b->mma_ok = false;
  for (Plan* pl : b->plans)
    if (!mma_supported(*pl)) return SG_OK;
  const size_t rec = mma_args_size();
  const size_t cap = lds_cap(kMmaBatchPerCu);
  std::vector<std::vector<BTile>> grid_tiles;
  std::vector<long long> grid_cost;
  size_t lds = 16;
  b->separate_m.clear();
  for (size_t gi = 0; gi < b->plans.size(); ++gi) {
    const Plan& pl = *b->plans[gi];
    size_t need = 0;
    for (int c0 = 0; c0 < pl.nout;) {
      const int nc = std::min(pl.nout - c0, pl.nurbs ? 3 : 4);
      need = std::max(need, mma_lds(pl, nc + (pl.nurbs ? 1 : 0), 1));
      c0 += nc;
    }
    if (need > cap) {
      b->separate_m.push_back(static_cast<int>(gi));
      continue;
    }
    const Tiling& t = pl.mma;
    const int m1 = pl.dims[0].m, m2 = pl.dims[1].m;
    const int m3 = pl.nin == 3 ? pl.dims[2].m : 1;
    const int chunk = t.chunk_rows();
    std::vector<BTile> tiles;
    long long cost = 0;
    for (int c0 = 0; c0 < pl.nout;) {
      const int nc = std::min(pl.nout - c0, pl.nurbs ? 3 : 4);
      const int r = static_cast<int>(b->mrec_grid.size());
      b->mrec_grid.push_back(static_cast<int>(gi));
      b->mrec_c0.push_back(c0);
      b->mrec_nc.push_back(nc);
      lds = std::max(lds, mma_lds(pl, nc + (pl.nurbs ? 1 : 0), 1));
      for (int z = 0; z < m3; ++z)
        for (int y = 0; y < t.n_chunks2; ++y)
          for (int x = 0; x < t.n_tiles1; ++x) {
            tiles.push_back(BTile{r, x, y, z});
            cost += static_cast<long long>(std::min(kLanes, m1 - x * kLanes)) *
                    std::min(chunk, m2 - y * chunk) * nc;
          }
      c0 += nc;
    }
    grid_tiles.push_back(std::move(tiles));
    grid_cost.push_back(cost);
  }
  std::vector<BTile> sorted;
  int host_q[kQueueInts] = {};
  pack_queues(grid_tiles, grid_cost, sorted, host_q);
  b->hmargs.assign(b->mrec_grid.size() * rec, 0);
  hipError_t e = hipSuccess;
  if (!sorted.empty()) {
    e = hipMalloc(&b->dtiles_m, sorted.size() * sizeof(BTile));
    if (e == hipSuccess)
      e = hipMemcpy(b->dtiles_m, sorted.data(), sorted.size() * sizeof(BTile),
                    hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&b->counters_m, kQueueInts * sizeof(int));
    if (e == hipSuccess)
      e = hipMemcpy(b->counters_m, host_q, sizeof(host_q), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&b->dmargs, b->hmargs.size());
  }
  if (e != hipSuccess) return hip_fail(e, "batch MFMA tile upload");
  b->n_tiles_m = static_cast<int>(sorted.size());
  b->lds_m = lds;
  b->mma_ok = true;
  return SG_OK;
}

// Tiles of every persistent grid in one queue per XCD.  Each grid takes the
// batch's rows per wave, halved until its tiles fit in the LDS share of one
// of `blocks_per_cu` workgroups; a grid whose tiles never fit is evaluated
// with its own launch instead.  Grids go whole to the least loaded queue,
// largest first (by outputs plus window work); a queue holds its grids in
// that order, each grid's tiles row by row, so neighbouring tiles share
// control-point planes in L2.
int build_tiles(sg_batch* b, int rows_per_wave, int blocks_per_cu) {
  // the kernel's row tables hold the rows of a wave for at most this many
  if (rows_per_wave < 1 || rows_per_wave > kMaxBatchRows)
    return fail(SG_ERR_INVALID, "batch rows per wave out of range");
  const size_t cap = lds_cap(blocks_per_cu);
  std::vector<std::vector<BTile>> grid_tiles;
  std::vector<long long> grid_cost;
  size_t lds = 16;
  std::vector<int> persistent;
  b->separate = b->fixed_separate;
  for (int gi : b->candidates) {
    const Plan& pl = *b->plans[gi];
    const GridTiling t = batch_grid_tiling(pl, rows_per_wave, cap);
    const int rows = t.rows;
    if (t.lds > cap) {
      b->separate.push_back(gi);
      continue;
    }
    persistent.push_back(gi);
    lds = std::max(lds, t.lds);
    const DimDev& d1 = pl.dims[0];
    const DimDev& d2 = pl.dims[1];
    const int chunk = rows * kWaves;
    // 3-D grids: one tile per kWaves samples j3 (a plane per wave)
    const int m3 = pl.nin == 3 ? pl.dims[2].m : 1;
    const int nz = pl.nin == 3 ? (m3 + kWaves - 1) / kWaves : 1;
    std::vector<BTile> tiles;
    long long cost = 0;
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < t.nty; ++y)
        for (int x = 0; x < t.ntx; ++x) {
          tiles.push_back(BTile{gi, x, y, z});
          const long long lanes = std::min(t.tw, d1.m - x * t.tw);
          const long long nrows = std::min(chunk, d2.m - y * chunk);
          const long long planes = pl.nin == 3 ? std::min(kWaves, m3 - z * kWaves) : 1;
          cost += lanes * nrows * planes * pl.nout + 64LL * planes * t.w1 * t.w2;
        }
    grid_tiles.push_back(std::move(tiles));
    grid_cost.push_back(cost);
    b->hgrids[gi].rows_per_wave = rows;
    b->hgrids[gi].vec = t.tw / kLanes;
  }
  std::vector<BTile> sorted;
  int host_q[kQueueInts] = {};
  pack_queues(grid_tiles, grid_cost, sorted, host_q);
  if (b->dtiles) (void)hipFree(b->dtiles);
  b->dtiles = nullptr;
  hipError_t e = hipSuccess;
  if (!sorted.empty()) {
    e = hipMalloc(&b->dtiles, sorted.size() * sizeof(BTile));
    if (e == hipSuccess)
      e = hipMemcpy(b->dtiles, sorted.data(), sorted.size() * sizeof(BTile), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && b->counters)
    e = hipMemcpy(b->counters, host_q, sizeof(host_q), hipMemcpyHostToDevice);
  if (e == hipSuccess && !b->hgrids.empty())
    e = hipMemcpy(b->dgrids, b->hgrids.data(), b->hgrids.size() * sizeof(BGrid),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "batch tile upload");
  b->persistent = persistent;
  b->n_tiles = static_cast<int>(sorted.size());
  b->lds = lds;
  b->rows_per_wave = rows_per_wave;
  b->blocks_per_cu = blocks_per_cu;
  return SG_OK;
}

hipError_t launch_batch(sg_batch* b, hipStream_t s) {
  if (b->n_tiles == 0 || !b->counters) return hipSuccess;  // nothing persistent to run
  hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, b->counters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int blocks = std::max(1, std::min(b->n_tiles, b->blocks_per_cu * b->n_cu));
  hipLaunchKernelGGL(batch_persistent, dim3(blocks), dim3(kThreads), b->lds, s, b->dgrids,
                     b->dtiles, b->counters);
  return hipGetLastError();
}

}  // namespace

extern "C" {

int sg_batch_create(const sg_grid_desc* grids, int32_t n_grids, void* stream, sg_batch** out) {
  if (!out) return fail(SG_ERR_INVALID, "batch output pointer is NULL");
  *out = nullptr;
  if (n_grids < 0 || (n_grids > 0 && !grids)) return fail(SG_ERR_INVALID, "bad grid list");
  const int device = n_grids > 0 ? grids[0].device : 0;
  for (int i = 0; i < n_grids; ++i) {
    int rc = validate_desc(&grids[i]);
    if (rc != SG_OK) return fail(rc, "grid " + std::to_string(i) + ": " + sg_last_error());
    if (grids[i].device != device) return fail(SG_ERR_INVALID, "all grids must share a device");
  }
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != device) (void)hipSetDevice(device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  sg_batch* b = new sg_batch();
  b->device = device;
  auto bail = [&](int rc) {
    free_batch(b);
    if (prev != device) (void)hipSetDevice(prev);
    return rc;
  };
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return bail(hip_fail(e, "hipGetDeviceProperties"));
  b->n_cu = prop.multiProcessorCount;
  b->plans.assign(n_grids, nullptr);
  b->hgrids.assign(n_grids, BGrid{});
  for (int i = 0; i < n_grids; ++i) {
    int rc = build_plan(&grids[i], s, false, &b->plans[i]);
    if (rc != SG_OK) return bail(rc);
    const Plan& pl = *b->plans[i];
    long long ns = 1;
    for (const DimDev& d : pl.dims) ns *= d.m;
    b->n_samples += ns;
    const bool ok = pl.stream_ok && pl.dims[1].p <= kMaxStreamDegree;
    (ok ? b->candidates : b->fixed_separate).push_back(i);
    BGrid& g = b->hgrids[i];
    g.nin = pl.nin;
    g.nout = pl.nout;
    g.nurbs = pl.nurbs;
    const DimDev& d1 = pl.dims[0];
    g.n1 = d1.n;
    g.p1 = d1.p;
    g.m1 = d1.m;
    g.tab1 = d1.ktab32;
    g.span1 = d1.span;
    g.ncp = d1.n;
    g.nsamp = d1.m;
    if (pl.nin >= 2) {
      const DimDev& d2 = pl.dims[1];
      g.n2 = d2.n;
      g.p2 = d2.p;
      g.m2 = d2.m;
      g.tab2 = d2.ktab32;
      g.span2 = d2.span;
      g.ncp *= d2.n;
      g.nsamp *= d2.m;
    }
    g.m3 = 1;
    if (pl.nin == 3) {
      const DimDev& d3 = pl.dims[2];
      g.p3 = d3.p;
      g.m3 = d3.m;
      g.ncp *= d3.n;
      g.nsamp *= d3.m;
    }
  }
  if (n_grids > 0) {
    e = hipMalloc(&b->dgrids, n_grids * sizeof(BGrid));
    if (e == hipSuccess) e = hipMalloc(&b->counters, kQueueInts * sizeof(int));
    if (e != hipSuccess) return bail(hip_fail(e, "batch allocation"));
  }
  if ((e = upload_plane_lists(b, s)) != hipSuccess) return bail(hip_fail(e, "batch plane lists"));
  int rc = build_tiles(b, 16, 3);
  if (rc != SG_OK) return bail(rc);
  if (n_grids > 0 && (rc = build_mma_tiles(b)) != SG_OK) return bail(rc);
  if (n_grids == 0) b->mma_ok = true;  // nothing to evaluate in any tier
  b->P.assign(n_grids, nullptr);
  b->W.assign(n_grids, nullptr);
  b->out.assign(n_grids, nullptr);
  if (prev != device) (void)hipSetDevice(prev);
  *out = b;
  return SG_OK;
}

int sg_batch_destroy(sg_batch* batch) {
  if (!batch) return SG_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != batch->device) (void)hipSetDevice(batch->device);
  const int device = batch->device;
  free_batch(batch);
  if (prev != device) (void)hipSetDevice(prev);
  return SG_OK;
}

int sg_batch_set_arrays(sg_batch* b, const void* const* control_points, const void* const* weights,
                        void* const* evals) {
  if (!b) return fail(SG_ERR_INVALID, "batch is NULL");
  const int n = static_cast<int>(b->plans.size());
  if (n > 0 && (!control_points || !evals)) return fail(SG_ERR_INVALID, "NULL array list");
  for (int i = 0; i < n; ++i) {
    const bool nurbs = b->plans[i]->nurbs != 0;
    const void* w = weights ? weights[i] : nullptr;
    if (nurbs != (w != nullptr))
      return fail(SG_ERR_INVALID, "grid " + std::to_string(i) +
                                      (nurbs ? ": NURBS grid needs weights" : ": B-spline grid takes no weights"));
    if (b->hgrids[i].nsamp > 0 && (!control_points[i] || !evals[i]))
      return fail(SG_ERR_INVALID, "grid " + std::to_string(i) + ": null device pointer");
    b->P[i] = control_points[i];
    b->W[i] = w;
    b->out[i] = evals[i];
    b->hgrids[i].P = static_cast<const float*>(control_points[i]);
    // B-spline grids never read W; it still points at valid memory
    b->hgrids[i].W = static_cast<const float*>(w ? w : control_points[i]);
    b->hgrids[i].out = static_cast<float*>(evals[i]);
  }
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != b->device) (void)hipSetDevice(b->device);
  hipError_t e = hipSuccess;
  if (n > 0)
    e = hipMemcpy(b->dgrids, b->hgrids.data(), n * sizeof(BGrid), hipMemcpyHostToDevice);
  if (e == hipSuccess && b->mma_ok && b->dmargs) {
    const size_t rec = mma_args_size();
    for (size_t r = 0; r < b->mrec_grid.size(); ++r) {
      const int gi = b->mrec_grid[r];
      mma_fill_args(*b->plans[gi], b->mrec_c0[r], b->mrec_nc[r], b->hgrids[gi].P,
                    b->plans[gi]->nurbs ? b->hgrids[gi].W : nullptr, b->hgrids[gi].out,
                    b->hmargs.data() + r * rec);
    }
    e = hipMemcpy(b->dmargs, b->hmargs.data(), b->hmargs.size(), hipMemcpyHostToDevice);
  }
  if (prev != b->device) (void)hipSetDevice(prev);
  return e == hipSuccess ? SG_OK : hip_fail(e, "batch array upload");
}

int sg_batch_evaluate(sg_batch* b, int32_t tier, void* stream) {
  if (!b) return fail(SG_ERR_INVALID, "batch is NULL");
  if (tier != SG_TIER_NATIVE && tier != SG_TIER_SPLIT_F16 && tier != SG_TIER_BF16)
    return fail(SG_ERR_INVALID, "unknown tier");
  if (tier != SG_TIER_NATIVE && !b->mma_ok)
    return fail(SG_ERR_UNSUPPORTED,
                "the split-FP16 and BF16 tiers need every grid of the batch to take them "
                "(Nin 2-3, monotone sample points, degrees <= 5 in dimensions 1-2)");
  for (size_t i = 0; i < b->out.size(); ++i)
    if (!b->out[i] && b->hgrids[i].nsamp > 0)
      return fail(SG_ERR_INVALID, "sg_batch_set_arrays was not called");
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != b->device) (void)hipSetDevice(b->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (s && hipStreamIsCapturing(s, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone)
    b->captured = true;
  if (tier != SG_TIER_NATIVE) {
    hipError_t e = hipSuccess;
    if (b->n_tiles_m > 0) {
      const int per_cu = std::max(1, batch_mma_resident(tier, b->lds_m));
      const int blocks = std::max(1, std::min(b->n_tiles_m, per_cu * b->n_cu));
      e = launch_batch_mma(tier, b->dmargs, b->dtiles_m, b->counters_m, blocks, b->lds_m, s);
    }
    const int32_t zero[SG_MAX_NIN] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t k = 0; e == hipSuccess && k < b->separate_m.size(); ++k) {
      const int i = b->separate_m[k];
      float* outs[1] = {static_cast<float*>(b->out[i])};
      e = launch_forward_mma(*b->plans[i], tier, 1, zero, static_cast<const float*>(b->P[i]),
                             static_cast<const float*>(b->W[i]), outs, s);
    }
    if (prev != b->device) (void)hipSetDevice(prev);
    return e == hipSuccess ? SG_OK : hip_fail(e, "batch MFMA launch");
  }
  hipError_t e = launch_batch(b, s);
  const int32_t zero[SG_MAX_NIN] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t k = 0; e == hipSuccess && k < b->separate.size(); ++k) {
    const int i = b->separate[k];
    void* outs[1] = {b->out[i]};
    e = launch_forward(*b->plans[i], 1, zero, b->P[i], b->W[i], outs, SG_F32, s);
  }
  if (prev != b->device) (void)hipSetDevice(prev);
  return e == hipSuccess ? SG_OK : hip_fail(e, "batch launch");
}

int sg_batch_evaluate_each(sg_batch* b, void* stream) {
  if (!b) return fail(SG_ERR_INVALID, "batch is NULL");
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != b->device) (void)hipSetDevice(b->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int32_t zero[SG_MAX_NIN] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipError_t e = hipSuccess;
  for (size_t i = 0; e == hipSuccess && i < b->plans.size(); ++i) {
    if (!b->out[i] && b->hgrids[i].nsamp > 0) {
      if (prev != b->device) (void)hipSetDevice(prev);
      return fail(SG_ERR_INVALID, "sg_batch_set_arrays was not called");
    }
    void* outs[1] = {b->out[i]};
    e = launch_forward(*b->plans[i], 1, zero, b->P[i], b->W[i], outs, SG_F32, s);
  }
  if (prev != b->device) (void)hipSetDevice(prev);
  return e == hipSuccess ? SG_OK : hip_fail(e, "per-grid launch");
}

int sg_batch_autotune(sg_batch* b, int32_t reps, void* stream) {
  if (!b) return fail(SG_ERR_INVALID, "batch is NULL");
  if (b->captured)
    return fail(SG_ERR_INVALID,
                "sg_batch_autotune after sg_batch_evaluate was captured into a graph: the "
                "captured launch keeps the tiling it was recorded with; autotune before "
                "capturing");
  for (size_t i = 0; i < b->out.size(); ++i)
    if (!b->out[i] && b->hgrids[i].nsamp > 0)
      return fail(SG_ERR_INVALID, "sg_batch_set_arrays was not called");
  if (reps < 1) reps = 3;
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != b->device) (void)hipSetDevice(b->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  float best = 1e30f;
  int best_rows = b->rows_per_wave, best_bpc = b->blocks_per_cu;
  int rc = SG_OK;
  for (int bpc : {2, 3, 4, 6}) {
    for (int rows : {8, 16, 32}) {
      if ((rc = build_tiles(b, rows, bpc)) != SG_OK) continue;
      hipError_t e = launch_batch(b, s);  // warm-up
      if (e == hipSuccess) e = hipEventRecord(t0, s);
      for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_batch(b, s);
      if (e == hipSuccess) e = hipEventRecord(t1, s);
      if (e == hipSuccess) e = hipEventSynchronize(t1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
      if (e != hipSuccess) {
        rc = hip_fail(e, "autotune launch");
        break;
      }
      if (ms < best) {
        best = ms;
        best_rows = rows;
        best_bpc = bpc;
      }
    }
    if (rc != SG_OK && rc != SG_ERR_UNSUPPORTED) break;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  const int rc2 = build_tiles(b, best_rows, best_bpc);
  if (prev != b->device) (void)hipSetDevice(prev);
  if (rc != SG_OK && rc != SG_ERR_UNSUPPORTED) return rc;
  return rc2;
}

int sg_batch_info(const sg_batch* b, int64_t* n_tiles, int64_t* n_samples, int32_t* n_persistent,
                  int32_t* rows_per_wave, int32_t* blocks_per_cu) {
  if (!b) return fail(SG_ERR_INVALID, "batch is NULL");
  if (n_tiles) *n_tiles = b->n_tiles;
  if (n_samples) *n_samples = b->n_samples;
  if (n_persistent) *n_persistent = static_cast<int32_t>(b->persistent.size());
  if (rows_per_wave) *rows_per_wave = b->rows_per_wave;
  if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
  return SG_OK;
}

int sg_batch_describe(const sg_batch* b, char* text, int32_t capacity) {
  if (!b || !text || capacity < 1) return fail(SG_ERR_INVALID, "NULL argument");
  int resident = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &resident, reinterpret_cast<const void*>(batch_persistent), kThreads, b->lds) !=
      hipSuccess)
    resident = -1;
  std::ostringstream os;
  os << "grids=" << b->plans.size() << " persistent=" << b->persistent.size()
     << " separate=" << b->separate.size() << " tiles=" << b->n_tiles
     << " samples=" << b->n_samples << " rows_per_wave=" << b->rows_per_wave
     << " blocks_per_cu=" << b->blocks_per_cu << " resident_per_cu=" << resident
     << " lds_bytes=" << b->lds << " cus=" << b->n_cu << " mfma_tiers=" << (b->mma_ok ? 1 : 0)
     << " mfma_tiles=" << b->n_tiles_m << " mfma_lds_bytes=" << b->lds_m
     << " mfma_separate=" << b->separate_m.size();
  const std::string s = os.str();
  const size_t n = std::min(s.size(), static_cast<size_t>(capacity - 1));
  std::memcpy(text, s.data(), n);
  text[n] = '\0';
  return SG_OK;
}

}  // extern "C"
