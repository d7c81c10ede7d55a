// sg_plan_host.hip -- host-side plan arithmetic (sg_plan_host.hpp): no HIP
// runtime calls, no kernels.  Built into the library, and on its own under
// the sanitizers by the CPU test of tests/native/plan_check.cpp.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "sg_plan_host.hpp"

namespace sg {

int env_int(const char* name, int fallback) {
  const char* v = std::getenv(name);
  if (!v || !*v) return fallback;
  const int x = std::atoi(v);
  return x > 0 ? x : fallback;
}

// Rows of dimension 1 whose length is not a multiple of 128 bytes make
// neighbouring tiles share a cache line; keep such tiles on one XCD.
int default_xcd_map(const Plan& pl) {
  return (pl.dims[0].m * sizeof(float)) % 128 == 0 ? 0 : 2;
}

void compute_tiling(const Plan& pl, Tiling& t, int rows_per_wave, int wx) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  const std::vector<int32_t>& s1 = d1.h_span;
  const std::vector<int32_t>& s2 = d2.h_span;
  t.rows_per_wave = rows_per_wave;
  t.wx = wx;
  const int tw = t.tile_width();
  t.n_tiles1 = (d1.m + tw - 1) / tw;
  t.w1max = 0;
  for (int i = 0; i < t.n_tiles1; ++i) {
    const int a = i * tw, b = std::min(a + tw, d1.m) - 1;
    t.w1max = std::max(t.w1max, s1[b] - s1[a] + d1.p + 1);
  }
  const int chunk = t.chunk_rows();
  t.n_chunks2 = (d2.m + chunk - 1) / chunk;
  t.w2max = 0;
  t.w2wave_max = 0;
  for (int c = 0; c < t.n_chunks2; ++c) {
    const int a = c * chunk, b = std::min(a + chunk, d2.m) - 1;
    t.w2max = std::max(t.w2max, s2[b] - s2[a] + d2.p + 1);
    for (int w = 0; w < kWaves / wx; ++w) {
      const int ra = a + w * rows_per_wave;
      const int rb = std::min(ra + rows_per_wave, b + 1) - 1;
      if (ra > rb) continue;
      t.w2wave_max = std::max(t.w2wave_max, s2[rb] - s2[ra] + d2.p + 1);
    }
  }
}

// the band window of fwd_stream; the MFMA kernels add their chunk's B2 rows
size_t forward_lds_worst(const Plan& pl, const Tiling& t) {
  const int nc = std::min(pl.nout, 4) + (pl.nurbs ? 1 : 0);
  const size_t p2 = static_cast<size_t>(pl.dims[1].p);
  return (static_cast<size_t>(nc) * t.w1max * t.w2max + t.chunk_rows() * (p2 + 1)) *
             sizeof(double) +
         static_cast<size_t>(t.chunk_rows()) * sizeof(int);
}

// Samples of dimension 3 per adjoint workgroup: enough workgroups to fill
// the chip (about 8 per CU), as few as possible beyond that so that each
// k3 window is shared by many j3.
void compute_ztiling(const Plan& pl, Tiling& t, int j3_per_block) {
  if (pl.nin != 3) {
    t.j3_per_block = 1;
    t.n_z = 1;
    t.w3max = 1;
    return;
  }
  const DimDev& d3 = pl.dims[2];
  if (j3_per_block <= 0) {
    const long long xy = static_cast<long long>(t.n_tiles1) * t.n_chunks2;
    const long long target = 1024;
    j3_per_block = static_cast<int>(std::max<long long>(1, (xy * d3.m) / target));
  }
  t.j3_per_block = std::min(j3_per_block, d3.m);
  t.n_z = (d3.m + t.j3_per_block - 1) / t.j3_per_block;
  t.w3max = 0;
  for (int z = 0; z < t.n_z; ++z) {
    const int a = z * t.j3_per_block, b = std::min(a + t.j3_per_block, d3.m) - 1;
    t.w3max = std::max(t.w3max, d3.h_span[b] - d3.h_span[a] + d3.p + 1);
  }
}

// adj_stream: each wave's partial window of its rows, the dimension-1 B
// rows, the (p3 + 1)-slab ring of the workgroup's window, the chunk's B2
// rows and the row / span bookkeeping
size_t adjoint_stream_lds(const Plan& pl, int nc, size_t elem) {
  const Tiling& t = pl.adj;
  const int p3 = pl.nin == 3 ? pl.dims[2].p : 0;
  const size_t ls = kLanes + 16 / elem;
  const size_t chunk = static_cast<size_t>(t.chunk_rows());
  return static_cast<size_t>(kWaves) * t.w2wave_max * nc * ls * elem +
         static_cast<size_t>(t.w1max) * ls * elem +
         static_cast<size_t>(p3 + 1) * t.w1max * t.w2max * nc * elem +
         chunk * (pl.dims[1].p + 1) * elem +
         (chunk + 2 * kWaves + 2 * t.w1max) * sizeof(int);
}

// mma_stream: G for nch channels, the chunk's B2 rows of nf fields and its
// spans
size_t mma_lds(const Plan& pl, int nch, int nf) {
  const Tiling& t = pl.mma;
  return (static_cast<size_t>(nch) * t.w1max * t.w2max +
          static_cast<size_t>(nf) * t.chunk_rows() * (pl.dims[1].p + 1)) * sizeof(float) +
         static_cast<size_t>(t.chunk_rows()) * sizeof(int);
}

namespace {

size_t adjoint_lds_min(const Plan& pl) {  // one channel, FP64
  return adjoint_stream_lds(pl, 1, sizeof(double));
}

// Matrix-core adjoint (FP32): 64-sample tiles along dimension 1 and chunks
// of 64, 32 or 16 rows along dimension 2 whose band windows fit one 16 x 16
// MFMA tile; samples of dimension 3 per task chosen so that all tasks (one
// wave each) are resident at once.  Leaves pl.adjm empty when the bands do
// not fit.
void plan_adjoint_mma(Plan& pl, const std::function<int(const Plan&)>& resident_waves) {
  Tiling& t = pl.adjm;
  t = Tiling();
  t.tile_w = kLanes;
  bool fits = false;
  for (int ch : {64, 32, 16}) {
    t.chunk_h = ch;
    compute_tiling(pl, t, ch / kWaves);
    if (t.w1max > 16) break;
    if (t.w2max <= 16) {
      fits = true;
      break;
    }
  }
  if (!fits) {
    t = Tiling();
    return;
  }
  int j3 = env_int("SG_ADJM_J3", 0);
  if (j3 <= 0 && pl.nin == 3) {
    const long long xy = static_cast<long long>(t.n_tiles1) * t.n_chunks2;
    const long long m3 = pl.dims[2].m;
    // about one task (wave) per SIMD: with seven row blocks in flight a wave
    // keeps HBM busy, and fewer, longer tasks leave fewer partial windows to
    // gather (400 x 400 x 200 samples: 94 us here against 100 us with every
    // resident slot filled)
    const long long slots = std::max(resident_waves(pl) / 2, 1);
    j3 = static_cast<int>(std::max<long long>(1, (xy * m3 + slots - 1) / slots));
    while (j3 < m3 && xy * ((m3 + j3 - 1) / j3) > slots) ++j3;
  }
  compute_ztiling(pl, t, j3);
}

}  // namespace

void plan_tilings(Plan& pl, bool want_adjoint, bool adjm_shape_ok,
                  const std::function<int(const Plan&)>& resident_waves) {
  // streaming kernels: Nin 2-3, monotone spans, p1, p2 <= 5, non-empty grid
  bool ok = (pl.nin == 2 || pl.nin == 3);
  for (int d = 0; ok && d < pl.nin; ++d) ok = pl.dims[d].monotone && pl.dims[d].m > 0;
  ok = ok && pl.dims[0].p <= kMaxStreamDegree && pl.dims[1].p <= kMaxStreamDegree;
  pl.nt_stores = env_int("SG_NT_STORES", 0) ? 1 : 0;
  if (ok) {
    const char* xm = std::getenv("SG_XCD_MAP");
    const int xcd = (xm && *xm) ? std::atoi(xm) : default_xcd_map(pl);
    // forward: one wave per 64 samples of dimension 1; 32 rows per wave when
    // dimension 2 splits into whole chunks of 128 rows, else 16 (a ragged
    // last chunk of 32-row waves leaves most of its workgroup idle)
    int wx = env_int("SG_FWD_WX", 1);
    if (wx != 1 && wx != 2 && wx != 4) wx = 1;
    int rows = env_int("SG_FWD_ROWS", pl.dims[1].m % (32 * kWaves) == 0 ? 32 : 16);
    compute_tiling(pl, pl.fwd, rows, wx);
    while (forward_lds_worst(pl, pl.fwd) > static_cast<size_t>(kMaxLds) &&
           (rows > 1 || wx > 1)) {
      if (wx > 1)
        wx /= 2;
      else
        rows /= 2;
      compute_tiling(pl, pl.fwd, rows, wx);
    }
    pl.fwd.xcd_map = xcd;
    // MFMA tiers: 64-sample tiles, each wave 16 samples of dimension 1
    // (LDS: the widest channel block, or four fields of one channel)
    const int mnch = std::max(std::min(pl.nout, 4) + (pl.nurbs ? 1 : 0), 1);
    auto mma_worst = [&]() {
      return std::max(mma_lds(pl, mnch, 1), mma_lds(pl, 1, SG_MAX_FIELDS));
    };
    int mrows = 32;
    compute_tiling(pl, pl.mma, mrows, 1);
    while (mma_worst() > static_cast<size_t>(kMaxLds) && mrows > 4) {
      mrows /= 2;
      compute_tiling(pl, pl.mma, mrows, 1);
    }
    if (mma_worst() > static_cast<size_t>(kMaxLds)) pl.mma = Tiling();
    pl.mma.xcd_map = xcd;
    ok = forward_lds_worst(pl, pl.fwd) <= static_cast<size_t>(kMaxLds);
  }
  // adjoint tilings for NURBS plans too: the rational adjoint runs the
  // B-spline adjoint on the same dimensions
  if (ok && want_adjoint) {
    int rows = env_int("SG_ADJ_ROWS", 16);
    const int j3 = env_int("SG_ADJ_J3", 0);
    compute_tiling(pl, pl.adj, rows);
    compute_ztiling(pl, pl.adj, j3);
    while (adjoint_lds_min(pl) > static_cast<size_t>(kMaxLds) && rows > 1) {
      rows /= 2;
      compute_tiling(pl, pl.adj, rows);
      compute_ztiling(pl, pl.adj, j3);
    }
    if (adjoint_lds_min(pl) > static_cast<size_t>(kMaxLds)) pl.adj = Tiling();
    const char* xm = std::getenv("SG_XCD_MAP");
    pl.adj.xcd_map = (xm && *xm) ? std::atoi(xm) : default_xcd_map(pl);
  }
  pl.stream_ok = ok;
  const char* adj_mma_env = std::getenv("SG_ADJ_MMA");  // "0": streaming kernel only
  if (ok && want_adjoint && !(adj_mma_env && std::strcmp(adj_mma_env, "0") == 0) &&
      adjm_shape_ok)
    plan_adjoint_mma(pl, resident_waves);
}

// Gather lists of a tiling whose partial windows are w1 x w2 x w3max per
// channel, laid out [block3][block2][block1][k3 - lo3][channel][k2 - lo2]
// [k1 - lo1]: the element offset of control point k in a window is the sum
// of one entry per dimension and channel * gchan.
GatherLists gather_lists(const Plan& pl, const Tiling& tl, int w1, int w2) {
  GatherLists g;
  const long long plane = static_cast<long long>(w1) * w2;
  const long long win = static_cast<long long>(tl.w3max) * pl.nout * plane;  // one block
  // (block, offset in window) of dimension d -> its part of the element offset
  const long long stride1 = win, stride2 = win * tl.n_tiles1,
                  stride3 = win * tl.n_tiles1 * tl.n_chunks2;
  auto lists = [&](int nblocks, int block_len, const DimDev& D, long long bstride,
                   long long ostride, std::vector<int32_t>& ptr, std::vector<long long>& off) {
    std::vector<std::vector<long long>> per(D.n);
    for (int b = 0; b < nblocks; ++b) {
      const int a = b * block_len, e = std::min(a + block_len, D.m) - 1;
      const int lo = D.h_span[a] - D.p;
      for (int k = lo; k <= D.h_span[e]; ++k) per[k].push_back(b * bstride + (k - lo) * ostride);
    }
    ptr.assign(D.n + 1, 0);
    off.clear();
    for (int k = 0; k < D.n; ++k) {
      off.insert(off.end(), per[k].begin(), per[k].end());
      ptr[k + 1] = static_cast<int32_t>(off.size());
    }
  };
  lists(tl.n_tiles1, tl.tile_width(), pl.dims[0], stride1, 1, g.ptr[0], g.off[0]);
  lists(tl.n_chunks2, tl.chunk_rows(), pl.dims[1], stride2, w1, g.ptr[1], g.off[1]);
  if (pl.nin == 3) {
    lists(tl.n_z, tl.j3_per_block, pl.dims[2], stride3, pl.nout * plane, g.ptr[2], g.off[2]);
  } else {
    g.ptr[2] = {0, 1};
    g.off[2] = {0};
  }
  g.gchan = plane;
  // the gather splits a window number f < n1 n2 n3 (per-point counts) into
  // (i1, i2, i3) with udiv_fast when f * n1 n2 and f12 * n1 stay below 2^32
  long long nmax[3] = {1, 1, 1};
  for (int d = 0; d < 3; ++d)
    for (size_t k = 0; k + 1 < g.ptr[d].size(); ++k)
      nmax[d] = std::max<long long>(nmax[d], g.ptr[d][k + 1] - g.ptr[d][k]);
  const long long n12 = nmax[0] * nmax[1];
  g.gfast = (n12 * n12 * nmax[2] < (1LL << 32) && nmax[0] * n12 < (1LL << 32)) ? 1 : 0;
  return g;
}

void band_lists(const DimDev& D, std::vector<int32_t>& ptr, std::vector<int32_t>& js) {
  ptr.assign(D.n + 1, 0);
  js.assign(static_cast<size_t>(D.m) * (D.p + 1), 0);
  for (int j = 0; j < D.m; ++j)
    for (int a = 0; a <= D.p; ++a) ptr[D.h_span[j] - D.p + a + 1]++;
  for (int k = 0; k < D.n; ++k) ptr[k + 1] += ptr[k];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int j = 0; j < D.m; ++j)
    for (int a = 0; a <= D.p; ++a) js[fill[D.h_span[j] - D.p + a]++] = j * (D.p + 1) + a;
}

// LDS of one CU; a persistent launch of `bpc` workgroups per CU gives each
// at most a bpc-th of it, or fewer workgroups than asked are resident.
constexpr size_t kLdsPerCu = 160 * 1024;

size_t lds_cap(int blocks_per_cu) {
  const size_t share = kLdsPerCu / static_cast<size_t>(std::max(1, blocks_per_cu));
  return std::min(static_cast<size_t>(kMaxLds), (share & ~size_t(511)) - 512);
}

int batch_vec(const Plan& pl) {
  const int m1 = pl.dims[0].m;
  return m1 > kLanes && pl.dims[0].p <= kVec2MaxDegree &&
                 pl.dims[1].p <= kVec2MaxDegree
             ? 2
             : 1;
}

GridTiling grid_tiling(const Plan& pl, int rows, int vec) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  const int tw = kLanes * vec;
  GridTiling t{rows, (d1.m + tw - 1) / tw, 0, 0, 0, 0, tw};
  const int chunk = rows * kWaves;
  t.nty = (d2.m + chunk - 1) / chunk;
  for (int i = 0; i < t.ntx; ++i) {
    const int a = i * tw, e = std::min(a + tw, d1.m) - 1;
    t.w1 = std::max(t.w1, d1.h_span[e] - d1.h_span[a] + d1.p + 1);
  }
  for (int c = 0; c < t.nty; ++c) {
    const int a = c * chunk, e = std::min(a + chunk, d2.m) - 1;
    t.w2 = std::max(t.w2, d2.h_span[e] - d2.h_span[a] + d2.p + 1);
  }
  const size_t nch = std::min(pl.nout, kBC) + (pl.nurbs ? 1 : 0);
  const size_t planes = pl.nin == 3 ? kWaves : 1;
  t.lds = planes * nch * t.w1 * t.w2 * sizeof(float);  // the window slabs
  return t;
}

GridTiling batch_grid_tiling(const Plan& pl, int rows_per_wave, size_t cap) {
  auto fit = [&](int vec) {
    int rows = rows_per_wave;
    GridTiling t = grid_tiling(pl, rows, vec);
    while (t.lds > cap && rows > 1) t = grid_tiling(pl, rows /= 2, vec);
    return t;
  };
  const GridTiling t1 = fit(1);
  if (batch_vec(pl) != 2) return t1;
  const GridTiling t2 = fit(2);
  return t2.lds <= cap ? t2 : t1;
}

int plane_list_host(const int32_t* span3, const float* tab3, int p3, int j3_0, int planes,
                    float* list) {
  const int cap = plane_list_floats(p3);
  std::fill(list, list + cap, 0.f);
  std::vector<int> ks;
  for (int pz = 0; pz < planes; ++pz)
    for (int d = 0; d <= p3; ++d) ks.push_back(span3[j3_0 + pz] - p3 + d);
  std::sort(ks.begin(), ks.end());
  ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
  const int nk = static_cast<int>(ks.size());
  auto put_k = [&](int i, int k3) { std::memcpy(list + i * kEntry + kEntryK, &k3, sizeof(int)); };
  for (int i = 0; i < nk; ++i) {
    put_k(i, ks[i]);
    for (int pz = 0; pz < planes; ++pz) {
      const int d = ks[i] - (span3[j3_0 + pz] - p3);
      if (d >= 0 && d <= p3) list[i * kEntry + pz] = tab3[static_cast<size_t>(j3_0 + pz) * (p3 + 1) + d];
    }
  }
  put_k(nk, ks.back());  // padding: zero weights, a plane of the grid
  std::memcpy(list + kEntryN, &nk, sizeof(int));
  return nk;
}

void pack_queues(const std::vector<std::vector<BTile>>& grid_tiles,
                 const std::vector<long long>& grid_cost, std::vector<BTile>& sorted,
                 int* host_q) {
  std::vector<size_t> order(grid_tiles.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t c) { return grid_cost[a] > grid_cost[c]; });
  std::vector<std::vector<size_t>> queue_grids(kXcds);
  long long load[kXcds] = {};
  for (size_t g : order) {
    const int q = static_cast<int>(std::min_element(load, load + kXcds) - load);
    load[q] += grid_cost[g];
    queue_grids[q].push_back(g);
  }
  sorted.clear();
  std::fill(host_q, host_q + kQueueInts, 0);
  int* start = host_q + kXcds * kQueueStride;
  for (int q = 0; q < kXcds; ++q) {
    start[q] = static_cast<int>(sorted.size());
    for (size_t g : queue_grids[q])
      sorted.insert(sorted.end(), grid_tiles[g].begin(), grid_tiles[g].end());
  }
  start[kXcds] = static_cast<int>(sorted.size());
}

}  // namespace sg
