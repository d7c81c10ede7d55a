// plan_check.cpp -- the host plan code (splinegrids.jl_amd/csrc/sg_plan_host.hip)
// on the CPU, built with the address and undefined-behaviour sanitizers.
//
// For the shapes the GPU tests use (the five benchmark grids, degree and
// size corner cases, samples on repeated knots, outside the knot extent and
// unsorted, and a batch of random uneven grids) it builds a Plan's host
// part -- spans by the reference's rule (src/util_kernels.jl:22-49: the
// number of knots <= t, clamped to [p+1, n], 1-based) -- runs every tiling,
// the gather lists, the transposed band lists and the batch's tile windows
// and queue packing, and checks what the kernels rely on:
//   * every tile's band window lies inside the control points and within
//     the tiling's widest window (the LDS size);
//   * gather offsets stay inside the partial-window buffer, and every
//     control point a window holds is listed once per window;
//   * band lists are a permutation of the (sample, term) pairs;
//   * the queues hold every tile exactly once, in order of their starts.
// Prints "plan_check ok ..." and exits 0, or names the first failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "sg_plan_host.hpp"

using namespace sg;

namespace {

int g_checks = 0;

void require(bool ok, const std::string& what) {
  ++g_checks;
  if (!ok) {
    std::fprintf(stderr, "plan_check FAILED: %s\n", what.c_str());
    std::exit(1);
  }
}

struct DimSpec {
  int n, p, m;
  int knots;    // 0 clamped uniform, 1 random interior with repeats
  int samples;  // 0 uniform over the extent, 1 on every knot, 2 past the extent, 3 unsorted
};

std::vector<double> make_knots(const DimSpec& d, std::mt19937_64& rng) {
  // clamped: p + 1 zeros, n - p - 1 interior values, p + 1 ones
  const int ni = d.n - d.p - 1;
  std::vector<double> all(d.p + 1, 0.0);
  std::vector<double> inner(ni);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (int i = 0; i < ni; ++i) inner[i] = d.knots == 1 ? u(rng) : (i + 1.0) / (ni + 1.0);
  std::sort(inner.begin(), inner.end());
  if (d.knots == 1 && d.p >= 1)  // every third interior knot doubled
    for (int i = 0; i + 1 < ni; i += 3) inner[i + 1] = inner[i];
  all.insert(all.end(), inner.begin(), inner.end());
  all.insert(all.end(), d.p + 1, 1.0);
  return all;
}

std::vector<double> make_samples(const DimSpec& d, const std::vector<double>& knots,
                                 std::mt19937_64& rng) {
  std::vector<double> t(d.m);
  const double a = knots.front(), b = knots.back();
  for (int i = 0; i < d.m; ++i) t[i] = d.m > 1 ? a + (b - a) * i / (d.m - 1) : a;
  if (d.samples == 1)
    for (int i = 0; i < d.m; ++i) t[i] = knots[i % knots.size()];
  if (d.samples == 2)
    for (int i = 0; i < d.m; ++i) t[i] = a - 0.25 + 1.5 * (b - a) * i / std::max(1, d.m - 1);
  if (d.samples == 1) std::sort(t.begin(), t.end());
  if (d.samples == 3) std::shuffle(t.begin(), t.end(), rng);
  return t;
}

// 0-based span: the last basis function whose support holds t
std::vector<int32_t> spans(const DimSpec& d, const std::vector<double>& knots,
                           const std::vector<double>& t) {
  std::vector<int32_t> s(d.m);
  for (int j = 0; j < d.m; ++j) {
    int c = 0;
    for (double k : knots) c += k <= t[j] ? 1 : 0;
    c = std::min(std::max(c, d.p + 1), d.n);
    s[j] = c - 1;
  }
  return s;
}

Plan make_plan(const std::vector<DimSpec>& dims, int nout, bool nurbs, std::mt19937_64& rng) {
  Plan pl;
  pl.nin = static_cast<int>(dims.size());
  pl.nout = nout;
  pl.nurbs = nurbs ? 1 : 0;
  pl.dims.resize(dims.size());
  for (size_t i = 0; i < dims.size(); ++i) {
    const DimSpec& d = dims[i];
    DimDev& D = pl.dims[i];
    D.n = d.n;
    D.p = d.p;
    D.m = d.m;
    D.D = 0;
    const std::vector<double> knots = make_knots(d, rng);
    D.h_span = spans(d, knots, make_samples(d, knots, rng));
    D.monotone = std::is_sorted(D.h_span.begin(), D.h_span.end());
  }
  return pl;
}

std::string name_of(const Plan& pl) {
  std::string s = "n=(";
  for (const DimDev& D : pl.dims) s += std::to_string(D.n) + ",";
  s += ") p=(";
  for (const DimDev& D : pl.dims) s += std::to_string(D.p) + ",";
  s += ") m=(";
  for (const DimDev& D : pl.dims) s += std::to_string(D.m) + ",";
  return s + ") nout=" + std::to_string(pl.nout) + (pl.nurbs ? " nurbs" : "");
}

// every block of block_len samples: its window [span(first) - p, span(last)]
// lies in [0, n) and is at most wmax wide
void check_windows(const DimDev& D, int nblocks, int block_len, int wmax, const std::string& what) {
  require(nblocks == (D.m + block_len - 1) / block_len, what + ": block count");
  for (int b = 0; b < nblocks; ++b) {
    const int a = b * block_len, e = std::min(a + block_len, D.m) - 1;
    const int lo = D.h_span[a] - D.p, hi = D.h_span[e];
    require(lo >= 0 && hi < D.n && lo <= hi, what + ": window inside the control points");
    require(hi - lo + 1 <= wmax, what + ": window within the widest");
  }
}

// (the forward tilings take one sample j3 per workgroup; the adjoint ones
// group j3_per_block of them)
void check_tiling(const Plan& pl, const Tiling& t, bool groups_j3, const std::string& what) {
  if (t.n_tiles1 == 0) return;
  check_windows(pl.dims[0], t.n_tiles1, t.tile_width(), t.w1max, what + " dim 1");
  check_windows(pl.dims[1], t.n_chunks2, t.chunk_rows(), t.w2max, what + " dim 2");
  if (pl.nin == 3 && groups_j3)
    check_windows(pl.dims[2], t.n_z, t.j3_per_block, t.w3max, what + " dim 3");
}

void check_gather(const Plan& pl, const Tiling& t, int w1, int w2, const std::string& what) {
  if (t.n_tiles1 == 0) return;
  const GatherLists g = gather_lists(pl, t, w1, w2);
  const long long total = static_cast<long long>(t.n_tiles1) * t.n_chunks2 * t.n_z *
                          t.w3max * pl.nout * w1 * w2;
  long long maxoff[3] = {0, 0, 0};
  const int nblocks[3] = {t.n_tiles1, t.n_chunks2, pl.nin == 3 ? t.n_z : 1};
  for (int d = 0; d < 3; ++d) {
    const int n = d < pl.nin ? pl.dims[d].n : 1;
    require(static_cast<int>(g.ptr[d].size()) == n + 1, what + ": pointer length");
    require(g.ptr[d][0] == 0 && g.ptr[d][n] == static_cast<int32_t>(g.off[d].size()),
            what + ": pointer ends");
    for (int k = 0; k < n; ++k) require(g.ptr[d][k] <= g.ptr[d][k + 1], what + ": pointers sorted");
    long long listed = 0;
    for (long long o : g.off[d]) {
      require(o >= 0, what + ": offset >= 0");
      maxoff[d] = std::max(maxoff[d], o);
      ++listed;
    }
    if (d < pl.nin) {  // each block lists every point of its window once
      const DimDev& D = pl.dims[d];
      const int len = d == 0 ? t.tile_width() : d == 1 ? t.chunk_rows() : t.j3_per_block;
      long long want = 0;
      for (int b = 0; b < nblocks[d]; ++b) {
        const int a = b * len, e = std::min(a + len, D.m) - 1;
        want += D.h_span[e] - (D.h_span[a] - D.p) + 1;
      }
      require(listed == want, what + ": every window point listed once");
    }
  }
  require(maxoff[0] + maxoff[1] + maxoff[2] + (pl.nout - 1) * g.gchan < total,
          what + ": offsets inside the partial buffer");
}

void check_bands(const DimDev& D, const std::string& what) {
  std::vector<int32_t> ptr, js;
  band_lists(D, ptr, js);
  const long long pairs = static_cast<long long>(D.m) * (D.p + 1);
  require(static_cast<int>(ptr.size()) == D.n + 1 && ptr[D.n] == pairs, what + ": band pointers");
  std::vector<char> seen(pairs, 0);
  for (int k = 0; k < D.n; ++k)
    for (int i = ptr[k]; i < ptr[k + 1]; ++i) {
      const int v = js[i];
      require(v >= 0 && v < pairs && !seen[v], what + ": band entry once");
      seen[v] = 1;
      require(D.h_span[v / (D.p + 1)] - D.p + v % (D.p + 1) == k, what + ": band entry holds k");
    }
}

// the batch's plane lists of a 3-D grid, every tile row: each band entry
// of each sample exactly once with its weight, k3 increasing, the padding
// entry, the stored length, nothing written past the list
void check_plane_lists(const Plan& pl, const std::string& nm) {
  const DimDev& D = pl.dims[2];
  const int p = D.p, stride = plane_list_floats(p);
  std::vector<float> tab(static_cast<size_t>(D.m) * (p + 1));
  for (size_t i = 0; i < tab.size(); ++i) tab[i] = 0.5f + static_cast<float>(i);  // distinct, nonzero
  std::vector<float> list(stride + kEntry, -7.f);
  for (int j0 = 0; j0 < D.m; j0 += kWaves) {
    const int planes = std::min(kWaves, D.m - j0);
    const int nk = plane_list_host(D.h_span.data(), tab.data(), p, j0, planes, list.data());
    const std::string w = nm + " plane list j3 " + std::to_string(j0);
    require(nk >= p + 1 && nk <= kWaves * (p + 1), w + ": length");
    int n_stored;
    std::memcpy(&n_stored, list.data() + kEntryN, sizeof(int));
    require(n_stored == nk, w + ": stored length");
    for (int i = 0; i < kEntry; ++i) require(list[stride + i] == -7.f, w + ": nothing past the list");
    auto k_of = [&](int i) {
      int k;
      std::memcpy(&k, list.data() + i * kEntry + kEntryK, sizeof(int));
      return k;
    };
    for (int i = 0; i + 1 < nk; ++i) require(k_of(i) < k_of(i + 1), w + ": k3 increasing");
    require(k_of(nk) == k_of(nk - 1), w + ": padding plane");
    for (int pz = 0; pz < kWaves; ++pz) require(list[nk * kEntry + pz] == 0.f, w + ": padding weights");
    for (int pz = 0; pz < kWaves; ++pz) {
      int found = 0;
      for (int i = 0; i < nk; ++i) {
        const float v = list[i * kEntry + pz];
        const int d = k_of(i) - (pz < planes ? D.h_span[j0 + pz] - p : 0);
        if (pz < planes && d >= 0 && d <= p) {
          require(v == tab[static_cast<size_t>(j0 + pz) * (p + 1) + d], w + ": weight");
          ++found;
        } else {
          require(v == 0.f, w + ": zero outside the band");
        }
      }
      require(found == (pz < planes ? p + 1 : 0), w + ": every band entry once");
    }
  }
}

void check_plan(Plan& pl) {
  const std::string nm = name_of(pl);
  if (pl.nin == 3 && pl.dims[2].m > 0) check_plane_lists(pl, nm);
  const bool adjm_ok = (pl.nin == 2 || pl.nin == 3) &&
                       (pl.nin == 2 || pl.dims[2].p <= kMaxStreamDegree);
  plan_tilings(pl, true, adjm_ok, [](const Plan&) { return 2048; });
  bool want_stream = (pl.nin == 2 || pl.nin == 3);
  for (const DimDev& D : pl.dims) want_stream = want_stream && D.monotone && D.m > 0;
  if (pl.nin >= 2)
    want_stream = want_stream && pl.dims[0].p <= kMaxStreamDegree &&
                  pl.dims[1].p <= kMaxStreamDegree;
  if (pl.stream_ok) require(want_stream, nm + ": streaming only where it applies");
  for (const DimDev& D : pl.dims) check_bands(D, nm);
  if (!pl.stream_ok) return;
  check_tiling(pl, pl.fwd, false, nm + " fwd");
  require(forward_lds_worst(pl, pl.fwd) <= static_cast<size_t>(kMaxLds), nm + ": fwd LDS");
  check_tiling(pl, pl.mma, false, nm + " mma");
  if (pl.mma.n_tiles1 > 0)
    require(mma_lds(pl, 1, 1) <= static_cast<size_t>(kMaxLds), nm + ": mma LDS");
  check_tiling(pl, pl.adj, true, nm + " adj");
  check_gather(pl, pl.adj, pl.adj.w1max, pl.adj.w2max, nm + " adj gather");
  if (pl.adjm.n_tiles1 > 0) {
    require(pl.adjm.w1max <= kAdjMmaWin && pl.adjm.w2max <= kAdjMmaWin, nm + ": adjm window");
    check_tiling(pl, pl.adjm, true, nm + " adjm");
    check_gather(pl, pl.adjm, kAdjMmaWin, kAdjMmaWin, nm + " adjm gather");
  }
}

// the batch: windows and LDS of every grid's tiles, and the queues
void check_batch(std::vector<Plan>& plans) {
  for (int rows : {8, 16, 32}) {
    std::vector<std::vector<BTile>> grid_tiles;
    std::vector<long long> cost;
    size_t ntiles = 0;
    for (size_t gi = 0; gi < plans.size(); ++gi) {
      const Plan& pl = plans[gi];
      if (!pl.stream_ok) continue;
      const GridTiling t = grid_tiling(pl, rows);
      const std::string nm = name_of(pl) + " batch rows " + std::to_string(rows);
      check_windows(pl.dims[0], t.ntx, t.tw, t.w1, nm + " dim 1");
      check_windows(pl.dims[1], t.nty, rows * kWaves, t.w2, nm + " dim 2");
      require(t.lds > 0, nm + ": LDS");
      // the choice of rows and samples per lane under each LDS cap
      for (int bpc : {2, 3, 4, 6}) {
        const size_t cap = lds_cap(bpc);
        const GridTiling c = batch_grid_tiling(pl, rows, cap);
        const std::string cn = nm + " cap " + std::to_string(cap);
        require(c.tw == kLanes || (c.tw == 2 * kLanes && batch_vec(pl) == 2), cn + ": samples per lane");
        if (c.tw > kLanes)
          require(pl.dims[0].m > kLanes && pl.dims[0].p <= kVec2MaxDegree &&
                      pl.dims[1].p <= kVec2MaxDegree && c.lds <= cap,
                  cn + ": two samples per lane only for long rows and where they fit");
        require(c.rows >= 1 && c.rows <= rows && rows % c.rows == 0, cn + ": rows");
        const GridTiling again = grid_tiling(pl, c.rows, c.tw / kLanes);
        require(again.lds == c.lds && again.w1 == c.w1 && again.w2 == c.w2, cn + ": consistent");
        check_windows(pl.dims[0], c.ntx, c.tw, c.w1, cn + " dim 1");
        check_windows(pl.dims[1], c.nty, c.rows * kWaves, c.w2, cn + " dim 2");
        if (c.lds > cap) require(grid_tiling(pl, 1, 1).lds > cap, cn + ": a fitting choice was missed");
      }
      std::vector<BTile> tiles;
      const int nz = pl.nin == 3 ? (pl.dims[2].m + kWaves - 1) / kWaves : 1;
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < t.nty; ++y)
          for (int x = 0; x < t.ntx; ++x) tiles.push_back(BTile{static_cast<int>(gi), x, y, z});
      ntiles += tiles.size();
      cost.push_back(static_cast<long long>(tiles.size()) * (1 + (gi % 7)));
      grid_tiles.push_back(std::move(tiles));
    }
    std::vector<BTile> sorted;
    std::vector<int> host_q(kQueueInts, -1);
    pack_queues(grid_tiles, cost, sorted, host_q.data());
    require(sorted.size() == ntiles, "queues hold every tile");
    const int* start = host_q.data() + kXcds * kQueueStride;
    require(start[0] == 0 && start[kXcds] == static_cast<int>(ntiles), "queue start ends");
    for (int q = 0; q < kXcds; ++q) {
      require(start[q] <= start[q + 1], "queue starts sorted");
      require(host_q[q * kQueueStride] == 0, "queue counters zero");
    }
    std::vector<std::vector<char>> seen(plans.size());
    for (size_t g = 0; g < plans.size(); ++g) seen[g].assign(1 << 20, 0);
    for (const BTile& t : sorted) {
      const Plan& pl = plans[t.grid];
      const GridTiling gt = grid_tiling(pl, rows);
      const long long id = (static_cast<long long>(t.bz) * gt.nty + t.by) * gt.ntx + t.bx;
      require(id >= 0 && id < (1 << 20) && !seen[t.grid][id], "each tile once");
      seen[t.grid][id] = 1;
    }
  }
  for (int bpc = 1; bpc <= 8; ++bpc)
    require(lds_cap(bpc) > 0 && lds_cap(bpc) <= static_cast<size_t>(kMaxLds), "LDS share");
}

}  // namespace

int main() {
  std::mt19937_64 rng(0x5EED);
  int n_plans = 0;
  auto run = [&](const std::vector<DimSpec>& dims, int nout, bool nurbs) {
    Plan pl = make_plan(dims, nout, nurbs, rng);
    check_plan(pl);
    ++n_plans;
  };
  // the benchmark grids
  run({{10, 2, 50, 0, 0}, {10, 3, 50, 0, 0}, {5, 2, 25, 0, 0}}, 4, false);
  run({{128, 3, 4096, 1, 0}, {128, 3, 4096, 1, 0}}, 1, false);
  run({{64, 3, 512, 0, 0}, {64, 3, 512, 0, 0}, {32, 2, 256, 0, 0}}, 3, true);
  run({{40, 3, 400, 0, 0}, {40, 3, 400, 0, 0}, {20, 3, 200, 0, 0}}, 3, false);
  // degree, size and sample corner cases
  for (int p = 0; p <= 6; ++p)
    for (int m : {1, 2, 3, 17, 65, 130})
      for (int kind : {0, 1, 2, 3}) {
        run({{p + 1, p, m, 0, kind}, {p + 3, std::min(p, 5), m + 1, 1, kind}}, 1 + p % 3, p % 2);
        run({{p + 2, p, m, 1, kind}, {9, 2, 2 * m, 0, kind}, {p + 1, p, 5, 0, kind}},
            1 + (m % 4), m % 2);
      }
  // wide windows: many control points, few samples
  run({{64, 3, 16, 0, 0}, {64, 4, 20, 1, 0}, {64, 2, 16, 0, 0}}, 3, true);
  run({{200, 5, 32, 1, 0}, {150, 5, 40, 0, 1}}, 4, false);
  // the uneven batch
  std::vector<Plan> batch;
  std::uniform_int_distribution<int> nd(8, 64), pd(1, 4), m2(32, 512), m3(16, 128);
  for (int g = 0; g < 256; ++g) {
    const int nin = g % 2 ? 3 : 2;
    std::vector<DimSpec> dims;
    for (int d = 0; d < nin; ++d) {
      const int n = nd(rng);
      dims.push_back({n, std::min(pd(rng), n - 1), nin == 2 ? m2(rng) : m3(rng), g % 3 == 0, 0});
    }
    batch.push_back(make_plan(dims, g % 2 ? 3 : 1, g % 4 == 0, rng));
    check_plan(batch.back());
    ++n_plans;
  }
  check_batch(batch);
  std::printf("plan_check ok: %d plans, %d checks\n", n_plans, g_checks);
  return 0;
}
