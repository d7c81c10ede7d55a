// sg_mfma.hip -- evaluate!(grid) with the contraction on MFMA (FP32 arrays).
//
// Tiers
//   SPLIT_F16  every operand x is split into hi = fp16(x) and lo = fp16(x-hi)
//              after power-of-two scaling; products use hi*hi + hi*lo + lo*hi
//              on v_mfma_f32_16x16x16_f16 with FP32 accumulation (the lo*lo
//              term is below 2^-22 relative).
//   BF16       one pass of v_mfma_f32_16x16x16bf16_1k on bf16(x).
//
// Structure (per workgroup, same tiling and stage 1 as fwd_stream): the
// tile's band window G (FP32, LDS) is scaled by 2^-eG so that max|G| < 1.
// Each wave owns 16 consecutive samples j1 of the 64-sample tile and walks
// the chunk's rows in blocks of 16.  For a block, with K = 16 consecutive
// control points k2 from kb,
//   V^T[k2][j1]  = sum_a B1[j1, a] * G[k2][s1-p1+a]        (FP32 FMA)
//   Out[j2][j1]  = sum_k2 B2d[j2][k2] * V^T[k2][j1]         (MFMA, 16x16x16)
// where B2d is the banded B2 made dense on the fly.  V^T is kept while the
// rows' bands stay inside [kb, kb+16).  MFMA operand layout (gfx9 16x16x16):
// lane l = 16 g + i holds A[i][4g..4g+3], B[4g..4g+3][i], D[4g..4g+3][i].
#include <cstring>
#include <vector>

#include "sg_mma_ops.hpp"
#include "sg_plan_host.hpp"
#include "sg_tiles.hpp"

namespace sg {

namespace {

struct MmaArgs {
  BandSrc<float> src;
  float* out[SG_MAX_FIELDS];
  const float* tab1[SG_MAX_FIELDS];  // [m1][p1+1] at each field's order
  const float* tab2[SG_MAX_FIELDS];  // [m2][p2+1]
  const int32_t* span1;
  const int32_t* span2;
  int m1, m2, p1, p2;
  int rows_per_chunk;
  int ntx, nty, xcd_map;
  int gmax;                      // LDS floats reserved per channel of G
  float s1[SG_MAX_FIELDS];       // 2^-e of the dimension 1 / 2 tables of each field
  float s2[SG_MAX_FIELDS];
  int e12[SG_MAX_FIELDS];        // e1 + e2 of each field
  long long nsamp;
  int nin, nc, nurbs;            // batched launches: what the record needs
};

// One tile (64 samples j1 x a chunk of rows x sample j3) of channels
// [src.c0, src.c0 + NC): stage 1, the chunk's B2 rows, the window's scale,
// then each wave's 16 samples j1 over the chunk's rows in blocks of 16.
template <int TIER, int NIN, int NF, int NC, bool NURBS>
__device__ __attribute__((always_inline)) void mma_tile(const MmaArgs& a, const FwdTile& tile,
                                                        unsigned char* smem) {
  constexpr int NCH = NC + (NURBS ? 1 : 0);
  constexpr int PM = kMaxStreamDegree;
  float* G = reinterpret_cast<float*>(smem);
  __shared__ float red[kWaves][NCH];

  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const int r0 = tile.r0, r1 = tile.r1, j3 = tile.j3;
  const int p1 = a.p1, p2 = a.p2;
  const int lo1 = a.span1[tile.j1_0] - p1;
  const int W1 = a.span1[tile.j1_last] - lo1 + 1;
  const int lo2 = a.span2[r0] - p2;
  const int hi2 = a.span2[r1 - 1];
  const int W2 = hi2 - lo2 + 1;
  const int cs = W1 * W2;
  // the chunk's B2 rows of every field (scaled) and spans in LDS
  const int chunk = a.rows_per_chunk;
  float* RT = G + static_cast<size_t>(NCH) * a.gmax;                        // [NF][chunk][p2+1]
  int* RS = reinterpret_cast<int*>(RT + static_cast<size_t>(NF) * chunk * (p2 + 1));

  stage1_band_window<float, NIN, NC, NURBS>(G, a.src, j3, lo1, W1, lo2, W2);
  {
    const size_t rbase = static_cast<size_t>(r0) * (p2 + 1);
    const size_t fstride = static_cast<size_t>(chunk) * (p2 + 1);
    stage_to_lds<1, 1>(r1 - r0, [&](int, int i) { return a.span2[r0 + i]; },
                       [&](int, int i, int v) { RS[i] = v; });
    stage_to_lds<NF, 4>((r1 - r0) * (p2 + 1), [&](int f, int i) { return a.tab2[f][rbase + i]; },
                        [&](int f, int i, float v) { RT[f * fstride + i] = v * a.s2[f]; });
  }
  __syncthreads();

  // per-channel power-of-two scale of the window: max|G_c| * 2^-e < 1
  float mx[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    float m = 0.f;
    for (int i = threadIdx.x; i < cs; i += kThreads) m = fmaxf(m, fabsf(G[ch * cs + i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    mx[ch] = m;
  }
  if (lane == 0)
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) red[wave][ch] = mx[ch];
  __syncthreads();
  float gs[NCH];  // 2^-eG
  int eg[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    float m = fmaxf(fmaxf(red[0][ch], red[1][ch]), fmaxf(red[2][ch], red[3][ch]));
    int e = 0;
    if (m > 0.f) frexpf(m, &e);
    eg[ch] = e;
    gs[ch] = ldexpf(1.f, -e);
  }

  // ---- per-lane constants: sample j1 (column) and k-quarter g ----
  const int g = lane >> 4, col = lane & 15;
  const int j1q = tile.j1_0 + 16 * wave + col;
  const bool active = j1q < a.m1;
  const int j1c = active ? j1q : a.m1 - 1;
  const int q1b = a.span1[j1c] - p1 - lo1;
  // b1 is zero past p1 (scaled by the field's 2^-e1)
  float b1[NF][PM + 1];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int t = 0; t <= PM; ++t) {
      const float v = a.tab1[f][static_cast<size_t>(j1c) * (p1 + 1) + min(t, p1)] * a.s1[f];
      b1[f][t] = t <= p1 ? v : 0.f;
    }

  Operand<TIER> vop[NF][NCH];
  int kb = -(1 << 30);  // k2 base of the cached V^T operands
  auto build_v = [&](int base) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k2 = base + 4 * g + j;
      const bool in = k2 >= lo2 && k2 <= hi2;
      const float* gr = G + (in ? (k2 - lo2) : 0) * W1 + q1b;
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        float gv[PM + 1];
#pragma unroll
        for (int t = 0; t <= PM; ++t) gv[t] = gr[ch * cs + min(t, p1)];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          float v = 0.f;
#pragma unroll
          for (int t = 0; t <= PM; ++t) v += b1[f][t] * gv[t];
          vop[f][ch].set(j, in ? v * gs[ch] : 0.f);
        }
      }
    }
  };

  const long long obase = static_cast<long long>(j1q) + static_cast<long long>(a.m1) * a.m2 * j3 +
                          a.nsamp * a.src.c0;
  for (int jb = r0; jb < r1; jb += 16) {
    const int jl = min(jb + 15, r1 - 1);
    const int klo = wave_uniform(RS[jb - r0]) - p2, khi = wave_uniform(RS[jl - r0]);
    // this lane's row of the A operands
    const bool rowok = jb + col < r1;
    const int j2a = rowok ? jb + col : r1 - 1;
    const int base2 = RS[j2a - r0] - p2;
    float4_t acc[NF][NCH];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) acc[f][ch] = float4_t{0.f, 0.f, 0.f, 0.f};
    int kc = (klo >= kb && khi <= kb + 15) ? kb : klo;
    for (; kc <= khi; kc += 16) {
      if (kc != kb) {
        kb = kc;
        build_v(kb);
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        // A[col][4g + j] = B2 of row j2a at k2 = kc + 4g + j (0 off the band)
        const float* rt = RT + (static_cast<size_t>(f) * chunk + (j2a - r0)) * (p2 + 1);
        Operand<TIER> aop;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int off = kc + 4 * g + j - base2;
          const float v = rt[min(max(off, 0), p2)];
          aop.set(j, (rowok && off >= 0 && off <= p2) ? v : 0.f);
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) acc[f][ch] = Operand<TIER>::mma(aop, vop[f][ch], acc[f][ch]);
      }
    }
    if (active) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int j2 = jb + 4 * g + j;
          if (j2 >= r1) continue;
          const long long o = obase + static_cast<long long>(a.m1) * j2;
          if constexpr (NURBS) {
            const float inv = nurbs_recip(ldexpf(acc[f][NC][j], eg[NC] + a.e12[f]));
#pragma unroll
            for (int c = 0; c < NC; ++c)
              a.out[f][o + a.nsamp * c] = ldexpf(acc[f][c][j], eg[c] + a.e12[f]) * inv;
          } else {
#pragma unroll
            for (int c = 0; c < NC; ++c)
              a.out[f][o + a.nsamp * c] = ldexpf(acc[f][c][j], eg[c] + a.e12[f]);
          }
        }
    }
  }
}

template <int TIER, int NIN, int NF, int NC, bool NURBS>
__global__ __launch_bounds__(kThreads) void mma_stream(MmaArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const FwdTile tile = fwd_tile(a.xcd_map, a.ntx, a.nty, kLanes, a.rows_per_chunk, a.m1, a.m2,
                                NIN);
  mma_tile<TIER, NIN, NF, NC, NURBS>(a, tile, smem);
}

// Batched launch of the MFMA tiers: a persistent kernel takes tiles of many
// grids from the work queues; each tile names the argument record of one
// grid's channel block (field: derivative order 0), the tile along
// dimension 1, the chunk along dimension 2 and the sample j3.
template <int TIER>
__device__ __attribute__((always_inline)) void batch_mma_dispatch(const MmaArgs& a, const FwdTile& tile,
                                                                  unsigned char* smem) {
  if (a.nin == 2) {
    if (a.nurbs) {
      switch (a.nc) {
        case 1: mma_tile<TIER, 2, 1, 1, true>(a, tile, smem); break;
        case 2: mma_tile<TIER, 2, 1, 2, true>(a, tile, smem); break;
        default: mma_tile<TIER, 2, 1, 3, true>(a, tile, smem); break;
      }
    } else {
      switch (a.nc) {
        case 1: mma_tile<TIER, 2, 1, 1, false>(a, tile, smem); break;
        case 2: mma_tile<TIER, 2, 1, 2, false>(a, tile, smem); break;
        case 3: mma_tile<TIER, 2, 1, 3, false>(a, tile, smem); break;
        default: mma_tile<TIER, 2, 1, 4, false>(a, tile, smem); break;
      }
    }
  } else {
    if (a.nurbs) {
      switch (a.nc) {
        case 1: mma_tile<TIER, 3, 1, 1, true>(a, tile, smem); break;
        case 2: mma_tile<TIER, 3, 1, 2, true>(a, tile, smem); break;
        default: mma_tile<TIER, 3, 1, 3, true>(a, tile, smem); break;
      }
    } else {
      switch (a.nc) {
        case 1: mma_tile<TIER, 3, 1, 1, false>(a, tile, smem); break;
        case 2: mma_tile<TIER, 3, 1, 2, false>(a, tile, smem); break;
        case 3: mma_tile<TIER, 3, 1, 3, false>(a, tile, smem); break;
        default: mma_tile<TIER, 3, 1, 4, false>(a, tile, smem); break;
      }
    }
  }
}

// At most 128 VGPRs: four workgroups per CU, as many as the batch's LDS cap
// allows (the BF16 / FP16 register demand otherwise holds it to three).
template <int TIER>
__global__ __launch_bounds__(kThreads, 4) void batch_mma(const MmaArgs* __restrict__ args,
                                                      const BTile* __restrict__ tiles,
                                                      int* counters) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_tile;
  const int home = static_cast<int>(blockIdx.x) % kXcds;
  int visited = 0;  // thread 0: queues found empty so far
  for (;;) {
    if (threadIdx.x == 0) s_tile = pop_tile(counters, home, visited);
    __syncthreads();
    const int ti = wave_uniform(s_tile);  // the same in every thread: a uniform exit
    if (ti < 0) break;
    const BTile t = tiles[ti];
    const MmaArgs a = args[wave_uniform(t.grid)];
    FwdTile tile;
    tile.j1_0 = t.bx * kLanes;
    tile.j1_last = min(tile.j1_0 + kLanes - 1, a.m1 - 1);
    tile.r0 = t.by * a.rows_per_chunk;
    tile.r1 = min(tile.r0 + a.rows_per_chunk, a.m2);
    tile.j3 = t.bz;
    batch_mma_dispatch<TIER>(a, tile, smem);
    __syncthreads();  // LDS and s_tile are rewritten for the next tile
  }
}

using MmaKernel = void (*)(MmaArgs);

template <int TIER, int NIN>
MmaKernel pick_mma_nc(int nf, int nc, bool nurbs) {
  if (nurbs) {
    switch (nc) {
      case 1: return mma_stream<TIER, NIN, 1, 1, true>;
      case 2: return mma_stream<TIER, NIN, 1, 2, true>;
      case 3: return mma_stream<TIER, NIN, 1, 3, true>;
    }
    return nullptr;
  }
  if (nf == 1) {
    switch (nc) {
      case 1: return mma_stream<TIER, NIN, 1, 1, false>;
      case 2: return mma_stream<TIER, NIN, 1, 2, false>;
      case 3: return mma_stream<TIER, NIN, 1, 3, false>;
      case 4: return mma_stream<TIER, NIN, 1, 4, false>;
    }
    return nullptr;
  }
  if (nc != 1) return nullptr;
  if (nf == 2) return mma_stream<TIER, NIN, 2, 1, false>;
  if (nf == 3) return mma_stream<TIER, NIN, 3, 1, false>;
  if (nf == 4) return mma_stream<TIER, NIN, 4, 1, false>;
  return nullptr;
}

MmaKernel pick_mma(int tier, int nin, int nf, int nc, bool nurbs) {
  if (tier == SG_TIER_SPLIT_F16)
    return nin == 2 ? pick_mma_nc<SG_TIER_SPLIT_F16, 2>(nf, nc, nurbs)
                    : pick_mma_nc<SG_TIER_SPLIT_F16, 3>(nf, nc, nurbs);
  return nin == 2 ? pick_mma_nc<SG_TIER_BF16, 2>(nf, nc, nurbs)
                  : pick_mma_nc<SG_TIER_BF16, 3>(nf, nc, nurbs);
}

}  // namespace

bool mma_supported(const Plan& pl) { return pl.stream_ok && pl.mma.n_tiles1 > 0; }

hipError_t launch_forward_mma(const Plan& pl, int tier, int nf, const int32_t* orders,
                              const float* P, const float* W, float* const* outs,
                              hipStream_t s) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  const int nin = pl.nin;
  const bool nurbs = pl.nurbs != 0;
  MmaArgs a{};
  a.src.P = P;
  a.src.W = W;
  a.src.n1 = d1.n;
  a.src.n2 = d2.n;
  a.span1 = d1.span;
  a.span2 = d2.span;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.p1 = d1.p;
  a.p2 = d2.p;
  const Tiling& tl = pl.mma;
  a.rows_per_chunk = tl.chunk_rows();
  a.ntx = tl.n_tiles1;
  a.nty = tl.n_chunks2;
  a.xcd_map = tl.xcd_map;
  a.gmax = tl.w1max * tl.w2max;
  a.src.ncp = 1;
  a.nsamp = 1;
  for (int d = 0; d < nin; ++d) {
    a.src.ncp *= pl.dims[d].n;
    a.nsamp *= pl.dims[d].m;
  }
  unsigned gz = 1;
  if (nin == 3) {
    a.src.span3 = pl.dims[2].span;
    a.src.p3 = pl.dims[2].p;
    gz = static_cast<unsigned>(pl.dims[2].m);
  }
  const dim3 grid(tl.n_tiles1 * tl.n_chunks2 * gz);
  // Fields that share the dimension-3 order share stage 1: launch each group
  // of up to 4 fields (B-splines; NURBS take one field per launch) and each
  // channel block.
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
    for (int gi = 0; gi < gnf; ++gi) {
      const int f = group[gi];
      const int o1 = orders[f * nin + 0], o2 = orders[f * nin + 1];
      a.tab1[gi] = d1.ktab32 + static_cast<size_t>(o1) * d1.m * (d1.p + 1);
      a.tab2[gi] = d2.ktab32 + static_cast<size_t>(o2) * d2.m * (d2.p + 1);
      const int e1 = d1.tab_exp[o1], e2 = d2.tab_exp[o2];
      a.s1[gi] = ldexpf(1.f, -e1);
      a.s2[gi] = ldexpf(1.f, -e2);
      a.e12[gi] = e1 + e2;
      a.out[gi] = outs[f];
    }
    if (nin == 3)
      a.src.tab3 = pl.dims[2].ktab32 + static_cast<size_t>(orders[group[0] * nin + 2]) *
                                           pl.dims[2].m * (pl.dims[2].p + 1);
    for (int c0 = 0; c0 < pl.nout;) {
      const int left = pl.nout - c0;
      const int nc = gnf > 1 ? 1 : (nurbs ? std::min(left, 3) : std::min(left, 4));
      MmaKernel k = pick_mma(tier, nin, gnf, nc, nurbs);
      if (!k) return hipErrorInvalidValue;
      a.src.c0 = c0;
      hipLaunchKernelGGL(k, grid, dim3(kThreads), mma_lds(pl, nc + (nurbs ? 1 : 0), gnf), s, a);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      c0 += nc;
    }
  }
  return hipSuccess;
}

size_t mma_args_size() { return sizeof(MmaArgs); }

// The argument record of channels [c0, c0 + nc) of a plan's order-0 field
// for the batched launch (P, W, out: the grid's arrays).
void mma_fill_args(const Plan& pl, int c0, int nc, const float* P, const float* W, float* out,
                   void* record) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  MmaArgs a{};
  a.src.P = P;
  a.src.W = W;
  a.src.n1 = d1.n;
  a.src.n2 = d2.n;
  a.src.c0 = c0;
  a.span1 = d1.span;
  a.span2 = d2.span;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.p1 = d1.p;
  a.p2 = d2.p;
  const Tiling& tl = pl.mma;
  a.rows_per_chunk = tl.chunk_rows();
  a.ntx = tl.n_tiles1;
  a.nty = tl.n_chunks2;
  a.gmax = tl.w1max * tl.w2max;
  a.src.ncp = 1;
  a.nsamp = 1;
  for (int d = 0; d < pl.nin; ++d) {
    a.src.ncp *= pl.dims[d].n;
    a.nsamp *= pl.dims[d].m;
  }
  if (pl.nin == 3) {
    a.src.span3 = pl.dims[2].span;
    a.src.p3 = pl.dims[2].p;
    a.src.tab3 = pl.dims[2].ktab32;
  }
  a.tab1[0] = d1.ktab32;
  a.tab2[0] = d2.ktab32;
  a.s1[0] = ldexpf(1.f, -d1.tab_exp[0]);
  a.s2[0] = ldexpf(1.f, -d2.tab_exp[0]);
  a.e12[0] = d1.tab_exp[0] + d2.tab_exp[0];
  a.out[0] = out;
  a.nin = pl.nin;
  a.nc = nc;
  a.nurbs = pl.nurbs;
  std::memcpy(record, &a, sizeof(a));
}

static const void* batch_mma_kernel(int tier) {
  return tier == SG_TIER_SPLIT_F16 ? reinterpret_cast<const void*>(batch_mma<SG_TIER_SPLIT_F16>)
                                   : reinterpret_cast<const void*>(batch_mma<SG_TIER_BF16>);
}

int batch_mma_resident(int tier, size_t lds) {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batch_mma_kernel(tier), kThreads,
                                                   lds) != hipSuccess)
    return 0;
  return blocks;
}

hipError_t launch_batch_mma(int tier, const void* args, const void* tiles, int* counters,
                            int blocks, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, counters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const auto* a = static_cast<const MmaArgs*>(args);
  const auto* t = static_cast<const BTile*>(tiles);
  if (tier == SG_TIER_SPLIT_F16)
    hipLaunchKernelGGL(batch_mma<SG_TIER_SPLIT_F16>, dim3(blocks), dim3(kThreads), lds, s, a, t,
                       counters);
  else
    hipLaunchKernelGGL(batch_mma<SG_TIER_BF16>, dim3(blocks), dim3(kThreads), lds, s, a, t,
                       counters);
  return hipGetLastError();
}

}  // namespace sg
