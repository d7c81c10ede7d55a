// sg_mfma.hip -- evaluate!(grid) with both mode products on MFMA (FP32 arrays).
//
// Tiers
//   SPLIT_F16  every operand x is split into hi = fp16(x) and lo = fp16(x-hi)
//              after power-of-two scaling; products use hi*hi + hi*lo + lo*hi
//              on v_mfma_f32_16x16x16_f16 with FP32 accumulation (the lo*lo
//              term is below 2^-22 relative).
//   BF16       one pass of v_mfma_f32_16x16x16bf16_1k on bf16(x).
//
// Structure.  A workgroup tile is 64 consecutive samples j1, a chunk of rows
// j2 and (3-D grids) four consecutive samples j3, one per wave; in 2-D grids,
// and 3-D grids whose four-plane windows do not fit the LDS, one plane and
// the four waves split the chunk's rows.  Stage 1 stages the tile's band
// window of control points in LDS, contracted along dimension 3 for each of
// the four planes at once (every control-point plane is loaded once for all
// four), G[plane][c][k2][k1] (NURBS: w P and w).  Each wave then owns all 64
// samples j1 of the tile as four MFMA column blocks of 16, and walks its rows
// in blocks of 16.  With K = 16 consecutive control points k2 from kb:
//
//   step 1  T1^T[k2][j1] = sum_k1 G[k2][k1] * B1[j1][k1]      (16 x 16 x 16 per
//           column block; recomputed only when a row block's band leaves
//           [kb, kb + 16))
//   step 2  Out[j2][j1]  = sum_k2 B2[j2][k2] * T1^T[k2][j1]    (per row block)
//
// The output registers of step 1 are, lane for lane, the B operand of step 2
// (gfx9 16x16x16 layout: lane l = 16 g + i holds A[i][4g..4g+3],
// B[4g..4g+3][i], D[4g..4g+3][i]), so no LDS round trip sits between them.
// Step 2 leaves row 4g + r of column block b in register r of lane (g, i);
// one v_permlane16_swap and one v_permlane32_swap per register transpose the
// four column blocks so that register r of row group g' holds row
// 4g' + r for all 64 samples: every store instruction writes 256 contiguous
// bytes of one row, as the FP32 kernel's do.
//
// Scaling: G of each plane and channel by 2^-eG (max|G| < 1), the tables of
// each field by 2^-e1 and 2^-e2 (per dimension and derivative order); the
// results are scaled back exactly with ldexp.  NURBS divide by the weight
// channel (one hardware reciprocal per sample): the table scales cancel, the
// channel scales leave 2^(eG_c - eG_w).
#include <cstring>
#include <vector>

#include "sg_mma_ops.hpp"
#include "sg_queue.hpp"
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
  int m1, m2, m3, p1, p2;
  int rows_per_chunk;            // rows of a tile
  int rows_per_wave;             // one plane per tile: rows of each wave
  int planes;                    // samples j3 per tile: 4 (one per wave) or 1
  int ntx, nty, xcd_map;
  int w1max, w2max;              // the plan's largest windows (LDS layout)
  float s1[SG_MAX_FIELDS];       // 2^-e of the dimension 1 / 2 tables of each field
  float s2[SG_MAX_FIELDS];
  int e12[SG_MAX_FIELDS];        // e1 + e2 of each field
  long long nsamp;
  int nin, nc, nurbs;            // batched launches: what the record needs
};

constexpr int kPlanes3 = kWaves;  // most samples j3 per 3-D tile (one per wave)

// Tile of the MFMA kernels: 64 samples j1 from j1_0, rows [r0, r1), samples
// j3 from j3_0 (3-D).
struct MTile {
  int j1_0, j1_last, r0, r1, j3_0;
};

template <int NIN>
__host__ __device__ constexpr int planes_of() {
  return NIN == 3 ? kPlanes3 : 1;
}

// LDS of a tile, in bytes from the start (each region 16-byte aligned):
//   Gs   the band window of every plane and channel, scaled and split into
//        the tier's 16-bit parts: [plane][ch][part][W2 + 1 rows][W1p] (the
//        last row zero; W1p = W1 rounded up to 16, the padding zero)
//   B1s  the tile's 64 rows of B1 made dense over the window, split:
//        [f][part][64][W1p] (rows past the grid zero)
//   B2s  the chunk's B2 rows (p2 + 1 values each), split: [f][part][chunk][p2+1]
//   RS   the chunk's spans (int)
//   red  per-wave partial maxima of every plane and channel (float)
//   Gf   the FP32 window of stage 1: [plane][ch][W1 * W2]
struct MmaLds {
  size_t gs, b1s, b2s, rs, red, gf, total;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline MmaLds mma_lds_layout(int tier, int nplanes, int nch, int w1max,
                                                  int w2max, int nf, int chunk, int p2) {
  const size_t parts = tier == SG_TIER_SPLIT_F16 ? 2 : 1;
  const size_t w1p = static_cast<size_t>((w1max + 15) / 16 * 16);
  MmaLds L;
  L.gs = 0;
  L.b1s = align16(L.gs + static_cast<size_t>(nplanes) * nch * parts * (w2max + 1) * w1p * 2);
  L.b2s = align16(L.b1s + static_cast<size_t>(nf) * parts * kLanes * w1p * 2);
  L.rs = align16(L.b2s + static_cast<size_t>(nf) * parts * chunk * (p2 + 1) * 2);
  L.red = align16(L.rs + static_cast<size_t>(chunk) * 4);
  L.gf = align16(L.red + static_cast<size_t>(kWaves) * kPlanes3 * 5 * 4);
  L.total = align16(L.gf + static_cast<size_t>(nplanes) * nch * w1max * w2max * 4);
  return L;
}

// The 16-bit parts of a tier's operands in LDS.
template <int TIER>
struct Parts;

template <>
struct Parts<SG_TIER_SPLIT_F16> {
  static constexpr int kParts = 2;
  // hi = fp16(x), lo = fp16(x - hi) at p[0] and p[stride]
  __device__ static void put(unsigned short* p, size_t stride, float x) {
    const _Float16 h = static_cast<_Float16>(x);
    p[0] = __builtin_bit_cast(unsigned short, h);
    p[stride] = __builtin_bit_cast(unsigned short, static_cast<_Float16>(x - static_cast<float>(h)));
  }
  // four consecutive K values (8-byte aligned) of both parts
  __device__ static Operand<SG_TIER_SPLIT_F16> get(const unsigned short* p, size_t stride) {
    Operand<SG_TIER_SPLIT_F16> o;
    o.hi = *reinterpret_cast<const half4_t*>(p);
    o.lo = *reinterpret_cast<const half4_t*>(p + stride);
    return o;
  }
  __device__ static Operand<SG_TIER_SPLIT_F16> zero() {
    Operand<SG_TIER_SPLIT_F16> o;
    o.hi = half4_t{0, 0, 0, 0};
    o.lo = half4_t{0, 0, 0, 0};
    return o;
  }
};

template <>
struct Parts<SG_TIER_BF16> {
  static constexpr int kParts = 1;
  __device__ static void put(unsigned short* p, size_t, float x) {
    p[0] = static_cast<unsigned short>(bf16_bits(x));
  }
  __device__ static Operand<SG_TIER_BF16> get(const unsigned short* p, size_t) {
    Operand<SG_TIER_BF16> o;
    o.v = *reinterpret_cast<const short4_t*>(p);
    return o;
  }
  __device__ static Operand<SG_TIER_BF16> zero() {
    Operand<SG_TIER_BF16> o;
    o.v = short4_t{0, 0, 0, 0};
    return o;
  }
};

// Four registers x[b] (column block b, row group g of lane (g, i)) become
// four registers y[g'] (row group g' of every column block): lane l of
// y[g'] holds column 16 (l / 16) + l % 16 = l of row group g'.
__device__ inline void transpose_blocks(float (&x)[4]) {
  auto s01 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x[0]), __float_as_uint(x[1]),
                                              false, false);
  auto s23 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x[2]), __float_as_uint(x[3]),
                                              false, false);
  auto t02 = __builtin_amdgcn_permlane32_swap(s01[0], s23[0], false, false);
  auto t13 = __builtin_amdgcn_permlane32_swap(s01[1], s23[1], false, false);
  x[0] = __uint_as_float(t02[0]);
  x[1] = __uint_as_float(t13[0]);
  x[2] = __uint_as_float(t02[1]);
  x[3] = __uint_as_float(t13[1]);
}

// One tile of channels [src.c0, src.c0 + NC): stage 1 (the window of every
// plane in FP32 and its per-channel maxima), then every operand the waves
// need in LDS, scaled and split once per tile -- the window, the tile's B1
// rows made dense over the window, the chunk's B2 rows -- and then each
// wave's rows in blocks of 16, every operand read with 8-byte LDS loads.
template <int TIER, int NIN, int NF, int NC, bool NURBS>
__device__ __attribute__((always_inline)) void mma_tile(const MmaArgs& a, const MTile& tile,
                                                        unsigned char* smem) {
  constexpr int NCH = NC + (NURBS ? 1 : 0);
  constexpr int PL = planes_of<NIN>();
  constexpr int PARTS = Parts<TIER>::kParts;
  using Op = Operand<TIER>;

  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const int r0 = tile.r0, r1 = tile.r1;
  const int p1 = a.p1, p2 = a.p2;
  const int lo1 = a.span1[tile.j1_0] - p1;
  const int W1 = a.span1[tile.j1_last] - lo1 + 1;
  const int lo2 = a.span2[r0] - p2;
  const int W2 = a.span2[r1 - 1] - lo2 + 1;
  const int cs = W1 * W2;
  const int W1p = (W1 + 15) & ~15;
  const int nk1 = W1p >> 4;  // K blocks of step 1
  const bool split_rows = NIN == 2 || a.planes == 1;  // waves split the rows
  const int nplanes = NIN == 3 ? min(a.planes, a.m3 - tile.j3_0) : 1;
  const int n1 = tile.j1_last - tile.j1_0 + 1;  // samples j1 of the tile
  const int chunk = a.rows_per_chunk;
  const int nrows = r1 - r0;
  const MmaLds L = mma_lds_layout(TIER, NIN == 3 ? a.planes : 1, NCH, a.w1max, a.w2max, NF,
                                  chunk, p2);
  unsigned short* Gs = reinterpret_cast<unsigned short*>(smem + L.gs);
  unsigned short* B1s = reinterpret_cast<unsigned short*>(smem + L.b1s);
  unsigned short* B2s = reinterpret_cast<unsigned short*>(smem + L.b2s);
  int* RS = reinterpret_cast<int*>(smem + L.rs);
  float* red = reinterpret_cast<float*>(smem + L.red);  // [wave][plane][ch]
  float* Gf = reinterpret_cast<float*>(smem + L.gf);
  // part strides (elements) and per-plane/channel block sizes of Gs
  const size_t gs_rows = static_cast<size_t>(W2 + 1) * W1p;          // one part of one channel
  const size_t b1_part = static_cast<size_t>(kLanes) * W1p;
  const size_t b2_part = static_cast<size_t>(chunk) * (p2 + 1);

  // ---- stage 1: the FP32 windows and their maxima ----
  stage1_planes<float, NIN, NC, NURBS, PL>(Gf, a.src, tile.j3_0, nplanes, lo1, W1, lo2, W2);
  // the tile's B1 rows, dense over the window (zero outside each band and
  // for samples past the grid), and the chunk's B2 rows, both split
  for (int idx = threadIdx.x; idx < NF * kLanes * W1p; idx += kThreads) {
    const int f = idx / (kLanes * W1p);
    const int rem = idx - f * kLanes * W1p;
    const int i = rem / W1p, q1 = rem - i * W1p;
    float v = 0.f;
    if (i < n1) {
      const int j1 = tile.j1_0 + i;
      const int o = q1 - (a.span1[j1] - p1 - lo1);
      if (o >= 0 && o <= p1) v = a.tab1[f][static_cast<size_t>(j1) * (p1 + 1) + o] * a.s1[f];
    }
    Parts<TIER>::put(B1s + static_cast<size_t>(f) * PARTS * b1_part + rem, b1_part, v);
  }
  for (int idx = threadIdx.x; idx < NF * nrows * (p2 + 1); idx += kThreads) {
    const int f = idx / (nrows * (p2 + 1));
    const int rem = idx - f * nrows * (p2 + 1);
    const float v = a.tab2[f][static_cast<size_t>(r0) * (p2 + 1) + rem] * a.s2[f];
    Parts<TIER>::put(B2s + static_cast<size_t>(f) * PARTS * b2_part + rem, b2_part, v);
  }
  for (int i = threadIdx.x; i < nrows; i += kThreads) RS[i] = a.span2[r0 + i];
  {
    // per-wave maxima of every plane and channel of the window
    float m[PL][NCH];
#pragma unroll
    for (int w = 0; w < PL; ++w)
#pragma unroll
      for (int c = 0; c < NCH; ++c) m[w][c] = 0.f;
    __syncthreads();  // Gf is complete
    for (int idx = threadIdx.x; idx < cs; idx += kThreads)
#pragma unroll
      for (int w = 0; w < PL; ++w)
        if (w < nplanes)
#pragma unroll
          for (int c = 0; c < NCH; ++c) m[w][c] = fmaxf(m[w][c], fabsf(Gf[(w * NCH + c) * cs + idx]));
#pragma unroll
    for (int w = 0; w < PL; ++w)
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const float mw = wave_max(m[w][c]);
        if (lane == 0) red[(wave * kPlanes3 + w) * 5 + c] = mw;
      }
  }
  __syncthreads();  // red is complete
  // the scale exponents of every plane and channel: max |G| 2^-e < 1
  int eg[PL][NCH];
#pragma unroll
  for (int w = 0; w < PL; ++w)
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      float mx = 0.f;
#pragma unroll
      for (int v = 0; v < kWaves; ++v) mx = fmaxf(mx, red[(v * kPlanes3 + w) * 5 + c]);
      eg[w][c] = scale_exponent(mx);
    }
  // the windows scaled and split, with a zero row W2 and zero columns W1 .. W1p-1
  {
    const int cells = (W2 + 1) * W1p;
    for (int idx = threadIdx.x; idx < cells; idx += kThreads) {
      const int q2 = idx / W1p, q1 = idx - q2 * W1p;
      const bool in = q2 < W2 && q1 < W1;
#pragma unroll
      for (int w = 0; w < PL; ++w)
        if (w < nplanes)
#pragma unroll
          for (int c = 0; c < NCH; ++c) {
            const float v = in ? ldexpf(Gf[(w * NCH + c) * cs + q2 * W1 + q1], -eg[w][c]) : 0.f;
            Parts<TIER>::put(Gs + (static_cast<size_t>(w) * NCH + c) * PARTS * gs_rows + idx,
                             gs_rows, v);
          }
    }
  }
  __syncthreads();

  // ---- this wave: its plane and rows (no barrier follows) ----
  const int w = split_rows ? 0 : wave;
  if (w >= nplanes) return;
  int rb = r0, re = r1;
  if (split_rows) {
    rb = r0 + wave * a.rows_per_wave;
    re = min(rb + a.rows_per_wave, r1);
    if (rb >= re) return;
  }
  int egw[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) egw[c] = eg[0][c];
#pragma unroll
  for (int v = 1; v < PL; ++v)
    if (w == v)
#pragma unroll
      for (int c = 0; c < NCH; ++c) egw[c] = eg[v][c];
  const unsigned short* Gw = Gs + static_cast<size_t>(w) * NCH * PARTS * gs_rows;
  const int j3 = NIN == 3 ? tile.j3_0 + w : 0;
  const int g = lane >> 4, col = lane & 15;

  // step 1 for one field, channel, column block and k2 base kb (kb - lo2 a
  // multiple of 4): T1^T[kb + 4g + r][j1_0 + 16 b + col] in FP32
  auto step1 = [&](int f, int ch, int b, int kb) {
    float4_t d = float4_t{0.f, 0.f, 0.f, 0.f};
    const int row = min(kb + col - lo2, W2);  // A operand: row k2 = kb + col (W2: zeros)
    const unsigned short* ga = Gw + static_cast<size_t>(ch) * PARTS * gs_rows +
                               static_cast<size_t>(row) * W1p + 4 * g;
    const unsigned short* bb = B1s + static_cast<size_t>(f) * PARTS * b1_part +
                               static_cast<size_t>(16 * b + col) * W1p + 4 * g;
    for (int kk = 0; kk < nk1; ++kk)
      d = Op::mma(Parts<TIER>::get(ga + 16 * kk, gs_rows), Parts<TIER>::get(bb + 16 * kk, b1_part),
                  d);
    return d;
  };
  // step-2 A operand of field f for rows jb .. jb + 15 and k2 from base:
  // lane (g, col) holds B2[jb + col][base + 4g + r] (0 off the band)
  auto b2_operand = [&](int f, int jb, int base) {
    const int j2 = jb + col;
    const bool rin = j2 < re;
    const int jr = (rin ? j2 : re - 1) - r0;
    const int o0 = base + 4 * g - (RS[jr] - p2);
    const unsigned short* rt = B2s + static_cast<size_t>(f) * PARTS * b2_part +
                               static_cast<size_t>(jr) * (p2 + 1);
    unsigned short v[2][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = o0 + r;
      const bool ok = rin && o >= 0 && o <= p2;
      const int oc = min(max(o, 0), p2);
#pragma unroll
      for (int part = 0; part < PARTS; ++part) {
        const unsigned short x = rt[part * b2_part + oc];
        v[part][r] = ok ? x : static_cast<unsigned short>(0);
      }
    }
    Op aop;
    if constexpr (TIER == SG_TIER_SPLIT_F16) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        aop.hi[r] = __builtin_bit_cast(_Float16, v[0][r]);
        aop.lo[r] = __builtin_bit_cast(_Float16, v[1][r]);
      }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) aop.v[r] = static_cast<short>(v[0][r]);
    }
    return aop;
  };

  // cached step-1 results (split for step 2), valid for k2 base kb
  Op t1[NF][NCH][4];
  int kb = -(1 << 30);
  auto refresh_t1 = [&](int base) {
    kb = base;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const float4_t d = step1(f, ch, b, base);
#pragma unroll
          for (int r = 0; r < 4; ++r) t1[f][ch][b].set(r, d[r]);
        }
  };

  // output rows: a wave-uniform pointer per row, each lane one sample
  const long long obase = static_cast<long long>(tile.j1_0) +
                          static_cast<long long>(a.m1) * a.m2 * j3 + a.nsamp * a.src.c0;
  const bool active = lane < n1;

  for (int jb = rb; jb < re; jb += 16) {
    const int jl = min(jb + 15, re - 1);
    const int klo = RS[jb - r0] - p2, khi = RS[jl - r0];
    const int kb_al = lo2 + ((klo - lo2) & ~3);  // aligned base for klo
    const bool narrow = khi <= kb_al + 15;
    if (narrow && !(kb <= klo && khi <= kb + 15)) refresh_t1(kb_al);

    // one field and channel of the row block: acc[r][q] = row jb + 4q + r
    auto rows_of = [&](int f, int ch, const Op& aop, float (&acc)[4][4]) {
      float4_t d[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) d[b] = float4_t{0.f, 0.f, 0.f, 0.f};
      if (narrow) {
#pragma unroll
        for (int b = 0; b < 4; ++b) d[b] = Op::mma(aop, t1[f][ch][b], d[b]);
      } else {
        // the rows' bands span more than 16 control points: step 1 for
        // every 16-wide part of the window, not cached
        for (int kc = kb_al; kc <= khi; kc += 16) {
          const Op aopk = b2_operand(f, jb, kc);
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const float4_t t = step1(f, ch, b, kc);
            Op top;
#pragma unroll
            for (int r = 0; r < 4; ++r) top.set(r, t[r]);
            d[b] = Op::mma(aopk, top, d[b]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float x[4] = {d[0][r], d[1][r], d[2][r], d[3][r]};
        transpose_blocks(x);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][q] = x[q];
      }
    };

#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const Op aop = narrow ? b2_operand(f, jb, kb) : Parts<TIER>::zero();
      float inv[4][4];
      if constexpr (NURBS) {
        rows_of(f, NC, aop, inv);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 4; ++q) inv[r][q] = nurbs_recip(inv[r][q]);
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        float v[4][4];
        rows_of(f, c, aop, v);
        const int e = wave_uniform(NURBS ? egw[c] - egw[NC] : egw[c] + a.e12[f]);
        Global<float>* out = global_ptr(a.out[f] + obase + a.nsamp * c);
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int j2 = jb + 4 * q + r;
            float x = NURBS ? v[r][q] * inv[r][q] : v[r][q];
            if (e != 0) x = ldexpf(x, e);
            // a wave-uniform row pointer and a 32-bit lane offset: the store
            // takes its base from scalar registers
            Global<float>* row = out + static_cast<long long>(a.m1) * j2;
            if (j2 < re && active) row[lane] = x;
          }
      }
    }
  }
}

template <int TIER, int NIN, int NF, int NC, bool NURBS>
__global__ __launch_bounds__(kThreads) void mma_stream(MmaArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = xcd_logical_block(a.xcd_map, a.ntx);
  const int bx = L % a.ntx, by = (L / a.ntx) % a.nty;
  MTile tile;
  tile.j1_0 = bx * kLanes;
  tile.j1_last = min(tile.j1_0 + kLanes - 1, a.m1 - 1);
  tile.r0 = by * a.rows_per_chunk;
  tile.r1 = min(tile.r0 + a.rows_per_chunk, a.m2);
  tile.j3_0 = NIN == 3 ? (L / (a.ntx * a.nty)) * a.planes : 0;
  mma_tile<TIER, NIN, NF, NC, NURBS>(a, tile, smem);
}

// Batched launch of the MFMA tiers: a persistent kernel takes tiles of many
// grids from the work queues; each tile names the argument record of one
// grid's channel block (field: derivative order 0), the tile along
// dimension 1, the chunk along dimension 2 and the group of four samples j3.
template <int TIER>
__device__ __attribute__((always_inline)) void batch_mma_dispatch(const MmaArgs& a,
                                                                  const MTile& tile,
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

template <int TIER>
__global__ __launch_bounds__(kThreads) void batch_mma(const MmaArgs* __restrict__ args,
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
    MTile tile;
    tile.j1_0 = t.bx * kLanes;
    tile.j1_last = min(tile.j1_0 + kLanes - 1, a.m1 - 1);
    tile.r0 = t.by * a.rows_per_chunk;
    tile.r1 = min(tile.r0 + a.rows_per_chunk, a.m2);
    tile.j3_0 = t.bz * a.planes;
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

// The record of a plan's MFMA tiling (everything but the arrays and fields).
void fill_common(const Plan& pl, MmaArgs& a) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  a.src.n1 = d1.n;
  a.src.n2 = d2.n;
  a.span1 = d1.span;
  a.span2 = d2.span;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.m3 = pl.nin == 3 ? pl.dims[2].m : 1;
  a.p1 = d1.p;
  a.p2 = d2.p;
  const Tiling& tl = pl.mma;
  a.rows_per_chunk = tl.chunk_rows();
  a.rows_per_wave = tl.rows_per_wave;
  a.planes = pl.nin == 3 ? tl.j3_per_block : 1;
  a.ntx = tl.n_tiles1;
  a.nty = tl.n_chunks2;
  a.xcd_map = tl.xcd_map;
  a.w1max = tl.w1max;
  a.w2max = tl.w2max;
  a.src.ncp = 1;
  a.nsamp = 1;
  for (int d = 0; d < pl.nin; ++d) {
    a.src.ncp *= pl.dims[d].n;
    a.nsamp *= pl.dims[d].m;
  }
  if (pl.nin == 3) {
    a.src.span3 = pl.dims[2].span;
    a.src.p3 = pl.dims[2].p;
  }
  a.nin = pl.nin;
  a.nurbs = pl.nurbs;
}

}  // namespace

bool mma_supported(const Plan& pl) { return pl.stream_ok && pl.mma.n_tiles1 > 0; }

int mma_max_planes(int nin) { return nin == 3 ? kPlanes3 : 1; }

// LDS of an MFMA launch (nch channels, nf fields; sized for the split-FP16
// tier, whose operands have two parts).
size_t mma_lds(const Plan& pl, int nch, int nf) {
  const Tiling& t = pl.mma;
  const int planes = pl.nin == 3 ? t.j3_per_block : 1;
  return mma_lds_layout(SG_TIER_SPLIT_F16, planes, nch, t.w1max, t.w2max, nf, t.chunk_rows(),
                        pl.dims[1].p)
      .total;
}

// Fields per MFMA launch: the most (up to SG_MAX_FIELDS) whose LDS, with one
// channel, fits a workgroup.
int mma_fields(const Plan& pl) {
  const size_t cap = mma_lds(pl, 1 + pl.nurbs, 1) <= static_cast<size_t>(kMaxLds)
                         ? static_cast<size_t>(kMaxLds) : static_cast<size_t>(kMaxLdsHw);
  int nf = SG_MAX_FIELDS;
  while (nf > 1 && mma_lds(pl, 1 + pl.nurbs, nf) > cap) --nf;
  return nf;
}

// Value channels per MFMA launch: the most (up to 4, NURBS 3) whose LDS
// fits a workgroup.
int mma_channels(const Plan& pl, int nf) {
  const int most = nf > 1 ? 1 : (pl.nurbs ? 3 : 4);
  for (int nc = std::min(most, pl.nout); nc > 1; --nc)
    if (mma_lds(pl, nc + pl.nurbs, nf) <= static_cast<size_t>(kMaxLds)) return nc;
  return 1;
}

hipError_t launch_forward_mma(const Plan& pl, int tier, int nf, const int32_t* orders,
                              const float* P, const float* W, float* const* outs,
                              hipStream_t s) {
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  const int nin = pl.nin;
  const bool nurbs = pl.nurbs != 0;
  MmaArgs a{};
  fill_common(pl, a);
  a.src.P = P;
  a.src.W = W;
  const Tiling& tl = pl.mma;
  const unsigned gz = nin == 3 ? static_cast<unsigned>(tl.n_z) : 1u;
  const dim3 grid(tl.n_tiles1 * tl.n_chunks2 * gz);
  // Fields that share the dimension-3 order share stage 1: launch each group
  // of up to 4 fields (B-splines; NURBS take one field per launch) and each
  // channel block.
  std::vector<int> done(nf, 0);
  for (int f0 = 0; f0 < nf; ++f0) {
    if (done[f0]) continue;
    std::vector<int> group;
    for (int f = f0; f < nf && static_cast<int>(group.size()) < mma_fields(pl); ++f) {
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
      const int nc = std::min(left, mma_channels(pl, gnf));
      MmaKernel k = pick_mma(tier, nin, gnf, nc, nurbs);
      if (!k) return hipErrorInvalidValue;
      a.src.c0 = c0;
      a.nc = nc;
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
  fill_common(pl, a);
  a.src.P = P;
  a.src.W = W;
  a.src.c0 = c0;
  if (pl.nin == 3) a.src.tab3 = pl.dims[2].ktab32;
  a.tab1[0] = d1.ktab32;
  a.tab2[0] = d2.ktab32;
  a.s1[0] = ldexpf(1.f, -d1.tab_exp[0]);
  a.s2[0] = ldexpf(1.f, -d2.tab_exp[0]);
  a.e12[0] = d1.tab_exp[0] + d2.tab_exp[0];
  a.out[0] = out;
  a.nc = nc;
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
