// sg_adjoint_mma.hip -- evaluate_adjoint!(grid) for FP32 arrays on matrix
// cores: Pbar = M^T ebar without atomics (src/adjoint.jl:1-40, :52-83).
//
// The two innermost contractions of the adjoint are small matrix products.
// For a tile of 64 samples j1 whose bands cover at most 16 control points k1
// (from lo1), 16 rows j2 (a row block) and one sample j3:
//
//   T1[j2][k1] = sum_j1 ebar[j1, j2, j3] * B1[j1, k1]     (16 x 64) (64 x 16)
//   T2[k2][k1] = sum_j2 B2[j2, k2] * T1[j2][k1]           (16 x 16) (16 x 16)
//
// Both run on v_mfma_f32_16x16x4f32 (FP32 products, FP32 sums).  With
// lane l = 16 g + i, that instruction takes A[i][g] and B[g][i] from lane l
// and leaves D[4 g + r][i] in register r.  So:
//   GEMM1, K-step s (16 steps):  A = ebar[j1_0 + j1_of(g, s), row i]
//                                B = B1[j1_0 + j1_of(g, s)][lo1 + i]
//     (j1_of(g, s) = 16 (s / 4) + 4 g + s % 4: lane l reads four 16-byte
//     pieces of its row, and the four lanes of a row read 64 contiguous
//     bytes per load instruction)
//   GEMM2, K-step r (4 steps, row j2 = 4 g + r of the block):
//                                A = B2[4 g + r][lo2 + i]
//                                B = T1[4 g + r][i] = register r of GEMM1
//   D2: T2[lo2 + 4 g + r][lo1 + i], summed over the row blocks of the chunk.
// T2 of each j3 is weighted by B3[j3, d] into a register ring of p3 + 1
// slabs of k3; a slab whose k3 leaves the band of the next j3 is final for
// this task and goes to its partial window.  adj_gather (sg_adjoint.hip)
// then sums the few partial windows that hold each control point.
//
// One wave is one task (64 samples x a chunk of up to 64 rows x a range of
// samples j3, every channel in turn); waves share nothing, so there is no
// LDS and no barrier.  Only one channel's T2 and ring are live at a time;
// the ebar buffers take most registers (two waves per SIMD): a wave keeps
// kBufs - 1 row blocks of ebar in flight ahead of the one it multiplies,
// across row blocks, samples j3 and channels.
#include "sg_mma_ops.hpp"

namespace sg {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kTile1 = 64;     // samples j1 per task
constexpr int kMmaWin = kAdjMmaWin;  // control points per MFMA tile edge
constexpr int kBufs = 8;  // ebar units per unrolled group; kBufs - 1 are in flight

struct AdjMmaArgs {
  const float* E;
  const float* S;  // scale field [m3][m2][m1] applied to E as it loads (SCALE)
  float* part;
  const float* tab1;  // [m1][p1+1]
  const float* tab2;  // [m2][p2+1]
  const float* tab3;  // [m3][P3+1]; 2-D grids pass m3 = 1, P3 = 0, B3 = 1
  const int32_t* span1;
  const int32_t* span2;
  const int32_t* span3;
  int m1, m2, m3, p1, p2;
  int c0, nc, nout;  // channels c0 .. c0 + nc - 1 of nout
  int chunk, j3_per_task;
  int ntx, nty, ntasks, w3max;
  long long m12, nsamp;
};

// K-step s of GEMM1 takes sample j1_0 + j1_of(g, s): the four lanes of a row
// (g = 0..3) read 64 contiguous bytes in each 16-byte load q = s / 4.
__device__ inline int j1_of(int g, int s) { return 16 * (s >> 2) + 4 * g + (s & 3); }

// ebar is read with raw buffer loads: one descriptor per (sample j3,
// channel) plane and fixed per-lane byte offsets.  An offset at or past the
// plane's size returns 0 without a memory access, which masks rows past the
// chunk and samples past m1 without a branch (branches around loads make
// the compiler wait for every load in flight at the next merge).
constexpr unsigned kOff = 0x80000000u;   // always past a plane (< 2 GiB)
constexpr int kRsrcWord3 = 0x00020000;   // raw buffer, 32-bit elements (gfx9)

// SCALE: ebar times the field S, loaded at the same offsets into buffers of
// its own; half as many units in flight keeps the registers of both.
// CM (every tier and the scaled form, p3 <= kCmMaxP3): the units go
// sample by sample, the channels (at most kCmChannels per launch) of a
// sample together, each channel with a ring of its own: B3 rows and ring
// moves once per sample rather than once per channel and sample, and a unit
// of S is read again by the next channel while it is still in L2 rather
// than once per channel pass over the task (C4: 95.3 -> 90.6 us; the split-FP16
// and BF16 tiers 100 -> 98 and 95 us).
constexpr int kCmMaxP3 = 3;
constexpr int kCmChannels = 3;

template <int P3, int NB, bool VEC, int TIER, bool SCALE>
__global__ __launch_bounds__(kThreads) void adj_mma(AdjMmaArgs a) {
  constexpr int KB = SCALE ? kBufs / 2 : kBufs;  // buffer positions
  constexpr bool CM = P3 <= kCmMaxP3;
  constexpr int NCM = CM ? kCmChannels : 1;
  static_assert(KB % NB == 0, "row block of a buffer position must be constant");
  const int lane = threadIdx.x & (kLanes - 1);
  const int task = static_cast<int>(blockIdx.x) * kWaves +
                   __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  if (task >= a.ntasks) return;  // whole wave; nothing below synchronises
  const int g = lane >> 4, i = lane & 15;
  const int bx = task % a.ntx, by = (task / a.ntx) % a.nty, bz = task / (a.ntx * a.nty);
  const int j1_0 = bx * kTile1;
  const int r0 = by * a.chunk, r1 = min(r0 + a.chunk, a.m2);
  const int j3a = bz * a.j3_per_task, j3b = min(j3a + a.j3_per_task, a.m3);
  const int p1 = a.p1, p2 = a.p2;
  const int lo1 = a.span1[j1_0] - p1;
  const int lo2 = a.span2[r0] - p2;
  const int W1 = a.span1[min(j1_0 + kTile1, a.m1) - 1] - lo1 + 1;  // <= 16
  const int W2 = a.span2[r1 - 1] - lo2 + 1;                          // <= 16
  const int lo3 = a.span3[j3a] - P3;

  // Operand tables.  Every load below is unconditional, at a clamped index,
  // and multiplied by a 0/1 mask afterwards (table values are finite), so
  // that each group of loads is issued at once and costs one memory latency;
  // a select would let the compiler move each load under its own branch and
  // wait for it there.
  // GEMM1 B operand: B1 at (j1_0 + j1_of(g, s), lo1 + i), 0 off the band
  float b1[16];
  {
    int sp[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) sp[s] = a.span1[min(j1_0 + j1_of(g, s), a.m1 - 1)];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int j1 = j1_0 + j1_of(g, s);
      const int off = lo1 + i - (sp[s] - p1);
      const float v = a.tab1[static_cast<size_t>(min(j1, a.m1 - 1)) * (p1 + 1) +
                             min(max(off, 0), p1)];
      b1[s] = v * ((j1 < a.m1 && off >= 0 && off <= p1) ? 1.f : 0.f);
    }
  }
  // MFMA tiers: GEMM1 on 16x16x32 FP16 / BF16 products.  K block h of a
  // unit holds samples j1_0 + 32 h + 0..31; element e of lane (g, i) is
  // sample 16 (2 h + e / 4) + 4 g + e % 4 in both operands -- element e % 4
  // of its 16-byte load 2 h + e / 4, and K-step 8 h + e of the FP32 path.
  // B1 is scaled by 2^-eb (split FP16: max|B1| of the task below 1) and
  // converted once per task.
  Operand8<TIER == SG_TIER_NATIVE ? SG_TIER_BF16 : TIER> b1op[2];
  int eb = 0;
  if constexpr (TIER != SG_TIER_NATIVE) {
    float sb = 1.f;
    if constexpr (TIER == SG_TIER_SPLIT_F16) {
      float m = 0.f;
#pragma unroll
      for (int s = 0; s < 16; ++s) m = fmaxf(m, fabsf(b1[s]));
      eb = scale_exponent(wave_max(m));
      sb = ldexpf(1.f, -eb);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < 8; ++e) b1op[h].set(e, b1[8 * h + e] * sb);
  }
  // GEMM2 A operand of row block b, K-step r: B2 at (r0 + 16 b + 4 g + r, lo2 + i)
  float b2[NB][4];
  {
    int sp[NB][4];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) sp[b][r] = a.span2[min(r0 + 16 * b + 4 * g + r, r1 - 1)];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j2 = r0 + 16 * b + 4 * g + r;
        const int off = lo2 + i - (sp[b][r] - p2);
        const float v = a.tab2[static_cast<size_t>(min(j2, r1 - 1)) * (p2 + 1) +
                               min(max(off, 0), p2)];
        b2[b][r] = v * ((j2 < r1 && off >= 0 && off <= p2) ? 1.f : 0.f);
      }
  }

  // Units in order: channel c outermost, then sample j3, then row block b;
  // unit (c, j3, b) is row r0 + 16 b + i (lane (g, i)) of sample j3 and
  // channel c0 + c, and load q holds samples j1_0 + 16 q + 4 g .. + 3.
  const int J = j3b - j3a;
  const int U = a.nc * J * NB;
  // byte offset of (row block b, load q) within a plane, or kOff
  unsigned voff[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j2 = r0 + 16 * b + i;
    voff[b] = j2 < r1 ? static_cast<unsigned>((j2 * a.m1 + j1_0 + 4 * g) * 4) : kOff;
  }
  const unsigned plane_bytes = static_cast<unsigned>(a.m12 * 4);
  // Load cursor: the (channel, sample) plane of the next unit to load, kept
  // as running scalar state (no division per load).  Units past the last
  // read nothing: their descriptor has no records.
  int lc = 0, ljl = 0;
  const float* lplane = a.E + static_cast<long long>(j3a) * a.m12 + a.nsamp * a.c0;
  const float* splane = SCALE ? a.S + static_cast<long long>(j3a) * a.m12 : nullptr;
  auto load_plane = [&](const float* plane, f4 (&e)[4], int b) {
    // units past the task's last read nothing: their descriptor has no records
    const bool live = CM ? ljl < J : lc < a.nc;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(plane), 0, static_cast<int>(live ? plane_bytes : 0u), kRsrcWord3);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j1 = j1_0 + 16 * q + 4 * g;
      if constexpr (VEC) {
        const unsigned o = j1 < a.m1 ? voff[b] + 64u * q : kOff;
        // the b128 / b32 loads return raw 32-bit words: reinterpret, not convert
        e[q] = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(rs, static_cast<int>(o), 0, 0));
      } else {
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          const unsigned o = j1 + w < a.m1 ? voff[b] + 64u * q + 4u * w : kOff;
          e[q][w] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(o), 0, 0));
        }
      }
    }
  };
  auto load = [&](f4 (&e)[4], f4 (&d)[4], int b) {
    load_plane(lplane, e, b);
    if constexpr (SCALE) load_plane(splane, d, b);
    if (b == NB - 1) {  // the plane's last row block: advance the cursor
      if constexpr (CM) {  // units (sample, channel, row block)
        if (++lc == a.nc) {
          lc = 0;
          ++ljl;
          splane += a.m12;
        }
        lplane = a.E + static_cast<long long>(j3a + ljl) * a.m12 + a.nsamp * (a.c0 + lc);
      } else {  // units (channel, sample, row block)
        lplane += a.m12;
        if constexpr (SCALE) splane += a.m12;
        if (++ljl == J) {
          ljl = 0;
          ++lc;
          lplane = a.E + static_cast<long long>(j3a) * a.m12 + a.nsamp * (a.c0 + lc);
          if constexpr (SCALE) splane = a.S + static_cast<long long>(j3a) * a.m12;
        }
      }
    }
  };

  // Units are taken in unrolled groups of KB with one register buffer per
  // position: at position k the unit KB - 1 ahead is loaded into the
  // buffer that position k - 1 has just used, so KB - 1 units are in
  // flight and no buffer is ever copied (a copy of a register whose load is
  // in flight waits for that load).  KB % NB == 0 keeps the row block of
  // position k, k % NB, a constant.
  f4 buf[KB][4];
  f4 sbuf[SCALE ? KB : 1][4];  // the scale of each buffered unit (SCALE)
#pragma unroll
  for (int k = 0; k < KB - 1; ++k) load(buf[k], sbuf[SCALE ? k : 0], k % NB);

  f4 R[P3 + 1];  // ring of the current channel: slot q holds k3 = cur3 - P3 + q
#pragma unroll
  for (int q = 0; q <= P3; ++q) R[q] = f4{0.f, 0.f, 0.f, 0.f};
  // CM: a ring per channel, and the B3 row of the current sample
  f4 RC[NCM][CM ? P3 + 1 : 1];
  float b3c[P3 + 1];
  if constexpr (CM) {
#pragma unroll
    for (int ch = 0; ch < NCM; ++ch)
#pragma unroll
      for (int q = 0; q <= P3; ++q) RC[ch][q] = f4{0.f, 0.f, 0.f, 0.f};
  }
  int cur3 = a.span3[j3a];
  f4 acc = {0.f, 0.f, 0.f, 0.f};  // T2 of the current sample j3
  // B3 row of the next sample, loaded one sample ahead (before that sample's
  // ebar loads, so that waiting for it drains none of them)
  float b3n[P3 + 1];
#pragma unroll
  for (int q = 0; q <= P3; ++q) b3n[q] = a.tab3[static_cast<size_t>(j3a) * (P3 + 1) + q];

  // slot k3 of channel c is final: lane (g, i) writes
  // part[task][k3 - lo3][c0 + c][4 g + r][i] inside the task's W2 x W1
  // window (the gather reads no other entry)
  auto flush = [&](const f4& slab, int c, int k3) {
    float* dst = a.part + ((static_cast<size_t>(task) * a.w3max + (k3 - lo3)) * a.nout +
                           a.c0 + c) * (kMmaWin * kMmaWin);
    if (i >= W1) return;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (4 * g + r < W2) dst[(4 * g + r) * kMmaWin + i] = slab[r];
  };

  // compute cursor: channel and sample of the unit being multiplied
  int cc = 0, cjl = 0;
  for (int u0 = 0; u0 < U; u0 += KB) {
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int u = u0 + k;
      load(buf[(k + KB - 1) % KB], sbuf[SCALE ? (k + KB - 1) % KB : 0], (k + KB - 1) % NB);
      if (u < U) {
        const int b = k % NB;
        if constexpr (SCALE) {
#pragma unroll
          for (int q = 0; q < 4; ++q) buf[k][q] *= sbuf[k][q];
        }
        f4 d1 = {0.f, 0.f, 0.f, 0.f};
        if constexpr (TIER == SG_TIER_NATIVE) {
#pragma unroll
          for (int s = 0; s < 16; ++s)
            d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(buf[k][s >> 2][s & 3], b1[s], d1, 0, 0, 0);
        } else {
          // split FP16: the unit's ebar scaled by 2^-eu (max below 1), and
          // T1 scaled back by 2^(eu + eb); BF16 needs no scaling
          int eu = 0;
          float su = 1.f;
          if constexpr (TIER == SG_TIER_SPLIT_F16) {
            float m = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
              for (int t = 0; t < 4; ++t) m = fmaxf(m, fabsf(buf[k][q][t]));
            eu = scale_exponent(wave_max(m));
            su = ldexpf(1.f, -eu);
          }
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            Operand8<TIER> op;
#pragma unroll
            for (int e = 0; e < 8; ++e) op.set(e, buf[k][2 * h + e / 4][e % 4] * su);
            d1 = Operand8<TIER>::mma(op, b1op[h], d1);
          }
          if constexpr (TIER == SG_TIER_SPLIT_F16) {
#pragma unroll
            for (int r = 0; r < 4; ++r) d1[r] = ldexpf(d1[r], eu + eb);
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(b2[b][r], d1[r], acc, 0, 0, 0);
        if (CM && b == NB - 1) {
          // T2 of sample j3, channel c is complete: at the sample's first
          // channel every ring moves to the sample's band; then channel c's
          // ring takes it
          const int c = cc, jl = cjl;
          const int j3 = j3a + jl;
          if (++cc == a.nc) {
            cc = 0;
            ++cjl;
          }
          if (c == 0) {
#pragma unroll
            for (int q = 0; q <= P3; ++q) b3c[q] = b3n[q];
            const int j3n = jl + 1 < J ? j3 + 1 : j3;
#pragma unroll
            for (int q = 0; q <= P3; ++q)
              b3n[q] = a.tab3[static_cast<size_t>(j3n) * (P3 + 1) + q];
            const int s3 = a.span3[j3];
            while (cur3 < s3) {
#pragma unroll
              for (int ch = 0; ch < NCM; ++ch) {
                if (ch < a.nc) flush(RC[ch][0], ch, cur3 - P3);
#pragma unroll
                for (int q = 0; q < P3; ++q) RC[ch][q] = RC[ch][q + 1];
                RC[ch][P3] = f4{0.f, 0.f, 0.f, 0.f};
              }
              ++cur3;
            }
          }
#pragma unroll
          for (int ch = 0; ch < NCM; ++ch)
            if (ch == c) {  // c is uniform: a scalar branch, no register indexing
#pragma unroll
              for (int q = 0; q <= P3; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) RC[ch][q][r] += b3c[q] * acc[r];
            }
          acc = f4{0.f, 0.f, 0.f, 0.f};
          if (jl == J - 1 && c == a.nc - 1) {  // the task's last unit: every ring is final
#pragma unroll
            for (int ch = 0; ch < NCM; ++ch)
              if (ch < a.nc) {
#pragma unroll
                for (int q = 0; q <= P3; ++q) flush(RC[ch][q], ch, cur3 - P3 + q);
              }
          }
        } else if (!CM && b == NB - 1) {
          // T2 of sample j3 is complete: weight it into the ring
          const int c = cc, jl = cjl;
          const int j3 = j3a + jl;
          if (++cjl == J) {
            cjl = 0;
            ++cc;
          }
          float b3[P3 + 1];
#pragma unroll
          for (int q = 0; q <= P3; ++q) b3[q] = b3n[q];
          const int j3n = jl + 1 < J ? j3 + 1 : j3a;  // the next channel restarts at j3a
#pragma unroll
          for (int q = 0; q <= P3; ++q) b3n[q] = a.tab3[static_cast<size_t>(j3n) * (P3 + 1) + q];
          const int s3 = a.span3[j3];
          while (cur3 < s3) {
            flush(R[0], c, cur3 - P3);
#pragma unroll
            for (int q = 0; q < P3; ++q) R[q] = R[q + 1];
            R[P3] = f4{0.f, 0.f, 0.f, 0.f};
            ++cur3;
          }
#pragma unroll
          for (int q = 0; q <= P3; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) R[q][r] += b3[q] * acc[r];
          acc = f4{0.f, 0.f, 0.f, 0.f};
          if (jl == J - 1) {  // last sample of channel c: the ring is final
#pragma unroll
            for (int q = 0; q <= P3; ++q) {
              flush(R[q], c, cur3 - P3 + q);
              R[q] = f4{0.f, 0.f, 0.f, 0.f};
            }
            cur3 = a.span3[j3a];
          }
        }
      }
    }
  }
}

using AdjMmaKernel = void (*)(AdjMmaArgs);

template <int P3, int NB, bool VEC>
AdjMmaKernel pick_tier(int tier, bool scale) {
  if (scale)  // the scaled form: native tier only
    return tier == SG_TIER_NATIVE ? adj_mma<P3, NB, VEC, SG_TIER_NATIVE, true> : nullptr;
  switch (tier) {
    case SG_TIER_NATIVE: return adj_mma<P3, NB, VEC, SG_TIER_NATIVE, false>;
    case SG_TIER_SPLIT_F16: return adj_mma<P3, NB, VEC, SG_TIER_SPLIT_F16, false>;
    case SG_TIER_BF16: return adj_mma<P3, NB, VEC, SG_TIER_BF16, false>;
  }
  return nullptr;
}

template <int P3, int NB>
AdjMmaKernel pick_vec(bool vec, int tier, bool scale) {
  return vec ? pick_tier<P3, NB, true>(tier, scale) : pick_tier<P3, NB, false>(tier, scale);
}

template <int P3>
AdjMmaKernel pick_nb(int nb, bool vec, int tier, bool scale) {
  switch (nb) {
    case 1: return pick_vec<P3, 1>(vec, tier, scale);
    case 2: return pick_vec<P3, 2>(vec, tier, scale);
    case 4: return pick_vec<P3, 4>(vec, tier, scale);
  }
  return nullptr;
}

AdjMmaKernel pick_adj_mma(int p3, int nb, bool vec, int tier, bool scale = false) {
  switch (p3) {
    case 0: return pick_nb<0>(nb, vec, tier, scale);
    case 1: return pick_nb<1>(nb, vec, tier, scale);
    case 2: return pick_nb<2>(nb, vec, tier, scale);
    case 3: return pick_nb<3>(nb, vec, tier, scale);
    case 4: return pick_nb<4>(nb, vec, tier, scale);
    case 5: return pick_nb<5>(nb, vec, tier, scale);
  }
  return nullptr;
}

int plan_p3(const Plan& pl) { return pl.nin == 3 ? pl.dims[2].p : 0; }

}  // namespace

bool adjoint_mma_shape_ok(const Plan& pl) {
  // one (sample j3, channel) plane of ebar must stay below the out-of-range
  // offset of the buffer loads
  const long long m12 = static_cast<long long>(pl.dims[0].m) * pl.dims[1].m;
  return (pl.nin == 2 || pl.nin == 3) && plan_p3(pl) <= kMaxStreamDegree &&
         m12 * 4 < static_cast<long long>(kOff) - 64;
}

int adjoint_mma_resident_waves(const Plan& pl) {
  AdjMmaKernel k = pick_adj_mma(plan_p3(pl), pl.adjm.chunk_rows() / 16, true, SG_TIER_NATIVE);
  if (!k) return 0;
  int blocks = 0, cus = 0, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(k),
                                                   kThreads, 0) != hipSuccess)
    return 0;
  return blocks * kWaves * cus;
}

size_t adjoint_mma_work_bytes(const Plan& pl) {
  const Tiling& t = pl.adjm;
  if (t.n_tiles1 == 0) return 0;
  const size_t units = static_cast<size_t>(t.n_tiles1) * t.n_chunks2 * t.n_z;
  return units * t.w3max * pl.nout * kMmaWin * kMmaWin * sizeof(float);
}

hipError_t launch_adjoint_mma(const Plan& pl, int tier, const int32_t* orders, const float* E,
                              float* Pbar, int accumulate, hipStream_t s, const float* scale) {
  const Tiling& tl = pl.adjm;
  const DimDev& d1 = pl.dims[0];
  const DimDev& d2 = pl.dims[1];
  AdjMmaArgs a{};
  a.E = E;
  a.S = scale;
  a.part = static_cast<float*>(pl.work);
  a.tab1 = d1.ktab32 + static_cast<size_t>(orders[0]) * d1.m * (d1.p + 1);
  a.tab2 = d2.ktab32 + static_cast<size_t>(orders[1]) * d2.m * (d2.p + 1);
  a.span1 = d1.span;
  a.span2 = d2.span;
  a.m1 = d1.m;
  a.m2 = d2.m;
  a.m3 = 1;
  a.p1 = d1.p;
  a.p2 = d2.p;
  if (pl.nin == 3) {
    const DimDev& d3 = pl.dims[2];
    a.tab3 = d3.ktab32 + static_cast<size_t>(orders[2]) * d3.m * (d3.p + 1);
    a.span3 = d3.span;
    a.m3 = d3.m;
  } else {  // one sample j3 = 0 with span 0 and B3 = 1
    a.tab3 = pl.unit_tab;
    a.span3 = pl.unit_span;
  }
  a.nout = pl.nout;
  a.chunk = tl.chunk_rows();
  a.j3_per_task = tl.j3_per_block;
  a.ntx = tl.n_tiles1;
  a.nty = tl.n_chunks2;
  a.ntasks = tl.n_tiles1 * tl.n_chunks2 * tl.n_z;
  a.w3max = tl.w3max;
  a.m12 = static_cast<long long>(d1.m) * d2.m;
  a.nsamp = a.m12 * a.m3;
  // 16-byte loads need 16-byte rows: m1 % 4 == 0 and an aligned base
  const bool vec = (d1.m % 4 == 0) && (reinterpret_cast<uintptr_t>(E) % 16 == 0) &&
                   (reinterpret_cast<uintptr_t>(scale) % 16 == 0);
  const dim3 grid(static_cast<unsigned>((a.ntasks + kWaves - 1) / kWaves));
  AdjMmaKernel k = pick_adj_mma(plan_p3(pl), tl.chunk_rows() / 16, vec, tier, scale != nullptr);
  if (!k) return hipErrorInvalidValue;
  // channels in the middle: kCmChannels per launch
  const int per = plan_p3(pl) <= kCmMaxP3 ? kCmChannels : pl.nout;
  hipError_t e = hipSuccess;
  for (int c0 = 0; c0 < pl.nout && e == hipSuccess; c0 += per) {
    a.c0 = c0;
    a.nc = std::min(per, pl.nout - c0);
    hipLaunchKernelGGL(k, grid, dim3(kThreads), 0, s, a);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return e;
  return launch_adjoint_gather(pl, tl, pl.work, Pbar, SG_F32, accumulate, s);
}

}  // namespace sg
