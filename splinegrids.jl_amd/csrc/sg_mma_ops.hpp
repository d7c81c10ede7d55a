// sg_mma_ops.hpp -- MFMA operands of the split-FP16 and BF16 tiers.
//
//   SPLIT_F16  x (scaled by a power of two so that |x| < 1) is split into
//              hi = fp16(x) and lo = fp16(x - hi); a product takes
//              hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x16_f16 with FP32
//              accumulation (the lo*lo term is below 2^-22 relative).
//   BF16       one pass of v_mfma_f32_16x16x16bf16_1k on bf16(x).
//
// Operand layout (gfx9 16x16x16): lane l = 16 g + i holds A[i][4g..4g+3],
// B[4g..4g+3][i] and D[4g..4g+3][i].
#pragma once

#include "sg_internal.hpp"

namespace sg {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ inline short bf16_bits(float x) {  // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<short>(u >> 16);
}

// The largest v over the wave (v >= 0), in every lane.
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

// e with m * 2^-e in [0.5, 1): the scale 2^-e brings values up to m below 1.
__device__ inline int scale_exponent(float m) {
  int e = 0;
  if (m > 0.f) frexpf(m, &e);
  return e;
}

template <int TIER>
struct Operand;

template <>
struct Operand<SG_TIER_SPLIT_F16> {
  half4_t hi, lo;
  __device__ void set(int j, float x) {
    const _Float16 h = static_cast<_Float16>(x);
    hi[j] = h;
    lo[j] = static_cast<_Float16>(x - static_cast<float>(h));
  }
  __device__ static float4_t mma(const Operand& a, const Operand& b, float4_t c) {
    c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.lo, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.lo, b.hi, c, 0, 0, 0);
    return c;
  }
};

template <>
struct Operand<SG_TIER_BF16> {
  short4_t v;
  __device__ void set(int j, float x) { v[j] = bf16_bits(x); }
  __device__ static float4_t mma(const Operand& a, const Operand& b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.v, b.v, c, 0, 0, 0);
  }
};

}  // namespace sg
