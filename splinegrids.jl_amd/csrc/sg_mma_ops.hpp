// sg_mma_ops.hpp -- MFMA operands of the split-FP16 and BF16 tiers.
//
//   SPLIT_F16  x (scaled by a power of two so that |x| < 1) is split into
//              hi = fp16(x) and lo = fp16(x - hi); a product takes
//              hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x16_f16 with FP32
//              accumulation (the lo*lo term is below 2^-22 relative).
//   BF16       one pass of v_mfma_f32_16x16x16bf16_1k on bf16(x).
//
// Operand layout (gfx9 16x16x16): lane l = 16 g + i holds A[i][4g..4g+3],
// B[4g..4g+3][i] and D[4g..4g+3][i].  The gfx950 16x16x32 forms (Operand8)
// take eight K values per lane; A and B pair their elements the same way,
// so a caller that puts the same K index in element e of both operands of a
// lane gets the product whatever the hardware's K order.
#pragma once

#include "sg_bf16.hpp"
#include "sg_internal.hpp"

namespace sg {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ inline short bf16_bits(float x) {  // round to nearest even, NaN stays NaN
  return static_cast<short>(bf16_round_bits(__float_as_uint(x)));
}

// The largest v over the wave (v >= 0), as a wave-uniform value.  DPP row
// shifts by 1, 2, 4 and 8 leave each row's maximum in its lane 15 (lanes
// without a source read 0), row_bcast:15 and row_bcast:31 carry it up to
// lane 63; all of it on the VALU, no LDS round trips.
template <int CTRL, int ROW_MASK>
__device__ inline float dpp_max(float v) {
  const int o = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, true);
  return fmaxf(v, __int_as_float(o));
}
__device__ inline float wave_max(float v) {
  v = dpp_max<0x111, 0xf>(v);  // row_shr:1
  v = dpp_max<0x112, 0xf>(v);  // row_shr:2
  v = dpp_max<0x114, 0xf>(v);  // row_shr:4
  v = dpp_max<0x118, 0xf>(v);  // row_shr:8
  v = dpp_max<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_max<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
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

// Eight K values per lane: v_mfma_f32_16x16x32_f16 / _bf16 (gfx950).
template <int TIER>
struct Operand8;

template <>
struct Operand8<SG_TIER_SPLIT_F16> {
  half8_t hi, lo;
  __device__ void set(int j, float x) {
    const _Float16 h = static_cast<_Float16>(x);
    hi[j] = h;
    lo[j] = static_cast<_Float16>(x - static_cast<float>(h));
  }
  __device__ static float4_t mma(const Operand8& a, const Operand8& b, float4_t c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.hi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.lo, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.lo, b.hi, c, 0, 0, 0);
    return c;
  }
};

template <>
struct Operand8<SG_TIER_BF16> {
  bf16x8_t v;
  __device__ void set(int j, float x) {
    v[j] = __builtin_bit_cast(__bf16, static_cast<unsigned short>(bf16_bits(x)));
  }
  __device__ static float4_t mma(const Operand8& a, const Operand8& b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
  }
};

}  // namespace sg
