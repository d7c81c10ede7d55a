// sg_bf16.hpp -- FP32 -> BF16 rounding (bits in, bits out).
//
// Round to nearest, ties to even, on the FP32 bit pattern; NaNs become a
// quiet NaN of the same sign (a NaN whose payload sits only in the low 16
// bits would otherwise carry into the exponent and turn into an infinity).
// constexpr, so both the kernels and host-side tests use this one definition.
#pragma once

#include <cstdint>

namespace sg {

constexpr uint16_t bf16_round_bits(uint32_t u) {
  return (u & 0x7fffffffu) > 0x7f800000u
             ? static_cast<uint16_t>((u >> 16) | 0x40u)
             : static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

}  // namespace sg
