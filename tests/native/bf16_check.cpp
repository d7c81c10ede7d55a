// Prints bf16_round_bits of every FP32 bit pattern given on stdin (one
// unsigned integer per line), for tests/test_native_host.py.
#include <cstdint>
#include <cstdio>

#include "sg_bf16.hpp"

int main() {
  unsigned long u = 0;
  while (std::scanf("%lu", &u) == 1)
    std::printf("%u\n", static_cast<unsigned>(sg::bf16_round_bits(static_cast<uint32_t>(u))));
  return 0;
}
