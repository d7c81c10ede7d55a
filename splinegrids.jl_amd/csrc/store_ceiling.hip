// store_ceiling.hip -- the write bandwidth an MI355X reaches for the byte
// count of the 512 x 512 x 256 x 3 FP32 forward (805 MB), written (a) as a
// contiguous fill with 4-, 8- and 16-byte stores per lane, and (b) in the
// forward kernel's own pattern: tiles of 64 samples j1 (256 contiguous
// bytes of a row) x rows j2 x one sample j3, every row storing the three
// channel planes, with nothing computed, with workgroups numbered as the
// forward numbers them by default (j1 tiles fastest, then chunks, then j3,
// dealt round-robin over the XCDs) and (c) with each XCD taking one
// contiguous eighth of the tiles (the forward's XCD map 1).  (b) is the
// ceiling the forward kernel's store stream can reach; bench.py --workload
// ceiling runs this program and prints its lines.  Output: one JSON object
// per measurement.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int M1 = 512, M2 = 512, M3 = 256, NC = 3;
constexpr long long N = static_cast<long long>(M1) * M2 * M3 * NC;

template <int V>
__global__ __launch_bounds__(256) void fill(float* __restrict__ out) {
  const long long i = (static_cast<long long>(blockIdx.x) * 256 + threadIdx.x) * V;
  if (i + V > N) return;
  if constexpr (V == 4)
    *reinterpret_cast<float4*>(out + i) = make_float4(1.f, 1.f, 1.f, 1.f);
  else if constexpr (V == 2)
    *reinterpret_cast<float2*>(out + i) = make_float2(1.f, 1.f);
  else
    out[i] = 1.f;
}

// the forward's tiles: 64 lanes along j1, 4 waves of `rows` rows along j2,
// one sample j3; consecutive tiles walk j1 tiles, then chunks, then j3; with
// xcd_contiguous, workgroup b (on XCD b mod 8) takes tile (b mod 8) n / 8 +
// b / 8 of the n tiles (n is a multiple of 8 here)
__global__ __launch_bounds__(256) void pattern(float* __restrict__ out, int rows,
                                               int xcd_contiguous) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ntx = M1 / 64, nty = M2 / (rows * 4);
  const int n = static_cast<int>(gridDim.x);
  const int b = blockIdx.x;
  const int L = xcd_contiguous ? (b % 8) * (n / 8) + b / 8 : b;
  const int bx = L % ntx, by = (L / ntx) % nty, j3 = L / (ntx * nty);
  const int r0 = by * rows * 4 + wave * rows;
  const long long plane = static_cast<long long>(M1) * M2 * M3;
  long long o = bx * 64 + lane + static_cast<long long>(M1) * M2 * j3 +
                static_cast<long long>(M1) * r0;
  for (int r = 0; r < rows; ++r, o += M1)
    for (int c = 0; c < NC; ++c) out[o + plane * c] = 1.f;
}

}  // namespace

int main() {
  float* d = nullptr;
  if (hipMalloc(&d, N * sizeof(float)) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1;
  int failures = 0;
  auto timeit = [&](auto launch, const char* what) {
    for (int i = 0; i < 3; ++i) launch();
    (void)hipEventRecord(a);
    const int reps = 20;
    for (int i = 0; i < reps; ++i) launch();
    (void)hipEventRecord(b);
    if (hipEventSynchronize(b) != hipSuccess || hipGetLastError() != hipSuccess) {
      ++failures;
      return;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    const double us = ms * 1e3 / reps;
    std::printf("{\"store\": \"%s\", \"bytes\": %lld, \"us\": %.2f, \"TB_per_s\": %.3f}\n", what,
                N * 4, us, N * 4 / (us * 1e-6) / 1e12);
  };
  const unsigned b1 = static_cast<unsigned>(N / 256), b2 = b1 / 2, b4 = b1 / 4;
  timeit([&] { hipLaunchKernelGGL(fill<1>, dim3(b1), dim3(256), 0, 0, d); }, "fill 4 B per lane");
  timeit([&] { hipLaunchKernelGGL(fill<2>, dim3(b2), dim3(256), 0, 0, d); }, "fill 8 B per lane");
  timeit([&] { hipLaunchKernelGGL(fill<4>, dim3(b4), dim3(256), 0, 0, d); }, "fill 16 B per lane");
  for (int xc : {0, 1})
    for (int rows : {8, 16, 32}) {
      const unsigned grid = (M1 / 64) * (M2 / (rows * 4)) * M3;
      char name[96];
      std::snprintf(name, sizeof name, "forward pattern, %d rows per wave%s", rows,
                    xc ? ", XCD-contiguous tiles" : "");
      timeit([&] { hipLaunchKernelGGL(pattern, dim3(grid), dim3(256), 0, 0, d, rows, xc); }, name);
    }
  (void)hipFree(d);
  return failures ? 1 : 0;
}
