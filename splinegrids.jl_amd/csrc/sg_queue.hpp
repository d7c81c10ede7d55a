// sg_queue.hpp -- device work queues of the batched launches.
//
// A batch's tiles sit in one queue per XCD: queue q holds tiles[start[q] ..
// start[q + 1]) and is drained through the atomic counter
// counters[q * kQueueStride] (one 256-byte line each); start[] is stored
// after the counters.  Workgroup b of a persistent launch runs on XCD b mod 8
// (workgroups are dispatched to the XCDs round-robin), takes tiles from that
// XCD's queue first and then from the other queues in turn.
#pragma once

#include "sg_internal.hpp"

namespace sg {

// A tile of a batched launch: grid (or argument record) index, tile along
// dimension 1, chunk along dimension 2, and the dimension-3 position.
struct BTile {
  int grid, bx, by, bz;
};

constexpr int kQueueStride = 64;
constexpr int kQueueInts = kXcds * kQueueStride + kXcds + 1;

namespace {

[[maybe_unused]] __global__ void queue_reset(int* counters) {
  if (blockIdx.x == 0 && threadIdx.x < kXcds) counters[threadIdx.x * kQueueStride] = 0;
}

// Thread 0 of a workgroup: the next tile, own queue first, or -1 once every
// queue is empty.  `visited` (kept by the caller) counts the queues already
// found empty; they are not asked again.
__device__ inline int pop_tile(int* counters, int home, int& visited) {
  const int* start = counters + kXcds * kQueueStride;
  for (; visited < kXcds; ++visited) {
    const int q = (home + visited) % kXcds;
    const int i = start[q] + atomicAdd(counters + q * kQueueStride, 1);
    if (i < start[q + 1]) return i;
  }
  return -1;
}

}  // namespace

}  // namespace sg
