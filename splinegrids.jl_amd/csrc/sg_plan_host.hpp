// sg_plan_host.hpp -- the host side of a plan: tilings, band windows, LDS
// budgets, the adjoint's gather lists, the transposed band lists and the
// batch's work queues.  Everything here is index arithmetic on host vectors
// (the spans of every dimension); nothing calls the HIP runtime, so the
// same code is built on its own under the address and undefined-behaviour
// sanitizers and driven on the CPU by tests/native/plan_check.cpp.
#pragma once

#include <functional>
#include <vector>

#include "sg_queue.hpp"

namespace sg {

// ---- environment knobs (positive integers; anything else: fallback) ----
int env_int(const char* name, int fallback);

// ---- tilings of the streaming kernels (sg_api.cpp builds the plan) ----

// Band windows of a tiling: tiles of kLanes * wx samples along dimension 1
// (or t.tile_w), chunks of kWaves / wx * rows_per_wave rows (or t.chunk_h)
// along dimension 2.
void compute_tiling(const Plan& pl, Tiling& t, int rows_per_wave, int wx = 1);
// Samples of dimension 3 per workgroup (0: about 1024 workgroups) and the
// widest k3 window of those groups.
void compute_ztiling(const Plan& pl, Tiling& t, int j3_per_block);
// Largest LDS claim of a forward launch (FP64, widest channel block).
size_t forward_lds_worst(const Plan& pl, const Tiling& t);
// LDS of an adjoint streaming launch of nc channels of elem bytes.
size_t adjoint_stream_lds(const Plan& pl, int nc, size_t elem);
// LDS of an MFMA forward launch: nch window channels, nf fields.
size_t mma_lds(const Plan& pl, int nch, int nf);
// Workgroup numbering of a plan's launches (see xcd_logical_block).
int default_xcd_map(const Plan& pl);

// Every tiling of a plan whose dimensions carry their host spans and
// monotone flags: stream_ok, fwd, mma, and with want_adjoint adj and (when
// adjm_shape_ok) adjm, whose task count uses resident_waves(plan) -- the
// waves of the matrix-core adjoint the device holds at once for the
// plan's adjm tiling as chosen so far.
void plan_tilings(Plan& pl, bool want_adjoint, bool adjm_shape_ok,
                  const std::function<int(const Plan&)>& resident_waves);

// For dimension d and control point k, the element offsets (in the partial
// windows of tiling t, w1 x w2 x t.w3max per channel) of the windows that
// hold k: ptr[d][k] .. ptr[d][k + 1] index off[d].  gchan: offset between
// channels; gfast: window counts small enough for udiv_fast in the gather.
struct GatherLists {
  std::vector<int32_t> ptr[3];
  std::vector<long long> off[3];
  long long gchan = 0;
  int gfast = 0;
};
GatherLists gather_lists(const Plan& pl, const Tiling& t, int w1, int w2);

// Transposed band lists of a dimension: the (sample, term) pairs whose band
// holds control point k, j * (p + 1) + a, in sample order, at
// js[ptr[k] .. ptr[k + 1]).
void band_lists(const DimDev& D, std::vector<int32_t>& ptr, std::vector<int32_t>& js);

// ---- the batch's persistent launch (sg_batch.hip) ----

constexpr int kBC = 4;  // value channels per pass; the weight is slab 4
// 3-D tiles' plane list: entries of kEntry floats, weights then k3 at
// kEntryK; entry 0 holds the list's length at kEntryN
constexpr int kEntry = 8;
constexpr int kEntryK = 4;
constexpr int kEntryN = 5;
__host__ __device__ constexpr int plane_list_floats(int p3) {
  return (kWaves * (p3 + 1) + 1) * kEntry;
}

// The plane list of the 3-D batch tile whose kWaves samples j3 start at j3_0
// (`planes` <= kWaves of them inside the grid): one entry per control-point
// plane k3 in the union of their dimension-3 bands, in increasing k3,
// {w[0..kWaves), k3}, where w[pz] is B3 of sample j3_0 + pz at k3 (0 outside
// its band), then one padding entry with zero weights and the last k3, so
// that entries can be taken in pairs.  span3 / tab3: the dimension's spans
// and FP32 order-0 table [m3][p3 + 1].  Writes plane_list_floats(p3) floats
// (the rest zero) and returns the length.
int plane_list_host(const int32_t* span3, const float* tab3, int p3, int j3_0, int planes,
                    float* list);

// LDS one workgroup may take when blocks_per_cu of them share a CU.
size_t lds_cap(int blocks_per_cu);

// Samples j1 per lane a grid's batch tiles may take: 2 for grids of degrees
// p1, p2 <= kVec2MaxDegree whose rows are longer than one wave (tiles of
// 2 kLanes samples, rows stored in 8-byte pairs: half the store
// instructions, and a row of up to 128 samples in one tile), else 1.  Such
// tiles take channel blocks of at most kVec2Slabs window slabs.
constexpr int kVec2MaxDegree = 4;
constexpr int kVec2Slabs = 4;
int batch_vec(const Plan& pl);

// Window sizes of a grid's batch tiles (kLanes * vec samples j1) for a
// rows-per-wave choice.
struct GridTiling {
  int rows, ntx, nty, w1, w2;
  size_t lds;  // bytes of LDS one tile of this grid needs
  int tw;      // samples j1 per tile
};
GridTiling grid_tiling(const Plan& pl, int rows, int vec = 1);

// The tiling of a grid in a batch of `rows_per_wave` under an LDS cap: rows
// halve from rows_per_wave until the window fits; two samples per lane
// where batch_vec allows them and their wider window fits at some number of
// rows (on C5 that beats one sample per lane even with fewer rows), else
// one.  lds > cap when nothing fits (the grid then needs its own launch).
GridTiling batch_grid_tiling(const Plan& pl, int rows_per_wave, size_t cap);

// The queues of a tile list: grids go whole to the least loaded queue,
// largest first; a queue holds its grids in that order.  Writes the tiles in
// queue order and the queue starts (host_q: zero counters, then starts).
void pack_queues(const std::vector<std::vector<BTile>>& grid_tiles,
                 const std::vector<long long>& grid_cost, std::vector<BTile>& sorted,
                 int* host_q);

}  // namespace sg
