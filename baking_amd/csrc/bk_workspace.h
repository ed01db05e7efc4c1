// Device layout of the substep workspace and the tile helpers shared by the
// substep kernels (bk_substep.hip) and the adjoint's tiled scatter
// (bk_adjoint.hip).  Internal: not part of the C ABI.
#pragma once
#include "bk_device.h"
#include "../../include/baking_abi.h"

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin (workgroup b
// to XCD b % 8), and each XCD has its own L2.  xcd_block() renumbers them so
// that every XCD runs one contiguous range of logical workgroups: neighbouring
// particles and mesh vertices -- which gather the same grid nodes and
// particles -- then share one L2 instead of being spread over all eight.  A
// bijection on [0, gridDim.x) for any grid size.  Used by the gather-bound
// kernels (pass B1, mesh reconstruction): measured on the batched workload it
// speeds them up, while it slows the tiled scatters (their halo flushes then
// collide on one XCD) and the deferred return maps (spatially clustered work
// lands on a few XCDs), which keep the round-robin order.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
  const int b = int(blockIdx.x), G = int(gridDim.x);
  const int x = b % kXcds, i = b / kXcds, per = G / kXcds, rem = G % kXcds;
  return x * per + (x < rem ? x : rem) + i;
}

constexpr int kTileThreads = 256;  // P2G workgroup size
constexpr int kRegion = 6;         // 4 cells + 2-node stencil reach
constexpr int kRegionNodes = kRegion * kRegion * kRegion;

struct bk_workspace {
  int max_n = 0, total_blocks = 0, max_doughs = 0, n_vertices = 0, n_slots = 0;
  bk_dough* doughs = nullptr;                   // device copy of the dough table
  uint32_t* keys_raw = nullptr;                 // cell keys in particle order
  uint32_t* keys_s[2] = {nullptr, nullptr};     // radix scratch pairs
  int32_t* vals_s[2] = {nullptr, nullptr};
  uint32_t* sorted_keys = nullptr;              // sorted keys / particle indices
  int32_t* sorted_vals = nullptr;
  int32_t* hist = nullptr;
  int32_t* digit_total = nullptr;
  int32_t* blk_start = nullptr;                 // per batch block: first sorted particle
  int32_t* blk_count = nullptr;                 // per batch block: particles keyed to it (0 between substeps)
  double2* dTW = nullptr;                       // per node {dT, dW} of the thermal grid, interleaved for pass B1's 27-node
                                                // gather (written by the diffusion next to the bk_grid arrays; allocated
                                                // with the first thermal dough table)
  int32_t* scan_part = nullptr;                 // partial sums of the block-count scan
  int32_t* blk_end = nullptr;
  int32_t* node_flag = nullptr;                 // per batch block: active this substep
  int32_t* blocks = nullptr;                    // non-empty particle blocks
  int4* tiles = nullptr;                        // P2G tiles: (block, dough, first sorted particle, count)
  int32_t* active = nullptr;                    // active node blocks
  int32_t* counters = nullptr;                  // [0] full tiles, [1] active blocks, [2] deferred, [3] particle blocks,
                                                // [4] emitted samples, [5] deferred queue not yet processed,
                                                // [6] substeps run, [7] sparse tiles
  int32_t* queue = nullptr;                     // particles whose coupled porosity needs the full solve
  float4* prep = nullptr;                       // per particle, pass C -> deferred pass: the first half of a deferred
                                                // particle's Ding update (DingPrep, kDingPrepVec float4)
  float* tau = nullptr;
  float* vertices = nullptr;                    // reconstructed positions of the boundary faces' vertices
  int32_t* vert_mark = nullptr;                 // per vertex: substep count when a boundary face last used it
  double* slot_exposed = nullptr;
  double* slot_contact = nullptr;
  int32_t* flags = nullptr;                     // BK_FLAG_COUNT per dough
  int32_t* dflags = nullptr;                    // BK_FLAG_COUNT per dough: a deferred pass running beside the next
                                                // substep raises its flags here; k_join merges them
  float* features = nullptr;                    // surrogate inputs of the queued particles, kMlpRow per queue slot
  float* predictions = nullptr;                 // surrogate outputs, kMlpOut per queue slot
  const uint8_t* mlp = nullptr;                 // surrogate inference blob (caller-owned), or null
  float* samples = nullptr;                     // emitted (features, exact outputs) rows of kSampleRow floats, or null
  int32_t sample_cap = 0;
  unsigned long long* stats = nullptr;          // [0] porosity evaluations [1] guided solves accepted [2] guided solves
                                                // rejected [3] queued return maps; summed over substeps
  bk_settling* settle = nullptr;                // settling controller record per dough
  float* settle_ref = nullptr;                  // caller's [2][max_n][9] settling reference plastic gradients, or null
  float* settle_part = nullptr;                 // settling metric maxima per dough and run of 256 particles
  int64_t settle_part_cap = 0;                  // floats in settle_part
  int32_t* freeze = nullptr;                    // per dough, this substep: 0 runs, 1 first substep after a flag (writes
                                                // nothing: the output buffer keeps the last accepted state), 2 later
                                                // substeps (copies its state through); set by k_freeze at substep start
  int32_t* freeze_age = nullptr;                // per dough: substeps run since it stopped (0 while it runs)
  int32_t freezing = 0;                         // 1: doughs with a flag raised stop (batches); 0: they run on
};

// This substep's freeze mode of dough d (0 when freezing is off): see bk_workspace::freeze.
__device__ __forceinline__ int frozen_mode(const bk_workspace* ws, int d) { return ws->freezing ? ws->freeze[d] : 0; }

// Whether dough d rests this substep (physics/thermal_water.py:296-355 with
// settling=True): params.settling 1 always, 2 while its controller record is in
// phase 1.
__device__ __forceinline__ bool resting(const bk_params& P, const bk_workspace* ws, int d) {
  return P.settling == 1 || (P.settling == 2 && ws->settle[d].phase != 0);
}

// Stable LSD radix sort of (key, value) pairs on the low key_bits bits (values
// null: 0..n-1), with its scratch (keys_s, vals_s, hist, digit_total) taken from
// `ws`, a host-side view whose scratch fields point to device memory of at least
// n keys and values each, 256 * ceil(n / kRadixTile) histogram entries and 256
// digit totals (bk_substep.hip).
constexpr int kRadixTile = 2048;
int bk_internal_radix_sort(const bk_workspace* ws, const uint32_t* keys_in, const int32_t* vals_in, uint32_t* keys_out,
                           int32_t* vals_out, int n, int key_bits, hipStream_t st);

// Enqueue the settling controller after a substep with input plastic
// gradients Fp_in and output (v, C, Fp) (bk_settling.hip).
int bk_internal_settling_launch(bk_workspace* ws, const bk_dough* doughs, int n_doughs, int max_dough_n, const float* v,
                                const float* C, const float* Fp_in, const float* Fp, hipStream_t st);
// Position of batch block b inside its dough's grid.
struct BlockGeom {
  int d, nb, bx, by, bz;
  int64_t node_base;  // first node of the dough's grid in the batch grid
};
__device__ __forceinline__ BlockGeom block_geom(const bk_dough* __restrict__ D, int d, int b) {
  BlockGeom g;
  g.d = d;
  g.nb = (D[d].params.res + 3) / 4;
  int lb = b - D[d].block_offset;
  g.bz = lb % g.nb;
  g.by = (lb / g.nb) % g.nb;
  g.bx = lb / (g.nb * g.nb);
  g.node_base = int64_t(D[d].block_offset) << 6;
  return g;
}

// One P2G tile: up to `chunk` sorted particles of one particle block.  Every
// field is wave-uniform.  The tile list holds the tiles of at least
// kFullTile particles from its front (counters[0] of them) and the sparser
// ones -- blocks on a dough's surface, a split block's remainder -- from its
// back (counters[7]), so tile index t walks the full tiles first: a tiled pass
// over more tiles than the GPU holds at once ends on its shortest work.
constexpr int kFullTile = 256;
__device__ __forceinline__ int tile_count(const bk_workspace* ws) { return ws->counters[0] + ws->counters[7]; }
__device__ __forceinline__ int tile_slot(const bk_workspace* ws, int t) {
  const int front = ws->counters[0];
  return t < front ? t : ws->max_n + ws->total_blocks - 1 - (t - front);
}
struct Tile {
  int d, start, count;
  BlockGeom g;
};
__device__ __forceinline__ Tile tile_at(const bk_dough* __restrict__ doughs, const bk_workspace* ws, int t) {
  const int4 e = ws->tiles[tile_slot(ws, t)];
  Tile T;
  const int blk = __builtin_amdgcn_readfirstlane(e.x);
  T.d = __builtin_amdgcn_readfirstlane(e.y);
  T.start = __builtin_amdgcn_readfirstlane(e.z);
  T.count = __builtin_amdgcn_readfirstlane(e.w);
  T.g = block_geom(doughs, T.d, blk);
  return T;
}

__device__ __forceinline__ int region_index(int li, int lj, int lk) { return (li * kRegion + lj) * kRegion + lk; }

// A tile's particles are sorted by cell (the low 6 bits of the key).  Lanes of
// the tiled scatters take one cell each (lane = cell index), and the waves of
// the workgroup take its particles by rank within the cell.  Then the lanes of
// one LDS atomic address 64 distinct stencil nodes: no address conflicts, and
// with the (36, 6, 1) region strides the f64 atomics spread evenly over the
// banks.  Fills lo[c], hi[c] with the tile-relative range of cell c (empty:
// lo = hi = 0).  Call with the whole workgroup after zeroing lo and hi and a
// barrier; ends with a barrier.
__device__ __forceinline__ void tile_cell_ranges(const bk_workspace* ws, const Tile& tg, int* lo, int* hi) {
  const uint32_t* k = ws->sorted_keys + tg.start;
  for (int j = threadIdx.x; j < tg.count; j += blockDim.x) {
    const uint32_t key = k[j];
    if (j == 0 || k[j - 1] != key) lo[key & 63] = j;
    if (j == tg.count - 1 || k[j + 1] != key) hi[key & 63] = j + 1;
  }
  __syncthreads();
}

// A tile's work list for K particles per lane: entry (c, q) is the rank group q (ranks K q .. K q + K - 1) of cell
// c, and the entries run group-major (every cell's group 0, then every cell's group 1, ...), so 64 consecutive
// entries are 64 distinct cells except where they cross from one group to the next.  Lanes taking consecutive
// entries stay busy where a tile has empty or short cells, which lane-per-cell leaves idle.  Entry e holds
// (tile-relative position of the group's first particle) << 6 | c.  Wavefront 0 fills groups q0, q0 + 1, ...
// while they fit kWorkCap entries and leaves the entry count in *n and the next group in *next (-1: none left).
// Call with lo / hi from tile_cell_ranges, the whole workgroup, then a barrier.
constexpr int kWorkCap = 1024;
constexpr int kMaxTileChunk = 1 << 20;  // tile-relative positions stay below 2^20, so (position << 6 | cell) fits 32 bits
template <int K>
__device__ __forceinline__ void tile_work_list(const int* lo, const int* hi, int q0, int* list, int* n, int* next) {
  if (threadIdx.x >= 64) return;
  const int c = int(threadIdx.x);
  const int first = lo[c], len = hi[c] - first;
  const unsigned long long below = c == 0 ? 0ull : (~0ull >> (64 - c));
  int e = 0, q = q0;
  for (;;) {
    const bool has = len > K * q;
    const unsigned long long m = __ballot(has);
    const int cnt = __popcll(m);
    if (cnt == 0) { q = -1; break; }
    if (e + cnt > kWorkCap) break;
    if (has) list[e + __popcll(m & below)] = ((first + K * q) << 6) | c;
    e += cnt;
    ++q;
  }
  if (c == 0) {
    *n = e;
    *next = q;
  }
}

