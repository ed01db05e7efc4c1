// Fused MLS-MPM baking substep for gfx950 (MI355X): workspace, cell keys,
// stable radix sort, and the substep of a batch of independent doughs.
//
// A batch is a table of doughs (a single simulator is a batch of one).  Their
// particles, tiled grids, meshes and boundary faces are concatenated: dough d
// owns particles [particle_offset, +n) and node blocks [block_offset, +nb^3)
// of the batch grid (4x4x4 node blocks of 64 contiguous nodes); mesh, face and
// slot indices in the scene are global.  Every stage is one launch for the
// whole batch.  Kernels take the table as a read-only argument and look up
// their dough per tile, node block or particle with a wave-uniform index, so a
// dough's parameters are read with scalar loads.
//
// One substep (reference physics/mpm.py:163-225) runs as:
//   clear    zero the node blocks the previous substep touched
//   keys     batch cell key of every particle: (batch block << 6) | cell in block,
//            and its rank among its block's particles (per-block counters)
//   sort     scan of the block counts (each non-empty block's range and the
//            block list), scatter by block, counting sort by cell per block
//   tiles    activate the 2x2x2 node blocks each particle block reaches; cut
//            blocks into P2G tiles of at most `tile_chunk` particles
//   mesh     reconstruct the boundary faces' vertices, then their areas per owner slot
//   pass A   thermal P2G (7 f64 scalars) accumulated per tile in LDS, flushed
//            with one global f64 atomic per touched node and scalar
//   diffuse  6-neighbour FV heat/water fluxes per active node block
//   pass B1  scalar G2P + gas source + gas and water stress (per particle)
//   pass B2  mechanical P2G (mass + momentum) per tile in f64 LDS, then flush
//   grid     momentum -> velocity, gravity, unilateral walls
//   pass C   G2P + Ding return map + gas refresh + saturations (per particle);
//            particles whose coupled porosity needs the full solve are queued
//   surrogate (optional) the return-map network on the queued particles' features
//   pass C'  the queued return maps, densely; with the surrogate, each coupled
//            porosity solve starts from the network's prediction
//   settling (doughs still resting before baking) the rest metrics and the
//            rest/bake decision on the device (bk_settling.hip)
#include "bk_device.h"
#include "bk_surrogate.h"
#include "bk_workspace.h"
#include "../../include/baking_abi.h"

#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

using namespace bk;

namespace {

thread_local std::string g_last_error;

int fail(const char* what) {
  g_last_error = what;
  return -1;
}
int check_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return 0;
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
  return -2;
}
#define BK_HIP(call)                                        \
  do {                                                      \
    int _rc = check_hip((call), #call);                     \
    if (_rc) return _rc;                                    \
  } while (0)
#define BK_LAUNCH(name) BK_HIP(hipGetLastError())

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;  // keys per workgroup
static_assert(kSortTile == kRadixTile, "the radix sort's scratch sizing in bk_workspace.h");
constexpr int kMaxDoughRes = 1024;                    // 3 * res^3 stays below 2^32 (stencil_terms)
constexpr int kParticleThreads = 256;                 // launch bound of the per-particle kernels

inline int div_up(long long a, long long b) { return int((a + b - 1) / b); }
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

}  // namespace

// The library's one error text (bk_last_error), also set by the other sources.
void bk_internal_set_error(const char* what) { g_last_error = what; }

constexpr int kSampleRow = 12;  // 7 features, 4 exact outputs, 1 unused

// Host side of a workspace: a copy of the device descriptor, the uploaded
// dough table with its totals, and the launch tuning.
struct HostWs {
  bk_workspace dev;
  std::vector<bk_dough> doughs;
  int n = 0, n_vertices = 0, n_slots = 0, key_bits = 6, max_dough_n = 0;
  bool thermal = false, gas = false, surrogate = false, settle = false;
  bk_tuning tuning{4096, 256, 4096, 256};
};
static std::mutex g_ws_mutex;
static std::unordered_map<const bk_workspace*, HostWs> g_ws_cache;

// ============================================================================ dough lookup
// Run body(d) for the wavefront's active lanes.  Doughs are contiguous in
// particle order, so nearly every wavefront holds one dough: then body gets it
// wave-uniform (an SGPR) and reads the dough's table entry with scalar loads.
// A wavefront straddling two doughs runs body with the per-lane dough (vector
// loads of the entry).  The test goes through a ballot: a branch on (d == u)
// would let the compiler substitute the per-lane d for u inside body.
template <class Body> __device__ __forceinline__ void per_dough(int d, Body&& body) {
  const int u = __builtin_amdgcn_readfirstlane(d);
  if (__ballot(d != u) == 0ull) body(u);
  else body(d);
}

__device__ __forceinline__ int dough_of_particle(const int32_t* dough_of, int p) { return dough_of ? dough_of[p] : 0; }

// Dough owning batch block b: binary search over the table's block offsets.
__device__ __forceinline__ int dough_of_block(const bk_dough* __restrict__ D, int nd, int b) {
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (D[mid].block_offset <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int32_t* dough_flags(const bk_workspace* ws, int d) { return ws->flags + d * BK_FLAG_COUNT; }

// ============================================================================ keys + radix sort
__device__ __forceinline__ uint32_t tiled_cell_key(const int b[3], int nb) {
  uint32_t blk = (uint32_t((b[0] >> 2) * nb + (b[1] >> 2)) * uint32_t(nb)) + uint32_t(b[2] >> 2);
  return (blk << 6) | uint32_t(((b[0] & 3) << 4) | ((b[1] & 3) << 2) | (b[2] & 3));
}

// Stand-alone cell keys (bk_cell_keys): row-major flatten_node, or tiled over one grid.
__global__ void k_cell_keys(const float* __restrict__ x, int n, float ox, float oy, float oz, float dx, int res, int tiled,
                            uint32_t* __restrict__ keys, int32_t* __restrict__ flags) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float o[3] = {ox, oy, oz};
  float g[3];
  grid_position(x + 3 * p, o, dx, g);
  int b[3];
  stencil_base(g, b);
  bool ok = valid_stencil(b, res);
  flag_max(flags, 1, !ok);
  uint32_t key = 0;
  if (ok) key = tiled ? tiled_cell_key(b, (res + 3) / 4) : uint32_t(flatten_node(b[0], b[1], b[2], res));
  keys[p] = key;
}

// Batch keys: the dough's first block plus its tiled key, so one sort orders
// every dough.  Thread 0 also resets the per-substep counters.  Off-grid
// particles take their dough's first cell (the P2G tiles skip them) and raise
// the stencil flag.
// Batch cell key of every particle and its rank among the particles of its
// batch block (the order in which their appends reached the block's counter).
// The lanes of a wavefront that share a block -- consecutive particles mostly
// do -- take their ranks from one atomic.
// Keys and per-block ranks of particles (one per lane, p).  Particle 0 resets
// the substep's list counters and advances its stamp; the active-block count
// (counters[1]) is reset by k_block_scan_partials, after the clear that reads it.
constexpr uint32_t kFrozenKey = 0xFFFFFFFFu;  // the key of a particle of a stopped dough: not sorted, no tile

__device__ __forceinline__ void batch_keys(const bk_dough* __restrict__ doughs, const int32_t* __restrict__ dough_of,
                                           const float* __restrict__ x, int n, uint32_t* __restrict__ keys,
                                           int32_t* __restrict__ ranks, bk_workspace* ws, int p, int reset_queue) {
  if (p == 0) {
    ws->counters[0] = 0;
    if (reset_queue) ws->counters[2] = 0;  // else k_join resets it, after the previous substep's deferred pass read it
    ws->counters[3] = 0;
    ws->counters[6] += 1;
    ws->counters[7] = 0;
  }
  const bool valid = p < n;
  bool live = valid;  // a stopped dough's particles get no key, so no tile, node block or grid work
  uint32_t key = 0u;
  if (valid) {
    per_dough(dough_of_particle(dough_of, p), [&](int d) {
      if (frozen_mode(ws, d)) {
        live = false;
        return;
      }
      const bk_params& P = doughs[d].params;
      const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
      float g[3];
      grid_position(x + 3 * p, o, P.dx, g);
      int b[3];
      stencil_base(g, b);
      bool ok = valid_stencil(b, P.res);
      flag_max(dough_flags(ws, d), BK_FLAG_STENCIL, !ok);
      key = (uint32_t(doughs[d].block_offset) << 6) + (ok ? tiled_cell_key(b, (P.res + 3) / 4) : 0u);
    });
  }
  // group the lanes by block (wave-uniform loop, no memory traffic), then every
  // group's first lane adds the group's size: all of the wavefront's atomics in
  // one instruction, so their returns are awaited once
  const unsigned lane = __lane_id();
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  unsigned long long pending = __ballot(live);
  int leader = 0, offset = 0, size = 0;
  while (pending) {
    const int first = __ffsll((long long)pending) - 1;
    const uint32_t blk = __shfl(key >> 6, first, 64);
    const unsigned long long same = __ballot(live && (key >> 6) == blk) & pending;
    if ((same >> lane) & 1ull) {
      leader = first;
      offset = __popcll(same & below);
    }
    if (int(lane) == first) size = __popcll(same);
    pending &= ~same;
  }
  int base = 0;
  if (size > 0) base = atomicAdd(&ws->blk_count[key >> 6], size);
  base = __shfl(base, leader, 64);
  if (valid) {
    keys[p] = live ? key : kFrozenKey;
    ranks[p] = base + offset;
  }
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                             int32_t* __restrict__ hist, int nwg) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int base = blockIdx.x * kSortTile;
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    int e = base + i * kSortThreads + threadIdx.x;
    if (e < n) atomicAdd(&h[(keys[e] >> shift) & 255], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nwg + blockIdx.x] = h[threadIdx.x];
}

// Per-digit exclusive scan over workgroups: workgroup d scans hist[d*nwg ..
// (d+1)*nwg) in 256-wide coalesced chunks and records the digit's total.
__global__ __launch_bounds__(256) void k_scan_digits(int32_t* __restrict__ hist, int32_t* __restrict__ digit_total, int nwg) {
  __shared__ int buf[256];
  const int d = blockIdx.x;
  int carry = 0;
  for (int base = 0; base < nwg; base += 256) {
    int i = base + threadIdx.x;
    int v = i < nwg ? hist[d * nwg + i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nwg) hist[d * nwg + i] = carry + buf[threadIdx.x] - v;
    carry += buf[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) digit_total[d] = carry;
}

// Stable scatter: each wave owns a contiguous 512-key segment of the
// workgroup's 2048 keys; ranks inside a wave come from 8 ballots (lanes with the
// same digit), across waves from per-wave digit counts.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout, int32_t* __restrict__ vout, int n,
                                                                int shift, const int32_t* __restrict__ offsets,
                                                                const int32_t* __restrict__ digit_total, int nwg) {
  __shared__ int run[4][256];
  __shared__ int dbase[256];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int base = blockIdx.x * kSortTile + w * (kSortTile / 4);
#pragma unroll
  for (int q = 0; q < 4; ++q) run[q][threadIdx.x] = 0;
  // exclusive scan of the 256 digit totals (every workgroup recomputes it)
  int tot = digit_total[threadIdx.x];
  dbase[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int t = threadIdx.x >= off ? dbase[threadIdx.x - off] : 0;
    __syncthreads();
    dbase[threadIdx.x] += t;
    __syncthreads();
  }
  dbase[threadIdx.x] -= tot;
  __syncthreads();
  for (int r = 0; r < kSortItems; ++r) {
    int e = base + r * 64 + lane;
    if (e < n) atomicAdd(&run[w][(kin[e] >> shift) & 255], 1);
  }
  __syncthreads();
  {
    int d = threadIdx.x;
    int c0 = run[0][d], c1 = run[1][d], c2 = run[2][d];
    int o = dbase[d] + offsets[d * nwg + blockIdx.x];
    run[0][d] = o;
    run[1][d] = o + c0;
    run[2][d] = o + c0 + c1;
    run[3][d] = o + c0 + c1 + c2;
  }
  __syncthreads();
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortItems; ++r) {
    int e = base + r * 64 + lane;
    bool valid = e < n;
    uint32_t key = valid ? kin[e] : 0u;
    int d = int((key >> shift) & 255);
    unsigned long long mask = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1;
      unsigned long long bb = __ballot(valid && bit);
      mask &= bit ? bb : ~bb;
    }
    int pos = 0;
    if (valid) pos = run[w][d] + __popcll(mask & lt);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    if (valid) {
      kout[pos] = key;
      vout[pos] = vin ? vin[e] : e;
      if ((mask & lt) == 0ull) run[w][d] += __popcll(mask);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// Stable LSD radix sort.  Pass p reads the previous pass's output and writes a
// scratch pair (A for even p, B for odd p); the last pass writes the caller's
// output, so neither the input nor the output is ever read and written in one pass.
static int radix_sort(const bk_workspace* ws, const uint32_t* keys_in, const int32_t* vals_in, uint32_t* keys_out,
                      int32_t* vals_out, int n, int key_bits, hipStream_t st) {
  if (n <= 0) return 0;
  int passes = (key_bits + 7) / 8;
  if (passes < 1) passes = 1;
  int nwg = div_up(n, kSortTile);
  const uint32_t* ksrc = keys_in;
  const int32_t* vsrc = vals_in;
  for (int p = 0; p < passes; ++p) {
    bool last = p == passes - 1;
    uint32_t* kdst = last ? keys_out : ws->keys_s[p & 1];
    int32_t* vdst = last ? vals_out : ws->vals_s[p & 1];
    hipLaunchKernelGGL(k_radix_hist, dim3(nwg), dim3(kSortThreads), 0, st, ksrc, n, 8 * p, ws->hist, nwg);
    hipLaunchKernelGGL(k_scan_digits, dim3(256), dim3(256), 0, st, ws->hist, ws->digit_total, nwg);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nwg), dim3(kSortThreads), 0, st, ksrc, vsrc, kdst, vdst, n, 8 * p, ws->hist,
                       ws->digit_total, nwg);
    BK_LAUNCH(radix);
    ksrc = kdst;
    vsrc = vdst;
  }
  return 0;
}

// radix_sort for the other sources of the library (declared in bk_workspace.h).
int bk_internal_radix_sort(const bk_workspace* ws, const uint32_t* keys_in, const int32_t* vals_in, uint32_t* keys_out,
                           int32_t* vals_out, int n, int key_bits, hipStream_t st) {
  return radix_sort(ws, keys_in, vals_in, keys_out, vals_out, n, key_bits, st);
}

// ============================================================================ block sort, tiles and active node blocks
// The particles are ordered by batch block and, inside each block, by cell:
//   k_block_scan_*   exclusive scan of the per-block counts the keys left:
//                    each non-empty block's [start, end) in the sorted order,
//                    the list of non-empty blocks, and the counts reset to 0
//   block_scatter    particle p to start[block] + rank (in k_activate's launch)
//   cell sort        inside each non-empty block, a counting sort by cell (k_sort_reconstruct)
// The result is the permutation of a stable sort of the keys (a cell's
// particles in index order), whatever order the atomics of the keys and the
// scatter left them in; the order of the P2G accumulations the tiles feed
// still follows their atomics.  Every pass reads and writes each particle
// once, against the four 8-bit passes of a radix sort of the whole key.
constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;  // block counts per scan workgroup

__device__ __forceinline__ int block_sum(int v, int* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = int(threadIdx.x) >> 6;
  if (__lane_id() == 0) lds[w] = v;
  __syncthreads();
  int t = 0;
  for (int k = 0; k < int(blockDim.x) >> 6; ++k) t += lds[k];
  __syncthreads();
  return t;
}

// Exclusive scan of v over the workgroup (256 threads); `lds` holds 4 ints.
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds) {
  const unsigned lane = __lane_id();
  int inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(inc, off, 64);
    if (int(lane) >= off) inc += t;
  }
  const int w = int(threadIdx.x) >> 6;
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  int before = 0;
  for (int k = 0; k < w; ++k) before += lds[k];
  __syncthreads();
  return before + inc - v;
}

__global__ __launch_bounds__(256) void k_block_scan_partials(const int32_t* __restrict__ counts, int nb, int32_t* __restrict__ part,
                                                             int32_t* __restrict__ counters) {
  __shared__ int lds[4];
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[1] = 0;  // the active-block count, read by the clear before this kernel
  const int base = blockIdx.x * kScanTile + int(threadIdx.x) * kScanItems;
  int sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < nb) sum += counts[base + i];
  sum = block_sum(sum, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

__global__ __launch_bounds__(256) void k_block_scan_apply(int32_t* __restrict__ counts, int nb, const int32_t* __restrict__ part,
                                                          bk_workspace* ws) {
  __shared__ int lds[17];
  const int base = blockIdx.x * kScanTile + int(threadIdx.x) * kScanItems;
  int c[kScanItems];
  int sum = 0, nonempty = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    c[i] = base + i < nb ? counts[base + i] : 0;
    sum += c[i];
    nonempty += c[i] > 0;
  }
  // the workgroups before this one: a few hundred partials, summed here instead of by a separate scan kernel
  int before = 0;
  for (int k = int(threadIdx.x); k < int(blockIdx.x); k += 256) before += part[k];
  before = block_sum(before, lds);
  int offset = before + block_exclusive_scan(sum, lds);
  int slot = block_reserve(&ws->counters[3], nonempty, lds);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (c[i] > 0) {
      ws->blk_start[base + i] = offset;
      ws->blk_end[base + i] = offset + c[i];
      ws->blocks[slot++] = base + i;
      counts[base + i] = 0;
      offset += c[i];
    }
  }
}

// Particle p to its block's start + its rank (run in k_activate's last workgroups).
__device__ __forceinline__ void block_scatter(const uint32_t* __restrict__ keys, const int32_t* __restrict__ ranks, int n,
                                              const bk_workspace* __restrict__ ws, uint32_t* __restrict__ kout,
                                              int32_t* __restrict__ vout, int p) {
  if (p >= n) return;
  const uint32_t key = keys[p];
  if (key == kFrozenKey) return;
  const int pos = ws->blk_start[key >> 6] + ranks[p];
  kout[pos] = key;
  vout[pos] = p;
}

// One wavefront per non-empty block: cell counts in LDS, their exclusive scan
// (lane c holds cell c), then every particle takes the next slot of its cell.
// Wavefront-scope barrier: the lanes' LDS accesses before it are visible to
// every lane after it (the wavefront's own LDS region, no other wavefront).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inside each non-empty block, a stable sort by cell: wavefront w of nw takes
// blocks w, w + nw, ...  Cell counts go to its 64-int LDS cursor and are
// scanned (lane c holds cell c); the block's particles are bucketed by cell in
// LDS as words cell << kIndexBits | particle, and each then takes the slot of
// its cell's first particle plus its rank by particle index within the cell.
// The result is the order of a stable sort of the keys -- (block, cell), then
// particle index -- whatever order the block scatter's atomics left the block
// in.  A block of more than kCellSortCap particles ranks each particle against
// its block's input segment in global memory instead (same order, slower).
constexpr int kCellSortCap = 1024;
constexpr uint32_t kIndexBits = 26;  // particle indices below 2^26 (checked at workspace creation)

__device__ __forceinline__ void cell_sort_wave(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                               const bk_workspace* __restrict__ ws, uint32_t* __restrict__ kout,
                                               int32_t* __restrict__ vout, int* cursor, int* cstart, uint32_t* buf, int w, int nw) {
  const int nblk = ws->counters[3];
  const int lane = int(__lane_id());
  for (int bi = w; bi < nblk; bi += nw) {
    const int b = ws->blocks[bi];
    const int start = ws->blk_start[b], end = ws->blk_end[b], size = end - start;
    cursor[lane] = 0;
    wave_sync();
    for (int j = start + lane; j < end; j += 64) atomicAdd(&cursor[kin[j] & 63u], 1);
    wave_sync();
    const int count = cursor[lane];
    int inc = count;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(inc, off, 64);
      if (lane >= off) inc += t;
    }
    wave_sync();
    cstart[lane] = inc - count;
    cursor[lane] = inc - count;
    wave_sync();
    const uint32_t block_key = uint32_t(b) << 6;
    if (size <= kCellSortCap) {
      for (int j = start + lane; j < end; j += 64) {
        const uint32_t c = kin[j] & 63u;
        buf[atomicAdd(&cursor[c], 1)] = (c << kIndexBits) | uint32_t(vin[j]);
      }
      wave_sync();  // cursor[c] now ends cell c's bucket
      for (int j = lane; j < size; j += 64) {
        const uint32_t word = buf[j];
        const uint32_t c = word >> kIndexBits;
        const int lo = cstart[c], hi = cursor[c];
        int rank = 0;
        for (int k = lo; k < hi; ++k) rank += buf[k] < word;
        kout[start + lo + rank] = block_key | c;
        vout[start + lo + rank] = int32_t(word & ((1u << kIndexBits) - 1u));
      }
    } else {
      for (int j = start + lane; j < end; j += 64) {
        const uint32_t c = kin[j] & 63u;
        const int32_t v = vin[j];
        int rank = 0;
        for (int k = start; k < end; ++k) rank += (kin[k] & 63u) == c && vin[k] < v;
        kout[start + cstart[c] + rank] = block_key | c;
        vout[start + cstart[c] + rank] = v;
      }
    }
    wave_sync();
  }
}

// Eight threads per non-empty particle block: each activates one of the 2x2x2
// node blocks the block's stencils reach; the first also cuts the block into
// tiles of at most `chunk` particles.  The loop's trip count is uniform per
// workgroup, so every thread reaches the per-workgroup appends.
__device__ __forceinline__ void mark_face_vertices(const bk_scene& sc, bk_workspace* ws, int s);
// Three independent jobs that all follow the block-count scan, in one launch:
// workgroups [0, L) build the tile and active-block lists, the next
// mark_blocks mark the boundary faces' vertices for the mesh stage, and the
// last scatter_blocks place each particle at its block's start + rank
// (L = gridDim.x - mark_blocks - scatter_blocks; L = 0 leaves the lists to
// another launch).  The split is per workgroup, so the list workgroups'
// collectives see whole workgroups.
__global__ __launch_bounds__(256) void k_activate(const bk_dough* __restrict__ doughs, int nd, bk_workspace* ws, int chunk,
                                                  bk_scene sc, int nslots, int mark_blocks, const uint32_t* __restrict__ keys,
                                                  const int32_t* __restrict__ ranks, int n, uint32_t* __restrict__ kout,
                                                  int32_t* __restrict__ vout, int scatter_blocks) {
  __shared__ int lds[17];
  const int list_blocks = int(gridDim.x) - mark_blocks - scatter_blocks;
  if (int(blockIdx.x) >= list_blocks + mark_blocks) {
    block_scatter(keys, ranks, n, ws, kout, vout, (int(blockIdx.x) - list_blocks - mark_blocks) * int(blockDim.x) + int(threadIdx.x));
    return;
  }
  if (int(blockIdx.x) >= list_blocks) {
    const int s = (int(blockIdx.x) - list_blocks) * int(blockDim.x) + int(threadIdx.x);
    if (s < nslots) mark_face_vertices(sc, ws, s);
    return;
  }
  const int total = ws->counters[3] * 8;
  for (int base = blockIdx.x * blockDim.x; base < total; base += list_blocks * blockDim.x) {
    const int t = base + int(threadIdx.x);
    int b = 0, d = 0, start = 0, end = 0, ntile = 0, nbk = 0;
    bool fresh = false;
    if (t < total) {
      b = ws->blocks[t >> 3];
      const int corner = t & 7;
      d = dough_of_block(doughs, nd, b);
      const BlockGeom g = block_geom(doughs, d, b);
      if (corner == 0) {
        start = ws->blk_start[b];
        end = ws->blk_end[b];
        ntile = (end - start + chunk - 1) / chunk;
      }
      const int cx = g.bx + (corner >> 2), cy = g.by + ((corner >> 1) & 1), cz = g.bz + (corner & 1);
      if (cx < g.nb && cy < g.nb && cz < g.nb) {
        nbk = doughs[d].block_offset + (cx * g.nb + cy) * g.nb + cz;
        fresh = atomicExch(&ws->node_flag[nbk], 1) == 0;
      }
    }
    int nfull = 0;  // tiles of at least kFullTile particles: every one but the last, and the last when it is
    for (int k = 0; k < ntile; ++k) nfull += min(chunk, end - start - k * chunk) >= kFullTile;
    int front = block_reserve(&ws->counters[0], nfull, lds);
    int back = block_reserve(&ws->counters[7], ntile - nfull, lds);
    const int last = ws->max_n + ws->total_blocks - 1;
    for (int k = 0; k < ntile; ++k) {
      const int s = start + k * chunk, c = min(chunk, end - s);
      ws->tiles[c >= kFullTile ? front++ : last - back++] = make_int4(b, d, s, c);
    }
    block_append(&ws->counters[1], ws->active, fresh, nbk, lds);
  }
}

// Zero every grid array over the previous substep's active blocks (64 nodes
// each), as workgroup b of nb.
__device__ __forceinline__ void clear_active_blocks(bk_workspace* ws, const bk_grid& g, int thermal, int b, int nb) {
  int nact = ws->counters[1];
  for (int t = b * int(blockDim.x) + int(threadIdx.x); t < nact * 64; t += nb * int(blockDim.x)) {
    int blk = ws->active[t >> 6];
    int64_t node = (int64_t(blk) << 6) | (t & 63);
    g.m[node] = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) { g.p[3 * node + a] = 0.0f; g.v[3 * node + a] = 0.0f; }
    if (thermal) {
      g.V[node] = 0.0; g.C[node] = 0.0; g.E[node] = 0.0; g.W[node] = 0.0;
      g.H[node] = 0.0; g.L[node] = 0.0; g.X[node] = 0.0; g.dT[node] = 0.0; g.dW[node] = 0.0;
    }
    if ((t & 63) == 0) ws->node_flag[blk] = 0;
  }
}

// Stages 1 and 2 in one launch: workgroups [0, clear_wgs) clear the previous
// substep's blocks, the rest form the keys (which need no cleared grid, and do
// not touch what the clear reads or writes).
__global__ void k_clear_keys(const bk_dough* __restrict__ doughs, const int32_t* __restrict__ dough_of, const float* __restrict__ x,
                             int n, uint32_t* __restrict__ keys, int32_t* __restrict__ ranks, bk_workspace* ws, bk_grid g,
                             int thermal, int clear_wgs, int reset_queue) {
  if (int(blockIdx.x) < clear_wgs)
    clear_active_blocks(ws, g, thermal, int(blockIdx.x), clear_wgs);
  else
    batch_keys(doughs, dough_of, x, n, keys, ranks, ws, (int(blockIdx.x) - clear_wgs) * int(blockDim.x) + int(threadIdx.x),
               reset_queue);
}

// ============================================================================ mesh + surface areas
// Only the boundary faces' areas enter the substep (the Robin terms), so only
// their vertices are reconstructed: each slot marks its faces' vertices with
// the substep count (in the last workgroups of k_activate), and the
// reconstruction skips the vertices left unmarked.
__device__ __forceinline__ void mark_face_vertices(const bk_scene& sc, bk_workspace* ws, int s) {
  const int32_t stamp = ws->counters[6];
  for (int k = sc.slot_face_offsets[s]; k < sc.slot_face_offsets[s + 1]; ++k) {
    const int32_t* f = sc.faces + 3 * sc.slot_faces[k];
    ws->vert_mark[f[0]] = stamp;
    ws->vert_mark[f[1]] = stamp;
    ws->vert_mark[f[2]] = stamp;
  }
}

// Two independent jobs in one launch: workgroups [0, sort_wgs) run the cell
// sort (four wavefronts each, one block per wavefront at a time), the next nv
// reconstruct one marked vertex each (geometry/mesh_tracking.py:13-38).
// sort_wgs is a multiple of 8, so vertex workgroups keep the XCD-contiguous
// order of xcd_block() within their range.
__global__ __launch_bounds__(256) void k_sort_reconstruct(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                                          bk_workspace* ws, uint32_t* __restrict__ kout, int32_t* __restrict__ vout,
                                                          int sort_wgs, const float* __restrict__ x, const float* __restrict__ F,
                                                          const float* __restrict__ ref_vertices, const int32_t* __restrict__ offs,
                                                          const int32_t* __restrict__ parts, const float* __restrict__ wts,
                                                          const float* __restrict__ ref_offs, float* __restrict__ out, int nv) {
  __shared__ int cursor[4][64];
  __shared__ int cstart[4][64];
  __shared__ uint32_t buf[4][kCellSortCap];
  __shared__ float red[16];
  const int bx = int(blockIdx.x);
  if (bx < sort_wgs) {
    const int wv = int(threadIdx.x >> 6);
    cell_sort_wave(kin, vin, ws, kout, vout, cursor[wv], cstart[wv], buf[wv], bx * 4 + wv, sort_wgs * 4);
    return;
  }
  const int r = bx - sort_wgs;  // XCD r % kXcds, as bx is
  const int xc = r % kXcds, i = r / kXcds, per = nv / kXcds, rem = nv % kXcds;
  const int v = xc * per + (xc < rem ? xc : rem) + i;
  if (v >= nv || ws->vert_mark[v] != ws->counters[6]) return;
  reconstruct_vertex_block(v, x, F, ref_vertices, offs, parts, wts, ref_offs, out, red);
}

// physics/thermal_water.py:77-92, summed per owner over its faces in face order.
__device__ __forceinline__ void face_area(const float* verts, const int32_t* f, double floor_cd, double& area, bool& contact) {
  double P[3][3];
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) P[k][a] = double(verts[3 * f[k] + a]);
  double e1[3], e2[3];
  for (int a = 0; a < 3; ++a) { e1[a] = P[1][a] - P[0][a]; e2[a] = P[2][a] - P[0][a]; }
  double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
  double len = sqrt(nx * nx + ny * ny + nz * nz);
  area = 0.5 * len;
  double cy = (P[0][1] + P[1][1] + P[2][1]) / 3.0;
  contact = cy <= floor_cd && ny < -0.2 * len;
}

__global__ void k_slot_areas(const bk_dough* __restrict__ doughs, bk_scene sc, const float* __restrict__ verts, bk_workspace* ws,
                             int nslots) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  per_dough(dough_of_particle(sc.dough_of, sc.slot_owner[s]), [&](int d) {
    const bk_params& P = doughs[d].params;
    const float floor_h = P.origin[1] + float(P.boundary_cells - 1) * P.dx;
    double fcd = double(floor_h + P.contact_distance);
    double e = 0.0, c = 0.0;
    for (int k = sc.slot_face_offsets[s]; k < sc.slot_face_offsets[s + 1]; ++k) {
      double a;
      bool isc;
      face_area(verts, sc.faces + 3 * sc.slot_faces[k], fcd, a, isc);
      if (isc) c += a; else e += a;
    }
    ws->slot_exposed[s] = e;
    ws->slot_contact[s] = c;
  });
}

// ============================================================================ per-particle transport terms
struct Surface {
  double exposed, contact;
};
__device__ __forceinline__ Surface particle_surface(const bk_scene& sc, const bk_workspace* ws, int p) {
  int slot = sc.particle_slot[p];
  Surface s{0.0, 0.0};
  if (slot >= 0) { s.exposed = ws->slot_exposed[slot]; s.contact = ws->slot_contact[slot]; }
  return s;
}

// Robin heat, water loss and exchange of one particle (thermal_water.py:108-120).
struct Robin {
  double vol, sat, loss, heat, exch;
};
// A resting substep has no oven, tray or surface-loss terms (thermal_water.py:306, 325-326).
__device__ __forceinline__ Robin robin_terms(const bk_params& P, bool rest, float heat_param, float Jf, float V0, double T,
                                             double W, const Surface& s) {
  const float h = rest ? 0.0f : heat_param;
  const float tray = rest ? 0.0f : P.tray_h;
  const float loss = rest ? 0.0f : P.loss_coef;
  Robin r;
  r.vol = double(V0) * double(Jf);
  r.sat = W / (double(P.rho_w) * r.vol);
  r.loss = 0.0;
  if (P.water_enabled) r.loss = double(loss) * T * r.sat * s.exposed;
  r.heat = double(h) * s.exposed * (double(P.T_oven) - T) + double(tray) * s.contact * (double(P.T_tray) - T) -
           double(P.latent) * r.loss;
  r.exch = double(h) * s.exposed + double(tray) * s.contact;
  if (P.water_enabled) r.exch += double(P.latent * loss) * r.sat * s.exposed;
  return r;
}

// In both P2G passes lane c of every wave takes the particles of cell c of
// the tile's block (tile_cell_ranges), so one LDS atomic instruction touches
// 64 distinct nodes.  A cell's particles share one stencil base (the cell is
// the key's, formed from that base), so a lane can take K of them at a time
// and add their summed contribution with one LDS atomic per node and field.
// Wave w takes ranks K*w .. K*w + K-1 of each cell, then every 4K-th.  Pass A
// (four f64 fields per node, bound by its LDS atomics) takes pairs;
// pass B2 (four fields, more arithmetic per particle) is fastest one at a
// time.  (A loop over groups of slots with different bases would let the
// compiler keep all 27 K weight products live across it: 200+ VGPRs.)
constexpr int kPackA = 2;
constexpr int kPackB2 = 1;

// Stencil base and axis weights of one pack slot, and whether it holds a
// particle of this tile (`in_range`: the slot's rank is inside the cell).  A
// slot that does not gets zero weights, so no NaN of a far-off particle can
// reach the sums.
__device__ __forceinline__ bool pack_member(const float g3[3], bool in_range, int res, const Tile& tg, int b[3], float wx[3],
                                            float wy[3], float wz[3]) {
  stencil_base(g3, b);
  const bool ok = in_range && valid_stencil(b, res) && (b[0] >> 2) == tg.g.bx && (b[1] >> 2) == tg.g.by && (b[2] >> 2) == tg.g.bz;
  axis_weights(g3[0] - float(b[0]), wx);
  axis_weights(g3[1] - float(b[1]), wy);
  axis_weights(g3[2] - float(b[2]), wz);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    wx[k] = ok ? wx[k] : 0.0f;
    wy[k] = ok ? wy[k] : 0.0f;
    wz[k] = ok ? wz[k] : 0.0f;
  }
  return ok;
}

// The pending pack slots that share the stencil base of the lowest pending one,
// and that base relative to the tile's block.  The keys are formed from the
// same positions and the same base computation, so every pending slot of a
// cell shares it; one that does not is left out and flagged.
template <int K>
__device__ __forceinline__ unsigned pack_group(const int (&base)[K][3], const Tile& tg, unsigned pending, int& li, int& lj, int& lk) {
  int rb[3] = {base[K - 1][0], base[K - 1][1], base[K - 1][2]};
#pragma unroll
  for (int u = K - 2; u >= 0; --u)
    if (pending >> u & 1u) {
      rb[0] = base[u][0];
      rb[1] = base[u][1];
      rb[2] = base[u][2];
    }
  unsigned sel = 0u;
#pragma unroll
  for (int u = 0; u < K; ++u)
    if ((pending >> u & 1u) && base[u][0] == rb[0] && base[u][1] == rb[1] && base[u][2] == rb[2]) sel |= 1u << u;
  li = rb[0] - 4 * tg.g.bx;
  lj = rb[1] - 4 * tg.g.by;
  lk = rb[2] - 4 * tg.g.bz;
  return sel;
}

// ============================================================================ pass A: thermal P2G
// The Robin heat, water loss and exchange of one face owner (thermal_water.py:108-120) added to the tile's H, L and
// X sums (thermal_water.py:121-137).  Only a face owner has exposed or contact area: for every other particle the
// three terms are 0 * finite = 0, so a tile's main loop adds only the volume, capacity, energy and water sums and
// lists its face owners, whose terms are added afterwards, one owner per lane (owners whose stencils share a node
// serialise on its LDS address).  A particle off the grid or off the tile's block (a key out of date, flagged by
// the main loop) adds nothing, as in the main loop.
constexpr int kSurfCap = 128;  // face owners listed per tile; a tile with more sweeps its particles instead
__device__ __forceinline__ void owner_to_tile(const bk_params& P, bool rest, float heat_param, const bk_scene& sc,
                                              const bk_state& in, const bk_workspace* ws, const Tile& tg, int p, int slot,
                                              double (*rob)[kRegionNodes]) {
  const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
  float g3[3];
  grid_position(in.x + 3 * p, o, P.dx, g3);
  int b[3];
  stencil_base(g3, b);
  if (!valid_stencil(b, P.res) || (b[0] >> 2) != tg.g.bx || (b[1] >> 2) != tg.g.by || (b[2] >> 2) != tg.g.bz) return;
  const float Jf = m_det(m_load(in.F + 9 * p));
  const Surface sf{ws->slot_exposed[slot], ws->slot_contact[slot]};
  const Robin r = robin_terms(P, rest, heat_param, Jf, sc.V0[p], in.T[p], in.W[p], sf);
  const double f[3] = {r.heat, r.loss, r.exch};
  float wx[3], wy[3], wz[3];
  axis_weights(g3[0] - float(b[0]), wx);
  axis_weights(g3[1] - float(b[1]), wy);
  axis_weights(g3[2] - float(b[2]), wz);
  const int li = b[0] - 4 * tg.g.bx, lj = b[1] - 4 * tg.g.by, lk = b[2] - 4 * tg.g.bz;
#pragma unroll
  for (int ox = 0; ox < 3; ++ox)
#pragma unroll
    for (int oy = 0; oy < 3; ++oy)
#pragma unroll
      for (int oz = 0; oz < 3; ++oz) {
        const double wt = double(wx[ox] * wy[oy] * wz[oz]);
        const int ri = region_index(li + ox, lj + oy, lk + oz);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const double v = wt * f[a];
          if (v != 0.0) atomicAdd(&rob[a][ri], v);
        }
      }
}

__global__ __launch_bounds__(kTileThreads) void k_pass_a(const bk_dough* __restrict__ doughs, bk_scene sc, bk_state in, bk_grid g,
                                                         bk_workspace* ws) {
  __shared__ double acc[7][kRegionNodes];  // V, C, E, W from the main loop; H, L, X from the face owners
  __shared__ int cell_lo[64], cell_hi[64];
  __shared__ int surf_p[kSurfCap], surf_s[kSurfCap];  // listed face owners: particle, slot
  __shared__ int surf_n;
  __shared__ int work[kWorkCap];
  __shared__ int work_n, work_next;
  const int ntiles = tile_count(ws);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tg = tile_at(doughs, ws, t);
    const bk_params& P = doughs[tg.d].params;
    if (!P.transport_enabled) continue;  // a dough without transport in a batch with some (uniform per tile)
    const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
    for (int i = threadIdx.x; i < 7 * kRegionNodes; i += kTileThreads) (&acc[0][0])[i] = 0.0;
    if (threadIdx.x < 64) {
      cell_lo[threadIdx.x] = 0;
      cell_hi[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) surf_n = 0;
    __syncthreads();
    tile_cell_ranges(ws, tg, cell_lo, cell_hi);
    for (int q0 = 0; q0 >= 0;) {
    tile_work_list<kPackA>(cell_lo, cell_hi, q0, work, &work_n, &work_next);
    __syncthreads();
    const int nwork = work_n;
    q0 = work_next;
    for (int e0 = 64 * w; e0 < nwork; e0 += 4 * 64) {
      const bool live = e0 + lane < nwork;  // a lane past the list reads the last entry and takes no particle
      const int ent = work[min(e0 + lane, nwork - 1)];
      const int hi = cell_hi[ent & 63];
      const int j0 = ent >> 6;
      float wx[kPackA][3], wy[kPackA][3], wz[kPackA][3];
      double f[kPackA][4];
      int base[kPackA][3], pid[kPackA], slot[kPackA];
      unsigned pending = 0u;
#pragma unroll
      for (int u = 0; u < kPackA; ++u) {
        const int j = j0 + u;
        const int p = ws->sorted_vals[tg.start + min(j, hi - 1)];
        pid[u] = p;
        slot[u] = sc.particle_slot[p];
        float g3[3];
        grid_position(in.x + 3 * p, o, P.dx, g3);
        const bool ok = pack_member(g3, live && j < hi, P.res, tg, base[u], wx[u], wy[u], wz[u]);
        const float Jf = m_det(m_load(in.F + 9 * p));
        const double T = in.T[p], W = in.W[p];
        const double cap = sc.capacities[p];
        const double v[4] = {double(sc.V0[p]) * double(Jf), cap, cap * T, W};
#pragma unroll
        for (int a = 0; a < 4; ++a) f[u][a] = ok ? v[a] : 0.0;
        if (ok) pending |= 1u << u;
      }
      int li, lj, lk;
      const unsigned sel = pack_group(base, tg, pending, li, lj, lk);
      flag_max(dough_flags(ws, tg.d), BK_FLAG_STENCIL, sel != pending);  // a slot off its cell's base: a key out of date
      if (sel) {
        double fs[kPackA][4];
#pragma unroll
        for (int u = 0; u < kPackA; ++u)
#pragma unroll
          for (int a = 0; a < 4; ++a) fs[u][a] = (sel >> u & 1u) ? f[u][a] : 0.0;
#pragma unroll
        for (int ox = 0; ox < 3; ++ox)
#pragma unroll
          for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
              double wt[kPackA];
#pragma unroll
              for (int u = 0; u < kPackA; ++u) wt[u] = double(wx[u][ox] * wy[u][oy] * wz[u][oz]);
              const int ri = region_index(li + ox, lj + oy, lk + oz);
              auto sum = [&](int a) {
                double t = wt[0] * fs[0][a];
#pragma unroll
                for (int u = 1; u < kPackA; ++u) t += wt[u] * fs[u][a];
                return t;
              };
#pragma unroll
              for (int a = 0; a < 4; ++a) atomicAdd(&acc[a][ri], sum(a));
            }
      }
#pragma unroll
      for (int u = 0; u < kPackA; ++u)
        if ((sel >> u & 1u) && slot[u] >= 0) {
          const int i = atomicAdd(&surf_n, 1);
          if (i < kSurfCap) {
            surf_p[i] = pid[u];
            surf_s[i] = slot[u];
          }
        }
    }
    __syncthreads();  // the list is read before the next window overwrites it, and every add is done after the last
    }
    const int ns = surf_n;  // uniform
    if (ns > 0) {
      const bool rest = resting(P, ws, tg.d);
      const float heat_param = sc.heat_param[tg.d];
      if (ns <= kSurfCap) {
        for (int i = threadIdx.x; i < ns; i += kTileThreads) owner_to_tile(P, rest, heat_param, sc, in, ws, tg, surf_p[i], surf_s[i], acc + 4);
      } else {  // more owners than the list holds: every thread takes the tile's owners among its particles
        for (int j = threadIdx.x; j < tg.count; j += kTileThreads) {
          const int p = ws->sorted_vals[tg.start + j];
          const int sl = sc.particle_slot[p];
          if (sl >= 0) owner_to_tile(P, rest, heat_param, sc, in, ws, tg, p, sl, acc + 4);
        }
      }
      __syncthreads();
    }
    const int nf = ns > 0 ? 7 : 4;
    for (int ri = threadIdx.x; ri < kRegionNodes; ri += kTileThreads) {
      int li = ri / 36, lj = (ri / 6) % 6, lk = ri % 6;
      int i = 4 * tg.g.bx + li, j = 4 * tg.g.by + lj, k = 4 * tg.g.bz + lk;
      if (i >= P.res || j >= P.res || k >= P.res) continue;
      int64_t node = tg.g.node_base + tiled_node(i, j, k, tg.g.nb);
      double* dst[7] = {g.V, g.C, g.E, g.W, g.H, g.L, g.X};
#pragma unroll
      for (int a = 0; a < 7; ++a) {
        double v = acc[a][ri];
        if (a < nf && v != 0.0) unsafeAtomicAdd(dst[a] + node, v);
      }
    }
    __syncthreads();
  }
}

// ============================================================================ diffusion (thermal_water.py:140-190, GridFields fused)
// 64 consecutive threads own one active node block, so the block (and its
// dough) is wave-uniform.  The wavefront stages {C, V, E, W} of its block's 64
// nodes and of the 96 nodes just outside its six faces in LDS (coalesced
// loads of the block, one load per halo node), then every lane takes its six
// neighbours from LDS.  Face f = 2 * axis + side holds the 16 nodes beyond
// that side, indexed by the two other axes' positions in the block (lower
// axis first); a node outside the grid is staged as zeros and never read.
constexpr int kDiffHalo = 96;
__device__ __forceinline__ int halo_slot(int ax, int side, const int lc[3]) {
  const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
  return (2 * ax + side) * 16 + lc[a1] * 4 + lc[a2];
}

__global__ __launch_bounds__(256) void k_diffuse(const bk_dough* __restrict__ doughs, int nd, bk_grid g, bk_workspace* ws) {
  // one array per field: a wavefront's 64 consecutive doubles of a field fall in distinct bank pairs
  __shared__ double stage[4][4][64 + kDiffHalo];
  const int wv = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
  double* const sC = stage[wv][0];
  double* const sV = stage[wv][1];
  double* const sE = stage[wv][2];
  double* const sW = stage[wv][3];
  const int nact = ws->counters[1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nact * 64; t += gridDim.x * blockDim.x) {
    const int blk = __builtin_amdgcn_readfirstlane(ws->active[t >> 6]);
    const BlockGeom bg = block_geom(doughs, dough_of_block(doughs, nd, blk), blk);
    const bk_params& P = doughs[bg.d].params;
    const int loc = lane;
    const int64_t node = (int64_t(blk) << 6) | loc;
    const int lc[3] = {loc >> 4, (loc >> 2) & 3, loc & 3};
    const int c[3] = {4 * bg.bx + lc[0], 4 * bg.by + lc[1], 4 * bg.bz + lc[2]};
    const double Ci = g.C[node];
    sC[loc] = Ci;
    sV[loc] = g.V[node];
    sE[loc] = g.E[node];
    sW[loc] = g.W[node];
    for (int h = lane; h < kDiffHalo; h += 64) {
      const int f = h >> 4, ax = f >> 1, side = f & 1;
      const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
      int q[3];
      q[ax] = side ? 4 : -1;
      q[a1] = (h >> 2) & 3;
      q[a2] = h & 3;
      const int o[3] = {4 * bg.bx + q[0], 4 * bg.by + q[1], 4 * bg.bz + q[2]};
      double c4 = 0.0, v4 = 0.0, e4 = 0.0, w4 = 0.0;
      if (o[ax] >= 0 && o[ax] < P.res) {
        const int64_t other = bg.node_base + tiled_node(o[0], o[1], o[2], bg.nb);
        c4 = g.C[other];
        v4 = g.V[other];
        e4 = g.E[other];
        w4 = g.W[other];
      }
      sC[64 + h] = c4;
      sV[64 + h] = v4;
      sE[64 + h] = e4;
      sW[64 + h] = w4;
    }
    wave_sync();
    bool heat_bad = false, water_bad = false;
    if (Ci > 0.0) {
      const double rho_w = double(P.rho_w);
      const double dt = double(P.dt);
      const double dx3 = double(P.dx * P.dx * P.dx);
      const bool rest = resting(P, ws, bg.d);
      const double hk = rest ? 0.0 : double(P.kappa * P.dx);
      const double wk = rest ? 0.0 : double(P.diffusivity * P.rho_w * P.dx);
      double Vi = sV[loc];
      double Ti = sE[loc] / Ci;
      double Si = sW[loc] / (rho_w * Vi);
      double hr = g.H[node], wr = -g.L[node], hd = g.X[node], wd = 0.0;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const int step = 2 * side - 1;
          if (c[ax] + step < 0 || c[ax] + step >= P.res) continue;
          int ln[3] = {lc[0], lc[1], lc[2]};
          ln[ax] += step;
          const int j = (ln[ax] >= 0 && ln[ax] < 4) ? ((ln[0] << 4) | (ln[1] << 2) | ln[2]) : 64 + halo_slot(ax, side, lc);
          double Cj = sC[j];
          if (Cj > 0.0) {
            double Vj = sV[j];
            double Tj = sE[j] / Cj;
            double fill = 2.0 * Vi * Vj / (Vi + Vj) / dx3;
            double hc = hk * fill;
            hr += hc * (Tj - Ti);
            hd += hc;
            if (P.water_enabled) {
              double Sj = sW[j] / (rho_w * Vj);
              double mob = wk * fill * 0.5 * (Ti + Tj) * 0.5 * (Si + Sj);
              wr += mob * (Sj - Si);
              wd += mob / (rho_w * Vi);
            }
          }
        }
      heat_bad = dt * hd > Ci;
      double change = dt * wr;
      const double Wi = sW[loc];
      water_bad = Wi + change < 0.0 || dt * wd > 1.0;
      const double dT = dt * hr / Ci;
      // the relative water change dW / W each particle gathers (thermal_water.py:222-223), once per node
      const double dW = Wi > 0.0 ? change / Wi : 0.0;
      g.dT[node] = dT;
      g.dW[node] = dW;
      ws->dTW[node] = make_double2(dT, dW);
    } else {
      ws->dTW[node] = make_double2(0.0, 0.0);  // as the cleared dT and dW of a node without heat capacity
    }
    flag_max(dough_flags(ws, bg.d), BK_FLAG_HEAT, heat_bad);
    flag_max(dough_flags(ws, bg.d), BK_FLAG_WATER, water_bad);
    wave_sync();  // every lane has read the stage before the next block overwrites it
  }
}

// ============================================================================ pass B1: scalar G2P, gas, stresses
// Five wavefronts per SIMD: the 27-node gather's terms fit in 88 VGPRs without spilling (128 unconstrained, four).
__global__ __launch_bounds__(kParticleThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_pass_b1(
    const bk_dough* __restrict__ doughs, bk_scene sc, bk_state in, bk_state out, bk_grid g, bk_workspace* ws, int n) {
  int p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n) return;
  per_dough(dough_of_particle(sc.dough_of, p), [&](int d) {
    if (frozen_mode(ws, d)) return;  // a stopped dough: pass C keeps or copies its state
    const bk_params& P = doughs[d].params;
    const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
    const int nb = (P.res + 3) / 4;
    const int64_t node_base = int64_t(doughs[d].block_offset) << 6;
    bool jbad = false, pbad = false, cflbad = false;
    mat3<float> F = m_load(in.F + 9 * p);
    float Jf = m_det(F);
    mat3<float> tau;
    float T_mat = 0.0f, dens = 0.0f;
    double moles = 0.0, sat = 0.0;
    if (P.transport_enabled) {
      float g3[3];
      grid_position(in.x + 3 * p, o, P.dx, g3);
      int b[3];
      stencil_base(g3, b);
      double tc = 0.0, mr = 0.0;
      if (valid_stencil(b, P.res)) {
        float l[3] = {g3[0] - float(b[0]), g3[1] - float(b[1]), g3[2] - float(b[2])};
        float wx[3], wy[3], wz[3];
        axis_weights(l[0], wx);
        axis_weights(l[1], wy);
        axis_weights(l[2], wz);
        const double2* __restrict__ tw = ws->dTW + node_base;
        const StencilTerms st = stencil_terms(b, nb);
#pragma unroll
        for (int ox = 0; ox < 3; ++ox)
#pragma unroll
          for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
              double wt = double(wx[ox] * wy[oy] * wz[oz]);
              const double2 q = tw[st.x[ox] + st.y[oy] + st.z[oz]];  // {dT, dW} in one 16-byte load
              tc += wt * q.x;
              mr += wt * q.y;
            }
      }
      double T_old = in.T[p], W_old = in.W[p];
      const bool rest = resting(P, ws, d);
      Robin r = robin_terms(P, rest, sc.heat_param[d], Jf, sc.V0[p], T_old, W_old, particle_surface(sc, ws, p));
      double T_new = T_old + tc;
      double W_new = W_old * (1.0 + mr);
      double dry = sc.dry[p];
      out.T[p] = T_new;
      T_mat = float(T_new);
      out.W[p] = W_new;
      out.losses[p] = in.losses[p] + double(P.dt) * r.loss;
      out.exchanges[p] = in.exchanges[p] + double(P.dt) * r.heat;
      out.masses[p] = float(dry + W_new);
      dens = float((dry + W_new) / double(sc.V0[p]));
      out.density[p] = dens;
      out.T_mat[p] = T_mat;
      sat = W_new / (double(P.rho_w) * double(sc.V0[p]) * double(Jf));
      if (P.model == 1 && P.gas_enabled) {
        moles = in.moles[p];
        if (!rest) {
          double tm = 0.5 * (T_old + T_new);
          double z = (tm - double(P.gas_peak)) / double(P.gas_width);
          moles += double(P.dt * P.density) * double(sc.prod_param[d]) * exp(-(z * z));
        }
        out.moles[p] = moles;
      }
    } else if (P.model == 1) {
      T_mat = in.T_mat[p];
      dens = in.density[p];
      moles = P.gas_enabled ? in.moles[p] : 0.0;
    }
    if (P.model == 0) {
      // physics/materials/elastic.py:13-26
      if (!isfinite(Jf) || Jf <= 0.0f) { jbad = true; tau = m_zero<float>(); }
      else tau = neo_hookean_kirchhoff(F, P.mu, P.lam, Jf);
    } else {
      tau = m_load(in.tau + 9 * p);
      if (P.gas_enabled) {
        // physics/materials/gas.py:16-35
        float nv = in.porosity[p];
        if (nv <= 0.0f || nv >= 1.0f || !isfinite(nv) || Jf <= 0.0f) pbad = true;
        else {
          float pr = float(gas_pressure(moles, T_mat, Jf, nv));
          float add = Jf * (P.p_ambient - pr);
          tau.a[0] += add; tau.a[4] += add; tau.a[8] += add;
          float bulk = pr * (1.0f + (1.0f - P.n0 + nv) / nv);
          float wave = sqrtf((P.lam + 2.0f * P.mu + bulk) * Jf / dens);
          cflbad = !isfinite(pr) || wave * P.dt / P.dx > P.cfl_limit;
        }
      }
    }
    if (P.transport_enabled) {
      // physics/thermal_water.py:248-253
      float k = Jf * (P.water_enabled ? P.water_stress : 0.0f) * (float(sat) - P.s0);
      tau.a[0] -= k; tau.a[4] -= k; tau.a[8] -= k;
    }
    m_store(ws->tau + 9 * p, tau);
    int32_t* fl = dough_flags(ws, d);
    flag_max(fl, BK_FLAG_J, jbad);
    flag_max(fl, BK_FLAG_POROSITY, pbad);
    flag_max(fl, BK_FLAG_GAS_CFL, cflbad);
  });
}

// ============================================================================ pass B2: mechanical P2G (physics/transfer.py:36-70)
__global__ __launch_bounds__(kTileThreads) void k_pass_b2(const bk_dough* __restrict__ doughs, bk_scene sc, bk_state in, bk_state out,
                                                          bk_grid g, bk_workspace* ws) {
  // f64 LDS accumulators: ds_add_f64 sustains far higher throughput than ds_add_f32 on gfx950
  // for this scatter, and each node's sum is rounded to float once, at the flush.
  __shared__ double acc[4][kRegionNodes];
  __shared__ int cell_lo[64], cell_hi[64];
  __shared__ int work[kWorkCap];
  __shared__ int work_n, work_next;
  const int ntiles = tile_count(ws);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tg = tile_at(doughs, ws, t);
    const bk_params& P = doughs[tg.d].params;
    const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
    const float* masses = P.transport_enabled ? out.masses : sc.m0;
    for (int i = threadIdx.x; i < 4 * kRegionNodes; i += kTileThreads) (&acc[0][0])[i] = 0.0;
    if (threadIdx.x < 64) {
      cell_lo[threadIdx.x] = 0;
      cell_hi[threadIdx.x] = 0;
    }
    __syncthreads();
    tile_cell_ranges(ws, tg, cell_lo, cell_hi);
    for (int q0 = 0; q0 >= 0;) {
    tile_work_list<kPackB2>(cell_lo, cell_hi, q0, work, &work_n, &work_next);
    __syncthreads();
    const int nwork = work_n;
    q0 = work_next;
    for (int e0 = 64 * w; e0 < nwork; e0 += 4 * 64) {
      const bool live = e0 + lane < nwork;  // a lane past the list reads the last entry and takes no particle
      const int ent = work[min(e0 + lane, nwork - 1)];
      const int hi = cell_hi[ent & 63];
      const int j0 = ent >> 6;
      float wx[kPackB2][3], wy[kPackB2][3], wz[kPackB2][3], l[kPackB2][3];
      float mass[kPackB2], mv[kPackB2][3];
      mat3<float> A[kPackB2];
      int base[kPackB2][3];
      unsigned pending = 0u;
#pragma unroll
      for (int u = 0; u < kPackB2; ++u) {
        const int j = j0 + u;
        const int p = ws->sorted_vals[tg.start + min(j, hi - 1)];
        float g3[3];
        grid_position(in.x + 3 * p, o, P.dx, g3);
        const bool ok = pack_member(g3, live && j < hi, P.res, tg, base[u], wx[u], wy[u], wz[u]);
#pragma unroll
        for (int a = 0; a < 3; ++a) l[u][a] = ok ? g3[a] - float(base[u][a]) : 0.0f;  // finite offsets for an empty slot
        const float m = masses[p];
        const float sc_tau = P.dt * sc.V0[p] * 4.0f / (P.dx * P.dx);
        const mat3<float> Cm = m_load(in.C + 9 * p);
        const mat3<float> tau = m_load(ws->tau + 9 * p);
#pragma unroll
        for (int i = 0; i < 9; ++i) A[u].a[i] = ok ? m * Cm.a[i] - sc_tau * tau.a[i] : 0.0f;
        mass[u] = ok ? m : 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) mv[u][a] = ok ? m * in.v[3 * p + a] : 0.0f;
        if (ok) pending |= 1u << u;
      }
      int li, lj, lk;
      const unsigned sel = pack_group(base, tg, pending, li, lj, lk);
      flag_max(dough_flags(ws, tg.d), BK_FLAG_STENCIL, sel != pending);  // a slot off its cell's base: a key out of date
      if (sel) {
#pragma unroll
        for (int ox = 0; ox < 3; ++ox)
#pragma unroll
          for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
              const int ri = region_index(li + ox, lj + oy, lk + oz);
              double sm = 0.0, sp[3] = {0.0, 0.0, 0.0};
#pragma unroll
              for (int u = 0; u < kPackB2; ++u) {
                const float wt = (sel >> u & 1u) ? wx[u][ox] * wy[u][oy] * wz[u][oz] : 0.0f;
                const float dd[3] = {(float(ox) - l[u][0]) * P.dx, (float(oy) - l[u][1]) * P.dx, (float(oz) - l[u][2]) * P.dx};
                sm += double(wt * mass[u]);
#pragma unroll
                for (int a = 0; a < 3; ++a)
                  sp[a] += double(wt * (mv[u][a] + (A[u].a[a * 3 + 0] * dd[0] + A[u].a[a * 3 + 1] * dd[1] + A[u].a[a * 3 + 2] * dd[2])));
              }
              atomicAdd(&acc[0][ri], sm);
#pragma unroll
              for (int a = 0; a < 3; ++a) atomicAdd(&acc[1 + a][ri], sp[a]);
            }
      }
    }
    __syncthreads();  // the list is read before the next window overwrites it, and every add is done before the flush
    }
    for (int ri = threadIdx.x; ri < kRegionNodes; ri += kTileThreads) {
      int li = ri / 36, lj = (ri / 6) % 6, lk = ri % 6;
      int i = 4 * tg.g.bx + li, j = 4 * tg.g.by + lj, k = 4 * tg.g.bz + lk;
      if (i >= P.res || j >= P.res || k >= P.res) continue;
      int64_t node = tg.g.node_base + tiled_node(i, j, k, tg.g.nb);
      float m = float(acc[0][ri]);
      if (m != 0.0f) unsafeAtomicAdd(g.m + node, m);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        float v = float(acc[1 + a][ri]);
        if (v != 0.0f) unsafeAtomicAdd(g.p + 3 * node + a, v);
      }
    }
    __syncthreads();
  }
}

// ============================================================================ grid update (physics/grid.py:20-38)
__global__ __launch_bounds__(256) void k_grid_update(const bk_dough* __restrict__ doughs, int nd, bk_grid g, bk_workspace* ws) {
  const int nact = ws->counters[1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nact * 64; t += gridDim.x * blockDim.x) {
    const int blk = __builtin_amdgcn_readfirstlane(ws->active[t >> 6]);
    const int loc = t & 63;
    const int64_t node = (int64_t(blk) << 6) | loc;
    float m = g.m[node];
    if (!(m > 0.0f)) continue;
    const BlockGeom bg = block_geom(doughs, dough_of_block(doughs, nd, blk), blk);
    const bk_params& P = doughs[bg.d].params;
    int c[3] = {4 * bg.bx + (loc >> 4), 4 * bg.by + ((loc >> 2) & 3), 4 * bg.bz + (loc & 3)};
    float vel[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      vel[a] = g.p[3 * node + a] / m + P.dt * P.gravity[a];
      if (c[a] < P.boundary_cells && vel[a] < 0.0f) vel[a] = 0.0f;
      if (c[a] >= P.res - P.boundary_cells && vel[a] > 0.0f) vel[a] = 0.0f;
    }
    if (P.sticky_floor && c[1] < P.boundary_cells) vel[0] = vel[1] = vel[2] = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) g.v[3 * node + a] = vel[a];
  }
}

// ============================================================================ material update after G2P
// Ding return map, gas refresh and saturation of particle p
// (ding.py:303-327, gas.py:38-45, thermal_water.py:256-260).
__device__ __forceinline__ DingParams ding_params(const bk_params& P) {
  DingParams D;
  D.mu = P.mu; D.lam = P.lam; D.yield_stress = P.yield_stress; D.n0 = P.n0; D.dt = P.dt;
  for (int i = 0; i < 4; ++i) { D.knots[i] = P.knots[i]; D.yf[i] = P.yield_factors[i]; D.vf[i] = P.viscosity_factors[i]; }
  D.temperature_enabled = P.temperature_enabled;
  D.coupled = P.gas_enabled;
  D.bracketed = P.return_solver;
  return D;
}

// The outputs of a finished Ding update, the gas pressure and the saturation.
__device__ inline void material_outputs(const bk_params& P, const bk_scene& sc, const bk_state& in, const bk_state& out, int p,
                                        const mat3<float>& F_new, int status, const DingOut& r) {
  if (status == kDingDone) {
    m_store(out.Fe + 9 * p, r.Fe);
    m_store(out.Fp + 9 * p, r.Fp);
    m_store(out.tau + 9 * p, r.tau);
    out.residual[p] = r.residual;
    out.porosity_residual[p] = r.porosity_residual;
    out.porosity[p] = r.porosity;
    out.T_mat[p] = r.T;
    out.gelled[p] = r.gelled;
    out.yield_stress[p] = r.yield_stress;
    out.viscosity[p] = r.viscosity;
  }
  if (!P.transport_enabled) out.density[p] = in.density[p];
  float Jn = m_det(F_new);
  if (P.gas_enabled) {
    double moles = P.transport_enabled ? out.moles[p] : in.moles[p];
    if (!P.transport_enabled) out.moles[p] = moles;
    // a finished update's porosity and temperature are the values just stored: no reload behind the stores
    const bool done = status == kDingDone;
    float nn = done ? r.porosity : out.porosity[p];
    if (nn > 0.0f && nn < 1.0f) out.pressure[p] = gas_pressure(moles, done ? r.T : out.T_mat[p], Jn, nn);
  }
  if (P.transport_enabled) out.s[p] = out.W[p] / (double(P.rho_w) * double(sc.V0[p]) * double(Jn));
}

// The per-particle inputs of pass C's Ding update.  Pass C loads them before
// its G2P gather, whose stores the compiler cannot move loads across, so that
// their latency overlaps the gather.
struct DingInputs {
  mat3<float> Fe;
  float porosity, T;
  int gelled;
  float sat, fp_dev;  // the network features of a deferral: saturation, |Fp - I| (with pore gas only)
};

__device__ __forceinline__ DingInputs ding_inputs(const bk_params& P, const bk_state& in, const bk_state& out, int p) {
  DingInputs q;
  q.Fe = m_load(in.Fe + 9 * p);
  q.porosity = in.porosity[p];
  q.gelled = in.gelled[p];
  q.T = P.transport_enabled ? out.T_mat[p] : in.T_mat[p];
  q.sat = 0.0f;
  q.fp_dev = 0.0f;
  if (P.gas_enabled) {  // only a coupled particle can be deferred to the network
    q.sat = P.transport_enabled ? float(in.s[p]) : 0.0f;
    const mat3<float> Fp = m_load(in.Fp + 9 * p);
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float e = Fp.a[i] - ((i % 4 == 0) ? 1.0f : 0.0f);
      acc += e * e;
    }
    q.fp_dev = sqrtf(acc);
  }
  return q;
}

// Pass C's update of particle p.  A particle deferred to the second pass
// (`defer`) leaves the first half's record at prep_out + kDingPrepVec * p.
__device__ inline int material_update(const bk_params& P, float visc_param, const bk_scene& sc, const bk_state& in,
                                      const bk_state& out, int p, const DingInputs& q, const mat3<float>& F_new,
                                      const mat3<float>& Cn, bool defer, bool& f0, bool& f1, DingAux* aux,
                                      float4* __restrict__ prep_out) {
  const DingParams D = ding_params(P);
  DingOut r;
  int status = ding_update(D, q.Fe, q.porosity, q.gelled, F_new, Cn, q.T, visc_param, r, defer, aux,
                           prep_out + kDingPrepVec * int64_t(p));
  f0 = r.flag0;
  f1 = r.flag1;
  if (status == kDingDeferred) return status;
  material_outputs(P, sc, in, out, p, F_new, status, r);
  return status;
}

// The deferred pass's update of particle p: the second half from pass C's record, whose stretches, yield
// stress and viscosity are read for the solve and its rotations only after it.
__device__ inline int material_finish(const bk_params& P, const bk_scene& sc, const bk_state& in, const bk_state& out, int p,
                                      const float4* __restrict__ prep, bool& f0, bool& f1, DingAux* aux) {
  const DingParams D = ding_params(P);
  const float4* rec = prep + kDingPrepVec * int64_t(p);
  const float porosity = in.porosity[p];
  DingOut r;
  r.flag0 = 0;
  r.flag1 = 0;
  float ret_n, cres;
  {
    const float4 e = rec[4], f = rec[5], g = rec[6];
    const float s[3] = {e.z, e.w, f.x};
    ding_solve(D, s, f.w, g.x, f.z, porosity, false, aux, ret_n, cres, r.flag1);
  }
  const float T_local = P.transport_enabled ? out.T_mat[p] : in.T_mat[p];
  const mat3<float> F_new = m_load(out.F + 9 * p);
  ding_complete(D, load_prep(rec), porosity, ret_n, cres, F_new, T_local, r, aux);
  f0 = r.flag0;
  f1 = r.flag1;
  material_outputs(P, sc, in, out, p, F_new, kDingDone, r);
  return kDingDone;
}

// Surrogate inputs of particle p from its Ding extras (see bk_surrogate.h).
__device__ __forceinline__ void surrogate_features(const DingInputs& q, const DingAux& aux, float f[kMlpRow]) {
  f[0] = aux.eps[0];
  f[1] = aux.eps[1];
  f[2] = aux.eps[2];
  f[3] = aux.T;
  f[4] = q.sat;
  f[5] = q.porosity;
  f[6] = q.fp_dev;
  f[7] = 0.0f;
}

// Surrogate network over the queued particles (between the two halves of pass C).
__global__ __launch_bounds__(256) void k_surrogate(bk_workspace* ws) {
  __shared__ __align__(16) uint8_t lds[kMlpLdsBytes];
  mlp_rows(ws->mlp, ws->features, ws->predictions, ws->counters[2], lds);
}

// Second half of pass C: the deferred coupled-porosity return maps, 64 per
// wavefront with no idle lanes waiting on a neighbour's solve.  The loop's
// trip count is uniform per workgroup (sample appends are per workgroup).
// fl: the flag block the pass raises its flags in (ws->flags, or ws->dflags beside the next substep); pending: take
// the queue only while it is marked unprocessed (counters[5]), so a deferred pass launched after it was taken is a
// no-op.
__global__ __launch_bounds__(256) void k_pass_c_deferred(const bk_dough* __restrict__ doughs, bk_scene sc, bk_state in, bk_state out,
                                                         bk_workspace* ws, int32_t* __restrict__ fl, int pending) {
  __shared__ int lds[17];
  __shared__ unsigned long long totals[4];
  if (threadIdx.x < 4) totals[threadIdx.x] = 0ull;
  const int nq = (!pending || ws->counters[5]) ? ws->counters[2] : 0;
  const bool emit_on = ws->samples != nullptr;
  const bool have_mlp = ws->mlp != nullptr;
  unsigned evals = 0, accepted = 0, rejected = 0, handled = 0;
  for (int base = blockIdx.x * blockDim.x; base < nq; base += gridDim.x * blockDim.x) {
    const int i = base + int(threadIdx.x);
    bool emit = false;
    float row[kSampleRow];
    if (i < nq) {
      const int p = ws->queue[i];
      ++handled;
      // per-lane dough: this kernel is bound by its fp64 solves, not by reading the table
      const int d = dough_of_particle(sc.dough_of, p);
      {
        const bk_params& P = doughs[d].params;
        bool f0 = false, f1 = false;
        DingAux aux;
        if (P.surrogate && have_mlp) {
          // start from the network's porosity; the overstress start maps the trial
          // deviator onto the predicted one (|dev tau_ret| = |dev tau_trial| / (1 + df gamma))
          const float* pr = ws->predictions + size_t(i) * kMlpOut;
          aux.guided = true;
          aux.standin = P.surrogate == 2;
          aux.n_start = double(in.porosity[p]) + double(pr[3]);
          const float m = (pr[0] + pr[1] + pr[2]) / 3.0f;
          aux.dev_target = 2.0f * P.mu *
                           sqrtf((pr[0] - m) * (pr[0] - m) + (pr[1] - m) * (pr[1] - m) + (pr[2] - m) * (pr[2] - m));
        }
        const int status = material_finish(P, sc, in, out, p, ws->prep, f0, f1, &aux);
        flag_max(fl + d * BK_FLAG_COUNT, BK_FLAG_POROSITY, f0);
        flag_max(fl + d * BK_FLAG_COUNT, BK_FLAG_COUPLED, f1);
        evals += unsigned(aux.evals);
        if (aux.guided) {
          if (aux.accepted) ++accepted; else ++rejected;
        }
        if (emit_on && P.gas_enabled && status == kDingDone) {
          const float* f = ws->features + size_t(i) * kMlpRow;
#pragma unroll
          for (int k = 0; k < kMlpIn; ++k) row[k] = f[k];
          row[7] = aux.eps_ret[0];
          row[8] = aux.eps_ret[1];
          row[9] = aux.eps_ret[2];
          row[10] = out.porosity[p] - in.porosity[p];
          row[11] = 0.0f;
          emit = true;
        }
      }
    }
    if (emit_on) {
      const int slot = block_reserve(&ws->counters[4], emit ? 1 : 0, lds);
      if (emit && slot < ws->sample_cap) {
#pragma unroll
        for (int k = 0; k < kSampleRow; ++k) ws->samples[size_t(slot) * kSampleRow + k] = row[k];
      }
    }
  }
  // statistics: wavefront sums, then one LDS atomic per wavefront and one global atomic per workgroup
  unsigned v[4] = {evals, accepted, rejected, handled};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  }
  __syncthreads();
  if (__lane_id() == 0)
    for (int k = 0; k < 4; ++k) atomicAdd(&totals[k], (unsigned long long)v[k]);
  __syncthreads();
  if (threadIdx.x < 4 && totals[threadIdx.x]) atomicAdd(&ws->stats[threadIdx.x], totals[threadIdx.x]);
}

// ============================================================================ stopped doughs
// Every state field of particle p from `in` to `out` (the fields the state has).
__device__ __forceinline__ void copy_particle_state(const bk_state& in, const bk_state& out, int p) {
#pragma unroll
  for (int a = 0; a < 3; ++a) { out.x[3 * p + a] = in.x[3 * p + a]; out.v[3 * p + a] = in.v[3 * p + a]; }
#pragma unroll
  for (int a = 0; a < 9; ++a) { out.F[9 * p + a] = in.F[9 * p + a]; out.C[9 * p + a] = in.C[9 * p + a]; }
  if (in.Fe) {
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      out.Fe[9 * p + a] = in.Fe[9 * p + a];
      out.Fp[9 * p + a] = in.Fp[9 * p + a];
      out.tau[9 * p + a] = in.tau[9 * p + a];
    }
    out.residual[p] = in.residual[p];
    out.porosity[p] = in.porosity[p];
    out.T_mat[p] = in.T_mat[p];
    out.gelled[p] = in.gelled[p];
    out.yield_stress[p] = in.yield_stress[p];
    out.viscosity[p] = in.viscosity[p];
    out.porosity_residual[p] = in.porosity_residual[p];
    out.density[p] = in.density[p];
  }
  if (in.moles) { out.moles[p] = in.moles[p]; out.pressure[p] = in.pressure[p]; }
  if (in.T) {
    out.T[p] = in.T[p]; out.W[p] = in.W[p]; out.s[p] = in.s[p];
    out.losses[p] = in.losses[p]; out.exchanges[p] = in.exchanges[p]; out.masses[p] = in.masses[p];
  }
}

// At the start of a substep, each dough's freeze mode from its flags (raised by earlier substeps) and its mode
// of the previous substep: no flag 0; the first substep with one 1; later substeps 2; and the substeps run
// since it stopped, this one included.
__global__ void k_freeze(bk_workspace* ws, int nd) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const int32_t* fl = ws->flags + d * BK_FLAG_COUNT;
  bool any = false;
#pragma unroll
  for (int k = 0; k < BK_FLAG_COUNT; ++k) any |= fl[k] != 0;
  const int old = ws->freeze[d];
  const int mode = !any ? 0 : (old == 0 ? 1 : 2);
  ws->freeze[d] = mode;
  ws->freeze_age[d] = mode ? ws->freeze_age[d] + 1 : 0;
}

// The join of a substep whose first stages ran beside the previous substep's deferred pass (after both): that
// pass raised its flags in dflags.  A dough the substep started unfrozen but that pass flagged stops now, as if
// k_freeze had seen the flags: mode 1, and its flags are the deferred pass's alone (what the substep's first
// stages raised for it does not count, since they would not have run for a stopped dough).  Otherwise the flags
// are merged.  Then the deferred queue is emptied for this substep's pass C and marked unprocessed.
__global__ void k_join(bk_workspace* ws, int nd) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d == 0) {
    ws->counters[2] = 0;
    ws->counters[5] = 1;
  }
  if (d >= nd) return;
  int32_t* fl = ws->flags + d * BK_FLAG_COUNT;
  int32_t* df = ws->dflags + d * BK_FLAG_COUNT;
  bool any = false;
#pragma unroll
  for (int k = 0; k < BK_FLAG_COUNT; ++k) any |= df[k] != 0;
  if (!any) return;
  const bool stopping = ws->freezing && ws->freeze[d] == 0;
#pragma unroll
  for (int k = 0; k < BK_FLAG_COUNT; ++k) {
    fl[k] = stopping ? df[k] : max(fl[k], df[k]);
    df[k] = 0;
  }
  if (stopping) {
    ws->freeze[d] = 1;
    ws->freeze_age[d] = 1;
  }
}

// After the deferred pass that ends a run of such substeps: the queue is processed.
__global__ void k_pending_done(bk_workspace* ws) {
  if (threadIdx.x == 0) ws->counters[5] = 0;
}

// ============================================================================ pass C: G2P + Ding + gas refresh + saturation
// Four wavefronts per SIMD (at most 128 VGPRs): left to itself the compiler takes 134 and the SIMD holds three,
// too few to cover the node gather and the fp64 SVD's dependent divide and square-root chains; five (96 VGPRs)
// spill to scratch and run slower than three.
__global__ __launch_bounds__(kParticleThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_pass_c(
    const bk_dough* __restrict__ doughs, bk_scene sc, bk_state in, bk_state out, bk_grid g, bk_workspace* ws, int n) {
  __shared__ int lds[17];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  bool deferred = false;
  float feat[kMlpRow];
  if (p < n) {
    per_dough(dough_of_particle(sc.dough_of, p), [&](int d) {
      if (const int mode = frozen_mode(ws, d)) {
        // a stopped dough: its output buffer already holds its last accepted state on the first substep
        // after the flag (mode 1: nothing is written); from then on the state is copied through (mode 2)
        if (mode == 2) copy_particle_state(in, out, p);
        return;
      }
      const bk_params& P = doughs[d].params;
      const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
      const int nb = (P.res + 3) / 4;
      const int64_t node_base = int64_t(doughs[d].block_offset) << 6;
      bool stencil_bad = false, f0 = false, f1 = false;
      float g3[3];
      grid_position(in.x + 3 * p, o, P.dx, g3);
      int b[3];
      stencil_base(g3, b);
      if (!valid_stencil(b, P.res)) {
        stencil_bad = true;
      } else {
        DingInputs q;
        if (P.model == 1) q = ding_inputs(P, in, out, p);
        float l[3] = {g3[0] - float(b[0]), g3[1] - float(b[1]), g3[2] - float(b[2])};
        float wx[3], wy[3], wz[3];
        axis_weights(l[0], wx);
        axis_weights(l[1], wy);
        axis_weights(l[2], wz);
        // C = 4/dx^2 sum_i w_i v_i (x_i - x_p) with x_i - x_p = (o_i - l) dx, o_i in {0, 1, 2}^3, formed as
        // 4/dx (sum_i w_i v_i o_i - v l): the offsets are constants of the unrolled loop, so the first sum
        // takes one add (o = 1) or fma (o = 2) per velocity component and axis instead of a product per pair
        float nv[3] = {0.f, 0.f, 0.f}, U[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        const float* __restrict__ gv = g.v + 3 * node_base;
        const StencilTerms st = stencil_terms(b, nb, 3u);
#pragma unroll
        for (int ox = 0; ox < 3; ++ox)
#pragma unroll
          for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
              const float wt = wx[ox] * wy[oy] * wz[oz];
              const float* vp = gv + (st.x[ox] + st.y[oy] + st.z[oz]);
              const int o[3] = {ox, oy, oz};
#pragma unroll
              for (int a = 0; a < 3; ++a) {
                const float wv = wt * vp[a];
                nv[a] += wv;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                  if (o[c] == 1) U[a][c] += wv;
                  if (o[c] == 2) U[a][c] = fmaf(2.0f, wv, U[a][c]);
                }
              }
            }
        mat3<float> Cn;
        const float s1 = 4.0f / P.dx;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int c = 0; c < 3; ++c) Cn.a[a * 3 + c] = s1 * (U[a][c] - nv[a] * l[c]);
        mat3<float> F_old = m_load(in.F + 9 * p);
        mat3<float> inc = m_scale(P.dt, Cn);
        inc.a[0] += 1.0f; inc.a[4] += 1.0f; inc.a[8] += 1.0f;
        mat3<float> F_new = m_mul(inc, F_old);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          out.v[3 * p + a] = nv[a];
          out.x[3 * p + a] = in.x[3 * p + a] + P.dt * nv[a];
        }
        m_store(out.C + 9 * p, Cn);
        m_store(out.F + 9 * p, F_new);
        if (P.model == 1) {
          DingAux aux;
          deferred = material_update(P, sc.visc_param[d], sc, in, out, p, q, F_new, Cn, true, f0, f1, P.gas_enabled ? &aux : nullptr,
                                     ws->prep) == kDingDeferred;
          if (deferred) surrogate_features(q, aux, feat);
        }
      }
      int32_t* fl = dough_flags(ws, d);
      flag_max(fl, BK_FLAG_STENCIL, stencil_bad);
      flag_max(fl, BK_FLAG_POROSITY, f0);
      flag_max(fl, BK_FLAG_COUPLED, f1);
    });
  }
  const int slot = block_reserve(&ws->counters[2], deferred ? 1 : 0, lds);
  if (deferred) {
    ws->queue[slot] = p;
#pragma unroll
    for (int k = 0; k < kMlpRow; ++k) ws->features[size_t(slot) * kMlpRow + k] = feat[k];
  }
}

// ============================================================================ acceptance aggregates
// Per-dough f64 sums (experiments/recording.py:46-112, physics/mpm.py:292-308):
// [0] sum V0 det F, [1] sum capacity * T, [2] sum capacity, [3] sum W,
// [4] sum water losses, [5] sum heat exchanges, [6] sum V0 * gas moles,
// [7] sum particle mass.  A wavefront inside one dough reduces by shuffles and
// adds once per sum; one straddling two doughs adds per lane.
constexpr int kMeasureSums = 8;
__global__ __launch_bounds__(256) void k_measure(bk_scene sc, bk_state s, int n, double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double v[kMeasureSums];
#pragma unroll
  for (int k = 0; k < kMeasureSums; ++k) v[k] = 0.0;
  int d = 0;
  if (p < n) {
    d = dough_of_particle(sc.dough_of, p);
    const double V0 = sc.V0[p];
    v[0] = V0 * m_det(m_cast<double>(m_load(s.F + 9 * p)));
    if (s.T && sc.capacities) {
      const double c = sc.capacities[p];
      v[1] = c * s.T[p];
      v[2] = c;
    }
    if (s.W) v[3] = s.W[p];
    if (s.losses) v[4] = s.losses[p];
    if (s.exchanges) v[5] = s.exchanges[p];
    if (s.moles) v[6] = V0 * s.moles[p];
    v[7] = s.masses ? double(s.masses[p]) : double(sc.m0[p]);
  }
  const int u = __builtin_amdgcn_readfirstlane(d);
  if (__ballot(p < n && d != u) == 0ull) {
#pragma unroll
    for (int k = 0; k < kMeasureSums; ++k) {
      double x = v[k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
      if (__lane_id() == 0 && x != 0.0) unsafeAtomicAdd(out + u * kMeasureSums + k, x);
    }
  } else if (p < n) {
#pragma unroll
    for (int k = 0; k < kMeasureSums; ++k)
      if (v[k] != 0.0) unsafeAtomicAdd(out + d * kMeasureSums + k, v[k]);
  }
}

__global__ void k_untile_f32(const float* __restrict__ tiled, float* __restrict__ dense, int res, int comps) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t total = int64_t(res) * res * res;
  if (i >= total) return;
  int k = int(i % res), j = int((i / res) % res), ii = int(i / (int64_t(res) * res));
  int64_t t = tiled_node(ii, j, k, (res + 3) / 4);
  for (int c = 0; c < comps; ++c) dense[i * comps + c] = tiled[t * comps + c];
}
__global__ void k_untile_f64(const double* __restrict__ tiled, double* __restrict__ dense, int res) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t total = int64_t(res) * res * res;
  if (i >= total) return;
  int k = int(i % res), j = int((i / res) % res), ii = int(i / (int64_t(res) * res));
  dense[i] = tiled[tiled_node(ii, j, k, (res + 3) / 4)];
}

// ============================================================================ host side
static int ws_host(bk_workspace* ws, bk_workspace& h) {
  BK_HIP(hipMemcpy(&h, ws, sizeof(h), hipMemcpyDeviceToHost));
  return 0;
}

static HostWs* host_view(bk_workspace* ws) {
  if (!ws) return nullptr;
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  auto it = g_ws_cache.find(ws);
  if (it != g_ws_cache.end()) return &it->second;
  HostWs h;
  if (ws_host(ws, h.dev)) return nullptr;
  return &(g_ws_cache[ws] = h);
}

// For the adjoint (bk_adjoint.hip): the device dough table of a workspace,
// its dough count, its P2G launch width and its device arrays.  The tiles,
// active blocks, vertices and slot areas are those of the last substep the
// workspace ran.
int bk_internal_tile_launch(bk_workspace* ws, const bk_dough** doughs, int* n_doughs, int* grid_cap,
                            const bk_workspace** device_view) {
  HostWs* hw = host_view(ws);
  if (!hw) return fail("workspace: unknown handle");
  *doughs = hw->dev.doughs;
  *n_doughs = int(hw->doughs.size());
  *grid_cap = hw->tuning.grid_cap;
  *device_view = &hw->dev;  // host copy of the device descriptor: device addresses of the workspace arrays
  return 0;
}

// Optional per-stage timing: HIP events recorded around each stage of the
// substep on the launch stream (no extra synchronisation while recording).
enum { kStageClear, kStageKeys, kStageSort, kStageTiles, kStageMesh, kStagePassA, kStageDiffuse, kStagePassB1, kStagePassB2,
       kStageGrid, kStagePassC, kStageSurrogate, kStagePassCDeferred, kStageSettling, kStageCount };
static const char* kStageNames[kStageCount] = {"clear", "keys", "sort", "tiles", "mesh", "pass_a", "diffuse", "pass_b1",
                                               "pass_b2", "grid", "pass_c", "surrogate", "pass_c_deferred", "settling"};
struct Profiler {
  int max_steps = 0, step = 0;
  std::vector<hipEvent_t> ev;        // [step][stage][start/stop]
  std::vector<uint8_t> recorded;     // [step][stage]: the stage ran (stages a substep skips have no events)
};
static std::unordered_map<const bk_workspace*, Profiler> g_prof;

struct StageTimer {
  Profiler* pr;
  hipStream_t st;
  int stage;
  StageTimer(Profiler* p, hipStream_t s, int k) : pr(p), st(s), stage(k) {
    if (pr) {
      (void)hipEventRecord(pr->ev[(size_t(pr->step) * kStageCount + stage) * 2 + 0], st);
      pr->recorded[size_t(pr->step) * kStageCount + stage] = 1;
    }
  }
  ~StageTimer() {
    if (pr) (void)hipEventRecord(pr->ev[(size_t(pr->step) * kStageCount + stage) * 2 + 1], st);
  }
};

// A substep whose first stages run beside the previous substep's deferred pass (captured graphs only): the side
// stream and the events of the fork and the join.
struct Pipe {
  hipStream_t side;
  hipEvent_t fork, join;
};
// Whether a batch's substeps run that way: tuned so (deferred_wgs > 0), a deferred pass exists (pore gas) and no
// settling controller reads the state after it.
static bool can_pipeline(const HostWs* hw) { return hw->tuning.deferred_wgs > 0 && hw->gas && !hw->settle; }

// One substep of the uploaded batch, enqueued on `st`.  With `pipe`, the previous substep's deferred pass runs on
// the side stream beside stages 1-7 (which neither read nor write what it writes: the Ding, gas and saturation
// fields of the queued particles, its flags), k_join follows both, and this substep's own deferred pass is left for
// the next substep or for bk_*_finish.
static int batch_substep(bk_workspace* ws, HostWs* hw, const bk_scene* sc, const bk_state* in, bk_state* out, bk_grid* g,
                         hipStream_t st, bool allow_profile, const Pipe* pipe = nullptr) {
  const bool pipelined = pipe != nullptr;
  if (hw->doughs.empty()) return fail("bk_batch_substep: no dough table uploaded");
  if (!sc || !in || !out || !g) return fail("bk_batch_substep: null argument");
  if (!sc->dough_of && hw->doughs.size() > 1) return fail("bk_batch_substep: a batch of several doughs needs scene.dough_of");
  const bk_workspace* h = &hw->dev;
  const bk_dough* D = h->doughs;
  const int nd = int(hw->doughs.size());
  const int n = hw->n;
  const bk_tuning& tu = hw->tuning;
  const int cap = tu.grid_cap;
  const int pblock = tu.particle_block;
  Profiler* pr = nullptr;
  if (allow_profile) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    auto it = g_prof.find(ws);
    if (it != g_prof.end() && it->second.step < it->second.max_steps) pr = &it->second;
  }
  // 0. with freezing on, each dough's freeze mode for this substep from the flags earlier substeps raised
  if (h->freezing) {
    hipLaunchKernelGGL(k_freeze, dim3(div_up(nd, 64)), dim3(64), 0, st, ws, nd);
    BK_LAUNCH(freeze);
  }
  if (pipelined) BK_HIP(hipEventRecord(pipe->fork, st));
  // The previous substep's deferred pass on the side stream (its input state is this substep's output buffers and
  // back), issued after stage 1-2's launch so that the stages keep the launch stream's hardware queue in the graph.
  // tu.deferred_wgs workgroups at most (the pass strides over its queue): fewer hold fewer of the wavefront slots
  // the stages beside it need, for longer.
  auto side_branch = [&]() -> int {
    if (!pipelined) return 0;
    BK_HIP(hipStreamWaitEvent(pipe->side, pipe->fork, 0));
    hipLaunchKernelGGL(k_pass_c_deferred, dim3(clampi(div_up(n, 256), 1, tu.deferred_wgs)), dim3(256), 0, pipe->side, D, *sc, *out,
                       *in, ws, h->dflags, 1);
    BK_LAUNCH(pass_c_deferred);
    BK_HIP(hipEventRecord(pipe->join, pipe->side));
    return 0;
  };
  // 1. clear the previous substep's blocks (counters[1] still holds their count)
  // 2. keys and ranks within blocks (also resets the per-substep counters)
  // One launch; under stage profiling, one launch per stage.
  const int key_wgs = div_up(n, 256);
  if (pr) {
    {
      StageTimer tm(pr, st, kStageClear);
      hipLaunchKernelGGL(k_clear_keys, dim3(cap), dim3(256), 0, st, D, sc->dough_of, in->x, n, h->keys_raw, h->vals_s[1], ws, *g,
                         int(hw->thermal), cap, int(!pipelined));
      BK_LAUNCH(clear);
    }
    {
      StageTimer tm(pr, st, kStageKeys);
      hipLaunchKernelGGL(k_clear_keys, dim3(key_wgs), dim3(256), 0, st, D, sc->dough_of, in->x, n, h->keys_raw, h->vals_s[1], ws,
                         *g, int(hw->thermal), 0, int(!pipelined));
      BK_LAUNCH(keys);
    }
  } else {
    hipLaunchKernelGGL(k_clear_keys, dim3(cap + key_wgs), dim3(256), 0, st, D, sc->dough_of, in->x, n, h->keys_raw, h->vals_s[1], ws,
                       *g, int(hw->thermal), cap, int(!pipelined));
    BK_LAUNCH(clear_keys);
  }
  if (int rc = side_branch()) return rc;
  {  // 3. order by block, then by cell inside each block
    StageTimer tm(pr, st, kStageSort);
    const int nb = h->total_blocks;
    const int nscan = div_up(nb, kScanTile);
    hipLaunchKernelGGL(k_block_scan_partials, dim3(nscan), dim3(256), 0, st, h->blk_count, nb, h->scan_part, h->counters);
    hipLaunchKernelGGL(k_block_scan_apply, dim3(nscan), dim3(256), 0, st, h->blk_count, nb, h->scan_part, ws);
    // the scatter, with the tile lists and the face-vertex marks in the same launch (under stage profiling, the
    // lists and marks get a launch of their own in stage 4)
    const int list_wgs = clampi(div_up((long long)n * 8, 256), 1, 1024);
    const int mark_slots = (hw->thermal && hw->n_vertices > 0 && hw->n_slots > 0) ? hw->n_slots : 0;
    const int mark_wgs = div_up(mark_slots, 256);
    const int scatter_wgs = div_up(n, 256);
    if (pr)
      hipLaunchKernelGGL(k_activate, dim3(scatter_wgs), dim3(256), 0, st, D, nd, ws, tu.tile_chunk, *sc, 0, 0, h->keys_raw,
                         h->vals_s[1], n, h->keys_s[0], h->vals_s[0], scatter_wgs);
    else
      hipLaunchKernelGGL(k_activate, dim3(list_wgs + mark_wgs + scatter_wgs), dim3(256), 0, st, D, nd, ws, tu.tile_chunk, *sc,
                         mark_slots, mark_wgs, h->keys_raw, h->vals_s[1], n, h->keys_s[0], h->vals_s[0], scatter_wgs);
    // the cell sort, with the mesh reconstruction in the same launch outside stage profiling (its vertex marks are set)
    const int sort_wgs = div_up(div_up(clampi(div_up(n, 64), 1, 16384), 4), kXcds) * kXcds;
    const int rec_nv = (!pr && hw->thermal && hw->n_vertices > 0 && hw->n_slots > 0) ? hw->n_vertices : 0;
    hipLaunchKernelGGL(k_sort_reconstruct, dim3(sort_wgs + rec_nv), dim3(256), 0, st, h->keys_s[0], h->vals_s[0], ws, h->sorted_keys,
                       h->sorted_vals, sort_wgs, in->x, in->F, sc->ref_vertices, sc->bind_offsets, sc->bind_particles,
                       sc->bind_weights, sc->bind_ref_offsets, h->vertices, rec_nv);
    BK_LAUNCH(sort);
  }
  if (pr) {  // 4. tiles, active node blocks (launched with the scatter outside stage profiling)
    StageTimer tm(pr, st, kStageTiles);
    const int mark_slots = (hw->thermal && hw->n_vertices > 0 && hw->n_slots > 0) ? hw->n_slots : 0;
    const int mark_wgs = div_up(mark_slots, 256);
    hipLaunchKernelGGL(k_activate, dim3(clampi(div_up((long long)n * 8, 256), 1, 1024) + mark_wgs), dim3(256), 0, st, D, nd, ws,
                       tu.tile_chunk, *sc, mark_slots, mark_wgs, h->keys_raw, h->vals_s[1], n, h->keys_s[0], h->vals_s[0], 0);
    BK_LAUNCH(tiles);
  }
  if (hw->thermal) {
    {  // 5. mesh reconstruction and surface areas
      StageTimer tm(pr, st, kStageMesh);
      if (hw->n_vertices > 0 && hw->n_slots > 0) {
        if (pr)  // outside stage profiling the reconstruction ran with the cell sort
          hipLaunchKernelGGL(k_sort_reconstruct, dim3(hw->n_vertices), dim3(256), 0, st, h->keys_s[0], h->vals_s[0], ws,
                             h->sorted_keys, h->sorted_vals, 0, in->x, in->F, sc->ref_vertices, sc->bind_offsets,
                             sc->bind_particles, sc->bind_weights, sc->bind_ref_offsets, h->vertices, hw->n_vertices);
      }
      if (hw->n_slots > 0)
        hipLaunchKernelGGL(k_slot_areas, dim3(div_up(hw->n_slots, 256)), dim3(256), 0, st, D, *sc, h->vertices, ws, hw->n_slots);
      BK_LAUNCH(mesh);
    }
    {  // 6. thermal P2G
      StageTimer tm(pr, st, kStagePassA);
      hipLaunchKernelGGL(k_pass_a, dim3(cap), dim3(kTileThreads), 0, st, D, *sc, *in, *g, ws);
      BK_LAUNCH(pass_a);
    }
    {  // 7. diffusion
      StageTimer tm(pr, st, kStageDiffuse);
      hipLaunchKernelGGL(k_diffuse, dim3(cap), dim3(256), 0, st, D, nd, *g, ws);
      BK_LAUNCH(diffuse);
    }
  }
  if (pipelined) {
    BK_HIP(hipStreamWaitEvent(st, pipe->join, 0));
    hipLaunchKernelGGL(k_join, dim3(div_up(nd, 64)), dim3(64), 0, st, ws, nd);
    BK_LAUNCH(join);
  }
  {  // 8. scalar G2P, gas and water stress
    StageTimer tm(pr, st, kStagePassB1);
    hipLaunchKernelGGL(k_pass_b1, dim3(div_up(n, pblock)), dim3(pblock), 0, st, D, *sc, *in, *out, *g, ws, n);
    BK_LAUNCH(pass_b1);
  }
  {  // 9. mechanical P2G
    StageTimer tm(pr, st, kStagePassB2);
    hipLaunchKernelGGL(k_pass_b2, dim3(cap), dim3(kTileThreads), 0, st, D, *sc, *in, *out, *g, ws);
    BK_LAUNCH(pass_b2);
  }
  {  // 10. grid update
    StageTimer tm(pr, st, kStageGrid);
    hipLaunchKernelGGL(k_grid_update, dim3(cap), dim3(256), 0, st, D, nd, *g, ws);
    BK_LAUNCH(grid_update);
  }
  {  // 11. G2P + material
    StageTimer tm(pr, st, kStagePassC);
    hipLaunchKernelGGL(k_pass_c, dim3(div_up(n, pblock)), dim3(pblock), 0, st, D, *sc, *in, *out, *g, ws, n);
    BK_LAUNCH(pass_c);
  }
  if (hw->gas && hw->surrogate && h->mlp) {  // 11b. return-map network on the queued particles
    StageTimer tm(pr, st, kStageSurrogate);
    hipLaunchKernelGGL(k_surrogate, dim3(clampi(div_up(n, 256), 1, 1024)), dim3(256), 0, st, ws);
    BK_LAUNCH(surrogate);
  }
  if (hw->gas && !pipelined) {  // 12. deferred coupled-porosity return maps
    StageTimer tm(pr, st, kStagePassCDeferred);
    hipLaunchKernelGGL(k_pass_c_deferred, dim3(clampi(div_up(n, 256), 1, 1024)), dim3(256), 0, st, D, *sc, *in, *out, ws, h->flags, 0);
    BK_LAUNCH(pass_c_deferred);
  }
  if (hw->settle) {  // 13. settling controller (rest metrics and the rest/bake decision)
    StageTimer tm(pr, st, kStageSettling);
    int rc = bk_internal_settling_launch(ws, D, nd, hw->max_dough_n, out->v, out->C, in->Fp, out->Fp, st);
    if (rc) return rc;
  }
  if (pr) ++pr->step;
  return 0;
}

static bool same_layout(const std::vector<bk_dough>& a, const bk_dough* b, int nd) {
  if (int(a.size()) != nd) return false;
  for (int d = 0; d < nd; ++d)
    if (a[d].block_offset != b[d].block_offset || a[d].params.res != b[d].params.res) return false;
  return true;
}

static bool same_table(const HostWs* hw, const bk_dough* table, int nd) {
  return int(hw->doughs.size()) == nd && memcmp(hw->doughs.data(), table, sizeof(bk_dough) * nd) == 0;
}

// Validate and upload a dough table.  Waits for the device first, since
// kernels in flight may read the table.  A table whose grid layout differs from
// the last one also forgets the last substep's active blocks: the first
// substep after it needs a zeroed grid.
static int upload_doughs(bk_workspace* ws, HostWs* hw, const bk_dough* table, int nd) {
  const bk_workspace& h = hw->dev;
  if (!table || nd < 1 || nd > h.max_doughs) return fail("bk_workspace_set_doughs: dough count exceeds the workspace");
  if (same_table(hw, table, nd)) return 0;
  long long n = 0, blocks = 0, nv = 0, ns = 0;
  int max_dough_n = 0;
  bool thermal = false, gas = false, surrogate = false, settle = false;
  for (int d = 0; d < nd; ++d) {
    const bk_dough& D = table[d];
    if (D.params.settling < 0 || D.params.settling > 2) return fail("bk_workspace_set_doughs: params.settling must be 0, 1 or 2");
    if (D.params.res < 4 || D.params.n < 1 || !(D.params.dx > 0.0f))
      return fail("bk_workspace_set_doughs: every dough needs res >= 4, n >= 1 and dx > 0");
    // the G2P gathers form a dough's float index 3 * node in 32 bits
    if (D.params.res > kMaxDoughRes) return fail("bk_workspace_set_doughs: a dough's grid resolution is at most 1024");
    if (D.particle_offset != n || D.block_offset != blocks)
      return fail("bk_workspace_set_doughs: doughs must be packed in order (particle_offset, block_offset)");
    long long nb = (D.params.res + 3) / 4;
    n += D.params.n;
    blocks += nb * nb * nb;
    nv += D.params.n_vertices;
    ns += D.params.n_slots;
    thermal |= D.params.transport_enabled != 0;
    gas |= D.params.model == 1 && D.params.gas_enabled != 0;
    surrogate |= D.params.surrogate != 0;
    settle |= D.params.settling == 2;
    max_dough_n = D.params.n > max_dough_n ? D.params.n : max_dough_n;
  }
  if (n > h.max_n || blocks > h.total_blocks || nv > h.n_vertices || ns > h.n_slots)
    return fail("bk_workspace_set_doughs: batch exceeds the workspace (particles, blocks, vertices or slots)");
  const bool relayout = !same_layout(hw->doughs, table, nd);
  BK_HIP(hipDeviceSynchronize());
  BK_HIP(hipMemcpy(h.doughs, table, sizeof(bk_dough) * nd, hipMemcpyHostToDevice));
  if (relayout) {
    BK_HIP(hipMemset(h.node_flag, 0, size_t(h.total_blocks) * 4));
    BK_HIP(hipMemset(h.counters, 0, 64));
  }
  if (settle) {  // the settling controller's partials: 3 floats per dough and run of 256 particles
    const int64_t need = int64_t(nd) * ((max_dough_n + 255) / 256) * 3;
    if (need > hw->dev.settle_part_cap) {
      if (hw->dev.settle_part) BK_HIP(hipFree(hw->dev.settle_part));
      hw->dev.settle_part = nullptr;
      hw->dev.settle_part_cap = 0;
      BK_HIP(hipMalloc(&hw->dev.settle_part, size_t(need) * sizeof(float)));
      hw->dev.settle_part_cap = need;
      BK_HIP(hipMemcpy(ws, &hw->dev, sizeof(bk_workspace), hipMemcpyHostToDevice));
    }
  }
  if (thermal && !hw->dev.dTW) {  // 16 B per node of the workspace's grid
    BK_HIP(hipMalloc(&hw->dev.dTW, size_t(h.total_blocks) * 64 * sizeof(double2)));
    BK_HIP(hipMemcpy(ws, &hw->dev, sizeof(bk_workspace), hipMemcpyHostToDevice));
  }
  hw->doughs.assign(table, table + nd);
  hw->n = int(n);
  hw->n_vertices = int(nv);
  hw->n_slots = int(ns);
  hw->thermal = thermal;
  hw->gas = gas;
  hw->surrogate = surrogate;
  hw->settle = settle;
  hw->max_dough_n = max_dough_n;
  int bits = 6;
  while ((1ll << bits) < blocks * 64) ++bits;
  hw->key_bits = bits;
  return 0;
}

constexpr int kMaxGraphSteps = 64;  // substeps per captured graph

struct bk_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bk_workspace* ws = nullptr;
  std::vector<bk_dough> table;  // the table the graph was captured with
};

// pipelined: the substep with its first stages beside the previous substep's deferred pass, when the batch allows
// it (can_pipeline); else the plain substep.  steps: that many substeps in one graph, alternating in -> out -> in ...
// (one graph launch instead of `steps`).
static bk_graph* capture(bk_workspace* ws, HostWs* hw, const bk_scene* sc, const bk_state* in, bk_state* out, bk_grid* g,
                         bool pipelined = false, int steps = 1) {
  hipStream_t cap = nullptr;
  if (check_hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreateWithFlags")) return nullptr;
  Pipe pipe{nullptr, nullptr, nullptr};
  const bool pipe_on = pipelined && can_pipeline(hw);
  int rc = 0;
  if (pipe_on) {
    rc = check_hip(hipStreamCreateWithFlags(&pipe.side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&pipe.fork, hipEventDisableTiming), "hipEventCreateWithFlags");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&pipe.join, hipEventDisableTiming), "hipEventCreateWithFlags");
  }
  bk_graph* gr = new bk_graph();
  gr->ws = ws;
  gr->table = hw->doughs;
  if (!rc) rc = check_hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (!rc) {
    bk_state a = *in, b = *out;
    int rc_sub = 0;
    for (int k = 0; k < steps && !rc_sub; ++k)
      rc_sub = batch_substep(ws, hw, sc, k % 2 ? &b : &a, k % 2 ? &a : &b, g, cap, false, pipe_on ? &pipe : nullptr);
    int rc_end = check_hip(hipStreamEndCapture(cap, &gr->graph), "hipStreamEndCapture");
    rc = rc_sub ? rc_sub : rc_end;
  }
  if (!rc) rc = check_hip(hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipStreamDestroy(cap);
  if (pipe.side) (void)hipStreamDestroy(pipe.side);
  if (pipe.fork) (void)hipEventDestroy(pipe.fork);
  if (pipe.join) (void)hipEventDestroy(pipe.join);
  if (rc) {
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    delete gr;
    return nullptr;
  }
  return gr;
}

static bk_dough single_dough(const bk_params* P) {
  bk_dough d;
  memset(&d, 0, sizeof(d));
  d.params = *P;
  return d;
}

// ============================================================================ C ABI
extern "C" {

int bk_abi_version(void) { return BK_ABI_VERSION; }
const char* bk_last_error(void) { return g_last_error.c_str(); }

int bk_device_info(char* name, int name_len, int* cu_count, int* xcd_count) {
  int dev = 0;
  BK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  BK_HIP(hipGetDeviceProperties(&prop, dev));
  if (name && name_len > 0) { strncpy(name, prop.gcnArchName, name_len - 1); name[name_len - 1] = 0; }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (xcd_count) *xcd_count = 8;
  return 0;
}

bk_workspace* bk_workspace_create_batch(int32_t max_particles, int32_t total_blocks, int32_t max_doughs, int32_t n_vertices,
                                        int32_t n_slots) {
  if (max_particles < 1 || max_particles > (1 << kIndexBits) || total_blocks < 1 || max_doughs < 1 || n_vertices < 0 ||
      n_slots < 0 || (long long)total_blocks * 64 >= (1ll << 32)) {
    fail("bk_workspace_create: bad sizes");
    return nullptr;
  }
  bk_workspace* ws = nullptr;
  if (hipMalloc(&ws, sizeof(bk_workspace)) != hipSuccess) { fail("hipMalloc workspace"); return nullptr; }
  bk_workspace h;
  h.max_n = max_particles;
  h.total_blocks = total_blocks;
  h.max_doughs = max_doughs;
  h.n_vertices = n_vertices;
  h.n_slots = n_slots;
  size_t n = size_t(max_particles);
  size_t nbk = size_t(total_blocks);
  int nwg = div_up(max_particles, kSortTile);
  bool ok = true;
  auto alloc = [&](void** p, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (hipMalloc(p, bytes) != hipSuccess) ok = false;
    else if (hipMemset(*p, 0, bytes) != hipSuccess) ok = false;
  };
  alloc((void**)&h.doughs, sizeof(bk_dough) * size_t(max_doughs));
  alloc((void**)&h.keys_raw, n * 4);
  alloc((void**)&h.sorted_keys, n * 4);
  alloc((void**)&h.sorted_vals, n * 4);
  for (int i = 0; i < 2; ++i) {
    alloc((void**)&h.keys_s[i], n * 4);
    alloc((void**)&h.vals_s[i], n * 4);
  }
  alloc((void**)&h.hist, size_t(256) * nwg * 4);
  alloc((void**)&h.digit_total, 256 * 4);
  alloc((void**)&h.blk_start, nbk * 4);
  alloc((void**)&h.blk_count, nbk * 4);
  alloc((void**)&h.scan_part, size_t(div_up((long long)nbk, kScanTile)) * 4);
  alloc((void**)&h.blk_end, nbk * 4);
  alloc((void**)&h.node_flag, nbk * 4);
  alloc((void**)&h.blocks, nbk * 4);
  alloc((void**)&h.tiles, (n + nbk) * sizeof(int4));  // a tile per chunk: at most one per particle plus one per block
  alloc((void**)&h.active, nbk * 4);
  alloc((void**)&h.counters, 64);
  alloc((void**)&h.tau, n * 36);
  alloc((void**)&h.queue, n * 4);
  alloc((void**)&h.prep, n * kDingPrepVec * sizeof(float4));
  alloc((void**)&h.vertices, size_t(n_vertices) * 12);
  alloc((void**)&h.vert_mark, size_t(n_vertices) * 4);
  alloc((void**)&h.slot_exposed, size_t(n_slots) * 8);
  alloc((void**)&h.slot_contact, size_t(n_slots) * 8);
  alloc((void**)&h.flags, size_t(BK_FLAG_COUNT) * max_doughs * 4);
  alloc((void**)&h.dflags, size_t(BK_FLAG_COUNT) * max_doughs * 4);
  alloc((void**)&h.features, n * kMlpRow * 4);
  alloc((void**)&h.predictions, n * kMlpOut * 4);
  alloc((void**)&h.stats, 4 * sizeof(unsigned long long));
  alloc((void**)&h.settle, sizeof(bk_settling) * size_t(max_doughs));
  alloc((void**)&h.freeze, size_t(max_doughs) * 4);
  alloc((void**)&h.freeze_age, size_t(max_doughs) * 4);
  if (!ok || hipMemcpy(ws, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) {
    fail("bk_workspace_create: allocation failed");
    return nullptr;
  }
  return ws;
}

bk_workspace* bk_workspace_create(int32_t max_particles, int32_t res, int32_t n_vertices, int32_t n_slots) {
  if (res < 4) { fail("bk_workspace_create: bad sizes"); return nullptr; }
  int nb = (res + 3) / 4;
  return bk_workspace_create_batch(max_particles, nb * nb * nb, 1, n_vertices, n_slots);
}

void bk_workspace_destroy(bk_workspace* ws) {
  if (!ws) return;
  bk_workspace h;
  {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    g_ws_cache.erase(ws);
    auto it = g_prof.find(ws);
    if (it != g_prof.end()) {
      for (hipEvent_t e : it->second.ev) (void)hipEventDestroy(e);
      g_prof.erase(it);
    }
  }
  if (ws_host(ws, h)) return;
  void* ptrs[] = {h.doughs, h.keys_raw, h.sorted_keys, h.sorted_vals, h.keys_s[0], h.keys_s[1], h.vals_s[0], h.vals_s[1],
                  h.hist, h.digit_total, h.blk_start, h.blk_count, h.scan_part, h.blk_end, h.node_flag, h.blocks, h.tiles, h.active, h.counters,
                  h.queue, h.prep, h.tau, h.vertices, h.vert_mark, h.slot_exposed, h.slot_contact, h.flags, h.dflags, h.features, h.predictions,
                  h.stats, h.settle, h.settle_part, h.dTW, h.freeze, h.freeze_age};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  (void)hipFree(ws);
}

int bk_workspace_set_doughs(bk_workspace* ws, const bk_dough* table, int32_t n_doughs) {
  HostWs* hw = host_view(ws);
  if (!hw) return fail("bk_workspace_set_doughs: bad workspace");
  return upload_doughs(ws, hw, table, n_doughs);
}

int bk_workspace_settling_reference(bk_workspace* ws, float* reference) {
  HostWs* hw = host_view(ws);
  if (!hw) return fail("bk_workspace_settling_reference: bad workspace");
  BK_HIP(hipDeviceSynchronize());
  hw->dev.settle_ref = reference;
  BK_HIP(hipMemcpy(ws, &hw->dev, sizeof(bk_workspace), hipMemcpyHostToDevice));
  return 0;
}

int bk_workspace_settling_set(bk_workspace* ws, int32_t first, int32_t count, const bk_settling* records, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !records || count < 1) return fail("bk_workspace_settling_set: bad argument");
  if (first < 0 || (long long)first + count > hw->dev.max_doughs)
    return fail("bk_workspace_settling_set: dough range outside the workspace");
  std::vector<bk_settling> r(records, records + count);
  for (bk_settling& x : r) {
    if (x.phase != 0 && x.phase != 1) return fail("bk_workspace_settling_set: phase must be 0 or 1");
    if (x.ref_slot != 0 && x.ref_slot != 1) return fail("bk_workspace_settling_set: ref_slot must be 0 or 1");
    if (!(x.time_step > 0.0)) return fail("bk_workspace_settling_set: time_step must be positive");
    if (x.auto_mode && !hw->dev.settle_ref)
      return fail("bk_workspace_settling_set: auto mode needs a reference buffer (bk_workspace_settling_reference)");
  }
  hipStream_t st = (hipStream_t)stream;
  BK_HIP(hipMemcpyAsync(hw->dev.settle + first, r.data(), sizeof(bk_settling) * size_t(count), hipMemcpyHostToDevice, st));
  BK_HIP(hipStreamSynchronize(st));
  return 0;
}

int bk_workspace_settling_get(bk_workspace* ws, int32_t first, int32_t count, bk_settling* records, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !records || count < 1) return fail("bk_workspace_settling_get: bad argument");
  if (first < 0 || (long long)first + count > hw->dev.max_doughs)
    return fail("bk_workspace_settling_get: dough range outside the workspace");
  hipStream_t st = (hipStream_t)stream;
  BK_HIP(hipMemcpyAsync(records, hw->dev.settle + first, sizeof(bk_settling) * size_t(count), hipMemcpyDeviceToHost, st));
  BK_HIP(hipStreamSynchronize(st));
  return 0;
}

int bk_workspace_set_tuning(bk_workspace* ws, const bk_tuning* tuning) {
  HostWs* hw = host_view(ws);
  if (!hw || !tuning) return fail("bk_workspace_set_tuning: bad argument");
  const bk_tuning t = *tuning;
  if (t.tile_chunk < 1 || t.tile_chunk > kMaxTileChunk || t.grid_cap < 1 || t.grid_cap > 65536 || t.particle_block < 64 ||
      t.particle_block > kParticleThreads || t.particle_block % 64 || t.deferred_wgs < 0 || t.deferred_wgs > 65536)
    return fail("bk_workspace_set_tuning: need tile_chunk in [1, 2^20], grid_cap in [1, 65536], particle_block in {64, 128, 192, "
                "256}, deferred_wgs in [0, 65536]");
  hw->tuning = t;
  return 0;
}

int bk_workspace_get_tuning(bk_workspace* ws, bk_tuning* tuning) {
  HostWs* hw = host_view(ws);
  if (!hw || !tuning) return fail("bk_workspace_get_tuning: bad argument");
  *tuning = hw->tuning;
  return 0;
}

// Device pointers stored in the device descriptor: written with a blocking copy after the device is idle.
static int set_device_field(bk_workspace* ws, size_t offset, const void* value, size_t bytes) {
  BK_HIP(hipDeviceSynchronize());
  BK_HIP(hipMemcpy((char*)ws + offset, value, bytes, hipMemcpyHostToDevice));
  return 0;
}

int bk_workspace_set_surrogate(bk_workspace* ws, const void* blob) {
  HostWs* hw = host_view(ws);
  if (!hw) return fail("bk_workspace_set_surrogate: bad workspace");
  const uint8_t* b = (const uint8_t*)blob;
  int rc = set_device_field(ws, offsetof(bk_workspace, mlp), &b, sizeof(b));
  if (rc) return rc;
  hw->dev.mlp = b;
  return 0;
}

int bk_workspace_set_samples(bk_workspace* ws, float* buffer, int32_t capacity) {
  HostWs* hw = host_view(ws);
  if (!hw || capacity < 0 || (buffer && capacity < 1)) return fail("bk_workspace_set_samples: bad argument");
  int rc = set_device_field(ws, offsetof(bk_workspace, samples), &buffer, sizeof(buffer));
  if (!rc) rc = set_device_field(ws, offsetof(bk_workspace, sample_cap), &capacity, sizeof(capacity));
  if (rc) return rc;
  BK_HIP(hipMemset(hw->dev.counters + 4, 0, 4));
  hw->dev.samples = buffer;
  hw->dev.sample_cap = capacity;
  return 0;
}

int bk_workspace_stats(bk_workspace* ws, int64_t* out, int32_t reset, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !out) return fail("bk_workspace_stats: bad argument");
  hipStream_t st = (hipStream_t)stream;
  unsigned long long s[4];
  int32_t samples = 0;
  BK_HIP(hipMemcpyAsync(s, hw->dev.stats, sizeof(s), hipMemcpyDeviceToHost, st));
  BK_HIP(hipMemcpyAsync(&samples, hw->dev.counters + 4, 4, hipMemcpyDeviceToHost, st));
  BK_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < 4; ++k) out[k] = int64_t(s[k]);
  out[4] = samples;
  if (reset) {
    BK_HIP(hipMemsetAsync(hw->dev.stats, 0, sizeof(s), st));
    BK_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

int bk_workspace_sort_order(bk_workspace* ws, uint32_t* keys, uint32_t* sorted_keys, int32_t* order, int32_t n, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || n < 0 || n > hw->dev.max_n) return fail("bk_workspace_sort_order: bad argument");
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (keys) BK_HIP(hipMemcpyAsync(keys, hw->dev.keys_raw, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
  if (sorted_keys) BK_HIP(hipMemcpyAsync(sorted_keys, hw->dev.sorted_keys, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
  if (order) BK_HIP(hipMemcpyAsync(order, hw->dev.sorted_vals, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
  return 0;
}

int32_t* bk_workspace_flags(bk_workspace* ws) {
  HostWs* hw = host_view(ws);
  return hw ? hw->dev.flags : nullptr;
}
float* bk_workspace_vertices(bk_workspace* ws) {
  HostWs* hw = host_view(ws);
  return hw ? hw->dev.vertices : nullptr;
}
int bk_workspace_counts(bk_workspace* ws, int32_t* out3, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !out3) return fail("bk_workspace_counts: bad argument");
  int32_t c[8];
  BK_HIP(hipMemcpyAsync(c, hw->dev.counters, sizeof(c), hipMemcpyDeviceToHost, (hipStream_t)stream));
  BK_HIP(hipStreamSynchronize((hipStream_t)stream));
  out3[0] = c[0] + c[7];  // full and sparse tiles
  out3[1] = c[1];
  out3[2] = c[2];
  return 0;
}

int bk_workspace_read_flags_all(bk_workspace* ws, int32_t* host, int32_t n_doughs, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !host || n_doughs < 1 || n_doughs > hw->dev.max_doughs) return fail("bk_workspace_read_flags: bad argument");
  BK_HIP(hipMemcpyAsync(host, hw->dev.flags, size_t(BK_FLAG_COUNT) * n_doughs * 4, hipMemcpyDeviceToHost, (hipStream_t)stream));
  BK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}
int bk_workspace_write_flags_all(bk_workspace* ws, const int32_t* host, int32_t n_doughs, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !host || n_doughs < 1 || n_doughs > hw->dev.max_doughs) return fail("bk_workspace_write_flags: bad argument");
  BK_HIP(hipMemcpyAsync(hw->dev.flags, host, size_t(BK_FLAG_COUNT) * n_doughs * 4, hipMemcpyHostToDevice, (hipStream_t)stream));
  // a dough whose flags are written clear runs again; one written with a flag keeps its current state from
  // the next substep on (mode 1: the next k_freeze makes it 2, copying its state through)
  std::vector<int32_t> modes(static_cast<size_t>(n_doughs), 0), ages(static_cast<size_t>(n_doughs), 0);
  for (int d = 0; d < n_doughs; ++d)
    for (int k = 0; k < BK_FLAG_COUNT; ++k)
      if (host[d * BK_FLAG_COUNT + k]) modes[d] = 1;
  BK_HIP(hipMemcpyAsync(hw->dev.freeze, modes.data(), size_t(n_doughs) * 4, hipMemcpyHostToDevice, (hipStream_t)stream));
  BK_HIP(hipMemcpyAsync(hw->dev.freeze_age, ages.data(), size_t(n_doughs) * 4, hipMemcpyHostToDevice, (hipStream_t)stream));
  BK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int bk_workspace_set_freezing(bk_workspace* ws, int32_t on) {
  HostWs* hw = host_view(ws);
  if (!hw || (on != 0 && on != 1)) return fail("bk_workspace_set_freezing: bad argument");
  int rc = set_device_field(ws, offsetof(bk_workspace, freezing), &on, sizeof(on));
  if (rc) return rc;
  hw->dev.freezing = on;
  return 0;
}

int bk_workspace_freeze_modes(bk_workspace* ws, int32_t* host, int32_t n_doughs, int32_t write, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !host || n_doughs < 1 || n_doughs > hw->dev.max_doughs) return fail("bk_workspace_freeze_modes: bad argument");
  hipStream_t st = (hipStream_t)stream;
  std::vector<int32_t> modes(static_cast<size_t>(n_doughs)), ages(static_cast<size_t>(n_doughs));
  if (write) {
    for (int d = 0; d < n_doughs; ++d) {
      modes[d] = host[2 * d];
      ages[d] = host[2 * d + 1];
      if (modes[d] < 0 || modes[d] > 2 || ages[d] < 0) return fail("bk_workspace_freeze_modes: modes are 0, 1 or 2, ages >= 0");
    }
    BK_HIP(hipMemcpyAsync(hw->dev.freeze, modes.data(), size_t(n_doughs) * 4, hipMemcpyHostToDevice, st));
    BK_HIP(hipMemcpyAsync(hw->dev.freeze_age, ages.data(), size_t(n_doughs) * 4, hipMemcpyHostToDevice, st));
    BK_HIP(hipStreamSynchronize(st));
  } else {
    BK_HIP(hipMemcpyAsync(modes.data(), hw->dev.freeze, size_t(n_doughs) * 4, hipMemcpyDeviceToHost, st));
    BK_HIP(hipMemcpyAsync(ages.data(), hw->dev.freeze_age, size_t(n_doughs) * 4, hipMemcpyDeviceToHost, st));
    BK_HIP(hipStreamSynchronize(st));
    for (int d = 0; d < n_doughs; ++d) {
      host[2 * d] = modes[d];
      host[2 * d + 1] = ages[d];
    }
  }
  return 0;
}int bk_workspace_read_flags(bk_workspace* ws, int32_t* host8, void* stream) { return bk_workspace_read_flags_all(ws, host8, 1, stream); }
int bk_workspace_write_flags(bk_workspace* ws, const int32_t* host8, void* stream) {
  return bk_workspace_write_flags_all(ws, host8, 1, stream);
}

int bk_profile_begin(bk_workspace* ws, int32_t max_steps) {
  if (max_steps < 1) return fail("bk_profile_begin: max_steps must be positive");
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  Profiler& pr = g_prof[ws];
  for (hipEvent_t e : pr.ev) (void)hipEventDestroy(e);
  pr.ev.assign(size_t(max_steps) * kStageCount * 2, nullptr);
  pr.recorded.assign(size_t(max_steps) * kStageCount, 0);
  for (auto& e : pr.ev) BK_HIP(hipEventCreate(&e));
  pr.max_steps = max_steps;
  pr.step = 0;
  return 0;
}

int bk_profile_end(bk_workspace* ws, float* ms_per_stage, int32_t* steps_recorded) {
  Profiler* pr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    auto it = g_prof.find(ws);
    if (it == g_prof.end()) return fail("bk_profile_end: profiling not started");
    pr = &it->second;
  }
  for (int k = 0; k < kStageCount; ++k) ms_per_stage[k] = 0.0f;
  for (int s = 0; s < pr->step; ++s)
    for (int k = 0; k < kStageCount; ++k) {
      if (!pr->recorded[size_t(s) * kStageCount + k]) continue;
      hipEvent_t a = pr->ev[(size_t(s) * kStageCount + k) * 2], b = pr->ev[(size_t(s) * kStageCount + k) * 2 + 1];
      BK_HIP(hipEventSynchronize(b));
      float ms = 0.0f;
      BK_HIP(hipEventElapsedTime(&ms, a, b));
      ms_per_stage[k] += ms;
    }
  *steps_recorded = pr->step;
  for (hipEvent_t e : pr->ev) (void)hipEventDestroy(e);
  {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    g_prof.erase(ws);
  }
  return 0;
}

const char* bk_profile_stage_name(int32_t k) { return (k >= 0 && k < kStageCount) ? kStageNames[k] : ""; }
int32_t bk_profile_stage_count(void) { return kStageCount; }

int bk_cell_keys(const float* x, int32_t n, const float* o, float dx, int32_t res, int32_t tiled, uint32_t* keys, int32_t* flags,
                 void* stream) {
  if (n < 0 || !o || res < 1) return fail("bk_cell_keys: bad argument");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_cell_keys, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, x, n, o[0], o[1], o[2], dx, res, tiled,
                     keys, flags);
  BK_LAUNCH(k_cell_keys);
  return 0;
}

int bk_sort_pairs(bk_workspace* ws, const uint32_t* keys_in, const int32_t* values_in, uint32_t* keys_out, int32_t* values_out,
                  int32_t n, int32_t key_bits, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw) return -2;
  if (n < 0 || n > hw->dev.max_n || key_bits < 1 || key_bits > 32) return fail("bk_sort_pairs: bad argument");
  return radix_sort(&hw->dev, keys_in, values_in, keys_out, values_out, n, key_bits, (hipStream_t)stream);
}

int bk_batch_substep(bk_workspace* ws, const bk_scene* scene, const bk_state* in, bk_state* out, bk_grid* grid, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw) return fail("bk_batch_substep: bad workspace");
  return batch_substep(ws, hw, scene, in, out, grid, (hipStream_t)stream, true);
}

bk_graph* bk_batch_graph_create(bk_workspace* ws, const bk_scene* scene, const bk_state* in, bk_state* out, bk_grid* grid) {
  HostWs* hw = host_view(ws);
  if (!hw) { fail("bk_batch_graph_create: bad workspace"); return nullptr; }
  return capture(ws, hw, scene, in, out, grid);
}

bk_graph* bk_batch_graph_create_pipelined(bk_workspace* ws, const bk_scene* scene, const bk_state* in, bk_state* out,
                                          bk_grid* grid) {
  HostWs* hw = host_view(ws);
  if (!hw) { fail("bk_batch_graph_create_pipelined: bad workspace"); return nullptr; }
  return capture(ws, hw, scene, in, out, grid, true);
}

bk_graph* bk_batch_graph_create_steps(bk_workspace* ws, const bk_scene* scene, const bk_state* a, bk_state* b, bk_grid* grid,
                                      int32_t steps, int32_t pipelined) {
  HostWs* hw = host_view(ws);
  if (!hw || steps < 1 || steps > kMaxGraphSteps) { fail("bk_batch_graph_create_steps: bad workspace or steps"); return nullptr; }
  return capture(ws, hw, scene, a, b, grid, pipelined != 0, steps);
}

// The deferred pass a run of pipelined substeps left (a no-op when none is pending), on the last substep's in / out.
static int finish(bk_workspace* ws, HostWs* hw, const bk_scene* sc, const bk_state* in, bk_state* out, hipStream_t st) {
  if (!sc || !in || !out) return fail("bk_*_finish: null argument");
  if (!hw->gas) return 0;
  const bk_workspace* h = &hw->dev;
  hipLaunchKernelGGL(k_pass_c_deferred, dim3(clampi(div_up(hw->n, 256), 1, 1024)), dim3(256), 0, st, h->doughs, *sc, *in, *out, ws,
                     h->flags, 1);
  BK_LAUNCH(pass_c_deferred);
  hipLaunchKernelGGL(k_pending_done, dim3(1), dim3(64), 0, st, ws);
  BK_LAUNCH(pending_done);
  return 0;
}

int bk_batch_finish(bk_workspace* ws, const bk_scene* scene, const bk_state* in, bk_state* out, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw) return fail("bk_batch_finish: bad workspace");
  if (hw->doughs.empty()) return fail("bk_batch_finish: no dough table uploaded");
  return finish(ws, hw, scene, in, out, (hipStream_t)stream);
}

int bk_substep(const bk_params* P, const bk_scene* sc, const bk_state* in, bk_state* out, bk_grid* g, bk_workspace* ws,
               void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !P) return fail("bk_substep: bad argument");
  bk_dough d = single_dough(P);
  int rc = upload_doughs(ws, hw, &d, 1);
  if (rc) return rc;
  return batch_substep(ws, hw, sc, in, out, g, (hipStream_t)stream, true);
}

bk_graph* bk_substep_graph_create(const bk_params* P, const bk_scene* sc, const bk_state* in, bk_state* out, bk_grid* g,
                                  bk_workspace* ws) {
  HostWs* hw = host_view(ws);
  if (!hw || !P) { fail("bk_substep_graph_create: bad argument"); return nullptr; }
  bk_dough d = single_dough(P);
  if (upload_doughs(ws, hw, &d, 1)) return nullptr;
  return capture(ws, hw, sc, in, out, g);
}

bk_graph* bk_substep_graph_create_pipelined(const bk_params* P, const bk_scene* sc, const bk_state* in, bk_state* out, bk_grid* g,
                                            bk_workspace* ws) {
  HostWs* hw = host_view(ws);
  if (!hw || !P) { fail("bk_substep_graph_create_pipelined: bad argument"); return nullptr; }
  bk_dough d = single_dough(P);
  if (upload_doughs(ws, hw, &d, 1)) return nullptr;
  return capture(ws, hw, sc, in, out, g, true);
}

bk_graph* bk_substep_graph_create_steps(const bk_params* P, const bk_scene* sc, const bk_state* a, bk_state* b, bk_grid* g,
                                        bk_workspace* ws, int32_t steps, int32_t pipelined) {
  HostWs* hw = host_view(ws);
  if (!hw || !P || steps < 1 || steps > kMaxGraphSteps) { fail("bk_substep_graph_create_steps: bad argument"); return nullptr; }
  bk_dough d = single_dough(P);
  if (upload_doughs(ws, hw, &d, 1)) return nullptr;
  return capture(ws, hw, sc, a, b, g, pipelined != 0, steps);
}

int bk_substep_finish(const bk_params* P, const bk_scene* sc, const bk_state* in, bk_state* out, bk_workspace* ws, void* stream) {
  HostWs* hw = host_view(ws);
  if (!hw || !P) return fail("bk_substep_finish: bad argument");
  bk_dough d = single_dough(P);
  int rc = upload_doughs(ws, hw, &d, 1);
  if (rc) return rc;
  return finish(ws, hw, sc, in, out, (hipStream_t)stream);
}

int bk_graph_launch(bk_graph* gr, void* stream) {
  if (!gr) return fail("bk_graph_launch: null graph");
  HostWs* hw = host_view(gr->ws);
  if (!hw) return fail("bk_graph_launch: the graph's workspace is gone");
  int rc = upload_doughs(gr->ws, hw, gr->table.data(), int(gr->table.size()));  // no-op unless another table was uploaded since
  if (rc) return rc;
  BK_HIP(hipGraphLaunch(gr->exec, (hipStream_t)stream));
  return 0;
}

void bk_graph_destroy(bk_graph* gr) {
  if (!gr) return;
  if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
  if (gr->graph) (void)hipGraphDestroy(gr->graph);
  delete gr;
}

int bk_measure(const bk_scene* scene, const bk_state* state, int32_t n, int32_t n_doughs, double* out, void* stream) {
  if (!scene || !state || !out || n < 0 || n_doughs < 1 || !scene->V0 || !state->F) return fail("bk_measure: bad argument");
  hipStream_t st = (hipStream_t)stream;
  BK_HIP(hipMemsetAsync(out, 0, sizeof(double) * kMeasureSums * size_t(n_doughs), st));
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_measure, dim3(div_up(n, 256)), dim3(256), 0, st, *scene, *state, n, out);
  BK_LAUNCH(measure);
  return 0;
}

int bk_grid_untile(const float* tiled, float* dense, int32_t res, int32_t comps, void* stream) {
  long long total = (long long)res * res * res;
  hipLaunchKernelGGL(k_untile_f32, dim3(div_up(total, 256)), dim3(256), 0, (hipStream_t)stream, tiled, dense, res, comps);
  BK_LAUNCH(untile);
  return 0;
}
int bk_grid_untile_f64(const double* tiled, double* dense, int32_t res, void* stream) {
  long long total = (long long)res * res * res;
  hipLaunchKernelGGL(k_untile_f64, dim3(div_up(total, 256)), dim3(256), 0, (hipStream_t)stream, tiled, dense, res);
  BK_LAUNCH(untile);
  return 0;
}

}  // extern "C"
