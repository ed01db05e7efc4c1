// Settling controller for gfx950: the rest before baking without a host round
// trip per substep (reference physics/quasistatic.py:11-21 SettlingMetrics and
// physics/mpm.py:77-109 UpdateSettling).
//
// The reference launches SettlingMetrics after every resting substep, copies
// its three maxima to the host and decides there whether the dough is quiet,
// whether to start baking or whether to give up.  Here the decision runs on
// the device, on a per-dough record (bk_settling) that the substep kernels
// read to choose resting or baking transport:
//   k_settling_update  per 256 particles of a dough (blockIdx.y = dough): the
//                      workgroup's three maxima, stored to a partials array
//   k_settling_finish  one workgroup per dough: the maxima over its partials,
//                      then the record update of UpdateSettling
// When a substep closes a plastic-rate window, the reference copies the
// plastic gradients it produced (mpm.py:95).  Those are the next substep's
// input, so that substep's k_settling_update takes them as its reference and
// stores them, from the copy it stages in LDS anyway, in the record's new
// half of the reference buffer: no copy kernel.
// No atomics: maxima of thousands of workgroups taken with same-address
// atomics serialise in L2 (~14 ns each, 40-160 us per substep on a 0.7M
// particle dough).  A NaN metric counts as +inf, so a state with a NaN speed
// or rate is never quiet (the reference's CPU max keeps or drops a NaN
// depending on the particles after it).
#include "bk_device.h"
#include "bk_workspace.h"
#include "../../include/baking_abi.h"

using namespace bk;

void bk_internal_set_error(const char* what);  // bk_last_error's text (bk_substep.hip)

namespace {
constexpr int kSettleThreads = 256;
constexpr int kSettleWaves = kSettleThreads / 64;
inline int div_up(long long a, long long b) { return int((a + b - 1) / b); }
}  // namespace

// physics/quasistatic.py:15-21 for one particle (its velocity, affine
// velocity, reference and new plastic gradient): |v|, |sym C|_F and
// |Fp - Fp_ref|_F / (elapsed |Fp_ref|_F), all in float like the reference.
__device__ __forceinline__ void particle_metrics(const float* v, const float* C, const float* old_p, const float* new_p,
                                                 float elapsed, float m[3]) {
  const float vx = v[0], vy = v[1], vz = v[2];
  m[0] = sqrtf(vx * vx + vy * vy + vz * vz);
  float a[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = C[i];
  float ss = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float s = 0.5f * (a[3 * i + j] + a[3 * j + i]);
      ss += s * s;
    }
  m[1] = sqrtf(ss);
  m[2] = 0.0f;
  if (old_p && new_p) {
    float cc = 0.0f, oo = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float o = old_p[i];
      const float c = new_p[i] - o;
      cc += c * c;
      oo += o * o;
    }
    m[2] = sqrtf(cc) / (elapsed * sqrtf(oo));
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (isnan(m[k])) m[k] = INFINITY;
}

// Metrics of particles [first, first + count), count <= kSettleThreads, one per
// thread.  The workgroup first copies their v, C and plastic gradients into LDS
// with coalesced loads (a thread reading its own 3x3 matrices would issue
// 36-byte-strided loads, each wave touching 18 cache lines per load); with
// `keep_old` it also stores the reference gradients it staged there.  Threads
// without a particle return zeros.  Contains barriers: every thread of the
// workgroup calls it.
__device__ __forceinline__ void staged_metrics(const float* __restrict__ v, const float* __restrict__ C,
                                               const float* __restrict__ old_p, const float* __restrict__ new_p, float elapsed,
                                               int64_t first, int count, float m[3], float* __restrict__ keep_old = nullptr) {
  __shared__ float sv[3 * kSettleThreads], sc[9 * kSettleThreads], so[9 * kSettleThreads], sn[9 * kSettleThreads];
  const bool plastic = old_p && new_p;
  const int nk = 9 * count;
  float rc[9], ro[9], rn[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {  // every load of the workgroup in flight at once
    const int k = threadIdx.x + j * kSettleThreads;
    if (k < nk) {
      rc[j] = C[9 * first + k];
      if (plastic) {
        ro[j] = old_p[9 * first + k];
        rn[j] = new_p[9 * first + k];
      }
    }
  }
  float rv[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int k = threadIdx.x + j * kSettleThreads;
    if (k < 3 * count) rv[j] = v[3 * first + k];
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const int k = threadIdx.x + j * kSettleThreads;
    if (k < nk) {
      sc[k] = rc[j];
      if (plastic) {
        so[k] = ro[j];
        sn[k] = rn[j];
        if (keep_old) keep_old[9 * first + k] = ro[j];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int k = threadIdx.x + j * kSettleThreads;
    if (k < 3 * count) sv[k] = rv[j];
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < count) {
    particle_metrics(sv + 3 * t, sc + 9 * t, plastic ? so + 9 * t : nullptr, plastic ? sn + 9 * t : nullptr, elapsed, m);
  } else {
    m[0] = m[1] = m[2] = 0.0f;
  }
}

// Maxima of the workgroup's three metrics, in thread 0.
__device__ __forceinline__ void block_max3(float m[3]) {
  __shared__ float part[3][kSettleWaves];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m[k] = fmaxf(m[k], __shfl_xor(m[k], o));
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) part[k][wave] = m[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float r = part[k][0];
      for (int w = 1; w < kSettleWaves; ++w) r = fmaxf(r, part[k][w]);
      m[k] = r;
    }
}

// physics/mpm.py:78-109 after the substep that brought the step count to
// `steps`, with this substep's metric maxima (auto mode).
__device__ void advance_record(bk_settling* S, int64_t steps, const float maxima[3], bool auto_mode) {
  S->step_count = steps;
  S->copy_pending = 0;
  if (!auto_mode) {  // mpm.py:80-84
    if (steps >= S->settling_steps) {
      S->baking_start_step = steps;
      S->reason = 1;
      S->phase = 0;
    }
    return;
  }
  const double dt = S->time_step;
  float values[3] = {maxima[0], maxima[1], maxima[2]};
  // mpm.py:93-97: a plastic-rate window closes after min(0.01 s, hold) of rest
  const bool close_window = double(steps - S->reference_step) * dt + 1e-12 >= fmin(0.01, S->hold_duration_s);
  if (close_window) {
    S->plastic_rate = double(values[2]);
    S->reference_step = steps;
  }
  values[2] = float(S->plastic_rate);
  S->diagnostics[0] = values[0];
  S->diagnostics[1] = values[1];
  S->diagnostics[2] = values[2];
  // mpm.py:100-109
  const bool eligible = double(steps - 1) * dt + 1e-12 >= S->min_duration_s;
  const bool small = double(values[0]) <= S->speed_tolerance && double(fmaxf(values[1], values[2])) <= S->rate_tolerance;
  S->quiet_steps = (eligible && small) ? S->quiet_steps + 1 : 0;
  if (double(S->quiet_steps) * dt + 1e-12 >= S->hold_duration_s) {
    S->baking_start_step = steps;
    S->reason = 2;
    S->phase = 0;
    return;  // the reference no longer reads the window's copy once baking has started
  }
  if (steps >= S->settling_steps) S->reason = 3;
  if (close_window) {  // the new reference goes to the other half; the next substep's update fills it
    S->ref_slot ^= 1;
    S->copy_pending = 1;
  }
}

// Metric maxima of one resting substep, per run of 256 particles of every
// dough whose params.settling is 2, whose record is in phase 1 and in auto
// mode: workgroup x of dough y stores those of particles [256 x, 256 x + 256)
// to part[(y * gridDim.x + x) * 3 + k].  Fp_in is the substep's input plastic
// gradients, the reference after a closed window.
__global__ __launch_bounds__(kSettleThreads) void k_settling_update(const bk_dough* __restrict__ D, const bk_workspace* ws,
                                                                   const float* __restrict__ v, const float* __restrict__ C,
                                                                   const float* __restrict__ Fp_in,
                                                                   const float* __restrict__ Fp) {
  const int d = blockIdx.y;
  if (D[d].params.settling != 2) return;
  const bk_settling* S = ws->settle + d;
  if (S->phase == 0 || !S->auto_mode) return;
  const int64_t steps = S->step_count + 1;
  float* ref = ws->settle_ref ? ws->settle_ref + size_t(S->ref_slot) * size_t(ws->max_n) * 9 : nullptr;
  const bool renew = ref && S->copy_pending && Fp_in;  // the previous substep closed a window
  const float elapsed = float(double(steps - S->reference_step) * S->time_step);
  const int i0 = blockIdx.x * kSettleThreads;
  const int left = D[d].params.n - i0;
  const int count = left < 0 ? 0 : (left > kSettleThreads ? kSettleThreads : left);
  float m[3];
  staged_metrics(v, C, renew ? Fp_in : ref, ref ? Fp : nullptr, elapsed, int64_t(D[d].particle_offset) + i0, count, m,
                 renew ? ref : nullptr);
  block_max3(m);
  if (threadIdx.x == 0) {
    float* out = ws->settle_part + (size_t(d) * gridDim.x + blockIdx.x) * 3;
    out[0] = m[0];
    out[1] = m[1];
    out[2] = m[2];
  }
}

// Workgroup d: dough d's maxima over its `runs` partials (auto mode), then its
// record update (physics/mpm.py:78-109).
__global__ __launch_bounds__(kSettleThreads) void k_settling_finish(const bk_dough* __restrict__ D, bk_workspace* ws, int runs) {
  const int d = blockIdx.x;
  if (D[d].params.settling != 2) return;
  bk_settling* S = ws->settle + d;
  if (S->phase == 0) return;
  const bool auto_mode = S->auto_mode != 0;
  float m[3] = {0.0f, 0.0f, 0.0f};
  if (auto_mode) {
    const float* part = ws->settle_part + size_t(d) * runs * 3;
    constexpr int kUnroll = 8;  // independent loads in flight per thread
    for (int r0 = threadIdx.x; r0 < runs; r0 += kUnroll * kSettleThreads) {
      float q[kUnroll][3];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int r = r0 + u * kSettleThreads;
#pragma unroll
        for (int k = 0; k < 3; ++k) q[u][k] = r < runs ? part[3 * r + k] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
#pragma unroll
        for (int k = 0; k < 3; ++k) m[k] = fmaxf(m[k], q[u][k]);
    }
  }
  block_max3(m);
  if (threadIdx.x == 0) advance_record(S, S->step_count + 1, m, auto_mode);
}

// Enqueue the controller after a substep with input plastic gradients Fp_in
// and output state (v, C, Fp);
// the workspace's partials array holds n_doughs * ceil(max_dough_n / 256) * 3
// floats.
int bk_internal_settling_launch(bk_workspace* ws, const bk_dough* doughs, int n_doughs, int max_dough_n, const float* v,
                                const float* C, const float* Fp_in, const float* Fp, hipStream_t st) {
  if (n_doughs < 1 || n_doughs > 65535 || max_dough_n < 1) {
    bk_internal_set_error("settling controller: bad dough table");
    return -1;
  }
  const int runs = div_up(max_dough_n, kSettleThreads);
  hipLaunchKernelGGL(k_settling_update, dim3(runs, n_doughs), dim3(kSettleThreads), 0, st, doughs, ws, v, C, Fp_in, Fp);
  hipLaunchKernelGGL(k_settling_finish, dim3(n_doughs), dim3(kSettleThreads), 0, st, doughs, ws, runs);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    bk_internal_set_error(hipGetErrorString(e));
    return -2;
  }
  return 0;
}

// physics/quasistatic.py:11-21 as a kernel of its own: running maxima into
// metrics[3] (non-negative starting values; the reference zeroes them).
__global__ __launch_bounds__(kSettleThreads) void k_SettlingMetrics(const float* __restrict__ v, const float* __restrict__ C,
                                                                   const float* __restrict__ old_p,
                                                                   const float* __restrict__ new_p, unsigned int* metrics,
                                                                   float time_step, int n) {
  const int i0 = blockIdx.x * kSettleThreads;
  const int count = n - i0 > kSettleThreads ? kSettleThreads : n - i0;
  float m[3];
  staged_metrics(v, C, old_p, new_p, time_step, i0, count, m);
  block_max3(m);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) atomicMax(metrics + k, __float_as_uint(m[k]));
}

extern "C" int bk_SettlingMetrics(const float* velocities, const float* affine, const float* old_plastic, const float* new_plastic,
                                  float* metrics, float time_step, int32_t n, void* stream) {
  if (n < 0 || !metrics || (n > 0 && (!velocities || !affine || !old_plastic || !new_plastic))) {
    bk_internal_set_error("bk_SettlingMetrics: bad argument");
    return -1;
  }
  if (n == 0) return 0;
  const int blocks = div_up(n, kSettleThreads);
  hipLaunchKernelGGL(k_SettlingMetrics, dim3(blocks), dim3(kSettleThreads), 0, (hipStream_t)stream, velocities, affine, old_plastic,
                     new_plastic, reinterpret_cast<unsigned int*>(metrics), time_step, n);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    bk_internal_set_error(hipGetErrorString(e));
    return -2;
  }
  return 0;
}
