// CheckState on the device: every statistic and health test of
// MpmSimulator.CheckState (reference physics/mpm.py:243-308, with
// DingMaterial.GasWaveSpeed, physics/materials/ding.py:285-292) in one pass,
// so a check after Advance copies BK_CHECK_COUNT doubles to the host instead
// of the particle state.
//
//   k_check_partials  one launch over four domains: particles, the dough's
//                     tiled node masses, mesh vertices and tets.  Each
//                     workgroup reduces its share and stores one partial of
//                     BK_CHECK_COUNT doubles (plain stores, no atomics).
//   k_check_finish    one workgroup: combines the partials in a fixed order
//                     (the sums are deterministic) into out[BK_CHECK_COUNT].
//
// Precision follows the reference's numpy: det F and the speed in float, the
// stencil test, the sums, the gas and water wave terms in double.  A NaN wave
// term is skipped, as Python's max(wave, nan) keeps wave.
#include "bk_device.h"
#include "../../include/baking_abi.h"

using namespace bk;

void bk_internal_set_error(const char* what);  // bk_last_error's text (bk_substep.hip)

namespace {
constexpr int kCheckThreads = 256;
constexpr int kCheckWaves = kCheckThreads / 64;
constexpr int kMaxParticleGroups = 2048;
constexpr int kMaxNodeGroups = 1024;
constexpr int kMaxMeshGroups = 64;
inline int div_up(long long a, long long b) { return int((a + b - 1) / b); }
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

enum Op { kMax, kMin, kSum };
__host__ __device__ constexpr Op op_of(int k) {
  return k == BK_CHECK_MIN_J ? kMin
         : (k == BK_CHECK_INVERTED_TETS || k == BK_CHECK_VOLUME || k == BK_CHECK_MASS || k == BK_CHECK_GRID_MASS ||
            k == BK_CHECK_KINETIC_ENERGY)
             ? kSum
             : kMax;
}
__host__ __device__ constexpr bool is_flag(int k) {
  return k == BK_CHECK_NONFINITE_STATE || k == BK_CHECK_OUTSIDE_STENCIL || k == BK_CHECK_NONFINITE_DING ||
         k == BK_CHECK_NONFINITE_VERTICES;
}
// Flags (max of 0/1) start at 0; other maxima at -inf, minima at +inf, sums at 0.
__device__ __forceinline__ double identity_of(int k) {
  return (op_of(k) == kSum || is_flag(k)) ? 0.0 : (op_of(k) == kMin ? INFINITY : -INFINITY);
}
__device__ __forceinline__ double combine(int k, double a, double b) {
  switch (op_of(k)) {
    case kSum: return a + b;
    case kMin: return fmin(a, b);
    default: return fmax(a, b);
  }
}

struct Groups {
  int particles, nodes, vertices, tets;
};
Groups groups_for(int32_t n, int64_t nodes, int32_t nv, int32_t nt) {
  Groups g;
  g.particles = clampi(div_up(n, kCheckThreads), 1, kMaxParticleGroups);
  g.nodes = nodes > 0 ? clampi(div_up(nodes, 4ll * kCheckThreads), 1, kMaxNodeGroups) : 0;
  g.vertices = nv > 0 ? clampi(div_up(nv, kCheckThreads), 1, kMaxMeshGroups) : 0;
  g.tets = nt > 0 ? clampi(div_up(nt, kCheckThreads), 1, kMaxMeshGroups) : 0;
  return g;
}

struct CheckArgs {
  bk_params P;
  double origin[3], dx;          // the simulation config's grid origin and spacing, as the reference's numpy uses them
  const float *x, *v, *F, *C, *V0, *masses;
  const float *Fe, *Fp, *residual, *porosity, *density;
  const double *pressure, *saturation;
  const float* node_masses;
  int64_t node_count;
  const float* vertices;
  const int32_t* tets;
  int32_t n, n_vertices, n_tets;
  Groups groups;
};
}  // namespace

__device__ __forceinline__ bool all_finite(const float* a, int count) {
  bool ok = true;
  for (int i = 0; i < count; ++i) ok &= isfinite(a[i]);
  return ok;
}

__device__ __forceinline__ float det3f(const float* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}
__device__ __forceinline__ double det3d(const double* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// Particle p's contribution (physics/mpm.py:249-283, 292-304).
__device__ __forceinline__ void particle_terms(const CheckArgs& a, int p, double acc[BK_CHECK_COUNT]) {
  float x[3], v[3], F[9], C[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    x[i] = a.x[3 * p + i];
    v[i] = a.v[3 * p + i];
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    F[i] = a.F[9 * p + i];
    C[i] = a.C[9 * p + i];
  }
  const bool finite = all_finite(x, 3) && all_finite(v, 3) && all_finite(F, 9) && all_finite(C, 9);
  if (!finite) acc[BK_CHECK_NONFINITE_STATE] = 1.0;
  const float J = det3f(F);
  acc[BK_CHECK_MIN_J] = fmin(acc[BK_CHECK_MIN_J], double(J));
  acc[BK_CHECK_MAX_J] = fmax(acc[BK_CHECK_MAX_J], double(J));
  const double hi = double(a.P.res) - 1.5;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double g = (double(x[i]) - a.origin[i]) / a.dx;
    if (g < 0.5 || g >= hi) acc[BK_CHECK_OUTSIDE_STENCIL] = 1.0;
  }
  const float v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  acc[BK_CHECK_MAX_SPEED] = fmax(acc[BK_CHECK_MAX_SPEED], double(sqrtf(v2)));
  const double m = double(a.masses[p]);
  acc[BK_CHECK_MASS] += m;
  acc[BK_CHECK_KINETIC_ENERGY] += m * double(v2);  // halved in the finish
  acc[BK_CHECK_VOLUME] += double(a.V0[p]) * double(J);
  if (a.P.model == 1) {
    if (a.Fe && a.Fp && !(all_finite(a.Fe + 9 * p, 9) && all_finite(a.Fp + 9 * p, 9))) acc[BK_CHECK_NONFINITE_DING] = 1.0;
    if (a.residual) acc[BK_CHECK_MAX_RESIDUAL] = fmax(acc[BK_CHECK_MAX_RESIDUAL], double(a.residual[p]));
    const double rho = a.density ? double(a.density[p]) : 0.0;
    if (a.P.gas_enabled && a.pressure && a.porosity && a.density) {  // ding.py:285-292
      double Fd[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) Fd[i] = double(F[i]);
      const double n = double(a.porosity[p]);
      const double bulk = a.pressure[p] * (1.0 + (1.0 - double(a.P.n0) + n) / n);
      const double modulus = double(a.P.lam) + 2.0 * double(a.P.mu);
      const double w = sqrt((modulus + bulk) * det3d(Fd) / rho);
      if (!isnan(w)) acc[BK_CHECK_GAS_WAVE] = fmax(acc[BK_CHECK_GAS_WAVE], w);
    }
    if (a.P.transport_enabled && a.P.water_enabled && a.saturation && a.density) {  // mpm.py:271-275
      const double w2 = double(a.P.water_stress) * a.saturation[p] * double(J) / rho;
      if (!isnan(w2)) acc[BK_CHECK_WATER_WAVE_SQ] = fmax(acc[BK_CHECK_WATER_WAVE_SQ], w2);
    }
  }
}

// The workgroup's reduction of acc into thread 0 (every thread calls it).
__device__ __forceinline__ void block_reduce(double acc[BK_CHECK_COUNT]) {
  __shared__ double part[BK_CHECK_COUNT][kCheckWaves];
#pragma unroll
  for (int k = 0; k < BK_CHECK_COUNT; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[k] = combine(k, acc[k], __shfl_xor(acc[k], o));
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < BK_CHECK_COUNT; ++k) part[k][wave] = acc[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < BK_CHECK_COUNT; ++k) {
      double r = part[k][0];
      for (int w = 1; w < kCheckWaves; ++w) r = combine(k, r, part[k][w]);
      acc[k] = r;
    }
}

__global__ __launch_bounds__(kCheckThreads) void k_check_partials(CheckArgs a, double* __restrict__ partials) {
  double acc[BK_CHECK_COUNT];
#pragma unroll
  for (int k = 0; k < BK_CHECK_COUNT; ++k) acc[k] = identity_of(k);
  const Groups& g = a.groups;
  int b = blockIdx.x;
  if (b < g.particles) {
    for (int p = b * kCheckThreads + threadIdx.x; p < a.n; p += g.particles * kCheckThreads) particle_terms(a, p, acc);
  } else if ((b -= g.particles) < g.nodes) {  // mpm.py:297 grid_mass
    for (int64_t i = int64_t(b) * kCheckThreads + threadIdx.x; i < a.node_count; i += int64_t(g.nodes) * kCheckThreads)
      acc[BK_CHECK_GRID_MASS] += double(a.node_masses[i]);
  } else if ((b -= g.nodes) < g.vertices) {  // mpm.py:286-288
    for (int i = b * kCheckThreads + threadIdx.x; i < a.n_vertices; i += g.vertices * kCheckThreads)
      if (!all_finite(a.vertices + 3 * i, 3)) acc[BK_CHECK_NONFINITE_VERTICES] = 1.0;
  } else {  // mpm.py:306-307 inverted_mesh_tets: det([v1-v0, v2-v0, v3-v0]) / 6 <= 0
    b -= g.vertices;
    for (int t = b * kCheckThreads + threadIdx.x; t < a.n_tets; t += g.tets * kCheckThreads) {
      const int32_t* q = a.tets + 4 * t;
      double e[9];
      const float* v0 = a.vertices + 3 * q[0];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const float* vr = a.vertices + 3 * q[r + 1];
#pragma unroll
        for (int c = 0; c < 3; ++c) e[3 * r + c] = double(vr[c]) - double(v0[c]);
      }
      if (det3d(e) / 6.0 <= 0.0) acc[BK_CHECK_INVERTED_TETS] += 1.0;
    }
  }
  block_reduce(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < BK_CHECK_COUNT; ++k) partials[size_t(blockIdx.x) * BK_CHECK_COUNT + k] = acc[k];
}

// Partials combined per thread in index order, then across the workgroup in
// a fixed tree: the same sums on every run.
__global__ __launch_bounds__(kCheckThreads) void k_check_finish(const double* __restrict__ partials, int count,
                                                               double* __restrict__ out) {
  double acc[BK_CHECK_COUNT];
#pragma unroll
  for (int k = 0; k < BK_CHECK_COUNT; ++k) acc[k] = identity_of(k);
  for (int i = threadIdx.x; i < count; i += kCheckThreads)
#pragma unroll
    for (int k = 0; k < BK_CHECK_COUNT; ++k) acc[k] = combine(k, acc[k], partials[size_t(i) * BK_CHECK_COUNT + k]);
  block_reduce(acc);
  if (threadIdx.x == 0) {
    acc[BK_CHECK_KINETIC_ENERGY] *= 0.5;
#pragma unroll
    for (int k = 0; k < BK_CHECK_COUNT; ++k) out[k] = acc[k];
  }
}

extern "C" {

int64_t bk_check_scratch_size(int32_t n, int64_t node_count, int32_t n_vertices, int32_t n_tets) {
  const Groups g = groups_for(n, node_count, n_vertices, n_tets);
  return int64_t(g.particles + g.nodes + g.vertices + g.tets) * BK_CHECK_COUNT;
}

int bk_check_state(const bk_params* P, const bk_scene* scene, const bk_state* state, const double* grid_origin_dx,
                   const float* node_masses, int64_t node_count, const float* vertices, int32_t n_vertices, const int32_t* tets,
                   int32_t n_tets, double* scratch, double* out, void* stream) {
  if (!P || !scene || !state || !grid_origin_dx || !scratch || !out) {
    bk_internal_set_error("bk_check_state: null argument");
    return -1;
  }
  if (P->n < 1 || !state->x || !state->v || !state->F || !state->C || !scene->V0 ||
      (P->transport_enabled ? !state->masses : !scene->m0) || node_count < 0 || (node_count > 0 && !node_masses) ||
      n_vertices < 0 || n_tets < 0 || ((n_vertices > 0 || n_tets > 0) && !vertices) || (n_tets > 0 && !tets) ||
      !(grid_origin_dx[3] > 0.0)) {
    bk_internal_set_error("bk_check_state: bad sizes or missing arrays");
    return -1;
  }
  CheckArgs a{};
  a.P = *P;
  for (int i = 0; i < 3; ++i) a.origin[i] = grid_origin_dx[i];
  a.dx = grid_origin_dx[3];
  a.x = state->x;
  a.v = state->v;
  a.F = state->F;
  a.C = state->C;
  a.V0 = scene->V0;
  a.masses = P->transport_enabled ? state->masses : scene->m0;
  a.Fe = state->Fe;
  a.Fp = state->Fp;
  a.residual = state->residual;
  a.porosity = state->porosity;
  a.density = state->density;
  a.pressure = state->pressure;
  a.saturation = state->s;
  a.node_masses = node_masses;
  a.node_count = node_count;
  a.vertices = vertices;
  a.tets = tets;
  a.n = P->n;
  a.n_vertices = n_vertices;
  a.n_tets = n_tets;
  a.groups = groups_for(P->n, node_count, n_vertices, n_tets);
  const int total = a.groups.particles + a.groups.nodes + a.groups.vertices + a.groups.tets;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_check_partials, dim3(total), dim3(kCheckThreads), 0, st, a, scratch);
  hipLaunchKernelGGL(k_check_finish, dim3(1), dim3(kCheckThreads), 0, st, scratch, total, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    bk_internal_set_error(hipGetErrorString(e));
    return -2;
  }
  return 0;
}

}  // extern "C"
