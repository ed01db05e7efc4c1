// Device-side building blocks of the MLS-MPM baking substep (gfx950).
//
// Everything here is __device__ inline code shared by the fused substep
// kernels and the per-kernel entry points.  Types follow the reference:
// its Warp `float` is float here, its wp.float64 is double, and every explicit
// float()/wp.float64() conversion of the reference is kept so results agree
// with the reference's arithmetic to rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bk {

// ------------------------------------------------------------------ 3x3 matrices (row-major)
template <typename T> struct mat3 { T a[9]; };

template <typename T> __device__ __forceinline__ mat3<T> m_load(const T* p) {
  mat3<T> m;
#pragma unroll
  for (int i = 0; i < 9; ++i) m.a[i] = p[i];
  return m;
}
template <typename T> __device__ __forceinline__ void m_store(T* p, const mat3<T>& m) {
#pragma unroll
  for (int i = 0; i < 9; ++i) p[i] = m.a[i];
}
template <typename T> __device__ __forceinline__ mat3<T> m_eye() {
  mat3<T> m;
#pragma unroll
  for (int i = 0; i < 9; ++i) m.a[i] = (i % 4 == 0) ? T(1) : T(0);
  return m;
}
template <typename T> __device__ __forceinline__ mat3<T> m_zero() {
  mat3<T> m;
#pragma unroll
  for (int i = 0; i < 9; ++i) m.a[i] = T(0);
  return m;
}
template <typename T> __device__ __forceinline__ mat3<T> m_mul(const mat3<T>& x, const mat3<T>& y) {
  mat3<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = x.a[i * 3 + 0] * y.a[j] + x.a[i * 3 + 1] * y.a[3 + j] + x.a[i * 3 + 2] * y.a[6 + j];
  return r;
}
template <typename T> __device__ __forceinline__ mat3<T> m_mul_bt(const mat3<T>& x, const mat3<T>& y) {  // x * y^T
  mat3<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = x.a[i * 3 + 0] * y.a[j * 3 + 0] + x.a[i * 3 + 1] * y.a[j * 3 + 1] + x.a[i * 3 + 2] * y.a[j * 3 + 2];
  return r;
}
template <typename T> __device__ __forceinline__ mat3<T> m_scale(T s, const mat3<T>& m) {
  mat3<T> r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.a[i] = s * m.a[i];
  return r;
}
template <typename T> __device__ __forceinline__ T m_det(const mat3<T>& m) {
  const T* a = m.a;
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}
template <typename T> __device__ __forceinline__ mat3<T> m_inverse(const mat3<T>& m) {
  const T* a = m.a;
  T d = m_det(m);
  mat3<T> r;
  r.a[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  r.a[1] = (a[2] * a[7] - a[1] * a[8]) / d;
  r.a[2] = (a[1] * a[5] - a[2] * a[4]) / d;
  r.a[3] = (a[5] * a[6] - a[3] * a[8]) / d;
  r.a[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  r.a[5] = (a[2] * a[3] - a[0] * a[5]) / d;
  r.a[6] = (a[3] * a[7] - a[4] * a[6]) / d;
  r.a[7] = (a[1] * a[6] - a[0] * a[7]) / d;
  r.a[8] = (a[0] * a[4] - a[1] * a[3]) / d;
  return r;
}
template <typename T, typename S> __device__ __forceinline__ mat3<T> m_cast(const mat3<S>& m) {
  mat3<T> r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.a[i] = T(m.a[i]);
  return r;
}

// ------------------------------------------------------------------ stencil (physics/transfer.py:4-33)
__device__ __forceinline__ float quadratic_weight(float l, int i) {
  if (i == 0) return 0.5f * (1.5f - l) * (1.5f - l);
  if (i == 1) return 0.75f - (l - 1.0f) * (l - 1.0f);
  return 0.5f * (l - 0.5f) * (l - 0.5f);
}
// All three weights of one axis.
__device__ __forceinline__ void axis_weights(float l, float w[3]) {
  w[0] = 0.5f * (1.5f - l) * (1.5f - l);
  w[1] = 0.75f - (l - 1.0f) * (l - 1.0f);
  w[2] = 0.5f * (l - 0.5f) * (l - 0.5f);
}
// grid_position = (x - origin) / dx with IEEE single-precision subtract and divide.
__device__ __forceinline__ void grid_position(const float* x, const float o[3], float dx, float g[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) g[a] = __fdiv_rn(__fsub_rn(x[a], o[a]), dx);
}
// base = int(floor(g - 1/2)).  A NaN or a value outside [-1, 2^30] (a particle
// far off the grid, e.g. after a blow-up that a flag reports) becomes -1, so
// it is never taken for a valid base and no index is formed from it.
__device__ __forceinline__ void stencil_base(const float g[3], int b[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float f = floorf(__fsub_rn(g[a], 0.5f));
    b[a] = (f >= -1.0f && f <= 1073741824.0f) ? int(f) : -1;
  }
}
// base >= 0 and base + 2 < res, written without the addition so no base can overflow it.
__device__ __forceinline__ bool valid_stencil(const int b[3], int res) {
  return b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[0] < res - 2 && b[1] < res - 2 && b[2] < res - 2;
}
__device__ __forceinline__ int64_t flatten_node(int i, int j, int k, int res) { return (int64_t(i) * res + j) * res + k; }

// Tiled (block-major) node numbering: 4x4x4 node blocks, nb blocks per axis.
__device__ __forceinline__ int64_t tiled_node(int i, int j, int k, int nb) {
  return ((int64_t((i >> 2) * nb + (j >> 2)) * nb + (k >> 2)) << 6) | ((i & 3) << 4) | ((j & 3) << 2) | (k & 3);
}
// The tiled number is a sum of one term per axis (block part times its stride plus the position inside the block),
// so the 27 nodes of a stencil at base b are x[o0] + y[o1] + z[o2], each term formed once per axis, in 32 bits
// (a dough's res^3 nodes stay below 2^30).  scale multiplies every term (3: float index into an xyz array).
struct StencilTerms {
  uint32_t x[3], y[3], z[3];
};
__device__ __forceinline__ StencilTerms stencil_terms(const int b[3], int nb, uint32_t scale = 1u) {
  StencilTerms t;
  const uint32_t sx = uint32_t(nb) * uint32_t(nb) * 64u, sy = uint32_t(nb) * 64u;
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    const uint32_t i = uint32_t(b[0] + o), j = uint32_t(b[1] + o), k = uint32_t(b[2] + o);
    t.x[o] = scale * ((i >> 2) * sx + ((i & 3u) << 4));
    t.y[o] = scale * ((j >> 2) * sy + ((j & 3u) << 2));
    t.z[o] = scale * (((k >> 2) << 6) + (k & 3u));
  }
  return t;
}

// ------------------------------------------------------------------ SVD (double)
// Symmetric cyclic Jacobi on A = F^T F (Golub & Van Loan, Matrix Computations,
// Alg. 8.5.3): V diagonalises A, sigma_i = sqrt(lambda_i) sorted descending,
// U's first two columns are the normalised F*v_i (Gram-Schmidt), the third is
// their cross product, and sigma_3 takes the sign of det F.  U and V are
// rotations and F = U diag(sigma) V^T.
__device__ __forceinline__ void jacobi_rotate(double A[3][3], double V[3][3], int p, int q) {
  double apq = A[p][q];
  if (apq == 0.0) return;
  // t = sgn(theta) / (|theta| + sqrt(1 + theta^2)) with theta = d / e, multiplied through by |e|: one division
  const double d = A[q][q] - A[p][p], e = 2.0 * apq;
  double t = (d * e >= 0.0 ? fabs(e) : -fabs(e)) / (fabs(d) + sqrt(d * d + e * e));
  double c = rsqrt(1.0 + t * t);
  double s = t * c;
  int r = 3 - p - q;
  double arp = A[r][p], arq = A[r][q];
  A[p][p] -= t * apq;
  A[q][q] += t * apq;
  A[p][q] = A[q][p] = 0.0;
  A[r][p] = A[p][r] = c * arp - s * arq;
  A[r][q] = A[q][r] = s * arp + c * arq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double vp = V[i][p], vq = V[i][q];
    V[i][p] = c * vp - s * vq;
    V[i][q] = s * vp + c * vq;
  }
}

// tol2: the sweeps stop once the squared off-diagonal of A is at most tol2 times its squared diagonal.
__device__ inline void svd3(const mat3<double>& F, mat3<double>& U, double sig[3], mat3<double>& Vout, double tol2 = 1e-34) {
  double A[3][3], V[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      A[i][j] = F.a[0 * 3 + i] * F.a[0 * 3 + j] + F.a[1 * 3 + i] * F.a[1 * 3 + j] + F.a[2 * 3 + i] * F.a[2 * 3 + j];
      V[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 10; ++sweep) {
    double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    double diag = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (off <= tol2 * diag) break;
    jacobi_rotate(A, V, 0, 1);
    jacobi_rotate(A, V, 0, 2);
    jacobi_rotate(A, V, 1, 2);
  }
  // sort descending (3-element network), swapping the eigenvalues with their columns of V
  double lam[3] = {A[0][0], A[1][1], A[2][2]};
  double Vs[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Vs[i][j] = V[i][j];
  auto order = [&](int p, int q) {
    const bool sw = lam[p] < lam[q];
    const double lp = lam[p], lq = lam[q];
    lam[p] = sw ? lq : lp;
    lam[q] = sw ? lp : lq;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double vp = Vs[i][p], vq = Vs[i][q];
      Vs[i][p] = sw ? vq : vp;
      Vs[i][q] = sw ? vp : vq;
    }
  };
  order(0, 1);
  order(1, 2);
  order(0, 1);
  double dv = Vs[0][0] * (Vs[1][1] * Vs[2][2] - Vs[1][2] * Vs[2][1]) - Vs[0][1] * (Vs[1][0] * Vs[2][2] - Vs[1][2] * Vs[2][0]) +
              Vs[0][2] * (Vs[1][0] * Vs[2][1] - Vs[1][1] * Vs[2][0]);
  if (dv < 0.0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) Vs[i][2] = -Vs[i][2];
  }
  // columns F*v_j
  double B[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) B[i][j] = F.a[i * 3 + 0] * Vs[0][j] + F.a[i * 3 + 1] * Vs[1][j] + F.a[i * 3 + 2] * Vs[2][j];
  double n0 = sqrt(B[0][0] * B[0][0] + B[1][0] * B[1][0] + B[2][0] * B[2][0]);
  double u0[3], u1[3], u2[3];
  if (n0 > 0.0) {
    const double r0 = 1.0 / n0;
#pragma unroll
    for (int i = 0; i < 3; ++i) u0[i] = B[i][0] * r0;
  } else {
    u0[0] = 1.0; u0[1] = 0.0; u0[2] = 0.0;
  }
  double d01 = u0[0] * B[0][1] + u0[1] * B[1][1] + u0[2] * B[2][1];
#pragma unroll
  for (int i = 0; i < 3; ++i) u1[i] = B[i][1] - d01 * u0[i];
  double n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  if (n1 > 1e-300) {
    const double r1 = 1.0 / n1;
#pragma unroll
    for (int i = 0; i < 3; ++i) u1[i] *= r1;
  } else {
    double e0 = fabs(u0[0]) < 0.9 ? 1.0 : 0.0, e1 = fabs(u0[0]) < 0.9 ? 0.0 : 1.0;
    double d = e0 * u0[0] + e1 * u0[1];
    u1[0] = e0 - d * u0[0]; u1[1] = e1 - d * u0[1]; u1[2] = -d * u0[2];
    double nn = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) u1[i] /= nn;
  }
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  sig[0] = n0;
  sig[1] = u1[0] * B[0][1] + u1[1] * B[1][1] + u1[2] * B[2][1];
  sig[2] = u2[0] * B[0][2] + u2[1] * B[1][2] + u2[2] * B[2][2];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    U.a[i * 3 + 0] = u0[i]; U.a[i * 3 + 1] = u1[i]; U.a[i * 3 + 2] = u2[i];
    Vout.a[i * 3 + 0] = Vs[i][0]; Vout.a[i * 3 + 1] = Vs[i][1]; Vout.a[i * 3 + 2] = Vs[i][2];
  }
}

// ------------------------------------------------------------------ Ding material (physics/materials/ding.py)
// ding.py:56-65
__device__ __forceinline__ float temperature_factor(float t, const float k[4], const float f[4]) {
  float r = f[0];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (t >= k[i]) {
      float frac = fminf((t - k[i]) / (k[i + 1] - k[i]), 1.0f);
      r = (1.0f - frac) * f[i] + frac * f[i + 1];
    }
  return r;
}
// ding.py:68-74
__device__ __forceinline__ float porosity_coefficient(float n) {
  float c = 0.0f;
  if (n > 0.0f) {
    float lg = logf(1.0f / n);
    c = 9.0f / (4.0f * (lg * lg) * (3.5f * n + 1.0f));
  }
  return c;
}
// ding.py:77-87 (32 fixed Newton steps on the scalar overstress equation)
__device__ __forceinline__ float solve_overstress(float q2, float p2, float df, float mf) {
  float root = 0.0f;
  if (q2 + p2 > 1.0f) {
    for (int it = 0; it < 32; ++it) {
      float dd = 1.0f + df * root;
      float md = 1.0f + mf * root;
      float dd2 = dd * dd, md2 = md * md;
      float res = root + 1.0f - q2 / dd2 - p2 / md2;
      float der = 1.0f + 2.0f * df * q2 / (dd2 * dd) + 2.0f * mf * p2 / (md2 * md);
      float next = root - res / der;
      // a fixed point of the iteration stays fixed, so stopping there returns
      // exactly what the remaining fixed steps would
      if (next == root) break;
      root = next;
    }
  }
  return root;
}
// porous_return.py:11-16
__device__ __forceinline__ double pressure_coefficient64(double n) {
  double c = 0.0;
  if (n > 0.0) {
    double lg = log(1.0 / n);
    c = 9.0 / (4.0 * (lg * lg) * (3.5 * n + 1.0));
  }
  return c;
}
// porous_return.py:19-31 (<= 48 Newton steps, stop when |r| < 1e-12)
__device__ __forceinline__ double overstress64(double q, double p, double df, double mf) {
  double root = 0.0;
  if (q + p > 1.0) {
    for (int it = 0; it < 48; ++it) {
      double u = 1.0 + df * root, v = 1.0 + mf * root;
      double u2 = u * u, v2 = v * v;
      double res = root + 1.0 - q / u2 - p / v2;
      double der = 1.0 + 2.0 * df * q / (u2 * u) + 2.0 * mf * p / (v2 * v);
      root -= res / der;
      if (fabs(res) < 1.0e-12) break;
    }
  }
  return root;
}
// porous_return.py:34-40
__device__ __forceinline__ double porosity_prediction(double n, float q, float p, float df, float ms, float ex, double pv, float n0) {
  double a = pressure_coefficient64(n);
  double root = overstress64(double(q), a * double(p), double(df), double(ms) * a);
  // no overstress: with finite ms * a and ex, den = 1 and pv * exp(ex * 0) is pv exactly
  if (root == 0.0 && isfinite(double(ms) * a) && isfinite(double(ex))) return pv - 1.0 + double(n0);
  double den = 1.0 + double(ms) * a * root;
  return pv * exp(double(ex) * (1.0 - 1.0 / den)) - 1.0 + double(n0);
}
// porous_return.py:43-61 (trial early exit, else 64 bisection steps)
__device__ __forceinline__ float solve_porosity(float q, float p, float df, float ms, float ex, double pv, float n0) {
  double trial = pv - 1.0 + double(n0);
  if (trial > 0.0 && trial < 1.0) {
    double tc = pressure_coefficient64(trial);
    if (double(q) + tc * double(p) <= 1.0) return float(trial);
  }
  double lo = 0.0, hi = 1.0 - 1.0e-9;
  for (int it = 0; it < 64; ++it) {
    double mid = (lo + hi) * 0.5;
    double pred = porosity_prediction(mid, q, p, df, ms, ex, pv, n0);
    if (mid < pred) lo = mid; else hi = mid;
  }
  return float((lo + hi) * 0.5);
}

// Overstress Newton of porous_return.py:19-31 started from `guess` when the
// residual there is not positive (left of the root on this increasing,
// concave residual Newton's iterates stay left and increase monotonically),
// else from 0 as the reference does.
__device__ __forceinline__ double overstress64_from(double q, double p, double df, double mf, double guess) {
  double root = 0.0;
  if (q + p > 1.0) {
    if (guess > 0.0) {
      double u = 1.0 + df * guess, v = 1.0 + mf * guess;
      if (guess + 1.0 - q / (u * u) - p / (v * v) <= 0.0) root = guess;
    }
    for (int it = 0; it < 48; ++it) {
      double u = 1.0 + df * root, v = 1.0 + mf * root;
      double u2 = u * u, v2 = v * v;
      double res = root + 1.0 - q / u2 - p / v2;
      double der = 1.0 + 2.0 * df * q / (u2 * u) + 2.0 * mf * p / (v2 * v);
      root -= res / der;
      if (fabs(res) < 1.0e-12) break;
    }
  }
  return root;
}

// f(n) = n - porosity_prediction(n) and f'(n), with the inner overstress
// solve warm-started from `gamma` (updated).  f' follows the implicit
// derivative of the converged scalar equations (porous_return.py:113-137).
__device__ __forceinline__ double porosity_gap(double n, float qf, float pf, float dff, float msf, float exf, double pv, float n0,
                                               double& gamma, double& slope) {
  const double q = qf, p = pf, d = dff, h = msf, ex = exf;
  double a = pressure_coefficient64(n);
  gamma = overstress64_from(q, a * p, d, h * a, gamma);
  double u = 1.0 + d * gamma, v = 1.0 + h * a * gamma;
  double grown = pv * exp(ex * (1.0 - 1.0 / v));
  double ra = 0.0;
  if (q + a * p > 1.0) {
    double rd = 1.0 + 2.0 * d * q / (u * u * u) + 2.0 * h * a * a * p / (v * v * v);
    ra = p * (1.0 - h * a * gamma) / (rd * v * v * v);
  }
  double ad = n > 0.0 ? a * (2.0 / (n * log(1.0 / n)) - 3.5 / (3.5 * n + 1.0)) : 0.0;
  slope = 1.0 - grown * ex / (v * v) * h * ad * (gamma + a * ra);
  return n - (grown - 1.0 + double(n0));
}

// The end-state porosity equation of solve_porosity (n = prediction(n) on
// [0, 1 - 1e-9]) solved by safeguarded Newton (Numerical Recipes' rtsafe):
// start at the old porosity, keep the bisection's bracket invariant
// f(lo) < 0 <= f(hi), take the Newton step when it stays inside the bracket
// and at least halves the previous step, else bisect; stop when the step or
// the bracket is a few ulps.  The root agrees with the reference's 64-step
// bisection to double rounding, hence to the float32 result it returns.
// Results at the endpoints (the reference's flagged no-solution cases) are
// recomputed with the literal bisection.
__device__ inline float solve_porosity_bracketed(float q, float p, float df, float ms, float ex, double pv, float n0,
                                                int* evals = nullptr) {
  double trial = pv - 1.0 + double(n0);
  if (trial > 0.0 && trial < 1.0) {
    double tc = pressure_coefficient64(trial);
    if (double(q) + tc * double(p) <= 1.0) return float(trial);
  }
  double lo = 0.0, hi = 1.0 - 1.0e-9;
  double x = (trial > 0.0 && trial < 1.0) ? trial : 0.5 * (lo + hi);
  double gamma = 0.0, slope = 1.0, last = hi - lo;
  bool done = false;
  for (int it = 0; it < 100 && !done; ++it) {
    double f = porosity_gap(x, q, p, df, ms, ex, pv, n0, gamma, slope);
    if (evals) ++*evals;
    if (f < 0.0) lo = x; else hi = x;
    if (hi - lo <= 4.0e-16 * hi) break;
    double step = slope > 0.0 ? f / slope : 2.0;
    double xn = x - step;
    if (!(xn > lo && xn < hi) || fabs(step) > 0.5 * last) {
      xn = 0.5 * (lo + hi);
      last = hi - lo;
    } else {
      last = fabs(step);
    }
    done = fabs(xn - x) <= 4.0e-16 * fabs(x);
    x = xn;
  }
  if (!(x > 1.0e-300) || x >= 1.0 - 2.0e-9) {
    if (evals) *evals += 64;
    return solve_porosity(q, p, df, ms, ex, pv, n0);
  }
  return float(x);
}

// The same root from a predicted start (n_start, with gamma_start for the
// inner overstress solve): plain Newton on f(n) = n - prediction(n), at most
// kGuidedEvals evaluations, every iterate inside (0, 1 - 1e-9).  Converged
// when a Newton step is below 1e-10: the next iterate is then within
// ~f''/f' * 1e-20 of the root, so float(root) is the float the bracketed
// solve returns.  Returns false when not converged (the caller then runs the
// bracketed solve).
constexpr int kGuidedEvals = 4;
__device__ inline bool solve_porosity_guided(float q, float p, float df, float ms, float ex, double pv, float n0, double n_start,
                                             double gamma_start, float& root, int& evals) {
  double x = n_start, gamma = gamma_start > 0.0 ? gamma_start : 0.0, slope = 1.0;
  for (int it = 0; it < kGuidedEvals; ++it) {
    if (!(x > 0.0 && x < 1.0 - 1.0e-9)) return false;
    double f = porosity_gap(x, q, p, df, ms, ex, pv, n0, gamma, slope);
    ++evals;
    if (!(slope > 0.0) || !isfinite(f)) return false;
    double step = f / slope;
    x -= step;
    if (fabs(step) <= 1.0e-10) {
      if (!(x > 0.0 && x < 1.0 - 2.0e-9)) return false;
      root = float(x);
      return true;
    }
  }
  return false;
}

// The network's porosity standing in for the coupled solve (params.surrogate 2): one evaluation of the
// porosity equation f(n) = n - prediction(n) at the network's value (inner overstress warm-started from its
// return), one Newton correction.  Accepted when the correction is at most kStandinStep: the corrected value
// is then within |f''/(2 f')| step^2 (about 1e-9 here) of the root the exact solve finds, about one float
// ulp of the porosity.  The porosity residual |root - prediction(root)| of the float result follows from the
// same linearisation, f(root) = f(x) + f'(x) (root - x) with f(x) of order step^2.  Returns false when the
// check fails (the caller then solves exactly).
constexpr double kStandinStep = 2.0e-5;
__device__ inline bool porosity_standin(float q, float p, float df, float ms, float ex, double pv, float n0, double n_net,
                                        double gamma_start, float& root, double& residual, int& evals) {
  if (!(n_net > 0.0 && n_net < 1.0 - 1.0e-9)) return false;
  double gamma = gamma_start > 0.0 ? gamma_start : 0.0, slope = 1.0;
  const double f = porosity_gap(n_net, q, p, df, ms, ex, pv, n0, gamma, slope);
  ++evals;
  if (!(slope > 0.0) || !isfinite(f)) return false;
  const double step = f / slope;
  if (!(fabs(step) <= kStandinStep)) return false;
  const double x = n_net - step;
  if (!(x > 0.0 && x < 1.0 - 2.0e-9)) return false;
  root = float(x);
  residual = fabs(slope * (double(root) - x));
  return true;
}

// Result of the Ding return map for one particle.
struct DingOut {
  mat3<float> Fe, Fp, tau;
  float residual, porosity, porosity_residual, T, yield_stress, viscosity;
  int gelled;
  int flag0, flag1;  // porosity domain, unresolved coupled porosity
};

struct DingParams {
  float mu, lam, yield_stress, n0, dt;
  float knots[4], yf[4], vf[4];
  int temperature_enabled, coupled;
  int bracketed;  // 1: solve_porosity_bracketed, 0: the reference's 64-step bisection
};

// ding.py:103-137 with the SVD of the trial gradient precomputed.
__device__ __forceinline__ void ding_return(float det_trial, const mat3<float>& U, const float s[3], const mat3<float>& V,
                                            float mu, float lam, float y, float eta, float a, float dt, mat3<float>& Fe,
                                            mat3<float>& stress, float& residual, float* eps_ret = nullptr) {
  float eps[3] = {logf(s[0]), logf(s[1]), logf(s[2])};
  float tr = eps[0] + eps[1] + eps[2];
  float ts[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) ts[i] = 2.0f * mu * eps[i] + lam * tr;
  float mean = (ts[0] + ts[1] + ts[2]) / 3.0f;
  float dev[3] = {ts[0] - mean, ts[1] - mean, ts[2] - mean};
  float q2 = 1.5f * (dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]) / (y * y);
  float p2 = a * mean * mean / (y * y);
  float df = 6.0f * mu * dt * y * y / eta;
  float mf = 2.0f * a / 3.0f * (2.0f * mu + 3.0f * lam) * dt * y * y / eta;
  float g = solve_overstress(q2, p2, df, mf);
  float cs[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) cs[i] = dev[i] / (1.0f + df * g) + mean / (1.0f + mf * g);
  float cm = (cs[0] + cs[1] + cs[2]) / 3.0f;
  float e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float le = (cs[i] - cm) / (2.0f * mu) + cm / (2.0f * mu + 3.0f * lam);
    if (eps_ret) eps_ret[i] = le;
    e[i] = expf(le);
  }
  // Fe = U diag(e) V^T ; stress = U diag(cs) U^T
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Fe.a[i * 3 + j] = U.a[i * 3 + 0] * e[0] * V.a[j * 3 + 0] + U.a[i * 3 + 1] * e[1] * V.a[j * 3 + 1] + U.a[i * 3 + 2] * e[2] * V.a[j * 3 + 2];
      stress.a[i * 3 + j] =
          U.a[i * 3 + 0] * cs[0] * U.a[j * 3 + 0] + U.a[i * 3 + 1] * cs[1] * U.a[j * 3 + 1] + U.a[i * 3 + 2] * cs[2] * U.a[j * 3 + 2];
    }
  if (a == 0.0f) Fe = m_scale(powf(det_trial / m_det(Fe), 1.0f / 3.0f), Fe);
  residual = 0.0f;
  if (q2 + p2 > 1.0f) {
    float u = 1.0f + df * g, v = 1.0f + mf * g;
    residual = fabsf(g + 1.0f - q2 / (u * u) - p2 / (v * v));
  }
}

// Outcome of ding_update.
enum DingStatus { kDingInvalid = 0, kDingDone = 1, kDingDeferred = 2 };

// Optional extras of ding_update for the coupled (gas) return map.
struct DingAux {
  // in: solve the coupled porosity from a predicted start first (see solve_porosity_guided); with `standin`,
  // accept the prediction after one corrected evaluation when it passes porosity_standin's check
  bool guided = false;
  bool standin = false;
  double n_start = 0.0, gamma_start = 0.0;
  float dev_target = 0.0f;  // predicted |deviatoric returned Kirchhoff stress| (0: start the overstress at 0)
  // out
  bool accepted = false;  // the guided solve converged (standin: the prediction was accepted)
  int evals = 0;          // evaluations of the porosity equation
  float eps[3];           // principal trial log-strains, descending
  int order[3];           // eps[k] is the principal value order[k] of the SVD
  float eps_ret[3];       // returned log-strains (before the volume fit), in the order of eps
  float T;                // temperature the yield and viscosity factors were taken at
  float dev_norm;         // |deviatoric trial Kirchhoff stress|
};

// The first half of the Ding update of one particle (ding.py:140-160): the
// temperature-scaled yield stress and viscosity, the gel state and the SVD of
// the elastic trial gradient.  Everything the second half needs, so a particle
// whose coupled solve is deferred to a second pass carries this record there
// instead of decomposing its trial gradient again.
struct DingPrep {
  mat3<float> U, V;
  float s[3];
  float det_trial;  // det of the elastic trial gradient (the volume fit of a return at zero porosity coefficient)
  float Teff;       // temperature the yield and viscosity factors were taken at
  float yv, ev;     // yield stress and viscosity at Teff
  int gel;
};
// The record as 7 float4 (pass C -> deferred pass queue): U, V, s, det_trial, Teff, yv, ev, gel (bits).
constexpr int kDingPrepVec = 7;
__device__ __forceinline__ void store_prep(float4* dst, const DingPrep& r) {
  const mat3<float>& U = r.U;
  const mat3<float>& V = r.V;
  dst[0] = make_float4(U.a[0], U.a[1], U.a[2], U.a[3]);
  dst[1] = make_float4(U.a[4], U.a[5], U.a[6], U.a[7]);
  dst[2] = make_float4(U.a[8], V.a[0], V.a[1], V.a[2]);
  dst[3] = make_float4(V.a[3], V.a[4], V.a[5], V.a[6]);
  dst[4] = make_float4(V.a[7], V.a[8], r.s[0], r.s[1]);
  dst[5] = make_float4(r.s[2], r.det_trial, r.Teff, r.yv);
  dst[6] = make_float4(r.ev, __int_as_float(r.gel), 0.0f, 0.0f);
}
__device__ __forceinline__ DingPrep load_prep(const float4* src) {
  DingPrep r;
  const float4 a = src[0], b = src[1], c = src[2], d = src[3], e = src[4], f = src[5], g = src[6];
  const float u[9] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x};
  const float v[9] = {c.y, c.z, c.w, d.x, d.y, d.z, d.w, e.x, e.y};
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    r.U.a[k] = u[k];
    r.V.a[k] = v[k];
  }
  r.s[0] = e.z; r.s[1] = e.w; r.s[2] = f.x;
  r.det_trial = f.y; r.Teff = f.z; r.yv = f.w;
  r.ev = g.x;
  r.gel = __float_as_int(g.y);
  return r;
}

__device__ inline void ding_prepare(const DingParams& P, const mat3<float>& Fe_old, int gelled_in, const mat3<float>& C_new, float T,
                                    float visc_param, DingPrep& r) {
  r.gel = gelled_in;
  r.Teff = T;
  r.yv = P.yield_stress;
  r.ev = visc_param;
  if (P.temperature_enabled) {
    if (T >= P.knots[3]) r.gel = 1;
    if (r.gel) r.Teff = P.knots[3];
    r.yv = P.yield_stress * temperature_factor(r.Teff, P.knots, P.yf);
    r.ev = visc_param * temperature_factor(r.Teff, P.knots, P.vf);
  }
  mat3<float> trial;
  {
    mat3<float> inc = m_scale(P.dt, C_new);
    inc.a[0] += 1.0f; inc.a[4] += 1.0f; inc.a[8] += 1.0f;
    trial = m_mul(inc, Fe_old);
  }
  r.det_trial = m_det(trial);
  mat3<double> Ud, Vd;
  double sd[3];
  // converged to |off-diagonal| <= 1e-9 |diagonal| of trial^T trial: the singular values' error is second order
  // (1e-18 relative) and U, V are within 1e-9, both below the float rounding of what the return map uses
  svd3(m_cast<double>(trial), Ud, sd, Vd, 1e-18);
  r.U = m_cast<float>(Ud);
  r.V = m_cast<float>(Vd);
  r.s[0] = float(sd[0]);
  r.s[1] = float(sd[1]);
  r.s[2] = float(sd[2]);
}

// The second half (ding.py:160-221) in two steps.  ding_solve: the coupled
// porosity of the return (porous_return.py:43-61) from the trial's principal
// stretches s and the yield stress and viscosity; with `defer` set, a particle
// whose porosity needs the bisection solve returns kDingDeferred, so a second
// pass can run the expensive solves densely.  ding_complete: the return map,
// the volume fit and the outputs.
__device__ inline int ding_solve(const DingParams& P, const float s[3], float yv, float ev, float Teff, float porosity, bool defer,
                                 DingAux* aux, float& ret_n, float& cres, int& flag1) {
  ret_n = porosity;
  cres = 0.0f;
  if (!P.coupled) return kDingDone;
  const float mu = P.mu, lam = P.lam, dt = P.dt;
  float eps[3] = {logf(s[0]), logf(s[1]), logf(s[2])};
  float ts[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) ts[i] = 2.0f * mu * eps[i] + lam * (eps[0] + eps[1] + eps[2]);
  float mean = (ts[0] + ts[1] + ts[2]) / 3.0f;
  float dev[3] = {ts[0] - mean, ts[1] - mean, ts[2] - mean};
  float q = 1.5f * (dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]) / (yv * yv);
  float pv = mean * mean / (yv * yv);
  float df = 6.0f * mu * dt * yv * yv / ev;
  float ms = 2.0f / 3.0f * (2.0f * mu + 3.0f * lam) * dt * yv * yv / ev;
  float ex = 3.0f * mean / (2.0f * mu + 3.0f * lam);
  double opv = 1.0 - double(P.n0) + double(porosity);
  if (aux) {
    // descending order by a 3-element network on (value, index) pairs, kept in registers
    float e0 = eps[0], e1 = eps[1], e2 = eps[2];
    int o0 = 0, o1 = 1, o2 = 2;
    if (e0 < e1) { float t = e0; e0 = e1; e1 = t; int u = o0; o0 = o1; o1 = u; }
    if (e1 < e2) { float t = e1; e1 = e2; e2 = t; int u = o1; o1 = o2; o2 = u; }
    if (e0 < e1) { float t = e0; e0 = e1; e1 = t; int u = o0; o0 = o1; o1 = u; }
    aux->order[0] = o0; aux->order[1] = o1; aux->order[2] = o2;
    aux->eps[0] = e0; aux->eps[1] = e1; aux->eps[2] = e2;
    aux->T = Teff;
    aux->dev_norm = sqrtf(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  }
  if (aux && aux->guided && aux->dev_target > 0.0f && df > 0.0f) {
    const double g = (double(aux->dev_norm) / double(aux->dev_target) - 1.0) / double(df);
    aux->gamma_start = g > 0.0 ? g : 0.0;
  }
  bool solved = false, stood_in = false, clear_inside = false;
  if (defer) {
    double trial_n = opv - 1.0 + double(P.n0);
    const double a_trial = (trial_n > 0.0 && trial_n < 1.0) ? pressure_coefficient64(trial_n) : 0.0;
    bool inside = trial_n > 0.0 && trial_n < 1.0 && double(q) + a_trial * double(pv) <= 1.0;
    if (!inside) return kDingDeferred;
    // the solvers' elastic exit, decided by the same test on the same values
    ret_n = float(trial_n);
    solved = true;
    // The residual below evaluates the prediction at float(trial_n), within 2^-24 of trial_n.  On (0, 0.5] the
    // pressure coefficient a(n) = 9 / (4 ln(1/n)^2 (3.5 n + 1)) has |d ln a / d ln n| = |2 / ln(1/n) - 3.5 n / (3.5 n + 1)|
    // < 3, so q + a p moves by less than 2e-7 of a p <= 1: with more margin than 1e-6 below 1 the prediction's
    // overstress is 0 and it returns pv - 1 + n0 without evaluating a again.
    clear_inside = trial_n > 1.0e-30 && trial_n <= 0.5 && 1.0 - (double(q) + a_trial * double(pv)) > 1.0e-6 && isfinite(ms) &&
                   isfinite(ex);
  }
  double standin_residual = 0.0;
  if (!solved && aux && aux->standin) {
    stood_in = porosity_standin(q, pv, df, ms, ex, opv, P.n0, aux->n_start, aux->gamma_start, ret_n, standin_residual,
                                aux->evals);
    aux->accepted = solved = stood_in;
  }
  if (!solved && aux && aux->guided) {
    solved = solve_porosity_guided(q, pv, df, ms, ex, opv, P.n0, aux->n_start, aux->gamma_start, ret_n, aux->evals);
    if (!aux->standin) aux->accepted = solved;
  }
  if (!solved) {
    if (P.bracketed) {
      ret_n = solve_porosity_bracketed(q, pv, df, ms, ex, opv, P.n0, aux ? &aux->evals : nullptr);
    } else {
      ret_n = solve_porosity(q, pv, df, ms, ex, opv, P.n0);
      if (aux) aux->evals += 64;
    }
  }
  if (stood_in) {
    cres = float(standin_residual);
  } else {
    double pred = clear_inside ? opv - 1.0 + double(P.n0) : porosity_prediction(double(ret_n), q, pv, df, ms, ex, opv, P.n0);
    if (aux) ++aux->evals;
    cres = float(fabs(double(ret_n) - pred));
  }
  if (cres > 1.0e-6f || ret_n <= 0.0f || ret_n >= 1.0f) flag1 = 1;
  return kDingDone;
}

__device__ inline void ding_complete(const DingParams& P, const DingPrep& pr, float porosity, float ret_n, float cres,
                                     const mat3<float>& F_total, float T, DingOut& o, DingAux* aux) {
  const mat3<float>& U = pr.U;
  const mat3<float>& V = pr.V;
  const float* s = pr.s;
  const float yv = pr.yv, ev = pr.ev;
  const float mu = P.mu, lam = P.lam, dt = P.dt;
  mat3<float> Fe, tau;
  float resid;
  float eps_ret[3];
  ding_return(pr.det_trial, U, s, V, mu, lam, yv, ev, porosity_coefficient(ret_n), dt, Fe, tau, resid, eps_ret);
  if (aux && P.coupled) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int o = aux->order[k];
      aux->eps_ret[k] = o == 0 ? eps_ret[0] : (o == 1 ? eps_ret[1] : eps_ret[2]);
    }
  }
  if (P.coupled) {
    double target = 1.0 - double(P.n0) + double(ret_n);
    double vs = m_det(m_cast<double>(F_total)) / (target * m_det(m_cast<double>(Fe)));
    Fe = m_scale(float(cbrt(vs)), Fe);
  }
  mat3<float> Fp = m_mul(m_inverse(Fe), F_total);
  float nn = m_det(Fp) - 1.0f + P.n0;
  if (P.coupled) nn = ret_n;
  if (P.n0 == 0.0f && porosity == 0.0f) nn = 0.0f;
  if (!isfinite(nn) || nn < 0.0f || nn >= 1.0f) o.flag0 = 1;
  o.Fe = Fe;
  o.Fp = Fp;
  o.tau = tau;
  o.residual = resid;
  o.porosity_residual = cres;
  o.porosity = nn;
  o.T = T;
  o.gelled = pr.gel;
  o.yield_stress = yv;
  o.viscosity = ev;
}

__device__ inline int ding_finish(const DingParams& P, const DingPrep& pr, float porosity, const mat3<float>& F_total, float T,
                                  DingOut& o, bool defer = false, DingAux* aux = nullptr, float4* defer_store = nullptr) {
  o.flag0 = 0;
  o.flag1 = 0;
  float ret_n, cres;
  if (ding_solve(P, pr.s, pr.yv, pr.ev, pr.Teff, porosity, defer, aux, ret_n, cres, o.flag1) == kDingDeferred) {
    if (defer_store) store_prep(defer_store, pr);
    return kDingDeferred;
  }
  ding_complete(P, pr, porosity, ret_n, cres, F_total, T, o, aux);
  return kDingDone;
}

// ding.py:140-221 (UpdateDing body for one particle).  Returns kDingInvalid when
// the incoming porosity is outside [0, 1) (flag 0, outputs untouched).  A
// deferred particle leaves the first half's record at `defer_store`.
__device__ inline int ding_update(const DingParams& P, const mat3<float>& Fe_old, float porosity, int gelled_in, const mat3<float>& F_total,
                                  const mat3<float>& C_new, float T, float visc_param, DingOut& o, bool defer = false,
                                  DingAux* aux = nullptr, float4* defer_store = nullptr) {
  o.flag0 = 0;
  o.flag1 = 0;
  if (!isfinite(porosity) || porosity < 0.0f || porosity >= 1.0f) { o.flag0 = 1; return kDingInvalid; }
  DingPrep pr;
  ding_prepare(P, Fe_old, gelled_in, C_new, T, visc_param, pr);
  return ding_finish(P, pr, porosity, F_total, T, o, defer, aux, defer_store);
}

// ------------------------------------------------------------------ gas (physics/materials/gas.py)
constexpr double kGasConstant = 8.31446261815324;
__device__ __forceinline__ double gas_pressure(double moles, float T, float J, float n) {
  return moles * kGasConstant * double(T) / (double(J) * double(n));
}

// Neo-Hookean Kirchhoff stress (physics/materials/elastic.py:4-10).
__device__ __forceinline__ mat3<float> neo_hookean_kirchhoff(const mat3<float>& F, float mu, float lam, float J) {
  mat3<float> t = m_mul_bt(F, F);
  float lj = lam * logf(J);
#pragma unroll
  for (int i = 0; i < 9; ++i) t.a[i] = mu * (t.a[i] - ((i % 4 == 0) ? 1.0f : 0.0f)) + ((i % 4 == 0) ? lj : 0.0f);
  return t;
}

// ------------------------------------------------------------------ mesh reconstruction (geometry/mesh_tracking.py:13-38)
// One 256-thread workgroup per vertex; the threads stride over the vertex's
// particle bindings.  The reference sums weighted deviations of the
// predictions x + F r from an anchor (the first binding's prediction) to keep
// float error small at high valence.  Here each thread sums deviations from
// its own first prediction a_t, with its weight total W_t, and adds
// W_t (a_t - A) once the anchor A (thread 0's a_t) is known: every summed term
// is still a small deviation, and no load waits for the anchor's dependent
// loads.  A vertex can hold ~1000 bindings and each is a dependent pair of
// loads (particle index, then its x and F), so a thread issues kRecUnroll
// bindings' loads before accumulating them in order; a slot past e reads
// binding e-1 (a valid address) with weight 0, so the loads are used
// unconditionally and cannot be sunk into per-slot branches.  Call with the
// whole workgroup and the same `v`; `red` holds 16 floats of LDS.
constexpr int kRecThreads = 256;
constexpr int kRecUnroll = 2;
__device__ inline void reconstruct_vertex_block(int v, const float* __restrict__ x, const float* __restrict__ F,
                                                const float* __restrict__ ref_vertices, const int32_t* __restrict__ offs,
                                                const int32_t* __restrict__ parts, const float* __restrict__ wts,
                                                const float* __restrict__ ref_offs, float* __restrict__ out, float* red) {
  const int t = threadIdx.x;
  const int s = offs[v], e = offs[v + 1];
  if (s == e) {
    if (t < 3) out[3 * v + t] = ref_vertices[3 * v + t];
    return;
  }
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;  // this thread's first prediction
  float d0 = 0.f, d1 = 0.f, d2 = 0.f, wsum = 0.f;
  bool first = true;
  for (int k0 = s + t; k0 < e; k0 += kRecThreads * kRecUnroll) {
    int pi[kRecUnroll];
#pragma unroll
    for (int u = 0; u < kRecUnroll; ++u) pi[u] = parts[min(k0 + kRecThreads * u, e - 1)];
#pragma unroll
    for (int u = 0; u < kRecUnroll; ++u) {
      const int k = min(k0 + kRecThreads * u, e - 1);
      const float* Fp = F + 9 * pi[u];
      const float* r = ref_offs + 3 * k;
      const float wk = k0 + kRecThreads * u < e ? wts[k] : 0.0f;
      const float p0 = x[3 * pi[u] + 0] + (Fp[0] * r[0] + Fp[1] * r[1] + Fp[2] * r[2]);
      const float p1 = x[3 * pi[u] + 1] + (Fp[3] * r[0] + Fp[4] * r[1] + Fp[5] * r[2]);
      const float p2 = x[3 * pi[u] + 2] + (Fp[6] * r[0] + Fp[7] * r[1] + Fp[8] * r[2]);
      if (u == 0 && first) {  // the loop runs only while k0 < e: slot 0 of the first pass is a binding
        a0 = p0;
        a1 = p1;
        a2 = p2;
        first = false;
      }
      d0 += wk * (p0 - a0);
      d1 += wk * (p1 - a1);
      d2 += wk * (p2 - a2);
      wsum += wk;
    }
  }
  if (t == 0) {  // thread 0's first binding is binding s: the reference's anchor
    red[12] = a0;
    red[13] = a1;
    red[14] = a2;
  }
  __syncthreads();
  const float A0 = red[12], A1 = red[13], A2 = red[14];
  if (wsum != 0.0f) {
    d0 += wsum * (a0 - A0);
    d1 += wsum * (a1 - A1);
    d2 += wsum * (a2 - A2);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    d0 += __shfl_xor(d0, off, 64);
    d1 += __shfl_xor(d1, off, 64);
    d2 += __shfl_xor(d2, off, 64);
  }
  if ((t & 63) == 0) {
    red[3 * (t >> 6) + 0] = d0;
    red[3 * (t >> 6) + 1] = d1;
    red[3 * (t >> 6) + 2] = d2;
  }
  __syncthreads();
  if (t < 3) {
    float d = red[t];
#pragma unroll
    for (int w = 1; w < kRecThreads / 64; ++w) d += red[3 * w + t];
    out[3 * v + t] = red[12 + t] + d;
  }
}

// ------------------------------------------------------------------ wave-level helpers
// Append `value` to a device queue for every lane with `take` set: one atomic
// per wavefront reserves the slots, lanes fill them in lane order.  Call with
// the whole wavefront converged.
__device__ __forceinline__ void wave_append(int32_t* counter, int32_t* queue, bool take, int32_t value) {
  unsigned long long m = __ballot(take);
  if (m == 0ull) return;
  const unsigned lane = __lane_id();
  const unsigned leader = unsigned(__ffsll((long long)m) - 1);
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, int(leader), 64);
  if (take) {
    unsigned long long below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
    queue[base + __popcll(below)] = value;
  }
}

// Reserve `count` consecutive slots of a device list for every thread of a
// workgroup with one atomic per workgroup: an inclusive scan of the counts
// within each wavefront (shuffles), then across the wavefronts in LDS.
// Returns the thread's first slot.  Every thread of the workgroup must call
// it (blockDim.x a multiple of 64, at most 1024); `lds` holds 17 ints.
// Appends to one counter from many wavefronts serialise at that address, so
// batching them per workgroup keeps the counter off the critical path.
__device__ __forceinline__ int block_reserve(int32_t* counter, int count, int* lds) {
  const int lane = int(__lane_id());
  const int w = int(threadIdx.x) >> 6;
  const int nw = int(blockDim.x) >> 6;
  int incl = count;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) lds[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < nw; ++i) {
      int c = lds[i];
      lds[i] = run;
      run += c;
    }
    lds[16] = run > 0 ? atomicAdd(counter, run) : 0;
  }
  __syncthreads();
  const int first = lds[16] + lds[w] + incl - count;
  __syncthreads();  // lds may be reused right after
  return first;
}

// block_reserve for one optional item per thread.
__device__ __forceinline__ void block_append(int32_t* counter, int32_t* list, bool take, int32_t value, int* lds) {
  const int slot = block_reserve(counter, take ? 1 : 0, lds);
  if (take) list[slot] = value;
}

__device__ __forceinline__ void flag_max(int* flags, int index, bool set) {
  // One atomic per wavefront: the lowest set lane reports.
  unsigned long long m = __ballot(set);
  if (m != 0ull && __lane_id() == (unsigned)(__ffsll((long long)m) - 1)) atomicMax(flags + index, 1);
}

}  // namespace bk
