// CPU oracle for the MLS-MPM baking substep.  TEST INFRASTRUCTURE ONLY.
//
// A serial, line-by-line restatement of the reference's Warp kernels
// (bbbpimasheep/baking, Python + Warp 1.18).  Only tests/, the smoke check and
// bench.py's cpu_baseline leg may load this library, and only as the checker or
// the timed CPU baseline; the product path (baking_amd/) never links it.
//
// Conventions mirror Warp's CPU device: each kernel is one serial loop over its
// launch dimension in index order, so every atomic_add becomes a plain add in
// particle order.  `R` is the reference's `float` (wp.float32); it is
// instantiated as float (the reference precision) and as double (the FP64 run
// of the same kernels used for error statements).  Quantities the reference
// keeps in wp.float64 are double in both instantiations, and every explicit
// float()/wp.float64() conversion of the reference is kept.
//
// vec3 arrays are contiguous [n][3]; mat33 arrays are contiguous row-major
// [n][3][3] (Warp's layout).  Error flags are int arrays with the reference's
// indices.
//
// The 3x3 SVD (wp.svd3 in the reference, third-party native code) is replaced
// by a one-sided Jacobi (Hestenes 1958) SVD in double precision; its outputs
// are checked against LAPACK in tests/.  Results of the return map are
// invariant to the SVD's sign/order conventions.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

// The timing build of bench.py's CPU baseline (-DORACLE_OPENMP -fopenmp) splits each kernel's loop over the
// host's threads and makes the scatters' adds and the flag raises atomic; the checker build (no define) is
// the serial restatement described above, in particle order.
#ifdef ORACLE_OPENMP
#include <omp.h>
#define ORACLE_FOR _Pragma("omp parallel for schedule(static, 256)")
#define ORACLE_ATOMIC _Pragma("omp atomic")
#else
#define ORACLE_FOR
#define ORACLE_ATOMIC
#endif

namespace {

// flags[k] = max(flags[k], 1): the reference's wp.atomic_max(flags, k, 1)
inline void raise_flag(int* flags, int k) {
#ifdef ORACLE_OPENMP
#pragma omp atomic write
  flags[k] = 1;
#else
  flags[k] = std::max(flags[k], 1);
#endif
}

// ---------------------------------------------------------------- small linear algebra
template <typename T> struct M3 { T a[9]; };
template <typename T> struct V3 { T a[3]; };

template <typename T> inline M3<T> load_m(const T* p) { M3<T> m; for (int i = 0; i < 9; ++i) m.a[i] = p[i]; return m; }
template <typename T> inline void store_m(T* p, const M3<T>& m) { for (int i = 0; i < 9; ++i) p[i] = m.a[i]; }
template <typename T> inline V3<T> load_v(const T* p) { return V3<T>{{p[0], p[1], p[2]}}; }
template <typename T> inline void store_v(T* p, const V3<T>& v) { p[0] = v.a[0]; p[1] = v.a[1]; p[2] = v.a[2]; }
template <typename T> inline M3<T> ident() { M3<T> m{}; m.a[0] = m.a[4] = m.a[8] = T(1); return m; }
template <typename T> inline M3<T> zero_m() { M3<T> m{}; return m; }
template <typename T> inline M3<T> mul(const M3<T>& x, const M3<T>& y) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T s = x.a[i * 3 + 0] * y.a[0 * 3 + j];
      s = s + x.a[i * 3 + 1] * y.a[1 * 3 + j];
      s = s + x.a[i * 3 + 2] * y.a[2 * 3 + j];
      r.a[i * 3 + j] = s;
    }
  return r;
}
template <typename T> inline V3<T> mulv(const M3<T>& x, const V3<T>& v) {
  V3<T> r;
  for (int i = 0; i < 3; ++i) {
    T s = x.a[i * 3 + 0] * v.a[0];
    s = s + x.a[i * 3 + 1] * v.a[1];
    s = s + x.a[i * 3 + 2] * v.a[2];
    r.a[i] = s;
  }
  return r;
}
template <typename T> inline M3<T> scale(T s, const M3<T>& m) { M3<T> r; for (int i = 0; i < 9; ++i) r.a[i] = s * m.a[i]; return r; }
template <typename T> inline M3<T> add(const M3<T>& x, const M3<T>& y) { M3<T> r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] + y.a[i]; return r; }
template <typename T> inline M3<T> sub(const M3<T>& x, const M3<T>& y) { M3<T> r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] - y.a[i]; return r; }
template <typename T> inline M3<T> transpose(const M3<T>& m) {
  M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = m.a[j * 3 + i]; return r;
}
template <typename T> inline T det(const M3<T>& m) {
  const T* a = m.a;
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}
template <typename T> inline M3<T> inverse(const M3<T>& m) {
  const T* a = m.a;
  T d = det(m);
  M3<T> r;
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
template <typename T, typename S> inline M3<T> cast_m(const M3<S>& m) { M3<T> r; for (int i = 0; i < 9; ++i) r.a[i] = T(m.a[i]); return r; }

// One-sided Jacobi SVD (Hestenes): rotate column pairs of A = F*V until they
// are mutually orthogonal; singular values are the column norms, U the
// normalised columns.  Converges quadratically; 12 sweeps bound the loop.
// Outputs U, sigma, V with F = U diag(sigma) V^T, U and V proper rotations,
// sigma[2] carrying the sign of det(F).
void svd3_jacobi(const M3<double>& F, M3<double>& U, double sigma[3], M3<double>& V) {
  M3<double> A = F;
  V = ident<double>();
  for (int sweep = 0; sweep < 12; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < 3; ++i) {
          alpha += A.a[i * 3 + p] * A.a[i * 3 + p];
          beta += A.a[i * 3 + q] * A.a[i * 3 + q];
          gamma += A.a[i * 3 + p] * A.a[i * 3 + q];
        }
        if (gamma == 0.0) continue;
        off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < 3; ++i) {
          double ap = A.a[i * 3 + p], aq = A.a[i * 3 + q];
          A.a[i * 3 + p] = c * ap - s * aq;
          A.a[i * 3 + q] = s * ap + c * aq;
          double vp = V.a[i * 3 + p], vq = V.a[i * 3 + q];
          V.a[i * 3 + p] = c * vp - s * vq;
          V.a[i * 3 + q] = s * vp + c * vq;
        }
      }
    if (off < 1e-17) break;
  }
  double norms[3];
  for (int j = 0; j < 3; ++j)
    norms[j] = std::sqrt(A.a[0 * 3 + j] * A.a[0 * 3 + j] + A.a[1 * 3 + j] * A.a[1 * 3 + j] + A.a[2 * 3 + j] * A.a[2 * 3 + j]);
  // Sort descending (columns of A and V together).
  int order[3] = {0, 1, 2};
  std::sort(order, order + 3, [&](int x, int y) { return norms[x] > norms[y]; });
  M3<double> As, Vs;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) { As.a[i * 3 + j] = A.a[i * 3 + order[j]]; Vs.a[i * 3 + j] = V.a[i * 3 + order[j]]; }
  for (int j = 0; j < 3; ++j) sigma[j] = norms[order[j]];
  if (det(Vs) < 0) { for (int i = 0; i < 3; ++i) { Vs.a[i * 3 + 2] = -Vs.a[i * 3 + 2]; As.a[i * 3 + 2] = -As.a[i * 3 + 2]; } }
  // U columns: normalise the two largest, complete with a cross product so U is a rotation.
  double u0[3], u1[3], u2[3];
  for (int i = 0; i < 3; ++i) u0[i] = sigma[0] > 0 ? As.a[i * 3 + 0] / sigma[0] : (i == 0 ? 1.0 : 0.0);
  double d01 = 0;
  for (int i = 0; i < 3; ++i) { u1[i] = As.a[i * 3 + 1]; d01 += u0[i] * u1[i]; }
  for (int i = 0; i < 3; ++i) u1[i] -= d01 * u0[i];
  double n1 = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  if (n1 > 1e-300) { for (int i = 0; i < 3; ++i) u1[i] /= n1; }
  else {
    double e[3] = {std::fabs(u0[0]) < 0.9 ? 1.0 : 0.0, std::fabs(u0[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    double d = e[0] * u0[0] + e[1] * u0[1] + e[2] * u0[2];
    for (int i = 0; i < 3; ++i) u1[i] = e[i] - d * u0[i];
    double nn = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    for (int i = 0; i < 3; ++i) u1[i] /= nn;
  }
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  // Signed third singular value: projection of A's third column on u2.
  sigma[2] = As.a[0 * 3 + 2] * u2[0] + As.a[1 * 3 + 2] * u2[1] + As.a[2 * 3 + 2] * u2[2];
  for (int i = 0; i < 3; ++i) { U.a[i * 3 + 0] = u0[i]; U.a[i * 3 + 1] = u1[i]; U.a[i * 3 + 2] = u2[i]; }
  V = Vs;
}

// ---------------------------------------------------------------- stencil helpers
// physics/transfer.py:4-13
template <typename R> inline R quadratic_weight(R l, int i) {
  if (i == 0) return R(0.5) * (R(1.5) - l) * (R(1.5) - l);
  if (i == 1) return R(0.75) - (l - R(1.0)) * (l - R(1.0));
  return R(0.5) * (l - R(0.5)) * (l - R(0.5));
}
// physics/transfer.py:16-20
template <typename R> inline void stencil_base(const R g[3], int b[3]) {
  for (int a = 0; a < 3; ++a) b[a] = int(std::floor(g[a] - R(0.5)));
}
// physics/transfer.py:23-28
inline bool valid_stencil(const int b[3], int res) {
  return b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[0] + 2 < res && b[1] + 2 < res && b[2] + 2 < res;
}
// physics/transfer.py:31-33
inline int64_t flatten_node(int i, int j, int k, int res) { return (int64_t(i) * res + j) * res + k; }

template <typename R> inline void grid_position(const R* x, const R* origin, R dx, R g[3]) {
  for (int a = 0; a < 3; ++a) g[a] = (x[a] - origin[a]) / dx;
}

// ---------------------------------------------------------------- Ding helpers (float)
// physics/materials/ding.py:56-65
template <typename R> R temperature_factor(R t, const R* knots, const R* factors) {
  R f = factors[0];
  for (int k = 0; k < 3; ++k)
    if (t >= knots[k]) {
      R frac = std::min((t - knots[k]) / (knots[k + 1] - knots[k]), R(1.0));
      f = (R(1.0) - frac) * factors[k] + frac * factors[k + 1];
    }
  return f;
}
// physics/materials/ding.py:68-74
template <typename R> R porosity_coefficient(R n) {
  R c = R(0.0);
  if (n > R(0.0)) c = R(9.0) / (R(4.0) * std::pow(std::log(R(1.0) / n), R(2.0)) * (R(3.5) * n + R(1.0)));
  return c;
}
// physics/materials/ding.py:77-87
template <typename R> R solve_overstress(R q2, R p2, R df, R mf) {
  R root = R(0.0);
  if (q2 + p2 > R(1.0)) {
    for (int it = 0; it < 32; ++it) {
      R dd = R(1.0) + df * root;
      R md = R(1.0) + mf * root;
      R res = root + R(1.0) - q2 / std::pow(dd, R(2.0)) - p2 / std::pow(md, R(2.0));
      R der = R(1.0) + R(2.0) * df * q2 / std::pow(dd, R(3.0)) + R(2.0) * mf * p2 / std::pow(md, R(3.0));
      root = root - res / der;
    }
  }
  return root;
}

// ---------------------------------------------------------------- porous return (double)
// physics/materials/porous_return.py:11-16
double pressure_coefficient64(double n) {
  double c = 0.0;
  if (n > 0.0) c = 9.0 / (4.0 * std::pow(std::log(1.0 / n), 2.0) * (3.5 * n + 1.0));
  return c;
}
// physics/materials/porous_return.py:19-31
double overstress64(double q, double p, double df, double mf) {
  double root = 0.0;
  if (q + p > 1.0) {
    for (int it = 0; it < 48; ++it) {
      double u = 1.0 + df * root, v = 1.0 + mf * root;
      double res = root + 1.0 - q / (u * u) - p / (v * v);
      double der = 1.0 + 2.0 * df * q / std::pow(u, 3.0) + 2.0 * mf * p / std::pow(v, 3.0);
      root -= res / der;
      if (std::fabs(res) < 1.0e-12) break;
    }
  }
  return root;
}
// physics/materials/porous_return.py:34-40
template <typename R> double porosity_prediction(double n, R q, R p, R df, R ms, R ex, double pv, R n0) {
  double a = pressure_coefficient64(n);
  double root = overstress64(double(q), a * double(p), double(df), double(ms) * a);
  double den = 1.0 + double(ms) * a * root;
  return pv * std::exp(double(ex) * (1.0 - 1.0 / den)) - 1.0 + double(n0);
}
// physics/materials/porous_return.py:43-61
template <typename R> R solve_porosity(R q, R p, R df, R ms, R ex, double pv, R n0) {
  double trial = pv - 1.0 + double(n0);
  if (trial > 0.0 && trial < 1.0) {
    double tc = pressure_coefficient64(trial);
    if (double(q) + tc * double(p) <= 1.0) return R(trial);
  }
  double lo = 0.0, hi = 1.0 - 1.0e-9;
  for (int it = 0; it < 64; ++it) {
    double mid = (lo + hi) * 0.5;
    double pred = porosity_prediction<R>(mid, q, p, df, ms, ex, pv, n0);
    if (mid < pred) lo = mid; else hi = mid;
  }
  return R((lo + hi) * 0.5);
}

// physics/materials/ding.py:103-137
template <typename R>
void ding_return(const M3<R>& trial, R mu, R lam, R y, R eta, R a, R dt, const M3<R>* svd_u, const R* svd_s,
                 const M3<R>* svd_v, M3<R>& Fe, M3<R>& stress, R& residual) {
  M3<R> U = *svd_u, V = *svd_v;
  R s[3] = {svd_s[0], svd_s[1], svd_s[2]};
  R eps[3] = {std::log(s[0]), std::log(s[1]), std::log(s[2])};
  R tr = eps[0] + eps[1] + eps[2];
  R ts[3];
  for (int i = 0; i < 3; ++i) ts[i] = R(2.0) * mu * eps[i] + lam * tr * R(1.0);
  R mean = (ts[0] + ts[1] + ts[2]) / R(3.0);
  R dev[3] = {ts[0] - mean * R(1.0), ts[1] - mean * R(1.0), ts[2] - mean * R(1.0)};
  R dd = dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2];
  R q2 = R(1.5) * dd / (y * y);
  R p2 = a * mean * mean / (y * y);
  R df = R(6.0) * mu * dt * y * y / eta;
  R mf = R(2.0) * a / R(3.0) * (R(2.0) * mu + R(3.0) * lam) * dt * y * y / eta;
  R g = solve_overstress(q2, p2, df, mf);
  R cs[3];
  for (int i = 0; i < 3; ++i) cs[i] = dev[i] / (R(1.0) + df * g) + mean / (R(1.0) + mf * g) * R(1.0);
  R cm = (cs[0] + cs[1] + cs[2]) / R(3.0);
  R ce[3];
  for (int i = 0; i < 3; ++i) ce[i] = (cs[i] - cm * R(1.0)) / (R(2.0) * mu) + cm / (R(2.0) * mu + R(3.0) * lam) * R(1.0);
  M3<R> D = zero_m<R>();
  D.a[0] = std::exp(ce[0]); D.a[4] = std::exp(ce[1]); D.a[8] = std::exp(ce[2]);
  Fe = mul(mul(U, D), transpose(V));
  if (a == R(0.0)) Fe = scale(std::pow(det(trial) / det(Fe), R(1.0) / R(3.0)), Fe);
  M3<R> S = zero_m<R>();
  S.a[0] = cs[0]; S.a[4] = cs[1]; S.a[8] = cs[2];
  stress = mul(mul(U, S), transpose(U));
  residual = R(0.0);
  if (q2 + p2 > R(1.0))
    residual = std::fabs(g + R(1.0) - q2 / std::pow(R(1.0) + df * g, R(2.0)) - p2 / std::pow(R(1.0) + mf * g, R(2.0)));
}

template <typename R> void svd_of_trial(const M3<R>& trial, M3<R>& U, R s[3], M3<R>& V, double sd[3]) {
  M3<double> Ud, Vd;
  svd3_jacobi(cast_m<double>(trial), Ud, sd, Vd);
  U = cast_m<R>(Ud);
  V = cast_m<R>(Vd);
  for (int i = 0; i < 3; ++i) s[i] = R(sd[i]);
}

}  // namespace

static const double kGasConstant = 8.31446261815324;  // physics/materials/gas.py:8

template <typename R> static inline double gas_pressure(double moles, R T, R J, R n) {
  // physics/materials/gas.py:11-13
  return moles * kGasConstant * double(T) / (double(J) * double(n));
}

// ============================================================================ exported kernels
template <typename R>
static void ParticleToGrid(int n, const R* x, const R* v, const R* C, const R* tau, const R* m, const R* V0, R* node_m,
                           R* node_p, const R* origin, R dx, int res, R dt, int* flags) {
  // physics/transfer.py:36-70
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    R g[3];
    grid_position(x + 3 * p, origin, dx, g);
    int b[3];
    stencil_base(g, b);
    if (!valid_stencil(b, res)) { raise_flag(flags, 1); continue; }
    R l[3] = {g[0] - R(b[0]), g[1] - R(b[1]), g[2] - R(b[2])};
    R mass = m[p];
    R s = dt * V0[p] * R(4.0) / (dx * dx);
    M3<R> A = sub(scale(mass, load_m(C + 9 * p)), scale(s, load_m(tau + 9 * p)));
    V3<R> vel = load_v(v + 3 * p);
    for (int ox = 0; ox < 3; ++ox)
      for (int oy = 0; oy < 3; ++oy)
        for (int oz = 0; oz < 3; ++oz) {
          R w = quadratic_weight(l[0], ox) * quadratic_weight(l[1], oy) * quadratic_weight(l[2], oz);
          int64_t node = flatten_node(b[0] + ox, b[1] + oy, b[2] + oz, res);
          V3<R> d = {{(R(ox) - l[0]) * dx, (R(oy) - l[1]) * dx, (R(oz) - l[2]) * dx}};
          V3<R> Ad = mulv(A, d);
          ORACLE_ATOMIC node_m[node] += w * mass;
          for (int a = 0; a < 3; ++a) {
            ORACLE_ATOMIC node_p[3 * node + a] += w * (mass * vel.a[a] + Ad.a[a]);
          }
        }
  }
}

template <typename R>
static void UpdateGrid(int64_t count, const R* node_m, const R* node_p, R* node_v, int res, int bc, const R* gravity, R dt) {
  // physics/grid.py:20-38
  ORACLE_FOR
  for (int64_t i = 0; i < count; ++i) {
    if (!(node_m[i] > R(0.0))) continue;
    R vel[3];
    for (int a = 0; a < 3; ++a) vel[a] = node_p[3 * i + a] / node_m[i] + dt * gravity[a];
    int64_t c[3] = {i / (int64_t(res) * res), (i / res) % res, i % res};
    for (int a = 0; a < 3; ++a) {
      if (c[a] < bc && vel[a] < R(0.0)) vel[a] = R(0.0);
      if (c[a] >= res - bc && vel[a] > R(0.0)) vel[a] = R(0.0);
    }
    for (int a = 0; a < 3; ++a) node_v[3 * i + a] = vel[a];
  }
}

template <typename R>
static void GridToParticle(int n, const R* x_in, const R* F_in, const R* node_v, R* x_out, R* v_out, R* C_out, R* F_out,
                           const R* origin, R dx, int res, R dt, int* flags) {
  // physics/transfer.py:73-109
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    R g[3];
    grid_position(x_in + 3 * p, origin, dx, g);
    int b[3];
    stencil_base(g, b);
    if (!valid_stencil(b, res)) { raise_flag(flags, 1); continue; }
    R l[3] = {g[0] - R(b[0]), g[1] - R(b[1]), g[2] - R(b[2])};
    V3<R> nv = {{0, 0, 0}};
    M3<R> nc = zero_m<R>();
    for (int ox = 0; ox < 3; ++ox)
      for (int oy = 0; oy < 3; ++oy)
        for (int oz = 0; oz < 3; ++oz) {
          R w = quadratic_weight(l[0], ox) * quadratic_weight(l[1], oy) * quadratic_weight(l[2], oz);
          int64_t node = flatten_node(b[0] + ox, b[1] + oy, b[2] + oz, res);
          R d[3] = {(R(ox) - l[0]) * dx, (R(oy) - l[1]) * dx, (R(oz) - l[2]) * dx};
          const R* vi = node_v + 3 * node;
          for (int a = 0; a < 3; ++a) nv.a[a] = nv.a[a] + w * vi[a];
          R s = R(4.0) / (dx * dx) * w;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) nc.a[i * 3 + j] = nc.a[i * 3 + j] + s * (vi[i] * d[j]);
        }
    store_v(v_out + 3 * p, nv);
    store_m(C_out + 9 * p, nc);
    for (int a = 0; a < 3; ++a) x_out[3 * p + a] = x_in[3 * p + a] + dt * nv.a[a];
    store_m(F_out + 9 * p, mul(add(ident<R>(), scale(dt, nc)), load_m(F_in + 9 * p)));
  }
}

template <typename R>
static void ComputeElasticStress(int n, const R* F, R* tau, R mu, R lam, int* flags) {
  // physics/materials/elastic.py:4-26
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    M3<R> f = load_m(F + 9 * p);
    R J = det(f);
    if (!std::isfinite(J) || J <= R(0.0)) { raise_flag(flags, 0); store_m(tau + 9 * p, zero_m<R>()); continue; }
    M3<R> t = add(scale(mu, sub(mul(f, transpose(f)), ident<R>())), scale(lam * std::log(J), ident<R>()));
    store_m(tau + 9 * p, t);
  }
}

template <typename R>
static void UpdateDing(int n, const R* old_Fe, const R* old_n, const int* old_gelled, const R* F_total, const R* C_aff, R* new_Fe,
                       R* new_Fp, R* new_tau, R* residuals, R* new_n, R* temps, int* new_gelled, R* yields, R* viscs,
                       const R* visc_param, int* flags, R* por_res, int coupled, R mu, R lam, R y0, R n0, int temp_on,
                       const R* knots, const R* yf, const R* vf, R dt, const R* local_T) {
  // physics/materials/ding.py:140-221
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    R porosity = old_n[p];
    if (!std::isfinite(porosity) || porosity < R(0.0) || porosity >= R(1.0)) { raise_flag(flags, 0); continue; }
    int gel = old_gelled[p];
    R T = local_T[p];
    int reached = int(T >= knots[3]);
    R Teff = T;
    R yv = y0;
    R ev = visc_param[0];
    if (temp_on) {
      if (reached) gel = 1;
      if (gel) Teff = knots[3];
      yv = y0 * temperature_factor(Teff, knots, yf);
      ev = visc_param[0] * temperature_factor(Teff, knots, vf);
    }
    M3<R> trial = mul(add(ident<R>(), scale(dt, load_m(C_aff + 9 * p))), load_m(old_Fe + 9 * p));
    M3<R> U, V;
    R s[3];
    double sd[3];
    svd_of_trial(trial, U, s, V, sd);
    R ret_n = porosity;
    R cres = R(0.0);
    if (coupled) {
      R eps[3] = {std::log(R(sd[0])), std::log(R(sd[1])), std::log(R(sd[2]))};
      R ts[3];
      for (int i = 0; i < 3; ++i) ts[i] = R(2.0) * mu * eps[i] + lam * (eps[0] + eps[1] + eps[2]) * R(1.0);
      R mean = (ts[0] + ts[1] + ts[2]) / R(3.0);
      R dev[3] = {ts[0] - mean * R(1.0), ts[1] - mean * R(1.0), ts[2] - mean * R(1.0)};
      R q = R(1.5) * (dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]) / (yv * yv);
      R pv = mean * mean / (yv * yv);
      R df = R(6.0) * mu * dt * yv * yv / ev;
      R ms = R(2.0) / R(3.0) * (R(2.0) * mu + R(3.0) * lam) * dt * yv * yv / ev;
      R ex = R(3.0) * mean / (R(2.0) * mu + R(3.0) * lam);
      double opv = 1.0 - double(n0) + double(porosity);
      ret_n = solve_porosity<R>(q, pv, df, ms, ex, opv, n0);
      double pred = porosity_prediction<R>(double(ret_n), q, pv, df, ms, ex, opv, n0);
      cres = R(std::fabs(double(ret_n) - pred));
      if (cres > R(1.0e-6) || ret_n <= R(0.0) || ret_n >= R(1.0)) raise_flag(flags, 1);
    }
    M3<R> Fe, tau;
    R resid;
    ding_return<R>(trial, mu, lam, yv, ev, porosity_coefficient(ret_n), dt, &U, s, &V, Fe, tau, resid);
    M3<R> Ft = load_m(F_total + 9 * p);
    if (coupled) {
      double target = 1.0 - double(n0) + double(ret_n);
      double vs = det(cast_m<double>(Ft)) / (target * det(cast_m<double>(Fe)));
      Fe = scale(R(std::pow(vs, 1.0 / 3.0)), Fe);
    }
    M3<R> Fp = mul(inverse(Fe), Ft);
    R nn = det(Fp) - R(1.0) + n0;
    if (coupled) nn = ret_n;
    if (n0 == R(0.0) && porosity == R(0.0)) nn = R(0.0);
    if (!std::isfinite(nn) || nn < R(0.0) || nn >= R(1.0)) raise_flag(flags, 0);
    store_m(new_Fe + 9 * p, Fe);
    store_m(new_Fp + 9 * p, Fp);
    store_m(new_tau + 9 * p, tau);
    residuals[p] = resid;
    por_res[p] = cres;
    new_n[p] = nn;
    temps[p] = T;
    new_gelled[p] = gel;
    yields[p] = yv;
    viscs[p] = ev;
  }
}

template <typename R>
static void CombineGasStress(int n, const R* solid, const R* F, const R* por, const double* moles, const R* T, R* total,
                             int* flags, R p_amb, R n0, const R* density, R emod, R dt, R dx, R cfl) {
  // physics/materials/gas.py:16-35
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    R J = det(load_m(F + 9 * p));
    R nv = por[p];
    if (nv <= R(0.0) || nv >= R(1.0) || !std::isfinite(nv) || J <= R(0.0)) { raise_flag(flags, 0); continue; }
    R pr = R(gas_pressure<R>(moles[p], T[p], J, nv));
    store_m(total + 9 * p, add(load_m(solid + 9 * p), scale(J * (p_amb - pr), ident<R>())));
    R bulk = pr * (R(1.0) + (R(1.0) - n0 + nv) / nv);
    R wave = std::sqrt((emod + bulk) * J / density[p]);
    if (!std::isfinite(pr) || wave * dt / dx > cfl) raise_flag(flags, 2);
  }
}

template <typename R>
static void RefreshGas(int n, const double* old_moles, double* new_moles, const R* F, const R* por, double* pres, const R* T) {
  // physics/materials/gas.py:38-45
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    new_moles[p] = old_moles[p];
    if (por[p] > R(0.0) && por[p] < R(1.0)) pres[p] = gas_pressure<R>(old_moles[p], T[p], det(load_m(F + 9 * p)), por[p]);
  }
}

template <typename R>
static void SurfaceAreas(int nf, const R* verts, const int* faces, const int* owners, double* exposed, double* contact, R floor_h,
                         R cdist) {
  // physics/thermal_water.py:77-92
  ORACLE_FOR
  for (int f = 0; f < nf; ++f) {
    double P[3][3];
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) P[k][a] = double(verts[3 * faces[3 * f + k] + a]);
    double e1[3], e2[3];
    for (int a = 0; a < 3; ++a) { e1[a] = P[1][a] - P[0][a]; e2[a] = P[2][a] - P[0][a]; }
    double nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    double area = 0.5 * len;
    double cy = (P[0][1] + P[1][1] + P[2][1]) / 3.0;
    if (cy <= double(floor_h + cdist) && nrm[1] < -0.2 * len) {
      ORACLE_ATOMIC contact[owners[f]] += area;
    } else {
      ORACLE_ATOMIC exposed[owners[f]] += area;
    }
  }
}

template <typename R>
static void ScalarToGrid(int n, const R* x, const R* F, const R* V0, const double* cap, const double* T, const double* W,
                         const double* exposed, const double* contact, const R* heat_param, R T_oven, R T_tray, R tray_h,
                         R loss_coef, R rho_w, R latent, int water_on, double* nV, double* nC, double* nE, double* nW, double* nH,
                         double* nL, double* nX, double* p_heat, double* p_loss, const R* origin, R dx, int res) {
  // physics/thermal_water.py:95-137
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    double vol = double(V0[p]) * double(det(load_m(F + 9 * p)));
    double t = T[p];
    double s = W[p] / (double(rho_w) * vol);
    double loss = 0.0;
    if (water_on) loss = double(loss_coef) * t * s * exposed[p];
    double heat = double(heat_param[0]) * exposed[p] * (double(T_oven) - t) + double(tray_h) * contact[p] * (double(T_tray) - t) -
                  double(latent) * loss;
    double exch = double(heat_param[0]) * exposed[p] + double(tray_h) * contact[p];
    if (water_on) exch += double(R(latent * loss_coef)) * s * exposed[p];
    p_heat[p] = heat;
    p_loss[p] = loss;
    R g[3];
    grid_position(x + 3 * p, origin, dx, g);
    int b[3];
    stencil_base(g, b);
    if (!valid_stencil(b, res)) continue;  // the reference writes unguarded; valid runs never reach this
    R l[3] = {g[0] - R(b[0]), g[1] - R(b[1]), g[2] - R(b[2])};
    for (int ox = 0; ox < 3; ++ox)
      for (int oy = 0; oy < 3; ++oy)
        for (int oz = 0; oz < 3; ++oz) {
          double w = double(quadratic_weight(l[0], ox) * quadratic_weight(l[1], oy) * quadratic_weight(l[2], oz));
          int64_t node = flatten_node(b[0] + ox, b[1] + oy, b[2] + oz, res);
          ORACLE_ATOMIC nV[node] += w * vol;
          ORACLE_ATOMIC nC[node] += w * cap[p];
          ORACLE_ATOMIC nE[node] += w * cap[p] * t;
          ORACLE_ATOMIC nW[node] += w * W[p];
          ORACLE_ATOMIC nH[node] += w * heat;
          ORACLE_ATOMIC nL[node] += w * loss;
          ORACLE_ATOMIC nX[node] += w * exch;
        }
  }
}

static void GridFields(int64_t count, const double* V, const double* C, const double* E, const double* W, double* T, double* S,
                       double rho_w) {
  // physics/thermal_water.py:140-147
  ORACLE_FOR
  for (int64_t i = 0; i < count; ++i)
    if (C[i] > 0.0) { T[i] = E[i] / C[i]; S[i] = W[i] / (rho_w * V[i]); }
}

template <typename R>
static void DiffuseGrid(int64_t count, const double* V, const double* C, const double* T, const double* S, const double* W,
                        const double* bh, const double* bl, const double* bx, double* dT, double* dW, int* errors, R kappa, R diff,
                        R rho_w, int water_on, R dx, int res, R dt) {
  // physics/thermal_water.py:150-190
  ORACLE_FOR
  for (int64_t i = 0; i < count; ++i) {
    if (C[i] <= 0.0) continue;
    int64_t c[3] = {i / (int64_t(res) * res), (i / res) % res, i % res};
    double hr = bh[i], wr = -bl[i], hd = bx[i], wd = 0.0;
    for (int ax = 0; ax < 3; ++ax)
      for (int side = 0; side < 2; ++side) {
        int64_t o[3] = {c[0], c[1], c[2]};
        o[ax] += 2 * side - 1;
        if (o[ax] >= 0 && o[ax] < res) {
          int64_t j = flatten_node(int(o[0]), int(o[1]), int(o[2]), res);
          if (C[j] > 0.0) {
            double fill = 2.0 * V[i] * V[j] / (V[i] + V[j]) / double(dx * dx * dx);
            double hc = double(kappa * dx) * fill;
            hr += hc * (T[j] - T[i]);
            hd += hc;
            if (water_on) {
              double mob = double(diff * rho_w * dx) * fill * 0.5 * (T[i] + T[j]) * 0.5 * (S[i] + S[j]);
              wr += mob * (S[j] - S[i]);
              wd += mob / (double(rho_w) * V[i]);
            }
          }
        }
      }
    if (double(dt) * hd > C[i]) raise_flag(errors, 0);
    double change = double(dt) * wr;
    if (W[i] + change < 0.0 || double(dt) * wd > 1.0) raise_flag(errors, 1);
    dT[i] = double(dt) * hr / C[i];
    dW[i] = change;
  }
}

template <typename R>
static void ScalarToParticles(int n, const R* x, const double* T_old, const double* W_old, const double* L_old, const double* X_old,
                              const double* dry, const R* V0, const R* F, const double* nW, const double* dT, const double* dW,
                              const double* p_heat, const double* p_loss, double* T_new, R* T_mat, double* W_new, double* L_new,
                              double* X_new, R* m_new, R* rho_ref, double* s_new, R rho_w, const R* origin, R dx, int res, R dt) {
  // physics/thermal_water.py:193-229
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    R g[3];
    grid_position(x + 3 * p, origin, dx, g);
    int b[3];
    stencil_base(g, b);
    double tc = 0.0, mr = 0.0;
    if (valid_stencil(b, res)) {  // the reference reads unguarded; valid runs always pass
      R l[3] = {g[0] - R(b[0]), g[1] - R(b[1]), g[2] - R(b[2])};
      for (int ox = 0; ox < 3; ++ox)
        for (int oy = 0; oy < 3; ++oy)
          for (int oz = 0; oz < 3; ++oz) {
            double w = double(quadratic_weight(l[0], ox) * quadratic_weight(l[1], oy) * quadratic_weight(l[2], oz));
            int64_t node = flatten_node(b[0] + ox, b[1] + oy, b[2] + oz, res);
            tc += w * dT[node];
            if (nW[node] > 0.0) mr += w * dW[node] / nW[node];
          }
    }
    double t = T_old[p] + tc;
    double mass = W_old[p] * (1.0 + mr);
    T_new[p] = t;
    T_mat[p] = R(t);
    W_new[p] = mass;
    L_new[p] = L_old[p] + double(dt) * p_loss[p];
    X_new[p] = X_old[p] + double(dt) * p_heat[p];
    m_new[p] = R(dry[p] + mass);
    rho_ref[p] = R((dry[p] + mass) / double(V0[p]));
    s_new[p] = mass / (double(rho_w) * double(V0[p]) * double(det(load_m(F + 9 * p))));
  }
}

template <typename R>
static void LocalGas(int n, const double* old_moles, const double* T_old, const double* T_new, double* new_moles, double* pres,
                     const R* F, const R* por, const R* param, R density, R peak, R width, R dt, int produce) {
  // physics/thermal_water.py:232-245
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    double moles = old_moles[p];
    if (produce) {
      double t = 0.5 * (T_old[p] + T_new[p]);
      moles += double(dt * density) * double(param[0]) * std::exp(-std::pow((t - double(peak)) / double(width), 2.0));
    }
    new_moles[p] = moles;
    if (por[p] > R(0.0)) pres[p] = gas_pressure<R>(moles, R(T_new[p]), det(load_m(F + 9 * p)), por[p]);
  }
}

template <typename R>
static void WaterStress(int n, const R* F, const double* sat, const R* old_s, R* new_s, R s0, R stiff) {
  // physics/thermal_water.py:248-253
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    R J = det(load_m(F + 9 * p));
    R k = J * stiff * (R(sat[p]) - s0);
    store_m(new_s + 9 * p, sub(load_m(old_s + 9 * p), scale(k, ident<R>())));
  }
}

template <typename R>
static void Saturations(int n, const double* W, const R* V0, const R* F, double* sat, R rho_w) {
  // physics/thermal_water.py:256-260
  ORACLE_FOR
  for (int p = 0; p < n; ++p) sat[p] = W[p] / (double(rho_w) * double(V0[p]) * double(det(load_m(F + 9 * p))));
}

template <typename R> inline R frobenius_sq(const M3<R>& m) {
  R s = R(0);
  for (int i = 0; i < 9; ++i) s = s + m.a[i] * m.a[i];
  return s;
}

template <typename R>
static void SettlingMetrics(int n, const R* v, const R* A, const R* old_p, const R* new_p, R* metrics, R dt) {
  // physics/quasistatic.py:11-21: running maxima of |v|, |sym C|_F and |Fp - Fp_ref|_F / (dt |Fp_ref|_F);
  // metrics[] keeps the caller's starting values (zero in the reference) like its atomic_max targets.
  for (int p = 0; p < n; ++p) {
    const M3<R> a = load_m(A + 9 * p), o = load_m(old_p + 9 * p), w = load_m(new_p + 9 * p);
    M3<R> s;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) s.a[i * 3 + j] = R(0.5) * (a.a[i * 3 + j] + a.a[j * 3 + i]);
    const R speed = std::sqrt(v[3 * p] * v[3 * p] + v[3 * p + 1] * v[3 * p + 1] + v[3 * p + 2] * v[3 * p + 2]);
    const R strain = std::sqrt(frobenius_sq(s));
    const R plastic = std::sqrt(frobenius_sq(sub(w, o))) / (dt * std::sqrt(frobenius_sq(o)));
    metrics[0] = std::max(metrics[0], speed);
    metrics[1] = std::max(metrics[1], strain);
    metrics[2] = std::max(metrics[2], plastic);
  }
}

template <typename R>
static void ReconstructMeshVertices(int nv, const R* x, const R* F, const R* ref_verts, const int* offs, const int* parts,
                                    const R* weights, const R* ref_offs, R* out) {
  // geometry/mesh_tracking.py:13-38
  ORACLE_FOR
  for (int v = 0; v < nv; ++v) {
    int s = offs[v], e = offs[v + 1];
    if (s == e) { for (int a = 0; a < 3; ++a) out[3 * v + a] = ref_verts[3 * v + a]; continue; }
    int ap = parts[s];
    V3<R> anchor = mulv(load_m(F + 9 * ap), load_v(ref_offs + 3 * s));
    for (int a = 0; a < 3; ++a) anchor.a[a] = x[3 * ap + a] + anchor.a[a];
    R disp[3] = {0, 0, 0};
    for (int k = s; k < e; ++k) {
      int pi = parts[k];
      V3<R> pr = mulv(load_m(F + 9 * pi), load_v(ref_offs + 3 * k));
      for (int a = 0; a < 3; ++a) disp[a] = disp[a] + weights[k] * ((x[3 * pi + a] + pr.a[a]) - anchor.a[a]);
    }
    for (int a = 0; a < 3; ++a) out[3 * v + a] = anchor.a[a] + disp[a];
  }
}

template <typename R>
static void CellKeys(int n, const R* x, const R* origin, R dx, int res, int tiled, uint32_t* keys, int* flags) {
  // Cell key of each particle: flatten_node(stencil_base((x - origin) / dx)) (physics/transfer.py:16-33),
  // or, with tiled != 0, the same cell numbered block-major over 4x4x4 cell blocks.
  int nb = (res + 3) / 4;
  ORACLE_FOR
  for (int p = 0; p < n; ++p) {
    R g[3];
    grid_position(x + 3 * p, origin, dx, g);
    int b[3];
    stencil_base(g, b);
    if (!valid_stencil(b, res)) { raise_flag(flags, 1); keys[p] = 0xffffffffu; continue; }
    if (!tiled) keys[p] = uint32_t(flatten_node(b[0], b[1], b[2], res));
    else {
      uint32_t blk = (uint32_t(b[0] >> 2) * nb + uint32_t(b[1] >> 2)) * nb + uint32_t(b[2] >> 2);
      keys[p] = (blk << 6) | uint32_t((b[0] & 3) << 4 | (b[1] & 3) << 2 | (b[2] & 3));
    }
  }
}

// ============================================================================ C ABI
#define ORACLE_BOTH(name, sig_f, sig_d, args) \
  extern "C" void oracle_##name##_f32 sig_f { name<float> args; } \
  extern "C" void oracle_##name##_f64 sig_d { name<double> args; }

extern "C" {

// Threads of the timing build's loops (1 in the checker build).
void oracle_set_threads(int threads) {
#ifdef ORACLE_OPENMP
  omp_set_num_threads(threads < 1 ? 1 : threads);
#else
  (void)threads;
#endif
}
int oracle_threads(void) {
#ifdef ORACLE_OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_svd3(const double* F, double* U, double* sigma, double* V) {
  M3<double> f = load_m(F), u, v;
  svd3_jacobi(f, u, sigma, v);
  store_m(U, u);
  store_m(V, v);
}

float oracle_porosity_coefficient(float n) { return porosity_coefficient<float>(n); }
float oracle_solve_overstress(float q2, float p2, float df, float mf) { return solve_overstress<float>(q2, p2, df, mf); }
double oracle_overstress64(double q, double p, double df, double mf) { return overstress64(q, p, df, mf); }
float oracle_solve_porosity(float q, float p, float df, float ms, float ex, double pv, float n0) {
  return solve_porosity<float>(q, p, df, ms, ex, pv, n0);
}
double oracle_porosity_prediction(double n, float q, float p, float df, float ms, float ex, double pv, float n0) {
  return porosity_prediction<float>(n, q, p, df, ms, ex, pv, n0);
}
float oracle_temperature_factor(float t, const float* knots, const float* factors) { return temperature_factor<float>(t, knots, factors); }
double oracle_gas_pressure(double moles, float T, float J, float n) { return gas_pressure<float>(moles, T, J, n); }

void oracle_GridFields(int64_t count, const double* V, const double* C, const double* E, const double* W, double* T, double* S,
                       double rho_w) {
  GridFields(count, V, C, E, W, T, S, rho_w);
}
}

ORACLE_BOTH(ParticleToGrid,
            (int n, const float* x, const float* v, const float* C, const float* tau, const float* m, const float* V0, float* nm,
             float* np_, const float* o, float dx, int res, float dt, int* fl),
            (int n, const double* x, const double* v, const double* C, const double* tau, const double* m, const double* V0,
             double* nm, double* np_, const double* o, double dx, int res, double dt, int* fl),
            (n, x, v, C, tau, m, V0, nm, np_, o, dx, res, dt, fl))
ORACLE_BOTH(UpdateGrid,
            (int64_t c, const float* nm, const float* np_, float* nv, int res, int bc, const float* g, float dt),
            (int64_t c, const double* nm, const double* np_, double* nv, int res, int bc, const double* g, double dt),
            (c, nm, np_, nv, res, bc, g, dt))
ORACLE_BOTH(GridToParticle,
            (int n, const float* x, const float* F, const float* nv, float* xo, float* vo, float* Co, float* Fo, const float* o,
             float dx, int res, float dt, int* fl),
            (int n, const double* x, const double* F, const double* nv, double* xo, double* vo, double* Co, double* Fo,
             const double* o, double dx, int res, double dt, int* fl),
            (n, x, F, nv, xo, vo, Co, Fo, o, dx, res, dt, fl))
ORACLE_BOTH(ComputeElasticStress, (int n, const float* F, float* tau, float mu, float lam, int* fl),
            (int n, const double* F, double* tau, double mu, double lam, int* fl), (n, F, tau, mu, lam, fl))
ORACLE_BOTH(UpdateDing,
            (int n, const float* oFe, const float* on, const int* og, const float* Ft, const float* Ca, float* nFe, float* nFp,
             float* nt, float* res, float* nn, float* temps, int* ng, float* ys, float* vs, const float* vp, int* fl, float* pr,
             int coupled, float mu, float lam, float y0, float n0, int ton, const float* kn, const float* yf, const float* vf,
             float dt, const float* lT),
            (int n, const double* oFe, const double* on, const int* og, const double* Ft, const double* Ca, double* nFe,
             double* nFp, double* nt, double* res, double* nn, double* temps, int* ng, double* ys, double* vs, const double* vp,
             int* fl, double* pr, int coupled, double mu, double lam, double y0, double n0, int ton, const double* kn,
             const double* yf, const double* vf, double dt, const double* lT),
            (n, oFe, on, og, Ft, Ca, nFe, nFp, nt, res, nn, temps, ng, ys, vs, vp, fl, pr, coupled, mu, lam, y0, n0, ton, kn, yf, vf,
             dt, lT))
ORACLE_BOTH(CombineGasStress,
            (int n, const float* so, const float* F, const float* por, const double* mo, const float* T, float* to, int* fl,
             float pa, float n0, const float* de, float em, float dt, float dx, float cfl),
            (int n, const double* so, const double* F, const double* por, const double* mo, const double* T, double* to, int* fl,
             double pa, double n0, const double* de, double em, double dt, double dx, double cfl),
            (n, so, F, por, mo, T, to, fl, pa, n0, de, em, dt, dx, cfl))
ORACLE_BOTH(RefreshGas, (int n, const double* om, double* nm, const float* F, const float* por, double* pr, const float* T),
            (int n, const double* om, double* nm, const double* F, const double* por, double* pr, const double* T),
            (n, om, nm, F, por, pr, T))
ORACLE_BOTH(SurfaceAreas,
            (int nf, const float* v, const int* f, const int* ow, double* ex, double* co, float fh, float cd),
            (int nf, const double* v, const int* f, const int* ow, double* ex, double* co, double fh, double cd),
            (nf, v, f, ow, ex, co, fh, cd))
ORACLE_BOTH(ScalarToGrid,
            (int n, const float* x, const float* F, const float* V0, const double* cap, const double* T, const double* W,
             const double* ex, const double* co, const float* hp, float To, float Tt, float th, float lc, float rw, float la,
             int won, double* nV, double* nC, double* nE, double* nW, double* nH, double* nL, double* nX, double* ph, double* pl,
             const float* o, float dx, int res),
            (int n, const double* x, const double* F, const double* V0, const double* cap, const double* T, const double* W,
             const double* ex, const double* co, const double* hp, double To, double Tt, double th, double lc, double rw,
             double la, int won, double* nV, double* nC, double* nE, double* nW, double* nH, double* nL, double* nX, double* ph,
             double* pl, const double* o, double dx, int res),
            (n, x, F, V0, cap, T, W, ex, co, hp, To, Tt, th, lc, rw, la, won, nV, nC, nE, nW, nH, nL, nX, ph, pl, o, dx, res))
ORACLE_BOTH(DiffuseGrid,
            (int64_t c, const double* V, const double* C, const double* T, const double* S, const double* W, const double* bh,
             const double* bl, const double* bx, double* dT, double* dW, int* er, float k, float d, float rw, int won, float dx,
             int res, float dt),
            (int64_t c, const double* V, const double* C, const double* T, const double* S, const double* W, const double* bh,
             const double* bl, const double* bx, double* dT, double* dW, int* er, double k, double d, double rw, int won,
             double dx, int res, double dt),
            (c, V, C, T, S, W, bh, bl, bx, dT, dW, er, k, d, rw, won, dx, res, dt))
ORACLE_BOTH(ScalarToParticles,
            (int n, const float* x, const double* To, const double* Wo, const double* Lo, const double* Xo, const double* dry,
             const float* V0, const float* F, const double* nW, const double* dT, const double* dW, const double* ph,
             const double* pl, double* Tn, float* Tm, double* Wn, double* Ln, double* Xn, float* mn, float* rr, double* sn,
             float rw, const float* o, float dx, int res, float dt),
            (int n, const double* x, const double* To, const double* Wo, const double* Lo, const double* Xo, const double* dry,
             const double* V0, const double* F, const double* nW, const double* dT, const double* dW, const double* ph,
             const double* pl, double* Tn, double* Tm, double* Wn, double* Ln, double* Xn, double* mn, double* rr, double* sn,
             double rw, const double* o, double dx, int res, double dt),
            (n, x, To, Wo, Lo, Xo, dry, V0, F, nW, dT, dW, ph, pl, Tn, Tm, Wn, Ln, Xn, mn, rr, sn, rw, o, dx, res, dt))
ORACLE_BOTH(LocalGas,
            (int n, const double* om, const double* To, const double* Tn, double* nm, double* pr, const float* F, const float* por,
             const float* pa, float de, float pk, float wi, float dt, int pro),
            (int n, const double* om, const double* To, const double* Tn, double* nm, double* pr, const double* F,
             const double* por, const double* pa, double de, double pk, double wi, double dt, int pro),
            (n, om, To, Tn, nm, pr, F, por, pa, de, pk, wi, dt, pro))
ORACLE_BOTH(WaterStress, (int n, const float* F, const double* s, const float* os, float* ns, float s0, float k),
            (int n, const double* F, const double* s, const double* os, double* ns, double s0, double k), (n, F, s, os, ns, s0, k))
ORACLE_BOTH(SettlingMetrics, (int n, const float* v, const float* A, const float* op, const float* np, float* m, float dt),
            (int n, const double* v, const double* A, const double* op, const double* np, double* m, double dt),
            (n, v, A, op, np, m, dt))
ORACLE_BOTH(Saturations, (int n, const double* W, const float* V0, const float* F, double* s, float rw),
            (int n, const double* W, const double* V0, const double* F, double* s, double rw), (n, W, V0, F, s, rw))
ORACLE_BOTH(ReconstructMeshVertices,
            (int nv, const float* x, const float* F, const float* rv, const int* of, const int* pa, const float* w, const float* ro,
             float* out),
            (int nv, const double* x, const double* F, const double* rv, const int* of, const int* pa, const double* w,
             const double* ro, double* out),
            (nv, x, F, rv, of, pa, w, ro, out))
ORACLE_BOTH(CellKeys, (int n, const float* x, const float* o, float dx, int res, int tiled, uint32_t* k, int* fl),
            (int n, const double* x, const double* o, double dx, int res, int tiled, uint32_t* k, int* fl),
            (n, x, o, dx, res, tiled, k, fl))
