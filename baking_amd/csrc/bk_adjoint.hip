// Adjoint (reverse-mode) kernels of one substep of the mechanical MLS-MPM
// path: G2P with the F update, the grid update, P2G, the neo-Hookean stress,
// and the Ding return map without pore gas (reference physics/transfer.py:36-109,
// grid.py:20-38, materials/elastic.py:4-10, materials/ding.py:77-221).
//
// Given the adjoint of the substep's output state (in the bk_adjoint arrays),
// the kernels replace it with the adjoint of its input state and add the
// gradient of the Ding viscosity parameter.  They read the forward values of
// that substep: the input and output states, the grid the forward left (mass,
// momentum, velocity) and the stress P2G used.  Adjoints are FP64.
//
// The return map's adjoint differentiates the map itself, not its iterations:
// the overstress by the implicit function theorem on its converged equation
// (as the reference's own adjoint, ding.py:90-100), the principal values by
// forward-mode dual numbers, and the rotation part by the closed forms of the
// derivative of an isotropic tensor function of the trial gradient, written
// with divided differences that stay finite (their limits) when two
// principal stretches coincide.
#include "bk_device.h"
#include "bk_workspace.h"
#include "../../include/baking_abi.h"

using namespace bk;

void bk_internal_set_error(const char* what);  // bk_last_error's text (bk_substep.hip)
int bk_internal_tile_launch(bk_workspace* ws, const bk_dough** doughs, int* n_doughs, int* grid_cap,
                            const bk_workspace** device_view);  // bk_substep.hip

namespace {
int fail(const char* what) {
  bk_internal_set_error(what);
  return -1;
}
inline int div_up(long long a, long long b) { return int((a + b - 1) / b); }
}  // namespace

// ------------------------------------------------------------------ dual numbers
// Value and derivatives along 6 directions: the 3 principal stretches, the
// old porosity, the viscosity and the yield stress.
constexpr int kDir = 6;
struct Dual {
  double v;
  double d[kDir];
};
__device__ __forceinline__ Dual dconst(double v) {
  Dual r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < kDir; ++k) r.d[k] = 0.0;
  return r;
}
__device__ __forceinline__ Dual dvar(double v, int k) {
  Dual r = dconst(v);
  r.d[k] = 1.0;
  return r;
}
__device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < kDir; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < kDir; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < kDir; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator*(double s, const Dual& a) {
  Dual r;
  r.v = s * a.v;
#pragma unroll
  for (int k = 0; k < kDir; ++k) r.d[k] = s * a.d[k];
  return r;
}
__device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
  Dual r;
  const double inv = 1.0 / b.v;
  r.v = a.v * inv;
#pragma unroll
  for (int k = 0; k < kDir; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
  return r;
}
__device__ __forceinline__ Dual dlog(const Dual& a) {
  Dual r;
  r.v = log(a.v);
  const double inv = 1.0 / a.v;
#pragma unroll
  for (int k = 0; k < kDir; ++k) r.d[k] = a.d[k] * inv;
  return r;
}
__device__ __forceinline__ Dual dexp(const Dual& a) {
  Dual r;
  r.v = exp(a.v);
#pragma unroll
  for (int k = 0; k < kDir; ++k) r.d[k] = r.v * a.d[k];
  return r;
}

// ------------------------------------------------------------------ coupled porosity root (porous_return.py:64-100)
// Derivatives of the converged end-state porosity n (the root of
// n = prediction(n), porous_return.py:34-40) with respect to q, p,
// dev_factor, mean_scale, exponent and plastic_volume (initial_porosity: 1/R'),
// by the implicit function theorem on that equation, as the reference's adjoint.
struct PorosityDerivs {
  double q, p, df, ms, ex, pv, n0;
};
// Overstress of the converged Newton iteration on R(g) = g + 1 - q2/(1+df g)^2 - p2/(1+mf g)^2
// (0 inside the yield surface), stopped one step after |R| falls below 1e-13 of its scale.
__device__ inline double overstress_converged(double q2, double p2, double df, double mf) {
  double root = 0.0;
  if (q2 + p2 > 1.0) {
    const double tol = 1e-13 * (1.0 + q2 + p2);
    for (int it = 0; it < 64; ++it) {
      const double u = 1.0 + df * root, w = 1.0 + mf * root;
      const double iu = 1.0 / u, iw = 1.0 / w;
      const double res = root + 1.0 - q2 * iu * iu - p2 * iw * iw;
      const double der = 1.0 + 2.0 * df * q2 * iu * iu * iu + 2.0 * mf * p2 * iw * iw * iw;
      root -= res / der;
      if (fabs(res) < tol) break;
    }
  }
  return root;
}

// a(n) (ding.py:68-74, porous_return.py:11-16) and its derivative.
__device__ __forceinline__ void coefficient_and_slope(double n, double& a, double& da) {
  a = pressure_coefficient64(n);
  da = 0.0;
  if (n > 0.0) {
    const double lg = log(1.0 / n);
    da = a * (2.0 / (n * lg) - 3.5 / (3.5 * n + 1.0));
  }
}

// g = the overstress at the root (overstress_converged(q, a p, d, h a)).
__device__ inline PorosityDerivs porosity_root_derivatives(double n, double g, double q, double p, double d, double h, double ex,
                                                           double pv) {
  double a, da;
  coefficient_and_slope(n, a, da);
  const double u = 1.0 + d * g, v = 1.0 + h * a * g;
  const double der = 1.0 + 2.0 * d * q / (u * u * u) + 2.0 * h * a * a * p / (v * v * v);
  double r_q = 0.0, r_p = 0.0, r_d = 0.0, r_h = 0.0, r_a = 0.0;
  if (q + a * p > 1.0) {
    r_q = 1.0 / (der * u * u);
    r_p = a / (der * v * v);
    r_d = -2.0 * q * g / (der * u * u * u);
    r_h = -2.0 * a * a * p * g / (der * v * v * v);
    r_a = p * (1.0 - h * a * g) / (der * v * v * v);
  }
  const double expo = exp(ex * (1.0 - 1.0 / v));
  const double nv = pv * expo;
  const double pref = nv * ex / (v * v);
  const double s = 1.0 / (1.0 - pref * h * da * (g + a * r_a));
  PorosityDerivs D;
  D.q = s * pref * h * a * r_q;
  D.p = s * pref * h * a * r_p;
  D.df = s * pref * h * a * r_d;
  D.ms = s * pref * (a * g + h * a * r_h);
  D.ex = s * nv * (1.0 - 1.0 / v);
  D.pv = s * expo;
  D.n0 = s;
  return D;
}

// ------------------------------------------------------------------ principal part of the return map
// From the principal stretches s, the old porosity and the viscosity eta: the
// corrected principal Kirchhoff stresses cs, the returned principal stretches
// e and the porosity n the map used (ding.py:103-131, 183-192), with their
// derivatives along (s0, s1, s2, old porosity, eta, yield stress y).  Non-coupled, n is the
// old porosity; coupled, the end-state root the forward solved
// (porous_return.py:43-61, its output porosity), with the derivatives above.
struct Principal {
  Dual cs[3], e[3], n;
};
__device__ inline Principal principal_return(const double s[3], double n_old, double n_root, double eta, double mu, double lam, double y,
                                             double dt, bool coupled, float n0) {
  Dual sv[3] = {dvar(s[0], 0), dvar(s[1], 1), dvar(s[2], 2)};
  const Dual ev = dvar(eta, 4);
  Dual eps[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) eps[i] = dlog(sv[i]);
  const Dual tr = eps[0] + eps[1] + eps[2];
  Dual ts[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) ts[i] = 2.0 * mu * eps[i] + lam * tr;
  const Dual mean = (1.0 / 3.0) * (ts[0] + ts[1] + ts[2]);
  Dual dev[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) dev[i] = ts[i] - mean;
  const Dual yd = dvar(y, 5);
  const Dual y2 = yd * yd;
  const Dual iy2 = dconst(1.0) / y2;
  const Dual q2 = 1.5 * ((dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]) * iy2);
  const Dual pm = (mean * mean) * iy2;
  const Dual rev = dconst(1.0) / ev;
  const Dual df = (6.0 * mu * dt) * (y2 * rev);
  const Dual ms = ((2.0 / 3.0) * (2.0 * mu + 3.0 * lam) * dt) * (y2 * rev);
  Principal P;
  P.n = dvar(coupled ? n_root : n_old, 3);
  if (coupled) P.n.d[3] = 0.0;
  double av, da;
  coefficient_and_slope(P.n.v, av, da);
  // one overstress solve serves the return map and, coupled, the porosity root's derivatives
  const double gv = overstress_converged(q2.v, av * pm.v, df.v, ms.v * av);
  if (coupled) {
    const Dual ex = (3.0 / (2.0 * mu + 3.0 * lam)) * mean;
    const double opv = 1.0 - double(n0) + n_old;
    const PorosityDerivs D = porosity_root_derivatives(P.n.v, gv, q2.v, pm.v, df.v, ms.v, ex.v, opv);
#pragma unroll
    for (int k = 0; k < kDir; ++k)
      P.n.d[k] = D.q * q2.d[k] + D.p * pm.d[k] + D.df * df.d[k] + D.ms * ms.d[k] + D.ex * ex.d[k] + (k == 3 ? D.pv : 0.0);
  }
  Dual ad = dconst(av);
#pragma unroll
  for (int k = 0; k < kDir; ++k) ad.d[k] = da * P.n.d[k];
  const Dual p2 = ad * pm;
  const Dual mf = ms * ad;
  // overstress derivative by the IFT on its converged equation (ding.py:90-100)
  Dual g = dconst(0.0);
  if (q2.v + p2.v > 1.0) {
    const double u = 1.0 + df.v * gv, w = 1.0 + mf.v * gv;
    const double iu = 1.0 / u, iw = 1.0 / w;
    const double der = 1.0 + 2.0 * df.v * q2.v * iu * iu * iu + 2.0 * mf.v * p2.v * iw * iw * iw;
    const double ider = 1.0 / der;
    g.v = gv;
#pragma unroll
    for (int k = 0; k < kDir; ++k) {
      const double dR = -q2.d[k] * iu * iu - p2.d[k] * iw * iw + 2.0 * q2.v * gv * df.d[k] * iu * iu * iu +
                        2.0 * p2.v * gv * mf.d[k] * iw * iw * iw;
      g.d[k] = -dR * ider;
    }
  }
  const Dual one = dconst(1.0);
  const Dual rdden = one / (one + df * g), rmden = one / (one + mf * g);
#pragma unroll
  for (int i = 0; i < 3; ++i) P.cs[i] = dev[i] * rdden + mean * rmden;
  const Dual cm = (1.0 / 3.0) * (P.cs[0] + P.cs[1] + P.cs[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i) P.e[i] = dexp((1.0 / (2.0 * mu)) * (P.cs[i] - cm) + (1.0 / (2.0 * mu + 3.0 * lam)) * cm);
  return P;
}

// Slope of the piecewise-linear temperature factor (ding.py:56-65) at t: 0 below the
// first knot and from the last knot on.
__device__ __forceinline__ double temperature_slope(double t, const float k[4], const float f[4]) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (t >= double(k[i]) && t < double(k[i + 1])) s = (double(f[i + 1]) - double(f[i])) / (double(k[i + 1]) - double(k[i]));
  return s;
}

// Divided difference (f_i - f_j) / (s_i - s_j) of a symmetric function of the
// principal values, or its limit df_i/ds_i - df_i/ds_j when they coincide.
__device__ __forceinline__ double divided(const Dual* f, const double* s, int i, int j) {
  const double ds = s[i] - s[j];
  if (fabs(ds) > 1e-6 * fmax(fabs(s[i]), fabs(s[j]))) return (f[i].v - f[j].v) / ds;
  return f[i].d[i] - f[i].d[j];
}

typedef double dmat[9];
__device__ __forceinline__ void mmul(const double* A, const double* B, double* R) {  // R = A B
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) R[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}
__device__ __forceinline__ void mmul_tn(const double* A, const double* B, double* R) {  // R = A^T B
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) R[i * 3 + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
__device__ __forceinline__ void mmul_nt(const double* A, const double* B, double* R) {  // R = A B^T
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) R[i * 3 + j] = A[i * 3] * B[j * 3] + A[i * 3 + 1] * B[j * 3 + 1] + A[i * 3 + 2] * B[j * 3 + 2];
}
__device__ __forceinline__ double mdet(const double* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}
// inverse transpose
__device__ __forceinline__ void minv_t(const double* a, double* r) {
  const double d = mdet(a);
  r[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  r[3] = (a[2] * a[7] - a[1] * a[8]) / d;
  r[6] = (a[1] * a[5] - a[2] * a[4]) / d;
  r[1] = (a[5] * a[6] - a[3] * a[8]) / d;
  r[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  r[7] = (a[2] * a[3] - a[0] * a[5]) / d;
  r[2] = (a[3] * a[7] - a[4] * a[6]) / d;
  r[5] = (a[1] * a[6] - a[0] * a[7]) / d;
  r[8] = (a[0] * a[4] - a[1] * a[3]) / d;
}

// Adjoint of the Ding return map (ding.py:140-221).  Outputs Fe, tau and the
// porosity, with adjoints gFe, gtau, gn.  Non-coupled: the porosity output is
// det(Fe^-1 F) - 1 + n0 (constant 0 when n0 = 0 and the old porosity is 0) and
// Fe takes the isochoric fit when a = 0; coupled: the porosity output is the
// end-state root and Fe is rescaled to det F / (1 - n0 + n).  Adds to gFe_old,
// gF (the new total gradient), gC (the new affine velocity), gporosity (the
// old porosity) and geta0 (the viscosity parameter).
__device__ inline void ding_return_adjoint(const DingParams& P, const double* Fe_old, const double* F, const double* C, double porosity,
                                           double n_root, int gelled_in, double T, double visc_param, const double* gFe_in,
                                           const double* gtau, double gn, double* gFe_old, double* gF, double* gC, double& gporosity,
                                           double& geta0, double& gT) {
  int gel = gelled_in;
  double Teff = T;
  double yv = P.yield_stress, vfac = 1.0;
  if (P.temperature_enabled) {
    if (T >= P.knots[3]) gel = 1;
    if (gel) Teff = P.knots[3];
    yv = double(P.yield_stress * temperature_factor(float(Teff), P.knots, P.yf));
    vfac = double(temperature_factor(float(Teff), P.knots, P.vf));
  }
  const double ev = visc_param * vfac;
  const double mu = P.mu, lam = P.lam, dt = P.dt;
  double A[9], trial[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i] = dt * C[i] + ((i % 4 == 0) ? 1.0 : 0.0);
  mmul(A, Fe_old, trial);
  mat3<double> Tm, Um, Vm;
#pragma unroll
  for (int i = 0; i < 9; ++i) Tm.a[i] = trial[i];
  double s[3];
  svd3(Tm, Um, s, Vm);
  const double* U = Um.a;
  const double* V = Vm.a;
  const bool coupled = P.coupled != 0;
  const Principal pr = principal_return(s, porosity, n_root, ev, mu, lam, yv, dt, coupled, P.n0);
  double e[3], cs[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e[i] = pr.e[i].v;
    cs[i] = pr.cs[i].v;
  }
  // Fe0 = U diag(e) V^T, and the output Fe = c Fe0
  double Fe0[9], tmp[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) tmp[i * 3 + j] = U[i * 3 + j] * e[j];
  mmul_nt(tmp, V, Fe0);
  const bool isochoric = !coupled && pr.n.v <= 0.0;  // a = 0
  const double target = 1.0 - double(P.n0) + pr.n.v;
  double c = 1.0;
  if (coupled) c = cbrt(mdet(F) / (target * mdet(Fe0)));
  else if (isochoric) c = cbrt(mdet(trial) / mdet(Fe0));
  double Fe[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fe[i] = c * Fe0[i];
  double gFe[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) gFe[i] = gFe_in[i];
  double gnroot = 0.0;  // adjoint of the porosity the map used
  if (coupled) {
    gnroot += gn;
  } else if (gn != 0.0 && !(P.n0 == 0.0f && porosity == 0.0)) {
    const double ratio = mdet(F) / mdet(Fe);
    double fit[9], feit[9];
    minv_t(F, fit);
    minv_t(Fe, feit);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      gF[i] += gn * ratio * fit[i];
      gFe[i] -= gn * ratio * feit[i];
    }
  }
  // the volume fit: c = (det X / det Fe0)^(1/3), X = trial (isochoric) or F / target (coupled)
  double gFe0[9], gTrial[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (coupled || isochoric) {
    double cbar = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) cbar += gFe[i] * Fe0[i];
    const double k3 = cbar * c / 3.0;
    double xit[9], fe0it[9];
    minv_t(coupled ? F : trial, xit);
    minv_t(Fe0, fe0it);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      gFe0[i] = c * gFe[i] - k3 * fe0it[i];
      if (coupled) gF[i] += k3 * xit[i];
      else gTrial[i] += k3 * xit[i];
    }
    if (coupled) gnroot -= k3 / target;
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) gFe0[i] = gFe[i];
  }
  // into the principal frame
  double Ne[9], Nt[9], t1[9];
  mmul_tn(U, gFe0, t1);
  mmul(t1, V, Ne);
  mmul_tn(U, gtau, t1);
  mmul(t1, U, Nt);
  double M[9];
  double gscal[kDir] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < kDir; ++k) gscal[k] += Ne[i * 4] * pr.e[i].d[k] + Nt[i * 4] * pr.cs[i].d[k];
#pragma unroll
  for (int k = 0; k < kDir; ++k) gscal[k] += gnroot * pr.n.d[k];
#pragma unroll
  for (int i = 0; i < 3; ++i) M[i * 4] = gscal[i];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i + 1; j < 3; ++j) {
      const double al = divided(pr.e, s, i, j), be = (e[i] + e[j]) / (s[i] + s[j]);
      const double pp = 0.5 * (al + be), pm = 0.5 * (al - be);
      const double kap = divided(pr.cs, s, i, j) / (s[i] + s[j]);
      const double nt = Nt[i * 3 + j] + Nt[j * 3 + i];
      M[i * 3 + j] = pp * Ne[i * 3 + j] + pm * Ne[j * 3 + i] + nt * kap * s[j];
      M[j * 3 + i] = pm * Ne[i * 3 + j] + pp * Ne[j * 3 + i] + nt * kap * s[i];
    }
  // back to the trial gradient: gTrial += U M V^T
  double UM[9], G[9];
  mmul(U, M, UM);
  mmul_nt(UM, V, G);
#pragma unroll
  for (int i = 0; i < 9; ++i) gTrial[i] += G[i];
  gporosity += gscal[3];
  geta0 += gscal[4] * vfac;
  // yield stress and viscosity factors of the temperature (ding.py:56-65, 160-170); a gelled particle
  // takes them at the last knot, where they no longer depend on it
  if (P.temperature_enabled && !gel)
    gT += gscal[4] * visc_param * temperature_slope(T, P.knots, P.vf) + gscal[5] * double(P.yield_stress) * temperature_slope(T, P.knots, P.yf);
  // trial = (I + dt C) Fe_old
  double t2[9];
  mmul_tn(A, gTrial, t2);
#pragma unroll
  for (int i = 0; i < 9; ++i) gFe_old[i] += t2[i];
  mmul_nt(gTrial, Fe_old, t2);
#pragma unroll
  for (int i = 0; i < 9; ++i) gC[i] += dt * t2[i];
}

// ------------------------------------------------------------------ transfers
__device__ __forceinline__ void axis_weight_grads(float l, double w[3], double dw[3]) {
  const double L = l;
  w[0] = 0.5 * (1.5 - L) * (1.5 - L);
  w[1] = 0.75 - (L - 1.0) * (L - 1.0);
  w[2] = 0.5 * (L - 0.5) * (L - 0.5);
  dw[0] = -(1.5 - L);
  dw[1] = -2.0 * (L - 1.0);
  dw[2] = L - 0.5;
}

struct AdjArgs {
  bk_params P;
  bk_state in, out;
  bk_grid g;
  const float* masses;      // particle masses P2G used
  const float* V0;
  const float* visc_param;  // per dough (one value)
  const float* prod_param;  // gas production parameter (one value)
  bk_adjoint adj;
  double* gm;  // grid adjoints (tiled nodes): mass, momentum (3), velocity (3)
  double* gp;
  double* gv;
  double* param_grad;  // [0] viscosity parameter, [1] gas production parameter, [2] heat-transfer parameter,
                       // [3] oven temperature
  int n, nb;
  // transport (when on): forward scene and workspace data
  const double* capacities;
  const float* heat_param;
  const int32_t *particle_slot, *slot_face_offsets, *slot_faces, *faces;
  const int32_t *bind_offsets, *bind_particles;
  const float *bind_weights, *bind_ref_offsets;
  const double *slot_exposed, *slot_contact;
  const float* vertices;
  // transport adjoint scratch: per tiled node, per particle, per slot, per vertex
  double *accT, *accM, *ghr, *gwr, *gV, *gC, *gE, *gW, *gH, *gL;
  double *tcbar, *mrbar;
  double *gSexp, *gScon;
  double* gvert;
};

// Per particle (Ding model): the return map's adjoint.  Its outputs' adjoints
// arrive in adj.Fe / adj.tau / adj.porosity; adds the new F's and C's
// adjoints to adj.F / adj.C and leaves the old Fe's and porosity's in
// adj.Fe / adj.porosity (adj.tau cleared for P2G).
__global__ __launch_bounds__(128) void k_adj_ding(AdjArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double geta = 0.0;
  if (p < a.n) {
    const bk_params& P = a.P;
    DingParams D;
    D.mu = P.mu; D.lam = P.lam; D.yield_stress = P.yield_stress; D.n0 = P.n0; D.dt = P.dt;
    for (int i = 0; i < 4; ++i) { D.knots[i] = P.knots[i]; D.yf[i] = P.yield_factors[i]; D.vf[i] = P.viscosity_factors[i]; }
    D.temperature_enabled = P.temperature_enabled;
    D.coupled = P.gas_enabled;
    D.bracketed = P.return_solver;
    double Fe_old[9], Fn[9], Cn[9], gFe[9], gtau[9], gFe_old[9], gF[9], gC[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      Fe_old[i] = a.in.Fe[9 * p + i];
      Fn[i] = a.out.F[9 * p + i];
      Cn[i] = a.out.C[9 * p + i];
      gFe[i] = a.adj.Fe[9 * p + i];
      gtau[i] = a.adj.tau[9 * p + i];
      gFe_old[i] = gF[i] = gC[i] = 0.0;
    }
    double gpor = 0.0, gT = 0.0;
    const float T_local = P.transport_enabled ? a.out.T_mat[p] : a.in.T_mat[p];
    ding_return_adjoint(D, Fe_old, Fn, Cn, double(a.in.porosity[p]), double(a.out.porosity[p]), a.in.gelled[p], double(T_local),
                        double(a.visc_param[0]), gFe, gtau, a.adj.porosity[p], gFe_old, gF, gC, gpor, geta, gT);
    if (P.transport_enabled) a.adj.T[p] += gT;  // T_mat = float(T_new)
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      a.adj.F[9 * p + i] += gF[i];
      a.adj.C[9 * p + i] += gC[i];
      a.adj.Fe[9 * p + i] = gFe_old[i];
      a.adj.tau[9 * p + i] = 0.0;
    }
    a.adj.porosity[p] = gpor;
  }
  // viscosity-parameter gradient: wavefront sum, one atomic per wavefront
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) geta += __shfl_xor(geta, off, 64);
  if (__lane_id() == 0 && geta != 0.0) unsafeAtomicAdd(a.param_grad, geta);
}

// Per particle: G2P's adjoint with the F update (transfer.py:73-109), its
// gather part.  Leaves the input state's partial adjoints of x and F in
// adj.x / adj.F, and the adjoints of the new velocity and affine velocity
// (with the position and F-update terms added) in adj.v / adj.C for the
// scatter to the grid (k_adj_scatter).
__global__ __launch_bounds__(128) void k_adj_g2p(AdjArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < a.n) {
    const bk_params& P = a.P;
    const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
    double gx[3], gv[3], gC[9], gF[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      gx[i] = a.adj.x[3 * p + i];
      gv[i] = a.adj.v[3 * p + i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      gC[i] = a.adj.C[9 * p + i];
      gF[i] = a.adj.F[9 * p + i];
    }
    // G2P (transfer.py:73-109)
    const float dx = P.dx;
    const double s4 = 4.0 / (double(dx) * double(dx));
    const double dt = P.dt;
    float g3[3];
    grid_position(a.in.x + 3 * p, o, dx, g3);
    int b[3];
    stencil_base(g3, b);
    double gx_in[3] = {gx[0], gx[1], gx[2]};
    double gF_in[9];
    double Cnew[9], Fold[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      Cnew[i] = a.out.C[9 * p + i];
      Fold[i] = a.in.F[9 * p + i];
    }
    // F' = (I + dt C') F  ->  gF_in = (I + dt C')^T gF ; gC' += dt gF F^T
    double Ainc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ainc[i] = dt * Cnew[i] + ((i % 4 == 0) ? 1.0 : 0.0);
    mmul_tn(Ainc, gF, gF_in);
    double t[9];
    mmul_nt(gF, Fold, t);
#pragma unroll
    for (int i = 0; i < 9; ++i) gC[i] += dt * t[i];
    // x' = x + dt v'
#pragma unroll
    for (int i = 0; i < 3; ++i) gv[i] += dt * gx[i];
    if (valid_stencil(b, P.res)) {
      const float l[3] = {g3[0] - float(b[0]), g3[1] - float(b[1]), g3[2] - float(b[2])};
      double wx[3], wy[3], wz[3], dwx[3], dwy[3], dwz[3];
      axis_weight_grads(l[0], wx, dwx);
      axis_weight_grads(l[1], wy, dwy);
      axis_weight_grads(l[2], wz, dwz);
      const double inv_dx = 1.0 / double(dx);
#pragma unroll 1
      for (int ox = 0; ox < 3; ++ox)
        for (int oy = 0; oy < 3; ++oy)
          for (int oz = 0; oz < 3; ++oz) {
            const double w = wx[ox] * wy[oy] * wz[oz];
            const double dw[3] = {dwx[ox] * wy[oy] * wz[oz] * inv_dx, wx[ox] * dwy[oy] * wz[oz] * inv_dx, wx[ox] * wy[oy] * dwz[oz] * inv_dx};
            const int64_t node = tiled_node(b[0] + ox, b[1] + oy, b[2] + oz, a.nb);
            const double vi[3] = {a.g.v[3 * node], a.g.v[3 * node + 1], a.g.v[3 * node + 2]};
            const double dd[3] = {(double(ox) - l[0]) * dx, (double(oy) - l[1]) * dx, (double(oz) - l[2]) * dx};
            // value the node velocity contributes: v' . gv + s4 v^T gC dd
            double gCdd[3], gCtv[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              gCdd[r] = gC[r * 3] * dd[0] + gC[r * 3 + 1] * dd[1] + gC[r * 3 + 2] * dd[2];
              gCtv[r] = gC[r] * vi[0] + gC[3 + r] * vi[1] + gC[6 + r] * vi[2];
            }
            const double val = vi[0] * gv[0] + vi[1] * gv[1] + vi[2] * gv[2] + s4 * (vi[0] * gCdd[0] + vi[1] * gCdd[1] + vi[2] * gCdd[2]);
#pragma unroll
            for (int r = 0; r < 3; ++r) gx_in[r] += dw[r] * val - w * s4 * gCtv[r];
          }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      a.adj.x[3 * p + i] = gx_in[i];
      a.adj.v[3 * p + i] = gv[i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      a.adj.F[9 * p + i] = gF_in[i];
      a.adj.C[9 * p + i] = gC[i];
    }
  }
}

// G2P's adjoint, its scatter part: node velocity adjoints
// sum_p w_ip (gv_p + 4/dx^2 gC_p (x_i - x_p)), over the P2G tiles of the
// forward substep just run from the same positions (the workspace's sorted
// particles, one particle block per tile).  As the forward P2G does, each tile
// accumulates in LDS (FP64) and flushes one global atomic per node and
// component, instead of 81 global atomics per particle.
__global__ __launch_bounds__(kTileThreads) void k_adj_scatter(const bk_dough* __restrict__ doughs, AdjArgs a, bk_workspace* ws) {
  __shared__ double acc[3][kRegionNodes];
  __shared__ int cell_lo[64], cell_hi[64];
  const int ntiles = tile_count(ws);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bk_params& P = a.P;
  const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
  const float dx = P.dx;
  const double s4 = 4.0 / (double(dx) * double(dx));
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tg = tile_at(doughs, ws, t);
    for (int i = threadIdx.x; i < 3 * kRegionNodes; i += kTileThreads) (&acc[0][0])[i] = 0.0;
    if (threadIdx.x < 64) {
      cell_lo[threadIdx.x] = 0;
      cell_hi[threadIdx.x] = 0;
    }
    __syncthreads();
    tile_cell_ranges(ws, tg, cell_lo, cell_hi);
    const int hi = cell_hi[lane];
    for (int j = cell_lo[lane] + w; j < hi; j += 4) {
      const int p = ws->sorted_vals[tg.start + j];
      float g3[3];
      grid_position(a.in.x + 3 * p, o, dx, g3);
      int b[3];
      stencil_base(g3, b);
      if (!valid_stencil(b, P.res) || (b[0] >> 2) != tg.g.bx || (b[1] >> 2) != tg.g.by || (b[2] >> 2) != tg.g.bz) continue;
      const float l[3] = {g3[0] - float(b[0]), g3[1] - float(b[1]), g3[2] - float(b[2])};
      float wxf[3], wyf[3], wzf[3];
      axis_weights(l[0], wxf);
      axis_weights(l[1], wyf);
      axis_weights(l[2], wzf);
      double gv[3], gC[9];
#pragma unroll
      for (int i = 0; i < 3; ++i) gv[i] = a.adj.v[3 * p + i];
#pragma unroll
      for (int i = 0; i < 9; ++i) gC[i] = s4 * a.adj.C[9 * p + i];
      const int li = b[0] - 4 * tg.g.bx, lj = b[1] - 4 * tg.g.by, lk = b[2] - 4 * tg.g.bz;
#pragma unroll
      for (int ox = 0; ox < 3; ++ox)
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
          for (int oz = 0; oz < 3; ++oz) {
            const double wt = double(wxf[ox]) * double(wyf[oy]) * double(wzf[oz]);
            const double dd[3] = {(double(ox) - l[0]) * dx, (double(oy) - l[1]) * dx, (double(oz) - l[2]) * dx};
            const int ri = region_index(li + ox, lj + oy, lk + oz);
#pragma unroll
            for (int r = 0; r < 3; ++r)
              atomicAdd(&acc[r][ri], wt * (gv[r] + gC[r * 3] * dd[0] + gC[r * 3 + 1] * dd[1] + gC[r * 3 + 2] * dd[2]));
          }
    }
    __syncthreads();
    for (int ri = threadIdx.x; ri < kRegionNodes; ri += kTileThreads) {
      const int li = ri / 36, lj = (ri / 6) % 6, lk = ri % 6;
      const int i = 4 * tg.g.bx + li, j = 4 * tg.g.by + lj, k = 4 * tg.g.bz + lk;
      if (i >= P.res || j >= P.res || k >= P.res) continue;
      const int64_t node = tg.g.node_base + tiled_node(i, j, k, tg.g.nb);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double v = acc[r][ri];
        if (v != 0.0) unsafeAtomicAdd(a.gv + 3 * node + r, v);
      }
    }
    __syncthreads();
  }
}

// Per node: the grid update's adjoint (grid.py:20-38).  Consumes the velocity
// adjoints G2P's adjoint scattered and sets them back to zero, so the scratch
// is clear for the next call without a pass over the whole grid: only nodes
// with mass receive scatters (every stencil node of a particle has mass).
__global__ __launch_bounds__(256) void k_adj_grid(AdjArgs a, int64_t nodes) {
  const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (node >= nodes) return;
  const double m = a.g.m[node];
  if (!(m > 0.0)) return;
  const bk_params& P = a.P;
  const int64_t blk = node >> 6;
  const int loc = int(node & 63);
  const int nb = a.nb;
  const int bz = int(blk % nb), by = int((blk / nb) % nb), bx = int(blk / (int64_t(nb) * nb));
  const int c[3] = {4 * bx + (loc >> 4), 4 * by + ((loc >> 2) & 3), 4 * bz + (loc & 3)};
  double gmass = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double pk = a.g.p[3 * node + k];
    // the forward zeroed this component (a boundary node whose velocity pointed outwards) iff it stored 0 there
    const bool boundary = c[k] < P.boundary_cells || c[k] >= P.res - P.boundary_cells;
    const bool zeroed = boundary && a.g.v[3 * node + k] == 0.0f;
    const double gvk = zeroed ? 0.0 : a.gv[3 * node + k];
    a.gv[3 * node + k] = 0.0;
    a.gp[3 * node + k] = gvk / m;
    gmass -= gvk * pk / (m * m);
  }
  a.gm[node] = gmass;
}

// Particle p of P2G's adjoint (transfer.py:36-70) with the stress's: neo-Hookean
// (elastic.py:4-10) or the Ding state's, plus the pore-gas pressure term
// (gas.py:16-35) and the water stress (thermal_water.py:248-253) when
// transport is on.  Completes the mechanical part of the input state's
// adjoint; the mass and the new temperature and water feed adj.W / adj.T,
// which hold the adjoints of the substep's new T and W until the transport
// adjoint (k_adj_b1) turns them into the old ones.
__device__ inline void adj_p2g_particle(const AdjArgs& a, int p) {
  const bk_params& P = a.P;
  const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
  const float dx = P.dx;
  float g3[3];
  grid_position(a.in.x + 3 * p, o, dx, g3);
  int b[3];
  stencil_base(g3, b);
  if (!valid_stencil(b, P.res)) return;
  const float l[3] = {g3[0] - float(b[0]), g3[1] - float(b[1]), g3[2] - float(b[2])};
  double wx[3], wy[3], wz[3], dwx[3], dwy[3], dwz[3];
  axis_weight_grads(l[0], wx, dwx);
  axis_weight_grads(l[1], wy, dwy);
  axis_weight_grads(l[2], wz, dwz);
  const bool transport = P.transport_enabled != 0;
  const double mass = transport ? double(a.out.masses[p]) : double(a.masses[p]);
  const double s_tau = double(P.dt) * double(a.V0[p]) * 4.0 / (double(dx) * double(dx));
  // the stress P2G used: the state's for the Ding model, mu (F F^T - I) + lam log(J) I of the input F otherwise
  double tau[9], F[9];
  bool stressed = true;
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.in.F[9 * p + i];
  const double J = mdet(F);
  // pore gas: tau += J (p_ambient - p) I = (J p_ambient - N R T / n) I, N the substep's new moles
  bool gas = false;
  double moles = 0.0, Tg = 0.0, nv = 0.0;
  if (P.model == 1) {
#pragma unroll
    for (int i = 0; i < 9; ++i) tau[i] = a.in.tau[9 * p + i];
    if (P.gas_enabled) {
      nv = a.in.porosity[p];
      moles = a.out.moles[p];
      Tg = transport ? a.out.T_mat[p] : a.in.T_mat[p];
      gas = nv > 0.0 && nv < 1.0 && J > 0.0;
      if (gas) {
        const double add = J * double(P.p_ambient) - moles * kGasConstant * Tg / nv;
        tau[0] += add;
        tau[4] += add;
        tau[8] += add;
      }
    }
  } else {
    stressed = isfinite(J) && J > 0.0;
    mmul_nt(F, F, tau);
#pragma unroll
    for (int i = 0; i < 9; ++i) tau[i] = stressed ? double(P.mu) * (tau[i] - ((i % 4 == 0) ? 1.0 : 0.0)) + ((i % 4 == 0) ? double(P.lam) * log(J) : 0.0) : 0.0;
  }
  // water stress: tau -= J c (s - s0) I with s = W_new / (rho_w V0 J): k = c (W_new / (rho_w V0) - J s0)
  const double cw = (transport && P.water_enabled) ? double(P.water_stress) : 0.0;
  const double rw_v0 = double(P.rho_w) * double(a.V0[p]);
  if (cw != 0.0) {
    const double k = cw * (a.out.W[p] / rw_v0 - J * double(P.s0));
    tau[0] -= k;
    tau[4] -= k;
    tau[8] -= k;
  }
  double Am[9], vv[3], Cp[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    Cp[i] = a.in.C[9 * p + i];
    Am[i] = mass * Cp[i] - s_tau * tau[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) vv[i] = a.in.v[3 * p + i];
  double gx[3] = {0, 0, 0}, gvp[3] = {0, 0, 0}, gA[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gmass = 0.0;
  const double inv_dx = 1.0 / double(dx);
  for (int ox = 0; ox < 3; ++ox)
    for (int oy = 0; oy < 3; ++oy)
      for (int oz = 0; oz < 3; ++oz) {
        const double w = wx[ox] * wy[oy] * wz[oz];
        const double dw[3] = {dwx[ox] * wy[oy] * wz[oz] * inv_dx, wx[ox] * dwy[oy] * wz[oz] * inv_dx, wx[ox] * wy[oy] * dwz[oz] * inv_dx};
        const int64_t node = tiled_node(b[0] + ox, b[1] + oy, b[2] + oz, a.nb);
        const double gpi[3] = {a.gp[3 * node], a.gp[3 * node + 1], a.gp[3 * node + 2]};
        const double gmi = a.gm[node];
        const double dd[3] = {(double(ox) - l[0]) * dx, (double(oy) - l[1]) * dx, (double(oz) - l[2]) * dx};
        double Add[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) Add[r] = Am[r * 3] * dd[0] + Am[r * 3 + 1] * dd[1] + Am[r * 3 + 2] * dd[2];
        const double val = gmi * mass + gpi[0] * (mass * vv[0] + Add[0]) + gpi[1] * (mass * vv[1] + Add[1]) + gpi[2] * (mass * vv[2] + Add[2]);
        if (transport) {
          // d/d mass of m_i = w m and p_i = w (m v + (m C - ...) dd)
          double vcd = gmi;
#pragma unroll
          for (int r = 0; r < 3; ++r) vcd += gpi[r] * (vv[r] + Cp[r * 3] * dd[0] + Cp[r * 3 + 1] * dd[1] + Cp[r * 3 + 2] * dd[2]);
          gmass += w * vcd;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          // d dd / dx = -I: -(A^T gp)_r
          const double Atg = Am[r] * gpi[0] + Am[3 + r] * gpi[1] + Am[6 + r] * gpi[2];
          gx[r] += dw[r] * val - w * Atg;
          gvp[r] += w * mass * gpi[r];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) gA[r * 3 + cc] += w * gpi[r] * dd[cc];
      }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    a.adj.x[3 * p + r] += gx[r];
    a.adj.v[3 * p + r] = gvp[r];
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) a.adj.C[9 * p + i] = mass * gA[i];
  double gtau[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) gtau[i] = -s_tau * gA[i];
  const double tr = gtau[0] + gtau[4] + gtau[8];
  double gJ = 0.0;  // adjoint of det F from the gas and water terms
  if (transport) {
    // masses = dry + W_new; the water stress through W_new
    a.adj.W[p] += gmass - tr * cw / rw_v0;
    gJ += tr * cw * double(P.s0);
  }
  if (P.model == 1) {
#pragma unroll
    for (int i = 0; i < 9; ++i) a.adj.tau[9 * p + i] = gtau[i];
    if (gas) {
      gJ += tr * double(P.p_ambient);
      a.adj.porosity[p] += tr * moles * kGasConstant * Tg / (nv * nv);
      a.adj.moles[p] -= tr * kGasConstant * Tg / nv;  // adjoint of the new moles
      if (transport) a.adj.T[p] -= tr * moles * kGasConstant / nv;  // T_mat = float(T_new)
    }
  } else if (stressed) {
    // neo-Hookean Kirchhoff stress mu (F F^T - I) + lam log(J) I: gF += mu (gtau + gtau^T) F + lam tr(gtau) F^-T
    double sym[9], t[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) sym[i * 3 + j] = gtau[i * 3 + j] + gtau[j * 3 + i];
    mmul(sym, F, t);
#pragma unroll
    for (int i = 0; i < 9; ++i) a.adj.F[9 * p + i] += double(P.mu) * t[i];
    gJ += double(P.lam) * tr / J;
  }
  if (gJ != 0.0) {
    double fit[9];
    minv_t(F, fit);
#pragma unroll
    for (int i = 0; i < 9; ++i) a.adj.F[9 * p + i] += gJ * J * fit[i];
  }
}

__global__ __launch_bounds__(128) void k_adj_p2g(AdjArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < a.n) adj_p2g_particle(a, p);
}

// ------------------------------------------------------------------ transport adjoint
// The forward transport of one substep (thermal_water.py:95-260, fused as the
// substep's mesh, pass A, diffusion and pass B1 stages), in reverse:
//   B1   T' = T + sum_i w dT_i;  W' = W (1 + sum_{W_i > 0} w dW_i / W_i);
//        N' = N + dt rho alpha exp(-z^2), z = ((T + T')/2 - T_peak) / width
//   diff dT_i = dt hr_i / C_i, dW_i = dt wr_i over the 6-neighbour FV stencil
//   A    V_i, C_i, E_i, W_i, H_i, L_i = sum_p w (V0 J, cap, cap T, W, heat, loss)
//        with the Robin heat and loss of each particle's exposed / contact area
//   mesh areas of the boundary faces from the vertices each particle binds.

// Particle p: the B1 adjoint.  In: adj.T / adj.W / adj.moles hold the adjoints
// of the new T, W and moles.  Out: those of the old ones (the thermal P2G
// adjoint adds more), the scatter weights of dT and W_i-scaled dW, the
// positions' share through the gather weights, and the production parameter's
// gradient (returned).
__device__ inline double adj_b1_particle(const AdjArgs& a, int p) {
  const bk_params& P = a.P;
  const double Told = a.in.T[p], Tnew = a.out.T[p], Wold = a.in.W[p];
  double gTn = a.adj.T[p], gWn = a.adj.W[p];
  double gT = 0.0, gprod = 0.0;
  if (P.model == 1 && P.gas_enabled && !P.settling) {
    const double gN = a.adj.moles[p];  // N' = N + source: the old moles' adjoint is the same
    const double z = (0.5 * (Told + Tnew) - double(P.gas_peak)) / double(P.gas_width);
    const double e = exp(-(z * z));
    const double src = double(P.dt * P.density) * double(a.prod_param[0]) * e;
    gprod = gN * double(P.dt * P.density) * e;
    const double gz = gN * src * (-2.0 * z) * 0.5 / double(P.gas_width);
    gT += gz;
    gTn += gz;
  }
  // gather: tc = sum w dT_i, mr = sum_{W_i > 0} w dW_i / W_i
  const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
  float g3[3];
  grid_position(a.in.x + 3 * p, o, P.dx, g3);
  int b[3];
  stencil_base(g3, b);
  double mr = 0.0, gx[3] = {0.0, 0.0, 0.0};
  if (valid_stencil(b, P.res)) {
    const float l[3] = {g3[0] - float(b[0]), g3[1] - float(b[1]), g3[2] - float(b[2])};
    double wx[3], wy[3], wz[3], dwx[3], dwy[3], dwz[3];
    axis_weight_grads(l[0], wx, dwx);
    axis_weight_grads(l[1], wy, dwy);
    axis_weight_grads(l[2], wz, dwz);
    const double inv_dx = 1.0 / double(P.dx);
    const double gmr = gWn * Wold;
    for (int ox = 0; ox < 3; ++ox)
      for (int oy = 0; oy < 3; ++oy)
        for (int oz = 0; oz < 3; ++oz) {
          const int64_t node = tiled_node(b[0] + ox, b[1] + oy, b[2] + oz, a.nb);
          const double q = a.g.dW[node];  // the forward stores dW / W per node
          mr += wx[ox] * wy[oy] * wz[oz] * q;
          const double val = gTn * a.g.dT[node] + gmr * q;
          gx[0] += dwx[ox] * wy[oy] * wz[oz] * inv_dx * val;
          gx[1] += wx[ox] * dwy[oy] * wz[oz] * inv_dx * val;
          gx[2] += wx[ox] * wy[oy] * dwz[oz] * inv_dx * val;
        }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) a.adj.x[3 * p + r] += gx[r];
  a.tcbar[p] = gTn;         // adjoint of tc
  a.mrbar[p] = gWn * Wold;  // adjoint of mr
  a.adj.T[p] = gT + gTn;    // T' = T + tc
  a.adj.W[p] = gWn * (1.0 + mr);
  return gprod;
}

__global__ __launch_bounds__(128) void k_adj_b1(AdjArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double gprod = p < a.n ? adj_b1_particle(a, p) : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) gprod += __shfl_xor(gprod, off, 64);
  if (__lane_id() == 0 && gprod != 0.0) unsafeAtomicAdd(a.param_grad + 1, gprod);
}

// The B1 gathers' scatter: per node sum_p w tc_bar and sum_p w mr_bar over the
// forward substep's P2G tiles (one cell per lane, FP64 LDS, one global atomic
// per node and sum); k_adj_diffuse_a applies the per-node factors.
__global__ __launch_bounds__(kTileThreads) void k_adj_scatter_b1(const bk_dough* __restrict__ doughs, AdjArgs a, bk_workspace* ws) {
  __shared__ double acc[2][kRegionNodes];
  __shared__ int cell_lo[64], cell_hi[64];
  const int ntiles = tile_count(ws);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bk_params& P = a.P;
  const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tg = tile_at(doughs, ws, t);
    for (int i = threadIdx.x; i < 2 * kRegionNodes; i += kTileThreads) (&acc[0][0])[i] = 0.0;
    if (threadIdx.x < 64) {
      cell_lo[threadIdx.x] = 0;
      cell_hi[threadIdx.x] = 0;
    }
    __syncthreads();
    tile_cell_ranges(ws, tg, cell_lo, cell_hi);
    const int hi = cell_hi[lane];
    for (int j = cell_lo[lane] + w; j < hi; j += 4) {
      const int p = ws->sorted_vals[tg.start + j];
      float g3[3];
      grid_position(a.in.x + 3 * p, o, P.dx, g3);
      int b[3];
      stencil_base(g3, b);
      if (!valid_stencil(b, P.res) || (b[0] >> 2) != tg.g.bx || (b[1] >> 2) != tg.g.by || (b[2] >> 2) != tg.g.bz) continue;
      const float l[3] = {g3[0] - float(b[0]), g3[1] - float(b[1]), g3[2] - float(b[2])};
      float wxf[3], wyf[3], wzf[3];
      axis_weights(l[0], wxf);
      axis_weights(l[1], wyf);
      axis_weights(l[2], wzf);
      const double tb = a.tcbar[p], mb = a.mrbar[p];
      const int li = b[0] - 4 * tg.g.bx, lj = b[1] - 4 * tg.g.by, lk = b[2] - 4 * tg.g.bz;
#pragma unroll
      for (int ox = 0; ox < 3; ++ox)
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
          for (int oz = 0; oz < 3; ++oz) {
            const double wt = double(wxf[ox]) * double(wyf[oy]) * double(wzf[oz]);
            const int ri = region_index(li + ox, lj + oy, lk + oz);
            atomicAdd(&acc[0][ri], wt * tb);
            atomicAdd(&acc[1][ri], wt * mb);
          }
    }
    __syncthreads();
    for (int ri = threadIdx.x; ri < kRegionNodes; ri += kTileThreads) {
      const int li = ri / 36, lj = (ri / 6) % 6, lk = ri % 6;
      const int i = 4 * tg.g.bx + li, j = 4 * tg.g.by + lj, k = 4 * tg.g.bz + lk;
      if (i >= P.res || j >= P.res || k >= P.res) continue;
      const int64_t node = tg.g.node_base + tiled_node(i, j, k, tg.g.nb);
      if (acc[0][ri] != 0.0) unsafeAtomicAdd(a.accT + node, acc[0][ri]);
      if (acc[1][ri] != 0.0) unsafeAtomicAdd(a.accM + node, acc[1][ri]);
    }
    __syncthreads();
  }
}

// Per active node with capacity: the diffusion step's own adjoint,
// dT_i = dt hr_i / C_i and dW_i = dt wr_i (thermal_water.py:150-190), and the
// W_i of B1's dW_i / W_i.  Leaves hr_bar, wr_bar for the neighbour gather and
// starts the C and W adjoints; clears the scatter sums.
__global__ __launch_bounds__(256) void k_adj_diffuse_a(AdjArgs a, bk_workspace* ws) {
  const int nact = ws->counters[1];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nact * 64; t += gridDim.x * blockDim.x) {
    const int blk = __builtin_amdgcn_readfirstlane(ws->active[t >> 6]);
    const int64_t node = (int64_t(blk) << 6) | (t & 63);
    const double Ci = a.g.C[node];
    if (!(Ci > 0.0)) continue;
    const double gdT = a.accT[node], sM = a.accM[node];
    a.accT[node] = 0.0;
    a.accM[node] = 0.0;
    const double Wi = a.g.W[node];
    const double gdW = Wi > 0.0 ? sM / Wi : 0.0;
    const double dt = double(a.P.dt);
    a.ghr[node] = dt * gdT / Ci;
    a.gwr[node] = dt * gdW;
    a.gC[node] = -gdT * a.g.dT[node] / Ci;
    a.gW[node] = Wi > 0.0 ? -sM * a.g.dW[node] / Wi : 0.0;  // q = change / W: dq / dW = -q / W
  }
}

// Per active node with capacity: the FV fluxes' adjoint, gathered from the
// node's own terms and its neighbours' (whose sums hold (T_i - T_j) and
// (S_i - S_j) with the same symmetric fill), then T = E / C and S = W / (rho_w V).
__global__ __launch_bounds__(256) void k_adj_diffuse_b(AdjArgs a, bk_workspace* ws) {
  const int nact = ws->counters[1];
  const bk_params& P = a.P;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nact * 64; t += gridDim.x * blockDim.x) {
    const int blk = __builtin_amdgcn_readfirstlane(ws->active[t >> 6]);
    const int loc = t & 63;
    const int64_t node = (int64_t(blk) << 6) | loc;
    const double Ci = a.g.C[node];
    if (!(Ci > 0.0)) continue;
    const int nb = a.nb;
    const int bz = blk % nb, by = (blk / nb) % nb, bx = blk / (nb * nb);
    const int c[3] = {4 * bx + (loc >> 4), 4 * by + ((loc >> 2) & 3), 4 * bz + (loc & 3)};
    const double rho_w = double(P.rho_w);
    const double dx3 = double(P.dx * P.dx * P.dx);
    const double hk = double(P.kappa * P.dx);
    const double wk = double(P.diffusivity * P.rho_w * P.dx);
    const double Vi = a.g.V[node], Wi = a.g.W[node];
    const double Ti = a.g.E[node] / Ci, Si = Wi / (rho_w * Vi);
    const double hri = a.ghr[node], wri = a.gwr[node];
    double gTi = 0.0, gSi = 0.0, gVi = 0.0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        int q[3] = {c[0], c[1], c[2]};
        q[ax] += 2 * side - 1;
        if (q[ax] < 0 || q[ax] >= P.res) continue;
        const int64_t other = tiled_node(q[0], q[1], q[2], nb);
        const double Cj = a.g.C[other];
        if (!(Cj > 0.0)) continue;
        const double Vj = a.g.V[other];
        const double Tj = a.g.E[other] / Cj;
        const double fill = 2.0 * Vi * Vj / (Vi + Vj) / dx3;
        const double dfill = 2.0 * Vj * Vj / ((Vi + Vj) * (Vi + Vj)) / dx3;  // d fill / d V_i
        const double hrj = a.ghr[other];
        // heat: node i's term hk fill (Tj - Ti), node j's term hk fill (Ti - Tj)
        gTi += hk * fill * (hrj - hri);
        gVi += hk * dfill * (Tj - Ti) * (hri - hrj);
        if (P.water_enabled) {
          const double Sj = a.g.W[other] / (rho_w * Vj);
          const double wrj = a.gwr[other];
          const double tm = 0.5 * (Ti + Tj), sm = 0.5 * (Si + Sj);
          const double mob = wk * fill * tm * sm;
          // node i's term mob (Sj - Si), node j's term mob (Si - Sj): both through mob's dependence on i
          const double A = (Sj - Si) * (wri - wrj);
          gSi += mob * (wrj - wri) + A * wk * fill * tm * 0.5;
          gTi += A * wk * fill * 0.5 * sm;
          gVi += A * wk * dfill * tm * sm;
        }
      }
    a.gH[node] = hri;
    a.gL[node] = -wri;
    a.gE[node] = gTi / Ci;
    a.gC[node] += -gTi * Ti / Ci;
    a.gW[node] += gSi / (rho_w * Vi);
    a.gV[node] = gVi - gSi * Si / Vi;
  }
}

// Particle p: the thermal P2G's adjoint (thermal_water.py:95-137) with the
// Robin terms' (108-120) -- also those of B1's loss and exchange accumulators,
// which use the same inputs.  Adds to adj.T / adj.W / adj.x / adj.F, sets the
// adjoints of p's exposed and contact areas, returns the gradients of the
// heat-transfer parameter h and of the oven temperature (heat = h S_exp
// (T_oven - T) + ...: d heat / d T_oven = h S_exp).
__device__ inline double2 adj_pass_a_particle(const AdjArgs& a, int p) {
  const bk_params& P = a.P;
  const float o[3] = {P.origin[0], P.origin[1], P.origin[2]};
  const double T = a.in.T[p], W = a.in.W[p], cap = a.capacities[p];
  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = a.in.F[9 * p + i];
  const double J = mdet(F);
  const double vol = double(a.V0[p]) * J;
  const int slot = a.particle_slot[p];
  const double Sexp = slot >= 0 ? a.slot_exposed[slot] : 0.0, Scon = slot >= 0 ? a.slot_contact[slot] : 0.0;
  const double h = P.settling ? 0.0 : double(a.heat_param[0]);
  const double rho_w = double(P.rho_w);
  const double sat = W / (rho_w * vol);
  const double beta = P.water_enabled ? double(P.loss_coef) : 0.0;
  const double loss = beta * T * sat * Sexp;
  const double heat = h * Sexp * (double(P.T_oven) - T) + double(P.tray_h) * Scon * (double(P.T_tray) - T) - double(P.latent) * loss;
  // accumulators: losses' = losses + dt loss, exchanges' = exchanges + dt heat
  double gheat = double(P.dt) * a.adj.exchanges[p], gloss = double(P.dt) * a.adj.losses[p];
  double gvol = 0.0, gT = 0.0, gW = 0.0, gx[3] = {0.0, 0.0, 0.0};
  float g3[3];
  grid_position(a.in.x + 3 * p, o, P.dx, g3);
  int b[3];
  stencil_base(g3, b);
  if (valid_stencil(b, P.res)) {
    const float l[3] = {g3[0] - float(b[0]), g3[1] - float(b[1]), g3[2] - float(b[2])};
    double wx[3], wy[3], wz[3], dwx[3], dwy[3], dwz[3];
    axis_weight_grads(l[0], wx, dwx);
    axis_weight_grads(l[1], wy, dwy);
    axis_weight_grads(l[2], wz, dwz);
    const double inv_dx = 1.0 / double(P.dx);
    for (int ox = 0; ox < 3; ++ox)
      for (int oy = 0; oy < 3; ++oy)
        for (int oz = 0; oz < 3; ++oz) {
          const int64_t node = tiled_node(b[0] + ox, b[1] + oy, b[2] + oz, a.nb);
          const double w = wx[ox] * wy[oy] * wz[oz];
          const double gVn = a.gV[node], gCn = a.gC[node], gEn = a.gE[node], gWn = a.gW[node], gHn = a.gH[node], gLn = a.gL[node];
          gvol += w * gVn;
          gT += w * cap * gEn;
          gW += w * gWn;
          gheat += w * gHn;
          gloss += w * gLn;
          const double val = vol * gVn + cap * gCn + cap * T * gEn + W * gWn + heat * gHn + loss * gLn;
          gx[0] += dwx[ox] * wy[oy] * wz[oz] * inv_dx * val;
          gx[1] += wx[ox] * dwy[oy] * wz[oz] * inv_dx * val;
          gx[2] += wx[ox] * wy[oy] * dwz[oz] * inv_dx * val;
        }
  }
  // Robin terms
  gloss -= double(P.latent) * gheat;
  gT += -(h * Sexp + double(P.tray_h) * Scon) * gheat + beta * sat * Sexp * gloss;
  const double gsat = beta * T * Sexp * gloss;
  const double gSexp = h * (double(P.T_oven) - T) * gheat + beta * T * sat * gloss;
  const double gScon = double(P.tray_h) * (double(P.T_tray) - T) * gheat;
  const double gh = Sexp * (double(P.T_oven) - T) * gheat;
  const double goven = h * Sexp * gheat;
  gW += gsat / (rho_w * vol);
  gvol -= gsat * sat / vol;
  a.adj.T[p] += gT;
  a.adj.W[p] += gW;
#pragma unroll
  for (int r = 0; r < 3; ++r) a.adj.x[3 * p + r] += gx[r];
  if (gvol != 0.0) {
    double fit[9];
    minv_t(F, fit);
    const double gJ = gvol * double(a.V0[p]);
#pragma unroll
    for (int i = 0; i < 9; ++i) a.adj.F[9 * p + i] += gJ * J * fit[i];
  }
  if (slot >= 0) {
    a.gSexp[slot] = gSexp;
    a.gScon[slot] = gScon;
  }
  return P.settling ? make_double2(0.0, 0.0) : make_double2(gh, goven);
}

__global__ __launch_bounds__(128) void k_adj_pass_a(AdjArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const double2 g = p < a.n ? adj_pass_a_particle(a, p) : make_double2(0.0, 0.0);
  double gh = g.x, go = g.y;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    gh += __shfl_xor(gh, off, 64);
    go += __shfl_xor(go, off, 64);
  }
  if (__lane_id() == 0 && gh != 0.0) unsafeAtomicAdd(a.param_grad + 2, gh);
  if (__lane_id() == 0 && go != 0.0) unsafeAtomicAdd(a.param_grad + 3, go);
}

// Per owner slot: the boundary faces' areas (thermal_water.py:77-92) in reverse,
// area = |(b - a) x (c - a)| / 2 summed into the slot's exposed or contact total
// (the contact test is a discrete choice): vertex adjoints, FP64 atomics.
__global__ __launch_bounds__(256) void k_adj_faces(AdjArgs a, int nslots) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const bk_params& P = a.P;
  const double ge = a.gSexp[s], gc = a.gScon[s];
  if (ge == 0.0 && gc == 0.0) return;
  const float floor_h = P.origin[1] + float(P.boundary_cells - 1) * P.dx;
  const double fcd = double(floor_h + P.contact_distance);
  for (int k = a.slot_face_offsets[s]; k < a.slot_face_offsets[s + 1]; ++k) {
    const int32_t* f = a.faces + 3 * a.slot_faces[k];
    double Pv[3][3];
    for (int q = 0; q < 3; ++q)
      for (int r = 0; r < 3; ++r) Pv[q][r] = double(a.vertices[3 * f[q] + r]);
    double e1[3], e2[3];
    for (int r = 0; r < 3; ++r) { e1[r] = Pv[1][r] - Pv[0][r]; e2[r] = Pv[2][r] - Pv[0][r]; }
    const double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
    const double len = sqrt(nx * nx + ny * ny + nz * nz);
    const double cy = (Pv[0][1] + Pv[1][1] + Pv[2][1]) / 3.0;
    const bool contact = cy <= fcd && ny < -0.2 * len;
    const double ga = contact ? gc : ge;
    if (ga == 0.0 || !(len > 0.0)) continue;
    const double u[3] = {nx / len, ny / len, nz / len};
    // d area / d e1 = (e2 x u) / 2, d area / d e2 = (u x e1) / 2
    const double d1[3] = {0.5 * (e2[1] * u[2] - e2[2] * u[1]), 0.5 * (e2[2] * u[0] - e2[0] * u[2]), 0.5 * (e2[0] * u[1] - e2[1] * u[0])};
    const double d2[3] = {0.5 * (u[1] * e1[2] - u[2] * e1[1]), 0.5 * (u[2] * e1[0] - u[0] * e1[2]), 0.5 * (u[0] * e1[1] - u[1] * e1[0])};
    for (int r = 0; r < 3; ++r) {
      unsafeAtomicAdd(a.gvert + 3 * f[1] + r, ga * d1[r]);
      unsafeAtomicAdd(a.gvert + 3 * f[2] + r, ga * d2[r]);
      unsafeAtomicAdd(a.gvert + 3 * f[0] + r, -ga * (d1[r] + d2[r]));
    }
  }
}

// Per particle: mesh reconstruction's adjoint (mesh_tracking.py:13-38),
// v = (1 - sum_b w_b) anchor + sum_b w_b (x_b + F_b r_b) with anchor = x_s + F_s r_s
// of the vertex's first binding, gathered over the particle's own bindings
// (the reverse index adj.rev_*): no atomics.
__global__ __launch_bounds__(256) void k_adj_bindings(AdjArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  double gx[3] = {0.0, 0.0, 0.0}, gF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool any = false;
  for (int idx = a.adj.rev_offsets[p]; idx < a.adj.rev_offsets[p + 1]; ++idx) {
    const int k = a.adj.rev_bindings[idx], v = a.adj.rev_vertices[idx];
    const double gv[3] = {a.gvert[3 * v], a.gvert[3 * v + 1], a.gvert[3 * v + 2]};
    if (gv[0] == 0.0 && gv[1] == 0.0 && gv[2] == 0.0) continue;
    const double c = double(a.bind_weights[k]) + (k == a.bind_offsets[v] ? a.adj.anchor_coef[v] : 0.0);
    const float* r = a.bind_ref_offsets + 3 * k;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      gx[q] += c * gv[q];
#pragma unroll
      for (int e = 0; e < 3; ++e) gF[3 * q + e] += c * gv[q] * double(r[e]);
    }
    any = true;
  }
  if (!any) return;
#pragma unroll
  for (int q = 0; q < 3; ++q) a.adj.x[3 * p + q] += gx[q];
#pragma unroll
  for (int i = 0; i < 9; ++i) a.adj.F[9 * p + i] += gF[i];
}

__global__ void k_adj_clear_vertices(AdjArgs a, int nv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * nv) a.gvert[i] = 0.0;
}

// ------------------------------------------------------------------ C ABI
extern "C" {

int64_t bk_adjoint_scratch_size(const bk_params* P) {
  if (!P || P->res < 4 || P->n < 1) return -1;
  const int64_t nb = (P->res + 3) / 4, nodes = nb * nb * nb * 64;
  int64_t size = 7 * nodes;
  if (P->transport_enabled) size += 10 * nodes + 2 * int64_t(P->n) + 2 * int64_t(P->n_slots) + 3 * int64_t(P->n_vertices);
  return size;
}

int bk_adjoint_substep(const bk_params* P, const bk_scene* sc, const bk_state* in, const bk_state* out, const bk_grid* grid,
                       bk_workspace* ws, const bk_adjoint* adj, double* scratch, double* param_grad, void* stream) {
  if (!P || !sc || !in || !out || !grid || !ws || !adj || !scratch || !param_grad) return fail("bk_adjoint_substep: null argument");
  const bk_dough* doughs = nullptr;
  const bk_workspace* wsv = nullptr;
  int n_doughs = 0, grid_cap = 0;
  if (bk_internal_tile_launch(ws, &doughs, &n_doughs, &grid_cap, &wsv)) return -1;
  if (n_doughs != 1) return fail("bk_adjoint_substep: the workspace must hold one dough (the simulator's own)");
  if (P->n < 1 || P->res < 4) return fail("bk_adjoint_substep: bad sizes");
  if (P->settling == 2)
    return fail("bk_adjoint_substep: pass the substep's phase (settling 0 baking, 1 resting), not the controller mode 2");
  if (P->model == 1 && (!sc->visc_param || !in->Fe || !adj->Fe || !adj->tau || !adj->porosity))
    return fail("bk_adjoint_substep: the Ding adjoint needs Fe, tau and porosity arrays");
  if (P->model == 1 && P->gas_enabled && (!P->transport_enabled || !in->moles || !out->moles || !adj->moles || !sc->prod_param))
    return fail("bk_adjoint_substep: the pore-gas adjoint needs transport, moles and the production parameter");
  const bool transport = P->transport_enabled != 0;
  if (transport && (!adj->T || !adj->W || !adj->losses || !adj->exchanges || !in->T || !in->W || !out->T || !out->W ||
                    !out->masses || !sc->capacities || !sc->heat_param || !grid->V || !grid->dT))
    return fail("bk_adjoint_substep: the transport adjoint needs T, W, losses and exchanges arrays and the thermal grid");
  if (transport && P->n_vertices > 0 && (!adj->rev_offsets || !adj->rev_bindings || !adj->rev_vertices || !adj->anchor_coef))
    return fail("bk_adjoint_substep: the transport adjoint needs the reverse binding index (adj.rev_*, adj.anchor_coef)");
  hipStream_t st = (hipStream_t)stream;
  AdjArgs a = {};
  a.P = *P;
  if (a.P.settling == 1) {  // a resting substep (thermal_water.py:306, 325-326, 333)
    a.P.tray_h = 0.0f;
    a.P.loss_coef = 0.0f;
    a.P.kappa = 0.0f;
    a.P.diffusivity = 0.0f;
  }
  a.in = *in;
  a.out = *out;
  a.g = *grid;
  a.masses = sc->m0;
  a.V0 = sc->V0;
  a.visc_param = sc->visc_param;
  a.prod_param = sc->prod_param;
  a.adj = *adj;
  a.nb = (P->res + 3) / 4;
  a.n = P->n;
  const int64_t nodes = int64_t(a.nb) * a.nb * a.nb * 64;
  a.gm = scratch;
  a.gp = scratch + nodes;
  a.gv = scratch + 4 * nodes;
  a.param_grad = param_grad;
  if (transport) {
    double* t = scratch + 7 * nodes;
    double** grid_arrays[10] = {&a.accT, &a.accM, &a.ghr, &a.gwr, &a.gV, &a.gC, &a.gE, &a.gW, &a.gH, &a.gL};
    for (int k = 0; k < 10; ++k) { *grid_arrays[k] = t; t += nodes; }
    a.tcbar = t; t += P->n;
    a.mrbar = t; t += P->n;
    a.gSexp = t; t += P->n_slots;
    a.gScon = t; t += P->n_slots;
    a.gvert = t;
    a.capacities = sc->capacities;
    a.heat_param = sc->heat_param;
    a.particle_slot = sc->particle_slot;
    a.slot_face_offsets = sc->slot_face_offsets;
    a.slot_faces = sc->slot_faces;
    a.faces = sc->faces;
    a.bind_offsets = sc->bind_offsets;
    a.bind_particles = sc->bind_particles;
    a.bind_weights = sc->bind_weights;
    a.bind_ref_offsets = sc->bind_ref_offsets;
    a.slot_exposed = wsv->slot_exposed;
    a.slot_contact = wsv->slot_contact;
    a.vertices = wsv->vertices;
  }
  if (P->model == 1) hipLaunchKernelGGL(k_adj_ding, dim3(div_up(a.n, 128)), dim3(128), 0, st, a);
  hipLaunchKernelGGL(k_adj_g2p, dim3(div_up(a.n, 128)), dim3(128), 0, st, a);
  hipLaunchKernelGGL(k_adj_scatter, dim3(grid_cap), dim3(kTileThreads), 0, st, doughs, a, ws);
  hipLaunchKernelGGL(k_adj_grid, dim3(div_up(nodes, 256)), dim3(256), 0, st, a, nodes);
  hipLaunchKernelGGL(k_adj_p2g, dim3(div_up(a.n, 128)), dim3(128), 0, st, a);
  if (transport) {
    hipLaunchKernelGGL(k_adj_b1, dim3(div_up(a.n, 128)), dim3(128), 0, st, a);
    hipLaunchKernelGGL(k_adj_scatter_b1, dim3(grid_cap), dim3(kTileThreads), 0, st, doughs, a, ws);
    hipLaunchKernelGGL(k_adj_diffuse_a, dim3(grid_cap), dim3(256), 0, st, a, ws);
    hipLaunchKernelGGL(k_adj_diffuse_b, dim3(grid_cap), dim3(256), 0, st, a, ws);
    hipLaunchKernelGGL(k_adj_pass_a, dim3(div_up(a.n, 128)), dim3(128), 0, st, a);
    if (P->n_slots > 0) hipLaunchKernelGGL(k_adj_faces, dim3(div_up(P->n_slots, 256)), dim3(256), 0, st, a, P->n_slots);
    if (P->n_vertices > 0) {
      hipLaunchKernelGGL(k_adj_bindings, dim3(div_up(a.n, 256)), dim3(256), 0, st, a);
      hipLaunchKernelGGL(k_adj_clear_vertices, dim3(div_up(3 * int64_t(P->n_vertices), 256)), dim3(256), 0, st, a, P->n_vertices);
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : fail("bk_adjoint_substep: launch failed");
}

int bk_overstress_adjoint(const double* args4, const double* gbar, double* out4, int32_t n, void* stream);

}  // extern "C"

// IFT adjoint of the overstress solve alone (ding.py:90-100), for scalar checks.
__global__ void k_overstress_adjoint(const double* args4, const double* gbar, double* out4, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q2 = args4[4 * i], p2 = args4[4 * i + 1], df = args4[4 * i + 2], mf = args4[4 * i + 3];
  double r[4] = {0, 0, 0, 0};
  if (q2 + p2 > 1.0) {
    double g = 0.0;
    for (int it = 0; it < 64; ++it) {
      const double u = 1.0 + df * g, w = 1.0 + mf * g;
      const double res = g + 1.0 - q2 / (u * u) - p2 / (w * w);
      const double der = 1.0 + 2.0 * df * q2 / (u * u * u) + 2.0 * mf * p2 / (w * w * w);
      const double next = g - res / der;
      if (next == g) break;
      g = next;
    }
    const double u = 1.0 + df * g, w = 1.0 + mf * g;
    const double der = 1.0 + 2.0 * df * q2 / (u * u * u) + 2.0 * mf * p2 / (w * w * w);
    const double gb = gbar[i];
    r[0] = gb / (der * u * u);
    r[1] = gb / (der * w * w);
    r[2] = -gb * 2.0 * q2 * g / (der * u * u * u);
    r[3] = -gb * 2.0 * p2 * g / (der * w * w * w);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) out4[4 * i + k] = r[k];
}

extern "C" int bk_overstress_adjoint(const double* args4, const double* gbar, double* out4, int32_t n, void* stream) {
  if (n < 0 || !args4 || !gbar || !out4) return fail("bk_overstress_adjoint: bad argument");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_overstress_adjoint, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, args4, gbar, out4, n);
  return hipGetLastError() == hipSuccess ? 0 : fail("bk_overstress_adjoint: launch failed");
}

// Adjoint of UpdateDing for particles without pore gas (one thread each).
__global__ __launch_bounds__(128) void k_update_ding_adjoint(DingParams D, const float* old_Fe, const float* old_porosity, const int32_t* old_gelled,
                                      const float* F_total, const float* C_affine, const float* temperatures,
                                      const float* viscosity_parameter, const double* adj_Fe, const double* adj_tau,
                                      const double* adj_porosity, double* adj_old_Fe, double* adj_old_porosity, double* adj_F_total,
                                      double* adj_C_affine, double* adj_viscosity_parameter, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double geta = 0.0;
  if (p < n) {
    double Fe_old[9], Fn[9], Cn[9], gFe[9], gtau[9], gFe_old[9], gF[9], gC[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      Fe_old[i] = old_Fe[9 * p + i];
      Fn[i] = F_total[9 * p + i];
      Cn[i] = C_affine[9 * p + i];
      gFe[i] = adj_Fe ? adj_Fe[9 * p + i] : 0.0;
      gtau[i] = adj_tau ? adj_tau[9 * p + i] : 0.0;
      gFe_old[i] = gF[i] = gC[i] = 0.0;
    }
    double gpor = 0.0, gT = 0.0;
    ding_return_adjoint(D, Fe_old, Fn, Cn, double(old_porosity[p]), 0.0, old_gelled[p], double(temperatures[p]),
                        double(viscosity_parameter[0]), gFe, gtau, adj_porosity ? adj_porosity[p] : 0.0, gFe_old, gF, gC, gpor, geta,
                        gT);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      if (adj_old_Fe) adj_old_Fe[9 * p + i] += gFe_old[i];
      if (adj_F_total) adj_F_total[9 * p + i] += gF[i];
      if (adj_C_affine) adj_C_affine[9 * p + i] += gC[i];
    }
    if (adj_old_porosity) adj_old_porosity[p] += gpor;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) geta += __shfl_xor(geta, off, 64);
  if (adj_viscosity_parameter && __lane_id() == 0 && geta != 0.0) unsafeAtomicAdd(adj_viscosity_parameter, geta);
}

extern "C" int bk_UpdateDing_adjoint(const float* old_Fe, const float* old_porosity, const int32_t* old_gelled, const float* F_total,
                                     const float* C_affine, const float* temperatures, const float* viscosity_parameter,
                                     const double* adj_Fe, const double* adj_tau, const double* adj_porosity, double* adj_old_Fe,
                                     double* adj_old_porosity, double* adj_F_total, double* adj_C_affine,
                                     double* adj_viscosity_parameter, float mu, float lam, float yield_stress, float n0,
                                     int32_t temperature_enabled, const float* knots4_host, const float* yield4_host,
                                     const float* visc4_host, float dt, int32_t n, void* stream) {
  if (n < 0) return fail("bk_UpdateDing_adjoint: negative count");
  if (n == 0) return 0;
  if (!old_Fe || !old_porosity || !old_gelled || !F_total || !C_affine || !temperatures || !viscosity_parameter)
    return fail("bk_UpdateDing_adjoint: null input");
  if (temperature_enabled && (!knots4_host || !yield4_host || !visc4_host))
    return fail("bk_UpdateDing_adjoint: temperature tables missing");
  DingParams D;
  D.mu = mu; D.lam = lam; D.yield_stress = yield_stress; D.n0 = n0; D.dt = dt;
  for (int i = 0; i < 4; ++i) {
    D.knots[i] = temperature_enabled ? knots4_host[i] : 0.0f;
    D.yf[i] = temperature_enabled ? yield4_host[i] : 1.0f;
    D.vf[i] = temperature_enabled ? visc4_host[i] : 1.0f;
  }
  D.temperature_enabled = temperature_enabled;
  D.coupled = 0;
  D.bracketed = 0;
  hipLaunchKernelGGL(k_update_ding_adjoint, dim3(div_up(n, 128)), dim3(128), 0, (hipStream_t)stream, D, old_Fe, old_porosity,
                     old_gelled, F_total, C_affine, temperatures, viscosity_parameter, adj_Fe, adj_tau, adj_porosity, adj_old_Fe,
                     adj_old_porosity, adj_F_total, adj_C_affine, adj_viscosity_parameter, n);
  return hipGetLastError() == hipSuccess ? 0 : fail("bk_UpdateDing_adjoint: launch failed");
}

// Adjoint of the coupled porosity solve alone (porous_return.py:64-100), at the bracketed solve's root.
__global__ __launch_bounds__(256) void k_solve_porosity_adjoint(const float* args6, const double* pv, const double* gbar, double* out7, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* q = args6 + 6 * i;
  const float root = solve_porosity_bracketed(q[0], q[1], q[2], q[3], q[4], pv[i], q[5]);
  double ac, dac;
  coefficient_and_slope(double(root), ac, dac);
  const double g0 = overstress_converged(q[0], ac * double(q[1]), q[2], double(q[3]) * ac);
  const PorosityDerivs D = porosity_root_derivatives(double(root), g0, q[0], q[1], q[2], q[3], q[4], pv[i]);
  const double g = gbar[i];
  const double r[7] = {D.q, D.p, D.df, D.ms, D.ex, D.n0, D.pv};
#pragma unroll
  for (int k = 0; k < 7; ++k) out7[7 * i + k] = g * r[k];
}

extern "C" int bk_solve_porosity_adjoint(const float* args6, const double* plastic_volume, const double* gbar, double* out7, int32_t n,
                                         void* stream) {
  if (n < 0 || (n > 0 && (!args6 || !plastic_volume || !gbar || !out7))) return fail("bk_solve_porosity_adjoint: bad argument");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_solve_porosity_adjoint, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, args6, plastic_volume, gbar,
                     out7, n);
  return hipGetLastError() == hipSuccess ? 0 : fail("bk_solve_porosity_adjoint: launch failed");
}
