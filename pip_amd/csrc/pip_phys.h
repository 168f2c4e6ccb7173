// pip_phys.h -- per-cell and per-face physics of the PIP two-fluid step,
// written once for registers and used by every HIP kernel of libpipgpu.
//
// Each function restates one reference routine for a single cell or face;
// the evaluation order of every expression follows the reference so that the
// device path reproduces the Fortran arithmetic operation for operation
// (built without FMA contraction; quotients below).
// State order: U_m = (rho, mx, my, mz, E, Bx, By, Bz[, psi]),
//              U_h = (rho, mx, my, mz, E)   (IO_rot.f90:362-374).
#pragma once

#if defined(__HIPCC__)
#define PIP_HD __host__ __device__ __forceinline__
#else
#define PIP_HD inline
#endif

#include <math.h>

// PIP_NS: the namespace of everything below (pip; a second build of the
// same headers uses another, so that both sets of kernels live in one
// library).  PIP_RELAXED = 1 selects the relaxed device arithmetic of that
// build (see "Relaxed arithmetic" below); 0, the default, the exact one.
#ifndef PIP_NS
#define PIP_NS pip
#endif
#ifndef PIP_RELAXED
#define PIP_RELAXED 0
#endif
#if PIP_RELAXED && defined(__HIP_DEVICE_COMPILE__)
#define PIP_RELAXED_DEVICE 1
#else
#define PIP_RELAXED_DEVICE 0
#endif

namespace PIP_NS {

constexpr double kTiny = 1.0e-20;  // Globalvar.f90:101
constexpr double kEpsi = 1.0e-10;  // HLL_rot.f90:12
constexpr double kAcc = 1.03;      // HLL_rot.f90:12
constexpr double kHllg = 1.001;    // HLL_rot.f90:701
constexpr double kMuP = 0.5;       // PIP_rot.f90:36
constexpr double kMuN = 1.0;       // PIP_rot.f90:37

// MIN / MAX of the reference.  The device uses the hardware instructions
// (v_min_f64 / v_max_f64: one instruction instead of a compare and two
// selects).  For operands that are not NaN they return the same value as
// the compare-and-select of the host build, except which zero comes back
// for min(-0, +0) and max(+0, -0); on this path such zeros only arise in
// variables that are identically zero (B_z, m_z of a planar flow), where
// no later operation distinguishes them, so states compare equal.
#if defined(__HIP_DEVICE_COMPILE__)
PIP_HD double dmin(double a, double b) { return fmin(a, b); }
PIP_HD double dmax(double a, double b) { return fmax(a, b); }
#else
PIP_HD double dmin(double a, double b) { return (a < b) ? a : b; }
PIP_HD double dmax(double a, double b) { return (a > b) ? a : b; }
#endif
PIP_HD double sgl(double x) { return (double)(float)x; }  // Fortran default-real literal

// ------------------------------------------------------------ quotients
// quo(n, d, qrcp(d)) is n / d.  On the device it runs the steps of the
// general double division sequence -- v_rcp_f64, two Newton-Raphson steps
// r += r (1 - d r), q = n r, q += r (n - d q), v_div_fixup_f64 -- without
// its two v_div_scale_f64 range scalings.  Those are the identity unless an
// operand or the quotient lies at the ends of the exponent range (|n| below
// 2^-969, d or 1/d subnormal, a subnormal quotient or |n/d| near 2^768);
// zero, infinite and NaN operands are settled by the fixup alone.  Outside
// those ends the result is therefore bitwise the IEEE quotient.  Quotients
// that share a denominator share its refined reciprocal: 4 instructions
// each instead of 11.  The host build divides.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double qrcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}
#if PIP_RELAXED
// relaxed: the product with the refined reciprocal (within an ulp or two of
// the quotient; no residual correction, no fixup)
__device__ __forceinline__ double quo(double n, double, double r) { return n * r; }
#else
__device__ __forceinline__ double quo(double n, double d, double r) {
  const double q = n * r;
  return __builtin_amdgcn_div_fixup(__builtin_fma(r, __builtin_fma(-d, q, n), q), d, n);
}
#endif
// qsqrt(x) is sqrt(x) the same way: the steps of the general double square
// root sequence -- v_rsq_f64 y, g = x y, h = y/2, r = 1/2 - h g, g += g r,
// h += h r, two corrections g += h (x - g g), x itself for +-0 and +inf --
// without its scaling of arguments below 2^-767.
__device__ __forceinline__ double qsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
  g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
  return __builtin_amdgcn_class(x, 0x260) ? x : g;  // 0x260: -0, +0, +inf
}
#else
PIP_HD double qrcp(double) { return 0.0; }
PIP_HD double quo(double n, double d, double) { return n / d; }
PIP_HD double qsqrt(double x) { return sqrt(x); }
#endif
PIP_HD double qdiv(double n, double d) { return quo(n, d, qrcp(d)); }

// fdiv(n, d) and fsqrt(x): the quotients and square roots of the set_IR,
// get_Pr and cfl terms.  Exact build: the IEEE operations (general sequences,
// valid over the whole range, as the rates can underflow); relaxed device
// build: the refined reciprocal times n and qsqrt.
#if PIP_RELAXED_DEVICE
PIP_HD double fdiv(double n, double d) { return n * qrcp(d); }
PIP_HD double fsqrt(double x) { return qsqrt(x); }
#else
PIP_HD double fdiv(double n, double d) { return n / d; }
PIP_HD double fsqrt(double x) { return sqrt(x); }
#endif

// ------------------------------------------------------ exp / log / pow
// exp, log and pow as fixed sequences of IEEE double additions and
// multiplications (no FMA, no libm), so that every compiler and device
// returns the same bits.  Accuracy: a few ulp of the correctly rounded
// value.  exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-13
// Taylor polynomial (truncation < 1e-17), scaling by 2^k.  log: x = m 2^e
// with m in [1/sqrt2, sqrt2), log m = 2 atanh(s), s = (m-1)/(m+1), series to
// s^25.  pow(x,y) = exp(y log x) for x > 0.
constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
constexpr double kLn2Lo = 1.90821492927058770002e-10;  // 0x3dea39ef35793c76
constexpr double kInvLn2 = 1.44269504088896338700e+00;

PIP_HD double pexp(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893383973096) return HUGE_VAL;
  if (x < -745.133219101941108420) return 0.0;
  double kd = rint(x * kInvLn2);
  double r = (x - kd * kLn2Hi) - kd * kLn2Lo;
  double q = 1.0 / 6227020800.0;
  q = q * r + 1.0 / 479001600.0;
  q = q * r + 1.0 / 39916800.0;
  q = q * r + 1.0 / 3628800.0;
  q = q * r + 1.0 / 362880.0;
  q = q * r + 1.0 / 40320.0;
  q = q * r + 1.0 / 5040.0;
  q = q * r + 1.0 / 720.0;
  q = q * r + 1.0 / 120.0;
  q = q * r + 1.0 / 24.0;
  q = q * r + 1.0 / 6.0;
  q = q * r + 0.5;
  q = q * r + 1.0;
  q = q * r + 1.0;
  return ldexp(q, (int)kd);
}

PIP_HD double plog(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? -HUGE_VAL : (x - x) / (x - x);
  if (x == HUGE_VAL) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = fdiv(m - 1.0, m + 1.0);
  double s2 = s * s;
  double t = 1.0 / 25.0;
  t = t * s2 + 1.0 / 23.0;
  t = t * s2 + 1.0 / 21.0;
  t = t * s2 + 1.0 / 19.0;
  t = t * s2 + 1.0 / 17.0;
  t = t * s2 + 1.0 / 15.0;
  t = t * s2 + 1.0 / 13.0;
  t = t * s2 + 1.0 / 11.0;
  t = t * s2 + 1.0 / 9.0;
  t = t * s2 + 1.0 / 7.0;
  t = t * s2 + 1.0 / 5.0;
  t = t * s2 + 1.0 / 3.0;
  t = t * s2 + 1.0;
  double lm = (2.0 * s) * t;
  double ed = (double)e;
  return ed * kLn2Hi + (ed * kLn2Lo + lm);
}

PIP_HD double ppow(double x, double y) {
  if (x == 0.0) return (y > 0.0) ? 0.0 : HUGE_VAL;
  return pexp(y * plog(x));
}

// Physical constants of one run (the reference's module globals).
struct Phys {
  double gm, ro_lim, pr_lim, safety;
  double T0, n0, beta, t_ir;
  int flag_col;      // 1 constant alpha; 2..7 set_collisional (parity selects the IR normalisation)
  int flag_ir;       // 0 or 3
  int flag_ir_type;  // 0, 1, 2
  int pip_imp;       // 1: implicit collision source (PIP_rot.f90:1333-1358)
};

constexpr double kPi = 3.1415926535897932;  // Parameters.f90:2

// set_collisional col_type (PIP_rot.f90:21-25): flag_col - 1 for 2..7.
PIP_HD int col_type(const Phys& p) { return (p.flag_col >= 2 && p.flag_col <= 7) ? p.flag_col - 1 : 0; }

// Stage coefficients: Solver_rot.f90:223-231 (float() conversions kept).
PIP_HD void stage_coeffs(int istep, int t_order, double dt, double& c0, double& c1,
                         double& dt_sub) {
  if (istep == 1) {
    c1 = 1.0; c0 = 1.0; dt_sub = dt;
  } else {
    c1 = (double)((float)(istep - 1) / (float)(2 * t_order - istep));
    c0 = 1.0 - c1;
    dt_sub = c1 * dt;
  }
}

// ---------------------------------------------------------------- limiter
// minmod_func: Util_rot.f90:59-76
PIP_HD double minmod(double u1, double u2, double u3) {
  double X = u3 - u2, Y = u2 - u1;
  double SX = copysign(1.0, X);
  return SX * dmax(0.0, dmin(fabs(X), SX * Y));
}

// ------------------------------------------------------------ conversions
// u2v_mhd, one cell: HLL_rot.f90:1159-1186
template <int NV>
PIP_HD void u2v_mhd(const double* u, double* v, const Phys& p) {
  double v1 = dmax(u[0], p.ro_lim);
  double roi = qdiv(1.0, v1);
  v[0] = v1;
  v[1] = u[1] * roi;
  v[2] = u[2] * roi;
  v[3] = u[3] * roi;
  double pr = (p.gm - 1.0) * ((u[4] - (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) * roi) -
                              0.5 * ((u[5] * u[5] + u[6] * u[6]) + u[7] * u[7]));
  if (pr < p.pr_lim) pr = p.pr_lim;
  v[4] = pr;
  v[5] = u[5];
  v[6] = u[6];
  v[7] = u[7];
  if (NV == 9) v[8] = u[8];
}

// u2v_hd, one cell: HLL_rot.f90:1136-1156
PIP_HD void u2v_hd(const double* u, double* v, const Phys& p) {
  double v1 = dmax(u[0], p.ro_lim);
  double roi = qdiv(1.0, v1);
  v[0] = v1;
  v[1] = u[1] * roi;
  v[2] = u[2] * roi;
  v[3] = u[3] * roi;
  double pr = (p.gm - 1.0) * (u[4] - (0.5 * ((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3])) * roi);
  if (pr < p.pr_lim) pr = p.pr_lim;
  v[4] = pr;
}

// cq2pv_mhd, one cell, rewrites u in place: Scheme_rot.f90:153-179.
// prim = (de, vx, vy, vz, pr, bx, by, bz) after the floors.
PIP_HD void cq2pv_mhd(double* u, double* prim, const Phys& p) {
  double de = u[0];
  const double rde = qrcp(de);
  double vx = quo(u[1], de, rde), vy = quo(u[2], de, rde), vz = quo(u[3], de, rde);
  double bx = u[5], by = u[6], bz = u[7];
  double pr = (p.gm - 1.0) * ((u[4] - (0.5 * de) * ((vx * vx + vy * vy) + vz * vz)) -
                              0.5 * ((bx * bx + by * by) + bz * bz));
  pr = dmax(pr, p.pr_lim);
  de = dmax(de, p.ro_lim);
  u[0] = de;
  u[1] = vx * de;
  u[2] = vy * de;
  u[3] = vz * de;
  u[4] = (qdiv(pr, p.gm - 1.0) + (0.5 * de) * ((vx * vx + vy * vy) + vz * vz)) +
         0.5 * ((bx * bx + by * by) + bz * bz);
  prim[0] = de; prim[1] = vx; prim[2] = vy; prim[3] = vz;
  prim[4] = pr; prim[5] = bx; prim[6] = by; prim[7] = bz;
}

// cq2pv_hd, one cell, rewrites u: Scheme_rot.f90:132-150
PIP_HD void cq2pv_hd(double* u, double* prim, const Phys& p) {
  double de = u[0];
  const double rde = qrcp(de);
  double vx = quo(u[1], de, rde), vy = quo(u[2], de, rde), vz = quo(u[3], de, rde);
  double pr = (p.gm - 1.0) * (u[4] - (0.5 * de) * ((vx * vx + vy * vy) + vz * vz));
  pr = dmax(pr, p.pr_lim);
  de = dmax(de, p.ro_lim);
  u[0] = de;
  u[1] = vx * de;
  u[2] = vy * de;
  u[3] = vz * de;
  u[4] = qdiv(pr, p.gm - 1.0) + (0.5 * de) * ((vx * vx + vy * vy) + vz * vz);
  prim[0] = de; prim[1] = vx; prim[2] = vy; prim[3] = vz; prim[4] = pr;
}

// get_Pr_HD (+get_Te_HD), rewrites u[4]: Scheme_rot.f90:83-102
PIP_HD void getpr_hd(double* u, double& pr, double& te, const Phys& p) {
  double s = (u[1] * u[1] + u[2] * u[2]) + u[3] * u[3];
  pr = (p.gm - 1.0) * (u[4] - fdiv(0.5 * s, u[0]));
  pr = dmax(pr, p.pr_lim);
  u[4] = fdiv(pr, p.gm - 1.0) + 0.5 * fdiv(s, u[0]);
  te = p.gm * fdiv(pr, u[0]);
}

// get_Pr_MHD (+get_Te_MHD), rewrites u[4]: Scheme_rot.f90:104-128
PIP_HD void getpr_mhd(double* u, double& pr, double& te, const Phys& p) {
  double q = ((fdiv((u[1] * u[1] + u[2] * u[2]) + u[3] * u[3], u[0]) + u[5] * u[5]) + u[6] * u[6]) +
             u[7] * u[7];
  pr = (p.gm - 1.0) * (u[4] - 0.5 * q);
  pr = dmax(pr, p.pr_lim);
  u[4] = fdiv(pr, p.gm - 1.0) + 0.5 * q;
  te = (0.5 * p.gm) * fdiv(pr, u[0]);
}

// v2u_1: HLL_rot.f90:1189-1215
// rg1 = qrcp(gm - 1.0).
PIP_HD void v2u_hd(const double* v, double* u, double gm, double rg1) {
  u[0] = v[0];
  u[1] = v[0] * v[1];
  u[2] = v[0] * v[2];
  u[3] = v[0] * v[3];
  u[4] = quo(v[4], gm - 1.0, rg1) + (0.5 * v[0]) * ((v[1] * v[1] + v[2] * v[2]) + v[3] * v[3]);
}

template <int NV>
PIP_HD void v2u_mhd(const double* v, double* u, double gm, double rg1) {
  u[0] = v[0];
  u[1] = v[0] * v[1];
  u[2] = v[0] * v[2];
  u[3] = v[0] * v[3];
  u[5] = v[5]; u[6] = v[6]; u[7] = v[7];
  u[4] = (quo(v[4], gm - 1.0, rg1) + (0.5 * v[0]) * ((v[1] * v[1] + v[2] * v[2]) + v[3] * v[3])) +
         0.5 * ((v[5] * v[5] + v[6] * v[6]) + v[7] * v[7]);
  if (NV == 9) u[8] = v[8];
}

// ---------------------------------------------------------- Riemann solvers
// hllc_flux: HLL_rot.f90:615-687
PIP_HD void hllc_flux(const double* v_l, const double* v_r, double gm, double* flx) {
  double u_l[5], u_r[5], flx_l[5], flx_r[5], st_l[5], st_r[5];
  const double rg1 = qrcp(gm - 1.0);
  v2u_hd(v_l, u_l, gm, rg1);
  v2u_hd(v_r, u_r, gm, rg1);
  const double ra = qrcp(u_l[0]), rb = qrcp(u_r[0]);  // every quotient by rho_L, rho_R
  double pr_l = v_l[4], pr_r = v_r[4];
  flx_l[0] = u_l[1];
  flx_r[0] = u_r[1];
  flx_l[1] = pr_l + quo(u_l[1] * u_l[1], u_l[0], ra);
  flx_r[1] = pr_r + quo(u_r[1] * u_r[1], u_r[0], rb);
  flx_l[2] = quo(u_l[1] * u_l[2], u_l[0], ra);
  flx_r[2] = quo(u_r[1] * u_r[2], u_r[0], rb);
  flx_l[3] = quo(u_l[1] * u_l[3], u_l[0], ra);
  flx_r[3] = quo(u_r[1] * u_r[3], u_r[0], rb);
  flx_l[4] = quo(u_l[1] * (u_l[4] + pr_l), u_l[0], ra);
  flx_r[4] = quo(u_r[1] * (u_r[4] + pr_r), u_r[0], rb);
  double cs_l = qsqrt(quo(gm * pr_l, u_l[0], ra));
  double cs_r = qsqrt(quo(gm * pr_r, u_r[0], rb));
  double vn_l = quo(u_l[1], u_l[0], ra);
  double vn_r = quo(u_r[1], u_r[0], rb);
  double sl = kAcc * dmin(vn_l - cs_l, vn_r - cs_r);
  double sr = kAcc * dmax(vn_l + cs_l, vn_r + cs_r);
  double sm = qdiv(u_r[1] * (sr - vn_r) - u_l[1] * (sl - vn_l) + pr_l - pr_r,
                   u_r[0] * (sr - vn_r) - u_l[0] * (sl - vn_l));
  double pr_m = sm * u_l[0] * (sl - vn_l) - u_l[1] * (sl - vn_l) + pr_l;
  const double rsl = qrcp(sl - sm), rsr = qrcp(sr - sm);
  st_l[0] = quo(u_l[0] * (sl - vn_l), sl - sm, rsl);
  st_r[0] = quo(u_r[0] * (sr - vn_r), sr - sm, rsr);
  st_l[1] = st_l[0] * sm;
  st_r[1] = st_r[0] * sm;
  st_l[2] = quo(st_l[0] * u_l[2], u_l[0], ra);
  st_r[2] = quo(st_r[0] * u_r[2], u_r[0], rb);
  st_l[3] = quo(st_l[0] * u_l[3], u_l[0], ra);
  st_r[3] = quo(st_r[0] * u_r[3], u_r[0], rb);
  st_l[4] = quo((sl - vn_l) * u_l[4] - pr_l * vn_l + pr_m * sm, sl - sm, rsl);
  st_r[4] = quo((sr - vn_r) * u_r[4] - pr_r * vn_r + pr_m * sm, sr - sm, rsr);
  double thl = qdiv(fabs(sl) + sl, dmax(2.0 * sl, kEpsi));
  double thm = qdiv(fabs(sm) + sm, dmax(2.0 * sm, kEpsi));
  double thr = 1.0 - qdiv(fabs(sr) + sr, dmax(2.0 * sr, kEpsi));
#pragma unroll
  for (int n = 0; n < 5; n++)
    flx[n] = thr * flx_r[n] + thl * flx_l[n] +
             (1.0 - thl - thr) * (thm * (flx_l[n] + sl * (st_l[n] - u_l[n])) +
                                  (1.0 - thm) * (flx_r[n] + sr * (st_r[n] - u_r[n])));
}

// hlld_flux (OpenMHD form with HLLC-G fallback): HLL_rot.f90:691-910.
// NV = 8 (no psi) or 9 (GLM psi flux with ch = cmax, :904-908).
template <int NV>
PIP_HD void hlld_flux(const double* v_l, const double* v_r, double gm, double cmax, double* flx) {
  double u_l[NV], u_r[NV], flx_l[8], flx_r[8], u_hll[8];
  const double rg1 = qrcp(gm - 1.0);
  v2u_mhd<NV>(v_l, u_l, gm, rg1);
  v2u_mhd<NV>(v_r, u_r, gm, rg1);

  double b2_l = (v_l[5] * v_l[5] + v_l[6] * v_l[6]) + v_l[7] * v_l[7];
  double pt_l = v_l[4] + 0.5 * b2_l;
  double vbL = (v_l[1] * v_l[5] + v_l[2] * v_l[6]) + v_l[3] * v_l[7];
  flx_l[0] = v_l[0] * v_l[1];
  flx_l[1] = ((v_l[0] * (v_l[1] * v_l[1]) + v_l[4]) + 0.5 * b2_l) - v_l[5] * v_l[5];
  flx_l[2] = v_l[0] * v_l[1] * v_l[2] - v_l[5] * v_l[6];
  flx_l[3] = v_l[0] * v_l[1] * v_l[3] - v_l[5] * v_l[7];
  flx_l[4] = (u_l[4] + pt_l) * v_l[1] - vbL * v_l[5];
  flx_l[5] = 0.0;
  flx_l[6] = v_l[1] * v_l[6] - v_l[5] * v_l[2];
  flx_l[7] = v_l[1] * v_l[7] - v_l[5] * v_l[3];

  double b2_r = (v_r[5] * v_r[5] + v_r[6] * v_r[6]) + v_r[7] * v_r[7];
  double pt_r = v_r[4] + 0.5 * b2_r;
  double vbR = (v_r[1] * v_r[5] + v_r[2] * v_r[6]) + v_r[3] * v_r[7];
  flx_r[0] = v_r[0] * v_r[1];
  flx_r[1] = ((v_r[0] * (v_r[1] * v_r[1]) + v_r[4]) + 0.5 * b2_r) - v_r[5] * v_r[5];
  flx_r[2] = v_r[0] * v_r[1] * v_r[2] - v_r[5] * v_r[6];
  flx_r[3] = v_r[0] * v_r[1] * v_r[3] - v_r[5] * v_r[7];
  flx_r[4] = (u_r[4] + pt_r) * v_r[1] - vbR * v_r[5];
  flx_r[5] = 0.0;
  flx_r[6] = v_r[1] * v_r[6] - v_r[5] * v_r[2];
  flx_r[7] = v_r[1] * v_r[7] - v_r[5] * v_r[3];

  double f1 = gm * v_l[4];
  double f2 = 4.0 * f1 * (v_l[5] * v_l[5]);
  double vf_l = qsqrt(qdiv((f1 + b2_l) + qsqrt(dmax((f1 + b2_l) * (f1 + b2_l) - f2, 0.0)), 2.0 * v_l[0]));
  f1 = gm * v_r[4];
  f2 = 4.0 * f1 * (v_r[5] * v_r[5]);
  double vf_r = qsqrt(qdiv((f1 + b2_r) + qsqrt(dmax((f1 + b2_r) * (f1 + b2_r) - f2, 0.0)), 2.0 * v_r[0]));

  double sl = dmin(v_l[1], v_r[1]) - dmax(vf_l, vf_r);
  double sr = dmax(v_l[1], v_r[1]) + dmax(vf_l, vf_r);

  if (sl >= 0.0) {
#pragma unroll
    for (int n = 0; n < 8; n++) flx[n] = flx_l[n];
  } else if (sr <= 0.0) {
#pragma unroll
    for (int n = 0; n < 8; n++) flx[n] = flx_r[n];
  } else {
    f1 = qdiv(1.0, sr - sl);
#pragma unroll
    for (int n = 0; n < 8; n++) u_hll[n] = f1 * (((sr * u_r[n] - sl * u_l[n]) - flx_r[n]) + flx_l[n]);
    const double rh = qrcp(u_hll[0]);
    double sm = quo(u_hll[1], u_hll[0], rh);
    double pt = pt_l + v_l[0] * (sl - v_l[1]) * (sm - v_l[1]);
    const double rsl = qrcp(sl - sm), rsr = qrcp(sr - sm);  // ro_l1 / u_l1[4], ro_r1 / u_r1[4]
    double ro_l1 = quo(v_l[0] * (sl - v_l[1]), sl - sm, rsl);
    double ro_r1 = quo(v_r[0] * (sr - v_r[1]), sr - sm, rsr);
    double sal = qdiv(fabs(u_hll[5]), qsqrt(ro_l1));
    double sar = qdiv(fabs(u_hll[5]), qsqrt(ro_r1));
    if ((sl >= (sm - kHllg * sal)) || ((sm + kHllg * sar) >= sr)) {
      // revert to HLLC-G :775-807
      double vy_2 = quo(u_hll[2], u_hll[0], rh);
      double vz_2 = quo(u_hll[3], u_hll[0], rh);
      const bool left = (sm >= 0.0);
      const double* v = left ? v_l : v_r;
      const double* uu = left ? u_l : u_r;
      const double* fl = left ? flx_l : flx_r;
      double s = left ? sl : sr;
      double ro1 = left ? ro_l1 : ro_r1;
      double ptk = left ? pt_l : pt_r;
      double vb = left ? vbL : vbR;
      double u2[5];
      u2[4] = quo((((s - v[1]) * uu[4] - ptk * v[1]) + pt * sm) +
                      u_hll[5] * (((vb - sm * u_hll[5]) - vy_2 * u_hll[6]) - vz_2 * u_hll[7]),
                  s - sm, left ? rsl : rsr);
      u2[1] = ro1 * sm;
      u2[2] = ro1 * vy_2;
      u2[3] = ro1 * vz_2;
      flx[0] = s * (ro1 - v[0]) + fl[0];
#pragma unroll
      for (int n = 1; n < 5; n++) flx[n] = s * (u2[n] - uu[n]) + fl[n];
#pragma unroll
      for (int n = 5; n < 8; n++) flx[n] = s * (u_hll[n] - v[n]) + fl[n];
    } else {
      // HLLD intermediate states :810-843
      double bh2 = u_hll[5] * u_hll[5];
      double u_l1[8], u_r1[8];
      f1 = qdiv(1.0, v_l[0] * (sl - v_l[1]) * (sl - sm) - bh2);
      u_l1[5] = u_hll[5];
      u_l1[6] = v_l[6] * f1 * (v_l[0] * ((sl - v_l[1]) * (sl - v_l[1])) - bh2);
      u_l1[7] = v_l[7] * f1 * (v_l[0] * ((sl - v_l[1]) * (sl - v_l[1])) - bh2);
      double vy_l1 = v_l[2] - f1 * u_hll[5] * v_l[6] * (sm - v_l[1]);
      double vz_l1 = v_l[3] - f1 * u_hll[5] * v_l[7] * (sm - v_l[1]);
      u_l1[0] = ro_l1;
      u_l1[4] = quo((((sl - v_l[1]) * u_l[4] - pt_l * v_l[1]) + pt * sm) +
                        u_hll[5] * (((vbL - sm * u_hll[5]) - vy_l1 * u_l1[6]) - vz_l1 * u_l1[7]),
                    sl - sm, rsl);
      u_l1[1] = ro_l1 * sm;
      u_l1[2] = ro_l1 * vy_l1;
      u_l1[3] = ro_l1 * vz_l1;

      f1 = qdiv(1.0, v_r[0] * (sr - v_r[1]) * (sr - sm) - bh2);
      u_r1[5] = u_hll[5];
      u_r1[6] = v_r[6] * f1 * (v_r[0] * ((sr - v_r[1]) * (sr - v_r[1])) - bh2);
      u_r1[7] = v_r[7] * f1 * (v_r[0] * ((sr - v_r[1]) * (sr - v_r[1])) - bh2);
      double vy_r1 = v_r[2] - f1 * u_hll[5] * v_r[6] * (sm - v_r[1]);
      double vz_r1 = v_r[3] - f1 * u_hll[5] * v_r[7] * (sm - v_r[1]);
      u_r1[0] = ro_r1;
      u_r1[4] = quo((((sr - v_r[1]) * u_r[4] - pt_r * v_r[1]) + pt * sm) +
                        u_hll[5] * (((vbR - sm * u_hll[5]) - vy_r1 * u_r1[6]) - vz_r1 * u_r1[7]),
                    sr - sm, rsr);
      u_r1[1] = ro_r1 * sm;
      u_r1[2] = ro_r1 * vy_r1;
      u_r1[3] = ro_r1 * vz_r1;

      double sl1 = sm - sal;
      double sr1 = sm + sar;
      if (sl1 >= 0.0) {
        flx[0] = sl * (u_l1[0] - v_l[0]) + flx_l[0];
#pragma unroll
        for (int n = 1; n < 5; n++) flx[n] = sl * (u_l1[n] - u_l[n]) + flx_l[n];
#pragma unroll
        for (int n = 5; n < 8; n++) flx[n] = sl * (u_l1[n] - v_l[n]) + flx_l[n];
      } else if (sr1 <= 0.0) {
        flx[0] = sr * (u_r1[0] - v_r[0]) + flx_r[0];
#pragma unroll
        for (int n = 1; n < 5; n++) flx[n] = sr * (u_r1[n] - u_r[n]) + flx_r[n];
#pragma unroll
        for (int n = 5; n < 8; n++) flx[n] = sr * (u_r1[n] - v_r[n]) + flx_r[n];
      } else {
        // central states :861-898
        double ro_ls = qsqrt(ro_l1);
        double ro_rs = qsqrt(ro_r1);
        f1 = qdiv(1.0, ro_ls + ro_rs);
        f2 = copysign(1.0, u_hll[5]);
        double vy_2 = f1 * ((ro_ls * vy_l1 + ro_rs * vy_r1) + (u_r1[6] - u_l1[6]) * f2);
        double vz_2 = f1 * ((ro_ls * vz_l1 + ro_rs * vz_r1) + (u_r1[7] - u_l1[7]) * f2);
        double u2[8];
        u2[5] = u_hll[5];
        u2[6] = f1 * ((ro_ls * u_r1[6] + ro_rs * u_l1[6]) + ro_ls * ro_rs * (vy_r1 - vy_l1) * f2);
        u2[7] = f1 * ((ro_ls * u_r1[7] + ro_rs * u_l1[7]) + ro_ls * ro_rs * (vz_r1 - vz_l1) * f2);
        const bool left = (sm >= 0.0);
        const double* v = left ? v_l : v_r;
        const double* uu = left ? u_l : u_r;
        const double* fl = left ? flx_l : flx_r;
        const double* us = left ? u_l1 : u_r1;
        double s = left ? sl : sr;
        double s1 = left ? sl1 : sr1;
        double ros = left ? ro_ls : ro_rs;
        double vy1 = left ? vy_l1 : vy_r1;
        double vz1 = left ? vz_l1 : vz_r1;
        double sg = left ? -f2 : f2;
        u2[0] = us[0];
        u2[4] = us[4] + ros * (((vy1 * us[6] + vz1 * us[7]) - vy_2 * u2[6]) - vz_2 * u2[7]) * sg;
        u2[1] = us[0] * sm;
        u2[2] = us[0] * vy_2;
        u2[3] = us[0] * vz_2;
        flx[0] = ((s1 * u2[0] - (s1 - s) * us[0]) - s * v[0]) + fl[0];
#pragma unroll
        for (int n = 1; n < 5; n++) flx[n] = ((s1 * u2[n] - (s1 - s) * us[n]) - s * uu[n]) + fl[n];
#pragma unroll
        for (int n = 5; n < 8; n++) flx[n] = ((s1 * u2[n] - (s1 - s) * us[n]) - s * v[n]) + fl[n];
      }
    }
  }
  if (NV == 9) {
    double ch = cmax;
    flx[5] = 0.5 * ((u_l[NV - 1] + u_r[NV - 1]) - ch * (v_r[5] - v_l[5]));
    flx[NV - 1] = 0.5 * (ch * ch * (v_l[5] + v_r[5]) - ch * (v_r[NV - 1] - v_l[NV - 1]));
  }
}

// ---------------------------------------------------------------- set_IR
// Run constants of set_IR case 3 (PIP_rot.f90:185-207); computed once.
struct IRConst {
  double Te_0, rec_fac, f_p, tfac, sqrt_tfac, k_ion, heat_fac;
};

PIP_HD IRConst ir_constants(const Phys& p) {
  IRConst c;
  c.Te_0 = p.T0 / sgl(1.1604e4);
  c.rec_fac = sgl(2.6e-19) * (p.n0 * sgl(1.0e6)) / sqrt(c.Te_0);
  double ioneq = (sgl(2.6e-19) / sqrt(c.Te_0)) /
                 (sgl(2.91e-14) / (0.232 + 13.6 / c.Te_0) * ppow(13.6 / c.Te_0, 0.39) * pexp(-13.6 / c.Te_0));
  double f_n = ioneq / (ioneq + 1.0);
  c.f_p = 1.0 - f_n;
  double f_p_p = 2.0 * c.f_p / (f_n + 2.0 * c.f_p);
  if (p.flag_col % 2 == 1)
    c.tfac = 0.5 * f_p_p / c.f_p;
  else
    c.tfac = p.beta / 2.0 * f_p_p * 5.0 / 6.0 / c.f_p;
  c.sqrt_tfac = sqrt(c.tfac);
  c.k_ion = sgl(2.91e-14) * (p.n0 * sgl(1.0e6));
  // ion_pot / arb_heat factor (PIP_rot.f90:222-224, 1387-1389)
  if (p.flag_col % 2 == 1)
    c.heat_fac = 13.6 / p.gm / p.T0 / sgl(8.6173e-5);
  else
    c.heat_fac = 13.6 * p.beta / p.T0 / 2.0 / sgl(8.6173e-5);
  return c;
}

// Gm_rec, Gm_ion of one cell from rho_p and T_p: PIP_rot.f90:215-217
PIP_HD void ir_rates(double rho_p, double te_p, const Phys& p, const IRConst& c, double& gr,
                     double& gi) {
  gr = fdiv(fdiv(rho_p, fsqrt(te_p)) * p.t_ir, c.f_p) * c.sqrt_tfac;
  double a = fdiv(13.6 / c.Te_0, te_p) * c.tfac;
  double g = c.k_ion * rho_p * pexp(-a) * ppow(a, 0.39);
  gi = fdiv(fdiv(fdiv(g, 0.232 + a), c.rec_fac), c.f_p) * p.t_ir;
}

// set_collisional, one cell, col_type 1..6 (PIP_rot.f90:44-97): get_Te_MHD
// and get_Te_HD (both rewrite E) and the drift of get_vel_diff
// (Scheme_rot.f90:988-999); returns alpha_c.
PIP_HD double set_col_cell(double* um, double* uh, double col, const Phys& p) {
  const int ct = col_type(p);
  double pr, te_m, te_h;
  getpr_mhd(um, pr, te_m, p);
  getpr_hd(uh, pr, te_h, p);
  const double tt = te_h + te_m;
  double a = col * sqrt(tt / 2.0);
  if (ct == 1) {
    a = a / sqrt(p.beta / 2.0 * p.gm);
  } else if (ct >= 3) {
    const double vd0 = uh[1] / uh[0] - um[1] / um[0];
    const double vd1 = uh[2] / uh[0] - um[2] / um[0];
    const double vd2 = uh[3] / uh[0] - um[3] / um[0];
    const double v2 = (vd0 * vd0 + vd1 * vd1) + vd2 * vd2;
    a = a * sqrt(1.0 + 9.0 * kPi / 64.0 * p.gm / 2.0 * v2 / tt);
    if (ct == 3) a = a / sqrt(p.beta / 2.0 * p.gm);
    else if (ct == 5)
      a = a / sqrt(p.beta / 2.0 * p.gm) * ppow(p.beta / 2.0 * p.gm / tt / 2.0 + p.gm * kPi / 16.0 * v2, 0.125);
    else if (ct == 6)
      a = a / ppow(tt / 2.0 + p.gm * kPi / 16.0 * v2, 0.125);
  }
  return a;
}

// set_IR on one cell: get_Te_HD / get_Te_MHD rewrites, then the rates.
template <int NV>
PIP_HD void set_ir_cell(double* um, double* uh, const Phys& p, const IRConst& c, double& gr,
                        double& gi) {
  double pr, te_n, te_p;
  getpr_hd(uh, pr, te_n, p);
  getpr_mhd(um, pr, te_p, p);
  ir_rates(um[0], te_p, p, c, gr, gi);
}

// ----------------------------------------------------------------- sources
// source_PIP, explicit branch, one cell: PIP_rot.f90:1309-1429.
// Rewrites um/uh through cq2pv_* (:1328-1329); writes S_h[5], S_m[5]
// (S_m(6:nvar) = 0) and ion_pot.
PIP_HD void source_pip(double* um, double* uh, double ac, double gr, double gi, double arb_heat,
                       const Phys& p, double heat_fac, double* Sh, double* Sm, double& ion_pot,
                       double dtc = 0.0, double* dU0 = nullptr) {
  double ph[5], pm[8];
  cq2pv_hd(uh, ph, p);
  cq2pv_mhd(um, pm, p);
  double nde = ph[0], nvx = ph[1], nvy = ph[2], nvz = ph[3], npr = ph[4];
  double de = pm[0], vx = pm[1], vy = pm[2], vz = pm[3], pr = pm[4];
  double te = qdiv(pr, de) * p.gm * kMuP;
  double nte = qdiv(npr, nde) * p.gm * kMuN;
  double dS[5];
  dS[0] = 0.0;
  if (p.pip_imp == 1) {
    // implicit exchange over dtc = dt_coll_i (:1333-1358): returned as the
    // increment dU0 = dtc dS of the plasma's stage base (the neutrals' is
    // -dU0); S gets no collision term
    const double temp = dtc * (uh[0] + um[0]);
    const double ex = pexp(-ac * temp);
    dS[1] = (um[0] * uh[1] - uh[0] * um[1]) * (1.0 - ex) / temp;
    dS[2] = (um[0] * uh[2] - uh[0] * um[2]) * (1.0 - ex) / temp;
    dS[3] = (um[0] * uh[3] - uh[0] * um[3]) * (1.0 - ex) / temp;
    const double lambda = ac * (nde + de);
    const double kapper = 3.0 * (p.gm - 1.0) * ac * (kMuP * nde + kMuN * de);
    const double Bq = -1.0 / 3.0 * p.gm * kapper / ((nde + de) * (kapper - lambda)) *
                      (de * ((nvx - vx) * vx + (nvy - vy) * vy + (nvz - vz) * nvz) +
                       nde * ((nvx - vx) * nvx + (nvy - vy) * nvy + (nvz - vz) * nvz));
    const double Dq = -p.gm / 6.0 * kapper * (de - nde) / ((de + nde) * (kapper - 2.0 * lambda)) *
                      (((nvx - vx) * (nvx - vx) + (nvy - vy) * (nvy - vy)) + (nvz - vz) * (nvz - vz));
    const double ek = pexp(-kapper * dtc);
    dS[4] = -(3.0 / p.gm) * ac * de * nde / kapper *
            ((((ek - 1.0) * (nte - te) - (Bq + Dq) * ek) + Bq * pexp(-lambda * dtc)) +
             Dq * pexp(-2.0 * lambda * dtc)) / dtc;
#pragma unroll
    for (int n = 0; n < 5; n++) {
      dU0[n] = dtc * dS[n];
      Sh[n] = 0.0;
      Sm[n] = 0.0;
    }
  } else {
    dS[1] = ac * (um[0] * uh[1] - uh[0] * um[1]);
    dS[2] = ac * (um[0] * uh[2] - uh[0] * um[2]);
    dS[3] = ac * (um[0] * uh[3] - uh[0] * um[3]);
    dS[4] = ac * nde * de *
            (0.5 * ((nvx * nvx - vx * vx) + (nvy * nvy - vy * vy) + (nvz * nvz - vz * vz)) +
             3.0 / p.gm / 2.0 * (nte - te));
#pragma unroll
    for (int n = 0; n < 5; n++) {
      Sh[n] = 0.0 - dS[n];
      Sm[n] = 0.0 + dS[n];
    }
  }
  ion_pot = 0.0;
  if (p.flag_ir >= 1) {
    double ds[5];
    ds[0] = gr * de - gi * nde;
    ds[1] = gr * de * vx - gi * nde * nvx;
    ds[2] = gr * de * vy - gi * nde * nvy;
    ds[3] = gr * de * vz - gi * nde * nvz;
    double kin = 0.5 * (gr * de * ((vx * vx + vy * vy) + vz * vz) -
                        gi * nde * ((nvx * nvx + nvy * nvy) + nvz * nvz));
    if (p.flag_ir_type == 2)
      ds[4] = kin - fdiv(gi * npr, p.gm - 1.0) + fdiv(gr * pr, p.gm - 1.0);
    else
      ds[4] = kin - fdiv(gi * npr, p.gm - 1.0) + fdiv(0.5 * gr * pr, p.gm - 1.0);
#pragma unroll
    for (int n = 0; n < 5; n++) {
      Sh[n] = Sh[n] + ds[n];
      Sm[n] = Sm[n] - ds[n];
    }
    if (p.flag_ir_type == 0) {
      ion_pot = gi * nde * heat_fac;
      Sm[4] = Sm[4] - ion_pot + arb_heat;
    }
  }
}

// ------------------------------------------------------------------- cfl
// Per-cell terms of cfl_mhd / cfl_hd after their cq2pv rewrites:
// Scheme_rot.f90:267-278 and :229-241.  Rewrites um/uh.
template <int NV>
PIP_HD double cfl_cell(double* um, double* uh, double gmin_safety, const Phys& p) {
  double pm[8], ph[5];
  cq2pv_mhd(um, pm, p);
  double de = pm[0], vx = pm[1], vy = pm[2], vz = pm[3], pr = pm[4];
  double bx = pm[5], by = pm[6], bz = pm[7];
  const double dd = dmax(de, kTiny), rdd = qrcp(dd);
  double cs2 = quo(p.gm * pr, dd, rdd);
  double valf2 = quo((bx * bx + by * by) + bz * bz, dd, rdd);
  double vabs = fsqrt((vx * vx + vy * vy) + vz * vz);
  double dtm = qdiv(gmin_safety, vabs + fsqrt(cs2 + valf2));
  cq2pv_hd(uh, ph, p);
  double cs = fsqrt(qdiv(p.gm * ph[4], ph[0]));
  double vabsn = fsqrt(dmax((ph[1] * ph[1] + ph[2] * ph[2]) + ph[3] * ph[3], kTiny));
  double dth = qdiv(gmin_safety, vabsn + cs);
  return dmin(dtm, dth);
}

}  // namespace PIP_NS
