// pip_adj_point.h -- per-point transposes used by the adjoint stage
// kernels and by the host checks: the minmod MUSCL face states and the
// conserved -> primitive conversion, at the branch of the forward point.
#pragma once
#include "pip_phys.h"

namespace pip {

// Gradient of the left face state V_L = v1 + 0.5 minmod(v0, v1, v2) of cell
// v1 (Util_rot.f90:59-76, HLL_rot.f90:33-51) w.r.t. (v0, v1, v2).
PIP_HD void recon_left_grad(double v0, double v1, double v2, double& g0, double& g1,
                                                double& g2) {
  const double X = v2 - v1, Y = v1 - v0;
  const double SX = copysign(1.0, X);
  const double m = dmin(fabs(X), SX * Y);
  g0 = 0.0; g1 = 1.0; g2 = 0.0;
  if (m > 0.0) {  // m <= 0: zero slope (also where X or Y vanishes)
    if (fabs(X) < SX * Y) { g1 -= 0.5; g2 += 0.5; }  // slope = X
    else { g1 += 0.5; g0 -= 0.5; }                    // slope = Y
  }
}
// ... and of the right face state V_R = v1 - 0.5 minmod(v0, v1, v2).
PIP_HD void recon_right_grad(double v0, double v1, double v2, double& g0, double& g1,
                                                 double& g2) {
  const double X = v2 - v1, Y = v1 - v0;
  const double SX = copysign(1.0, X);
  const double m = dmin(fabs(X), SX * Y);
  g0 = 0.0; g1 = 1.0; g2 = 0.0;
  if (m > 0.0) {
    if (fabs(X) < SX * Y) { g1 += 0.5; g2 -= 0.5; }
    else { g1 -= 0.5; g0 += 0.5; }
  }
}

// Branch of minmod(v0, v1, v2) taken at the forward point: 0 zero slope,
// 1 slope X = v2 - v1, 2 slope Y = v1 - v0 (the routing of both face states).
PIP_HD unsigned recon_code(double v0, double v1, double v2) {
  const double X = v2 - v1, Y = v1 - v0;
  const double SX = copysign(1.0, X);
  const double m = dmin(fabs(X), SX * Y);
  if (!(m > 0.0)) return 0u;
  return (fabs(X) < SX * Y) ? 1u : 2u;
}

// Weights of the three cells in V_L = v1 + mm/2 (left) and V_R = v1 - mm/2
// (right) for a branch code, as d/d(v0, v1, v2).
PIP_HD double recon_w(unsigned code, bool right, int q) {
  if (code == 0u) return (q == 1) ? 1.0 : 0.0;
  if (code == 1u) {  // slope v2 - v1
    if (q == 0) return 0.0;
    if (q == 1) return right ? 1.5 : 0.5;
    return right ? -0.5 : 0.5;
  }
  if (q == 2) return 0.0;  // slope v1 - v0
  if (q == 1) return right ? 0.5 : 1.5;
  return right ? 0.5 : -0.5;
}

// Transpose of u2v_mhd (HLL_rot.f90:1159-1186) at u: lu += (dV/dU)^T lv,
// with the density and pressure floors frozen as at u.
template <int NV>
PIP_HD void u2v_vjp(const double* u, const double* lv, const Phys& p, double* lu) {
  const double g1 = p.gm - 1.0;
  const bool rfl = !(u[0] > p.ro_lim);
  const double roi = 1.0 / dmax(u[0], p.ro_lim);
  const double m2 = (u[1] * u[1] + u[2] * u[2]) + u[3] * u[3];
  const double pr = g1 * ((u[4] - (0.5 * m2) * roi) - 0.5 * ((u[5] * u[5] + u[6] * u[6]) + u[7] * u[7]));
  double lroi = 0.0;
  for (int d = 1; d <= 3; d++) {  // v_d = m_d / rho
    lu[d] += lv[d] * roi;
    lroi += lv[d] * u[d];
  }
  if (!(pr < p.pr_lim)) {  // p = (gm-1) (E - m^2/(2 rho) - B^2/2)
    const double lp = lv[4] * g1;
    lu[4] += lp;
    for (int d = 1; d <= 3; d++) lu[d] -= lp * u[d] * roi;
    lroi -= lp * 0.5 * m2;
    for (int d = 5; d <= 7; d++) lu[d] -= lp * u[d];
  }
  if (!rfl) lu[0] += lv[0] - lroi * roi * roi;
  for (int d = 5; d < NV; d++) lu[d] += lv[d];
}

// Transpose of u2v_hd (HLL_rot.f90:1136-1156).
PIP_HD void u2v_vjp_hd(const double* u, const double* lv, const Phys& p, double* lu) {
  const double g1 = p.gm - 1.0;
  const bool rfl = !(u[0] > p.ro_lim);
  const double roi = 1.0 / dmax(u[0], p.ro_lim);
  const double m2 = (u[1] * u[1] + u[2] * u[2]) + u[3] * u[3];
  const double pr = g1 * (u[4] - (0.5 * m2) * roi);
  double lroi = 0.0;
  for (int d = 1; d <= 3; d++) {
    lu[d] += lv[d] * roi;
    lroi += lv[d] * u[d];
  }
  if (!(pr < p.pr_lim)) {
    const double lp = lv[4] * g1;
    lu[4] += lp;
    for (int d = 1; d <= 3; d++) lu[d] -= lp * u[d] * roi;
    lroi -= lp * 0.5 * m2;
  }
  if (!rfl) lu[0] += lv[0] - lroi * roi * roi;
}

}  // namespace pip
