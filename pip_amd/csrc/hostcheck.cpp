// hostcheck.cpp -- host build of the per-face/per-cell device functions
// (pip_phys.h, pip_adj_gen.h) for CPU tests: forward fluxes and their
// vector-Jacobian products.  Test infrastructure; not used by libpipgpu.
#include "pip_adj_gen.h"
#include "pip_phys.h"

using namespace pip;

extern "C" {
void hc_hllc(const double* vl, const double* vr, double gm, double* f) { hllc_flux(vl, vr, gm, f); }
void hc_hlld9(const double* vl, const double* vr, double gm, double ch, double* f) {
  hlld_flux<9>(vl, vr, gm, ch, f);
}
void hc_hlld8(const double* vl, const double* vr, double gm, double* f) { hlld_flux<8>(vl, vr, gm, 0.0, f); }
void hc_hllc_vjp(const double* vl, const double* vr, double gm, const double* lam, double* lvl, double* lvr) {
  hllc_flux_vjp(vl, vr, gm, lam, lvl, lvr);
}
void hc_hlld9_vjp(const double* vl, const double* vr, double gm, double ch, const double* lam, double* lvl,
                  double* lvr) {
  hlld9_flux_vjp(vl, vr, gm, ch, lam, lvl, lvr);
}
void hc_source(const double* um, const double* uh, double ac, double gm, double* sm) {
  // collision part of source_pip on copies (the cq2pv rewrites are applied)
  double m[8], h[5], Sh[5], Sm[5], ip;
  for (int k = 0; k < 8; k++) m[k] = um[k];
  for (int k = 0; k < 5; k++) h[k] = uh[k];
  Phys p{};
  p.gm = gm;
  p.ro_lim = 1e-8;
  p.pr_lim = 1e-8;
  p.flag_ir = 0;
  source_pip(m, h, ac, 0.0, 0.0, 0.0, p, 0.0, Sh, Sm, ip);
  for (int k = 0; k < 5; k++) sm[k] = Sm[k];
}
void hc_source_vjp(const double* um, const double* uh, double ac, double gm, const double* lam, double* lum,
                   double* luh, double* lac) {
  source_vjp(um, uh, ac, gm, lam, lum, luh, lac);
}
double hc_exp(double x) { return pexp(x); }
double hc_pow(double x, double y) { return ppow(x, y); }
}
