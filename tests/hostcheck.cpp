// hostcheck.cpp -- host build of the per-face/per-cell device functions
// (pip_phys.h, pip_adj_gen.h, pip_adj_point.h) for CPU tests: forward
// fluxes, reconstruction and conversions, and their vector-Jacobian
// products.  Test infrastructure (built by tests/test_adjoint_cpu.py with
// -I pip_amd/csrc); not used by libpipgpu.
#include <vector>

#include "pip_adj_gen.h"
#include "pip_adj_items.h"
#include "pip_adj_point.h"
#include "pip_bnd.h"
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
double hc_minmod(double a, double b, double c) { return minmod(a, b, c); }
void hc_recon_grad(double v0, double v1, double v2, int right, double* g) {
  if (right) recon_right_grad(v0, v1, v2, g[0], g[1], g[2]);
  else recon_left_grad(v0, v1, v2, g[0], g[1], g[2]);
}
static Phys hc_phys(double gm, double ro_lim, double pr_lim) {
  Phys p{};
  p.gm = gm;
  p.ro_lim = ro_lim;
  p.pr_lim = pr_lim;
  return p;
}
void hc_u2v9(const double* u, double gm, double ro_lim, double pr_lim, double* v) {
  u2v_mhd<9>(u, v, hc_phys(gm, ro_lim, pr_lim));
}
void hc_u2v9_vjp(const double* u, const double* lv, double gm, double ro_lim, double pr_lim, double* lu) {
  u2v_vjp<9>(u, lv, hc_phys(gm, ro_lim, pr_lim), lu);
}
void hc_u2v_hd(const double* u, double gm, double ro_lim, double pr_lim, double* v) {
  u2v_hd(u, v, hc_phys(gm, ro_lim, pr_lim));
}
void hc_u2v_hd_vjp(const double* u, const double* lv, double gm, double ro_lim, double pr_lim, double* lu) {
  u2v_vjp_hd(u, lv, hc_phys(gm, ro_lim, pr_lim), lu);
}
int hc_src_index(int g, int npad, int tlo, int thi, int* flip) {
  bool f;
  int s = src_index(g, npad, tlo, thi, f);
  *flip = f ? 1 : 0;
  return s;
}
// ghost images of interior cell (ci, cj): out[3k..3k+2] = (xi, yj, mask)
int hc_images(int ci, int cj, int ix, int jx, const int* type, const unsigned* sym, int* out) {
  Bnd b;
  for (int s = 0; s < 4; s++) {
    b.type[s] = type[s];
    b.sym[s] = sym[s];
  }
  int n = 0;
  for_each_image(ci, cj, ix, jx, b, [&](int xi, int yj, unsigned m) {
    out[3 * n] = xi;
    out[3 * n + 1] = yj;
    out[3 * n + 2] = (int)m;
    n++;
  });
  return n;
}
// One adjoint step on the host (the same work items as the HIP kernels, run
// in sequence): p = {gm, ro_lim, pr_lim, dx, dy, col, safety, db_clean, dt,
// cmax}; flag_bnd[4]; on entry lam = dJ/dU_{n+1}, on exit dJ/dU_n.
void hc_adj_step(int ix, int jx, const double* p, const int* flag_bnd, const double* Un, const double* U1,
                 const double* ac, double* lam, double* lam_ac) {
  const int NVT = ADJ_NVT;
  const long long N = (long long)ix * jx;
  Phys ph{};
  ph.gm = p[0];
  ph.ro_lim = p[1];
  ph.pr_lim = p[2];
  ph.safety = p[6];
  std::vector<double> V(NVT * N), P(4 * NVT * N + 4 * N, 0.0), lamB(NVT * N, 0.0);
  const double dtc[2] = {p[8], p[9]};
  const Bnd bnd = make_bnd(flag_bnd, 2, ADJ_NV, 1);
  for (long long k = 0; k < N; k++) adj_fold_seed_item(lam, NVT, ix, jx, N, bnd, k);
  for (long long k = 0; k < N; k++) zero_ghosts_item(lam, NVT, ix, jx, N, k);
  for (int stage = 2; stage >= 1; stage--) {
    AdjArgs a;
    a.Uin = (stage == 2) ? U1 : Un;
    a.lam_out = (stage == 2) ? lam : lamB.data();
    a.lam_in = (stage == 2) ? lamB.data() : lam;
    a.lam_next = (stage == 2) ? nullptr : lam;
    a.lam_ac = lam_ac;
    a.ac = ac;
    a.V = V.data();
    a.make_V = 1;
    a.P = P.data();
    a.dtc = dtc;
    a.ix = ix;
    a.jx = jx;
    a.nx = ix - 4;
    a.ny = jx - 4;
    a.N = N;
    a.istep = stage;
    a.t_order = 2;
    a.dx = p[3];
    a.dy = p[4];
    a.col = p[5];
    a.safety = p[6];
    a.db_clean = p[7];
    a.ph = ph;
    a.bnd = bnd;
    for (long long k = 0; k < N; k++) adj_cell_item(a, k);
    for (int dir = 0; dir < 2; dir++) {
      const long long nf = (dir == 0) ? (long long)(a.nx + 1) * a.ny : (long long)a.nx * (a.ny + 1);
      for (long long t = 0; t < nf; t++) {
        adj_face_item<0>(a, dir, t);
        adj_face_item<1>(a, dir, t);
      }
    }
    for (long long k = 0; k < N; k++) adj_gather_item(a, k);
  }
}
double hc_exp(double x) { return pexp(x); }
double hc_pow(double x, double y) { return ppow(x, y); }
}
