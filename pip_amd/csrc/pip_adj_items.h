// pip_adj_items.h -- per-cell and per-face work items of the discrete
// adjoint stage (host and device): the HIP kernels of pip_adjoint.h run one
// item per thread, the host checks loop over them.
//
// The adjoint covers one SSP-RK2 step of the 2-D two-fluid update
// (nvar_m = 9 with Dedner psi, explicit collisions, no
// ionisation/recombination), for gradients with respect to the per-cell
// collision coefficient alpha_c and the initial state.
//
// Forward stage s (Solver_rot.f90:222-245, HLL_rot.f90:94-121):
//   Uout_c = base_c - a (Fx_{c+1/2} - Fx_{c-1/2}) - b (Fy_{c+1/2} - Fy_{c-1/2})
//            + dt_sub S(Uin_c, ac_c);   psi_out *= f  (Scheme_rot.f90:546-554)
//   base = U0 (s = 1) or c0 U0 + c1 Uin (s = 2), F = H(recon(u2v(Uin))).
// Reverse, with dt and cmax frozen at the values of the forward step and
// every branch (minmod slope, HLLD region, floors) taken as at the forward
// point; the cq2pv / get_Pr rewrites only re-round the state and pass
// adjoints through unchanged:
//   mu = lam_out on interior cells (psi component times f)
//   lam_U0 = c0 mu,  lam_in = c1 mu (s = 2); s = 1: lam_in = lam_U0 + mu
//   (U_n enters stage 1 as U0 and as Uin; lam_U0 of stage 2 is formed in
//   stage 1 from the step's output adjoint)
//   lam_in += dt_sub (dS/dUin)^T mu,  lam_ac += dt_sub (dS/dac)^T mu
//   per face: lamF = a (mu_R - mu_L) -> H^T: the adjoints of the face's
//             left and right states, written to planes of their own with
//             the face's minmod branch codes (no atomics, so the result is
//             deterministic)
//   per cell: recon^T -- gather those planes from the faces whose stencils
//             hold the cell and its ghost images (periodic / mirror sign
//             maps) -- then u2v^T, added to the pointwise part in lam_in.
// The host checks run these items one after another.  The HIP kernels run
// the plasma's face items over every face and the neutrals' only over the
// faces next to the boundary (adj_edge_face_item); the gather kernel
// computes the other neutral face adjoints of its tile itself (adj_face_vjp)
// and keeps them in shared memory.  Every value is formed by the same
// operations either way.
#pragma once

#include "pip_adj_gen.h"
#include "pip_adj_point.h"
#include "pip_bnd.h"
#include "pip_phys.h"

namespace pip {

struct AdjArgs {
  const double* Uin;      // stage input (nvt planes, ghosts filled)
  const double* lam_out;  // adjoint of the stage output
  double* lam_in;         // adjoint of Uin (ghosts zeroed by the cell item, interior assigned by the gather)
  const double* lam_next; // s = 1: adjoint of the step's output (stage 2's lam_out; may
                          // alias lam_in: read at a cell before the cell is written)
  double* lam_ac;         // += adjoint of alpha_c (nullable)
  const double* ac;       // per-cell alpha_c (nullable: constant col)
  double* V;              // nvt planes: primitives of Uin (all cells)
  int make_V;             // 1: the cell item computes V; 0: V was saved by the replay
  double* P;              // 2 dirs x 2 sides x nvt planes: face-state adjoints,
                          // then 2 dirs x 2 fluids planes of packed routing codes
  const double* dtc;      // device: {dt, cmax} of the forward step
  int ix, jx, nx, ny;
  long long N;
  int istep, t_order;
  double dx, dy, col, safety, db_clean;
  Phys ph;
  Bnd bnd;
};

constexpr int ADJ_NV = 9;
constexpr int ADJ_NVT = ADJ_NV + 5;

PIP_HD void adj_coeffs(const AdjArgs& a, double& c0, double& c1, double& dt_sub,
                                           double& psi_fac) {
  const double dt = a.dtc[0];
  stage_coeffs(a.istep, a.t_order, dt, c0, c1, dt_sub);
  psi_fac = pexp(-dt_sub / dt * a.safety * a.db_clean);
}

PIP_HD bool adj_interior(const AdjArgs& a, int i, int j) {
  return i >= 2 && i < a.ix - 2 && j >= 2 && j < a.jx - 2;
}

// adjoint of the stage output at an interior cell, psi scaled by the damping
PIP_HD double adj_mu_of(int n, double m, double psi_fac) { return (n == 8) ? m * psi_fac : m; }
PIP_HD double adj_mu(const AdjArgs& a, int n, long long o, double psi_fac) {
  return adj_mu_of(n, a.lam_out[n * a.N + o], psi_fac);
}

// Per cell (all cells): primitives of Uin for the face items (unless the
// replay saved them) and, in stage 1, zero ghost entries of lam_in (it
// receives the step's result).  The interior of lam_in is assigned by the
// gather item.
PIP_HD void adj_cell_item(const AdjArgs& a, long long k) {
  constexpr int NV = ADJ_NV, NVT = ADJ_NVT;
  if (k >= a.N) return;
  const long long N = a.N;
  if (a.make_V) {
    double um[NV], uh[5], vm[NV], vh[5];
#pragma unroll
    for (int n = 0; n < NV; n++) um[n] = a.Uin[n * N + k];
#pragma unroll
    for (int n = 0; n < 5; n++) uh[n] = a.Uin[(NV + n) * N + k];
    u2v_mhd<NV>(um, vm, a.ph);
    u2v_hd(uh, vh, a.ph);
#pragma unroll
    for (int n = 0; n < NV; n++) a.V[n * N + k] = vm[n];
#pragma unroll
    for (int n = 0; n < 5; n++) a.V[(NV + n) * N + k] = vh[n];
  }
  const int i = (int)(k % a.ix), j = (int)(k / a.ix);
  if (a.istep == 1 && !adj_interior(a, i, j)) {
#pragma unroll
    for (int n = 0; n < NVT; n++) a.lam_in[n * N + k] = 0.0;
  }
}

// Pointwise part of lam_in at an interior cell: the SSP blend plus the
// collision source VJP, S_m[0..4] = +dS, S_h[0..4] = -dS
// (PIP_rot.f90:1366-1367); lam_ac accumulates the source's alpha_c part.
// s = 2: c1 mu.  s = 1: (c0 mu') + mu, with mu' the adjoint of the step's
// output and c0 the stage-2 coefficient (the U0 path of stage 2).
PIP_HD void adj_point_part(const AdjArgs& a, long long k, const double* um, const double* uh, double* t) {
  constexpr int NV = ADJ_NV, NVT = ADJ_NVT;
  const long long N = a.N;
  double c0, c1, dt_sub, psi_fac;
  adj_coeffs(a, c0, c1, dt_sub, psi_fac);
  double mu[NVT];
#pragma unroll
  for (int n = 0; n < NVT; n++) mu[n] = adj_mu(a, n, k, psi_fac);
  if (a.istep == 1) {
    double d0, d1, dsub2;
    stage_coeffs(2, a.t_order, a.dtc[0], d0, d1, dsub2);
    const double pf2 = pexp(-dsub2 / a.dtc[0] * a.safety * a.db_clean);
#pragma unroll
    for (int n = 0; n < NVT; n++) {
      const double m2 = a.lam_next[n * N + k];
      t[n] = d0 * ((n == 8) ? m2 * pf2 : m2) + mu[n];
    }
  } else {
#pragma unroll
    for (int n = 0; n < NVT; n++) t[n] = c1 * mu[n];
  }
  double lam[5];
#pragma unroll
  for (int n = 0; n < 5; n++) lam[n] = dt_sub * (mu[n] - mu[NV + n]);
  double lum[8] = {0, 0, 0, 0, 0, 0, 0, 0}, luh[5] = {0, 0, 0, 0, 0}, lac = 0.0;
  const double acv = a.ac ? a.ac[k] : a.col;
  source_vjp(um, uh, acv, a.ph.gm, lam, lum, luh, &lac);
#pragma unroll
  for (int n = 0; n < 8; n++) t[n] += lum[n];
#pragma unroll
  for (int n = 0; n < 5; n++) t[NV + n] += luh[n];
  if (a.lam_ac) a.lam_ac[k] += lac;
}

// Swap of the normal/tangential components for y faces (HLL_rot.f90:414-434
// and :315-325); an involution.
PIP_HD int adj_perm(int n, int dir) {
  if (dir == 0) return n;
  if (n == 1) return 2;
  if (n == 2) return 1;
  if (n == 5) return 6;
  if (n == 6) return 5;
  if (n == ADJ_NV + 1) return ADJ_NV + 2;
  if (n == ADJ_NV + 2) return ADJ_NV + 1;
  return n;
}

// Packed routing codes of the faces of one direction and fluid: bits
// 4k..4k+1 the left-state branch, 4k+2..4k+3 the right-state branch of the
// fluid's k-th variable (recon_code), at the face's + cell.
// The codes of both fluids at a face sit side by side (one 16-byte
// access): code word (dir, cell, fluid) at ((dir * N + cell) * 2 + fluid).
struct alignas(16) CodePair {
  unsigned long long c[2];
};
PIP_HD CodePair* adj_code_pairs(const AdjArgs& a, int dir) {
  return reinterpret_cast<CodePair*>(a.P + 4LL * ADJ_NVT * a.N) + (long long)dir * a.N;
}

// The adjoints of a face's left and right states of one variable, side by
// side (one 16-byte access).
struct alignas(16) FacePair {
  double l, r;
};

// A face's VJP in three parts, so that a thread can issue the stencil loads
// of a second face before the first face's VJP (k_adj_plasma_faces):
//   adj_face_raw       the loads: the 4 primitive values of the stencil and
//                      the stage-output adjoints of the two cells (0 at
//                      cells that are not interior) per variable;
//   adj_face_from_raw  the reconstructed face states, lamF = coef (mu_R -
//                      mu_L) (psi scaled by the damping; a zero stays zero),
//                      the packed 2-bit minmod codes and the density floors
//                      of get_Side;
//   adj_face_compute   the flux VJP (HLLD with GLM for the plasma, HLLC for
//                      the neutrals) and the floors' zeroing.
template <int FL>
struct FaceRaw {
  static constexpr int NF = FL ? 5 : ADJ_NV;
  double v[NF][4], mR[NF], mL[NF];
};

template <int FL>
struct FaceIn {
  static constexpr int NF = FL ? 5 : ADJ_NV;
  double vl[NF], vr[NF], lamF[NF];
  unsigned long long codes;
  bool flL, flR;
};

template <int FL>
PIP_HD void adj_face_raw(const AdjArgs& a, int dir, int fi, int fj, FaceRaw<FL>& r) {
  constexpr int NF = FaceRaw<FL>::NF;
  constexpr int OFF = FL ? ADJ_NV : 0;  // first plane of this fluid
  const long long N = a.N;
  const long long st = (dir == 0) ? 1 : a.ix;
  const long long oR = (long long)fj * a.ix + fi;
  const long long oL = oR - st;
  const bool intL = (dir == 0) ? (fi - 1 >= 2) : (fj - 1 >= 2);
  const bool intR = (dir == 0) ? (fi < a.ix - 2) : (fj < a.jx - 2);
#pragma unroll
  for (int n = 0; n < NF; n++) {
    const int s = adj_perm(OFF + n, dir);
    const double* p = a.V + s * N + oR;
    r.v[n][0] = p[-2 * st];
    r.v[n][1] = p[-st];
    r.v[n][2] = p[0];
    r.v[n][3] = p[st];
    r.mR[n] = intR ? a.lam_out[s * N + oR] : 0.0;
    r.mL[n] = intL ? a.lam_out[s * N + oL] : 0.0;
  }
}

template <int FL>
PIP_HD void adj_face_from_raw(const AdjArgs& a, int dir, const FaceRaw<FL>& r, FaceIn<FL>& in) {
  constexpr int NF = FaceRaw<FL>::NF;
  constexpr int OFF = FL ? ADJ_NV : 0;
  double c0, c1, dt_sub, psi_fac;
  adj_coeffs(a, c0, c1, dt_sub, psi_fac);
  const double coef = dt_sub / ((dir == 0) ? a.dx : a.dy);
  unsigned long long codes = 0ull;
#pragma unroll
  for (int n = 0; n < NF; n++) {
    const int s = adj_perm(OFF + n, dir);
    const double v0 = r.v[n][0], v1 = r.v[n][1], v2 = r.v[n][2], v3 = r.v[n][3];
    in.vl[n] = v1 + 0.5 * minmod(v0, v1, v2);
    in.vr[n] = v2 - 0.5 * minmod(v1, v2, v3);
    in.lamF[n] = coef * (adj_mu_of(s, r.mR[n], psi_fac) - adj_mu_of(s, r.mL[n], psi_fac));
    const unsigned cl = recon_code(v0, v1, v2);
    const unsigned cr = recon_code(v1, v2, v3);
    codes |= (unsigned long long)(cl | (cr << 2)) << (4 * (s - OFF));
  }
  in.codes = codes;
  // density floor of get_Side (HLL_rot.f90:60-61)
  in.flL = !(in.vl[0] > a.ph.ro_lim);
  in.flR = !(in.vr[0] > a.ph.ro_lim);
  in.vl[0] = dmax(in.vl[0], a.ph.ro_lim);
  in.vr[0] = dmax(in.vr[0], a.ph.ro_lim);
}

template <int FL>
PIP_HD void adj_face_compute(const AdjArgs& a, FaceIn<FL>& in, double* lvl, double* lvr) {
  constexpr int NF = FaceIn<FL>::NF;
#pragma unroll
  for (int n = 0; n < NF; n++) { lvl[n] = 0.0; lvr[n] = 0.0; }
  if (FL == 0) hlld9_flux_vjp(in.vl, in.vr, a.ph.gm, a.dtc[1], in.lamF, lvl, lvr);
  else hllc_flux_vjp(in.vl, in.vr, a.ph.gm, in.lamF, lvl, lvr);
  if (in.flL) lvl[0] = 0.0;
  if (in.flR) lvr[0] = 0.0;
}

// The adjoints of the left and right states of the face with + cell
// (fi, fj) in direction dir for the fluid FL (0: plasma, HLLD with GLM; 1:
// neutrals, HLLC), density floor applied, into lvl / lvr (the fluid's
// variables in the x-sweep order); the face's packed routing code goes to
// code_out before the flux VJP (nothing of the stencil stays live across it).
template <int FL, class CodeOut>
PIP_HD void adj_face_vjp(const AdjArgs& a, int dir, int fi, int fj, double* lvl, double* lvr, CodeOut code_out) {
  FaceIn<FL> in;
  {
    FaceRaw<FL> r;
    adj_face_raw<FL>(a, dir, fi, fj, r);
    adj_face_from_raw<FL>(a, dir, r, in);
  }
  code_out(in.codes);
  adj_face_compute<FL>(a, in, lvl, lvr);
}

// The pair planes (dir * NVT + n) at the face's + cell offset oR: the
// adjoints of the left and right face states of variable n; adj_face_store
// also writes the face's routing code.
template <int FL>
PIP_HD void adj_face_pairs(const AdjArgs& a, int dir, long long oR, const double* lvl, const double* lvr) {
  constexpr int NF = FL ? 5 : ADJ_NV;
  constexpr int OFF = FL ? ADJ_NV : 0;
  FacePair* P = reinterpret_cast<FacePair*>(a.P) + (long long)dir * ADJ_NVT * a.N;
#pragma unroll
  for (int n = 0; n < NF; n++) {
    const int s = adj_perm(OFF + n, dir);
    P[s * a.N + oR] = FacePair{lvl[n], lvr[n]};
  }
}
template <int FL>
PIP_HD void adj_face_store(const AdjArgs& a, int dir, long long oR, unsigned long long codes, const double* lvl,
                           const double* lvr) {
  adj_code_pairs(a, dir)[oR].c[FL] = codes;
  adj_face_pairs<FL>(a, dir, oR, lvl, lvr);
}

// One thread per face and fluid; R = the cell on the + side of the face.
// dir 0: R in i = 2..nx+2, j = 2..ny+1; dir 1: R in i = 2..nx+1, j = 2..ny+2.
// Writes the face's pairs and routing code at R.
template <int FL>
PIP_HD void adj_face_item(const AdjArgs& a, int dir, long long t) {
  constexpr int NF = FL ? 5 : ADJ_NV;
  const int wx = (dir == 0) ? a.nx + 1 : a.nx;
  const int wy = (dir == 0) ? a.ny : a.ny + 1;
  if (t >= (long long)wx * wy) return;
  const int fi = (int)(t % wx) + 2, fj = (int)(t / wx) + 2;
  const long long oR = (long long)fj * a.ix + fi;
  double lvl[NF], lvr[NF];
  adj_face_vjp<FL>(a, dir, fi, fj, lvl, lvr,
                   [&](unsigned long long c) { adj_code_pairs(a, dir)[oR].c[FL] = c; });
  adj_face_pairs<FL>(a, dir, oR, lvl, lvr);
}

// The neutral faces the gather reads at the ghost images of boundary cells:
// x faces with + cell i in {2, 3, nx+1, nx+2} (rows 2..ny+1) and y faces
// with + cell j in {2, 3, ny+1, ny+2} (columns 2..nx+1); item t of
// adj_edge_faces(nx, ny).  Writes them like adj_face_item<1>.
PIP_HD long long adj_edge_faces(int nx, int ny) { return 4LL * ny + 4LL * nx; }
PIP_HD void adj_edge_face_item(const AdjArgs& a, long long t) {
  if (t >= adj_edge_faces(a.nx, a.ny)) return;
  const int dir = (t < 4LL * a.ny) ? 0 : 1;
  if (dir) t -= 4LL * a.ny;
  const int k = (int)(t % 4), r = (int)(t / 4);  // k-th edge face of row / column r
  const int w = dir ? a.ny : a.nx;
  const int c = (k < 2) ? 2 + k : w + k - 1;     // 2, 3, w+1, w+2
  if (k >= 2 && c <= 3) return;                  // w <= 2: already among 2, 3
  const int fi = dir ? r + 2 : c, fj = dir ? c : r + 2;
  const int wx = (dir == 0) ? a.nx + 1 : a.nx;
  adj_face_item<1>(a, dir, (long long)(fj - 2) * wx + (fi - 2));
}

// Left / right face-state weights of the cell at position q (0, 1, 2 =
// v0, v1, v2 of the stencil) for a recorded minmod branch code.
PIP_HD double wl(unsigned code, int q) {
  if (q == 0) return (code == 2u) ? -0.5 : 0.0;
  if (q == 1) return (code == 0u) ? 1.0 : ((code == 1u) ? 0.5 : 1.5);
  return (code == 1u) ? 0.5 : 0.0;
}
PIP_HD double wr(unsigned code, int q) {
  if (q == 0) return (code == 2u) ? 0.5 : 0.0;
  if (q == 1) return (code == 0u) ? 1.0 : ((code == 1u) ? 1.5 : 0.5);
  return (code == 1u) ? -0.5 : 0.0;
}

// Contributions of the faces of one direction and fluid to the cell:
// variables n0..n0+NF-1, faces with + cells c+2, c+1, c, c-1 (offsets
// o2, o1, o0, om, clamped to valid faces) scaled by v2..vm (1 valid, 0 not),
// with the fluid's branch codes k2..km at those faces.
template <int NF>
PIP_HD void adj_gather_fluid(const AdjArgs& a, int dir, int n0, long long o2, long long o1, long long o0,
                             long long om, double v2, double v1, double v0, double vm,
                             unsigned long long k2, unsigned long long k1, unsigned long long k0,
                             unsigned long long km, unsigned mask, double* lv) {
  constexpr int NVT = ADJ_NVT;
  const long long N = a.N;
  const FacePair* P = reinterpret_cast<const FacePair*>(a.P) + (long long)dir * NVT * N;
#pragma unroll
  for (int q = 0; q < NF; q++) {
    const int n = n0 + q, sh = 4 * q;
    const FacePair p2 = P[n * N + o2], p1 = P[n * N + o1], p0 = P[n * N + o0], pm = P[n * N + om];
    const double acc = v2 * (p2.l * wl((unsigned)(k2 >> sh) & 3u, 0)) +
                       v1 * (p1.l * wl((unsigned)(k1 >> sh) & 3u, 1) +
                             p1.r * wr((unsigned)(k1 >> (sh + 2)) & 3u, 0)) +
                       v0 * (p0.l * wl((unsigned)(k0 >> sh) & 3u, 2) +
                             p0.r * wr((unsigned)(k0 >> (sh + 2)) & 3u, 1)) +
                       vm * (pm.r * wr((unsigned)(km >> (sh + 2)) & 3u, 2));
    lv[n] += ((mask >> n) & 1u) ? -acc : acc;
  }
}

// The pointwise part and the transpose of u2v at an interior cell, after the
// face contributions lv: lam_in = pointwise + u2v^T lv.
PIP_HD void adj_cell_finish(const AdjArgs& a, long long k, double* lv) {
  constexpr int NV = ADJ_NV, NVT = ADJ_NVT;
  const long long N = a.N;
  double u[NVT], lu[NVT], t[NVT];
#pragma unroll
  for (int n = 0; n < NVT; n++) u[n] = a.Uin[n * N + k];
  adj_point_part(a, k, u, u + NV, t);
#pragma unroll
  for (int n = 0; n < NVT; n++) lu[n] = 0.0;
  u2v_vjp<NV>(u, lv, a.ph, lu);
  u2v_vjp_hd(u + NV, lv + NV, a.ph, lu + NV);
#pragma unroll
  for (int n = 0; n < NVT; n++) a.lam_in[n * N + k] = t[n] + lu[n];
}

// Sum of the face contributions to the primitives of the cell at (i, j),
// with the sign map `mask` of a ghost image applied (the plasma's only when
// NEUTRAL is false).  A face with + cell r
// (stride st) takes V_L from cells r-2, r-1, r and V_R from r-1, r, r+1,
// with the weights of the recorded minmod branches.
template <bool NEUTRAL = true>
PIP_HD void adj_gather_at(const AdjArgs& a, int i, int j, unsigned mask, double* lv) {
  constexpr int NV = ADJ_NV;
  const long long o = (long long)j * a.ix + i;
#pragma unroll
  for (int dir = 0; dir < 2; dir++) {
    const int c = dir ? j : i;  // coordinate along the sweep
    const int hi = dir ? a.ny + 2 : a.nx + 2;  // faces have + cells 2..hi
    const bool row_ok = dir ? (i >= 2 && i <= a.nx + 1) : (j >= 2 && j <= a.ny + 1);
    if (!row_ok) continue;
    const long long st = dir ? a.ix : 1;
    auto clampf = [&](int r) { return r < 2 ? 2 : (r > hi ? hi : r); };
    const long long o2 = o + (long long)(clampf(c + 2) - c) * st, o1 = o + (long long)(clampf(c + 1) - c) * st,
                    o0 = o + (long long)(clampf(c) - c) * st, om = o + (long long)(clampf(c - 1) - c) * st;
    const double v2 = (c + 2 >= 2 && c + 2 <= hi) ? 1.0 : 0.0, v1 = (c + 1 >= 2 && c + 1 <= hi) ? 1.0 : 0.0,
                 v0 = (c >= 2 && c <= hi) ? 1.0 : 0.0, vm = (c - 1 >= 2 && c - 1 <= hi) ? 1.0 : 0.0;
    const CodePair* C = adj_code_pairs(a, dir);
    const CodePair k2 = C[o2], k1 = C[o1], k0 = C[o0], km = C[om];
    adj_gather_fluid<NV>(a, dir, 0, o2, o1, o0, om, v2, v1, v0, vm, k2.c[0], k1.c[0], k0.c[0], km.c[0], mask, lv);
    if (NEUTRAL)
      adj_gather_fluid<5>(a, dir, NV, o2, o1, o0, om, v2, v1, v0, vm, k2.c[1], k1.c[1], k0.c[1], km.c[1], mask, lv);
  }
}

// Per interior cell: gather the face contributions at the cell and at its
// ghost images, then the transpose of u2v (HLL_rot.f90:1136-1186) with the
// forward point's floors; lam_in = pointwise part + that.
PIP_HD void adj_gather_item(const AdjArgs& a, long long k) {
  constexpr int NVT = ADJ_NVT;
  if (k >= a.N) return;
  const int ci = (int)(k % a.ix), cj = (int)(k / a.ix);
  if (!adj_interior(a, ci, cj)) return;
  double lv[NVT];
#pragma unroll
  for (int n = 0; n < NVT; n++) lv[n] = 0.0;
  adj_gather_at(a, ci, cj, 0u, lv);
  for_each_image(ci, cj, a.ix, a.jx, a.bnd,
                 [&](int xi, int yj, unsigned m) { adj_gather_at(a, xi, yj, m, lv); });
  adj_cell_finish(a, k, lv);
}

// Adjoint seed on the ghost cells of a state: add each ghost's value (with
// its boundary sign map) onto the interior cell it images, so that the
// adjoint kernels only read interior cells.
PIP_HD void adj_fold_seed_item(double* lam, int nvt, int ix, int jx, long long N, const Bnd& bnd, long long k) {
  if (k >= N) return;
  const int ci = (int)(k % ix), cj = (int)(k / ix);
  if (!(ci >= 2 && ci < ix - 2 && cj >= 2 && cj < jx - 2)) return;
  for_each_image(ci, cj, ix, jx, bnd, [&](int xi, int yj, unsigned m) {
    const long long g = (long long)yj * ix + xi;
    for (int n = 0; n < nvt; n++) {
      const double v = lam[n * N + g];
      lam[n * N + k] += ((m >> n) & 1u) ? -v : v;
    }
  });
}

// Zero the ghost cells of an nvt-plane state (the adjoint carries values on
// interior cells only).
PIP_HD void zero_ghosts_item(double* x, int nvt, int ix, int jx, long long N, long long k) {
  if (k >= N) return;
  const int i = (int)(k % ix), j = (int)(k / ix);
  if (i >= 2 && i < ix - 2 && j >= 2 && j < jx - 2) return;
  for (int n = 0; n < nvt; n++) x[n * N + k] = 0.0;
}

}  // namespace pip
