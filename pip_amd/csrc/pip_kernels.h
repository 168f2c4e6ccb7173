// pip_kernels.h -- HIP kernels of the fused two-fluid step for gfx950.
//
// Device layout (identical to the reference arrays, so uploads are memcpy):
// one buffer of NVT = nvar_m + 5 planes, plane n holding ix*jx doubles
// (x fastest); planes 0..nvar_m-1 are U_m, planes nvar_m.. are U_h.
//
// One launch of k_stage performs one SSP-RK stage for both fluids on a
// TX x TY tile of interior cells per workgroup:
//   load  : (TX+4)x(TY+4) cells -> primitives in LDS (u2v_*, HLL_rot.f90:1136-1186)
//   x     : (TX+1)*TY faces: minmod MUSCL + HLLD, then minmod MUSCL + HLLC
//           (one fluid at a time, so the neutral face states are not live
//           while HLLD runs) -> flux tile in LDS
//   cell  : source_PIP + SSP blend + x flux divergence (Solver_rot.f90:464-478)
//   y     : TX*(TY+1) faces (permuted HLLD/HLLC) -> flux tile in LDS
//   cell  : y flux divergence, add_source, Dedner psi damping, then (last
//           stage of a fused step) the set_IR/cfl rewrite chain of main.f90
//           with the dt reduction, and the store of the cell and its ghost
//           images (Boundary_rot.f90:323-402 applied by index mapping).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pip_bnd.h"
#include "pip_phys.h"

namespace PIP_NS {

// Reduction slots (bit patterns of non-negative doubles; min/max as uint64).
enum { RED_DT = 0, RED_RHOAC = 1, RED_GMAX = 2, RED_IONPOT = 3, RED_NAN = 4, RED_N = 5 };

struct DevScalars {
  double dt;        // dt of the current step (global dt)
  double time;
  double cmax;      // GLM ch (Scheme_rot.f90:205-211)
  double tend;
  double dt_out;    // last cfl result
  unsigned long long red[RED_N];
  long long nt;
  long long nmax;
  int active;       // 1 while time < tend and no stop condition
  int last;         // current step is the last one (skip the look-ahead rewrite)
  int status;       // stop_sim flags
  int ion_pot_valid;
  int pad[4];
};

// Stage kernel modes: a step of the time loop; a step of a frozen-dt replay
// ({dt, cmax} from StageArgs::dtc, always active, never the loop's last
// step, no dt reduction) that also writes the primitives of its input to
// Vout; a step of the time loop that writes them as well (the last segment of
// an adjoint's recording run).
enum { STAGE_PLAIN = 0, STAGE_REPLAY = 1, STAGE_SAVE = 2 };

// Boundary description: per side type (1 periodic, 2 mirror, 10 free) and
// sign masks (bit n set -> variable n flips sign), Boundary_rot.f90:239-320.
struct StageArgs {
  const double* Uin;  // stage input (stencil)
  const double* U0;   // state at run_solver entry (pointwise; may alias Uout)
  double* Uout;
  const double* ac;       // per-cell alpha_c or nullptr
  double* ac_out;         // set_collisional col_type >= 1: alpha_c of this stage (nullable)
  double* U0out;          // implicit collisions: updated stage base for the next stage (nullable)
  const double* arb_heat; // IR type 0
  double* ion_pot_out;    // nullable: ion_pot of this stage
  double* Vout;           // replay and saving kernels, nullable: primitives of Uin at the
                          // tile's cells and their ghost images
  const double* dtc;      // replay kernels: frozen {dt, cmax} of the replayed step
  DevScalars* sc;
  int ix, jx, nx, ny;
  long long N;            // plane stride ix*jx
  int istep, t_order, epi;  // epi: 0 none, 1 one cfl round, 2 two rounds + dt reduction
  int xcd_remap;          // 1: contiguous tile ranges per XCD
  double col;
  double gmin_safety;     // safety*min(dx,dy,dz)
  double dxyz_min;        // min(dx,dy,dz)
  double dx, dy;
  double db_clean;
  Phys ph;
  IRConst irc;
  Bnd bnd;
};

PIP_HD unsigned long long as_u64(double x) {
  union { double d; unsigned long long u; } c;
  c.d = x;
  return c.u;
}
PIP_HD double as_f64(unsigned long long x) {
  union { double d; unsigned long long u; } c;
  c.u = x;
  return c.d;
}

// Store one cell value set (NVT vars) at (ci,cj) and at all its ghost images.
template <int NDIM, int NVT>
__device__ __forceinline__ void store_cell(double* U, long long N, int ix, int jx, int ci, int cj,
                                           const double* u, const Bnd& b) {
#pragma unroll
  for (int n = 0; n < NVT; n++) U[n * N + (long long)cj * ix + ci] = u[n];
  for_each_image(ci, cj, ix, (NDIM == 2) ? jx : 1, b, [&](int xi, int yj, unsigned m) {
    for (int n = 0; n < NVT; n++) U[n * N + (long long)yj * ix + xi] = ((m >> n) & 1u) ? -u[n] : u[n];
  });
}

// One round of the set_IR/cfl chain on one cell (main.f90:32-36 / :44-47):
// set_IR rewrites (+ rates), cfl rewrites and the cell's dt terms.
struct RoundOut {
  double dtc, rhoac, gmax, ratio;
};

template <int NV, bool IR>
__device__ __forceinline__ RoundOut cfl_round(double* um, double* uh, double ac, double ion_pot,
                                              const StageArgs& a) {
  RoundOut r;
  double gr = 0.0, gi = 0.0;
  if (IR) set_ir_cell<NV>(um, uh, a.ph, a.irc, gr, gi);
  r.dtc = cfl_cell<NV>(um, uh, a.gmin_safety, a.ph);
  r.rhoac = dmax(um[0] * ac, uh[0] * ac);
  r.gmax = dmax(gr, gi);
  r.ratio = 0.0;
  if (IR && a.ph.flag_ir_type == 0) {
    double pr, te;
    getpr_mhd(um, pr, te, a.ph);  // cfl_pip_ir :321
    r.ratio = fdiv(ion_pot, fdiv(pr, a.ph.gm - 1.0));
  }
  return r;
}

// The state rewrites of one round alone, for a round whose dt terms are not
// used (the first of the two rounds of a fused step): set_IR's get_Pr
// rewrites, cfl's cq2pv_mhd / cq2pv_hd and, for IR type 0, cfl_pip_ir's
// get_Pr_MHD, in the order cfl_round applies them.
template <int NV, bool IR>
__device__ __forceinline__ void cfl_rewrite(double* um, double* uh, const StageArgs& a) {
  if (IR) {
    double gr, gi;
    set_ir_cell<NV>(um, uh, a.ph, a.irc, gr, gi);
  }
  double pm[8], ph[5];
  cq2pv_mhd(um, pm, a.ph);
  cq2pv_hd(uh, ph, a.ph);
  if (IR && a.ph.flag_ir_type == 0) {
    double pr, te;
    getpr_mhd(um, pr, te, a.ph);
  }
}

// Workgroup min/max reduction of the dt terms, then one atomic per slot.
template <int NT>
__device__ __forceinline__ void block_reduce_store(RoundOut r, int nanflag, bool valid, DevScalars* sc,
                                                   double* red_lds) {
  double v0 = valid ? r.dtc : 1.0e300;
  double v1 = valid ? r.rhoac : 0.0;
  double v2 = valid ? r.gmax : 0.0;
  double v3 = valid ? r.ratio : 0.0;
  int fl = valid ? nanflag : 0;
  for (int off = 32; off > 0; off >>= 1) {
    v0 = dmin(v0, __shfl_xor(v0, off, 64));
    v1 = dmax(v1, __shfl_xor(v1, off, 64));
    v2 = dmax(v2, __shfl_xor(v2, off, 64));
    v3 = dmax(v3, __shfl_xor(v3, off, 64));
    fl |= __shfl_xor(fl, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int NW = (NT + 63) / 64;
  __syncthreads();
  if (lane == 0) {
    red_lds[wid * 5 + 0] = v0;
    red_lds[wid * 5 + 1] = v1;
    red_lds[wid * 5 + 2] = v2;
    red_lds[wid * 5 + 3] = v3;
    red_lds[wid * 5 + 4] = (double)fl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NW; w++) {
      v0 = dmin(v0, red_lds[w * 5 + 0]);
      v1 = dmax(v1, red_lds[w * 5 + 1]);
      v2 = dmax(v2, red_lds[w * 5 + 2]);
      v3 = dmax(v3, red_lds[w * 5 + 3]);
      fl |= (int)red_lds[w * 5 + 4];
    }
    atomicMin(&sc->red[RED_DT], as_u64(v0));
    atomicMax(&sc->red[RED_RHOAC], as_u64(v1));
    atomicMax(&sc->red[RED_GMAX], as_u64(v2));
    atomicMax(&sc->red[RED_IONPOT], as_u64(v3));
    if (fl) atomicOr(&sc->red[RED_NAN], 1ull);
  }
}

// ------------------------------------------------------------ stage tile
// LDS footprint of one tile (doubles): primitives (TX+4)x(TY+4) x NVT, then
// the face-flux tile NVT x max((TX+1)TY, TX(TY+1)).
template <int NDIM, int NV, int TX, int TY>
struct TileGeom {
  static constexpr int NVT = NV + 5;
  static constexpr int NCELL = TX * TY;
  static constexpr int HX = TX + 4;
  static constexpr int HY = (NDIM == 2) ? TY + 4 : 1;
  static constexpr int NFX = (TX + 1) * TY;
  static constexpr int NFY = (NDIM == 2) ? TX * (TY + 1) : 0;
  static constexpr int NF = (NFX > NFY) ? NFX : NFY;
  static constexpr int LDS_DOUBLES = NVT * HY * HX + NVT * NF;
};

// One SSP-RK stage on the tile of interior cells [i0,i0+TX) x [j0,j0+TY)
// (interior coordinates), executed by all NT threads of the workgroup.
// NT >= TX*TY; shapes with (TX+1)*TY <= NT and TX*(TY+1) <= NT compute every
// face phase in a single pass of the workgroup (e.g. 16x15 on 256 threads).
// Returns whether the calling thread owns a cell; r / nanflag receive the
// cell's dt terms when epi >= 1.
template <int NDIM, int NV, bool IR, bool EXT, int TX, int TY, int NT, int MODE = STAGE_PLAIN>
__device__ __forceinline__ bool stage_tile(const StageArgs& a, double dt, double cmax, int epi, int i0,
                                           int j0, double* smem, RoundOut& r, int& nanflag) {
  using G = TileGeom<NDIM, NV, TX, TY>;
  constexpr int NVT = G::NVT;
  constexpr int HX = G::HX;
  constexpr int HY = G::HY;
  constexpr int OY = (NDIM == 2) ? 2 : 0;
  constexpr int NFX = G::NFX;
  constexpr int NFY = G::NFY;
  constexpr int NF = G::NF;
  double* V = smem;                      // NVT * HY * HX
  double* F = smem + NVT * HY * HX;      // NVT * NF
  double c0, c1, dt_sub;
  stage_coeffs(a.istep, a.t_order, dt, c0, c1, dt_sub);
  const int tid = threadIdx.x;
  const long long N = a.N;
  const int ix = a.ix;

  // ---- load halo tile -> primitives in LDS
  for (int c = tid; c < HX * HY; c += NT) {
    int hx = c % HX, hy = c / HX;
    int gi = i0 + hx, gj = (NDIM == 2) ? j0 + hy : 0;
    if (gi < a.ix && gj < a.jx) {
      long long o = (long long)gj * ix + gi;
      double um[NV], uh[5];
#pragma unroll
      for (int n = 0; n < NV; n++) um[n] = a.Uin[n * N + o];
#pragma unroll
      for (int n = 0; n < 5; n++) uh[n] = a.Uin[(NV + n) * N + o];
      if (EXT && col_type(a.ph) >= 1) {  // set_collisional's get_Te_* rewrite E (PIP_rot.f90:57-58)
        double pr, te;
        getpr_mhd(um, pr, te, a.ph);
        getpr_hd(uh, pr, te, a.ph);
      }
      if (IR) {  // set_IR of set_coefficients(.,1) rewrites E (Solver_rot.f90:111-113)
        double pr, te;
        getpr_hd(uh, pr, te, a.ph);
        getpr_mhd(um, pr, te, a.ph);
      }
      double vm[NV], vh[5];
      u2v_mhd<NV>(um, vm, a.ph);
      u2v_hd(uh, vh, a.ph);
#pragma unroll
      for (int n = 0; n < NV; n++) V[(n * HY + hy) * HX + hx] = vm[n];
#pragma unroll
      for (int n = 0; n < 5; n++) V[((NV + n) * HY + hy) * HX + hx] = vh[n];
      if (MODE != STAGE_PLAIN && a.Vout && hx >= 2 && hx < TX + 2 && gi < a.nx + 2 &&
          (NDIM == 1 || (hy >= 2 && hy < TY + 2 && gj < a.ny + 2))) {
        // a cell of this tile: its primitives and their ghost images (u2v
        // commutes with the boundary sign maps: v = m / rho, the pressure
        // takes squares)
        double vv[NVT];
#pragma unroll
        for (int n = 0; n < NV; n++) vv[n] = vm[n];
#pragma unroll
        for (int n = 0; n < 5; n++) vv[NV + n] = vh[n];
        store_cell<NDIM, NVT>(a.Vout, N, ix, a.jx, gi, gj, vv, a.bnd);
      }
    }
  }
  __syncthreads();

  // ---- x faces: face fi (0..TX) between tile cells fi-1 and fi
  for (int f = tid; f < NFX; f += NT) {
    int fi = f % (TX + 1), fj = f / (TX + 1);
    if (i0 + fi > a.nx || j0 + fj >= a.ny) continue;
    int hy = fj + OY;
    const double* row = V + hy * HX;
    // one fluid at a time: the neutral face states are not live while the
    // plasma solver runs (same operations, fewer registers at its peak)
    {
      double vl[NV], vr[NV], fl[NV];
#pragma unroll
      for (int n = 0; n < NV; n++) {
        const double* r = row + n * HY * HX + fi;
        double a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3];
        vl[n] = a1 + 0.5 * minmod(a0, a1, a2);
        vr[n] = a2 - 0.5 * minmod(a1, a2, a3);
      }
      vl[0] = dmax(vl[0], a.ph.ro_lim);
      vr[0] = dmax(vr[0], a.ph.ro_lim);
      hlld_flux<NV>(vl, vr, a.ph.gm, cmax, fl);
#pragma unroll
      for (int n = 0; n < NV; n++) F[n * NF + f] = fl[n];
    }
    {
      double vl[5], vr[5], fl[5];
#pragma unroll
      for (int n = 0; n < 5; n++) {
        const double* r = row + (NV + n) * HY * HX + fi;
        double a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3];
        vl[n] = a1 + 0.5 * minmod(a0, a1, a2);
        vr[n] = a2 - 0.5 * minmod(a1, a2, a3);
      }
      vl[0] = dmax(vl[0], a.ph.ro_lim);
      vr[0] = dmax(vr[0], a.ph.ro_lim);
      hllc_flux(vl, vr, a.ph.gm, fl);
#pragma unroll
      for (int n = 0; n < 5; n++) F[(NV + n) * NF + f] = fl[n];
    }
  }
  __syncthreads();

  // ---- own cell: sources and x update
  const int ti = tid % TX, tj = tid / TX;
  const int ci = i0 + ti + 2, cj = (NDIM == 2) ? j0 + tj + 2 : 0;
  const bool valid = (tid < TX * TY) && (i0 + ti < a.nx) && (j0 + tj < a.ny);
  double un[NVT];
  double Sm[5], Sh[5];
  double ion_pot = 0.0;
  double acc = a.col;
  const long long oc = (long long)cj * ix + ci;
  if (valid) {
    double um[NV], uh[5];
#pragma unroll
    for (int n = 0; n < NV; n++) um[n] = a.Uin[n * N + oc];
#pragma unroll
    for (int n = 0; n < 5; n++) uh[n] = a.Uin[(NV + n) * N + oc];
    if (a.ac) acc = a.ac[oc];
    if (EXT && col_type(a.ph) >= 1) {  // set_collisional (PIP_rot.f90:44-97), before set_IR
      acc = set_col_cell(um, uh, a.col, a.ph);
      if (a.ac_out) {
        a.ac_out[oc] = acc;
        for_each_image(ci, cj, ix, (NDIM == 2) ? a.jx : 1, a.bnd,
                       [&](int xi, int yj, unsigned) { a.ac_out[(long long)yj * ix + xi] = acc; });
      }
    }
    double gr = 0.0, gi = 0.0, arb = 0.0;
    if (IR) {
      set_ir_cell<NV>(um, uh, a.ph, a.irc, gr, gi);
      if (a.ph.flag_ir_type == 0) arb = a.arb_heat[oc];
    }
    // set_dt_coll (Solver_rot.f90:83-93): dt/2 per stage for t_order = 2
    const double dtc = (a.t_order == 2) ? dt * 0.5 : dt;
    double dU0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    Phys phs = a.ph;
    if (!EXT) phs.pip_imp = 0;  // the default kernels carry no implicit branch
    source_pip(um, uh, acc, gr, gi, arb, phs, a.irc.heat_fac, Sh, Sm, ion_pot, dtc, dU0);
    if (a.ion_pot_out) a.ion_pot_out[oc] = ion_pot;
    const double ax = dt_sub / a.dx;
    const int fL = tj * (TX + 1) + ti;
#pragma unroll
    for (int n = 0; n < NVT; n++) {
      double u0 = a.U0[n * N + oc];
      if (EXT && a.ph.pip_imp == 1) {  // U_m0 += dt_c dS, U_h0 -= dt_c dS (PIP_rot.f90:1355-1356)
        if (n < 5) u0 = u0 + dU0[n];
        else if (n >= NV && n < NV + 5) u0 = u0 - dU0[n - NV];
        if (a.U0out) a.U0out[n * N + oc] = u0;
      }
      if (a.istep == 1) {
        un[n] = u0 - ax * (F[n * NF + fL + 1] - F[n * NF + fL]);
      } else {
        double ur = (n < NV) ? um[n] : uh[n - NV];
        un[n] = c0 * u0 + c1 * ur - ax * (F[n * NF + fL + 1] - F[n * NF + fL]);
      }
    }
  }

  // ---- y faces
  if (NDIM == 2) {
    __syncthreads();
    for (int f = tid; f < NFY; f += NT) {
      int fi = f % TX, fj = f / TX;
      if (i0 + fi >= a.nx || j0 + fj > a.ny) continue;
      int hx = fi + 2;
      // permutation (1,3,2,4,5,7,6,8,9)/(1,3,2,4,5): HLL_rot.f90:414-434,
      // 315-325; one fluid at a time as in the x faces
      {
        double vl[NV], vr[NV], fl[NV];
#pragma unroll
        for (int n = 0; n < NV; n++) {
          int src = n;
          if (n == 1) src = 2; else if (n == 2) src = 1;
          else if (n == 5) src = 6; else if (n == 6) src = 5;
          const double* col = V + src * HY * HX + hx;
          double a0 = col[(fj + 0) * HX], a1 = col[(fj + 1) * HX], a2 = col[(fj + 2) * HX],
                 a3 = col[(fj + 3) * HX];
          vl[n] = a1 + 0.5 * minmod(a0, a1, a2);
          vr[n] = a2 - 0.5 * minmod(a1, a2, a3);
        }
        vl[0] = dmax(vl[0], a.ph.ro_lim);
        vr[0] = dmax(vr[0], a.ph.ro_lim);
        hlld_flux<NV>(vl, vr, a.ph.gm, cmax, fl);
#pragma unroll
        for (int n = 0; n < NV; n++) {
          int dst = n;
          if (n == 1) dst = 2; else if (n == 2) dst = 1;
          else if (n == 5) dst = 6; else if (n == 6) dst = 5;
          F[dst * NF + f] = fl[n];
        }
      }
      {
        double vl[5], vr[5], fl[5];
#pragma unroll
        for (int n = 0; n < 5; n++) {
          int src = (n == 1) ? 2 : (n == 2) ? 1 : n;
          const double* col = V + (NV + src) * HY * HX + hx;
          double a0 = col[(fj + 0) * HX], a1 = col[(fj + 1) * HX], a2 = col[(fj + 2) * HX],
                 a3 = col[(fj + 3) * HX];
          vl[n] = a1 + 0.5 * minmod(a0, a1, a2);
          vr[n] = a2 - 0.5 * minmod(a1, a2, a3);
        }
        vl[0] = dmax(vl[0], a.ph.ro_lim);
        vr[0] = dmax(vr[0], a.ph.ro_lim);
        hllc_flux(vl, vr, a.ph.gm, fl);
#pragma unroll
        for (int n = 0; n < 5; n++) {
          int dst = (n == 1) ? 2 : (n == 2) ? 1 : n;
          F[(NV + dst) * NF + f] = fl[n];
        }
      }
    }
    __syncthreads();
    if (valid) {
      const double ay = dt_sub / a.dy;
      const int fB = tj * TX + ti;
#pragma unroll
      for (int n = 0; n < NVT; n++) un[n] = un[n] - ay * (F[n * NF + fB + TX] - F[n * NF + fB]);
    }
  }

  r = RoundOut{1.0e300, 0.0, 0.0, 0.0};
  nanflag = 0;
  if (valid) {
    // add_source (Solver_rot.f90:498-503) for the five fluid variables
#pragma unroll
    for (int n = 0; n < 5; n++) {
      un[n] = un[n] + Sm[n] * dt_sub;
      un[NV + n] = un[NV + n] + Sh[n] * dt_sub;
    }
    if (NV == 9) {  // divb_cleaning_Dedner_source (Scheme_rot.f90:546-554)
      double fac = pexp(-dt_sub / dt * a.ph.safety * a.db_clean);
      un[8] = un[8] * fac;
    }
    if (epi >= 1) {
      double* um = un;
      double* uh = un + NV;
      if (epi >= 2) cfl_rewrite<NV, IR>(um, uh, a);  // main.f90:44-47, its dt unused
      r = cfl_round<NV, IR>(um, uh, acc, ion_pot, a);
#pragma unroll
      for (int n = 0; n < NV; n++) nanflag |= !(un[n] * 0.0 == 0.0);
    }
    store_cell<NDIM, NVT>(a.Uout, N, ix, a.jx, ci, cj, un, a.bnd);
  }
  return valid;
}

// ------------------------------------------------------------ stage kernel
// One workgroup per tile.  The look-ahead round of the fused step (epi=2)
// is skipped on the loop's last step (sc->last), as main.f90 ends there.
template <int NDIM, int NV, bool IR, bool EXT, int TX, int TY, int NT, int MODE>
__global__ void __launch_bounds__(NT, 2) k_stage(StageArgs a) {
  extern __shared__ double smem[];
  DevScalars* sc = a.sc;
  constexpr bool REPLAY = (MODE == STAGE_REPLAY);
  if (!REPLAY && !sc->active) return;
  const double dt = REPLAY ? a.dtc[0] : sc->dt;
  const double cmax = REPLAY ? a.dtc[1] : sc->cmax;
  const int epi = (!REPLAY && a.epi >= 2 && sc->last) ? 1 : a.epi;
  const int ntx = (a.nx + TX - 1) / TX;
  int bid = blockIdx.x;
  if (a.xcd_remap) {
    // workgroups are dealt round-robin to the 8 XCDs: give each XCD a
    // contiguous range of tiles, so that neighbouring tiles (which share
    // halo cells) meet in the same L2
    const int nb = gridDim.x, q = nb / 8, rr = nb % 8, x = bid % 8, k = bid / 8;
    bid = x * q + (x < rr ? x : rr) + k;
  }
  const int i0 = (bid % ntx) * TX;
  const int j0 = (bid / ntx) * TY;
  RoundOut r;
  int nanflag;
  bool valid = stage_tile<NDIM, NV, IR, EXT, TX, TY, NT, MODE>(a, dt, cmax, epi, i0, j0, smem, r, nanflag);
  if (!REPLAY && a.epi >= 1) {
    __syncthreads();
    block_reduce_store<NT>(r, nanflag, valid, sc, smem);
  }
}

// ------------------------------------------------ two-pass stage tile
// The stage of stage_tile for the explicit-collision configuration without
// ionisation/recombination (2-D), one fluid at a time: the plasma's
// primitives and fluxes (NV planes), then the neutrals' (5 planes), in one
// LDS tile of NV planes.  The tile (46 KB at 16x15, against 71 KB for all
// 14 planes) and the register set of one flux function at a time let three
// 256-thread workgroups share a CU.  Every value is computed by the same
// operations as in stage_tile, so the two kernels agree bit for bit.
template <int NV, int TX, int TY>
struct TileGeom2 {
  static constexpr int HX = TX + 4;
  static constexpr int HY = TY + 4;
  static constexpr int NFX = (TX + 1) * TY;
  static constexpr int NFY = TX * (TY + 1);
  static constexpr int NF = (NFX > NFY) ? NFX : NFY;
  static constexpr int LDS_DOUBLES = NV * HY * HX + NV * NF;
};

// One fluid (FL 0: plasma, NV variables from plane 0, HLLD with GLM; FL 1:
// neutrals, 5 variables from plane NV, HLLC) over the tile: halo ->
// primitives, x faces, the SSP blend with the x divergence, y faces, the y
// divergence.  un receives the fluid's part of the updated cell.  The
// plasma's partial cell value waits out the y faces in the cell's own
// output entries (read back by the same thread; in place, the stage base
// there was read first) rather than in registers next to HLLD's.
template <int FL, int NV, bool IR, int TX, int TY, int NT, int MODE>
__device__ __forceinline__ void fluid_pass(const StageArgs& a, double cmax, double c0, double c1,
                                           double dt_sub, int i0, int j0, double* smem, bool valid,
                                           long long oc, int ti, int tj, double* un) {
  constexpr int NFL = FL ? 5 : NV;  // variables of this fluid
  constexpr int OFF = FL ? NV : 0;  // its first plane
  using G = TileGeom2<NV, TX, TY>;
  constexpr int HX = G::HX, HY = G::HY, NFX = G::NFX, NFY = G::NFY, NF = G::NF;
  double* V = smem;
  double* F = smem + NV * HY * HX;
  const int tid = threadIdx.x;
  const long long N = a.N;
  const int ix = a.ix;

  // ---- halo -> primitives (u2v_mhd / u2v_hd act on one fluid each)
  for (int c = tid; c < HX * HY; c += NT) {
    int hx = c % HX, hy = c / HX;
    int gi = i0 + hx, gj = j0 + hy;
    if (gi < a.ix && gj < a.jx) {
      long long o = (long long)gj * ix + gi;
      double u[NFL], v[NFL];
#pragma unroll
      for (int n = 0; n < NFL; n++) u[n] = a.Uin[(OFF + n) * N + o];
      if (IR) {  // set_IR of set_coefficients(.,1) rewrites E (Solver_rot.f90:111-113), per fluid
        double pr, te;
        if (FL == 0) getpr_mhd(u, pr, te, a.ph);
        else getpr_hd(u, pr, te, a.ph);
      }
      if (FL == 0) u2v_mhd<NV>(u, v, a.ph);
      else u2v_hd(u, v, a.ph);
#pragma unroll
      for (int n = 0; n < NFL; n++) V[(n * HY + hy) * HX + hx] = v[n];
      if (MODE != STAGE_PLAIN && a.Vout && hx >= 2 && hx < TX + 2 && gi < a.nx + 2 && hy >= 2 && hy < TY + 2 &&
          gj < a.ny + 2) {
        // a cell of this tile: its primitives and their ghost images
#pragma unroll
        for (int n = 0; n < NFL; n++) a.Vout[(OFF + n) * N + o] = v[n];
        for_each_image(gi, gj, ix, a.jx, a.bnd, [&](int xi, int yj, unsigned m) {
          for (int n = 0; n < NFL; n++)
            a.Vout[(OFF + n) * N + (long long)yj * ix + xi] = ((m >> (OFF + n)) & 1u) ? -v[n] : v[n];
        });
      }
    }
  }
  __syncthreads();

  // ---- x faces
  for (int f = tid; f < NFX; f += NT) {
    int fi = f % (TX + 1), fj = f / (TX + 1);
    if (i0 + fi > a.nx || j0 + fj >= a.ny) continue;
    const double* row = V + (fj + 2) * HX;
    double vl[NFL], vr[NFL];
#pragma unroll
    for (int n = 0; n < NFL; n++) {
      const double* r = row + n * HY * HX + fi;
      double a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3];
      vl[n] = a1 + 0.5 * minmod(a0, a1, a2);
      vr[n] = a2 - 0.5 * minmod(a1, a2, a3);
    }
    vl[0] = dmax(vl[0], a.ph.ro_lim);
    vr[0] = dmax(vr[0], a.ph.ro_lim);
    double fl[NFL];
    if (FL == 0) hlld_flux<NV>(vl, vr, a.ph.gm, cmax, fl);
    else hllc_flux(vl, vr, a.ph.gm, fl);
#pragma unroll
    for (int n = 0; n < NFL; n++) F[n * NF + f] = fl[n];
  }
  __syncthreads();

  // ---- the SSP blend and the x divergence; the blend takes the stage
  // input after set_IR's get_Pr rewrite (with IR) and source_PIP's cq2pv
  // rewrite of this fluid (PIP_rot.f90:1328)
  if (valid) {
    double u[NFL], pv[8];
#pragma unroll
    for (int n = 0; n < NFL; n++) u[n] = a.Uin[(OFF + n) * N + oc];
    if (IR) {
      double pr, te;
      if (FL == 0) getpr_mhd(u, pr, te, a.ph);
      else getpr_hd(u, pr, te, a.ph);
    }
    if (FL == 0) cq2pv_mhd(u, pv, a.ph);
    else cq2pv_hd(u, pv, a.ph);
    const double ax = dt_sub / a.dx;
    const int fL = tj * (TX + 1) + ti;
#pragma unroll
    for (int n = 0; n < NFL; n++) {
      const double u0 = a.U0[(OFF + n) * N + oc];
      double x;
      if (a.istep == 1) x = u0 - ax * (F[n * NF + fL + 1] - F[n * NF + fL]);
      else x = c0 * u0 + c1 * u[n] - ax * (F[n * NF + fL + 1] - F[n * NF + fL]);
      if (FL == 0) a.Uout[(OFF + n) * N + oc] = x;
      else un[n] = x;
    }
  }
  __syncthreads();

  // ---- y faces: permutation (1,3,2,4,5,7,6,8,9) / (1,3,2,4,5)
  // (HLL_rot.f90:414-434, 315-325) within the fluid
  for (int f = tid; f < NFY; f += NT) {
    int fi = f % TX, fj = f / TX;
    if (i0 + fi >= a.nx || j0 + fj > a.ny) continue;
    int hx = fi + 2;
    double vl[NFL], vr[NFL];
#pragma unroll
    for (int n = 0; n < NFL; n++) {
      int src = n;
      if (n == 1) src = 2; else if (n == 2) src = 1;
      else if (FL == 0 && n == 5) src = 6; else if (FL == 0 && n == 6) src = 5;
      const double* col = V + src * HY * HX + hx;
      double a0 = col[(fj + 0) * HX], a1 = col[(fj + 1) * HX], a2 = col[(fj + 2) * HX],
             a3 = col[(fj + 3) * HX];
      vl[n] = a1 + 0.5 * minmod(a0, a1, a2);
      vr[n] = a2 - 0.5 * minmod(a1, a2, a3);
    }
    vl[0] = dmax(vl[0], a.ph.ro_lim);
    vr[0] = dmax(vr[0], a.ph.ro_lim);
    double fl[NFL];
    if (FL == 0) hlld_flux<NV>(vl, vr, a.ph.gm, cmax, fl);
    else hllc_flux(vl, vr, a.ph.gm, fl);
#pragma unroll
    for (int n = 0; n < NFL; n++) {
      int dst = n;
      if (n == 1) dst = 2; else if (n == 2) dst = 1;
      else if (FL == 0 && n == 5) dst = 6; else if (FL == 0 && n == 6) dst = 5;
      F[dst * NF + f] = fl[n];
    }
  }
  __syncthreads();
  if (valid) {
    const double ay = dt_sub / a.dy;
    const int fB = tj * TX + ti;
#pragma unroll
    for (int n = 0; n < NFL; n++) {
      const double x = (FL == 0) ? a.Uout[(OFF + n) * N + oc] : un[n];
      un[n] = x - ay * (F[n * NF + fB + TX] - F[n * NF + fB]);
    }
  }
}

template <int NV, bool IR, int TX, int TY, int NT, int MODE>
__global__ void __launch_bounds__(NT, 3) k_stage2p(StageArgs a) {
  extern __shared__ double smem[];
  constexpr int NVT = NV + 5;
  DevScalars* sc = a.sc;
  constexpr bool REPLAY = (MODE == STAGE_REPLAY);
  if (!REPLAY && !sc->active) return;
  const double dt = REPLAY ? a.dtc[0] : sc->dt;
  const double cmax = REPLAY ? a.dtc[1] : sc->cmax;
  const int epi = (!REPLAY && a.epi >= 2 && sc->last) ? 1 : a.epi;
  const int ntx = (a.nx + TX - 1) / TX;
  int bid = blockIdx.x;
  if (a.xcd_remap) {  // contiguous tile ranges per XCD (as k_stage)
    const int nb = gridDim.x, q = nb / 8, rr = nb % 8, x = bid % 8, k = bid / 8;
    bid = x * q + (x < rr ? x : rr) + k;
  }
  const int i0 = (bid % ntx) * TX;
  const int j0 = (bid / ntx) * TY;
  double c0, c1, dt_sub;
  stage_coeffs(a.istep, a.t_order, dt, c0, c1, dt_sub);
  const int tid = threadIdx.x;
  const int ti = tid % TX, tj = tid / TX;
  const int ci = i0 + ti + 2, cj = j0 + tj + 2;
  const bool valid = (tid < TX * TY) && (i0 + ti < a.nx) && (j0 + tj < a.ny);
  const long long N = a.N;
  const long long oc = (long long)cj * a.ix + ci;
  double un[NVT];
  fluid_pass<0, NV, IR, TX, TY, NT, MODE>(a, cmax, c0, c1, dt_sub, i0, j0, smem, valid, oc, ti, tj, un);
  __syncthreads();  // the neutral pass reuses the tile
  fluid_pass<1, NV, IR, TX, TY, NT, MODE>(a, cmax, c0, c1, dt_sub, i0, j0, smem, valid, oc, ti, tj, un + NV);

  RoundOut r = {1.0e300, 0.0, 0.0, 0.0};
  int nanflag = 0;
  if (valid) {
    // source_PIP on the stage input (it rewrites its copy through cq2pv, as
    // the passes did for the blend), add_source, Dedner damping
    double um[NV], uh[5];
#pragma unroll
    for (int n = 0; n < NV; n++) um[n] = a.Uin[n * N + oc];
#pragma unroll
    for (int n = 0; n < 5; n++) uh[n] = a.Uin[(NV + n) * N + oc];
    const double acc = a.ac ? a.ac[oc] : a.col;
    double gr = 0.0, gi = 0.0, arb = 0.0;
    if (IR) {  // set_IR (rewrites and rates) before source_PIP, as in stage_tile
      set_ir_cell<NV>(um, uh, a.ph, a.irc, gr, gi);
      if (a.ph.flag_ir_type == 0) arb = a.arb_heat[oc];
    }
    double Sm[5], Sh[5], ion_pot;
    Phys phs = a.ph;
    phs.pip_imp = 0;
    source_pip(um, uh, acc, gr, gi, arb, phs, a.irc.heat_fac, Sh, Sm, ion_pot);
    if (a.ion_pot_out) a.ion_pot_out[oc] = ion_pot;
#pragma unroll
    for (int n = 0; n < 5; n++) {
      un[n] = un[n] + Sm[n] * dt_sub;
      un[NV + n] = un[NV + n] + Sh[n] * dt_sub;
    }
    if (NV == 9) {
      double fac = pexp(-dt_sub / dt * a.ph.safety * a.db_clean);
      un[8] = un[8] * fac;
    }
    if (epi >= 1) {
      if (epi >= 2) cfl_rewrite<NV, IR>(un, un + NV, a);  // main.f90:44-47, its dt unused
      r = cfl_round<NV, IR>(un, un + NV, acc, ion_pot, a);
#pragma unroll
      for (int n = 0; n < NV; n++) nanflag |= !(un[n] * 0.0 == 0.0);
    }
    store_cell<2, NVT>(a.Uout, N, a.ix, a.jx, ci, cj, un, a.bnd);
  }
  if (!REPLAY && a.epi >= 1) {
    __syncthreads();
    block_reduce_store<NT>(r, nanflag, valid, sc, smem);
  }
}

// ------------------------------------------------------- pointwise kernels
// Rewrite chain on every cell (ghosts included, as the reference's array
// statements do): flags select set_IR (+rates/arb_heat) and cfl.
struct RoundArgs {
  double* U;
  const double* ac;
  double* gm_rec;
  double* gm_ion;
  double* arb_heat;
  const double* ion_pot;
  DevScalars* sc;
  int ix, jx, nx, ny;
  long long N;
  int do_ir, init_arb, do_cfl, read_rates, do_col;
  double* ac_out;  // do_col with col_type >= 1: alpha_c of every cell
  double col, gmin_safety;
  Phys ph;
  IRConst irc;
};

template <int NV, bool IR>
__global__ void __launch_bounds__(256) k_round(RoundArgs a) {
  constexpr int NVT = NV + 5;
  __shared__ double red_lds[4 * 5];
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = a.N;
  bool inb = k < N;
  RoundOut r = {1.0e300, 0.0, 0.0, 0.0};
  bool interior = false;
  int nanflag = 0;
  if (inb) {
    int i = (int)(k % a.ix), j = (int)(k / a.ix);
    interior = (i >= 2 && i < a.ix - 2) && (a.jx == 1 || (j >= 2 && j < a.jx - 2));
    double u[NVT];
#pragma unroll
    for (int n = 0; n < NVT; n++) u[n] = a.U[n * N + k];
    double* um = u;
    double* uh = u + NV;
    double acc = a.ac ? a.ac[k] : a.col;
    if (a.do_col && col_type(a.ph) >= 1) {  // set_collisional of set_coefficients(.,0)
      acc = set_col_cell(um, uh, a.col, a.ph);
      a.ac_out[k] = acc;
    }
    double gr = 0.0, gi = 0.0;
    if (IR && a.do_ir) {
      set_ir_cell<NV>(um, uh, a.ph, a.irc, gr, gi);
      a.gm_rec[k] = gr;
      a.gm_ion[k] = gi;
      if (a.init_arb) a.arb_heat[k] = gi * uh[0] * a.irc.heat_fac;
    } else if (IR && a.read_rates) {
      gr = a.gm_rec[k];
      gi = a.gm_ion[k];
    }
    if (a.do_cfl) {
      double dtc = cfl_cell<NV>(um, uh, a.gmin_safety, a.ph);
      r.dtc = interior ? dtc : 1.0e300;
      r.rhoac = dmax(um[0] * acc, uh[0] * acc);
      r.gmax = dmax(gr, gi);
      if (IR && a.ph.flag_ir_type == 0) {
        double pr, te;
        getpr_mhd(um, pr, te, a.ph);
        r.ratio = (a.ion_pot && interior) ? a.ion_pot[k] / (pr / (a.ph.gm - 1.0)) : 0.0;
      }
      if (interior) {
#pragma unroll
        for (int n = 0; n < NV; n++) nanflag |= !(u[n] * 0.0 == 0.0);
      }
    }
#pragma unroll
    for (int n = 0; n < NVT; n++) a.U[n * N + k] = u[n];
  }
  if (a.do_cfl) block_reduce_store<256>(r, nanflag, inb, a.sc, red_lds);
}

// Ghost fill by the composite x-then-y map (Boundary_rot.f90:34-125).
__global__ void __launch_bounds__(256) k_bnd(double* U, int nvt, int ix, int jx, long long N, Bnd b,
                                             int ndim) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  int i = (int)(k % ix), j = (int)(k / ix);
  bool gx = (i < 2 || i >= ix - 2);
  bool gy = (ndim == 2) && (j < 2 || j >= jx - 2);
  if (!gx && !gy) return;
  bool fx = false, fy = false;
  int si = gx ? src_index(i, ix, b.type[0], b.type[1], fx) : i;
  int sj = gy ? src_index(j, jx, b.type[2], b.type[3], fy) : j;
  unsigned m = 0;
  if (fx) m ^= b.sym[(i < 2) ? 0 : 1];
  if (fy) m ^= b.sym[(j < 2) ? 2 : 3];
  long long s = (long long)sj * ix + si;
  for (int n = 0; n < nvt; n++) {
    double v = U[n * N + s];
    U[n * N + k] = ((m >> n) & 1u) ? -v : v;
  }
}

// End of cfl: combine the reduction slots into dt (Scheme_rot.f90:183-213).
// mode 0: pip_cfl (optionally into the global dt); mode 2: the same at the
// top of a fused loop (also sets active/last); mode 1: end of a fused step
// (time += dt, stop_sim, next dt and cmax, activity flags).  In mode 1 a
// recorded run also passes hist: the next step's {dt, cmax} go to
// hist[2 hist_k], hist[2 hist_k + 1], as k_record_dt would write them.
struct FinArgs {
  DevScalars* sc;
  int mode, into_global, flag_ir, flag_ir_type, flag_divb, pip_imp;
  double safety, dxyz_min, pip_imp_factor;
  double* hist;
  int hist_k;
};

__device__ __forceinline__ void record_dt(const DevScalars* sc, double* hist, int k) {
  hist[2 * k] = sc->active ? sc->dt : 0.0;
  hist[2 * k + 1] = sc->cmax;
}

__global__ void k_finalize(FinArgs f) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DevScalars* sc = f.sc;
  if (f.mode == 1 && !sc->active) {
    if (f.hist) record_dt(sc, f.hist, f.hist_k);
    return;
  }
  if (f.mode == 1) sc->ion_pot_valid = 1;  // source_PIP ran in this step
  double dttemp = dmin(10.0, as_f64(sc->red[RED_DT]));
  // cfl_pn_col (Scheme_rot.f90:305-313)
  if (f.pip_imp == 1) dttemp = dmin(dttemp, f.pip_imp_factor / as_f64(sc->red[RED_RHOAC]));
  else dttemp = dmin(dttemp, dmin(f.safety, 0.3) / as_f64(sc->red[RED_RHOAC]));
  if (f.flag_ir >= 1) {
    dttemp = dmin(dttemp, dmin(f.safety, 0.1) / dmax(as_f64(sc->red[RED_GMAX]), 1.0e-5));
    if (f.flag_ir_type == 0) {
      double mx = sc->ion_pot_valid ? as_f64(sc->red[RED_IONPOT]) : 0.0;
      dttemp = dmin(dttemp, f.safety / dmax(mx, 1.0e-5));
    }
  }
  int nan = (int)(sc->red[RED_NAN] != 0ull);
  sc->red[RED_DT] = as_u64(1.0e300);
  sc->red[RED_RHOAC] = 0ull;
  sc->red[RED_GMAX] = 0ull;
  sc->red[RED_IONPOT] = 0ull;
  sc->red[RED_NAN] = 0ull;
  sc->dt_out = dttemp;
  if (f.mode == 0 || f.mode == 2) {
    if (f.into_global) sc->dt = dttemp;
    if (f.flag_divb >= 1) sc->cmax = f.safety * f.dxyz_min / sc->dt;
    sc->status = nan ? 2 : 0;
    if (f.mode == 2) {  // top of the time loop: decide whether this step is the last
      bool go = (sc->time < sc->tend) && (sc->nt < sc->nmax);
      sc->active = go ? 1 : 0;
      sc->last = ((sc->time + sc->dt >= sc->tend) || (sc->nt + 1 >= sc->nmax) || (sc->dt <= 1.0e-12)) ? 1 : 0;
    }
    return;
  }
  // fused step end: main.f90:58-62 then the next iteration's cfl(dt)
  double dt_used = sc->dt;
  sc->time = sc->time + dt_used;
  sc->nt = sc->nt + 1;
  int st = 0;
  if (dt_used <= 1.0e-12) st |= 1;
  if (nan) st |= 2;
  sc->status = st;
  bool more = (sc->time < sc->tend) && (sc->nt < sc->nmax) && (st == 0);
  sc->active = more ? 1 : 0;
  if (more) {
    sc->dt = dttemp;
    if (f.flag_divb >= 1) sc->cmax = f.safety * f.dxyz_min / sc->dt;
    sc->last = ((sc->time + sc->dt >= sc->tend) || (sc->nt + 1 >= sc->nmax) || (sc->dt <= 1.0e-12)) ? 1 : 0;
  }
  if (f.hist) record_dt(sc, f.hist, f.hist_k);
}

}  // namespace PIP_NS

namespace PIP_NS {
// Small control kernel: set loop bookkeeping fields on the device.
__global__ void k_set_loop(DevScalars* sc, double tend, long long nmax, int active, int ion_valid) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sc->tend = tend;
  sc->nmax = nmax;
  if (active >= 0) sc->active = active;
  if (ion_valid >= 0) sc->ion_pot_valid = ion_valid;
}
}  // namespace PIP_NS

namespace PIP_NS {
// ------------------------------------------------------ batched ensemble
// One persistent launch integrates an ensemble of independent 1-D runs of
// different sizes: each workgroup pulls a member index from a device queue
// (longest expected run first) and runs that member's whole time loop
// (main.f90:31-68) in a loop over TX-cell tiles of stage_tile, with its own
// dt, time and step count.  The arithmetic per cell is that of the plain
// path, so every member equals its plain run bit for bit.
struct EnsMember {
  long long off;  // offset (doubles) of the member's block: nv + 5 planes of ix * jx
  int nx, ix;
  double dx, col, gmin_safety, tend;
  long long nmax;
  // 2-D members (ny = 0 in 1-D): rows, padded rows, dy, min(dx, dy, dz)
  int ny, jx;
  double dy, dxyz_min;
};
struct EnsOut {
  double time, dt;
  long long nt;
  int status, pad;
};
struct EnsArgs {
  double* U;
  double* U1;
  const EnsMember* mem;
  EnsOut* out;
  const int* order;
  unsigned* counter;
  int n_members;
  Phys ph;
  Bnd bnd;
  double db_clean;
};

__global__ void k_reset_counter(unsigned* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

// Workgroup min/max of the dt terms; every thread gets the result.
template <int NT>
__device__ __forceinline__ RoundOut wg_reduce(RoundOut r, int& fl, double* lds) {
  for (int off = 32; off > 0; off >>= 1) {
    r.dtc = dmin(r.dtc, __shfl_xor(r.dtc, off, 64));
    r.rhoac = dmax(r.rhoac, __shfl_xor(r.rhoac, off, 64));
    fl |= __shfl_xor(fl, off, 64);
  }
  constexpr int NW = (NT + 63) / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    lds[wid * 3 + 0] = r.dtc;
    lds[wid * 3 + 1] = r.rhoac;
    lds[wid * 3 + 2] = (double)fl;
  }
  __syncthreads();
  RoundOut o = {lds[0], lds[1], 0.0, 0.0};
  int f = (int)lds[2];
  for (int w = 1; w < NW; w++) {
    o.dtc = dmin(o.dtc, lds[w * 3 + 0]);
    o.rhoac = dmax(o.rhoac, lds[w * 3 + 1]);
    f |= (int)lds[w * 3 + 2];
  }
  __syncthreads();
  fl = f;
  return o;
}

// cfl rewrite round over all cells of a member (ghosts included), as
// k_round does for the plain path; returns the reduced dt terms.
template <int TX>
__device__ __forceinline__ RoundOut ens_cfl_cells(double* U, const StageArgs& a, double* lds) {
  RoundOut acc = {1.0e300, 0.0, 0.0, 0.0};
  int fl = 0;
  for (int k = threadIdx.x; k < a.ix; k += TX) {
    double u[13];
#pragma unroll
    for (int n = 0; n < 13; n++) u[n] = U[n * a.N + k];
    double dtc = cfl_cell<8>(u, u + 8, a.gmin_safety, a.ph);
    bool interior = (k >= 2 && k < a.ix - 2);
    if (interior) acc.dtc = dmin(acc.dtc, dtc);
    acc.rhoac = dmax(acc.rhoac, dmax(u[0] * a.col, u[8] * a.col));
#pragma unroll
    for (int n = 0; n < 13; n++) U[n * a.N + k] = u[n];
  }
  return wg_reduce<TX>(acc, fl, lds);
}

__device__ __forceinline__ double ens_dt(const RoundOut& o, double safety) {
  // cfl without IR: Scheme_rot.f90:186-201
  double dttemp = dmin(10.0, o.dtc);
  return dmin(dttemp, dmin(safety, 0.3) / o.rhoac);
}

// StageArgs of ensemble member M (1-D, nvar_m = 8, constant alpha_c,
// explicit collisions, no IR); the state pointers are set per stage.
__device__ __forceinline__ StageArgs ens_stage_args(const EnsArgs& e, const EnsMember& M) {
  StageArgs a;
  a.ac = nullptr;
  a.ac_out = nullptr;
  a.U0out = nullptr;
  a.xcd_remap = 0;
  a.arb_heat = nullptr;
  a.ion_pot_out = nullptr;
  a.Vout = nullptr;
  a.dtc = nullptr;
  a.sc = nullptr;
  a.ix = M.ix;
  a.jx = 1;
  a.nx = M.nx;
  a.ny = 1;
  a.N = M.ix;
  a.t_order = 2;
  a.col = M.col;
  a.gmin_safety = M.gmin_safety;
  a.dxyz_min = M.dx;
  a.dx = M.dx;
  a.dy = 100.0;
  a.db_clean = e.db_clean;
  a.ph = e.ph;
  a.irc = IRConst{0, 0, 0, 0, 0, 0, 0};
  a.bnd = e.bnd;
  return a;
}

// Start of a member's run by one workgroup: the boundary map of the
// initial state (Initial_rot.f90:160), the prologue cfl and the loop-top
// cfl of the first iteration; returns the first dt.
template <int TX>
__device__ __forceinline__ double ens_start(const EnsArgs& e, const EnsMember& M, const StageArgs& a,
                                            double* U, double* lds_red) {
  if (threadIdx.x < 4) {
    int g = (threadIdx.x < 2) ? threadIdx.x : M.ix - 4 + threadIdx.x;
    bool flip;
    int sidx = src_index(g, M.ix, e.bnd.type[0], e.bnd.type[1], flip);
    unsigned msk = flip ? e.bnd.sym[(g < 2) ? 0 : 1] : 0u;
    for (int n = 0; n < 13; n++) {
      double v = U[n * a.N + sidx];
      U[n * a.N + g] = ((msk >> n) & 1u) ? -v : v;
    }
  }
  __syncthreads();
  ens_cfl_cells<TX>(U, a, lds_red);
  __syncthreads();
  RoundOut o = ens_cfl_cells<TX>(U, a, lds_red);
  return ens_dt(o, e.ph.safety);
}

template <int TX>
__global__ void __launch_bounds__(TX) k_ensemble(EnsArgs e) {
  extern __shared__ double smem[];
  __shared__ int s_member;
  constexpr int TC = TX - 1;  // cells per tile: TC+1 faces = one pass of TX threads
  using G = TileGeom<1, 8, TC, 1>;
  double* lds_red = smem + G::LDS_DOUBLES;  // 3 * (TX/64) doubles past the tile
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_member = (int)atomicAdd(e.counter, 1u);
    __syncthreads();
    const int q = s_member;
    if (q >= e.n_members) break;
    const int m = e.order[q];
    const EnsMember M = e.mem[m];
    double* U = e.U + M.off;
    double* U1 = e.U1 + M.off;
    StageArgs a = ens_stage_args(e, M);
    double dt = ens_start<TX>(e, M, a, U, lds_red);
    double time = 0.0;
    long long nt = 0;
    int status = 0;
    while (time < M.tend && nt < M.nmax && status == 0) {
      const bool last = (time + dt >= M.tend) || (nt + 1 >= M.nmax) || (dt <= 1.0e-12);
      RoundOut r;
      int nanflag;
      a.istep = 1;
      a.Uin = U;
      a.U0 = U;
      a.Uout = U1;
      a.epi = 0;
      for (int i0 = 0; i0 < M.nx; i0 += TC) {
        stage_tile<1, 8, false, false, TC, 1, TX>(a, dt, 0.0, 0, i0, 0, smem, r, nanflag);
        __syncthreads();
      }
      __threadfence_block();
      __syncthreads();
      a.istep = 2;
      a.Uin = U1;
      a.U0 = U;
      a.Uout = U;
      const int epi = last ? 1 : 2;
      RoundOut acc = {1.0e300, 0.0, 0.0, 0.0};
      int fl = 0;
      for (int i0 = 0; i0 < M.nx; i0 += TC) {
        bool v = stage_tile<1, 8, false, false, TC, 1, TX>(a, dt, 0.0, epi, i0, 0, smem, r, nanflag);
        if (v) {
          acc.dtc = dmin(acc.dtc, r.dtc);
          acc.rhoac = dmax(acc.rhoac, r.rhoac);
          fl |= nanflag;
        }
        __syncthreads();
      }
      __threadfence_block();
      RoundOut red = wg_reduce<TX>(acc, fl, lds_red);
      double dt_new = ens_dt(red, e.ph.safety);
      time = time + dt;
      nt = nt + 1;
      if (dt <= 1.0e-12) status |= 1;
      if (fl) status |= 2;
      if (time < M.tend && nt < M.nmax && status == 0) dt = dt_new;
    }
    if (threadIdx.x == 0) {
      EnsOut out;
      out.time = time;
      out.dt = dt;
      out.nt = nt;
      out.status = status;
      out.pad = 0;
      e.out[m] = out;
    }
  }
}

// ------------------------------------------------------ lockstep ensemble
// The ensemble with every member spread over workgroups instead of one: a
// start kernel (one workgroup per member), then per step two stage launches
// over a table of (member, tile) pairs of the running members.  Per-member
// dt, time, step count, flags and reductions live in DevScalars, two
// buffers by step parity: stage 1 of step k applies the end of step k-1
// (read from buffer (k-1)&1, in every workgroup of the member alike) and the
// member's first tile writes the result to buffer k&1, which stage 2 then
// reads and reduces into; at the end of a host chunk, k_lock_end applies
// the last step's end.  Every operation is the persistent loop's
// (k_ensemble) and the plain path's, so the results agree with both bit for
// bit; the members' tiles run in parallel, as in the plain path, instead of
// one after another on one CU.
struct LockTile {
  int m, i0, j0, pad;  // member, first cell of the tile (column, row)
};

// The persistent loop's step end for one member: dt of the next step from
// the reductions, time, step count, status, the active / last flags, and
// reset slots.  A member that has stopped keeps its state; *stops says
// whether it stops at this step end.
__device__ __forceinline__ DevScalars lock_step_end(const DevScalars& sc, const EnsMember& M, double safety,
                                                   bool* stops) {
  DevScalars o = sc;
  *stops = false;
  if (!sc.active) return o;
  const RoundOut red = {as_f64(sc.red[RED_DT]), as_f64(sc.red[RED_RHOAC]), 0.0, 0.0};
  const bool fl = sc.red[RED_NAN] != 0ull;
  o.red[RED_DT] = as_u64(1.0e300);
  o.red[RED_RHOAC] = 0ull;
  o.red[RED_GMAX] = 0ull;
  o.red[RED_IONPOT] = 0ull;
  o.red[RED_NAN] = 0ull;
  const double dt = sc.dt;
  const double dt_new = ens_dt(red, safety);
  o.time = sc.time + dt;
  o.nt = sc.nt + 1;
  int status = sc.status;
  if (dt <= 1.0e-12) status |= 1;
  if (fl) status |= 2;
  o.status = status;
  if (o.time < M.tend && o.nt < M.nmax && status == 0) {
    o.dt = dt_new;
    if (M.ny > 0) o.cmax = safety * M.dxyz_min / dt_new;  // GLM speed of 2-D members (psi)
    o.last = ((o.time + dt_new >= M.tend) || (o.nt + 1 >= M.nmax) || (dt_new <= 1.0e-12)) ? 1 : 0;
  } else {
    o.active = 0;
    *stops = true;
  }
  return o;
}

template <int TX>
__global__ void __launch_bounds__(TX) k_lock_start(EnsArgs e, DevScalars* scs) {
  extern __shared__ double smem[];
  constexpr int TC = TX - 1;
  using G = TileGeom<1, 8, TC, 1>;
  double* lds_red = smem + G::LDS_DOUBLES;
  const int m = blockIdx.x;
  if (m >= e.n_members) return;
  const EnsMember M = e.mem[m];
  const StageArgs a = ens_stage_args(e, M);
  const double dt = ens_start<TX>(e, M, a, e.U + M.off, lds_red);
  if (threadIdx.x == 0) {
    DevScalars sc;
    sc.dt = dt;
    sc.time = 0.0;
    sc.cmax = 0.0;
    sc.tend = M.tend;
    sc.dt_out = dt;
    sc.nt = 0;
    sc.nmax = M.nmax;
    sc.status = 0;
    sc.ion_pot_valid = 0;
    sc.red[RED_DT] = as_u64(1.0e300);
    sc.red[RED_RHOAC] = 0ull;
    sc.red[RED_GMAX] = 0ull;
    sc.red[RED_IONPOT] = 0ull;
    sc.red[RED_NAN] = 0ull;
    const bool go = (0.0 < M.tend) && (0 < M.nmax);
    sc.active = go ? 1 : 0;
    sc.last = ((0.0 + dt >= M.tend) || (1 >= M.nmax) || (dt <= 1.0e-12)) ? 1 : 0;
    scs[m] = sc;  // both parity buffers
    scs[e.n_members + m] = sc;
    if (!go) e.out[m] = EnsOut{0.0, dt, 0, 0, 0};
  }
}

// Stage ISTEP (1 or 2) of step k of the running members, one tile per
// workgroup; p = k & 1.  Stage 1 with apply_end first applies step k-1's
// end; stage 2 carries the cfl rounds and reduces into the member's slots.
template <int TX, int ISTEP>
__global__ void __launch_bounds__(TX) k_lock_stage(EnsArgs e, DevScalars* scs, const LockTile* tiles, int p,
                                                   int apply_end) {
  extern __shared__ double smem[];
  constexpr int TC = TX - 1;
  const LockTile t = tiles[blockIdx.x];
  const EnsMember M = e.mem[t.m];
  DevScalars* sc = scs + (long long)p * e.n_members + t.m;
  double dt;
  int last;
  if (ISTEP == 1 && apply_end) {
    bool stops;
    const DevScalars nw = lock_step_end(scs[(long long)(p ^ 1) * e.n_members + t.m], M, e.ph.safety, &stops);
    if (t.i0 == 0 && threadIdx.x == 0) {
      *sc = nw;
      if (stops) e.out[t.m] = EnsOut{nw.time, nw.dt, nw.nt, nw.status, 0};
    }
    if (!nw.active) return;
    dt = nw.dt;
    last = nw.last;
  } else {
    if (!sc->active) return;
    dt = sc->dt;
    last = sc->last;
  }
  StageArgs a = ens_stage_args(e, M);
  double* U = e.U + M.off;
  double* U1 = e.U1 + M.off;
  a.istep = ISTEP;
  a.Uin = (ISTEP == 1) ? U : U1;
  a.U0 = U;
  a.Uout = (ISTEP == 1) ? U1 : U;
  a.epi = 0;
  const int epi = (ISTEP == 1) ? 0 : (last ? 1 : 2);
  RoundOut r;
  int nanflag;
  const bool valid = stage_tile<1, 8, false, false, TC, 1, TX>(a, dt, 0.0, epi, t.i0, 0, smem, r, nanflag);
  if (ISTEP == 2) {
    __syncthreads();
    block_reduce_store<TX>(r, nanflag, valid, sc, smem);
  }
}

// Number of members still running, from scalar buffer 0 (where a chunk's
// closing k_lock_end leaves them); one workgroup.
__global__ void k_lock_count(const DevScalars* scs, int n, int* active) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int c = 0;
  for (int m = (int)threadIdx.x; m < n; m += (int)blockDim.x) c += scs[m].active ? 1 : 0;
  if (c) atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) *active = cnt;
}

// ---------------------------------------------------- 2-D ensemble members
// Members of a 2-D ensemble: the plain path's 2-D stage (psi, no IR,
// constant alpha_c, explicit collisions) on 16x15 tiles of 256 threads,
// every member's tiles side by side in the lockstep launches.
constexpr int kEns2TX = 16, kEns2TY = 15, kEns2NT = 256;

__device__ __forceinline__ StageArgs ens_stage_args2(const EnsArgs& e, const EnsMember& M) {
  StageArgs a = ens_stage_args(e, M);
  a.jx = M.jx;
  a.ny = M.ny;
  a.N = (long long)M.ix * M.jx;
  a.dy = M.dy;
  a.dxyz_min = M.dxyz_min;
  return a;
}

// cfl rewrite round over all cells of a 2-D member (ghosts included), as
// k_round does for the plain path; returns the reduced dt terms.
template <int NT>
__device__ __forceinline__ RoundOut ens_cfl_cells2(double* U, const StageArgs& a, double* lds) {
  RoundOut acc = {1.0e300, 0.0, 0.0, 0.0};
  int fl = 0;
  for (long long k = threadIdx.x; k < a.N; k += NT) {
    double u[14];
#pragma unroll
    for (int n = 0; n < 14; n++) u[n] = U[n * a.N + k];
    const double dtc = cfl_cell<9>(u, u + 9, a.gmin_safety, a.ph);
    const int i = (int)(k % a.ix), j = (int)(k / a.ix);
    if (i >= 2 && i < a.ix - 2 && j >= 2 && j < a.jx - 2) acc.dtc = dmin(acc.dtc, dtc);
    acc.rhoac = dmax(acc.rhoac, dmax(u[0] * a.col, u[9] * a.col));
#pragma unroll
    for (int n = 0; n < 14; n++) U[n * a.N + k] = u[n];
  }
  return wg_reduce<NT>(acc, fl, lds);
}

// Start of a 2-D member by one workgroup: the boundary map of the initial
// state (ghosts from interior sources, k_bnd's composite map), the prologue
// cfl and the loop-top cfl; writes both scalar buffers.
template <int NT>
__global__ void __launch_bounds__(NT) k_lock_start2(EnsArgs e, DevScalars* scs) {
  __shared__ double lds_red[3 * (NT / 64)];
  const int m = blockIdx.x;
  if (m >= e.n_members) return;
  const EnsMember M = e.mem[m];
  const StageArgs a = ens_stage_args2(e, M);
  double* U = e.U + M.off;
  for (long long k = threadIdx.x; k < a.N; k += NT) {
    const int i = (int)(k % a.ix), j = (int)(k / a.ix);
    const bool gx = (i < 2 || i >= a.ix - 2), gy = (j < 2 || j >= a.jx - 2);
    if (!gx && !gy) continue;
    bool fx = false, fy = false;
    const int si = gx ? src_index(i, a.ix, e.bnd.type[0], e.bnd.type[1], fx) : i;
    const int sj = gy ? src_index(j, a.jx, e.bnd.type[2], e.bnd.type[3], fy) : j;
    unsigned msk = 0u;
    if (fx) msk ^= e.bnd.sym[(i < 2) ? 0 : 1];
    if (fy) msk ^= e.bnd.sym[(j < 2) ? 2 : 3];
    const long long src = (long long)sj * a.ix + si;
    for (int n = 0; n < 14; n++) {
      const double v = U[n * a.N + src];
      U[n * a.N + k] = ((msk >> n) & 1u) ? -v : v;
    }
  }
  __syncthreads();
  ens_cfl_cells2<NT>(U, a, lds_red);
  __syncthreads();
  const RoundOut o = ens_cfl_cells2<NT>(U, a, lds_red);
  const double dt = ens_dt(o, e.ph.safety);
  if (threadIdx.x == 0) {
    DevScalars sc;
    sc.dt = dt;
    sc.time = 0.0;
    sc.cmax = e.ph.safety * M.dxyz_min / dt;
    sc.tend = M.tend;
    sc.dt_out = dt;
    sc.nt = 0;
    sc.nmax = M.nmax;
    sc.status = 0;
    sc.ion_pot_valid = 0;
    sc.red[RED_DT] = as_u64(1.0e300);
    sc.red[RED_RHOAC] = 0ull;
    sc.red[RED_GMAX] = 0ull;
    sc.red[RED_IONPOT] = 0ull;
    sc.red[RED_NAN] = 0ull;
    const bool go = (0.0 < M.tend) && (0 < M.nmax);
    sc.active = go ? 1 : 0;
    sc.last = ((0.0 + dt >= M.tend) || (1 >= M.nmax) || (dt <= 1.0e-12)) ? 1 : 0;
    scs[m] = sc;
    scs[e.n_members + m] = sc;
    if (!go) e.out[m] = EnsOut{0.0, dt, 0, 0, 0};
  }
}

// Stage ISTEP of step k (p = k & 1) for one 16x15 tile of a 2-D member, as
// k_lock_stage does for 1-D members.
template <int ISTEP>
__global__ void __launch_bounds__(kEns2NT, 2) k_lock_stage2(EnsArgs e, DevScalars* scs, const LockTile* tiles,
                                                            int p, int apply_end) {
  extern __shared__ double smem[];
  const LockTile t = tiles[blockIdx.x];
  const EnsMember M = e.mem[t.m];
  DevScalars* sc = scs + (long long)p * e.n_members + t.m;
  double dt, cmax;
  int last;
  if (ISTEP == 1 && apply_end) {
    bool stops;
    const DevScalars nw = lock_step_end(scs[(long long)(p ^ 1) * e.n_members + t.m], M, e.ph.safety, &stops);
    if (t.i0 == 0 && t.j0 == 0 && threadIdx.x == 0) {
      *sc = nw;
      if (stops) e.out[t.m] = EnsOut{nw.time, nw.dt, nw.nt, nw.status, 0};
    }
    if (!nw.active) return;
    dt = nw.dt;
    cmax = nw.cmax;
    last = nw.last;
  } else {
    if (!sc->active) return;
    dt = sc->dt;
    cmax = sc->cmax;
    last = sc->last;
  }
  StageArgs a = ens_stage_args2(e, M);
  double* U = e.U + M.off;
  double* U1 = e.U1 + M.off;
  a.istep = ISTEP;
  a.Uin = (ISTEP == 1) ? U : U1;
  a.U0 = U;
  a.Uout = (ISTEP == 1) ? U1 : U;
  a.epi = 0;
  const int epi = (ISTEP == 1) ? 0 : (last ? 1 : 2);
  RoundOut r;
  int nanflag;
  const bool valid = stage_tile<2, 9, false, false, kEns2TX, kEns2TY, kEns2NT>(a, dt, cmax, epi, t.i0, t.j0, smem,
                                                                               r, nanflag);
  if (ISTEP == 2) {
    __syncthreads();
    block_reduce_store<kEns2NT>(r, nanflag, valid, sc, smem);
  }
}

// Step end of the listed members (those of the chunk's table) after step k,
// p = k & 1: from buffer p into buffer p ^ 1; a member that stops writes its
// EnsOut.
__global__ void k_lock_end(EnsArgs e, DevScalars* scs, const int* members, int count, int p) {
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= count) return;
  const int m = members[q];
  bool stops;
  const DevScalars nw = lock_step_end(scs[(long long)p * e.n_members + m], e.mem[m], e.ph.safety, &stops);
  scs[(long long)(p ^ 1) * e.n_members + m] = nw;
  if (stops) e.out[m] = EnsOut{nw.time, nw.dt, nw.nt, nw.status, 0};
}
// Step bookkeeping for recorded / replayed runs: hist[2k] = dt, hist[2k+1]
// = cmax of step k.
__global__ void k_record_dt(const DevScalars* sc, double* hist, int k) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  record_dt(sc, hist, k);
}
// End of a frozen-dt replay of n steps: the scalars a step-by-step run
// leaves (time and step count advanced, dt / cmax of the last step).
__global__ void k_replay_end(DevScalars* sc, const double* hist, int n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t = sc->time;
  for (int k = 0; k < n; k++) t = t + hist[2 * k];
  sc->time = t;
  sc->nt = sc->nt + n;
  sc->dt = hist[2 * (n - 1)];
  sc->cmax = hist[2 * (n - 1) + 1];
  sc->active = 1;
  sc->last = 0;
}
__global__ void k_set_dt(DevScalars* sc, const double* hist, int k) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sc->dt = hist[2 * k];
  sc->cmax = hist[2 * k + 1];
  sc->active = 1;
  sc->last = 0;
}


}  // namespace PIP_NS
