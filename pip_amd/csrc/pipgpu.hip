// pipgpu.hip -- C-ABI of libpipgpu.so (declared in include/pipgpu.h).
//
// Owns the device state of one domain per context, dispatches the stage,
// rewrite/cfl, boundary and finalize kernels of pip_kernels.h on one HIP
// stream (default priority) and mirrors the reference's procedures:
// prologue, set_IR, cfl, run_solver, PIPbnd and the main.f90 time loop.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pipgpu.h"
#include "pip_adj_launch.h"
#include "pip_kernels.h"
#include "pip_relaxed.h"

using namespace pip;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(PIP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

// The device quotient and square root of the face phases, elementwise.
__global__ void k_check_arith(const double* n, const double* d, double* q, double* r, long long count) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  q[k] = qdiv(n[k], d[k]);
  r[k] = qsqrt(n[k]);
}

// Work memory of pip_adj_step_noalloc: the face-state adjoint pairs and
// branch codes of both directions and the adjoint of the stage-1 result,
// (5 nvt + 4) doubles per padded cell, in one fixed device array per GPU
// (held until the process ends).  Host code never passes the array's name
// to a kernel: it is a host-side stand-in for the symbol.  The device
// address comes from hipGetSymbolAddress, taken once per device when the
// first context is created on it (pip_create).
constexpr size_t kAdjScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) double scratch[kAdjScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];
static bool g_scratch_done[kMaxDevices];

// Once per device, with that device current: the device address of scratch.
static void init_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[dev]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_ptr[dev] = (e == hipSuccess) ? static_cast<double*>(ptr) : nullptr;
  g_scratch_err[dev] = e;
  g_scratch_done[dev] = true;
}

struct pip_ctx {
  pip_cfg cfg;
  int nv, nvt, nx, ny;
  long long N;
  double* U = nullptr;   // state (U_m planes then U_h planes)
  double* U1 = nullptr;  // stage buffers
  double* U2 = nullptr;
  double* ac = nullptr;
  double* gm_rec = nullptr;
  double* gm_ion = nullptr;
  double* arb_heat = nullptr;
  double* ion_pot = nullptr;
  double* Ub = nullptr;  // implicit collisions: updated stage base
  DevScalars* sc = nullptr;
  hipStream_t stream = nullptr;
  int tx = 16, ty = 15;
  int xcd_remap = 0;  // contiguous tile ranges per XCD (autotuned)
  int two_pass = 0;   // bit s-1: stage s takes the fluids one at a time (autotuned)
  int arith = PIP_ARITH_EXACT;  // stage kernels: exact (reference order) or relaxed build
  Phys ph;
  IRConst irc;
  Bnd bnd;
  double gmin_safety, dxyz_min;
  bool arb_done = false;
  bool arb_given = false;  // arb_heat uploaded for a restart: the next prologue keeps it
  bool top_pending = true;
  int device = 0;
  // adjoint work buffers (allocated on the first pip_adj_step)
  double* adjV = nullptr;   // nvt planes
  double* adjP = nullptr;   // 2 dirs x nvt planes of (left, right) face-state pairs + 4 planes of routing codes
  double* lamB = nullptr;   // nvt planes
  // pip_advance replays a captured chunk of fused steps (hipGraph); the key
  // names the stage kernels it was captured with
  int step_graph = 1;
  hipGraphExec_t sgraph = nullptr;
  long long sgraph_key = -1;
  long long eager_key = -1;  // kernel choice of the last chunk run from the stream (attributes set)
};

// --------------------------------------------------------------- helpers

static void build_bnd(pip_ctx* c) { c->bnd = make_bnd(c->cfg.flag_bnd, c->cfg.ndim, c->nv, c->cfg.flag_divb); }

static DevScalars host_scalars_init() {
  DevScalars s;
  memset(&s, 0, sizeof(s));
  s.dt = 0.0;
  s.time = 0.0;
  s.cmax = 0.0;
  s.tend = 1.0e300;
  s.red[RED_DT] = as_u64(1.0e300);
  s.nmax = (long long)1 << 62;
  s.active = 1;
  return s;
}

static StageArgs stage_args(pip_ctx* c, int istep, const double* Uin, const double* U0, double* Uout,
                            int epi, double* ion_out) {
  StageArgs a;
  a.Uin = Uin;
  a.U0 = U0;
  a.Uout = Uout;
  a.U0out = nullptr;
  if (c->cfg.flag_pip_imp == 1) {  // the stage base is updated by every stage
    a.U0 = (istep == 1) ? U0 : c->Ub;
    a.U0out = (istep < c->cfg.t_order) ? c->Ub : nullptr;
  }
  a.ac = c->cfg.ac_field ? c->ac : nullptr;
  a.ac_out = (col_type(c->ph) >= 1) ? c->ac : nullptr;
  a.arb_heat = c->arb_heat;
  a.ion_pot_out = ion_out;
  a.Vout = nullptr;
  a.dtc = nullptr;
  a.sc = c->sc;
  a.ix = c->cfg.ix;
  a.jx = c->cfg.jx;
  a.nx = c->nx;
  a.ny = c->ny;
  a.N = c->N;
  a.istep = istep;
  a.t_order = c->cfg.t_order;
  a.epi = epi;
  a.xcd_remap = c->xcd_remap;
  a.col = c->cfg.col;
  a.gmin_safety = c->gmin_safety;
  a.dxyz_min = c->dxyz_min;
  a.dx = c->cfg.dx;
  a.dy = c->cfg.dy;
  a.db_clean = c->cfg.db_clean;
  a.ph = c->ph;
  a.irc = c->irc;
  a.bnd = c->bnd;
  return a;
}

template <int NDIM, int NV, bool IR, bool EXT, int TX, int TY, int NT, int MODE>
static int launch_stage_t(pip_ctx* c, const StageArgs& a) {
  constexpr size_t lds = sizeof(double) * (size_t)TileGeom<NDIM, NV, TX, TY>::LDS_DOUBLES;
  static bool attr_done = false;
  auto kern = k_stage<NDIM, NV, IR, EXT, TX, TY, NT, MODE>;
  if (!attr_done) {
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_done = true;
  }
  int ntx = (c->nx + TX - 1) / TX;
  int nty = (NDIM == 2) ? (c->ny + TY - 1) / TY : 1;
  hipLaunchKernelGGL(kern, dim3(ntx * nty), dim3(NT), lds, c->stream, a);
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

template <int NDIM, int NV, bool IR, bool EXT, int MODE = STAGE_PLAIN>
static int launch_stage_tile(pip_ctx* c, const StageArgs& a) {
  if constexpr (NDIM == 1) {
    return launch_stage_t<1, NV, IR, EXT, 255, 1, 256, MODE>(c, a);
  } else {
    // tile shapes (cells) and workgroup sizes; the single-pass shapes keep
    // every face phase within one pass of the workgroup
    if (c->tx == 16 && c->ty == 15) return launch_stage_t<2, NV, IR, EXT, 16, 15, 256, MODE>(c, a);
    if (c->tx == 32 && c->ty == 7) return launch_stage_t<2, NV, IR, EXT, 32, 7, 256, MODE>(c, a);
    if (c->tx == 32 && c->ty == 15) return launch_stage_t<2, NV, IR, EXT, 32, 15, 512, MODE>(c, a);
    if (c->tx == 64 && c->ty == 7) return launch_stage_t<2, NV, IR, EXT, 64, 7, 512, MODE>(c, a);
    if (c->tx == 32 && c->ty == 8) return launch_stage_t<2, NV, IR, EXT, 32, 8, 256, MODE>(c, a);
    if (c->tx == 16 && c->ty == 7) return launch_stage_t<2, NV, IR, EXT, 16, 7, 128, MODE>(c, a);
    return fail(PIP_ERR_ARG, "unsupported tile shape");
  }
}

template <int NV, bool IR, int TX, int TY, int NT, int MODE>
static int launch_stage2p_t(pip_ctx* c, const StageArgs& a) {
  constexpr size_t lds = sizeof(double) * (size_t)TileGeom2<NV, TX, TY>::LDS_DOUBLES;
  static bool attr_done = false;
  auto kern = k_stage2p<NV, IR, TX, TY, NT, MODE>;
  if (!attr_done) {
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_done = true;
  }
  const int ntx = (c->nx + TX - 1) / TX, nty = (c->ny + TY - 1) / TY;
  hipLaunchKernelGGL(kern, dim3(ntx * nty), dim3(NT), lds, c->stream, a);
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

// the fluid-by-fluid stage kernel: 2-D, explicit constant/per-cell alpha,
// with or without set_IR (replays: without)
template <int NV, bool IR, int MODE>
static int launch_stage2p(pip_ctx* c, const StageArgs& a) {
  if (c->tx == 16 && c->ty == 15) return launch_stage2p_t<NV, IR, 16, 15, 256, MODE>(c, a);
  if (c->tx == 32 && c->ty == 7) return launch_stage2p_t<NV, IR, 32, 7, 256, MODE>(c, a);
  if (c->tx == 16 && c->ty == 7) return launch_stage2p_t<NV, IR, 16, 7, 128, MODE>(c, a);
  return fail(PIP_ERR_ARG, "the two-pass stage kernel takes tiles 16x15, 32x7 and 16x7");
}

template <int NDIM, int NV, bool IR>
static int launch_stage_ext(pip_ctx* c, const StageArgs& a) {
  const bool ext = col_type(c->ph) >= 1 || c->cfg.flag_pip_imp == 1;
  if constexpr (NDIM == 2) {
    if (((c->two_pass >> (a.istep - 1)) & 1) && !ext) {
      if (!a.dtc && !a.Vout) return launch_stage2p<NV, IR, STAGE_PLAIN>(c, a);
      if constexpr (NV == 9 && !IR) {
        if (a.dtc) return launch_stage2p<9, false, STAGE_REPLAY>(c, a);
      }
    }
  }
  if (a.dtc || a.Vout) {  // frozen-dt replay and saving kernels: the adjoint configuration only
    if constexpr (NDIM == 2 && NV == 9 && !IR) {
      if (!ext && a.dtc) return launch_stage_tile<2, 9, false, false, STAGE_REPLAY>(c, a);
      if (!ext) return launch_stage_tile<2, 9, false, false, STAGE_SAVE>(c, a);
    }
    return fail(PIP_ERR_UNSUPPORTED, "replay and saving kernels cover 2-D, flag_divb = 1, flag_IR = 0, explicit collisions");
  }
  return ext ? launch_stage_tile<NDIM, NV, IR, true>(c, a) : launch_stage_tile<NDIM, NV, IR, false>(c, a);
}

static int launch_stage(pip_ctx* c, const StageArgs& a) {
  const bool ir = c->cfg.flag_ir >= 2;
  if (c->arith == PIP_ARITH_RELAXED && !a.dtc) {  // frozen-dt replays stay exact (adjoint)
    const int tx = (c->cfg.ndim == 1) ? 255 : c->tx, ty = (c->cfg.ndim == 1) ? 1 : c->ty;
    const char* err = "";
    const int rc = pip_relaxed_stage(&a, sizeof(a), c->cfg.ndim, c->nv, ir ? 1 : 0, tx, ty, (void*)c->stream, &err);
    if (rc == -1) return fail(PIP_ERR_UNSUPPORTED, "relaxed arithmetic: tiles 16x15, 32x7 and 16x7 (2-D)");
    if (rc == -2) return fail(PIP_ERR_STATE, "relaxed arithmetic: stage argument block mismatch");
    if (rc != 0) return fail(PIP_ERR_HIP, std::string("relaxed stage kernel: ") + err);
    return PIP_OK;
  }
  if (c->cfg.ndim == 1) {
    if (c->nv != 8) return fail(PIP_ERR_UNSUPPORTED, "1-D runs have nvar_m = 8");
    return ir ? launch_stage_ext<1, 8, true>(c, a) : launch_stage_ext<1, 8, false>(c, a);
  }
  if (c->nv == 9)
    return ir ? launch_stage_ext<2, 9, true>(c, a) : launch_stage_ext<2, 9, false>(c, a);
  return ir ? launch_stage_ext<2, 8, true>(c, a) : launch_stage_ext<2, 8, false>(c, a);
}

static RoundArgs round_args(pip_ctx* c, int do_ir, int init_arb, int do_cfl, int read_rates,
                            const double* ion_pot, int do_col = 0) {
  RoundArgs r;
  r.U = c->U;
  const bool col_field = col_type(c->ph) >= 1;
  r.ac = (c->cfg.ac_field || (col_field && !do_col)) ? c->ac : nullptr;
  r.do_col = do_col;
  r.ac_out = col_field ? c->ac : nullptr;
  r.gm_rec = c->gm_rec;
  r.gm_ion = c->gm_ion;
  r.arb_heat = c->arb_heat;
  r.ion_pot = ion_pot;
  r.sc = c->sc;
  r.ix = c->cfg.ix;
  r.jx = c->cfg.jx;
  r.nx = c->nx;
  r.ny = c->ny;
  r.N = c->N;
  r.do_ir = do_ir;
  r.init_arb = init_arb;
  r.do_cfl = do_cfl;
  r.read_rates = read_rates;
  r.col = c->cfg.col;
  r.gmin_safety = c->gmin_safety;
  r.ph = c->ph;
  r.irc = c->irc;
  return r;
}

static int launch_round(pip_ctx* c, const RoundArgs& r) {
  dim3 grid((unsigned)((c->N + 255) / 256));
  const bool ir = c->cfg.flag_ir >= 1;
  if (c->nv == 9) {
    if (ir) hipLaunchKernelGGL((k_round<9, true>), grid, dim3(256), 0, c->stream, r);
    else hipLaunchKernelGGL((k_round<9, false>), grid, dim3(256), 0, c->stream, r);
  } else {
    if (ir) hipLaunchKernelGGL((k_round<8, true>), grid, dim3(256), 0, c->stream, r);
    else hipLaunchKernelGGL((k_round<8, false>), grid, dim3(256), 0, c->stream, r);
  }
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

// hist / hist_k (mode 1): record the next step's {dt, cmax} at hist_k.
static int launch_finalize(pip_ctx* c, int mode, int into_global, double* hist = nullptr, int hist_k = 0) {
  FinArgs f;
  f.sc = c->sc;
  f.hist = hist;
  f.hist_k = hist_k;
  f.mode = mode;
  f.into_global = into_global;
  f.flag_ir = c->cfg.flag_ir;
  f.flag_ir_type = c->cfg.flag_ir_type;
  f.flag_divb = c->cfg.flag_divb;
  f.safety = c->cfg.safety;
  f.dxyz_min = c->dxyz_min;
  f.pip_imp = c->cfg.flag_pip_imp;
  f.pip_imp_factor = c->cfg.pip_imp_factor;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, c->stream, f);
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

static int launch_bnd(pip_ctx* c, double* U) {
  dim3 grid((unsigned)((c->N + 255) / 256));
  hipLaunchKernelGGL(k_bnd, grid, dim3(256), 0, c->stream, U, c->nvt, c->cfg.ix, c->cfg.jx, c->N, c->bnd,
                     c->cfg.ndim);
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

static int set_loop(pip_ctx* c, double tend, long long nmax, int active, int ion_valid) {
  hipLaunchKernelGGL(k_set_loop, dim3(1), dim3(64), 0, c->stream, c->sc, tend, nmax, active, ion_valid);
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

#define TRY(x)            \
  do {                    \
    int rc_ = (x);        \
    if (rc_ != PIP_OK) return rc_; \
  } while (0)

// ------------------------------------------------------------------ C-ABI
extern "C" {

const char* pip_last_error(void) { return g_err.c_str(); }

int pip_check_arith(const double* n, const double* d, double* q, double* r, long long count) {
  if (count < 0 || count > (1LL << 31) || (count > 0 && !(n && d && q && r)))
    return fail(PIP_ERR_ARG, "pip_check_arith: bad arguments (0 <= count <= 2^31, non-null arrays)");
  if (count == 0) return PIP_OK;
  const size_t b = (size_t)count * sizeof(double);
  double* buf = nullptr;
  HIPCHK(hipMalloc(&buf, 4 * b));
  double *dn = buf, *dd = buf + count, *dq = buf + 2 * count, *dr = buf + 3 * count;
  hipError_t e = hipMemcpy(dn, n, b, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dd, d, b, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const unsigned nb = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL(k_check_arith, dim3(nb), dim3(256), 0, 0, dn, dd, dq, dr, count);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(q, dq, b, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(r, dr, b, hipMemcpyDeviceToHost);
  const hipError_t ef = hipFree(buf);
  if (e != hipSuccess) return fail(PIP_ERR_HIP, std::string("pip_check_arith: ") + hipGetErrorString(e));
  if (ef != hipSuccess) return fail(PIP_ERR_HIP, std::string("pip_check_arith: ") + hipGetErrorString(ef));
  return PIP_OK;
}
const char* pip_version(void) { return "pipgpu 0.1 (gfx950, fp64)"; }

int pip_create(const pip_cfg* cfg, pip_ctx** out) {
  if (!cfg || !out) return fail(PIP_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->ndim != 1 && cfg->ndim != 2) return fail(PIP_ERR_ARG, "ndim must be 1 or 2");
  if (cfg->ix < 8) return fail(PIP_ERR_ARG, "ix must be >= 8 (nx >= 4)");
  if (cfg->ndim == 1 && cfg->jx != 1) return fail(PIP_ERR_ARG, "jx must be 1 in 1-D");
  if (cfg->ndim == 2 && cfg->jx < 8) return fail(PIP_ERR_ARG, "jx must be >= 8 (ny >= 4)");
  if (cfg->ix > kMaxPad || cfg->jx > kMaxPad) return fail(PIP_ERR_ARG, "ix and jx must be <= 32767");
  int want_nv = 8 + ((cfg->ndim == 2 && cfg->flag_divb == 1) ? 1 : 0);
  if (cfg->nvar_m != want_nv) return fail(PIP_ERR_ARG, "nvar_m inconsistent with ndim/flag_divb");
  if (cfg->t_order != 2 && cfg->t_order != 3) return fail(PIP_ERR_UNSUPPORTED, "t_order must be 2 or 3");
  if (cfg->flag_ir != 0 && cfg->flag_ir != 3)
    return fail(PIP_ERR_UNSUPPORTED, "flag_IR must be 0 or 3 (set_IR case 3)");
  if (cfg->flag_ir_type < 0 || cfg->flag_ir_type > 2) return fail(PIP_ERR_ARG, "flag_IR_type in 0..2");
  if (cfg->flag_col < 1 || cfg->flag_col > 7) return fail(PIP_ERR_UNSUPPORTED, "flag_col must be 1..7");
  if (cfg->flag_col >= 2 && cfg->ac_field)
    return fail(PIP_ERR_ARG, "flag_col 2..7 computes alpha_c; ac_field needs flag_col = 1");
  if (cfg->flag_pip_imp != 0 && cfg->flag_pip_imp != 1) return fail(PIP_ERR_ARG, "flag_pip_imp must be 0 or 1");
  if (cfg->flag_pip_imp == 1 && !(cfg->pip_imp_factor > 0)) return fail(PIP_ERR_ARG, "pip_imp_factor must be > 0");
  for (int s = 0; s < 2 * cfg->ndim; s++) {
    int t = bnd_kind(cfg->flag_bnd[s]);
    if (t != 1 && t != 2 && t != 10) return fail(PIP_ERR_UNSUPPORTED, "boundary flag must be 1, 2-7, 10 or 11");
  }
  if (!(cfg->dx > 0) || (cfg->ndim == 2 && !(cfg->dy > 0))) return fail(PIP_ERR_ARG, "dx, dy must be > 0");
  pip_ctx* c = new pip_ctx();
  c->cfg = *cfg;
  c->nv = cfg->nvar_m;
  c->nvt = c->nv + 5;
  c->nx = cfg->ix - 4;
  c->ny = (cfg->ndim == 2) ? cfg->jx - 4 : 1;
  c->N = (long long)cfg->ix * cfg->jx;
  c->ph.gm = cfg->gm;
  c->ph.ro_lim = cfg->ro_lim;
  c->ph.pr_lim = cfg->pr_lim;
  c->ph.safety = cfg->safety;
  c->ph.T0 = cfg->T0;
  c->ph.n0 = cfg->n0;
  c->ph.beta = cfg->beta;
  c->ph.t_ir = cfg->t_ir;
  c->ph.flag_col = cfg->flag_col;
  c->ph.flag_ir = cfg->flag_ir;
  c->ph.flag_ir_type = cfg->flag_ir_type;
  c->ph.pip_imp = cfg->flag_pip_imp;
  c->irc = ir_constants(c->ph);
  double dy = (cfg->ndim >= 2) ? cfg->dy : 100.0, dz = 100.0;  // Model_rot.f90:424-429
  c->dxyz_min = dmin(dmin(cfg->dx, dy), dz);
  c->gmin_safety = cfg->safety * c->dxyz_min;
  build_bnd(c);
  hipError_t e;
  size_t bytes = sizeof(double) * (size_t)c->N;
  auto cleanup = [&](const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    pip_destroy(c);
    return fail(PIP_ERR_HIP, m);
  };
  if ((e = hipGetDevice(&c->device)) != hipSuccess) return cleanup("hipGetDevice");
  init_scratch(c->device);
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return cleanup("hipStreamCreate");
  if ((e = hipMalloc(&c->U, bytes * c->nvt)) != hipSuccess) return cleanup("hipMalloc U");
  if ((e = hipMalloc(&c->U1, bytes * c->nvt)) != hipSuccess) return cleanup("hipMalloc U1");
  if (cfg->t_order >= 3 && (e = hipMalloc(&c->U2, bytes * c->nvt)) != hipSuccess) return cleanup("hipMalloc U2");
  if ((cfg->ac_field || col_type(c->ph) >= 1) && (e = hipMalloc(&c->ac, bytes)) != hipSuccess)
    return cleanup("hipMalloc ac");
  if (cfg->flag_pip_imp == 1 && (e = hipMalloc(&c->Ub, bytes * c->nvt)) != hipSuccess) return cleanup("hipMalloc Ub");
  if (cfg->flag_ir) {
    if ((e = hipMalloc(&c->gm_rec, bytes)) != hipSuccess) return cleanup("hipMalloc gm_rec");
    if ((e = hipMalloc(&c->gm_ion, bytes)) != hipSuccess) return cleanup("hipMalloc gm_ion");
    if ((e = hipMalloc(&c->arb_heat, bytes)) != hipSuccess) return cleanup("hipMalloc arb_heat");
    if ((e = hipMalloc(&c->ion_pot, bytes)) != hipSuccess) return cleanup("hipMalloc ion_pot");
    if ((e = hipMemset(c->arb_heat, 0, bytes)) != hipSuccess) return cleanup("hipMemset");
    if ((e = hipMemset(c->ion_pot, 0, bytes)) != hipSuccess) return cleanup("hipMemset");
  }
  if ((e = hipMalloc(&c->sc, sizeof(DevScalars))) != hipSuccess) return cleanup("hipMalloc scalars");
  DevScalars hs = host_scalars_init();
  if ((e = hipMemcpy(c->sc, &hs, sizeof(hs), hipMemcpyHostToDevice)) != hipSuccess) return cleanup("hipMemcpy");
  if ((e = hipMemset(c->U, 0, bytes * c->nvt)) != hipSuccess) return cleanup("hipMemset");
  if ((e = hipMemset(c->U1, 0, bytes * c->nvt)) != hipSuccess) return cleanup("hipMemset");
  *out = c;
  return PIP_OK;
}

int pip_destroy(pip_ctx* c) {
  if (!c) return PIP_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->U);
  (void)hipFree(c->U1);
  (void)hipFree(c->U2);
  (void)hipFree(c->ac);
  (void)hipFree(c->gm_rec);
  (void)hipFree(c->gm_ion);
  (void)hipFree(c->arb_heat);
  (void)hipFree(c->ion_pot);
  (void)hipFree(c->Ub);
  (void)hipFree(c->sc);
  (void)hipFree(c->adjV);
  (void)hipFree(c->adjP);
  (void)hipFree(c->lamB);
  if (c->sgraph) (void)hipGraphExecDestroy(c->sgraph);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PIP_OK;
}

int pip_upload(pip_ctx* c, const double* U_h, const double* U_m, const double* ac) {
  if (!c || !U_h || !U_m) return fail(PIP_ERR_ARG, "null argument");
  size_t pb = sizeof(double) * (size_t)c->N;
  HIPCHK(hipMemcpyAsync(c->U, U_m, pb * c->nv, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->U + (size_t)c->nv * c->N, U_h, pb * 5, hipMemcpyHostToDevice, c->stream));
  if (c->cfg.ac_field) {
    if (!ac) return fail(PIP_ERR_ARG, "ac_field=1 needs the ac array");
    HIPCHK(hipMemcpyAsync(c->ac, ac, pb, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->top_pending = true;
  return PIP_OK;
}

int pip_download(pip_ctx* c, double* U_h, double* U_m) {
  if (!c || !U_h || !U_m) return fail(PIP_ERR_ARG, "null argument");
  size_t pb = sizeof(double) * (size_t)c->N;
  HIPCHK(hipMemcpyAsync(U_m, c->U, pb * c->nv, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(U_h, c->U + (size_t)c->nv * c->N, pb * 5, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PIP_OK;
}

int pip_bnd(pip_ctx* c) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  TRY(launch_bnd(c, c->U));
  return PIP_OK;
}

static int read_dt_out(pip_ctx* c, double* dt) {
  DevScalars hs;
  HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (dt) *dt = hs.dt_out;
  return PIP_OK;
}

int pip_set_ir(pip_ctx* c) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (c->cfg.flag_ir == 0) return PIP_OK;
  TRY(launch_round(c, round_args(c, 1, c->arb_done ? 0 : 1, 0, 0, nullptr)));
  c->arb_done = true;
  return PIP_OK;
}

int pip_cfl(pip_ctx* c, double* dt, int into_global) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  TRY(launch_round(c, round_args(c, 0, 0, 1, c->cfg.flag_ir ? 1 : 0, c->ion_pot)));
  TRY(launch_finalize(c, 0, into_global));
  return read_dt_out(c, dt);
}

int pip_prologue(pip_ctx* c, double* dt) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  DevScalars hs = host_scalars_init();
  HIPCHK(hipMemcpyAsync(c->sc, &hs, sizeof(hs), hipMemcpyHostToDevice, c->stream));
  // set_IR fixes arb_heat on its first call of a fresh run only
  // (PIP_rot.f90:219-229: nout = 0 and no restart)
  c->arb_done = c->arb_given;
  c->arb_given = false;
  TRY(launch_bnd(c, c->U));        // boundary map of the initial state
  // set_coefficients(U_h,U_m,0) (Initial_rot.f90:140): set_collisional, set_IR
  if (col_type(c->ph) >= 1) TRY(launch_round(c, round_args(c, 0, 0, 0, 0, nullptr, 1)));
  TRY(pip_set_ir(c));
  TRY(pip_cfl(c, dt, 1));          // cfl(U_h,U_m,dt) (Initial_rot.f90:152)
  c->top_pending = true;
  return PIP_OK;
}

// Stage s reads U (s=1), U1 (s=2) or U2 (s=3) and writes U1, U2 or, for the
// last stage, U in place (each cell reads U0 only at itself).
static int run_stages(pip_ctx* c, int epi_last, double* ion_out) {
  const int T = c->cfg.t_order;
  for (int s = 1; s <= T; s++) {
    const double* Uin = (s == 1) ? c->U : ((s == 2) ? c->U1 : c->U2);
    double* Uout = (s == T) ? c->U : ((s == 1) ? c->U1 : c->U2);
    int epi = (s == T) ? epi_last : 0;
    StageArgs a = stage_args(c, s, Uin, c->U, Uout, epi, ion_out);
    TRY(launch_stage(c, a));
  }
  return PIP_OK;
}

int pip_run_solver(pip_ctx* c) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  TRY(set_loop(c, 1.0e300, (long long)1 << 62, 1, -1));
  TRY(run_stages(c, 0, c->ion_pot));
  TRY(set_loop(c, 1.0e300, (long long)1 << 62, -1, 1));  // ion_pot now defined
  return PIP_OK;
}

static int loop_top(pip_ctx* c, double tend, long long nmax) {
  // first pass of main.f90:32-36 (set_IR, cfl(dt)) when no fused step has
  // already applied it as its look-ahead round
  TRY(set_loop(c, tend, nmax, 1, -1));
  if (c->cfg.flag_ir) TRY(launch_round(c, round_args(c, 1, c->arb_done ? 0 : 1, 0, 0, nullptr)));
  c->arb_done = true;
  TRY(launch_round(c, round_args(c, 0, 0, 1, c->cfg.flag_ir ? 1 : 0, c->ion_pot)));
  TRY(launch_finalize(c, 2, 1));
  c->top_pending = false;
  return PIP_OK;
}

static int enqueue_fused(pip_ctx* c, int n) {
  for (int k = 0; k < n; k++) {
    TRY(run_stages(c, 2, nullptr));
    TRY(launch_finalize(c, 1, 0));
  }
  return PIP_OK;
}

// The stage-kernel choice a captured chunk depends on.
static long long step_graph_key(const pip_ctx* c) {
  return ((((long long)c->tx * 64 + c->ty) * 2 + c->xcd_remap) * 8 + c->two_pass) * 2 + c->arith;
}

// One chunk of `n` fused steps from the captured graph (captured on first
// use, after a chunk with the same kernel choice has run from the stream so
// that no kernel attribute is set during capture, and again when the choice
// changed); the kernels predicate on the device's active flag, so the chunk
// is the same work as enqueue_fused(c, n).
static int launch_step_graph(pip_ctx* c, int n) {
  const long long key = step_graph_key(c);
  if (!c->sgraph || c->sgraph_key != key) {
    if (c->sgraph) {
      HIPCHK(hipGraphExecDestroy(c->sgraph));
      c->sgraph = nullptr;
    }
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_fused(c, n);
    const hipError_t ce = hipStreamEndCapture(c->stream, &g);
    if (rc != PIP_OK) {
      if (ce == hipSuccess) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ce != hipSuccess) return fail(PIP_ERR_HIP, std::string("step graph capture: ") + hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&c->sgraph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      c->sgraph = nullptr;
      return fail(PIP_ERR_HIP, std::string("step graph instantiate: ") + hipGetErrorString(ie));
    }
    c->sgraph_key = key;
  }
  HIPCHK(hipGraphLaunch(c->sgraph, c->stream));
  return PIP_OK;
}

int pip_advance(pip_ctx* c, int max_steps, double tend, int* steps_done, double* time) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (max_steps < 0) return fail(PIP_ERR_ARG, "max_steps < 0");
  DevScalars hs;
  HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  long long nt0 = hs.nt;
  long long nmax = nt0 + max_steps;
  TRY(loop_top(c, tend, nmax));
  const int chunk = 64;
  int left = max_steps;
  while (left > 0) {
    int n = left < chunk ? left : chunk;
    if (n == chunk && c->step_graph && c->eager_key == step_graph_key(c)) {
      TRY(launch_step_graph(c, n));
    } else {
      TRY(enqueue_fused(c, n));
      c->eager_key = step_graph_key(c);
    }
    left -= n;
    HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!hs.active) break;
  }
  c->top_pending = true;
  if (steps_done) *steps_done = (int)(hs.nt - nt0);
  if (time) *time = hs.time;
  if (hs.status) g_err = (hs.status & 2) ? "NaN in U_m" : "dt too small";
  return PIP_OK;
}

int pip_enqueue_steps(pip_ctx* c, int n, double tend) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (c->top_pending) {
    DevScalars hs;
    HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    TRY(loop_top(c, tend, (long long)1 << 62));
  }
  return enqueue_fused(c, n);
}

int pip_sync(pip_ctx* c) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return PIP_OK;
}

int pip_get_scalars(pip_ctx* c, double* dt, double* time, double* cmax, long* nt) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  DevScalars hs;
  HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (dt) *dt = hs.dt;
  if (time) *time = hs.time;
  if (cmax) *cmax = hs.cmax;
  if (nt) *nt = (long)hs.nt;
  return PIP_OK;
}

int pip_set_time(pip_ctx* c, double time, double dt) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  DevScalars hs;
  HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  hs.time = time;
  hs.dt = dt;
  if (c->cfg.flag_divb >= 1) hs.cmax = c->cfg.safety * c->dxyz_min / dt;
  HIPCHK(hipMemcpyAsync(c->sc, &hs, sizeof(hs), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PIP_OK;
}

int pip_status(pip_ctx* c, int* flags) {
  if (!c || !flags) return fail(PIP_ERR_ARG, "null argument");
  DevScalars hs;
  HIPCHK(hipMemcpyAsync(&hs, c->sc, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *flags = hs.status;
  return PIP_OK;
}

int pip_get_rates(pip_ctx* c, double* gm_rec, double* gm_ion, double* arb_heat) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (!c->cfg.flag_ir) return fail(PIP_ERR_STATE, "flag_IR = 0: no rate fields");
  size_t pb = sizeof(double) * (size_t)c->N;
  if (gm_rec) HIPCHK(hipMemcpyAsync(gm_rec, c->gm_rec, pb, hipMemcpyDeviceToHost, c->stream));
  if (gm_ion) HIPCHK(hipMemcpyAsync(gm_ion, c->gm_ion, pb, hipMemcpyDeviceToHost, c->stream));
  if (arb_heat) HIPCHK(hipMemcpyAsync(arb_heat, c->arb_heat, pb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PIP_OK;
}

int pip_set_arb_heat(pip_ctx* c, const double* arb_heat) {
  if (!c || !arb_heat) return fail(PIP_ERR_ARG, "null argument");
  if (!c->arb_heat || c->cfg.flag_ir_type != 0)
    return fail(PIP_ERR_STATE, "no arb_heat field (flag_IR = 0 or flag_IR_type != 0)");
  HIPCHK(hipMemcpyAsync(c->arb_heat, arb_heat, sizeof(double) * (size_t)c->N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->arb_given = true;
  c->arb_done = true;
  return PIP_OK;
}

int pip_get_ac(pip_ctx* c, double* ac) {
  if (!c || !ac) return fail(PIP_ERR_ARG, "null argument");
  if (c->ac) {
    HIPCHK(hipMemcpyAsync(ac, c->ac, sizeof(double) * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {  // flag_col = 1 without a per-cell field: ac = col everywhere (Solver_rot.f90:105)
    for (long long k = 0; k < c->N; k++) ac[k] = c->cfg.col;
  }
  return PIP_OK;
}

int pip_device_state(pip_ctx* c, double** U, double** ac) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (U) *U = c->U;
  if (ac) *ac = c->ac;
  return PIP_OK;
}

int pip_get_stream(pip_ctx* c, void** stream) {
  if (!c || !stream) return fail(PIP_ERR_ARG, "null argument");
  *stream = (void*)c->stream;
  return PIP_OK;
}

int pip_set_tile(pip_ctx* c, int tx, int ty) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  const int ok[][2] = {{16, 15}, {32, 7}, {32, 15}, {64, 7}, {32, 8}, {16, 7}};
  for (auto& t : ok)
    if (t[0] == tx && t[1] == ty) {
      c->tx = tx;
      c->ty = ty;
      return PIP_OK;
    }
  return fail(PIP_ERR_ARG, "unsupported tile shape");
}

int pip_set_two_pass(pip_ctx* c, int stage_mask) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (stage_mask < 0 || stage_mask > 7) return fail(PIP_ERR_ARG, "stage mask must be in 0..7");
  c->two_pass = stage_mask;
  return PIP_OK;
}

int pip_set_arith(pip_ctx* c, int mode) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (mode != PIP_ARITH_EXACT && mode != PIP_ARITH_RELAXED) return fail(PIP_ERR_ARG, "arith mode must be 0 or 1");
  if (mode == PIP_ARITH_RELAXED) {
    const bool ext = col_type(c->ph) >= 1 || c->cfg.flag_pip_imp == 1;
    const bool ir = c->cfg.flag_ir >= 2;
    if (ext || !pip_relaxed_covers(c->cfg.ndim, c->nv, ir ? 1 : 0, c->cfg.ndim == 1 ? 255 : 16, c->cfg.ndim == 1 ? 1 : 15))
      return fail(PIP_ERR_UNSUPPORTED,
                  "relaxed arithmetic covers 2-D runs with psi (with or without set_IR) and 1-D runs without "
                  "set_IR, with constant or per-cell alpha_c and explicit collisions");
  }
  c->arith = mode;
  return PIP_OK;
}

int pip_set_step_graph(pip_ctx* c, int on) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  c->step_graph = on ? 1 : 0;
  return PIP_OK;
}

int pip_set_xcd_remap(pip_ctx* c, int on) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  c->xcd_remap = on ? 1 : 0;
  return PIP_OK;
}

int pip_time_steps(pip_ctx* c, int nsteps, double* ms_total, double* ms_stage1, double* ms_stage2) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (c->cfg.t_order != 2) return fail(PIP_ERR_UNSUPPORTED, "timing helper covers t_order = 2");
  if (c->top_pending) TRY(loop_top(c, 1.0e300, (long long)1 << 62));
  const int NE = 3 * nsteps + 1;
  hipEvent_t* ev = new hipEvent_t[NE];
  for (int k = 0; k < NE; k++) HIPCHK(hipEventCreate(&ev[k]));
  HIPCHK(hipEventRecord(ev[0], c->stream));
  int rc = PIP_OK;
  for (int k = 0; k < nsteps && rc == PIP_OK; k++) {
    StageArgs a1 = stage_args(c, 1, c->U, c->U, c->U1, 0, nullptr);
    rc = launch_stage(c, a1);
    if (rc != PIP_OK) break;
    HIPCHK(hipEventRecord(ev[3 * k + 1], c->stream));
    StageArgs a2 = stage_args(c, 2, c->U1, c->U, c->U, 2, nullptr);
    rc = launch_stage(c, a2);
    if (rc != PIP_OK) break;
    HIPCHK(hipEventRecord(ev[3 * k + 2], c->stream));
    rc = launch_finalize(c, 1, 0);
    if (rc != PIP_OK) break;
    HIPCHK(hipEventRecord(ev[3 * k + 3], c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  double t1 = 0, t2 = 0, tt = 0;
  float ms;
  for (int k = 0; k < nsteps && rc == PIP_OK; k++) {
    HIPCHK(hipEventElapsedTime(&ms, ev[3 * k], ev[3 * k + 1]));
    t1 += ms;
    HIPCHK(hipEventElapsedTime(&ms, ev[3 * k + 1], ev[3 * k + 2]));
    t2 += ms;
  }
  HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[NE - 1]));
  tt = ms;
  for (int k = 0; k < NE; k++) (void)hipEventDestroy(ev[k]);
  delete[] ev;
  if (ms_total) *ms_total = tt;
  if (ms_stage1) *ms_stage1 = t1;
  if (ms_stage2) *ms_stage2 = t2;
  c->top_pending = true;
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------- adjoint
static dim3 cell_grid(const pip_ctx* c) { return dim3((unsigned)((c->N + 255) / 256)); }

static int adj_check(const pip_ctx* c) {
  if (c->cfg.ndim != 2 || c->nv != ADJ_NV) return fail(PIP_ERR_UNSUPPORTED, "adjoint needs ndim = 2 and flag_divb = 1");
  if (c->cfg.flag_ir != 0) return fail(PIP_ERR_UNSUPPORTED, "adjoint covers flag_IR = 0");
  if (c->cfg.flag_col != 1 || c->cfg.flag_pip_imp != 0)
    return fail(PIP_ERR_UNSUPPORTED, "adjoint covers flag_col = 1 (constant or per-cell alpha) and explicit collisions");
  if (c->cfg.t_order != 2) return fail(PIP_ERR_UNSUPPORTED, "adjoint covers SSP-RK2 (t_order = 2)");
  for (int s = 0; s < 4; s++)
    if (c->bnd.type[s] == 10) return fail(PIP_ERR_UNSUPPORTED, "adjoint covers periodic and mirror boundaries");
  return PIP_OK;
}

static int adj_alloc(pip_ctx* c) {
  if (c->adjV) return PIP_OK;
  const size_t pb = sizeof(double) * (size_t)c->N * c->nvt;
  HIPCHK(hipMalloc(&c->adjV, pb));
  HIPCHK(hipMalloc(&c->adjP, 4 * pb + 4 * sizeof(double) * (size_t)c->N));
  HIPCHK(hipMalloc(&c->lamB, pb));
  return PIP_OK;
}

static AdjArgs adj_args(pip_ctx* c, int istep, const double* Uin, const double* V, const double* lam_out,
                        double* lam_in, const double* lam_next, double* lam_ac, const double* dtc) {
  AdjArgs a;
  a.Uin = Uin;
  a.lam_out = lam_out;
  a.lam_in = lam_in;
  a.lam_next = lam_next;
  a.lam_ac = lam_ac;
  a.ac = c->cfg.ac_field ? c->ac : nullptr;
  // primitives saved by the replay (read only), or computed by k_adj_cell
  a.V = V ? const_cast<double*>(V) : c->adjV;
  a.make_V = V ? 0 : 1;
  a.P = c->adjP;
  a.dtc = dtc;
  a.ix = c->cfg.ix;
  a.jx = c->cfg.jx;
  a.nx = c->nx;
  a.ny = c->ny;
  a.N = c->N;
  a.istep = istep;
  a.t_order = c->cfg.t_order;
  a.dx = c->cfg.dx;
  a.dy = c->cfg.dy;
  a.col = c->cfg.col;
  a.safety = c->cfg.safety;
  a.db_clean = c->cfg.db_clean;
  a.ph = c->ph;
  a.bnd = c->bnd;
  return a;
}

static int adj_stage(pip_ctx* c, const AdjArgs& a) {
  // the primitives, when the replay did not save them, and (then) the stage-1
  // ghost zeroing of lam_in; otherwise the gather's extra blocks zero them
  HIPCHK(adj_stage_launch(a, c->stream));
  return PIP_OK;
}


extern "C" {

int pip_set_state_device(pip_ctx* c, const double* U) {
  if (!c || !U) return fail(PIP_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(c->U, U, sizeof(double) * (size_t)c->N * c->nvt, hipMemcpyDeviceToDevice, c->stream));
  c->top_pending = true;
  return PIP_OK;
}

int pip_get_state_device(pip_ctx* c, double* U) {
  if (!c || !U) return fail(PIP_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(U, c->U, sizeof(double) * (size_t)c->N * c->nvt, hipMemcpyDeviceToDevice, c->stream));
  return PIP_OK;
}

int pip_set_ac_device(pip_ctx* c, const double* ac) {
  if (!c || !ac) return fail(PIP_ERR_ARG, "null argument");
  if (!c->cfg.ac_field) return fail(PIP_ERR_STATE, "context was created with ac_field = 0");
  HIPCHK(hipMemcpyAsync(c->ac, ac, sizeof(double) * (size_t)c->N, hipMemcpyDeviceToDevice, c->stream));
  return PIP_OK;
}

int pip_record_steps(pip_ctx* c, int n, double* hist) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (n < 0 || (n > 0 && !hist)) return fail(PIP_ERR_ARG, "n < 0 or null history");
  if (c->top_pending) TRY(loop_top(c, 1.0e300, (long long)1 << 62));
  // step 0's {dt, cmax} here, step k+1's by the end of step k
  if (n > 0) hipLaunchKernelGGL(k_record_dt, dim3(1), dim3(64), 0, c->stream, c->sc, hist, 0);
  for (int k = 0; k < n; k++) {
    TRY(run_stages(c, 2, nullptr));
    TRY(launch_finalize(c, 1, 0, (k + 1 < n) ? hist : nullptr, k + 1));
  }
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

int pip_record_steps_saved(pip_ctx* c, int n, double* hist, double* Un, double* U1, double* Vn, double* V1) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (n < 0 || (n > 0 && !hist)) return fail(PIP_ERR_ARG, "n < 0 or null history");
  if (n > 0 && !(Un && U1 && Vn && V1)) return fail(PIP_ERR_ARG, "null save buffer");
  if (!(c->cfg.ndim == 2 && c->nv == 9 && c->cfg.flag_ir == 0 && c->cfg.t_order == 2 && col_type(c->ph) == 0 &&
        c->cfg.flag_pip_imp == 0))
    return fail(PIP_ERR_UNSUPPORTED, "saving steps cover 2-D, flag_divb = 1, flag_IR = 0, flag_col = 1, "
                                     "explicit collisions, SSP-RK2");
  if (c->arith != PIP_ARITH_EXACT) return fail(PIP_ERR_UNSUPPORTED, "saving steps run the exact arithmetic");
  if (c->top_pending) TRY(loop_top(c, 1.0e300, (long long)1 << 62));
  if (n == 0) return PIP_OK;
  // the steps of pip_record_steps with the stage outputs written straight
  // into the save buffers (as pip_replay_steps does) and the primitives of
  // each stage input saved by the stage kernels
  const size_t S = (size_t)c->N * c->nvt;
  HIPCHK(hipMemcpyAsync(Un, c->U, sizeof(double) * S, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_record_dt, dim3(1), dim3(64), 0, c->stream, c->sc, hist, 0);
  for (int k = 0; k < n; k++) {
    double* un = Un + (size_t)k * S;
    double* u1 = U1 + (size_t)k * S;
    double* out = (k + 1 < n) ? Un + (size_t)(k + 1) * S : c->U;
    StageArgs a1 = stage_args(c, 1, un, un, u1, 0, nullptr);
    StageArgs a2 = stage_args(c, 2, u1, un, out, 2, nullptr);
    a1.Vout = Vn + (size_t)k * S;
    a2.Vout = V1 + (size_t)k * S;
    TRY(launch_stage(c, a1));
    TRY(launch_stage(c, a2));
    TRY(launch_finalize(c, 1, 0, (k + 1 < n) ? hist : nullptr, k + 1));
  }
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

int pip_replay_steps(pip_ctx* c, int n, const double* hist, double* Un, double* U1, double* Vn, double* V1) {
  if (!c) return fail(PIP_ERR_ARG, "null context");
  if (n < 0 || (n > 0 && !hist)) return fail(PIP_ERR_ARG, "n < 0 or null history");
  if (c->cfg.t_order != 2 && U1) return fail(PIP_ERR_UNSUPPORTED, "stage saves cover t_order = 2");
  if ((Vn || V1) && !(Un && U1)) return fail(PIP_ERR_ARG, "primitive saves need Un and U1");
  const size_t sb = sizeof(double) * (size_t)c->N * c->nvt;
  TRY(set_loop(c, 1.0e300, (long long)1 << 62, 1, -1));
  c->top_pending = false;
  // direct: stage outputs go straight into the save buffers (step k reads
  // Un[k], writes U1[k] and then Un[k+1]; the last step writes the state).
  // frozen: the replay kernels take the step's {dt, cmax} from hist, so no
  // per-step scalar kernels run (the adjoint configuration).
  const bool direct = Un && U1 && c->cfg.t_order == 2 && c->cfg.flag_pip_imp == 0;
  const bool frozen = direct && c->cfg.ndim == 2 && c->nv == 9 && c->cfg.flag_ir == 0 && col_type(c->ph) == 0;
  if ((Vn || V1) && !frozen)
    return fail(PIP_ERR_UNSUPPORTED, "primitive saves cover 2-D, flag_divb = 1, flag_IR = 0, flag_col = 1, explicit collisions");
  if (direct && n > 0) {
    const size_t S = (size_t)c->N * c->nvt;
    HIPCHK(hipMemcpyAsync(Un, c->U, sb, hipMemcpyDeviceToDevice, c->stream));
    for (int k = 0; k < n; k++) {
      double* un = Un + (size_t)k * S;
      double* u1 = U1 + (size_t)k * S;
      double* out = (k + 1 < n) ? Un + (size_t)(k + 1) * S : c->U;
      StageArgs a1 = stage_args(c, 1, un, un, u1, 0, nullptr);
      StageArgs a2 = stage_args(c, 2, u1, un, out, 2, nullptr);
      if (frozen) {
        a1.dtc = a2.dtc = hist + 2 * k;
        a1.Vout = Vn ? Vn + (size_t)k * S : nullptr;
        a2.Vout = V1 ? V1 + (size_t)k * S : nullptr;
      } else {
        hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(64), 0, c->stream, c->sc, hist, k);
      }
      TRY(launch_stage(c, a1));
      TRY(launch_stage(c, a2));
      if (!frozen) TRY(launch_finalize(c, 1, 0));
    }
    if (frozen) hipLaunchKernelGGL(k_replay_end, dim3(1), dim3(64), 0, c->stream, c->sc, hist, n);
    HIPCHK(hipGetLastError());
    return PIP_OK;
  }
  for (int k = 0; k < n; k++) {
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(64), 0, c->stream, c->sc, hist, k);
    if (Un) HIPCHK(hipMemcpyAsync(Un + (size_t)k * c->N * c->nvt, c->U, sb, hipMemcpyDeviceToDevice, c->stream));
    TRY(run_stages(c, 2, nullptr));
    if (U1) HIPCHK(hipMemcpyAsync(U1 + (size_t)k * c->N * c->nvt, c->U1, sb, hipMemcpyDeviceToDevice, c->stream));
    TRY(launch_finalize(c, 1, 0));
  }
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

int pip_adj_seed(pip_ctx* c, double* lam) {
  if (!c || !lam) return fail(PIP_ERR_ARG, "null argument");
  HIPCHK(adj_seed_launch(lam, c->nvt, c->cfg.ix, c->cfg.jx, c->N, c->bnd, c->stream));
  return PIP_OK;
}

// One adjoint step with the face-state work buffer P and the stage-1
// adjoint lamB given.
static int adj_step_with(pip_ctx* c, const double* Un, const double* U1, const double* Vn, const double* V1,
                         const double* dtc, double* lam, double* lam_ac, double* P, double* lamB) {
  // stage 2: U_{n+1} = c0 U_n + c1 U1 - dt_sub div F(U1) + dt_sub S(U1):
  // lam -> lamB (adjoint of U1)
  AdjArgs a2 = adj_args(c, 2, U1, V1, lam, lamB, nullptr, lam_ac, dtc);
  a2.P = P;
  TRY(adj_stage(c, a2));
  // stage 1: U1 = U_n - dt div F(U_n) + dt S(U_n): lamB and the c0 part of
  // lam (the U0 path of stage 2) -> lam
  AdjArgs a1 = adj_args(c, 1, Un, Vn, lamB, lam, lam, lam_ac, dtc);
  a1.P = P;
  TRY(adj_stage(c, a1));
  return PIP_OK;
}

int pip_adj_step(pip_ctx* c, const double* Un, const double* U1, const double* Vn, const double* V1,
                 const double* dtc, double* lam, double* lam_ac) {
  if (!c || !Un || !U1 || !dtc || !lam) return fail(PIP_ERR_ARG, "null argument");
  if (!Vn != !V1) return fail(PIP_ERR_ARG, "give both saved primitive states or neither");
  TRY(adj_check(c));
  TRY(adj_alloc(c));
  return adj_step_with(c, Un, U1, Vn, V1, dtc, lam, lam_ac, c->adjP, c->lamB);
}

int pip_adj_step_noalloc(pip_ctx* c, const double* Un, const double* U1, const double* Vn, const double* V1,
                         const double* dtc, double* lam, double* lam_ac) {
  if (!c || !Un || !U1 || !Vn || !V1 || !dtc || !lam)
    return fail(PIP_ERR_ARG, "pip_adj_step_noalloc: null argument (Vn and V1 are required)");
  TRY(adj_check(c));
  const size_t need = sizeof(double) * (size_t)(5 * c->nvt + 4) * (size_t)c->N;
  if (need > kAdjScratchBytes)
    return fail(PIP_ERR_ARG, "pip_adj_step_noalloc: the grid needs " + std::to_string(need) +
                                 " bytes of work memory, more than its " + std::to_string(kAdjScratchBytes) +
                                 "-byte device array; use pip_adj_step");
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_done[c->device])
      return fail(PIP_ERR_STATE, "pip_adj_step_noalloc: no work array on this device");
    if (!g_scratch_ptr[c->device])
      return fail(PIP_ERR_HIP, std::string("pip_adj_step_noalloc: hipGetSymbolAddress: ") +
                                   hipGetErrorString(g_scratch_err[c->device]));
    base = g_scratch_ptr[c->device];
  }
  // P: 2 directions x nvt planes of (left, right) pairs, then 2 directions
  // of code pairs; lamB after it
  double* P = base;
  double* lamB = base + (size_t)(4 * c->nvt + 4) * (size_t)c->N;
  return adj_step_with(c, Un, U1, Vn, V1, dtc, lam, lam_ac, P, lamB);
}

}  // extern "C"

// ---------------------------------------------------------------- ensemble
struct pip_ens {
  pip_cfg base;
  int n;
  bool twod = false;  // 2-D members (psi): lockstep modes on 16x15 tiles
  int nvt = 13;       // planes per member: nvar_m + 5
  std::vector<EnsMember> hmem;
  EnsMember* mem = nullptr;
  EnsOut* out = nullptr;
  int* order = nullptr;
  unsigned* counter = nullptr;
  double* U = nullptr;
  double* U1 = nullptr;
  double* Uinit = nullptr;
  long long total = 0;
  hipStream_t stream = nullptr;
  int tx = 256;
  Phys ph;
  Bnd bnd;
  hipGraphExec_t gexec = nullptr;
  int graph_tx = 0;
  int mode = 0;                     // 0: persistent launch, 1: lockstep launches
  DevScalars* scs = nullptr;        // lockstep: per-member scalars
  LockTile* tiles = nullptr;        // lockstep: (member, tile) table of the running members
  LockTile* h_tiles = nullptr;      // pinned staging of the table
  DevScalars* h_scs = nullptr;      // pinned copy of the scalars (which members still run)
  int* list = nullptr;              // lockstep: members of the current table
  int* h_list = nullptr;            // pinned staging of the list
  int max_tiles = 0;
  // graph-captured lockstep (mode 2): a 64-step chunk over the tile table of
  // a set of members, captured once per set and replayed; the set shrinks to
  // the members still running when half of it has stopped
  struct LockGraph {
    std::vector<char> run;          // the members the table lists
    hipGraphExec_t exec = nullptr;
    LockTile* tiles = nullptr;      // this graph's own table and member list
    int* list = nullptr;
    int nm = 0;
  };
  std::vector<LockGraph> lgraphs;
  int lgraph_key = 0;               // kernel choice the cached graphs were built for
  int* d_active = nullptr;          // members still running after the chunk
  int* h_active = nullptr;          // pinned copy (written by the graph)
};

static void ens_drop_graphs(pip_ens* e) {
  for (auto& g : e->lgraphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    (void)hipFree(g.tiles);
    (void)hipFree(g.list);
  }
  e->lgraphs.clear();
}

template <int TX>
static int ens_launch(pip_ens* e) {
  using G = TileGeom<1, 8, TX - 1, 1>;
  size_t lds = sizeof(double) * (size_t)(G::LDS_DOUBLES + 3 * (TX / 64));
  static bool attr = false;
  auto kern = k_ensemble<TX>;
  if (!attr) {
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  int dev = 0, ncu = 0, nb = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)kern, TX, lds));
  if (nb < 1) return fail(PIP_ERR_HIP, "ensemble kernel does not fit on a CU");
  int grid = ncu * nb;
  if (grid > e->n) grid = e->n;
  EnsArgs a;
  a.U = e->U;
  a.U1 = e->U1;
  a.mem = e->mem;
  a.out = e->out;
  a.order = e->order;
  a.counter = e->counter;
  a.n_members = e->n;
  a.ph = e->ph;
  a.bnd = e->bnd;
  a.db_clean = e->base.db_clean;
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, e->stream, e->counter);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(TX), lds, e->stream, a);
  HIPCHK(hipGetLastError());
  return PIP_OK;
}

// Lockstep drivers.  Members' tiles are listed in a (member, tile) table:
// 1-D members in tiles of tc cells, 2-D members in 16x15 tiles.
static int ens_table(pip_ens* e, const std::vector<char>& run, int tc, int* nm_out) {
  int nt = 0, nm = 0;
  for (int m = 0; m < e->n; m++) {
    if (!run[m]) continue;
    e->h_list[nm++] = m;
    const EnsMember& M = e->hmem[m];
    const int ty = e->twod ? kEns2TY : 1, tx = e->twod ? kEns2TX : tc, ny = e->twod ? M.ny : 1;
    for (int j0 = 0; j0 < ny; j0 += ty)
      for (int i0 = 0; i0 < M.nx; i0 += tx) {
        if (nt >= e->max_tiles) return -1;
        e->h_tiles[nt++] = LockTile{m, i0, j0, 0};
      }
  }
  *nm_out = nm;
  return nt;
}

static EnsArgs ens_args(pip_ens* e) {
  EnsArgs a;
  a.U = e->U;
  a.U1 = e->U1;
  a.mem = e->mem;
  a.out = e->out;
  a.order = e->order;
  a.counter = e->counter;
  a.n_members = e->n;
  a.ph = e->ph;
  a.bnd = e->bnd;
  a.db_clean = e->base.db_clean;
  return a;
}

// Kernels of one kind of member: start(a), stage(a, table, istep, ntiles, p,
// apply_end).
struct EnsKernels {
  std::function<void(const EnsArgs&)> start;
  std::function<void(const EnsArgs&, const LockTile*, int, int, int, int)> stage;
};

// Host-driven lockstep: start kernel, then chunks of steps over the tile
// table of the members still running, each step two stage launches (stage 1
// applies the previous step's end), each chunk closed by k_lock_end over its
// members; the table is rebuilt between chunks from the members' active
// flags.  A member once stopped is never listed again (its other parity
// buffer may still hold its last running state).
static int ens_lockstep_run(pip_ens* e, int tc, const EnsKernels& K) {
  const EnsArgs a = ens_args(e);
  K.start(a);
  HIPCHK(hipGetLastError());
  std::vector<char> run(e->n, 1);  // members that may still run (all, until the start kernel has said)
  constexpr int chunk = 64;        // even: a chunk ends on the parity it started with
  long long k = 0;                 // global step index (parity of the scalar buffers)
  for (;;) {
    int nm = 0;
    const int nt = ens_table(e, run, tc, &nm);
    if (nt < 0) return fail(PIP_ERR_STATE, "lockstep tile table overflow");
    if (nt == 0) break;
    HIPCHK(hipMemcpyAsync(e->tiles, e->h_tiles, sizeof(LockTile) * (size_t)nt, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->list, e->h_list, sizeof(int) * (size_t)nm, hipMemcpyHostToDevice, e->stream));
    for (int s = 0; s < chunk; s++) {
      const int p = (int)((k + s) & 1);
      K.stage(a, e->tiles, 1, nt, p, s > 0 ? 1 : 0);
      K.stage(a, e->tiles, 2, nt, p, 0);
    }
    const int pl = (int)((k + chunk - 1) & 1);
    hipLaunchKernelGGL(k_lock_end, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, e->stream, a, e->scs, e->list,
                       nm, pl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(e->h_scs, e->scs + (size_t)(pl ^ 1) * e->n, sizeof(DevScalars) * (size_t)e->n,
                          hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    k += chunk;
    for (int q = 0; q < nm; q++) {
      const int m = e->h_list[q];
      run[m] = (char)(e->h_scs[m].active != 0);
    }
  }
  return PIP_OK;
}

// Graph-captured lockstep: the start kernel, then replays of a captured
// chunk -- kLockChunk steps of the two stage launches over the tile table of
// a set of members (members of the set that have stopped return at once:
// their scalars say inactive), the chunk's closing step end over the set,
// the count of all members still running and its copy to pinned host
// memory.  The host replays kGraphBatch chunks between looks at the count.
// When half of the set has stopped, the set shrinks to the members still
// running: their graph is taken from the cache or captured (with its own
// table), so a rerun of the same ensemble captures nothing.  Same kernels and
// order as lockstep mode, so the same results.
constexpr int kLockChunk = 64;  // even: a chunk ends on the parity it started with
constexpr int kGraphBatch = 4;
constexpr int kGraphCacheMax = 64;

static int ens_lockgraph_get(pip_ens* e, const EnsArgs& a, int tc, const EnsKernels& K,
                             const std::vector<char>& run, int* idx) {
  for (size_t g = 0; g < e->lgraphs.size(); g++)
    if (e->lgraphs[g].run == run) {
      *idx = (int)g;
      return PIP_OK;
    }
  if ((int)e->lgraphs.size() >= kGraphCacheMax) {
    HIPCHK(hipStreamSynchronize(e->stream));
    ens_drop_graphs(e);
  }
  int nm = 0;
  const int nt = ens_table(e, run, tc, &nm);
  if (nt < 0) return fail(PIP_ERR_STATE, "lockstep tile table overflow");
  pip_ens::LockGraph G;
  G.run = run;
  G.nm = nm;
  if (nt > 0) {
    HIPCHK(hipMalloc(&G.tiles, sizeof(LockTile) * (size_t)nt));
    HIPCHK(hipMalloc(&G.list, sizeof(int) * (size_t)nm));
    HIPCHK(hipMemcpyAsync(G.tiles, e->h_tiles, sizeof(LockTile) * (size_t)nt, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(G.list, e->h_list, sizeof(int) * (size_t)nm, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  if (nt > 0) {
    for (int s = 0; s < kLockChunk; s++) {
      const int p = s & 1;
      K.stage(a, G.tiles, 1, nt, p, s > 0 ? 1 : 0);
      K.stage(a, G.tiles, 2, nt, p, 0);
    }
    hipLaunchKernelGGL(k_lock_end, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, e->stream, a, e->scs, G.list,
                       nm, (kLockChunk - 1) & 1);
  }
  hipLaunchKernelGGL(k_lock_count, dim3(1), dim3(256), 0, e->stream, e->scs, e->n, e->d_active);
  const hipError_t le = hipGetLastError();
  const hipError_t me = hipMemcpyAsync(e->h_active, e->d_active, sizeof(int), hipMemcpyDeviceToHost, e->stream);
  const hipError_t ce = hipStreamEndCapture(e->stream, &g);
  hipError_t bad = le != hipSuccess ? le : (me != hipSuccess ? me : ce);
  if (bad == hipSuccess) {
    bad = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
    if (bad != hipSuccess) G.exec = nullptr;
  }
  if (ce == hipSuccess) (void)hipGraphDestroy(g);
  if (bad != hipSuccess) {
    (void)hipFree(G.tiles);
    (void)hipFree(G.list);
    return fail(PIP_ERR_HIP, std::string("lockstep graph capture: ") + hipGetErrorString(bad));
  }
  e->lgraphs.push_back(std::move(G));
  *idx = (int)e->lgraphs.size() - 1;
  return PIP_OK;
}

static int ens_lockgraph_run(pip_ens* e, int tc, int key, const EnsKernels& K) {
  const EnsArgs a = ens_args(e);
  if (e->lgraph_key != key) {
    HIPCHK(hipStreamSynchronize(e->stream));
    ens_drop_graphs(e);
    e->lgraph_key = key;
  }
  std::vector<char> run(e->n, 1);  // the set: every member, until half of them have stopped
  int cur = 0;
  TRY(ens_lockgraph_get(e, a, tc, K, run, &cur));
  int listed = e->lgraphs[cur].nm;
  K.start(a);
  HIPCHK(hipGetLastError());
  for (;;) {
    for (int b = 0; b < kGraphBatch; b++) HIPCHK(hipGraphLaunch(e->lgraphs[cur].exec, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    const int active = *e->h_active;
    if (active == 0) break;
    if (2 * active <= listed) {
      // the chunk's end wrote every listed member's scalars to buffer 0
      HIPCHK(hipMemcpyAsync(e->h_scs, e->scs, sizeof(DevScalars) * (size_t)e->n, hipMemcpyDeviceToHost,
                            e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
      for (int m = 0; m < e->n; m++) run[m] = (char)(run[m] && e->h_scs[m].active != 0);
      TRY(ens_lockgraph_get(e, a, tc, K, run, &cur));
      listed = e->lgraphs[cur].nm;
    }
  }
  return PIP_OK;
}

// The kernels of 1-D members at width TX (tiles of TX - 1 cells).
template <int TX>
static int ens_kernels_1d(pip_ens* e, EnsKernels* K) {
  using G = TileGeom<1, 8, TX - 1, 1>;
  const size_t lds_stage = sizeof(double) * (size_t)G::LDS_DOUBLES;
  const size_t lds_start = sizeof(double) * (size_t)(G::LDS_DOUBLES + 3 * (TX / 64));
  static bool attr = false;
  if (!attr) {
    HIPCHK(hipFuncSetAttribute((const void*)k_lock_start<TX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_start));
    HIPCHK(hipFuncSetAttribute((const void*)k_lock_stage<TX, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_stage));
    HIPCHK(hipFuncSetAttribute((const void*)k_lock_stage<TX, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_stage));
    attr = true;
  }
  hipStream_t st = e->stream;
  DevScalars* scs = e->scs;
  const int n = e->n;
  K->start = [=](const EnsArgs& a) {
    hipLaunchKernelGGL(k_lock_start<TX>, dim3((unsigned)n), dim3(TX), lds_start, st, a, scs);
  };
  K->stage = [=](const EnsArgs& a, const LockTile* tiles, int istep, int nt, int p, int apply_end) {
    if (istep == 1)
      hipLaunchKernelGGL((k_lock_stage<TX, 1>), dim3((unsigned)nt), dim3(TX), lds_stage, st, a, scs, tiles, p, apply_end);
    else
      hipLaunchKernelGGL((k_lock_stage<TX, 2>), dim3((unsigned)nt), dim3(TX), lds_stage, st, a, scs, tiles, p, apply_end);
  };
  return PIP_OK;
}

// The kernels of 2-D members (16x15 tiles).
static int ens_kernels_2d(pip_ens* e, EnsKernels* K) {
  const size_t lds_stage = sizeof(double) * (size_t)TileGeom<2, 9, kEns2TX, kEns2TY>::LDS_DOUBLES;
  static bool attr = false;
  if (!attr) {
    HIPCHK(hipFuncSetAttribute((const void*)k_lock_stage2<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_stage));
    HIPCHK(hipFuncSetAttribute((const void*)k_lock_stage2<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_stage));
    attr = true;
  }
  hipStream_t st = e->stream;
  DevScalars* scs = e->scs;
  const int n = e->n;
  K->start = [=](const EnsArgs& a) {
    hipLaunchKernelGGL(k_lock_start2<256>, dim3((unsigned)n), dim3(256), 0, st, a, scs);
  };
  K->stage = [=](const EnsArgs& a, const LockTile* tiles, int istep, int nt, int p, int apply_end) {
    if (istep == 1)
      hipLaunchKernelGGL(k_lock_stage2<1>, dim3((unsigned)nt), dim3(kEns2NT), lds_stage, st, a, scs, tiles, p, apply_end);
    else
      hipLaunchKernelGGL(k_lock_stage2<2>, dim3((unsigned)nt), dim3(kEns2NT), lds_stage, st, a, scs, tiles, p, apply_end);
  };
  return PIP_OK;
}

// Lockstep (graph = false) or graph-captured lockstep (graph = true) run of
// all members.
static int ens_lock_any(pip_ens* e, bool graph) {
  EnsKernels K;
  int tc = 0;
  if (e->twod) {
    TRY(ens_kernels_2d(e, &K));
  } else {
    switch (e->tx) {
      case 64: TRY(ens_kernels_1d<64>(e, &K)); break;
      case 128: TRY(ens_kernels_1d<128>(e, &K)); break;
      case 256: TRY(ens_kernels_1d<256>(e, &K)); break;
      case 512: TRY(ens_kernels_1d<512>(e, &K)); break;
      default: return fail(PIP_ERR_ARG, "ensemble tile must be 64, 128, 256 or 512");
    }
    tc = e->tx - 1;
  }
  const int key = e->twod ? 1 : e->tx;  // what the captured graph was built for
  return graph ? ens_lockgraph_run(e, tc, key, K) : ens_lockstep_run(e, tc, K);
}

static int ens_lockgraph(pip_ens* e) { return ens_lock_any(e, true); }
static int ens_lockstep(pip_ens* e) { return ens_lock_any(e, false); }

static int ens_enqueue(pip_ens* e) {
  switch (e->tx) {
    case 64: return ens_launch<64>(e);
    case 128: return ens_launch<128>(e);
    case 256: return ens_launch<256>(e);
    case 512: return ens_launch<512>(e);
  }
  return fail(PIP_ERR_ARG, "ensemble tile must be 64, 128, 256 or 512");
}

// The ensemble run between events a and b on the ensemble's stream (the
// graph, when asked for, is captured on first use and kept); *ms = the
// device time between them.
static int ens_run_between(pip_ens* e, int use_graph, hipEvent_t a, hipEvent_t b, float* ms) {
  if (e->twod && e->mode == 0)
    return fail(PIP_ERR_UNSUPPORTED, "2-D ensembles run in lockstep (mode 1) or graph-captured lockstep (mode 2)");
  if (e->mode == 1) {  // host-driven chunks: no graph
    HIPCHK(hipEventRecord(a, e->stream));
    TRY(ens_lockstep(e));
    HIPCHK(hipEventRecord(b, e->stream));
  } else if (e->mode == 2) {  // graph-captured lockstep chunks
    HIPCHK(hipEventRecord(a, e->stream));
    TRY(ens_lockgraph(e));
    HIPCHK(hipEventRecord(b, e->stream));
  } else if (use_graph) {
    if (!e->gexec || e->graph_tx != e->tx) {
      if (e->gexec) {
        HIPCHK(hipGraphExecDestroy(e->gexec));
        e->gexec = nullptr;
      }
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      const int rc = ens_enqueue(e);
      const hipError_t ce = hipStreamEndCapture(e->stream, &g);
      if (rc != PIP_OK) {
        if (ce == hipSuccess) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ce != hipSuccess) return fail(PIP_ERR_HIP, std::string("capture: ") + hipGetErrorString(ce));
      const hipError_t ie = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        e->gexec = nullptr;
        return fail(PIP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
      }
      e->graph_tx = e->tx;
    }
    HIPCHK(hipEventRecord(a, e->stream));
    HIPCHK(hipGraphLaunch(e->gexec, e->stream));
    HIPCHK(hipEventRecord(b, e->stream));
  } else {
    HIPCHK(hipEventRecord(a, e->stream));
    TRY(ens_enqueue(e));
    HIPCHK(hipEventRecord(b, e->stream));
  }
  HIPCHK(hipEventSynchronize(b));
  HIPCHK(hipEventElapsedTime(ms, a, b));
  return PIP_OK;
}

extern "C" {

int pip_ens_create(const pip_cfg* base, const pip_ens_member* members, int n, pip_ens** out) {
  if (!base || !members || !out || n <= 0) return fail(PIP_ERR_ARG, "null argument or n <= 0");
  *out = nullptr;
  const bool twod = base->ndim == 2;
  if (!((base->ndim == 1 && base->nvar_m == 8) || (twod && base->nvar_m == 9 && base->flag_divb == 1)))
    return fail(PIP_ERR_UNSUPPORTED, "ensemble members are 1-D (nvar_m = 8) or 2-D with psi (nvar_m = 9, flag_divb = 1)");
  if (base->flag_ir != 0 || base->ac_field || base->flag_col != 1 || base->flag_pip_imp != 0)
    return fail(PIP_ERR_UNSUPPORTED, "ensemble: flag_IR = 0, constant alpha_c, explicit collisions");
  if (base->t_order != 2) return fail(PIP_ERR_UNSUPPORTED, "ensemble: t_order = 2");
  for (int s = 0; s < 2 * base->ndim; s++) {
    int t = bnd_kind(base->flag_bnd[s]);
    if (t != 1 && t != 2 && t != 10) return fail(PIP_ERR_UNSUPPORTED, "boundary flag must be 1, 2-7, 10 or 11");
  }
  for (int m = 0; m < n; m++) {
    if (members[m].nx < 4 || !(members[m].dx > 0)) return fail(PIP_ERR_ARG, "member nx >= 4 and dx > 0 required");
    if (members[m].nx + 4 > kMaxPad) return fail(PIP_ERR_ARG, "member nx must be <= 32763");
    if (twod && (members[m].ny < 4 || !(members[m].dy > 0) || members[m].ny + 4 > kMaxPad))
      return fail(PIP_ERR_ARG, "2-D members need 4 <= ny <= 32763 and dy > 0");
  }
  pip_ens* e = new pip_ens();
  e->base = *base;
  e->n = n;
  e->twod = twod;
  e->nvt = base->nvar_m + 5;
  e->ph.gm = base->gm;
  e->ph.ro_lim = base->ro_lim;
  e->ph.pr_lim = base->pr_lim;
  e->ph.safety = base->safety;
  e->ph.T0 = base->T0;
  e->ph.n0 = base->n0;
  e->ph.beta = base->beta;
  e->ph.t_ir = base->t_ir;
  e->ph.flag_col = base->flag_col;
  e->ph.flag_ir = 0;
  e->ph.flag_ir_type = 0;
  {  // boundary masks of a member context
    pip_ctx tmp;
    tmp.cfg = *base;
    tmp.nv = base->nvar_m;
    build_bnd(&tmp);
    e->bnd = tmp.bnd;
  }
  e->hmem.resize(n);
  long long off = 0;
  for (int m = 0; m < n; m++) {
    EnsMember& M = e->hmem[m];
    M.off = off;
    M.nx = members[m].nx;
    M.ix = members[m].nx + 4;
    M.dx = members[m].dx;
    M.col = members[m].col;
    M.ny = twod ? members[m].ny : 0;
    M.jx = twod ? members[m].ny + 4 : 1;
    M.dy = twod ? members[m].dy : 100.0;  // Model_rot.f90:424-429
    M.dxyz_min = dmin(dmin(M.dx, M.dy), 100.0);
    M.gmin_safety = base->safety * M.dxyz_min;
    M.tend = members[m].tend;
    M.nmax = members[m].max_steps > 0 ? members[m].max_steps : ((long long)1 << 62);
    off += (long long)e->nvt * M.ix * M.jx;
  }
  e->total = off;
  std::vector<int> order(n);
  for (int m = 0; m < n; m++) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    double ca = members[a].cost > 0 ? members[a].cost : (double)members[a].nx;
    double cb = members[b].cost > 0 ? members[b].cost : (double)members[b].nx;
    return ca > cb;
  });
  hipError_t er;
  auto bail = [&](const char* what) {
    std::string msg = std::string(what) + ": " + hipGetErrorString(er);
    pip_ens_destroy(e);
    return fail(PIP_ERR_HIP, msg);
  };
  if ((er = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream");
  if ((er = hipMalloc(&e->U, sizeof(double) * off)) != hipSuccess) return bail("hipMalloc U");
  if ((er = hipMalloc(&e->U1, sizeof(double) * off)) != hipSuccess) return bail("hipMalloc U1");
  if ((er = hipMalloc(&e->Uinit, sizeof(double) * off)) != hipSuccess) return bail("hipMalloc Uinit");
  if ((er = hipMalloc(&e->mem, sizeof(EnsMember) * n)) != hipSuccess) return bail("hipMalloc mem");
  if ((er = hipMalloc(&e->out, sizeof(EnsOut) * n)) != hipSuccess) return bail("hipMalloc out");
  if ((er = hipMalloc(&e->order, sizeof(int) * n)) != hipSuccess) return bail("hipMalloc order");
  if ((er = hipMalloc(&e->counter, sizeof(unsigned) * 4)) != hipSuccess) return bail("hipMalloc counter");
  for (int m = 0; m < n; m++)  // tiles of the narrowest 1-D width, or 16x15 tiles
    e->max_tiles += twod ? ((members[m].nx + kEns2TX - 1) / kEns2TX) * ((members[m].ny + kEns2TY - 1) / kEns2TY)
                         : (members[m].nx + 62) / 63;
  if ((er = hipMalloc(&e->scs, 2 * sizeof(DevScalars) * n)) != hipSuccess) return bail("hipMalloc scalars");
  if ((er = hipMalloc(&e->list, sizeof(int) * n)) != hipSuccess) return bail("hipMalloc list");
  if ((er = hipHostMalloc((void**)&e->h_list, sizeof(int) * n, hipHostMallocDefault)) != hipSuccess)
    return bail("hipHostMalloc list");
  if ((er = hipMalloc(&e->tiles, sizeof(LockTile) * e->max_tiles)) != hipSuccess) return bail("hipMalloc tiles");
  if ((er = hipHostMalloc((void**)&e->h_tiles, sizeof(LockTile) * e->max_tiles, hipHostMallocDefault)) != hipSuccess)
    return bail("hipHostMalloc tiles");
  if ((er = hipHostMalloc((void**)&e->h_scs, sizeof(DevScalars) * n, hipHostMallocDefault)) != hipSuccess)
    return bail("hipHostMalloc scalars");
  if ((er = hipMalloc(&e->d_active, sizeof(int))) != hipSuccess) return bail("hipMalloc active");
  if ((er = hipHostMalloc((void**)&e->h_active, sizeof(int), hipHostMallocDefault)) != hipSuccess)
    return bail("hipHostMalloc active");
  if ((er = hipMemcpy(e->mem, e->hmem.data(), sizeof(EnsMember) * n, hipMemcpyHostToDevice)) != hipSuccess) return bail("copy");
  if ((er = hipMemcpy(e->order, order.data(), sizeof(int) * n, hipMemcpyHostToDevice)) != hipSuccess) return bail("copy");
  if ((er = hipMemset(e->U, 0, sizeof(double) * off)) != hipSuccess) return bail("memset");
  if ((er = hipMemset(e->U1, 0, sizeof(double) * off)) != hipSuccess) return bail("memset");
  if ((er = hipMemset(e->Uinit, 0, sizeof(double) * off)) != hipSuccess) return bail("memset");
  if ((er = hipMemset(e->out, 0, sizeof(EnsOut) * n)) != hipSuccess) return bail("memset");
  *out = e;
  return PIP_OK;
}

int pip_ens_destroy(pip_ens* e) {
  if (!e) return PIP_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  ens_drop_graphs(e);
  (void)hipFree(e->d_active);
  if (e->h_active) (void)hipHostFree(e->h_active);
  (void)hipFree(e->U);
  (void)hipFree(e->U1);
  (void)hipFree(e->Uinit);
  (void)hipFree(e->mem);
  (void)hipFree(e->out);
  (void)hipFree(e->order);
  (void)hipFree(e->counter);
  (void)hipFree(e->scs);
  (void)hipFree(e->tiles);
  (void)hipFree(e->list);
  if (e->h_tiles) (void)hipHostFree(e->h_tiles);
  if (e->h_list) (void)hipHostFree(e->h_list);
  if (e->h_scs) (void)hipHostFree(e->h_scs);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return PIP_OK;
}

int pip_ens_upload(pip_ens* e, int m, const double* U_h, const double* U_m) {
  if (!e || !U_h || !U_m || m < 0 || m >= e->n) return fail(PIP_ERR_ARG, "bad member or null array");
  const EnsMember& M = e->hmem[m];
  const long long N = (long long)M.ix * M.jx;
  const int nv = e->nvt - 5;
  const size_t pb = sizeof(double) * (size_t)N;
  HIPCHK(hipMemcpy(e->Uinit + M.off, U_m, pb * nv, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->Uinit + M.off + nv * N, U_h, pb * 5, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->U + M.off, e->Uinit + M.off, pb * e->nvt, hipMemcpyDeviceToDevice));
  return PIP_OK;
}

int pip_ens_download(pip_ens* e, int m, double* U_h, double* U_m) {
  if (!e || !U_h || !U_m || m < 0 || m >= e->n) return fail(PIP_ERR_ARG, "bad member or null array");
  const EnsMember& M = e->hmem[m];
  const long long N = (long long)M.ix * M.jx;
  const int nv = e->nvt - 5;
  const size_t pb = sizeof(double) * (size_t)N;
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipMemcpy(U_m, e->U + M.off, pb * nv, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(U_h, e->U + M.off + nv * N, pb * 5, hipMemcpyDeviceToHost));
  return PIP_OK;
}

int pip_ens_set_tile(pip_ens* e, int tx) {
  if (!e) return fail(PIP_ERR_ARG, "null ensemble");
  if (tx != 64 && tx != 128 && tx != 256 && tx != 512) return fail(PIP_ERR_ARG, "tile must be 64/128/256/512");
  e->tx = tx;
  return PIP_OK;
}

int pip_ens_set_mode(pip_ens* e, int mode) {
  if (!e) return fail(PIP_ERR_ARG, "null ensemble");
  if (mode < 0 || mode > 2)
    return fail(PIP_ERR_ARG, "ensemble mode must be 0 (persistent), 1 (lockstep) or 2 (graph-captured lockstep)");
  e->mode = mode;
  return PIP_OK;
}

int pip_ens_reset(pip_ens* e) {
  if (!e) return fail(PIP_ERR_ARG, "null ensemble");
  HIPCHK(hipMemcpyAsync(e->U, e->Uinit, sizeof(double) * e->total, hipMemcpyDeviceToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return PIP_OK;
}

int pip_ens_run(pip_ens* e, int use_graph, double* ms) {
  if (!e) return fail(PIP_ERR_ARG, "null ensemble");
  hipEvent_t a = nullptr, b = nullptr;
  HIPCHK(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail(PIP_ERR_HIP, "pip_ens_run: hipEventCreate");
  }
  float t = 0.f;
  const int rc = ens_run_between(e, use_graph, a, b, &t);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc == PIP_OK && ms) *ms = t;
  return rc;
}

int pip_ens_result(pip_ens* e, int m, double* time, long* nt, int* status) {
  if (!e || m < 0 || m >= e->n) return fail(PIP_ERR_ARG, "bad member");
  EnsOut o;
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipMemcpy(&o, e->out + m, sizeof(o), hipMemcpyDeviceToHost));
  if (time) *time = o.time;
  if (nt) *nt = (long)o.nt;
  if (status) *status = o.status;
  return PIP_OK;
}

}  // extern "C"
