/*
 * pipgpu.h -- C-ABI of libpipgpu.so, the MI355X (gfx950) implementation of
 * PIP's per-timestep two-fluid update (flag_eqs='PIP', HLL branch with HLLD
 * for the plasma and HLLC for the neutrals, minmod MUSCL, SSP-RK, explicit
 * ion-neutral collisions, set_IR case 3 ionisation/recombination, CFL dt).
 *
 * The reference exposes this path as Fortran module procedures operating on
 * module globals; each entry point below names the procedure it replaces.
 * Conventions:
 *   - arrays are the reference's explicit-shape U(ix,jx,kx=1,nvar) in
 *     column-major order (x fastest, variable slowest), ghosts included;
 *   - every call returns 0 on success or a negative code; pip_last_error()
 *     returns the message of the last failure on the calling thread;
 *   - the state lives on the GPU between calls; host copies only happen in
 *     pip_upload / pip_download (and the 8-byte dt of pip_cfl);
 *   - calls on one context must come from one host thread at a time.
 */
#ifndef PIPGPU_H
#define PIPGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define PIP_OK 0
#define PIP_ERR_ARG -1
#define PIP_ERR_HIP -2
#define PIP_ERR_STATE -3
#define PIP_ERR_UNSUPPORTED -4

/* stop_sim flags (IO_rot.f90:625-674) */
#define PIP_STOP_DT_SMALL 1
#define PIP_STOP_NAN 2

/* Configuration: the subset of module globalvar (Globalvar.f90:42-101) and
 * of the derived values of get_parameters (Model_rot.f90:255-333) that the
 * two-fluid step reads. */
typedef struct pip_cfg {
    int ndim;          /* 1 or 2 (kx = 1)                                 */
    int ix, jx;        /* padded sizes, margin = 2 (jx = 1 when ndim = 1)  */
    int nvar_m;        /* 8, or 9 with flag_divb = 1                       */
    int flag_divb;     /* Dedner GLM cleaning (2-D only)                   */
    int t_order;       /* SSP-RK order: 2 or 3 (flag_time = 1)             */
    int flag_ir;       /* 0 off, 3 = Popescu et al. empirical rates        */
    int flag_ir_type;  /* 0 (ion_pot/arb_heat), 1, 2                       */
    int flag_col;      /* 1 constant alpha_c = col; 2..7 set_collisional   */
    int ac_field;      /* 1: per-cell alpha_c uploaded with the state      */
    int flag_bnd[4];   /* x-, x+, y-, y+: 1 periodic, 2-7/11 mirror, 10 free */
    double gm, safety, ro_lim, pr_lim, dx, dy;
    double col, db_clean, t_ir, T0, n0, beta;
    int flag_pip_imp;       /* 1: implicit collision source                */
    double pip_imp_factor;  /* its dt factor (cfl_pn_col)                  */
} pip_cfg;

typedef struct pip_ctx pip_ctx;

/* allocate_vars + initialize_bnd (Model_rot.f90:376-392,
 * Boundary_rot.f90:20-32): device state for one domain. */
int pip_create(const pip_cfg *cfg, pip_ctx **out);
int pip_destroy(pip_ctx *ctx);

/* Host -> device copy of U_h(ix,jx,1,5), U_m(ix,jx,1,nvar_m) and, when
 * cfg.ac_field = 1, ac(ix,jx,1) (Globalvar.f90:17,26). */
int pip_upload(pip_ctx *ctx, const double *U_h, const double *U_m, const double *ac);
/* Device -> host copy of the state (the arrays IO_rot.f90:362-374 writes). */
int pip_download(pip_ctx *ctx, double *U_h, double *U_m);

/* Prologue tail, Initial_rot.f90:140-160: ghost fill, set_coefficients(.,0)
 * (set_collisional for flag_col 2..7, set_IR incl. arb_heat), cfl(dt).
 * Returns the first dt. */
int pip_prologue(pip_ctx *ctx, double *dt);

/* set_IR (PIP_rot.f90:127-311, case 3): get_Te rewrites + Gm_ion, Gm_rec. */
int pip_set_ir(pip_ctx *ctx);
/* cfl (Scheme_rot.f90:183-213): rewrites U, returns dt; when into_global = 1
 * the context's dt and cmax are updated as `call cfl(U_h,U_m,dt)` does. */
int pip_cfl(pip_ctx *ctx, double *dt, int into_global);
/* run_solver (Solver_rot.f90:42-81) with the context's dt. */
int pip_run_solver(pip_ctx *ctx);
/* PIPbnd (Boundary_rot.f90:34-52) single rank. */
int pip_bnd(pip_ctx *ctx);

/* The whole time loop of main.f90:31-68 on the device, fused: runs until
 * time >= tend, max_steps steps, or a stop_sim condition.  steps_done and
 * time may be NULL. */
int pip_advance(pip_ctx *ctx, int max_steps, double tend, int *steps_done, double *time);
/* 1 (default): pip_advance replays its 64-step chunks from a hipGraph
 * captured on first use (one graph launch instead of 3 launches per step;
 * the same kernels, so the same results); 0: plain stream launches. */
int pip_set_step_graph(pip_ctx *ctx, int on);
/* Enqueue n fused steps without waiting (no host sync); for timing loops. */
int pip_enqueue_steps(pip_ctx *ctx, int n, double tend);
int pip_sync(pip_ctx *ctx);

int pip_get_scalars(pip_ctx *ctx, double *dt, double *time, double *cmax, long *nt);
int pip_set_time(pip_ctx *ctx, double time, double dt);
/* stop_sim flags of the last step (PIP_STOP_*). */
int pip_status(pip_ctx *ctx, int *flags);

/* Rate fields after pip_set_ir (Gm_rec, Gm_ion, arb_heat), for inspection. */
int pip_get_rates(pip_ctx *ctx, double *gm_rec, double *gm_ion, double *arb_heat);
/* The collision coefficient field ac(ix,jx,1) (Globalvar.f90:26) as the
 * output routines see it: the per-cell field, the alpha of the last stage
 * for flag_col 2..7, or col everywhere for flag_col = 1. */
int pip_get_ac(pip_ctx *ctx, double *ac);
/* Restart (IO_rot.f90:585-589): host -> device copy of arb_heat(ix,jx,1)
 * read from a snapshot; the next pip_prologue keeps it instead of fixing it
 * from the state (set_IR does that only for a fresh run, PIP_rot.f90:219). */
int pip_set_arb_heat(pip_ctx *ctx, const double *arb_heat);

/* Device pointers and stream for zero-copy interop (torch.autograd wrapper,
 * ensemble driver).  U: nvar_m + 5 planes of ix*jx doubles (U_m then U_h). */
int pip_device_state(pip_ctx *ctx, double **U, double **ac);
int pip_get_stream(pip_ctx *ctx, void **stream);
/* Stage-kernel tile shape (autotuner): 16x15, 32x7, 32x15, 64x7, 32x8 or
 * 16x7 cells per workgroup. */
int pip_set_tile(pip_ctx *ctx, int tx, int ty);
/* Bit s-1 of stage_mask set: SSP-RK stage s takes the two fluids one after
 * the other in a tile of nvar_m planes (three workgroups per CU; tiles 16x15,
 * 32x7, 16x7; 2-D runs with explicit constant or per-cell alpha, with or
 * without set_IR; other runs keep the one-pass kernel); clear: both fluids
 * at once.  Bitwise the same results (autotuner). */
int pip_set_two_pass(pip_ctx *ctx, int stage_mask);
/* Arithmetic of the stage kernels.  PIP_ARITH_EXACT (the default): the
 * reference's operations in the reference's order, no FMA contraction,
 * correctly rounded quotients and square roots -- bitwise the CPU reference
 * path.  PIP_ARITH_RELAXED: the same operations with FMA contraction,
 * quotients as products with a refined reciprocal and the hardware min/max.
 * Not bitwise: after 100 smooth steps the variables of order one
 * (densities, energies, the background field, a shear flow's momentum) stay
 * within 1e-10 relative L1 of the reference path, but quantities set by a
 * small perturbation or by round-off (transverse momenta and fields, psi)
 * differ by up to ~1e-5 of their own size; shocks stay at the same cells.
 * Relaxed covers 2-D runs with psi (with or without set_IR, tiles 16x15,
 * 32x7, 16x7) and 1-D runs without set_IR, with constant or per-cell
 * alpha_c and explicit collisions (PIP_ERR_UNSUPPORTED otherwise); it
 * ignores the two-pass stage mask, and the adjoint's frozen-dt replays stay
 * exact. */
#define PIP_ARITH_EXACT 0
#define PIP_ARITH_RELAXED 1
int pip_set_arith(pip_ctx *ctx, int mode);
/* 1: map workgroups to tiles so that each of the 8 XCDs gets a contiguous
 * range of tiles (neighbouring halos in one L2); 0: row-major (autotuner). */
int pip_set_xcd_remap(pip_ctx *ctx, int on);
/* Average device time (ms) of the last timed enqueue of the dominant kernel;
 * measured with hipEvents on the context's stream. */
int pip_time_steps(pip_ctx *ctx, int nsteps, double *ms_total, double *ms_stage1, double *ms_stage2);

/* ---- discrete adjoint (no counterpart in the reference: PIP has no
 * derivatives).  Supports the reverse pass of a torch.autograd.Function:
 * record a run's dt/cmax sequence, replay it with frozen dt from a
 * checkpoint while saving the stage inputs, and pull an adjoint back
 * through one step.  All pointers are device pointers with the nvar_m + 5
 * plane layout of pip_device_state; calls are stream-ordered on the
 * context's stream (pip_sync before reading results from another stream).
 * Adjoint steps cover ndim = 2, flag_divb = 1, flag_IR = 0, SSP-RK2 and
 * periodic or mirror boundaries. */
/* device -> device copies of the state and of the per-cell alpha_c */
int pip_set_state_device(pip_ctx *ctx, const double *U);
int pip_get_state_device(pip_ctx *ctx, double *U);
int pip_set_ac_device(pip_ctx *ctx, const double *ac);
/* n fused steps of the time loop (main.f90:31-68, no end time) recording
 * hist[2k] = dt and hist[2k+1] = cmax of step k (device array, 2n doubles) */
int pip_record_steps(pip_ctx *ctx, int n, double *hist);
/* n steps with dt/cmax taken from hist (frozen-dt replay of a recorded run,
 * bitwise equal to it from the same state); when Un / U1 are non-NULL the
 * stage-1 input U_n and the stage-2 input U1 of step k are saved at plane
 * offset k*(nvar_m+5) of each array, and when Vn / V1 are non-NULL (with Un
 * and U1, in the configuration the adjoint covers) so are their primitives
 * (u2v, HLL_rot.f90:1136-1186) for pip_adj_step.  On return time and the
 * step count have advanced by the n steps. */
int pip_replay_steps(pip_ctx *ctx, int n, const double *hist, double *Un, double *U1, double *Vn,
                     double *V1);
/* pip_record_steps that also keeps what pip_replay_steps would save: the
 * stage inputs U_n, U1 of step k and their primitives at plane offset
 * k*(nvar_m+5) of Un, U1, Vn, V1 (all required).  The same steps and bits as
 * pip_record_steps; for the last checkpoint segment of a recorded run, whose
 * replay the backward sweep can then skip.  Covers the configuration of
 * pip_replay_steps's primitive saves in the exact arithmetic. */
int pip_record_steps_saved(pip_ctx *ctx, int n, double *hist, double *Un, double *U1, double *Vn,
                           double *V1);
/* adjoint seed: add the ghost-cell entries of lam onto the interior cells
 * they image (boundary sign maps), then zero the ghosts */
int pip_adj_seed(pip_ctx *ctx, double *lam);
/* one step backwards: on entry lam = dJ/dU_{n+1} (interior cells), on exit
 * lam = dJ/dU_n; lam_ac (may be NULL) accumulates dJ/dalpha_c.  Un, U1 are
 * the saved stage inputs of step n, Vn, V1 their saved primitives (both
 * NULL: computed here) and dtc = &hist[2n]. */
int pip_adj_step(pip_ctx *ctx, const double *Un, const double *U1, const double *Vn, const double *V1,
                 const double *dtc, double *lam, double *lam_ac);

/* pip_adj_step without allocation.  No counterpart in the reference (PIP
 * has no derivatives); for callers that cannot allocate device memory --
 * code under HIP graph capture, tight loops of an inverse solver, harnesses
 * that forbid hipMalloc.  pip_adj_step allocates its work buffers on its
 * first call; this variant keeps them (the face-state adjoints and branch
 * codes, and the adjoint of the stage-1 result: (5 (nvar_m + 5) + 4) * 8
 * bytes per padded cell, 592 B with psi) in one fixed 64 MiB device array
 * per GPU, so grids are limited to ix * jx <= 113359 (up to 332^2 interior
 * cells); larger grids return PIP_ERR_ARG.  Vn and V1 (the replay's saved
 * primitives) are required.  Same kernels and results as pip_adj_step.
 * Every call on one GPU uses that GPU's array: calls that overlap on one
 * GPU (other streams, other contexts) race. */
int pip_adj_step_noalloc(pip_ctx *ctx, const double *Un, const double *U1, const double *Vn,
                         const double *V1, const double *dtc, double *lam, double *lam_ac);

/* ---- batched ensemble (no counterpart in the reference: PIP runs one
 * domain per process).  Independent 1-D or 2-D two-fluid runs of different
 * sizes, each with its own dt/time/step count, integrated together (one
 * persistent launch, lockstep launches, or lockstep chunks replayed from a
 * hipGraph); every member equals its plain pip_advance run bit for bit. */
typedef struct pip_ens_member {
    int nx;            /* interior cells (ix = nx + 4)                     */
    double dx;         /* cell width                                      */
    double col;        /* alpha_c (flag_col = 1)                           */
    double tend;       /* end time of this run (main.f90:31)               */
    int max_steps;     /* step cap                                        */
    double cost;       /* expected work, for longest-first scheduling     */
    int ny;            /* 2-D ensembles: interior rows (jx = ny + 4)       */
    double dy;         /* 2-D ensembles: cell height                       */
} pip_ens_member;
typedef struct pip_ens pip_ens;
/* base supplies gm, safety, limits, flag_bnd, db_clean and the dimension:
 * ndim = 1 (nvar_m = 8; members nx, dx) or ndim = 2 with psi (nvar_m = 9,
 * flag_divb = 1; members nx, ny, dx, dy; lockstep modes only).  Every member
 * has the layout of a plain context: nvar_m + 5 planes of ix * jx. */
int pip_ens_create(const pip_cfg *base, const pip_ens_member *members, int n, pip_ens **out);
int pip_ens_upload(pip_ens *e, int m, const double *U_h, const double *U_m);
int pip_ens_download(pip_ens *e, int m, double *U_h, double *U_m);
/* workgroup width (cells per tile = threads): 64, 128, 256 or 512 */
int pip_ens_set_tile(pip_ens *e, int tx);
/* 0: one persistent launch, a workgroup per member (default; use_graph
 * replays it from a hipGraph); 1: lockstep, each member's tiles spread over
 * workgroups, two stage launches per step for all running members (stage 1
 * applies the previous step's end), host-driven in 64-step chunks with the
 * tile table rebuilt between chunks (use_graph ignored); 2: graph-captured
 * lockstep, a 64-step chunk over the tile table of a set of members captured
 * as a hipGraph and replayed until no member runs (stopped members' tiles
 * return at once); when half of the set has stopped the set shrinks to the
 * members still running, each set's graph captured once and cached
 * (use_graph ignored).  All give every member's plain pip_advance run bit
 * for bit. */
int pip_ens_set_mode(pip_ens *e, int mode);
/* restore every member to its uploaded state (device copy, not timed) */
int pip_ens_reset(pip_ens *e);
/* run all members to their end; ms (may be NULL) = device time of the launch */
int pip_ens_run(pip_ens *e, int use_graph, double *ms);
int pip_ens_result(pip_ens *e, int m, double *time, long *nt, int *status);
int pip_ens_destroy(pip_ens *e);

/* Diagnostic, no counterpart in the reference: q[k] = n[k] / d[k] and
 * r[k] = sqrt(n[k]) computed by the device sequences the Riemann solvers,
 * conversions and cfl terms use (host arrays of count doubles; synchronous;
 * allocates). */
int pip_check_arith(const double *n, const double *d, double *q, double *r, long long count);

const char *pip_last_error(void);
const char *pip_version(void);

#ifdef __cplusplus
}
#endif
#endif
