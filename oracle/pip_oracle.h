/*
 * pip_oracle.h -- CPU oracle for the PIP two-fluid timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran path (AstroSnow/PIP) used as the parity checker for the HIP kernels.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library (libpipgpu.so) never links or calls it.
 *
 * Every routine follows the reference's whole-array semantics literally
 * (ghost cells, sweep ranges, in-place state rewrites, evaluation order of
 * each expression, single-precision literals) and cites the file:line it
 * restates.  Arrays use the reference layout U(ix,jx,kx=1,nvar): x fastest,
 * variable slowest; index (i,j,n) -> n*ix*jx + j*ix + i with 0-based i,j.
 */
#ifndef PIP_ORACLE_H
#define PIP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int ndim;          /* 1 or 2 (kx = 1)                                   */
    int ix, jx;        /* padded sizes incl. margin=2 ghosts (jx=1 in 1-D)  */
    int nvar_m;        /* 8, or 9 with flag_divb=1 (psi)                    */
    int flag_divb;     /* 0 / 1 (Dedner GLM, 2-D only)                      */
    int t_order;       /* SSP-RK order, flag_time=1 (2 or 3)                */
    int flag_ir;       /* 0 off, 3 = set_IR case 3                          */
    int flag_ir_type;  /* 0 new (ion_pot/arb_heat), 1 Khomenko, 2 Singh     */
    int flag_col;      /* 1 constant alpha; 2..7 set_collisional col_type 1..6 */
    int ac_field;      /* 1: per-cell ac supplied by caller (kept as is)    */
    int flag_bnd[4];   /* x-, x+, y-, y+ : 1 periodic, 2..7/11 mirror, 10 free */
    double gm, safety, ro_lim, pr_lim, dx, dy;
    double col, db_clean, t_ir, T0, n0, beta;
    int flag_pip_imp;      /* 1: implicit collision source (PIP_rot:1333-1358) */
    double pip_imp_factor; /* its dt limit factor (Scheme_rot.f90:308-309)  */
} orc_cfg;

typedef struct orc_state orc_state;

orc_state *orc_create(const orc_cfg *cfg);
void orc_destroy(orc_state *s);

/* state access (padded arrays, reference layout) */
void orc_set_state(orc_state *s, const double *U_h, const double *U_m);
void orc_get_state(const orc_state *s, double *U_h, double *U_m);
void orc_set_ac(orc_state *s, const double *ac);
void orc_get_ac(const orc_state *s, double *ac);
void orc_get_rates(const orc_state *s, double *gm_rec, double *gm_ion,
                   double *arb_heat, double *ion_pot);
void orc_set_rates(orc_state *s, const double *gm_rec, const double *gm_ion);
void orc_get_scalars(const orc_state *s, double *out4); /* dt,time,cmax,nt */
/* test hooks: set dt and the GLM speed cmax used by orc_run_solver; copy
 * the stage-1 result (U_m planes then U_h planes) into buf (NULL: off) */
void orc_set_scalars(orc_state *s, double dt, double cmax);
void orc_set_stage_save(orc_state *s, double *buf);

/* reference entry points (Fortran names) */
void orc_pipbnd(orc_state *s);                 /* Boundary_rot.f90:34-52   */
void orc_set_coefficients(orc_state *s, int flag); /* Solver_rot.f90:95-121 */
void orc_set_ir(orc_state *s);                 /* PIP_rot.f90:127-311      */
double orc_cfl(orc_state *s, int into_global_dt); /* Scheme_rot.f90:183-213 */
void orc_run_solver(orc_state *s);             /* Solver_rot.f90:42-81     */
void orc_prologue(orc_state *s);               /* Initial_rot.f90:140-160  */
int orc_step(orc_state *s);                    /* main.f90:31-68, one pass */
int orc_stop_sim(const orc_state *s);          /* IO_rot.f90:625-674       */
/* fixed rates mode for source tests (no set_IR): IR sources use given rates */
void orc_set_fixed_rates(orc_state *s, int on);
/* test hook: arb_heat field (as if set_IR had fixed it) */
void orc_set_arb_heat(orc_state *s, const double *arb_heat);

/* threads of the elementwise loops (OpenMP; 1 = serial); returns the
 * previous setting */
int orc_set_threads(int n);

/* point functions (pure) */
double orc_exp(double x);
double orc_log(double x);
double orc_pow(double x, double y);
double orc_minmod(double u1, double u2, double u3);        /* Util_rot:59-76 */
void orc_hllc_flux(const double *vl, const double *vr, double gm, double *flx);
void orc_hlld_flux(const double *vl, const double *vr, int nvar, int divb,
                   double cmax, double gm, double *flx);

/* whole-array pieces for unit tests */
void orc_set_flux(orc_state *s, double *F_h, double *F_m);  /* Solver_rot:124 */
void orc_source_pip(orc_state *s, double *S_h, double *S_m); /* PIP_rot:1309 */

#ifdef __cplusplus
}
#endif
#endif
