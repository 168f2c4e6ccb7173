/*
 * pip_oracle.c -- CPU oracle (test infrastructure only) for the PIP two-fluid
 * timestep.  A literal C restatement of the reference Fortran routines on the
 * hot path; see pip_oracle.h for the layout and the usage restriction.
 *
 * Conventions that follow the Fortran exactly:
 *   - x**2 is x*x; a*b*c is (a*b)*c; a/b*c is (a/b)*c; left-to-right sums.
 *   - single-precision literals (e.g. 2.6e-19, 8.6173e-5) are (double)(float).
 *   - SIGN(1,x) is copysign(1,x); MIN/MAX of finite values.
 *   - whole-array semantics: every routine sweeps ghost cells as the
 *     reference does, and U is rewritten in place wherever the reference
 *     rewrites it (cq2pv_*, get_Pr_*).
 * Compile with -O2 -ffp-contract=off (no FMA contraction, like the reference
 * built for x86-64 without FMA).
 *
 * Elementwise loops (each iteration writes only its own cells) run on
 * OpenMP threads when the grid has more than PAR_MIN cells; the min/max
 * reductions of cfl stay serial, so every result is independent of the
 * thread count.
 */
#include "pip_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TINY 1.0e-20 /* Globalvar.f90:101 */
#define EPSI 1.0e-10 /* HLL_rot.f90:12 */
#define ACC 1.03     /* HLL_rot.f90:12 */
#define HLLG 1.001   /* HLL_rot.f90:701 */
#define MU_P 0.5     /* PIP_rot.f90:36 */
#define MU_N 1.0     /* PIP_rot.f90:37 */
#define SGL(x) ((double)(float)(x))
#define PI 3.1415926535897932 /* Parameters.f90:2 */
#define PAR_MIN 16384         /* cells below which loops stay serial */

struct orc_state {
    orc_cfg c;
    int N;                 /* ix*jx */
    double *Um, *Uh;       /* U_m(ix,jx,1,nvar_m), U_h(ix,jx,1,5) */
    double *ac, *gm_rec, *gm_ion, *arb_heat, *ion_pot;
    int arb_set, ion_pot_set, fixed_rates;
    double dt, time, cmax;
    long nt;
    int sym_m[4][9], sym_h[4][5];
    double *stage_save;    /* optional copy of the stage-1 result (test hook) */
    double dt_coll_i;      /* set_dt_coll value of the current stage        */
    /* scratch */
    double *V, *VL, *VR, *Fm, *Fh, *Sm, *Sh, *Um0, *Uh0, *tmp;
};

static double dmin(double a, double b) { return (a < b) ? a : b; }

/* exp / log / pow of the reference's dexp and ** (PIP_rot.f90:216,
 * Scheme_rot.f90:552), evaluated as fixed sequences of IEEE additions and
 * multiplications so that the HIP path can reproduce them bit for bit; they
 * are within a few ulp of the C library's (tested).
 * exp: Cody-Waite reduction x = k ln2 + r, degree-13 Taylor, scale 2^k.
 * log: x = m 2^e, m in [1/sqrt2, sqrt2), 2 atanh((m-1)/(m+1)) to s^25. */
#define LN2_HI 6.93147180369123816490e-01
#define LN2_LO 1.90821492927058770002e-10
#define INV_LN2 1.44269504088896338700e+00
double orc_exp(double x) {
    if (!(x == x)) return x;
    if (x > 709.782712893383973096) return HUGE_VAL;
    if (x < -745.133219101941108420) return 0.0;
    double kd = rint(x * INV_LN2);
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    static const double c[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0,
                                 1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0,
                                 1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0, 1.0};
    double q = c[0];
    for (int n = 1; n < 14; n++) q = q * r + c[n];
    return ldexp(q, (int)kd);
}
double orc_log(double x) {
    if (!(x > 0.0)) return (x == 0.0) ? -HUGE_VAL : (x - x) / (x - x);
    if (x == HUGE_VAL) return x;
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double t = 1.0 / 25.0;
    for (int k = 23; k >= 3; k -= 2) t = t * s2 + 1.0 / (double)k;
    t = t * s2 + 1.0;
    double lm = (2.0 * s) * t;
    double ed = (double)e;
    return ed * LN2_HI + (ed * LN2_LO + lm);
}
double orc_pow(double x, double y) {
    if (x == 0.0) return (y > 0.0) ? 0.0 : HUGE_VAL;
    return orc_exp(y * orc_log(x));
}
static double dmax(double a, double b) { return (a > b) ? a : b; }

/* ---------------------------------------------------------------- set-up */
/* set_sym_mhd / set_sym_hd: Boundary_rot.f90:239-320 */
static void set_sym(orc_state *s) {
    const orc_cfg *c = &s->c;
    for (int i = 0; i < 4; i++) {
        for (int n = 0; n < 9; n++) s->sym_m[i][n] = 1;
        for (int n = 0; n < 5; n++) s->sym_h[i][n] = 1;
    }
    for (int i = 1; i <= 2 * c->ndim; i++) {
        int dir = (i + 1) / 2, f = c->flag_bnd[i - 1];
        int *sm = s->sym_m[i - 1], *sh = s->sym_h[i - 1];
        /* MHD (Fortran indices n -> sm[n-1]) */
        if (f == 2) { sm[1 + dir - 1] = -1; sm[5 + dir - 1] = -1; }
        else if (f == 3) { sm[dir] = -1; sm[6 + dir % 3 - 1] = -1; sm[2 + (dir + 1) % 3 - 1] = -1; }
        else if (f == 4) { sm[dir] = -1; sm[6 + (dir + 1) % 3 - 1] = -1; sm[2 + dir % 3 - 1] = -1; }
        else if (f == 5) { sm[dir] = -1; sm[5 + dir - 1] = -1; sm[2 + (dir + 1) % 3 - 1] = -1; sm[6 + (dir + 1) % 3 - 1] = -1; }
        else if (f == 6) { sm[dir] = -1; sm[6 + dir % 3 - 1] = -1; sm[6 + (dir + 1) % 3 - 1] = -1; }
        else if (f == 7) { sm[1] = -1; sm[2] = -1; sm[3] = -1; }
        if (c->flag_divb == 1) {
            if (f == 2 || f == 20 || f == 5) sm[8] = 1;
            else if (f == 3 || f == 4 || f == 6 || f == 7) sm[8] = -1;
        }
        /* HD */
        if (f == 2 || f == 20) sh[dir] = -1;
        else if (f == 3) { sh[dir] = -1; sh[2 + (dir + 1) % 3 - 1] = -1; }
        else if (f == 4) { sh[dir] = -1; sh[2 + dir % 3 - 1] = -1; }
        else if (f == 5) { sh[dir] = -1; sh[2 + (dir + 1) % 3 - 1] = -1; }
        else if (f == 6) sh[dir] = -1;
    }
}

orc_state *orc_create(const orc_cfg *cfg) {
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->c = *cfg;
    s->N = cfg->ix * cfg->jx;
    size_t N = (size_t)s->N, nv = (size_t)cfg->nvar_m;
    s->Um = calloc(N * nv, sizeof(double));
    s->Uh = calloc(N * 5, sizeof(double));
    s->ac = calloc(N, sizeof(double));
    s->gm_rec = calloc(N, sizeof(double));
    s->gm_ion = calloc(N, sizeof(double));
    s->arb_heat = calloc(N, sizeof(double));
    s->ion_pot = calloc(N, sizeof(double));
    s->V = calloc(N * 9, sizeof(double));
    s->VL = calloc(N * 9, sizeof(double));
    s->VR = calloc(N * 9, sizeof(double));
    s->Fm = calloc(N * 9 * 2, sizeof(double));
    s->Fh = calloc(N * 5 * 2, sizeof(double));
    s->Sm = calloc(N * 9, sizeof(double));
    s->Sh = calloc(N * 5, sizeof(double));
    s->Um0 = calloc(N * 9, sizeof(double));
    s->Uh0 = calloc(N * 5, sizeof(double));
    s->tmp = calloc(N * 10, sizeof(double));
    set_sym(s);
    if (!cfg->ac_field)
        for (size_t k = 0; k < N; k++) s->ac[k] = cfg->col;
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    free(s->Um); free(s->Uh); free(s->ac); free(s->gm_rec); free(s->gm_ion);
    free(s->arb_heat); free(s->ion_pot); free(s->V); free(s->VL); free(s->VR);
    free(s->Fm); free(s->Fh); free(s->Sm); free(s->Sh); free(s->Um0);
    free(s->Uh0); free(s->tmp); free(s);
}

void orc_set_state(orc_state *s, const double *U_h, const double *U_m) {
    memcpy(s->Uh, U_h, sizeof(double) * s->N * 5);
    memcpy(s->Um, U_m, sizeof(double) * s->N * s->c.nvar_m);
}
void orc_get_state(const orc_state *s, double *U_h, double *U_m) {
    memcpy(U_h, s->Uh, sizeof(double) * s->N * 5);
    memcpy(U_m, s->Um, sizeof(double) * s->N * s->c.nvar_m);
}
void orc_set_ac(orc_state *s, const double *ac) { memcpy(s->ac, ac, sizeof(double) * s->N); }
void orc_get_ac(const orc_state *s, double *ac) { memcpy(ac, s->ac, sizeof(double) * s->N); }
void orc_get_rates(const orc_state *s, double *r, double *i, double *a, double *p) {
    size_t b = sizeof(double) * s->N;
    if (r) memcpy(r, s->gm_rec, b);
    if (i) memcpy(i, s->gm_ion, b);
    if (a) memcpy(a, s->arb_heat, b);
    if (p) memcpy(p, s->ion_pot, b);
}
void orc_set_rates(orc_state *s, const double *r, const double *i) {
    memcpy(s->gm_rec, r, sizeof(double) * s->N);
    memcpy(s->gm_ion, i, sizeof(double) * s->N);
}
void orc_set_fixed_rates(orc_state *s, int on) { s->fixed_rates = on; }
void orc_set_arb_heat(orc_state *s, const double *a) {
    memcpy(s->arb_heat, a, sizeof(double) * s->N);
    s->arb_set = 1;
}
void orc_get_scalars(const orc_state *s, double *o) {
    o[0] = s->dt; o[1] = s->time; o[2] = s->cmax; o[3] = (double)s->nt;
}

void orc_set_scalars(orc_state *s, double dt, double cmax) {
    s->dt = dt;
    s->cmax = cmax;
}

void orc_set_stage_save(orc_state *s, double *buf) { s->stage_save = buf; }

/* Threads of the elementwise loops (1: serial); returns the previous
 * setting, or 1 when the library was built without OpenMP. */
int orc_set_threads(int n) {
#ifdef _OPENMP
    int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
#else
    (void)n;
    return 1;
#endif
}

/* -------------------------------------------------------------- limiter */
/* minmod_func: Util_rot.f90:59-76 */
double orc_minmod(double U1, double U2, double U3) {
    double X = U3 - U2, Y = U2 - U1;
    double SX = copysign(1.0, X);
    return SX * dmax(0.0, dmin(fabs(X), SX * Y));
}

/* ------------------------------------------------------- conversions */
/* u2v_mhd: HLL_rot.f90:1159-1186 (whole array, all cells) */
static void u2v_mhd(const orc_state *s, const double *U, double *V) {
    const orc_cfg *c = &s->c;
    int N = s->N;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double u1 = U[k], u2 = U[N + k], u3 = U[2 * N + k], u4 = U[3 * N + k];
        double u5 = U[4 * N + k], u6 = U[5 * N + k], u7 = U[6 * N + k], u8 = U[7 * N + k];
        double v1 = dmax(u1, c->ro_lim);
        double roi = 1.0 / v1;
        V[k] = v1;
        V[N + k] = u2 * roi;
        V[2 * N + k] = u3 * roi;
        V[3 * N + k] = u4 * roi;
        double p = (c->gm - 1.0) * ((u5 - (0.5 * ((u2 * u2 + u3 * u3) + u4 * u4)) * roi)
                                    - 0.5 * ((u6 * u6 + u7 * u7) + u8 * u8));
        if (p < c->pr_lim) p = c->pr_lim;
        V[4 * N + k] = p;
        V[5 * N + k] = u6;
        V[6 * N + k] = u7;
        V[7 * N + k] = u8;
        if (c->flag_divb == 1) V[8 * N + k] = U[8 * N + k];
    }
}

/* u2v_hd: HLL_rot.f90:1136-1156 */
static void u2v_hd(const orc_state *s, const double *U, double *V) {
    const orc_cfg *c = &s->c;
    int N = s->N;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double u1 = U[k], u2 = U[N + k], u3 = U[2 * N + k], u4 = U[3 * N + k];
        double u5 = U[4 * N + k];
        double v1 = dmax(u1, c->ro_lim);
        double roi = 1.0 / v1;
        V[k] = v1;
        V[N + k] = u2 * roi;
        V[2 * N + k] = u3 * roi;
        V[3 * N + k] = u4 * roi;
        double p = (c->gm - 1.0) * (u5 - (0.5 * ((u2 * u2 + u3 * u3) + u4 * u4)) * roi);
        if (p < c->pr_lim) p = c->pr_lim;
        V[4 * N + k] = p;
    }
}

/* v2u_1: HLL_rot.f90:1189-1215 */
static void v2u_1(const double *v1, double *u1, int flag0, int divb, double gm) {
    u1[0] = v1[0];
    u1[1] = v1[0] * v1[1];
    u1[2] = v1[0] * v1[2];
    u1[3] = v1[0] * v1[3];
    if (flag0 == 0) {
        u1[4] = v1[4] / (gm - 1.0) + (0.5 * v1[0]) * ((v1[1] * v1[1] + v1[2] * v1[2]) + v1[3] * v1[3]);
    } else {
        u1[5] = v1[5]; u1[6] = v1[6]; u1[7] = v1[7];
        u1[4] = (v1[4] / (gm - 1.0) + (0.5 * v1[0]) * ((v1[1] * v1[1] + v1[2] * v1[2]) + v1[3] * v1[3]))
                + 0.5 * ((v1[5] * v1[5] + v1[6] * v1[6]) + v1[7] * v1[7]);
        if (divb == 1) u1[8] = v1[8];
    }
}

/* cq2pv_mhd: Scheme_rot.f90:153-179 -- REWRITES U_m in place.
 * Outputs de,vx,vy,vz,pr,bx,by,bz into prim[0..7] planes. */
static void cq2pv_mhd(orc_state *s, double *U, double *prim) {
    const orc_cfg *c = &s->c;
    int N = s->N;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double de = U[k];
        double vx = U[N + k] / de, vy = U[2 * N + k] / de, vz = U[3 * N + k] / de;
        double bx = U[5 * N + k], by = U[6 * N + k], bz = U[7 * N + k];
        double pr = (c->gm - 1.0) * ((U[4 * N + k] - (0.5 * de) * ((vx * vx + vy * vy) + vz * vz))
                                     - 0.5 * ((bx * bx + by * by) + bz * bz));
        pr = dmax(pr, c->pr_lim); /* :171 (equivalent to the minval guard) */
        de = dmax(de, c->ro_lim); /* :172 */
        U[k] = de;
        U[N + k] = vx * de;
        U[2 * N + k] = vy * de;
        U[3 * N + k] = vz * de;
        U[4 * N + k] = (pr / (c->gm - 1.0) + (0.5 * de) * ((vx * vx + vy * vy) + vz * vz))
                       + 0.5 * ((bx * bx + by * by) + bz * bz);
        if (prim) {
            prim[k] = de; prim[N + k] = vx; prim[2 * N + k] = vy; prim[3 * N + k] = vz;
            prim[4 * N + k] = pr; prim[5 * N + k] = bx; prim[6 * N + k] = by; prim[7 * N + k] = bz;
        }
    }
}

/* cq2pv_hd: Scheme_rot.f90:132-150 -- REWRITES U_h in place. */
static void cq2pv_hd(orc_state *s, double *U, double *prim) {
    const orc_cfg *c = &s->c;
    int N = s->N;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double de = U[k];
        double vx = U[N + k] / de, vy = U[2 * N + k] / de, vz = U[3 * N + k] / de;
        double pr = (c->gm - 1.0) * (U[4 * N + k] - (0.5 * de) * ((vx * vx + vy * vy) + vz * vz));
        pr = dmax(pr, c->pr_lim);
        de = dmax(de, c->ro_lim);
        U[k] = de;
        U[N + k] = vx * de;
        U[2 * N + k] = vy * de;
        U[3 * N + k] = vz * de;
        U[4 * N + k] = pr / (c->gm - 1.0) + (0.5 * de) * ((vx * vx + vy * vy) + vz * vz);
        if (prim) {
            prim[k] = de; prim[N + k] = vx; prim[2 * N + k] = vy; prim[3 * N + k] = vz;
            prim[4 * N + k] = pr;
        }
    }
}

/* get_Pr_HD (+get_Te_HD): Scheme_rot.f90:83-102 -- rewrites U_h(5). */
static void get_pr_hd(orc_state *s, double *U, double *pr_out, double *te_out) {
    const orc_cfg *c = &s->c;
    int N = s->N;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double u1 = U[k], u2 = U[N + k], u3 = U[2 * N + k], u4 = U[3 * N + k];
        double pr = (c->gm - 1.0) * (U[4 * N + k] - (0.5 * ((u2 * u2 + u3 * u3) + u4 * u4)) / u1);
        pr = dmax(pr, c->pr_lim);
        U[4 * N + k] = pr / (c->gm - 1.0) + 0.5 * (((u2 * u2 + u3 * u3) + u4 * u4) / u1);
        if (pr_out) pr_out[k] = pr;
        if (te_out) te_out[k] = c->gm * (pr / u1);
    }
}

/* get_Pr_MHD (+get_Te_MHD): Scheme_rot.f90:104-128 -- rewrites U_m(5). */
static void get_pr_mhd(orc_state *s, double *U, double *pr_out, double *te_out) {
    const orc_cfg *c = &s->c;
    int N = s->N;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double u1 = U[k], u2 = U[N + k], u3 = U[2 * N + k], u4 = U[3 * N + k];
        double u6 = U[5 * N + k], u7 = U[6 * N + k], u8 = U[7 * N + k];
        double q = (((((u2 * u2 + u3 * u3) + u4 * u4) / u1 + u6 * u6) + u7 * u7) + u8 * u8);
        double pr = (c->gm - 1.0) * (U[4 * N + k] - 0.5 * q);
        pr = dmax(pr, c->pr_lim);
        U[4 * N + k] = pr / (c->gm - 1.0) + 0.5 * q;
        if (pr_out) pr_out[k] = pr;
        if (te_out) te_out[k] = (0.5 * c->gm) * (pr / u1);
    }
}

/* ------------------------------------------------------ reconstruction */
/* get_Side, s_order=2: HLL_rot.f90:15-63 (nvar planes) */
static void get_side(const orc_state *s, const double *V, double *VL, double *VR, int dir, int nvar) {
    const orc_cfg *c = &s->c;
    int N = s->N, ix = c->ix, jx = c->jx;
    memset(VL, 0, sizeof(double) * N * nvar);
    memset(VR, 0, sizeof(double) * N * nvar);
    for (int n = 0; n < nvar; n++) {
        const double *U = V + (size_t)n * N;
        double *L = VL + (size_t)n * N, *R = VR + (size_t)n * N;
        if (dir == 1) {
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
            for (int j = 0; j < jx; j++)
                for (int i = 1; i <= ix - 2; i++) {
                    int k = j * ix + i;
                    double du = orc_minmod(U[k - 1], U[k], U[k + 1]);
                    L[k] = U[k] + 0.5 * du;
                    R[k - 1] = U[k] - 0.5 * du;
                }
        } else {
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
            for (int j = 1; j <= jx - 2; j++)
                for (int i = 0; i < ix; i++) {
                    int k = j * ix + i;
                    double du = orc_minmod(U[k - ix], U[k], U[k + ix]);
                    L[k] = U[k] + 0.5 * du;
                    R[k - ix] = U[k] - 0.5 * du;
                }
        }
    }
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        VL[k] = dmax(VL[k], c->ro_lim);
        VR[k] = dmax(VR[k], c->ro_lim);
    }
}

/* ------------------------------------------------------ Riemann solvers */
/* hllc_flux: HLL_rot.f90:615-687 (nvar=5) */
void orc_hllc_flux(const double *v_l, const double *v_r, double gm, double *flx) {
    double u_l[5], u_r[5], flx_l[5], flx_r[5], st_l[5], st_r[5];
    v2u_1(v_l, u_l, 0, 0, gm);
    v2u_1(v_r, u_r, 0, 0, gm);
    double pr_l = v_l[4], pr_r = v_r[4];
    flx_l[0] = u_l[1]; flx_r[0] = u_r[1];
    flx_l[1] = pr_l + (u_l[1] * u_l[1]) / u_l[0]; flx_r[1] = pr_r + (u_r[1] * u_r[1]) / u_r[0];
    flx_l[2] = u_l[1] * u_l[2] / u_l[0]; flx_r[2] = u_r[1] * u_r[2] / u_r[0];
    flx_l[3] = u_l[1] * u_l[3] / u_l[0]; flx_r[3] = u_r[1] * u_r[3] / u_r[0];
    flx_l[4] = u_l[1] * (u_l[4] + pr_l) / u_l[0]; flx_r[4] = u_r[1] * (u_r[4] + pr_r) / u_r[0];
    double cs_l = sqrt(gm * pr_l / u_l[0]);
    double cs_r = sqrt(gm * pr_r / u_r[0]);
    double vn_l = u_l[1] / u_l[0];
    double vn_r = u_r[1] / u_r[0];
    double sl = ACC * dmin(vn_l - cs_l, vn_r - cs_r);
    double sr = ACC * dmax(vn_l + cs_l, vn_r + cs_r);
    double sm = (u_r[1] * (sr - vn_r) - u_l[1] * (sl - vn_l) + pr_l - pr_r)
                / (u_r[0] * (sr - vn_r) - u_l[0] * (sl - vn_l));
    double pr_m = sm * u_l[0] * (sl - vn_l) - u_l[1] * (sl - vn_l) + pr_l;
    st_l[0] = u_l[0] * (sl - vn_l) / (sl - sm);
    st_r[0] = u_r[0] * (sr - vn_r) / (sr - sm);
    st_l[1] = st_l[0] * sm;
    st_r[1] = st_r[0] * sm;
    st_l[2] = st_l[0] * u_l[2] / u_l[0];
    st_r[2] = st_r[0] * u_r[2] / u_r[0];
    st_l[3] = st_l[0] * u_l[3] / u_l[0];
    st_r[3] = st_r[0] * u_r[3] / u_r[0];
    st_l[4] = ((sl - vn_l) * u_l[4] - pr_l * vn_l + pr_m * sm) / (sl - sm);
    st_r[4] = ((sr - vn_r) * u_r[4] - pr_r * vn_r + pr_m * sm) / (sr - sm);
    double thl = (fabs(sl) + sl) / dmax(2.0 * sl, EPSI);
    double thm = (fabs(sm) + sm) / dmax(2.0 * sm, EPSI);
    double thr = 1.0 - (fabs(sr) + sr) / dmax(2.0 * sr, EPSI);
    for (int n = 0; n < 5; n++)
        flx[n] = thr * flx_r[n] + thl * flx_l[n]
                 + (1.0 - thl - thr) * (thm * (flx_l[n] + sl * (st_l[n] - u_l[n]))
                                        + (1.0 - thm) * (flx_r[n] + sr * (st_r[n] - u_r[n])));
}

/* hlld_flux: HLL_rot.f90:691-910 (nvar = 8 or 9) */
void orc_hlld_flux(const double *v_l, const double *v_r, int nvar, int divb, double cmax,
                   double gm, double *flx) {
    double u_l[9], u_r[9], flx_l[9], flx_r[9], u_hll[9], u_l1[9], u_r1[9], u2[9];
    double f1, f2;
    for (int n = 0; n < 9; n++) { flx[n] = 0.0; flx_l[n] = 0.0; flx_r[n] = 0.0; }
    v2u_1(v_l, u_l, 1, divb, gm);
    v2u_1(v_r, u_r, 1, divb, gm);

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

    /* fast mode speed :736-744 */
    f1 = gm * v_l[4];
    f2 = 4.0 * f1 * (v_l[5] * v_l[5]);
    double vf_l = sqrt(((f1 + b2_l) + sqrt(dmax((f1 + b2_l) * (f1 + b2_l) - f2, 0.0))) / (2.0 * v_l[0]));
    f1 = gm * v_r[4];
    f2 = 4.0 * f1 * (v_r[5] * v_r[5]);
    double vf_r = sqrt(((f1 + b2_r) + sqrt(dmax((f1 + b2_r) * (f1 + b2_r) - f2, 0.0))) / (2.0 * v_r[0]));

    double sl = dmin(v_l[1], v_r[1]) - dmax(vf_l, vf_r);
    double sr = dmax(v_l[1], v_r[1]) + dmax(vf_l, vf_r);

    if (sl >= 0.0) {
        for (int n = 0; n < 8; n++) flx[n] = flx_l[n];
    } else if (sr <= 0.0) {
        for (int n = 0; n < 8; n++) flx[n] = flx_r[n];
    } else {
        f1 = 1.0 / (sr - sl);
        for (int n = 0; n < 8; n++)
            u_hll[n] = f1 * (((sr * u_r[n] - sl * u_l[n]) - flx_r[n]) + flx_l[n]);
        double sm = u_hll[1] / u_hll[0];
        double pt = pt_l + v_l[0] * (sl - v_l[1]) * (sm - v_l[1]);
        double ro_l1 = v_l[0] * (sl - v_l[1]) / (sl - sm);
        double ro_r1 = v_r[0] * (sr - v_r[1]) / (sr - sm);
        double sal = fabs(u_hll[5]) / sqrt(ro_l1);
        double sar = fabs(u_hll[5]) / sqrt(ro_r1);
        if ((sl >= (sm - HLLG * sal)) || ((sm + HLLG * sar) >= sr)) {
            /* revert to HLLC-G :775-807 */
            double vy_2 = u_hll[2] / u_hll[0];
            double vz_2 = u_hll[3] / u_hll[0];
            if (sm >= 0.0) {
                u2[4] = ((((sl - v_l[1]) * u_l[4] - pt_l * v_l[1]) + pt * sm)
                         + u_hll[5] * (((vbL - sm * u_hll[5]) - vy_2 * u_hll[6]) - vz_2 * u_hll[7]))
                        / (sl - sm);
                u2[1] = ro_l1 * sm;
                u2[2] = ro_l1 * vy_2;
                u2[3] = ro_l1 * vz_2;
                flx[0] = sl * (ro_l1 - v_l[0]) + flx_l[0];
                for (int n = 1; n < 5; n++) flx[n] = sl * (u2[n] - u_l[n]) + flx_l[n];
                for (int n = 5; n < 8; n++) flx[n] = sl * (u_hll[n] - v_l[n]) + flx_l[n];
            } else {
                u2[4] = ((((sr - v_r[1]) * u_r[4] - pt_r * v_r[1]) + pt * sm)
                         + u_hll[5] * (((vbR - sm * u_hll[5]) - vy_2 * u_hll[6]) - vz_2 * u_hll[7]))
                        / (sr - sm);
                u2[1] = ro_r1 * sm;
                u2[2] = ro_r1 * vy_2;
                u2[3] = ro_r1 * vz_2;
                flx[0] = sr * (ro_r1 - v_r[0]) + flx_r[0];
                for (int n = 1; n < 5; n++) flx[n] = sr * (u2[n] - u_r[n]) + flx_r[n];
                for (int n = 5; n < 8; n++) flx[n] = sr * (u_hll[n] - v_r[n]) + flx_r[n];
            }
        } else {
            /* HLLD :810-899 */
            double bh2 = u_hll[5] * u_hll[5];
            f1 = 1.0 / (v_l[0] * (sl - v_l[1]) * (sl - sm) - bh2);
            u_l1[5] = u_hll[5];
            u_l1[6] = v_l[6] * f1 * (v_l[0] * ((sl - v_l[1]) * (sl - v_l[1])) - bh2);
            u_l1[7] = v_l[7] * f1 * (v_l[0] * ((sl - v_l[1]) * (sl - v_l[1])) - bh2);
            double vy_l1 = v_l[2] - f1 * u_hll[5] * v_l[6] * (sm - v_l[1]);
            double vz_l1 = v_l[3] - f1 * u_hll[5] * v_l[7] * (sm - v_l[1]);
            u_l1[0] = ro_l1;
            u_l1[4] = ((((sl - v_l[1]) * u_l[4] - pt_l * v_l[1]) + pt * sm)
                       + u_hll[5] * (((vbL - sm * u_hll[5]) - vy_l1 * u_l1[6]) - vz_l1 * u_l1[7]))
                      / (sl - sm);
            u_l1[1] = ro_l1 * sm;
            u_l1[2] = ro_l1 * vy_l1;
            u_l1[3] = ro_l1 * vz_l1;

            f1 = 1.0 / (v_r[0] * (sr - v_r[1]) * (sr - sm) - bh2);
            u_r1[5] = u_hll[5];
            u_r1[6] = v_r[6] * f1 * (v_r[0] * ((sr - v_r[1]) * (sr - v_r[1])) - bh2);
            u_r1[7] = v_r[7] * f1 * (v_r[0] * ((sr - v_r[1]) * (sr - v_r[1])) - bh2);
            double vy_r1 = v_r[2] - f1 * u_hll[5] * v_r[6] * (sm - v_r[1]);
            double vz_r1 = v_r[3] - f1 * u_hll[5] * v_r[7] * (sm - v_r[1]);
            u_r1[0] = ro_r1;
            u_r1[4] = ((((sr - v_r[1]) * u_r[4] - pt_r * v_r[1]) + pt * sm)
                       + u_hll[5] * (((vbR - sm * u_hll[5]) - vy_r1 * u_r1[6]) - vz_r1 * u_r1[7]))
                      / (sr - sm);
            u_r1[1] = ro_r1 * sm;
            u_r1[2] = ro_r1 * vy_r1;
            u_r1[3] = ro_r1 * vz_r1;

            double sl1 = sm - sal;
            double sr1 = sm + sar;
            if (sl1 >= 0.0) {
                flx[0] = sl * (u_l1[0] - v_l[0]) + flx_l[0];
                for (int n = 1; n < 5; n++) flx[n] = sl * (u_l1[n] - u_l[n]) + flx_l[n];
                for (int n = 5; n < 8; n++) flx[n] = sl * (u_l1[n] - v_l[n]) + flx_l[n];
            } else if (sr1 <= 0.0) {
                flx[0] = sr * (u_r1[0] - v_r[0]) + flx_r[0];
                for (int n = 1; n < 5; n++) flx[n] = sr * (u_r1[n] - u_r[n]) + flx_r[n];
                for (int n = 5; n < 8; n++) flx[n] = sr * (u_r1[n] - v_r[n]) + flx_r[n];
            } else {
                double ro_ls = sqrt(ro_l1);
                double ro_rs = sqrt(ro_r1);
                f1 = 1.0 / (ro_ls + ro_rs);
                f2 = copysign(1.0, u_hll[5]);
                double vy_2 = f1 * ((ro_ls * vy_l1 + ro_rs * vy_r1) + (u_r1[6] - u_l1[6]) * f2);
                double vz_2 = f1 * ((ro_ls * vz_l1 + ro_rs * vz_r1) + (u_r1[7] - u_l1[7]) * f2);
                u2[5] = u_hll[5];
                u2[6] = f1 * ((ro_ls * u_r1[6] + ro_rs * u_l1[6]) + ro_ls * ro_rs * (vy_r1 - vy_l1) * f2);
                u2[7] = f1 * ((ro_ls * u_r1[7] + ro_rs * u_l1[7]) + ro_ls * ro_rs * (vz_r1 - vz_l1) * f2);
                if (sm >= 0.0) {
                    u2[0] = ro_l1;
                    u2[4] = u_l1[4] - ro_ls * (((vy_l1 * u_l1[6] + vz_l1 * u_l1[7]) - vy_2 * u2[6]) - vz_2 * u2[7]) * f2;
                    u2[1] = ro_l1 * sm;
                    u2[2] = ro_l1 * vy_2;
                    u2[3] = ro_l1 * vz_2;
                    flx[0] = ((sl1 * u2[0] - (sl1 - sl) * u_l1[0]) - sl * v_l[0]) + flx_l[0];
                    for (int n = 1; n < 5; n++)
                        flx[n] = ((sl1 * u2[n] - (sl1 - sl) * u_l1[n]) - sl * u_l[n]) + flx_l[n];
                    for (int n = 5; n < 8; n++)
                        flx[n] = ((sl1 * u2[n] - (sl1 - sl) * u_l1[n]) - sl * v_l[n]) + flx_l[n];
                } else {
                    u2[0] = ro_r1;
                    u2[4] = u_r1[4] + ro_rs * (((vy_r1 * u_r1[6] + vz_r1 * u_r1[7]) - vy_2 * u2[6]) - vz_2 * u2[7]) * f2;
                    u2[1] = ro_r1 * sm;
                    u2[2] = ro_r1 * vy_2;
                    u2[3] = ro_r1 * vz_2;
                    flx[0] = ((sr1 * u2[0] - (sr1 - sr) * u_r1[0]) - sr * v_r[0]) + flx_r[0];
                    for (int n = 1; n < 5; n++)
                        flx[n] = ((sr1 * u2[n] - (sr1 - sr) * u_r1[n]) - sr * u_r[n]) + flx_r[n];
                    for (int n = 5; n < 8; n++)
                        flx[n] = ((sr1 * u2[n] - (sr1 - sr) * u_r1[n]) - sr * v_r[n]) + flx_r[n];
                }
            }
        }
    }
    /* divB cleaning :904-908 */
    if (divb == 1 && nvar >= 9) {
        double ch = cmax;
        flx[5] = 0.5 * ((u_l[8] + u_r[8]) - ch * (v_r[5] - v_l[5]));
        flx[8] = 0.5 * (ch * ch * (v_l[5] + v_r[5]) - ch * (v_r[8] - v_l[8]));
    }
}

/* ------------------------------------------------------------- fluxes */
/* hlld_fluxes_ideal + set_HLLD_fluxes: HLL_rot.f90:109-121, 345-461.
 * F layout: F[(dir-1)*nv*N + n*N + k], F(i) = face i+1/2. */
static void hlld_fluxes_ideal(orc_state *s, double *F) {
    const orc_cfg *c = &s->c;
    int N = s->N, ix = c->ix, jx = c->jx, nv = c->nvar_m;
    memset(F, 0, sizeof(double) * (size_t)N * nv * 2);
    for (int dir = 1; dir <= c->ndim; dir++) {
        u2v_mhd(s, s->Um, s->V);
        get_side(s, s->V, s->VL, s->VR, dir, nv);
        double *Fd = F + (size_t)(dir - 1) * nv * N;
        if (dir == 1) {
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
            for (int j = 0; j < jx; j++)
                for (int i = 0; i <= ix - 2; i++) {
                    int k = j * ix + i;
                    double vl[9], vr[9], fl[9];
                    for (int n = 0; n < nv; n++) { vl[n] = s->VL[n * N + k]; vr[n] = s->VR[n * N + k]; }
                    orc_hlld_flux(vl, vr, nv, c->flag_divb, s->cmax, c->gm, fl);
                    for (int n = 0; n < nv; n++) Fd[n * N + k] = fl[n];
                }
        } else {
            static const int perm[9] = {0, 2, 1, 3, 4, 6, 5, 7, 8};
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
            for (int j = 0; j <= jx - 2; j++)
                for (int i = 0; i < ix; i++) {
                    int k = j * ix + i;
                    double vl[9], vr[9], fl[9];
                    for (int n = 0; n < nv; n++) {
                        vl[n] = s->VL[perm[n] * N + k];
                        vr[n] = s->VR[perm[n] * N + k];
                    }
                    orc_hlld_flux(vl, vr, nv, c->flag_divb, s->cmax, c->gm, fl);
                    for (int n = 0; n < nv; n++) Fd[perm[n] * N + k] = fl[n];
                }
        }
    }
}

/* hllc_fluxes_ideal + set_HLLC_fluxes: HLL_rot.f90:94-107, 295-343 */
static void hllc_fluxes_ideal(orc_state *s, double *F) {
    const orc_cfg *c = &s->c;
    int N = s->N, ix = c->ix, jx = c->jx;
    memset(F, 0, sizeof(double) * (size_t)N * 5 * 2);
    for (int dir = 1; dir <= c->ndim; dir++) {
        u2v_hd(s, s->Uh, s->V);
        get_side(s, s->V, s->VL, s->VR, dir, 5);
        double *Fd = F + (size_t)(dir - 1) * 5 * N;
        static const int perm[5] = {0, 2, 1, 3, 4};
        int imax = (dir == 1) ? ix - 2 : ix - 1;
        int jmax = (dir == 1) ? jx - 1 : jx - 2;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
        for (int j = 0; j <= jmax; j++)
            for (int i = 0; i <= imax; i++) {
                int k = j * ix + i;
                double vl[5], vr[5], fl[5];
                for (int n = 0; n < 5; n++) {
                    int p = (dir == 1) ? n : perm[n];
                    vl[n] = s->VL[p * N + k];
                    vr[n] = s->VR[p * N + k];
                }
                orc_hllc_flux(vl, vr, c->gm, fl);
                for (int n = 0; n < 5; n++) {
                    int p = (dir == 1) ? n : perm[n];
                    Fd[p * N + k] = fl[n];
                }
            }
    }
}

/* set_flux, flag_sch=0 / flag_hll=2: Solver_rot.f90:124-159 */
void orc_set_flux(orc_state *s, double *F_h, double *F_m) {
    hlld_fluxes_ideal(s, F_m ? F_m : s->Fm);
    hllc_fluxes_ideal(s, F_h ? F_h : s->Fh);
}

/* --------------------------------------------------------- boundaries */
/* boundary_control: Boundary_rot.f90:323-402, for one side of one fluid */
static void boundary_control(orc_state *s, double *U, int nvar, int dir, int ul,
                             const int *sym, int type) {
    const orc_cfg *c = &s->c;
    int ix = c->ix, jx = c->jx, N = s->N, m = 2;
    int ig = (dir == 1) ? ix : jx;
    /* ghost g (1..m, Fortran) <- source index per type */
    for (int g = 1; g <= m; g++) {
        int dst, src;
        if (ul == 0) {
            dst = g;
            if (type == 1) src = ig - 2 * m + g;
            else if (type == 2) src = 2 * m + 1 - g;
            else if (type == 10) src = m + 1;
            else return;
        } else {
            dst = ig - m + g;
            if (type == 1) src = m + g;
            else if (type == 2) src = ig - m + 1 - g;
            else if (type == 10) src = ig - m;
            else return;
        }
        int use_sym = (type != 10);
        for (int n = 0; n < nvar; n++) {
            double sg = use_sym ? (double)sym[n] : 1.0;
            double *P = U + (size_t)n * N;
            if (dir == 1) {
                for (int j = 0; j < jx; j++) P[j * ix + dst - 1] = sg * P[j * ix + src - 1];
            } else {
                for (int i = 0; i < ix; i++) P[(dst - 1) * ix + i] = sg * P[(src - 1) * ix + i];
            }
        }
    }
}

static int bnd_type(int f) {
    if (f >= 3 && f <= 7) return 2;
    if (f == 11) return 2;
    return f;
}

/* PIPbnd -> MHDbnd, HDbnd (single rank, neighbor=-1): Boundary_rot.f90:34-125 */
void orc_pipbnd(orc_state *s) {
    const orc_cfg *c = &s->c;
    for (int i = 1; i <= 2 * c->ndim; i++) {
        int dir = (i + 1) / 2, ul = (i + 1) % 2;
        boundary_control(s, s->Um, c->nvar_m, dir, ul, s->sym_m[i - 1], bnd_type(c->flag_bnd[i - 1]));
    }
    for (int i = 1; i <= 2 * c->ndim; i++) {
        int dir = (i + 1) / 2, ul = (i + 1) % 2;
        boundary_control(s, s->Uh, 5, dir, ul, s->sym_h[i - 1], bnd_type(c->flag_bnd[i - 1]));
    }
}

/* ------------------------------------------------------------ set_IR */
/* set_IR case 3 (Popescu+2019 empirical rates): PIP_rot.f90:178-229 */
void orc_set_ir(orc_state *s) {
    const orc_cfg *c = &s->c;
    int N = s->N;
    if (c->flag_ir == 0) return;
    double *Te_n = s->tmp, *Te_p = s->tmp + N;
    get_pr_hd(s, s->Uh, NULL, Te_n);  /* get_Te_HD  :182 */
    get_pr_mhd(s, s->Um, NULL, Te_p); /* get_Te_MHD :183 */
    if (s->fixed_rates) return;
    if (c->flag_ir != 3) return;
    double Te_0 = c->T0 / SGL(1.1604e4);
    double rec_fac = SGL(2.6e-19) * (c->n0 * SGL(1.0e6)) / sqrt(Te_0);
    double ioneq = (SGL(2.6e-19) / sqrt(Te_0))
                   / (SGL(2.91e-14) / (0.232 + 13.6 / Te_0) * orc_pow(13.6 / Te_0, 0.39) * orc_exp(-13.6 / Te_0));
    double f_n = ioneq / (ioneq + 1.0);
    double f_p = 1.0 - f_n;
    double f_p_p = 2.0 * f_p / (f_n + 2.0 * f_p);
    double tfac;
    if (c->flag_col % 2 == 1) tfac = 0.5 * f_p_p / f_p;
    else tfac = c->beta / 2.0 * f_p_p * 5.0 / 6.0 / f_p;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double rho = s->Um[k];
        s->gm_rec[k] = rho / sqrt(Te_p[k]) * c->t_ir / f_p * sqrt(tfac);
        double a = 13.6 / Te_0 / Te_p[k] * tfac;
        double g = SGL(2.91e-14) * (c->n0 * SGL(1.0e6)) * rho * orc_exp(-a) * orc_pow(a, 0.39);
        s->gm_ion[k] = g / (0.232 + a) / rec_fac / f_p * c->t_ir;
    }
    if (!s->arb_set) { /* first call only: nout==0 .and. flag_restart==-1 (:219-229) */
        for (int k = 0; k < N; k++) {
            if (c->flag_col % 2 == 1)
                s->arb_heat[k] = s->gm_ion[k] * s->Uh[k] * (13.6 / c->gm / c->T0 / SGL(8.6173e-5));
            else
                s->arb_heat[k] = s->gm_ion[k] * s->Uh[k] * (13.6 * c->beta / c->T0 / 2.0 / SGL(8.6173e-5));
        }
        s->arb_set = 1;
    }
}

/* set_collisional: PIP_rot.f90:44-97, col_type = flag_col-1 (:21-25).
 * Cases 1..6 take T_n, T_p from get_Te_* (which rewrite E on every cell)
 * and the drift v_n - v_p from get_vel_diff (Scheme_rot.f90:988-999). */
static void set_collisional(orc_state *s) {
    const orc_cfg *c = &s->c;
    int N = s->N;
    int ct = (c->flag_col >= 2 && c->flag_col <= 7) ? c->flag_col - 1 : 0;
    if (ct == 0) {
        if (!c->ac_field)
            for (int k = 0; k < N; k++) s->ac[k] = c->col; /* case 0 */
        return;
    }
    double *Te_m = s->tmp, *Te_h = s->tmp + N;
    get_pr_mhd(s, s->Um, NULL, Te_m); /* get_Te_MHD :57 */
    get_pr_hd(s, s->Uh, NULL, Te_h);  /* get_Te_HD  :58 */
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double tt = Te_h[k] + Te_m[k];
        double a = c->col * sqrt(tt / 2.0);
        if (ct == 1) a = a / sqrt(c->beta / 2.0 * c->gm);
        else if (ct >= 3) {
            double vd[3];
            for (int i = 0; i < 3; i++)
                vd[i] = s->Uh[(i + 1) * N + k] / s->Uh[k] - s->Um[(i + 1) * N + k] / s->Um[k];
            double v2 = (vd[0] * vd[0] + vd[1] * vd[1]) + vd[2] * vd[2];
            a = a * sqrt(1.0 + 9.0 * PI / 64.0 * c->gm / 2.0 * v2 / tt);
            if (ct == 3) a = a / sqrt(c->beta / 2.0 * c->gm);
            else if (ct == 5)
                a = a / sqrt(c->beta / 2.0 * c->gm) *
                    orc_pow(c->beta / 2.0 * c->gm / tt / 2.0 + c->gm * PI / 16.0 * v2, 0.125);
            else if (ct == 6)
                a = a / orc_pow(tt / 2.0 + c->gm * PI / 16.0 * v2, 0.125);
        }
        s->ac[k] = a;
    }
}

/* set_coefficients: Solver_rot.f90:95-121 */
void orc_set_coefficients(orc_state *s, int flag) {
    const orc_cfg *c = &s->c;
    if (c->flag_col >= 2 || (flag == 0 && c->flag_col == 1)) set_collisional(s);
    if (c->flag_ir >= 2 || (flag == 0 && c->flag_ir == 1)) orc_set_ir(s);
}

/* ---------------------------------------------------------------- cfl */
/* cfl + cfl_mhd + cfl_hd + cfl_pn_col + cfl_pip_ir: Scheme_rot.f90:183-327.
 * into_global_dt=1 mimics `call cfl(U_h,U_m,dt)` where the dummy aliases the
 * global dt that the cmax line then reads (:208). */
double orc_cfl(orc_state *s, int into_global_dt) {
    const orc_cfg *c = &s->c;
    int N = s->N, ix = c->ix, jx = c->jx;
    double *P = s->tmp; /* 8 planes */
    double dx = c->dx, dy = (c->ndim >= 2) ? c->dy : 100.0, dz = 100.0;
    double gmin = dmin(dmin(dx, dy), dz);
    int jlo = (c->ndim >= 2) ? 2 : 0, jhi = (c->ndim >= 2) ? jx - 3 : 0;
    /* cfl_mhd :245-289 */
    cq2pv_mhd(s, s->Um, P);
    double dt_min = 10.0;
    for (int j = jlo; j <= jhi; j++)
        for (int i = 2; i <= ix - 3; i++) {
            int k = j * ix + i;
            double de = P[k], vx = P[N + k], vy = P[2 * N + k], vz = P[3 * N + k], pr = P[4 * N + k];
            double bx = P[5 * N + k], by = P[6 * N + k], bz = P[7 * N + k];
            double cs2 = c->gm * pr / dmax(de, TINY);
            double valf2 = ((bx * bx + by * by) + bz * bz) / dmax(de, TINY);
            double vabs = sqrt((vx * vx + vy * vy) + vz * vz);
            dt_min = dmin(dt_min, c->safety * gmin / (vabs + sqrt(cs2 + valf2)));
        }
    double dttemp = dt_min;
    /* cfl_hd :215-243 */
    cq2pv_hd(s, s->Uh, P);
    dt_min = dttemp;
    for (int j = jlo; j <= jhi; j++)
        for (int i = 2; i <= ix - 3; i++) {
            int k = j * ix + i;
            double de = P[k], vx = P[N + k], vy = P[2 * N + k], vz = P[3 * N + k], pr = P[4 * N + k];
            double cs = sqrt(c->gm * pr / de);
            double vabs = sqrt(dmax((vx * vx + vy * vy) + vz * vz, TINY));
            dt_min = dmin(dt_min, c->safety * gmin / (vabs + cs));
        }
    dttemp = dt_min;
    /* cfl_pn_col :305-313 (all cells, ghosts included) */
    double mm = -HUGE_VAL, mh = -HUGE_VAL;
    for (int k = 0; k < N; k++) {
        mm = dmax(mm, s->Um[k] * s->ac[k]);
        mh = dmax(mh, s->Uh[k] * s->ac[k]);
    }
    if (c->flag_pip_imp == 1) dttemp = dmin(dttemp, c->pip_imp_factor / dmax(mm, mh));
    else dttemp = dmin(dttemp, dmin(c->safety, 0.3) / dmax(mm, mh));
    /* cfl_pip_ir :315-327 */
    if (c->flag_ir >= 1) {
        double mr = -HUGE_VAL, mi = -HUGE_VAL;
        for (int k = 0; k < N; k++) { mr = dmax(mr, s->gm_rec[k]); mi = dmax(mi, s->gm_ion[k]); }
        dttemp = dmin(dttemp, dmin(c->safety, 0.1) / dmax(dmax(mr, mi), 1.0e-5));
        if (c->flag_ir_type == 0) {
            double *pr = s->tmp + 8 * (size_t)N;
            get_pr_mhd(s, s->Um, pr, NULL);
            /* ion_pot is only defined after the first source_PIP call (it is
             * never allocated for IR case 3 before that); before it, the
             * empty maxval leaves the 1e-5 floor in charge. */
            if (s->ion_pot_set) {
                double mx = -HUGE_VAL;
                for (int k = 0; k < N; k++) mx = dmax(mx, s->ion_pot[k] / (pr[k] / (c->gm - 1.0)));
                dttemp = dmin(dttemp, c->safety / dmax(mx, 1.0e-5));
            } else {
                dttemp = dmin(dttemp, c->safety / 1.0e-5);
            }
        }
    }
    /* mpi_double_interface(dttemp,1): single rank -> identity */
    if (into_global_dt) s->dt = dttemp;
    if (c->flag_divb >= 1) s->cmax = c->safety * dmin(dmin(dx, dy), dz) / s->dt;
    return dttemp;
}

/* ----------------------------------------------------------- sources */
/* source_PIP, explicit branch (flag_pip_imp=0): PIP_rot.f90:1309-1429.
 * Rewrites U_h, U_m via cq2pv (:1328-1329).  S_h,S_m: 5 planes each. */
void orc_source_pip(orc_state *s, double *S_h, double *S_m) {
    const orc_cfg *c = &s->c;
    int N = s->N, nv = c->nvar_m;
    double *ph = s->tmp, *pm = s->tmp + 5 * (size_t)N; /* 5 + (>=5) planes */
    double *pmx = s->V; /* use V for the 8 mhd prim planes */
    (void)pm;
    memset(S_h, 0, sizeof(double) * N * 5);
    memset(S_m, 0, sizeof(double) * N * nv);
    cq2pv_hd(s, s->Uh, ph);
    cq2pv_mhd(s, s->Um, pmx);
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
    for (int k = 0; k < N; k++) {
        double nde = ph[k], nvx = ph[N + k], nvy = ph[2 * N + k], nvz = ph[3 * N + k], npr = ph[4 * N + k];
        double de = pmx[k], vx = pmx[N + k], vy = pmx[2 * N + k], vz = pmx[3 * N + k], pr = pmx[4 * N + k];
        double te = pr / de * c->gm * MU_P;
        double nte = npr / nde * c->gm * MU_N;
        double ac = s->ac[k];
        double um1 = s->Um[k], uh1 = s->Uh[k];
        double dS[5];
        dS[0] = 0.0;
        if (c->flag_pip_imp == 1) {
            /* implicit exchange (:1333-1358): integrated over dt_coll_i and
             * applied to the stage base U0, not to S */
            double dtc = s->dt_coll_i;
            double temp = dtc * (uh1 + um1);
            double ex = orc_exp(-ac * temp);
            dS[1] = (um1 * s->Uh[N + k] - uh1 * s->Um[N + k]) * (1.0 - ex) / temp;
            dS[2] = (um1 * s->Uh[2 * N + k] - uh1 * s->Um[2 * N + k]) * (1.0 - ex) / temp;
            dS[3] = (um1 * s->Uh[3 * N + k] - uh1 * s->Um[3 * N + k]) * (1.0 - ex) / temp;
            double lambda = ac * (nde + de);
            double kapper = 3.0 * (c->gm - 1.0) * ac * (MU_P * nde + MU_N * de);
            double B = -1.0 / 3.0 * c->gm * kapper / ((nde + de) * (kapper - lambda)) *
                       (de * ((nvx - vx) * vx + (nvy - vy) * vy + (nvz - vz) * nvz) +
                        nde * ((nvx - vx) * nvx + (nvy - vy) * nvy + (nvz - vz) * nvz));
            double D = -c->gm / 6.0 * kapper * (de - nde) / ((de + nde) * (kapper - 2.0 * lambda)) *
                       (((nvx - vx) * (nvx - vx) + (nvy - vy) * (nvy - vy)) + (nvz - vz) * (nvz - vz));
            double ek = orc_exp(-kapper * dtc);
            dS[4] = -(3.0 / c->gm) * ac * de * nde / kapper *
                    ((((ek - 1.0) * (nte - te) - (B + D) * ek) + B * orc_exp(-lambda * dtc)) +
                     D * orc_exp(-2.0 * lambda * dtc)) / dtc;
            for (int n = 0; n < 5; n++) {
                s->Uh0[n * N + k] = s->Uh0[n * N + k] - dtc * dS[n];
                s->Um0[n * N + k] = s->Um0[n * N + k] + dtc * dS[n];
            }
        } else {
            dS[1] = ac * (um1 * s->Uh[N + k] - uh1 * s->Um[N + k]);
            dS[2] = ac * (um1 * s->Uh[2 * N + k] - uh1 * s->Um[2 * N + k]);
            dS[3] = ac * (um1 * s->Uh[3 * N + k] - uh1 * s->Um[3 * N + k]);
            dS[4] = ac * nde * de * (0.5 * ((nvx * nvx - vx * vx) + (nvy * nvy - vy * vy) + (nvz * nvz - vz * vz))
                                     + 3.0 / c->gm / 2.0 * (nte - te));
            for (int n = 0; n < 5; n++) {
                S_h[n * N + k] = S_h[n * N + k] - dS[n];
                S_m[n * N + k] = S_m[n * N + k] + dS[n];
            }
        }
        if (c->flag_ir >= 1) {
            double gr = s->gm_rec[k], gi = s->gm_ion[k];
            double ds[5];
            ds[0] = gr * de - gi * nde;
            ds[1] = gr * de * vx - gi * nde * nvx;
            ds[2] = gr * de * vy - gi * nde * nvy;
            ds[3] = gr * de * vz - gi * nde * nvz;
            if (c->flag_ir_type == 2)
                ds[4] = 0.5 * (gr * de * ((vx * vx + vy * vy) + vz * vz) - gi * nde * ((nvx * nvx + nvy * nvy) + nvz * nvz))
                        - gi * npr / (c->gm - 1.0) + gr * pr / (c->gm - 1.0);
            else
                ds[4] = 0.5 * (gr * de * ((vx * vx + vy * vy) + vz * vz) - gi * nde * ((nvx * nvx + nvy * nvy) + nvz * nvz))
                        - gi * npr / (c->gm - 1.0) + 0.5 * gr * pr / (c->gm - 1.0);
            for (int n = 0; n < 5; n++) {
                S_h[n * N + k] = S_h[n * N + k] + ds[n];
                S_m[n * N + k] = S_m[n * N + k] - ds[n];
            }
            if (c->flag_ir_type == 0) {
                double ip;
                if (c->flag_col % 2 == 1) ip = gi * nde * (13.6 / c->gm / c->T0 / SGL(8.6173e-5));
                else ip = gi * nde * (13.6 * c->beta / c->T0 / 2.0 / SGL(8.6173e-5));
                s->ion_pot[k] = ip;
                S_m[4 * N + k] = S_m[4 * N + k] - ip + s->arb_heat[k];
            }
        }
    }
    if (c->flag_ir >= 1 && c->flag_ir_type == 0) s->ion_pot_set = 1;
}

/* -------------------------------------------------------- integrator */
/* add_flux_ssprk (HLL branch): Solver_rot.f90:410-495; ig(dir)=ix/jx for a
 * single rank (MPI_rot.f90:32-38). */
static void add_flux_ssprk(orc_state *s, const double *F, const double *U0, double *U, int nvar,
                           double dt, double c0, double c1, int istep) {
    const orc_cfg *c = &s->c;
    int N = s->N, ix = c->ix, jx = c->jx;
    const double *Fx = F, *Fy = F + (size_t)nvar * N;
    double a = dt / c->dx;
    for (int n = 0; n < nvar; n++) {
        const double *f = Fx + (size_t)n * N, *u0 = U0 + (size_t)n * N;
        double *u = U + (size_t)n * N;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
        for (int j = 0; j < jx; j++)
            for (int i = 1; i <= ix - 1; i++) {
                int k = j * ix + i;
                if (istep == 1) u[k] = u0[k] - a * (f[k] - f[k - 1]);
                else u[k] = c0 * u0[k] + c1 * u[k] - a * (f[k] - f[k - 1]);
            }
    }
    if (c->ndim >= 2) {
        double b = dt / c->dy;
        for (int n = 0; n < nvar; n++) {
            const double *f = Fy + (size_t)n * N;
            double *u = U + (size_t)n * N;
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
            for (int j = 1; j <= jx - 1; j++)
                for (int i = 0; i < ix; i++) {
                    int k = j * ix + i;
                    u[k] = u[k] - b * (f[k] - f[k - ix]);
                }
        }
    }
}

/* run_solver: Solver_rot.f90:42-81 with time_integral SSPRK branch :222-245 */
void orc_run_solver(orc_state *s) {
    const orc_cfg *c = &s->c;
    int N = s->N, nv = c->nvar_m;
    memcpy(s->Um0, s->Um, sizeof(double) * N * nv);
    memcpy(s->Uh0, s->Uh, sizeof(double) * N * 5);
    for (int istep = 1; istep <= c->t_order; istep++) {
        /* set_dt_coll (:83-93): dt/2 per stage for t_order=2, dt otherwise */
        s->dt_coll_i = (c->t_order == 2) ? s->dt * 0.5 : s->dt;
        orc_set_coefficients(s, 1);
        orc_set_flux(s, s->Fh, s->Fm);
        orc_source_pip(s, s->Sh, s->Sm);
        double c0, c1, dt_sub;
        if (istep == 1) { c1 = 1.0; c0 = 1.0; dt_sub = s->dt; }
        else {
            c1 = (double)((float)(istep - 1) / (float)(2 * c->t_order - istep)); /* float() :228 */
            c0 = 1.0 - c1;
            dt_sub = c1 * s->dt;
        }
        add_flux_ssprk(s, s->Fh, s->Uh0, s->Uh, 5, dt_sub, c0, c1, istep);
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
        for (int k = 0; k < N * 5; k++) s->Uh[k] = s->Uh[k] + s->Sh[k] * dt_sub; /* add_source :498 */
        add_flux_ssprk(s, s->Fm, s->Um0, s->Um, nv, dt_sub, c0, c1, istep);
#pragma omp parallel for schedule(static) if (N > PAR_MIN)
        for (int k = 0; k < N * nv; k++) s->Um[k] = s->Um[k] + s->Sm[k] * dt_sub;
        if (c->flag_divb == 1) { /* divb_cleaning_Dedner_source Scheme_rot.f90:546-554 */
            double f = orc_exp(-dt_sub / s->dt * c->safety * c->db_clean);
            for (int k = 0; k < N; k++) s->Um[8 * N + k] = s->Um[8 * N + k] * f;
        }
        orc_pipbnd(s); /* PIPbnd(U_h,U_m,istep) :70 */
        if (istep == 1 && s->stage_save) {
            memcpy(s->stage_save, s->Um, sizeof(double) * N * nv);
            memcpy(s->stage_save + (size_t)N * nv, s->Uh, sizeof(double) * N * 5);
        }
    }
    orc_pipbnd(s); /* post_step is a no-op for HLL; PIPbnd(.,0) :77 */
}

/* prologue tail: Initial_rot.f90:140-160 (IC ghosts filled by the boundary
 * map first; see DESIGN notes in the repository documentation). */
void orc_prologue(orc_state *s) {
    s->time = 0.0; s->nt = 0; s->dt = 0.0;
    orc_pipbnd(s);
    orc_set_coefficients(s, 0);
    orc_cfl(s, 1);
    orc_pipbnd(s);
}

/* stop_sim: IO_rot.f90:625-674 (flag_mhd=1 -> only U_m is scanned) */
int orc_stop_sim(const orc_state *s) {
    int st = 0;
    if (s->dt <= 1.0e-12) st |= 1;
    for (size_t k = 0; k < (size_t)s->N * s->c.nvar_m; k++)
        if (s->Um[k] * 0.0 != 0.0) { st |= 2; break; }
    return st;
}

/* one pass of the time loop: main.f90:31-62 */
int orc_step(orc_state *s) {
    const orc_cfg *c = &s->c;
    if (c->flag_ir >= 1) orc_set_ir(s);
    orc_cfl(s, 1);
    orc_run_solver(s);
    if (c->flag_ir >= 1) orc_set_ir(s);
    orc_cfl(s, 0); /* dtnew: unused, side effects only */
    s->time = s->time + s->dt;
    s->nt++;
    return orc_stop_sim(s);
}
