/*
 * pipio.h -- C-ABI of libpipio.so: snapshots of the two-fluid state in the
 * HDF5 layout the reference's output module writes and its restart reads
 * (t####.h5, IO_rot.f90:39-110 output, :140-360 datasets, :405-603 restart).
 * Single process: one file holds the whole padded domain, ghosts included,
 * as the reference's single-rank run writes it.
 *
 * Layout of a file:
 *   xgrid, ygrid, zgrid, dx, dy, dz   1-D, ix / jx / kx doubles
 *   time                              scalar double
 *   3-D fields                        dataspace (kx, jx, ix) in C order, i.e.
 *                                     the Fortran array U(ix,jx,kx) as it
 *                                     lies in memory
 *   plasma:   ro_p mx_p my_p mz_p en_p bx by bz [ps]
 *   neutrals: ro_n mx_n my_n mz_n en_n
 *   per-cell: ac, xi, col (= ac), ion, rec, aheat by the rules of
 *             def_varfiles / save_varfiles (IO_rot.f90:231-330)
 * Every call returns 0 on success or a negative code; pipio_last_error()
 * returns the message of the last failure on the calling thread.
 */
#ifndef PIPIO_H
#define PIPIO_H

#ifdef __cplusplus
extern "C" {
#endif

/* What is written: sizes, the flags that select datasets, and the
 * setting.txt *_sav keys (0 = write, the reference's default). */
typedef struct pipio_layout {
    int ix, jx, kx;     /* padded sizes (kx = 1)                          */
    int nvar_m;         /* 8, or 9 with the GLM psi                       */
    int flag_divb;      /* 1: psi written as 'ps' (when ps_sav = 0)        */
    int flag_col;       /* 1: 'col' (= ac) written                         */
    int flag_ir;        /* >= 1: 'ion' and 'rec' written                   */
    int flag_ir_type;   /* 0 with flag_ir != 0: 'aheat' written            */
    int ac_sav, xi_sav, col_sav, ion_sav, rec_sav, heat_sav, ps_sav;
} pipio_layout;

/* One snapshot.  Arrays are host memory; per-cell fields are (ix,jx,kx)
 * column-major and may be NULL (not written). */
typedef struct pipio_fields {
    double time;
    const double *x, *y, *z, *dx, *dy, *dz;
    const double *U_h;  /* (ix,jx,kx,5)      */
    const double *U_m;  /* (ix,jx,kx,nvar_m) */
    const double *ac, *xi, *gm_ion, *gm_rec, *arb_heat;
} pipio_fields;

/* output (IO_rot.f90:39-110): writes <outdir>/t<nout as %04d>.h5 */
int pipio_write(const char *outdir, int nout, const pipio_layout *layout, const pipio_fields *f);

/* restart (IO_rot.f90:405-603): reads <indir>/t<nout>.h5 into the non-NULL
 * outputs.  U_h, U_m and time are required datasets; ac, ion, rec and aheat
 * are read when their output is non-NULL (the reference reads them for
 * flag_col >= 1 and flag_IR >= 1). */
int pipio_read(const char *indir, int nout, const pipio_layout *layout, double *time, double *x, double *y,
               double *z, double *dx, double *dy, double *dz, double *U_h, double *U_m, double *ac,
               double *gm_ion, double *gm_rec, double *arb_heat);

const char *pipio_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
