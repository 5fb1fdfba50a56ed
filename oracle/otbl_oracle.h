/*
 * OpenTBL RK3 substep -- CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C restatement of the reference Fortran hot path
 * (rhsp.F90, pois.f90, lip.f90, coeft.f90, boun.f90, genflu.F90, ffts.f90,
 * change.f90, bl_1.F90, stotf.F90:ministats).  It exists to check the HIP
 * product; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and never as the thing measured or shipped.
 *
 * Memory layout (same as the reference's planes):
 *   a complex field (0:nz2, R, 1:nx) is stored as doubles with the real index
 *   kr = 0..nz1-1 fastest (kr = 2k -> Re mode k, 2k+1 -> Im mode k), then the
 *   row j = 1..R, then the plane i = 1..nx.  R = ny+1 for u, w (u-type) and
 *   R = ny for v, p (v-type).
 */
#ifndef OTBL_ORACLE_H
#define OTBL_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_box oc_box;

/* Create a box.  y has ny+2 entries, y[0..ny+1] (reference y(0:ny+1)).
 * xout is the recycling plane (1-based).  fp32_transposes=1 emulates the
 * reference's FP32 rounding inside chp2x/chx2p (change.f90:565-686);
 * 0 is the FP64-transpose variant.  freeze_dt=1 keeps dt after it has been
 * computed once (fixed-dt runs); 0 recomputes it at substep 1 of every step
 * as bl_1.F90:91 does. */
oc_box *oc_create(int nx, int ny, int nz, int xout, double re, double ax,
                  double az, double cfl, const double *y, int fp32_transposes,
                  int freeze_dt);
void oc_destroy(oc_box *b);

/* Field ids for oc_field / oc_set_field / oc_get_field. */
enum {
  OC_U = 0, OC_V, OC_W, OC_P, OC_RHSUPA, OC_RHSVPA, OC_RHSWPA, OC_NFIELDS
};
size_t oc_field_len(const oc_box *b, int id); /* number of doubles */
double *oc_field(oc_box *b, int id);          /* direct pointer */

/* Inflow data and genflu state: u0(1:ny+1), v0(1:ny), um(1:ny+1). */
void oc_set_inflow(oc_box *b, const double *u0, const double *v0,
                   const double *um);
void oc_get_um(const oc_box *b, double *um);
/* Run the setup that bl_1.F90:56-58 runs after the restart read:
 * coef, inlet_retheta, magicnumber.  Called by oc_create (coef) and
 * oc_set_inflow (inlet_retheta, magicnumber). */
void oc_set_time(oc_box *b, double tiempo, double timeinit, double dt);
void oc_get_time(const oc_box *b, double *tiempo, double *dt);
void oc_get_genflu(const oc_box *b, double *out6); /* rthin,din,rthout,dout,g1,g2 */
void oc_get_ener(const oc_box *b, double *ener15);

/* Coefficient export: name -> pointer to the array and its shape (rows,cols).
 * The array is in Fortran order; the first column is index 1 (or the
 * Fortran lower bound) of the reference array. */
const double *oc_coef(const oc_box *b, const char *name, int *rows, int *cols);
double oc_scalar(const oc_box *b, const char *name);

/* One RK3 substep m (1..3): rhsp, outflow_correction, pois
 * (bl_1.F90:94-125), and a whole step (3 substeps, tiempo += dt). */
void oc_rhsp(oc_box *b, int m);
void oc_outflow_correction(oc_box *b);
void oc_pois(oc_box *b, double varstep);
void oc_substep(oc_box *b, int m);
void oc_step(oc_box *b);

/* Operator-level entry points (lip.f90 / ffts.f90), on plane-layout fields
 * of the box's size.  which: 0 = u-type coefficient set, 1 = v/w-type. */
void oc_op_interpxx(oc_box *b, const double *gen, double *geni, int R,
                    int which_u);
void oc_op_difvisxx(oc_box *b, const double *prod, const double *vel,
                    double *out, int R, int which_u);
void oc_op_interpyy(oc_box *b, const double *gen, double *geni, int loguv,
                    int nplanes);
void oc_op_differzy(oc_box *b, const double *geni, double *res, int which,
                    int nplanes);
void oc_op_vistyy(oc_box *b, const double *u, double *res, int R, int nplanes);
void oc_op_implzy(oc_box *b, double *u, const double *res, int R, double rkk,
                  int nplanes);
/* z transforms of one row: c2r (fou: nz1 doubles -> phys: nz doubles) and
 * r2c (phys -> fou) with the reference scaling (ffts.f90:48-58). */
void oc_c2r(const oc_box *b, const double *fou, double *phys);
void oc_r2c(const oc_box *b, const double *phys, double *fou);
/* DCT-II x 1/n (iopt<0) and DCT-III x 0.5 (iopt>0) of nlines lines of
 * length n stored contiguously (ffts.f90:147-181). */
void oc_cosft(int n, int nlines, double *c, int iopt);
/* Per-x diagnostics from the kz=0 mode of u: Re_theta(x), Cf(x). */
void oc_retheta_cf(const oc_box *b, double *rth, double *cf);
/* Statistics (stotf.F90:51-291 built with -DNOCORR -DNOSPECTRA,
 * makefile-OMP-vul:11; the x prologue of rhsp.F90:72-102).  oc_statsp adds
 * one sample of the current u, v, w, p to the OC_NSTAT accumulators, each
 * (ny+1, nx) in Fortran order (row ny+1 and the rows the reference leaves
 * unset stay 0), writes the enstrophy field ens(nz, ny, nx) (FP32) and
 * counts the sample (ical).  Ids follow kStatNames: us vs ws uv ua va wa pp
 * pm vortx vorty vortz vortxa vortya vortza uw vw dispu dispv dispw dispuv
 * pup pvp pdudx pdvdy pdudy pdvdx pdwdz v3 u2v v2u w2v dudx0 dudy0 dvdx0
 * dvdy0 dwdx0 dwdy0 u3 w2u. */
#define OC_NSTAT 40
void oc_statsp(oc_box *b);
const char *oc_stat_name(int id);
void oc_stats_get(const oc_box *b, int id, double *out);
void oc_stats_reset(oc_box *b);
int oc_stats_ical(const oc_box *b);
const float *oc_stats_ens(const oc_box *b);
/* fd_dvdy, fd_dx, fd_ix, fd_iy, fd_iyp as (8, n) Fortran arrays from column 1 */
const double *oc_stats_coef(const oc_box *b, const char *name);
/* op 0 diffy_inplace, 1 interpy_new (tipo 0), 2 interpy_new (tipo 1) on n
 * planes; 3 diffx_inplace, 4 interpx_new on n pencils of nx points */
void oc_stats_op(const oc_box *b, int op, const double *in, double *out, int n);

/* OpenMP threads the oracle's parallel loops use. */
int oc_max_threads(void);
/* coefb (coeft.f90:1360-1789), flags 0, 1, 3..7: n closure weights */
void oc_coefb(int n, int flagv, double a, double b, double c, double d, double *bcc);

#ifdef __cplusplus
}
#endif
#endif
