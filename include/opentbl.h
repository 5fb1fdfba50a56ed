/*
 * opentbl.h -- C ABI of libopentbl_amd.so, the MI355X RK3 substep of the
 * OpenTBL zero-pressure-gradient boundary-layer DNS.
 *
 * One otbl_ctx holds one box: sizes, FP64 coefficients, every field resident
 * in device memory (HBM), the HIP stream the work is enqueued on and the
 * captured substep graphs.  All entry points return OTBL_OK (0) or a non-zero
 * status; otbl_last_error() gives the message.  Host arrays are plain
 * doubles; complex fields use the reference's planes layout: the complex*16
 * array (0:nz2, R, 1:nx) seen as doubles, i.e. real index kr = 0..nz1-1
 * fastest (2k = Re of mode k, 2k+1 = Im), then row j = 1..R, then plane
 * i = 1..nx, with R = ny+1 for u, w and R = ny for v, p.
 *
 * The reference's plane<->pencil transposes (change.f90:224-275, chp2x and
 * chx2p) have no counterpart: every x-line kernel reads and writes the
 * planes layout directly.
 */
#ifndef OPENTBL_H
#define OPENTBL_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct otbl_ctx otbl_ctx;

enum { OTBL_OK = 0, OTBL_EINVAL = 1, OTBL_EHIP = 2, OTBL_ENOMEM = 3, OTBL_ESTATE = 4 };
enum { OTBL_FP64 = 0, OTBL_MIXED = 1 };
/* dt every step: recomputed from the CFL limits at substep 1 of every step
 * (bl_1.F90:91, rhsp.F90:214-262).  Frozen: computed once, then kept. */
enum { OTBL_DT_EVERY_STEP = 0, OTBL_DT_FROZEN = 1 };
/* field ids: state the reference keeps between substeps (main_val) */
enum { OTBL_U = 0, OTBL_V, OTBL_W, OTBL_P, OTBL_RHSUPA, OTBL_RHSVPA, OTBL_RHSWPA,
       OTBL_NFIELDS };

typedef struct otbl_params {
  int nx, ny, nz;            /* grid: mod.F90:22 (parameter nx, ny, nz) */
  int xout;                  /* recycling plane: mod.F90:23 (xout) */
  double re, ax, az, cfl;    /* hrem_1.dat values read at bl_1.F90:239-244 */
  int precision;             /* OTBL_FP64, or OTBL_MIXED: FP32 fields and
                                FP32 derivative / transform / product
                                arithmetic with FP64 statistics, CFL, mass
                                flux, coefficient factorisation, mean-flow
                                (kz = 0) pressure solve and Poisson residuals;
                                z transforms on the register engine
                                (nz 256..4096) or the LDS Stockham kernels
                                (other nz), needs a register-engine
                                Bluestein length (nx-1 >= 129) and the
                                chunked y solver (ny <= 384).  The mean-flow (kz = 0) u and p
                                are kept in FP64 on the device.  Fields
                                cross otbl_set_field / otbl_get_field as FP64
                                either way; the operator entry points
                                (otbl_interpxx ...) stay FP64. */
  int dt_mode;               /* OTBL_DT_EVERY_STEP or OTBL_DT_FROZEN */
} otbl_params;

/* Box setup.  Replaces iniciap_1 (bl_1.F90:179-367: sizes, wavenumbers),
 * alloavar/alloabuff (alloa.F90:206-327: field and buffer allocation, here
 * in device memory) and coef (coeft.f90:10-1234: every compact-scheme
 * coefficient in FP64, factorised on the GPU).  y has ny+2 entries,
 * y(0:ny+1) as getstartzy reads it (getstartzy-bg.F90:180-295). */
int otbl_create(const otbl_params *prm, const double *y, otbl_ctx **out);
int otbl_destroy(otbl_ctx *ctx);
const char *otbl_last_error(void);

/* State transfer (the restart read of getstartzy-bg.F90:180-295 and the
 * write of escru.F90:9-251 without the file format). */
int otbl_set_field(otbl_ctx *ctx, int id, const double *host);
int otbl_get_field(otbl_ctx *ctx, int id, double *host);
size_t otbl_field_len(const otbl_ctx *ctx, int id); /* doubles */
/* Inflow profiles u0(1:ny+1), v0(1:ny) and the running mean um(1:ny+1)
 * (genmod, mod.F90:86-93), then inlet_retheta (bl_1.F90:372-396) and
 * magicnumber (boun.f90:102-143). */
int otbl_set_inflow(otbl_ctx *ctx, const double *u0, const double *v0, const double *um);
int otbl_get_um(otbl_ctx *ctx, double *um);
/* genflu diagnostics: rthin, din, rthout, dout, gamma(1), gamma(2) */
int otbl_get_genflu(otbl_ctx *ctx, double *out6);
/* energies (rhsp.F90:493-509, stotf.F90:395-440) of the last substep 1 */
int otbl_get_ener(otbl_ctx *ctx, double *ener15);

/* Per-substep diagnostics: info_per_step (genflu.F90:247-295), which the
 * reference writes as one line of 49 reals to unit 62 at every substep when
 * built without NOINFOSTEP (genflu.F90:157-160, 287-289).  Here each substep
 * appends the same 49 values, in the same order, to a ring of records in
 * device memory, with no host synchronisation: tiempo, rthout, utauout,
 * gamma(1:2), dout, jtop, ntop(1:2), then x_plane(8), u_inf(8), u_tau(8),
 * d_out(8), r_thout(8) at the sample planes ((nx-1) i)/8 + 1, i = 1..8 (the
 * planes the reference samples with one plane per rank). */
#define OTBL_DIAG_LEN 49
/* Ring of `capacity` records (0 turns the diagnostics off).  Captured
 * substep graphs are dropped and captured again on next use. */
int otbl_diag_enable(otbl_ctx *ctx, int capacity);
/* Waits for the context's stream, then copies the records written since the
 * last read, oldest first, at most max_records of them, into out
 * (max_records x OTBL_DIAG_LEN doubles).  *n_records receives the number
 * copied; *dropped the number lost to the ring's wrap-around since the last
 * read (records beyond max_records stay for the next read). */
int otbl_diag_read(otbl_ctx *ctx, double *out, int max_records, int *n_records, long long *dropped);
/* Re_theta(x) and Cf(x) at every plane x = 1..nx from the kz = 0 mode of u
 * (the FP64 mean-flow u in the mixed substep): Re_theta by the sums of
 * genflu.F90:77-87, which compute-Re_theta.f90:34-49 repeats as a
 * post-processing tool, and Cf = 2 u_tau^2 with u_tau^2 from genflu.F90:76.
 * rth and cf receive nx doubles each (host memory). */
int otbl_retheta_cf(otbl_ctx *ctx, double *rth, double *cf);
int otbl_set_time(otbl_ctx *ctx, double tiempo, double timeinit, double dt);
int otbl_get_time(otbl_ctx *ctx, double *tiempo, double *dt);
/* Coefficient arrays after setup, by their reference names (inxu, dcbx,
 * vyui, ayp, kmod, ..., and the stats-only schemes fd_dvdy, fd_iy, fd_iyp,
 * fd_dx, fd_ix of coeft.f90:56-336, LU-factored), in Fortran order;
 * rows/cols give the shape. */
int otbl_get_coef(otbl_ctx *ctx, const char *name, double *out, int *rows, int *cols);

/* ---- the RK3 substep (bl_1.F90:94-125) ---- */
/* rhsp (rhsp.F90:22-489): right-hand side, CFL dt (substep 1), genflu,
 * pressure-gradient predictor, boundary conditions, explicit RK3
 * combination and the implicit y-viscous solve.  m = 1..3. */
int otbl_rhsp(otbl_ctx *ctx, int m);
/* outflow_correction (boun.f90:63-93), FP64 net-flux sum on the GPU */
int otbl_outflow_correction(otbl_ctx *ctx);
/* pois (pois.f90:11-189) with an explicit varstep, and with the
 * varstep = 0.5/dt/rkdv(m) of bl_1.F90:117 evaluated on the GPU */
int otbl_pois(otbl_ctx *ctx, double varstep);
int otbl_pois_m(otbl_ctx *ctx, int m);
/* one substep: rhsp, outflow_correction, pois (bl_1.F90:98-119) */
int otbl_substep(otbl_ctx *ctx, int m);
/* nsteps whole steps: 3 substeps each, then tiempo += dt (bl_1.F90:85-128) */
int otbl_step(otbl_ctx *ctx, int nsteps);
/* 1: each substep runs as one captured hipGraph (captured on first use);
 * 0: kernels are launched one by one. */
int otbl_use_graphs(otbl_ctx *ctx, int on);
int otbl_sync(otbl_ctx *ctx);

/* ---- operators on device buffers of this box's size (lip.f90, ffts.f90) ----
 * in/out are device pointers to planes-layout fields (R rows per plane). */
/* interpxx (lip.f90:248-300): which_u=1 -> u coefficients (inxu, cofiux,
 * loguv=1), 0 -> v/w coefficients (inxv, cofivx, loguv=0) */
int otbl_interpxx(otbl_ctx *ctx, const double *gen, double *geni, int R, int which_u);
/* difvisxx (lip.f90:52-142): convective d/dx of prod plus rex d2/dx2 of vel */
int otbl_difvisxx(otbl_ctx *ctx, const double *prod, const double *vel, double *out, int R,
                  int which_u);
/* interpyy (lip.f90:359-440), spectral rows: loguv 0 = u/w rows (ny+1) to v
 * rows, 1 = v rows to u rows; output has ny+1 rows per plane (row ny+1 is
 * not written) */
int otbl_interpyy(otbl_ctx *ctx, const double *gen, double *geni, int loguv, int nplanes);
/* differzy (lip.f90:152-232): which 0 = v rows -> u rows (dcyu), 1 = u rows
 * -> v rows (dcyv); in and out have ny+1 rows per plane */
int otbl_differzy(otbl_ctx *ctx, const double *geni, double *res, int which, int nplanes);
/* vistyy (lip.f90:447-513) and implzy (lip.f90:528-597) */
int otbl_vistyy(otbl_ctx *ctx, const double *u, double *res, int R, int nplanes);
int otbl_implzy(otbl_ctx *ctx, double *u, const double *res, int R, double rkk, int nplanes);
/* fourxz (ffts.f90:27-67) on nrows spectral rows (nz1 doubles each) and
 * physical rows (nz doubles each): iopt >= 0 c2r, iopt < 0 r2c */
int otbl_fourxz(otbl_ctx *ctx, double *fou, double *phys, int iopt, int nrows);
/* cosftx (ffts.f90:147-181) along x over planes 2..nx of a v-type field,
 * rows 1..ny: iopt < 0 DCT-II x 1/(nx-1), iopt > 0 DCT-III x 0.5 */
int otbl_cosftx(otbl_ctx *ctx, const double *in, double *out, int iopt);

/* cosftx (ffts.f90:147-181) on nlines contiguous device lines of nx-1
 * doubles, in place: iopt < 0 FFTW REDFT10 (DCT-II) x 1/(nx-1), iopt > 0
 * REDFT01 (DCT-III) x 0.5.  The length is the box's: cosfti(n) is called
 * with n = nx - 1 (ffts.f90:124, bl_1.F90:304). */
int otbl_cosft_lines(otbl_ctx *ctx, double *lines, int nlines, int iopt);
/* soltriy (pois.f90:211-239): a is the caller's host array (3, 1:ny1),
 * LU-factored in place as the reference does; then column ii (0..nz2) of
 * the complex device plane u(0:nz2, 1:ny1) is solved into the device plane
 * us (us may be u). */
int otbl_soltriy(otbl_ctx *ctx, const double *u, double *us, int ii, double *a, int ny1);
/* otbl_soltriy without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * graph capture, per-mode solve loops of a serving path, harnesses that
 * forbid hipMalloc).  Same arguments, results and error behaviour as
 * otbl_soltriy; the tridiagonal is staged in a fixed device scratch of the
 * library (one per GPU, held until the process ends), so ny1 <= 8192 (a
 * larger ny1 is rejected with OTBL_EINVAL) and calls that overlap on one
 * GPU (from several threads) race on that scratch. */
int otbl_soltriy_noalloc(otbl_ctx *ctx, const double *u, double *us, int ii, double *a, int ny1);
/* coefinv (coeft.f90:1338-1353): in-place LU of the caller's host array
 * a(3, npter1:neq), computed on the GPU. */
int otbl_coefinv(otbl_ctx *ctx, double *a, int npter1, int neq);
/* soltrix (coeft.f90:1926-1967): the tridiagonal a(3, npter1:nfin),
 * nfin = npter1 + neq - 1, of the caller's host array a(3, *) LU-factored
 * in place when *iflag is 0 (then *iflag = 1, as coefinv factors), and the
 * columns npter2 .. npter2 + nlen - 1 of the host array u(nstr, *), rows
 * npter1 .. nfin, solved into the same rows of us (us may be u; its other
 * rows are untouched); on the GPU. */
int otbl_soltrix(otbl_ctx *ctx, const double *u, int nstr, int nlen, int neq, double *us, int npter1, int npter2,
                 double *a, int *iflag);
/* coefb (coeft.f90:1360-1789): the n weights bcc(1:n) of a boundary
 * closure from the spacings a, b, c, d by its Taylor-moment system, for the
 * flags the reference calls it with (flagv 0: n = 3; 1: n = 5; 3 to 7:
 * n = 4; coeft.f90:589-1023); other flags, or another n, return
 * OTBL_EINVAL.  Host arithmetic, as the box's own coefficient setup; no
 * context needed. */
int otbl_coefb(int n, int flagv, double a, double b, double c, double d, double *bcc);

/* ---- statistics (stotf.F90:51-291) ----
 * The reference's statsp as its production build compiles it
 * (makefile-OMP-vul:11: -DNOCORR -DNOSPECTRA, so no correlations and no
 * z-spectra; dissipation and triple products on), with the x interpolation
 * and x derivatives of rhsp.F90:72-102 that feed it.  A sample is taken
 * from the state at the start of substep 1, in the same captured graph as
 * the substep.  The 40 accumulators are (ny+1, nx) Fortran arrays (row ny+1
 * is never written, as in the reference, and rows the reference sizes ny
 * are returned with a zero row ny+1), named and ordered as the reference's
 * statistics files write them (escru.F90:318-341, 490-516):
 *   us vs ws uv ua va wa pp pm vortx vorty vortz vortxa vortya vortza uw vw
 *   dispu dispv dispw dispuv pup pvp pdudx pdvdy pdudy pdvdx pdwdz v3 u2v
 *   v2u w2v dudx0 dudy0 dvdx0 dvdy0 dwdx0 dwdy0 u3 w2u
 * The enstrophy field of the last sample, ens(nz, ny, nx) in 4-byte reals
 * (stotf.F90:150, written by escru.F90:159-170), is kept beside them.  FP64
 * arithmetic; a mixed box samples its FP32 fields. */
#define OTBL_NSTAT 40
const char *otbl_stats_name(int id); /* null outside 0..OTBL_NSTAT-1 */
/* allocates (on) or frees (off) the accumulators, cleared; nz <= 4096 */
int otbl_stats_enable(otbl_ctx *ctx, int on);
/* the sampling rule of bl_1.F90:92: a sample at substep 1 of every step
 * istep with mod(istep, avcrt) == 0 and istep > stest, istep counting the
 * steps of otbl_step / otbl_ensemble_step from 1 after this call
 * (avcrt = 0: no scheduled samples) */
int otbl_stats_schedule(otbl_ctx *ctx, int avcrt, int stest);
/* dostat = .TRUE. (bl_1.F90:92): the next substep 1 (otbl_rhsp with m = 1,
 * otbl_substep, otbl_step) takes a sample first */
int otbl_stats_request(otbl_ctx *ctx);
/* one sample of the current state now (statsp, with the x prologue of
 * rhsp.F90:72-90; uses the box's work fields, so not during a substep) */
int otbl_statsp(otbl_ctx *ctx);
/* accumulator id as (ny+1) x nx doubles, Fortran order (host memory) */
int otbl_stats_get(otbl_ctx *ctx, int id, double *out);
/* samples accumulated since the last reset (ical, rhsp.F90:92) */
int otbl_stats_ical(otbl_ctx *ctx, int *ical);
/* clears the accumulators and ical (escru.F90:412-414, 579-585, after the
 * reference writes its statistics files) */
int otbl_stats_reset(otbl_ctx *ctx);
/* the enstrophy field of the last sample: nz * ny * nx floats (host) */
int otbl_stats_get_ens(otbl_ctx *ctx, float *out);

/* ---- ensemble of independent boxes ----
 * No counterpart in the reference (one OpenTBL run advances one box): the
 * uneven-box ensemble whose boxes are advanced together to accumulate
 * converged statistics.  Each RK3 substep of all the boxes runs as one
 * captured hipGraph (one graph launch per substep for the whole ensemble)
 * whose independent branches are the boxes' substeps, forked from and
 * joined into the ensemble's own stream, so the kernels of every box overlap
 * on the GPU.  The boxes stay owned by the caller, keep their own state,
 * precision, dt and time, and give the same results as otbl_step on each of
 * them; they may still be stepped or read alone between ensemble steps
 * (their streams are ordered after the ensemble's work).  Destroy the
 * ensemble before any of its boxes.  At most 64 boxes. */
typedef struct otbl_ensemble otbl_ensemble;
int otbl_ensemble_create(otbl_ctx *const *boxes, int nbox, otbl_ensemble **out);
/* nsteps whole RK3 steps of every box (3 substeps, then tiempo += dt) */
int otbl_ensemble_step(otbl_ensemble *e, int nsteps);
int otbl_ensemble_sync(otbl_ensemble *e);
int otbl_ensemble_destroy(otbl_ensemble *e);

/* ---- device memory helpers (plain pointers, no framework types) ---- */
int otbl_dev_alloc(size_t bytes, void **ptr);
int otbl_dev_free(void *ptr);
int otbl_h2d(void *dst, const void *src, size_t bytes);
int otbl_d2h(void *dst, const void *src, size_t bytes);
/* device-to-device copy on the device's default stream (also the plain copy
 * of known size that calibrates hardware byte counters) */
int otbl_d2d(void *dst, const void *src, size_t bytes);

/* Kernel paths the context chose at setup, as text: "y:chunk<L>|march0
 * z:engine<nz>|stockham<nz> dct:engine<M>|stockham<M>|gemm<M> x:window|march
 * xseg:<threads per x line> pois:chunk<L>|march0 prun:<planes per thread>
 * tuned:0|1 streams:1|2" (streams:2: rhsp's plane loop 2 on a second stream
 * beside plane loop 1)
 * (no counterpart in the reference; for tests and benchmark records). */
int otbl_get_paths(otbl_ctx *ctx, char *buf, size_t len);

/* Autotuner (no counterpart in the reference): times, on this box's sizes,
 * the candidate launch shapes of its kernels and keeps the fastest: the
 * segments per x line of the windowed x solves (1 up to the most the x
 * systems' decay allows), the planes per thread of the element-wise
 * predictor and projection (1, 2, 4, 8), and for the mixed substep the
 * half-wave or full-wave chunks of rk_final.  Each candidate runs `reps`
 * times (after one untimed run) on the box's work fields, which hold nothing
 * between substeps, so the state is unchanged.  The plane run changes no
 * arithmetic; the x segments and the y chunk shapes change only round-off
 * (a segment's start error is below 1e-17, checked at setup).  report (may
 * be null) receives the timings; otbl_get_paths shows the choices
 * ("tuned:1").  Captured graphs are dropped and captured again. */
int otbl_autotune(otbl_ctx *ctx, int reps, char *report, size_t len);

/* ---- measurement ----
 * With kprof on, every kernel launch is bracketed by HIP events on the
 * context's stream; the report lists name, launches, total ms.  Kernels run
 * one by one (no graphs) while kprof is on. */
int otbl_kprof(otbl_ctx *ctx, int on);
int otbl_kprof_report(otbl_ctx *ctx, char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
