/*
 * OpenTBL RK3 substep -- CPU FP64 oracle.  TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP product, never the product.  See otbl_oracle.h for
 * the layout.  Each function cites the reference file:line it restates.
 *
 * Parity status of each part (how it is pinned, see tests/):
 *   coefficients (coef)          polynomial exactness on the actual grids,
 *                                Lele textbook interior values, kmod closed form
 *   tridiagonal sweeps (lip)     dense numpy.linalg.solve of assembled matrices
 *   c2r / r2c (fourxz)           numpy.fft.irfft / rfft with reference scaling
 *   cosft (cosftx)               scipy.fft.dct type 2 / type 3
 *   pois + soltriy + mtr         discrete divergence at round-off after the
 *                                projection; mtr == div(grad) identity
 *   outflow_correction           net flux exactly zero afterwards
 *   boun                         boundary rows satisfy their defining identities
 *   genflu (dobleu, profile)     gamma ~ 0.5: a profile linear in the rescaled
 *                                coordinate comes back as the closed form with
 *                                the dobleu weight integral by quadrature;
 *                                gamma_1 = gamma_2^(1/8); mode 0 = u0, v0
 *   difvisxx                     exact on cubics (convective) and quartics
 *                                (viscous) with the boundary closures; outflow
 *                                row U_inf dq/dx
 *   implzy                       dense solve of (A + rkk B) u = A wki; inverse
 *                                of I + rkk D_yy applied by vistyy
 *   rhsp composition             manufactured solution: the rhs converges to
 *                                d(u_j u_i)/dx_j - nu lap u_i at second order
 *                                under grid doubling; one RK3 step multiplies
 *                                a z-viscous mode by 1 + z + z^2/2 + z^3/6
 *   Re_theta(x), Cf(x)           inlet values of inlet_retheta; ramp closed forms
 */
#include "otbl_oracle.h"
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define PI_D (4.0 * atan(1.0))
/* Fortran-style accessors.  (3,n) / (2,n) coefficient arrays are stored with
 * column i at offset 3*i (resp. 2*i) so that 1-based i can be used directly. */
#define C3(a, r, i) (a)[3 * (size_t)(i) + (r)-1]
#define C2(a, r, i) (a)[2 * (size_t)(i) + (r)-1]
#define M4(a, r, c) (a)[4 * ((c)-1) + (r)-1]
#define M5(a, r, c) (a)[5 * ((c)-1) + (r)-1]
#define M2(a, r, c) (a)[2 * ((c)-1) + (r)-1]

struct oc_box {
  int nx, ny, nz, nz1, nz2, ngz, xout;
  double re, ax, az, cfl, pi;
  int fp32_transposes, freeze_dt;
  /* grid and metrics */
  double *y, *dy, *dyi, *idy, *idyy, *dymin, *x, *dx, *dxi, *hy;
  double idx, idxx, dxmin, dzmin, dtret, rex;
  double *kaz, *kaz2, *kvis, *kmod; /* kaz = imaginary part of reference kaz */
  double rkc[4], rkd[4], rkcv[4], rkdv[4];
  /* coefficient arrays (reference names) */
  double *inxu, *inxv, *inyu, *inyv, inbx[16], inby[16];
  double *cofiux, *cofivx, *cofiuy, *cofivy;
  double *dcxu, *dcxv, *dcyu, *dcyv, dcbx[20], dcby[16], ccon[8];
  double *cofcxu, *cofcxv, *cofcyu, *cofcyv;
  double *vixu, *vixv, *viyu, *viyv, *vyui, *vyvi, cofvbx[16], cofvby[16];
  double *cofvxu, *cofvxv, *cofvyu, *cofvyv;
  double *ayp, *vmagic;
  /* inflow / genflu state (genmod) */
  double *um, *u0, *v0;
  double rthin, din, rthout, dout, utauout, gam[2], timeinit, timec;
  double tiempo, dt;
  int dt_valid, setstep;
  double ener[16];
  /* fields */
  double *fld[OC_NFIELDS];
  size_t nu, nv; /* doubles in a u-type / v-type field */
  /* work arrays */
  double *pen_u, *pen_v, *pen_w, *pen_a; /* pencils */
  double *resu, *resv, *resw, *rhsu, *rhsv, *rhsw, *wki1, *wki2, *wki3;
  double *rest, *tbuf;
  double *vm; /* 1..ny */
  /* FFT */
  double *fft_tw;
  /* per-thread scratch of the plane loops (OpenMP over planes; every
   * parallel loop writes disjoint data and the only reductions are maxima,
   * so results do not depend on the thread count) */
  int nth;
  struct oc_tls *tls;
  /* statistics (stotf.F90:51-291): the stats-only compact schemes of
   * coeft.f90:56-285 as (8, n) arrays (column i at offset 8 i), the
   * accumulators (OC_NSTAT of (ny+1, nx), Fortran order), the last
   * enstrophy field ens(nz, ny, nx) in FP32 (escru.F90:159-170 writes it)
   * and the number of accumulated samples (ical, rhsp.F90:92) */
  double *fd_dvdy, *fd_dx, *fd_ix, *fd_iy, *fd_iyp;
  double *st;
  float *ens;
  int ical;
};
struct oc_tls {
  double *wkp, *wkpo, *wkf, *wkfo, *wk2r, *fft_a, *fft_b, *fft_t, *vm;
};
#define TLS(b) (&(b)->tls[omp_get_thread_num()])

static double *zalloc(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu doubles)\n", n);
    abort();
  }
  return p;
}
static double f32(double v) { return (double)(float)v; }

/* ------------------------------------------------------------------------
 * Small dense solve with partial pivoting (replaces the reference's
 * ludcmp/lubksb pair, coeft.f90:1818-1922, with an independent Gaussian
 * elimination).  a is n x n row-major: sum_j a[i][j] x_j = b_i.
 * ---------------------------------------------------------------------- */
static void dense_solve(int n, double a[5][5], double b[5]) {
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
    if (a[piv][c] == 0.0) {
      fprintf(stderr, "oracle: singular boundary-closure system\n");
      abort();
    }
    if (piv != c) {
      for (int j = 0; j < n; j++) {
        double t = a[c][j]; a[c][j] = a[piv][j]; a[piv][j] = t;
      }
      double t = b[c]; b[c] = b[piv]; b[piv] = t;
    }
    for (int r = c + 1; r < n; r++) {
      double f = a[r][c] / a[c][c];
      for (int j = c; j < n; j++) a[r][j] -= f * a[c][j];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; r--) {
    double s = b[r];
    for (int j = r + 1; j < n; j++) s -= a[r][j] * b[j];
    b[r] = s / a[r][r];
  }
}

/* coefb (coeft.f90:1360-1789): boundary-closure weights from Taylor
 * matching.  Only the flag values the RK3 path uses (0,1,3,4,5,6,7). */
static void coefb(int n, double a, double b, double c, double d, double bcc[5],
                  int flagv) {
  double m[5][5];
  memset(m, 0, sizeof m);
  for (int i = 0; i < 5; i++) bcc[i] = 0.0;
  switch (flagv) {
  case 1: /* coeft.f90:1369-1411 */
    bcc[0] = 1.0;
    m[0][0] = -1.0; m[0][1] = 0.0; m[0][2] = 0.5 * a * a;
    m[0][3] = 0.5 * (a + b) * (a + b); m[0][4] = 0.5 * (a + b + c) * (a + b + c);
    m[1][1] = 1.0; m[1][2] = 1.0; m[1][3] = 1.0; m[1][4] = 1.0;
    m[2][2] = a; m[2][3] = a + b; m[2][4] = a + b + c;
    m[3][0] = -a; m[3][2] = pow(a, 3) / 6.0; m[3][3] = pow(a + b, 3) / 6.0;
    m[3][4] = pow(a + b + c, 3) / 6.0;
    m[4][0] = -0.5 * a * a; m[4][2] = pow(a, 4) / 24.0;
    m[4][3] = pow(a + b, 4) / 24.0; m[4][4] = pow(a + b + c, 4) / 24.0;
    break;
  case 0: /* coeft.f90:1413-1437 */
    bcc[0] = 1.0;
    m[0][0] = 1.0; m[0][1] = 1.0; m[0][2] = 1.0;
    m[1][0] = -0.5 * a; m[1][1] = -(0.5 * a + b); m[1][2] = -(0.5 * a + b + c);
    m[2][0] = a * a / 8.0; m[2][1] = 0.5 * (0.5 * a + b) * (0.5 * a + b);
    m[2][2] = 0.5 * (0.5 * a + b + c) * (0.5 * a + b + c);
    break;
  case 3: { /* coeft.f90:1487-1523 */
    bcc[0] = 1.0;
    double s1 = 0.5 * a, s2 = 0.5 * a + b, s3 = 0.5 * a + b + c,
           s4 = 0.5 * a + b + c + d;
    m[0][0] = s1; m[0][1] = s2; m[0][2] = s3; m[0][3] = s4;
    m[1][0] = 1.0; m[1][1] = 1.0; m[1][2] = 1.0; m[1][3] = 1.0;
    m[2][0] = a * a / 8.0; m[2][1] = 0.5 * s2 * s2; m[2][2] = 0.5 * s3 * s3;
    m[2][3] = 0.5 * s4 * s4;
    m[3][0] = a * a * a / 48.0; m[3][1] = s2 * s2 * s2 / 6.0;
    m[3][2] = s3 * s3 * s3 / 6.0; m[3][3] = s4 * s4 * s4 / 6.0;
    break;
  }
  case 4: { /* coeft.f90:1525-1557 */
    bcc[0] = 1.0;
    double s3 = 0.5 * b + c, s4 = 0.5 * b + c + d;
    m[0][0] = 1.0; m[0][1] = 1.0; m[0][2] = 1.0; m[0][3] = 1.0;
    m[1][0] = -0.5 * a; m[1][1] = 0.5 * b; m[1][2] = s3; m[1][3] = s4;
    m[2][0] = a * a / 8.0; m[2][1] = b * b / 8.0; m[2][2] = 0.5 * s3 * s3;
    m[2][3] = 0.5 * s4 * s4;
    m[3][0] = -a * a * a / 48.0; m[3][1] = b * b * b / 48.0;
    m[3][2] = s3 * s3 * s3 / 6.0; m[3][3] = s4 * s4 * s4 / 6.0;
    break;
  }
  case 5: { /* coeft.f90:1559-1595 */
    bcc[0] = 1.0;
    double s3 = -(0.5 * b + c), s4 = -(0.5 * b + c + d);
    m[0][0] = 0.5 * a; m[0][1] = -0.5 * b; m[0][2] = s3; m[0][3] = s4;
    m[1][0] = 1.0; m[1][1] = 1.0; m[1][2] = 1.0; m[1][3] = 1.0;
    m[2][0] = a * a / 8.0; m[2][1] = b * b / 8.0; m[2][2] = 0.5 * s3 * s3;
    m[2][3] = 0.5 * s4 * s4;
    m[3][0] = a * a * a / 48.0; m[3][1] = -b * b * b / 48.0;
    m[3][2] = s3 * s3 * s3 / 6.0; m[3][3] = s4 * s4 * s4 / 6.0;
    break;
  }
  case 6: { /* coeft.f90:1597-1633 */
    bcc[1] = 1.0;
    double s3 = -(0.5 * a + b), s4 = -(0.5 * a + b + c);
    m[0][0] = 1.0; m[0][1] = 1.0; m[0][2] = 1.0; m[0][3] = 1.0;
    m[1][1] = -0.5 * a; m[1][2] = s3; m[1][3] = s4;
    m[2][1] = a * a / 8.0; m[2][2] = 0.5 * s3 * s3; m[2][3] = 0.5 * s4 * s4;
    m[3][1] = -a * a * a / 48.0; m[3][2] = s3 * s3 * s3 / 6.0;
    m[3][3] = s4 * s4 * s4 / 6.0;
    break;
  }
  case 7: { /* coeft.f90:1635-1672; note row 3 has no 1/2 on columns 3,4 */
    bcc[1] = 1.0;
    double s3 = 0.5 * b + c, s4 = 0.5 * b + c + d;
    m[0][0] = 1.0; m[0][1] = 1.0; m[0][2] = 1.0; m[0][3] = 1.0;
    m[1][0] = 0.5 * a; m[1][1] = -0.5 * b; m[1][2] = -s3; m[1][3] = -s4;
    m[2][0] = a * a / 8.0; m[2][1] = b * b / 8.0; m[2][2] = s3 * s3;
    m[2][3] = s4 * s4;
    m[3][0] = a * a * a / 48.0; m[3][1] = -b * b * b / 48.0;
    m[3][2] = -s3 * s3 * s3 / 6.0; m[3][3] = -s4 * s4 * s4 / 6.0;
    break;
  }
  default:
    fprintf(stderr, "oracle: coefb flag %d not on the RK3 path\n", flagv);
    abort();
  }
  dense_solve(n, m, bcc);
}

/* coefb for the tests of the library's closure weights */
void oc_coefb(int n, int flagv, double a, double b, double c, double d, double *bcc) {
  coefb(n, a, b, c, d, bcc, flagv);
}

/* coefinv (coeft.f90:1338-1353) with npter1 = 1: in-place LU of a
 * tridiagonal stored as (3, 1:neq). */
static void coefinv(double *a, int neq) {
  for (int j = 2; j <= neq; j++) {
    double d = 1.0 / C3(a, 2, j - 1);
    C3(a, 2, j - 1) = d;
    C3(a, 1, j) = -C3(a, 1, j) * d;
    C3(a, 2, j) = C3(a, 2, j) + C3(a, 1, j) * C3(a, 3, j - 1);
  }
  C3(a, 2, neq) = 1.0 / C3(a, 2, neq);
}

/* mtr (coeft.f90:1244-1332): Poisson operator in y, rows 1..ny-1. */
static void mtr(oc_box *b) {
  const double *yy = b->y;
  int lt = b->ny - 1;
  double ac2 = 0.0, bc2 = 0.0, ac1, bc1, dy1, dy2, a3, b2;
  ac1 = ac2; bc1 = bc2;
  dy1 = yy[2] - yy[1]; dy2 = yy[3] - yy[2];
  a3 = 1.0 / dy1; b2 = -1.0 / dy1;
  ac2 = 2.0 / (dy1 + dy2); bc2 = -2.0 / (dy1 + dy2);
  C3(b->ayp, 1, 1) = b2 * bc1;
  C3(b->ayp, 3, 1) = a3 * ac2;
  C3(b->ayp, 2, 1) = a3 * bc2 + b2 * ac1;
  for (int j = 2; j <= lt - 1; j++) {
    ac1 = ac2; bc1 = bc2;
    dy1 = yy[j + 1] - yy[j]; dy2 = yy[j + 2] - yy[j + 1];
    a3 = 1.0 / dy1; b2 = -1.0 / dy1;
    ac2 = 2.0 / (dy1 + dy2); bc2 = -2.0 / (dy1 + dy2);
    C3(b->ayp, 1, j) = b2 * bc1;
    C3(b->ayp, 3, j) = a3 * ac2;
    C3(b->ayp, 2, j) = a3 * bc2 + b2 * ac1;
  }
  b2 = -1.0 / dy2;
  ac1 = ac2; bc1 = bc2;
  C3(b->ayp, 1, lt) = b2 * bc1;
  C3(b->ayp, 3, lt) = 0.0;
  C3(b->ayp, 2, lt) = b2 * ac1;
}

/* coef (coeft.f90:10-1234), the arrays the RK3 step reads. */
static void coef(oc_box *b) {
  const int nx = b->nx, ny = b->ny, nz = b->nz;
  const double pi = b->pi, cfl = b->cfl, re = b->re;
  double *y = b->y, *dy = b->dy, *dyi = b->dyi, *dx = b->dx, *dxi = b->dxi,
         *x = b->x;
  double a, bb, c, aux, bcc[5];

  /* hy (coeft.f90:32-37), used by the energies diagnostic */
  for (int j = 1; j <= ny - 1; j++) b->hy[j] = (y[j + 1] - y[j - 1]) / 2.5 / 2.0;
  b->hy[0] = (y[1] - y[0]) / 2.5;
  b->hy[ny] = (y[ny] - y[ny - 1]) / 2.5;

  /* metrics (coeft.f90:397-424) */
  for (int j = 0; j <= ny; j++) dy[j] = y[j + 1] - y[j];
  for (int j = 1; j <= ny; j++) dyi[j] = 0.5 * (dy[j] + dy[j - 1]);
  double point0 = b->ax * pi / (nx + 1);
  for (int i = 0; i <= nx + 1; i++) x[i] = (double)i * point0;
  double daux = 1000.0;
  for (int i = 0; i <= nx; i++) {
    dx[i] = x[i + 1] - x[i];
    if (dx[i] < daux) daux = dx[i];
  }
  b->dxmin = daux;
  for (int i = 1; i <= nx; i++) dxi[i] = 0.5 * (dx[i] + dx[i - 1]);
  b->idx = 1.0 / dx[1];
  for (int j = 0; j <= ny; j++) b->idy[j] = 1.0 / dy[j];
  b->idxx = 1.0 / dx[1];
  for (int j = 1; j <= ny; j++) b->idyy[j] = 2.0 / (dy[j - 1] + dy[j]);

  /* time-step limits (coeft.f90:432-454) */
  b->dzmin = b->az * 2 * pi / nz;
  double dtrex = re * (cfl * b->dxmin * b->dxmin) / 6.0;
  double dtrez = re * (cfl * b->dzmin * b->dzmin) / (pi * pi);
  b->dtret = dtrex < dtrez ? dtrex : dtrez;
  for (int j = 1; j <= ny; j++) b->dymin[j] = dy[j] * cfl / sqrt(3.0);
  b->dxmin = b->dxmin * cfl / sqrt(3.0);
  b->dzmin = b->dzmin * cfl / pi;

  b->rex = -1.0 / re;
  for (int k = 0; k <= b->nz2; k++) b->kvis[k] = b->rex * b->kaz2[k];

  /* RK3 constants (coeft.f90:494-508) */
  b->rkc[1] = 8.0 / 15.0; b->rkc[2] = 5.0 / 12.0; b->rkc[3] = 3.0 / 4.0;
  b->rkd[1] = 0.0; b->rkd[2] = -17.0 / 60.0; b->rkd[3] = -5.0 / 12.0;
  b->rkcv[1] = 4.0 / 15.0; b->rkcv[2] = 1.0 / 15.0; b->rkcv[3] = 1.0 / 6.0;
  b->rkdv[1] = 4.0 / 15.0; b->rkdv[2] = 1.0 / 15.0; b->rkdv[3] = 1.0 / 6.0;

  /* midpoint interpolation, implicit part (coeft.f90:514-580) */
  for (int i = 2; i <= nx - 1; i++) {
    c = dxi[i + 1]; a = dxi[i]; bb = dx[i];
    C3(b->inxu, 1, i) = c * bb * bb / ((2 * a + bb) * (-bb + 2 * a) * (c + a));
    C3(b->inxu, 2, i) = 1.0;
    C3(b->inxu, 3, i) = -a * bb * bb / ((bb - 2 * c) * (bb + 2 * c) * (c + a));
  }
  for (int i = 2; i <= ny - 1; i++) {
    a = dy[i]; bb = dy[i - 1];
    C3(b->inyu, 1, i) = a * a / ((a + bb) * (a + 2 * bb));
    C3(b->inyu, 2, i) = 1.0;
    C3(b->inyu, 3, i) = bb * bb / (bb * bb + 2 * a * a + 3 * a * bb);
  }
  for (int i = 2; i <= nx - 1; i++) {
    a = dx[i]; bb = dx[i - 1];
    C3(b->inxv, 1, i) = a * a / ((bb + a) * (a + 2.0 * bb));
    C3(b->inxv, 2, i) = 1.0;
    C3(b->inxv, 3, i) = bb * bb / (2.0 * a * a + bb * bb + 3.0 * a * bb);
  }
  for (int i = 2; i <= ny - 1; i++) {
    a = dyi[i]; bb = dy[i - 1]; c = dyi[i - 1];
    C3(b->inyv, 1, i) = c * bb * bb / ((2 * a + bb) * (-bb + 2 * a) * (c + a));
    C3(b->inyv, 2, i) = 1.0;
    C3(b->inyv, 3, i) = -a * bb * bb / ((bb - 2 * c) * (bb + 2 * c) * (c + a));
  }
  double *ends[4] = {b->inxu, b->inxv, b->inyu, b->inyv};
  int nends[4] = {nx, nx, ny, ny};
  for (int e = 0; e < 4; e++) {
    C3(ends[e], 1, 1) = 0.0; C3(ends[e], 2, 1) = 1.0; C3(ends[e], 3, 1) = 0.0;
    C3(ends[e], 1, nends[e]) = 0.0; C3(ends[e], 2, nends[e]) = 1.0;
    C3(ends[e], 3, nends[e]) = 0.0;
  }
  /* boundary closures in x (coeft.f90:583-624) */
  coefb(3, dx[nx - 1], dxi[nx - 1], dxi[nx - 2], dxi[nx - 3], bcc, 0);
  M4(b->inbx, 1, 1) = bcc[0]; M4(b->inbx, 1, 2) = bcc[1];
  M4(b->inbx, 1, 3) = bcc[2]; M4(b->inbx, 1, 4) = 0.0;
  coefb(3, dx[nx], dx[nx - 1], dx[nx - 2], dx[nx - 3], bcc, 0);
  M4(b->inbx, 2, 1) = bcc[0]; M4(b->inbx, 2, 2) = bcc[1];
  M4(b->inbx, 2, 3) = bcc[2]; M4(b->inbx, 2, 4) = 0.0;
  coefb(4, dx[1], dx[1], dx[2], dx[3], bcc, 4);
  for (int q = 1; q <= 4; q++) M4(b->inbx, 3, q) = bcc[q - 1];
  coefb(4, dx[0], dx[1], dxi[2], dxi[3], bcc, 4);
  for (int q = 1; q <= 4; q++) M4(b->inbx, 4, q) = bcc[q - 1];
  /* boundary closures in y (coeft.f90:631-675) */
  coefb(4, dy[ny], dy[ny - 1], dyi[ny - 1], dyi[ny - 2], bcc, 4);
  for (int q = 1; q <= 4; q++) M4(b->inby, 1, q) = bcc[q - 1];
  coefb(3, -dy[0], -dy[1], -dy[2], dy[4], bcc, 0);
  M4(b->inby, 3, 1) = bcc[0]; M4(b->inby, 3, 2) = bcc[1];
  M4(b->inby, 3, 3) = bcc[2]; M4(b->inby, 3, 4) = 0.0;
  coefb(4, dy[0], dy[1], dyi[2], dyi[3], bcc, 4);
  for (int q = 1; q <= 4; q++) M4(b->inby, 2, q) = bcc[q - 1];
  coefb(4, -dy[ny - 1], -dy[ny - 1], -dy[ny - 2], -dy[ny - 3], bcc, 4);
  for (int q = 1; q <= 4; q++) M4(b->inby, 4, q) = bcc[q - 1];
  coefinv(b->inxu, nx);
  coefinv(b->inxv, nx);
  coefinv(b->inyu, ny);
  coefinv(b->inyv, ny);
  /* midpoint interpolation, explicit part (coeft.f90:685-713) */
  for (int i = 2; i <= ny - 1; i++) {
    a = dyi[i]; bb = dy[i - 1]; c = dyi[i - 1];
    C2(b->cofivy, 2, i) = (-2.0 * c * a) / (2 * a * bb + bb * bb - 4 * c * a - 2 * c * bb);
    C2(b->cofivy, 1, i) = (2.0 * c * a) / (-bb * bb + 2 * a * bb - 2 * c * bb + 4 * c * a);
    a = dy[i]; bb = dy[i - 1];
    C2(b->cofiuy, 2, i) = 4 * bb * bb / (2 * bb * bb + a * a + 3 * a * bb);
    C2(b->cofiuy, 1, i) = 4 * a * a / ((a + bb) * (2 * a + bb));
  }
  for (int i = 2; i <= nx - 1; i++) {
    a = dx[i]; bb = dx[i - 1];
    C2(b->cofivx, 2, i) = (4.0 * bb * bb) / (2.0 * bb * bb + a * a + 3.0 * a * bb);
    C2(b->cofivx, 1, i) = (4.0 * a * a) / ((bb + a) * (2.0 * a + bb));
    c = dxi[i + 1]; a = dxi[i]; bb = dx[i];
    C2(b->cofiux, 2, i) = (-2.0 * c * a) / (2 * a * bb + bb * bb - 4 * c * a - 2 * c * bb);
    C2(b->cofiux, 1, i) = (2.0 * c * a) / (-bb * bb + 2 * a * bb - 2 * c * bb + 4 * c * a);
  }

  /* first derivative, implicit part (coeft.f90:722-784) */
  for (int i = 2; i <= nx - 1; i++) {
    a = dx[i]; bb = dx[i - 1];
    double den = 2.0 * bb * bb * bb + 2.0 * a * a * a + 9.0 * bb * bb * a + 9.0 * bb * a * a;
    C3(b->dcxu, 1, i) = -a * (-2.0 * a * bb - bb * bb + 2.0 * a * a) / den;
    C3(b->dcxu, 2, i) = 1.0;
    C3(b->dcxu, 3, i) = ((a * a + 2.0 * a * bb - 2.0 * bb * bb) * bb) / den;
  }
  C3(b->dcxu, 2, 1) = 1.0; C3(b->dcxu, 1, 1) = 0.0; C3(b->dcxu, 3, 1) = 0.0;
  C3(b->dcxu, 2, nx) = 1.0; C3(b->dcxu, 1, nx) = 0.0; C3(b->dcxu, 3, nx) = 0.0;
  for (int i = 2; i <= ny; i++) {
    a = dyi[i]; bb = dyi[i - 1]; c = dy[i - 1];
    double den = -c * c * a - c * c * bb + 12.0 * bb * bb * a + 12.0 * bb * a * a;
    C3(b->dcyu, 1, i) = (bb * c * c) / den;
    C3(b->dcyu, 2, i) = 1.0;
    C3(b->dcyu, 3, i) = (c * c * a) / den;
  }
  C3(b->dcyu, 2, 1) = 1.0; C3(b->dcyu, 3, 1) = 0.0; C3(b->dcyu, 1, 1) = 0.0;
  C3(b->dcyu, 2, ny + 1) = 1.0; C3(b->dcyu, 3, ny + 1) = 0.0;
  C3(b->dcyu, 1, ny + 1) = 0.0;
  for (int i = 2; i <= nx - 1; i++) {
    a = dxi[i]; bb = dxi[i - 1]; c = dx[i - 1];
    double den = -c * c * a - c * c * bb + 12.0 * bb * bb * a + 12.0 * bb * a * a;
    C3(b->dcxv, 1, i) = (bb * c * c) / den;
    C3(b->dcxv, 2, i) = 1.0;
    C3(b->dcxv, 3, i) = (c * c * a) / den;
  }
  C3(b->dcxv, 2, 1) = 1.0; C3(b->dcxv, 3, 1) = 0.0; C3(b->dcxv, 1, 1) = 0.0;
  C3(b->dcxv, 2, nx) = 1.0; C3(b->dcxv, 3, nx) = 0.0; C3(b->dcxv, 1, nx) = 0.0;
  for (int i = 2; i <= ny - 1; i++) {
    a = dy[i]; bb = dy[i - 1];
    aux = (2.0 * bb * bb * bb + 2.0 * a * a * a + 9.0 * bb * bb * a + 9.0 * bb * a * a);
    C3(b->dcyv, 1, i) = -a * (-2.0 * a * bb - bb * bb + 2.0 * a * a) / aux;
    C3(b->dcyv, 2, i) = 1.0;
    C3(b->dcyv, 3, i) = ((a * a + 2.0 * a * bb - 2.0 * bb * bb) * bb) / aux;
  }
  C3(b->dcyv, 2, 1) = 1.0; C3(b->dcyv, 3, 1) = 0.0; C3(b->dcyv, 1, 1) = 0.0;
  C3(b->dcyv, 2, ny) = 1.0; C3(b->dcyv, 3, ny) = 0.0; C3(b->dcyv, 1, ny) = 0.0;
  /* first derivative boundary closures (coeft.f90:787-924) */
  coefb(4, dx[1], dxi[2], dxi[3], dxi[4], bcc, 3);
  for (int q = 1; q <= 4; q++) M5(b->dcbx, 1, q) = bcc[q - 1];
  coefb(4, dx[0], dx[1], dx[2], dx[3], bcc, 3);
  for (int q = 1; q <= 4; q++) M5(b->dcbx, 2, q) = bcc[q - 1];
  coefb(4, dx[nx], dx[nx - 1], dxi[nx - 1], dxi[nx - 2], bcc, 5);
  for (int q = 1; q <= 4; q++) M5(b->dcbx, 3, q) = bcc[q - 1];
  coefb(4, dx[nx - 1], dx[nx - 1], dx[nx - 2], dx[nx - 3], bcc, 5);
  for (int q = 1; q <= 4; q++) M5(b->dcbx, 4, q) = bcc[q - 1];
  coefb(4, -dx[nx - 1], -dxi[nx - 1], -dxi[nx - 2], -dxi[nx - 3], bcc, 5);
  for (int q = 1; q <= 4; q++) M5(b->dcbx, 5, q) = bcc[q - 1];
  coefb(4, -dy[0], -dy[1], -dyi[2], -dyi[3], bcc, 5);
  for (int q = 1; q <= 4; q++) M4(b->dcby, 1, q) = bcc[q - 1];
  coefb(4, dy[0], dy[1], dy[2], dy[3], bcc, 3);
  for (int q = 1; q <= 4; q++) M4(b->dcby, 3, q) = bcc[q - 1];
  coefb(4, -dy[ny], -dy[ny - 1], -dy[ny - 2], -dy[ny - 3], bcc, 3);
  for (int q = 1; q <= 4; q++) M4(b->dcby, 4, q) = bcc[q - 1];
  coefb(4, -dy[ny - 1], -dyi[ny - 1], -dyi[ny - 2], -dyi[ny - 3], bcc, 3);
  for (int q = 1; q <= 4; q++) M4(b->dcby, 2, q) = bcc[q - 1];
  coefb(4, dy[ny], dy[ny - 1], dyi[ny - 1], dyi[ny - 2], bcc, 7);
  for (int q = 1; q <= 4; q++) M2(b->ccon, 1, q) = bcc[q - 1];
  coefb(4, -dy[0], -dyi[1], -dyi[2], 0.0, bcc, 6);
  for (int q = 1; q <= 4; q++) M2(b->ccon, 2, q) = bcc[q - 1];
  coefinv(b->dcxu, nx);
  coefinv(b->dcxv, nx);
  coefinv(b->dcyu, ny + 1);
  coefinv(b->dcyv, ny);
  /* first derivative, explicit part (coeft.f90:933-965) */
  for (int i = 2; i <= ny - 1; i++) {
    a = dy[i]; bb = dy[i - 1];
    double den = 2.0 * bb * bb * bb + 2.0 * a * a * a + 9.0 * bb * bb * a + 9.0 * bb * a * a;
    C2(b->cofcyv, 1, i) = -24 * a * bb / den;
    C2(b->cofcyv, 2, i) = 24 * a * bb / den;
  }
  for (int i = 2; i <= ny; i++) {
    a = dyi[i]; bb = dyi[i - 1]; c = dy[i - 1];
    C2(b->cofcyu, 1, i) = -12.0 * a * bb / ((12.0 * a * bb - c * c) * c);
    C2(b->cofcyu, 2, i) = 12.0 * a * bb / ((12.0 * a * bb - c * c) * c);
  }
  for (int i = 2; i <= nx - 1; i++) {
    a = dxi[i]; bb = dxi[i - 1]; c = dx[i - 1];
    C2(b->cofcxv, 1, i) = -12.0 * a * bb / ((12.0 * a * bb - c * c) * c);
    C2(b->cofcxv, 2, i) = 12.0 * a * bb / ((12.0 * a * bb - c * c) * c);
    a = dx[i]; bb = dx[i - 1];
    double den = 2.0 * bb * bb * bb + 2.0 * a * a * a + 9.0 * bb * bb * a + 9.0 * bb * a * a;
    C2(b->cofcxu, 1, i) = -24 * a * bb / den;
    C2(b->cofcxu, 2, i) = 24 * a * bb / den;
  }

  /* second derivative, implicit part (coeft.f90:971-1019) */
#define VI_ROW(arr, i, A, B)                                                   \
  do {                                                                         \
    double a_ = (A), b_ = (B);                                                 \
    double den_ = b_ * b_ * b_ + a_ * a_ * a_ + 4 * a_ * a_ * b_ + 4 * a_ * b_ * b_; \
    C3(arr, 1, i) = b_ * (a_ * a_ + a_ * b_ - b_ * b_) / den_;                 \
    C3(arr, 2, i) = 1.0;                                                       \
    C3(arr, 3, i) = -(a_ * a_ - a_ * b_ - b_ * b_) * a_ / den_;                \
  } while (0)
  for (int i = 2; i <= nx - 1; i++) VI_ROW(b->vixu, i, dx[i], dx[i - 1]);
  C3(b->vixu, 1, 1) = 0.0; C3(b->vixu, 2, 1) = 1.0;
  C3(b->vixu, 3, nx) = 0.0; C3(b->vixu, 2, nx) = 1.0;
  for (int i = 2; i <= ny; i++) VI_ROW(b->viyu, i, dyi[i], dyi[i - 1]);
  C3(b->viyu, 1, 1) = 0.0; C3(b->viyu, 2, 1) = 1.0;
  C3(b->viyu, 3, ny + 1) = 0.0; C3(b->viyu, 2, ny + 1) = 1.0;
  for (int i = 2; i <= nx - 1; i++) VI_ROW(b->vixv, i, dxi[i], dxi[i - 1]);
  C3(b->vixv, 1, 1) = 0.0; C3(b->vixv, 2, 1) = 1.0;
  C3(b->vixv, 3, nx) = 0.0; C3(b->vixv, 2, nx) = 1.0;
  for (int i = 2; i <= ny - 1; i++) VI_ROW(b->viyv, i, dy[i], dy[i - 1]);
  C3(b->viyv, 1, 1) = 0.0; C3(b->viyv, 2, 1) = 1.0;
  C3(b->viyv, 3, ny) = 0.0; C3(b->viyv, 2, ny) = 1.0;
#undef VI_ROW
  /* second derivative boundary closures (coeft.f90:1023-1113) */
  coefb(5, dxi[1], dxi[2], dxi[3], 0.0, bcc, 1);
  C3(b->vixv, 3, 1) = bcc[0];
  for (int q = 1; q <= 4; q++) M4(b->cofvbx, 1, q) = bcc[q];
  coefb(5, dxi[nx - 1], dxi[nx - 2], dxi[nx - 3], 0.0, bcc, 1);
  C3(b->vixv, 1, nx) = bcc[0];
  for (int q = 1; q <= 4; q++) M4(b->cofvbx, 2, q) = bcc[q];
  coefb(5, dx[1], dx[2], dx[3], 0.0, bcc, 1);
  C3(b->vixu, 3, 1) = bcc[0];
  for (int q = 1; q <= 4; q++) M4(b->cofvbx, 3, q) = bcc[q];
  coefb(5, dx[nx - 1], dx[nx - 2], dx[nx - 3], 0.0, bcc, 1);
  C3(b->vixu, 1, nx) = bcc[0];
  for (int q = 1; q <= 4; q++) M4(b->cofvbx, 4, q) = bcc[q];
  coefb(5, dy[1], dy[2], dy[3], 0.0, bcc, 1);
  C3(b->viyv, 3, 1) = bcc[0];
  for (int q = 1; q <= 4; q++) M4(b->cofvby, 1, q) = bcc[q];
  coefb(5, dy[ny - 1], dy[ny - 2], dy[ny - 3], 0.0, bcc, 1);
  C3(b->viyv, 1, ny) = bcc[0];
  for (int q = 1; q <= 4; q++) M4(b->cofvby, 2, q) = bcc[q];
  coefb(5, dyi[1], dyi[2], dyi[3], 0.0, bcc, 1);
  C3(b->viyu, 3, 1) = bcc[0];
  for (int q = 1; q <= 4; q++) M4(b->cofvby, 3, q) = bcc[q];
  coefb(5, dyi[ny], dyi[ny - 1], dyi[ny - 2], 0.0, bcc, 1);
  C3(b->viyu, 1, ny + 1) = bcc[0];
  for (int q = 1; q <= 4; q++) M4(b->cofvby, 4, q) = bcc[q];
  /* unfactored copies for the implicit solve (coeft.f90:1116-1127) */
  for (int i = 1; i <= ny; i++)
    for (int r = 1; r <= 3; r++) {
      C3(b->vyui, r, i) = C3(b->viyu, r, i);
      C3(b->vyvi, r, i) = C3(b->viyv, r, i);
    }
  for (int r = 1; r <= 3; r++) C3(b->vyui, r, ny + 1) = C3(b->viyu, r, ny + 1);
  coefinv(b->vixu, nx);
  coefinv(b->vixv, nx);
  coefinv(b->viyu, ny + 1);
  coefinv(b->viyv, ny);
  /* second derivative, explicit part (coeft.f90:1136-1168) */
#define CV_ROW(arr, i, A, B)                                                   \
  do {                                                                         \
    double a_ = (A), b_ = (B);                                                 \
    double den_ = b_ * b_ * b_ + a_ * a_ * a_ + 4 * a_ * a_ * b_ + 4 * a_ * b_ * b_; \
    C3(arr, 1, i) = 12 * b_ / den_;                                            \
    C3(arr, 2, i) = -12 / (a_ * a_ + 3 * a_ * b_ + b_ * b_);                   \
    C3(arr, 3, i) = 12 * a_ / den_;                                            \
  } while (0)
  for (int i = 2; i <= ny - 1; i++) CV_ROW(b->cofvyv, i, dy[i], dy[i - 1]);
  for (int i = 2; i <= ny; i++) CV_ROW(b->cofvyu, i, dyi[i], dyi[i - 1]);
  for (int i = 2; i <= nx - 1; i++) {
    CV_ROW(b->cofvxv, i, dxi[i], dxi[i - 1]);
    CV_ROW(b->cofvxu, i, dx[i], dx[i - 1]);
  }
#undef CV_ROW
  /* Poisson operators (coeft.f90:1181-1186) */
  for (int i = 1; i <= nx - 1; i++)
    b->kmod[i] = 2.0 / (dx[1] * dx[1]) * (1 - cos((pi * (i - 1)) / (double)(nx - 1)));
  mtr(b);
}

/* inlet_retheta (bl_1.F90:372-396).  0.99 is a default-real literal. */
static void inlet_retheta(oc_box *b) {
  int ny = b->ny;
  double uinf = 0.0;
  for (int j = ny + 1 - 20; j <= ny + 1 - 5; j++) uinf += b->u0[j];
  uinf /= 16.0;
  double u99 = (double)0.99f * uinf;
  int jtop = 0;
  for (int j = 1; j <= ny - 2; j++) {
    if (b->u0[j] >= u99) break;
    jtop = j + 1;
  }
  b->din = b->y[jtop];
  double s = 0.0;
  for (int j = 2; j <= jtop + 3; j++) s += (uinf - b->u0[j]) * b->u0[j] * b->dy[j - 1];
  b->rthin = s * b->re / uinf;
}

/* magicnumber (boun.f90:102-143): mean v at the top from the Re_theta(x)
 * correlation.  dx here is ax*pi/(nx-1), unlike coef. */
static void magicnumber(oc_box *b) {
  int nx = b->nx;
  double e1 = 0.8659, c1 = 0.01277, ei1 = 1 / e1, kap = 0.384, ckap = 4.127,
         cprim = 7.135;
  double dxm = b->ax * b->pi / (nx - 1);
  double *xx = zalloc(nx + 1), *h = zalloc(nx + 1);
  for (int i = 1; i <= nx; i++) {
    xx[i] = dxm * (i - 1);
    double reth = pow(pow(b->rthin, ei1) + pow(c1, ei1) * b->re * xx[i], e1);
    double uep = log(reth) / kap + ckap;
    h[i] = reth / ((1 - cprim / uep) * b->re);
  }
  b->vmagic[1] = (h[2] - h[1]) / (xx[2] - xx[1]);
  for (int i = 2; i <= nx - 1; i++)
    b->vmagic[i] = (h[i + 1] - h[i - 1]) / (xx[i + 1] - xx[i - 1]);
  b->vmagic[nx] = (h[nx] - h[nx - 1]) / (xx[nx] - xx[nx - 1]);
  free(xx);
  free(h);
}

/* ------------------------------------------------------------------------
 * z transforms (ffts.f90:27-98): mixed-radix complex DFT (own code; the
 * reference calls FFTW).  c2r drops Im of mode 0 as a real inverse does.
 * ---------------------------------------------------------------------- */
static int small_factor(int n) {
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0) return p;
  return n;
}
/* y[k] = sum_j x[j*xs] exp(sgn 2 pi i j k / n); interleaved complex. */
static void fft_rec(int n, int N, const double *tw, const double *x, int xs,
                    double *y, int sgn, double *tmp) {
  if (n == 1) {
    y[0] = x[0];
    y[1] = x[1];
    return;
  }
  int p = small_factor(n), m = n / p, step = N / n;
  for (int r = 0; r < p; r++) fft_rec(m, N, tw, x + 2 * (size_t)r * xs, xs * p, y + 2 * r * m, sgn, tmp);
  if (p == 2) { /* radix 2: y0[k] +- w^k y1[k] */
    for (int k = 0; k < m; k++) {
      double wr = tw[2 * k * step], wi = sgn * tw[2 * k * step + 1];
      double ar = y[2 * k], ai = y[2 * k + 1], br = y[2 * (m + k)], bi = y[2 * (m + k) + 1];
      double tr = wr * br - wi * bi, ti = wr * bi + wi * br;
      tmp[2 * k] = ar + tr;
      tmp[2 * k + 1] = ai + ti;
      tmp[2 * (k + m)] = ar - tr;
      tmp[2 * (k + m) + 1] = ai - ti;
    }
    memcpy(y, tmp, 2 * sizeof(double) * n);
    return;
  }
  for (int q = 0; q < p; q++)
    for (int k = 0; k < m; k++) {
      int kk = k + m * q;
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; r++) {
        int t = (int)(((long)r * kk) % n) * step;
        double wr = tw[2 * t], wi = sgn * tw[2 * t + 1];
        double ar = y[2 * (r * m + k)], ai = y[2 * (r * m + k) + 1];
        sr += wr * ar - wi * ai;
        si += wr * ai + wi * ar;
      }
      tmp[2 * kk] = sr;
      tmp[2 * kk + 1] = si;
    }
  memcpy(y, tmp, 2 * sizeof(double) * n);
}

/* The same transform (x contiguous, N == n) by the iterative radix-2
 * decimation in time for powers of two, else by fft_rec. */
static void fft_any(int n, const double *tw, const double *x, double *y, int sgn, double *tmp) {
  if (n & (n - 1)) {
    fft_rec(n, n, tw, x, 1, y, sgn, tmp);
    return;
  }
  int lg = 0;
  while ((1 << lg) < n) lg++;
  for (int i = 0; i < n; i++) {
    int r = 0;
    for (int q = 0; q < lg; q++) r |= ((i >> q) & 1) << (lg - 1 - q);
    y[2 * r] = x[2 * i];
    y[2 * r + 1] = x[2 * i + 1];
  }
  for (int len = 2; len <= n; len *= 2) {
    const int half = len / 2, step = n / len;
    for (int i0 = 0; i0 < n; i0 += len)
      for (int k = 0; k < half; k++) {
        double wr = tw[2 * k * step], wi = sgn * tw[2 * k * step + 1];
        double *a = y + 2 * (i0 + k), *c = y + 2 * (i0 + k + half);
        double tr = wr * c[0] - wi * c[1], ti = wr * c[1] + wi * c[0];
        c[0] = a[0] - tr;
        c[1] = a[1] - ti;
        a[0] += tr;
        a[1] += ti;
      }
  }
}

void oc_c2r(const oc_box *b, const double *fou, double *phys) {
  int nz = b->nz;
  struct oc_tls *t = TLS(b);
  double *a = t->fft_a, *o = t->fft_b;
  memset(a, 0, 2 * sizeof(double) * nz);
  a[0] = fou[0]; /* Im of mode 0 does not enter a real inverse transform */
  for (int k = 1; k <= b->nz2; k++) {
    a[2 * k] = fou[2 * k];
    a[2 * k + 1] = fou[2 * k + 1];
    a[2 * (nz - k)] = fou[2 * k];
    a[2 * (nz - k) + 1] = -fou[2 * k + 1];
  }
  fft_any(nz, b->fft_tw, a, o, +1, t->fft_t);
  for (int n = 0; n < nz; n++) phys[n] = o[2 * n];
}

void oc_r2c(const oc_box *b, const double *phys, double *fou) {
  int nz = b->nz;
  struct oc_tls *t = TLS(b);
  double *a = t->fft_a, *o = t->fft_b;
  for (int n = 0; n < nz; n++) {
    a[2 * n] = phys[n];
    a[2 * n + 1] = 0.0;
  }
  fft_any(nz, b->fft_tw, a, o, -1, t->fft_t);
  double dnf = 1.0 / nz;
  for (int k = 0; k <= b->nz2; k++) {
    fou[2 * k] = dnf * o[2 * k];
    fou[2 * k + 1] = dnf * o[2 * k + 1];
  }
  fou[1] = 0.0;
}

/* Complex DFT of any length n, y[k] = sum_j x[j] exp(sgn 2 pi i j k / n)
 * (interleaved complex): the mixed-radix fft_rec when every prime factor of
 * n is at most 64, else Bluestein's chirp-z identity
 *   y[k] = conj(c_k) sum_j (x_j conj(c_j)) c_{k-j},  c_t = exp(-sgn i pi t^2 / n)
 * (2jk = j^2 + k^2 - (k-j)^2)
 * evaluated as a cyclic convolution of power-of-two length M >= 2n - 1 with
 * fft_rec (own code; the reference calls FFTW). */
typedef struct {
  int n, M, blue;
  double *tw, *twM, *chirp, *cf, *a, *o, *t;
} dft_plan;
static int max_prime(int n) {
  int m = 1;
  for (int p = 2; p * p <= n; p++)
    while (n % p == 0) {
      m = p;
      n /= p;
    }
  return n > 1 ? (n > m ? n : m) : m;
}
static void dft_init(dft_plan *q, int n, int sgn) {
  memset(q, 0, sizeof *q);
  q->n = n;
  q->blue = max_prime(n) > 64;
  int L = n;
  if (q->blue) {
    q->M = 1;
    while (q->M < 2 * n - 1) q->M *= 2;
    L = q->M;
  }
  q->tw = zalloc(2 * (size_t)n);
  for (int t = 0; t < n; t++) {
    q->tw[2 * t] = cos(2.0 * PI_D * t / n);
    q->tw[2 * t + 1] = sin(2.0 * PI_D * t / n);
  }
  q->a = zalloc(2 * (size_t)L);
  q->o = zalloc(2 * (size_t)L);
  q->t = zalloc(2 * (size_t)L);
  if (!q->blue) return;
  const int M = q->M;
  q->twM = zalloc(2 * (size_t)M);
  for (int t = 0; t < M; t++) {
    q->twM[2 * t] = cos(2.0 * PI_D * t / M);
    q->twM[2 * t + 1] = sin(2.0 * PI_D * t / M);
  }
  /* c_t = exp(-sgn i pi t^2 / n), t^2 reduced mod 2n */
  q->chirp = zalloc(2 * (size_t)n);
  for (int t = 0; t < n; t++) {
    long r = ((long)t * t) % (2L * n);
    q->chirp[2 * t] = cos(PI_D * r / n);
    q->chirp[2 * t + 1] = -sgn * sin(PI_D * r / n);
  }
  /* forward transform of the wrapped chirp, divided by M */
  double *h = zalloc(2 * (size_t)M);
  for (int t = 0; t < n; t++) {
    h[2 * t] = q->chirp[2 * t];
    h[2 * t + 1] = q->chirp[2 * t + 1];
    if (t) {
      h[2 * (M - t)] = q->chirp[2 * t];
      h[2 * (M - t) + 1] = q->chirp[2 * t + 1];
    }
  }
  q->cf = zalloc(2 * (size_t)M);
  fft_any(M, q->twM, h, q->cf, -1, q->t);
  for (int t = 0; t < 2 * M; t++) q->cf[t] /= M;
  free(h);
}
static void dft_free(dft_plan *q) {
  free(q->tw);
  free(q->twM);
  free(q->chirp);
  free(q->cf);
  free(q->a);
  free(q->o);
  free(q->t);
}
/* x -> y (both 2n doubles; may alias) */
static void dft_run(dft_plan *q, const double *x, double *y, int sgn) {
  const int n = q->n;
  if (!q->blue) {
    fft_any(n, q->tw, x, q->o, sgn, q->t);
    memcpy(y, q->o, 2 * sizeof(double) * n);
    return;
  }
  const int M = q->M;
  double *a = q->a, *o = q->o;
  memset(a, 0, 2 * sizeof(double) * M);
  for (int j = 0; j < n; j++) { /* x_j conj(c_j) */
    double cr = q->chirp[2 * j], ci = q->chirp[2 * j + 1];
    a[2 * j] = x[2 * j] * cr + x[2 * j + 1] * ci;
    a[2 * j + 1] = x[2 * j + 1] * cr - x[2 * j] * ci;
  }
  fft_any(M, q->twM, a, o, -1, q->t);
  for (int m = 0; m < M; m++) {
    double ar = o[2 * m], ai = o[2 * m + 1], br = q->cf[2 * m], bi = q->cf[2 * m + 1];
    a[2 * m] = ar * br - ai * bi;
    a[2 * m + 1] = ar * bi + ai * br;
  }
  fft_any(M, q->twM, a, o, +1, q->t);
  for (int k = 0; k < n; k++) { /* conj(c_k) times the convolution */
    double cr = q->chirp[2 * k], ci = q->chirp[2 * k + 1], zr = o[2 * k], zi = o[2 * k + 1];
    y[2 * k] = zr * cr + zi * ci;
    y[2 * k + 1] = zi * cr - zr * ci;
  }
}

/* cosftx (ffts.f90:147-181): FFTW REDFT10 x 1/n (iopt<0) and REDFT01 x 0.5
 * (iopt>0), with REDFT10 Y_k = 2 sum_j x_j cos(pi k (2j+1) / 2n) and REDFT01
 * Y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (2k+1) / 2n) (FFTW's definitions).
 * Each line is one complex DFT of length n of Makhoul's even/odd reordering
 * v_j = x_{2j}, v_{n-1-j} = x_{2j+1}:
 *   sum_j x_j cos(pi k (2j+1) / 2n) = Re(exp(-i pi k / 2n) V_k),
 * and REDFT01 is the unnormalised inverse of that map.  Line l starts at
 * c + l*stride. */
static void cosft_lines(int n, int nlines, double *c, size_t stride, int iopt) {
  double *e = zalloc(2 * (size_t)n);
  for (int k = 0; k < n; k++) {
    e[2 * k] = cos(PI_D * k / (2.0 * n));
    e[2 * k + 1] = sin(PI_D * k / (2.0 * n));
  }
#pragma omp parallel
  {
  double *z = zalloc(2 * (size_t)n);
  dft_plan q;
  dft_init(&q, n, iopt < 0 ? -1 : +1);
#pragma omp for schedule(static)
  for (int l = 0; l < nlines; l++) {
    double *x = c + (size_t)l * stride;
    if (iopt < 0) {
      for (int j = 0; 2 * j < n; j++) {
        z[2 * j] = x[2 * j];
        z[2 * j + 1] = 0.0;
      }
      for (int j = 0; 2 * j + 1 < n; j++) {
        z[2 * (n - 1 - j)] = x[2 * j + 1];
        z[2 * (n - 1 - j) + 1] = 0.0;
      }
      dft_run(&q, z, z, -1);
      /* Re(exp(-i pi k/2n) V_k), times 2 (REDFT10) and 1/n (cosftx) */
      for (int k = 0; k < n; k++) x[k] = 2.0 / n * (e[2 * k] * z[2 * k] + e[2 * k + 1] * z[2 * k + 1]);
    } else {
      /* V_k = exp(i pi k/2n) (X_k - i X_{n-k}), V_0 = X_0 */
      z[0] = x[0];
      z[1] = 0.0;
      for (int k = 1; k < n; k++) {
        double ar = x[k], ai = -x[n - k];
        z[2 * k] = e[2 * k] * ar - e[2 * k + 1] * ai;
        z[2 * k + 1] = e[2 * k] * ai + e[2 * k + 1] * ar;
      }
      dft_run(&q, z, z, +1);
      /* y_{2j} = z_j, y_{2j+1} = z_{n-1-j}, times 0.5 (cosftx) */
      for (int j = 0; 2 * j < n; j++) x[2 * j] = 0.5 * z[2 * j];
      for (int j = 0; 2 * j + 1 < n; j++) x[2 * j + 1] = 0.5 * z[2 * (n - 1 - j)];
    }
  }
  dft_free(&q);
  free(z);
  }
  free(e);
}
void oc_cosft(int n, int nlines, double *c, int iopt) { cosft_lines(n, nlines, c, (size_t)n, iopt); }

/* ------------------------------------------------------------------------
 * Plane <-> pencil transposes (change.f90:224-275).  Pencil p (0-based) is
 * plane element p of every plane; pencil storage is (nx, nlin).  With
 * fp32_transposes the values pass through single precision
 * (change.f90:565-686).
 * ---------------------------------------------------------------------- */
static void chp2x(const oc_box *b, double *pen, const double *plan, int R) {
  size_t P = (size_t)b->nz1 * R;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < b->nx; i++)
    for (size_t p = 0; p < P; p++) {
      double v = plan[(size_t)i * P + p];
      pen[p * b->nx + i] = b->fp32_transposes ? f32(v) : v;
    }
}
static void chx2p(const oc_box *b, double *plan, const double *pen, int R) {
  size_t P = (size_t)b->nz1 * R;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < b->nx; i++)
    for (size_t p = 0; p < P; p++) {
      double v = pen[p * b->nx + i];
      plan[(size_t)i * P + p] = b->fp32_transposes ? f32(v) : v;
    }
}

/* ------------------------------------------------------------------------
 * Compact line operators (lip.f90)
 * ---------------------------------------------------------------------- */
/* interpxx (lip.f90:248-300) on pencils (nx, nlin). */
static void interpxx(const oc_box *b, const double *gen, double *geni,
                     const double *icon, const double *econ, const double *inbx,
                     int chi, int nlin, int loguv) {
  int nx = b->nx, chi2 = (loguv == 1) ? chi + 1 : chi + 3;
#pragma omp parallel
  {
  double *bb = zalloc(nx + 1);
#pragma omp for schedule(static)
  for (int k = 0; k < nlin; k++) {
    const double *g = gen + (size_t)k * nx - 1; /* g[1..nx] */
    for (int i = 2; i <= nx - 1; i++) bb[i] = C2(econ, 1, i) * g[i] + C2(econ, 2, i) * g[i + 1];
    bb[1] = M4(inbx, chi2, 1) * g[1] + M4(inbx, chi2, 2) * g[2] + M4(inbx, chi2, 3) * g[3] +
            M4(inbx, chi2, 4) * g[4];
    bb[nx] = M4(inbx, chi, 1) * g[nx] + M4(inbx, chi, 2) * g[nx - 1] +
             M4(inbx, chi, 3) * g[nx - 2] + M4(inbx, chi, 4) * g[nx - 3];
    for (int i = 2; i <= nx; i++) bb[i] = bb[i] + C3(icon, 1, i) * bb[i - 1];
    bb[nx] = bb[nx] * C3(icon, 2, nx);
    for (int i = nx - 1; i >= 1; i--) bb[i] = (bb[i] - C3(icon, 3, i) * bb[i + 1]) * C3(icon, 2, i);
    double *o = geni + (size_t)k * nx - 1;
    for (int i = 1; i <= nx; i++) o[i] = bb[i];
  }
  free(bb);
  }
}

/* difvisxx (lip.f90:52-142) on pencils; result may alias upen_int. */
static void difvisxx(const oc_box *b, const double *upen_int, const double *upencil,
                     double *result, const double *icon, const double *iconv,
                     const double *dcbx, const double *econ, const double *econv,
                     const double *cofvbx, int chv, int nlin, int vel, double rex) {
  int nx = b->nx, chv1 = chv + 2, ic = (vel == 0) ? 1 : 3;
#pragma omp parallel
  {
  double *bb = zalloc(nx + 1), *cv = zalloc(nx + 1);
#pragma omp for schedule(static)
  for (int k = 0; k < nlin; k++) {
    const double *g = upen_int + (size_t)k * nx - 1;
    const double *q = upencil + (size_t)k * nx - 1;
    for (int i = 2; i <= nx - 1; i++) bb[i] = C2(econ, 1, i) * g[i - 1] + C2(econ, 2, i) * g[i];
    bb[1] = M5(dcbx, chv, 1) * g[1] + M5(dcbx, chv, 2) * g[2] + M5(dcbx, chv, 3) * g[3] +
            M5(dcbx, chv, 4) * g[4];
    bb[nx] = M5(dcbx, chv1, 1) * g[nx] + M5(dcbx, chv1, 2) * g[nx - 1] +
             M5(dcbx, chv1, 3) * g[nx - 2] + M5(dcbx, chv1, 4) * g[nx - 3];
    for (int j = 2; j <= nx; j++) bb[j] = bb[j] + C3(icon, 1, j) * bb[j - 1];
    bb[nx] = bb[nx] * C3(icon, 2, nx);
    for (int j = nx - 1; j >= 1; j--) bb[j] = (bb[j] - C3(icon, 3, j) * bb[j + 1]) * C3(icon, 2, j);
    bb[nx] = 1.0 * (q[nx] - q[nx - 1]) * b->idx; /* Uinfinity = 1 */
    for (int i = 1; i <= nx; i++) cv[i] = bb[i];
    for (int i = 2; i <= nx - 1; i++)
      bb[i] = C3(econv, 1, i) * q[i + 1] + C3(econv, 2, i) * q[i] + C3(econv, 3, i) * q[i - 1];
    bb[1] = M4(cofvbx, ic, 1) * q[1] + M4(cofvbx, ic, 2) * q[2] + M4(cofvbx, ic, 3) * q[3] +
            M4(cofvbx, ic, 4) * q[4];
    bb[nx] = M4(cofvbx, ic + 1, 1) * q[nx] + M4(cofvbx, ic + 1, 2) * q[nx - 1] +
             M4(cofvbx, ic + 1, 3) * q[nx - 2] + M4(cofvbx, ic + 1, 4) * q[nx - 3];
    for (int i = 2; i <= nx; i++) bb[i] = bb[i] + C3(iconv, 1, i) * bb[i - 1];
    bb[nx] = bb[nx] * C3(iconv, 2, nx);
    for (int i = nx - 1; i >= 1; i--) bb[i] = (bb[i] - C3(iconv, 3, i) * bb[i + 1]) * C3(iconv, 2, i);
    bb[nx] = 0.0;
    double *o = result + (size_t)k * nx - 1;
    for (int i = 1; i <= nx; i++) o[i] = cv[i] + rex * bb[i];
  }
  free(bb);
  free(cv);
  }
}

/* Row access for y-operators: element (k, j) of a (stride, rows) block. */
#define RW(a, s, k, j) (a)[(size_t)((j)-1) * (s) + (k)]

/* differzy (lip.f90:152-232); columns k = 0..ncol-1. */
static void differzy(const oc_box *b, const double *geni, int gs, double *result,
                     int rs, int ncol, const double *icon, const double *dcby,
                     const double *econ, int chv, int fwb1, int fwb2, int jee) {
  int ny = b->ny, cha = chv + 1;
  for (int k = 0; k < ncol; k++) {
    double bot = M4(dcby, chv, 1) * RW(geni, gs, k, 1) + M4(dcby, chv, 2) * RW(geni, gs, k, 2) +
                 M4(dcby, chv, 3) * RW(geni, gs, k, 3) + M4(dcby, chv, 4) * RW(geni, gs, k, 4);
    double top = M4(dcby, cha, 1) * RW(geni, gs, k, ny) + M4(dcby, cha, 2) * RW(geni, gs, k, ny - 1) +
                 M4(dcby, cha, 3) * RW(geni, gs, k, ny - 2) + M4(dcby, cha, 4) * RW(geni, gs, k, ny - 3);
    /* in place is allowed: keep the loop direction of the reference */
    if (fwb2 == 1) {
      for (int j = 2; j <= jee - 1; j++)
        RW(result, rs, k, j) = C2(econ, 2, j) * RW(geni, gs, k, j + fwb2) + C2(econ, 1, j) * RW(geni, gs, k, j + fwb1);
    } else {
      for (int j = jee - 1; j >= 2; j--)
        RW(result, rs, k, j) = C2(econ, 2, j) * RW(geni, gs, k, j + fwb2) + C2(econ, 1, j) * RW(geni, gs, k, j + fwb1);
    }
    RW(result, rs, k, 1) = bot;
    RW(result, rs, k, jee) = top;
    for (int j = 2; j <= jee; j++) RW(result, rs, k, j) += C3(icon, 1, j) * RW(result, rs, k, j - 1);
    RW(result, rs, k, jee) *= C3(icon, 2, jee);
    for (int j = jee - 1; j >= 1; j--)
      RW(result, rs, k, j) = (RW(result, rs, k, j) - C3(icon, 3, j) * RW(result, rs, k, j + 1)) * C3(icon, 2, j);
  }
}

/* interpyy (lip.f90:359-440); columns k = 0..m-1, strides gs/os. */
static void interpyy(const oc_box *b, const double *gen, int gs, double *geni, int os,
                     const double *icon, const double *econ, const double *inby,
                     int loguv, int m) {
  int ny = b->ny;
  for (int k = 0; k < m; k++) {
    if (loguv == 0) {
      double buf = M4(inby, 1, 1) * RW(gen, gs, k, ny + 1) + M4(inby, 1, 2) * RW(gen, gs, k, ny) +
                   M4(inby, 1, 3) * RW(gen, gs, k, ny - 1) + M4(inby, 1, 4) * RW(gen, gs, k, ny - 2);
      RW(geni, os, k, 1) = 0.0;
      for (int j = 2; j <= ny - 1; j++)
        RW(geni, os, k, j) = C2(econ, 1, j) * RW(gen, gs, k, j) + C2(econ, 2, j) * RW(gen, gs, k, j + 1);
      RW(geni, os, k, ny) = buf;
    } else {
      double buf = M4(inby, 3, 1) * RW(gen, gs, k, 1) + M4(inby, 3, 2) * RW(gen, gs, k, 2) +
                   M4(inby, 3, 3) * RW(gen, gs, k, 3) + M4(inby, 3, 4) * RW(gen, gs, k, 4);
      RW(geni, os, k, ny) = M4(inby, 4, 1) * RW(gen, gs, k, ny) + M4(inby, 4, 2) * RW(gen, gs, k, ny - 1) +
                            M4(inby, 4, 3) * RW(gen, gs, k, ny - 2) + M4(inby, 4, 4) * RW(gen, gs, k, ny - 3);
      for (int j = ny - 1; j >= 2; j--)
        RW(geni, os, k, j) = C2(econ, 2, j) * RW(gen, gs, k, j) + C2(econ, 1, j) * RW(gen, gs, k, j - 1);
      RW(geni, os, k, 1) = buf;
    }
    for (int j = 2; j <= ny; j++) RW(geni, os, k, j) += C3(icon, 1, j) * RW(geni, os, k, j - 1);
    RW(geni, os, k, ny) *= C3(icon, 2, ny);
    for (int j = ny - 1; j >= 1; j--)
      RW(geni, os, k, j) = (RW(geni, os, k, j) - C3(icon, 3, j) * RW(geni, os, k, j + 1)) * C3(icon, 2, j);
  }
}

/* vistyy (lip.f90:447-513) on one plane (nz1 columns). */
static void vistyy(const oc_box *b, const double *u, double *result, const double *icon,
                   const double *econ, const double *cofvby, int jee) {
  int s = b->nz1, ic = (jee == b->ny) ? 1 : 3;
  for (int k = 0; k < s; k++) {
    RW(result, s, k, 1) = M4(cofvby, ic, 1) * RW(u, s, k, 1) + M4(cofvby, ic, 2) * RW(u, s, k, 2) +
                          M4(cofvby, ic, 3) * RW(u, s, k, 3) + M4(cofvby, ic, 4) * RW(u, s, k, 4);
    for (int j = 2; j <= jee - 1; j++)
      RW(result, s, k, j) = C3(econ, 1, j) * RW(u, s, k, j + 1) + C3(econ, 2, j) * RW(u, s, k, j) +
                            C3(econ, 3, j) * RW(u, s, k, j - 1);
    RW(result, s, k, jee) = M4(cofvby, ic + 1, 1) * RW(u, s, k, jee) + M4(cofvby, ic + 1, 2) * RW(u, s, k, jee - 1) +
                            M4(cofvby, ic + 1, 3) * RW(u, s, k, jee - 2) + M4(cofvby, ic + 1, 4) * RW(u, s, k, jee - 3);
    for (int j = 2; j <= jee; j++) RW(result, s, k, j) += C3(icon, 1, j) * RW(result, s, k, j - 1);
    RW(result, s, k, jee) *= C3(icon, 2, jee);
    for (int j = jee - 1; j >= 1; j--)
      RW(result, s, k, j) = (RW(result, s, k, j) - C3(icon, 3, j) * RW(result, s, k, j + 1)) * C3(icon, 2, j);
  }
}

/* implzy (lip.f90:528-597) on one plane: rows 1 and jee of u are kept. */
static void implzy(const oc_box *b, double *u, const double *res, const double *icon,
                   const double *econ, int jee, double var1) {
  int s = b->nz1;
  double *a1 = zalloc(jee + 2), *a2 = zalloc(jee + 2), *a3 = zalloc(jee + 2);
  for (int j = 2; j <= jee - 1; j++) {
    for (int k = 0; k < s; k++)
      RW(u, s, k, j) = C3(icon, 1, j) * RW(res, s, k, j - 1) + C3(icon, 2, j) * RW(res, s, k, j) +
                       C3(icon, 3, j) * RW(res, s, k, j + 1);
    a1[j] = C3(icon, 1, j) + var1 * C3(econ, 3, j);
    a2[j] = C3(icon, 2, j) + var1 * C3(econ, 2, j);
    a3[j] = C3(icon, 3, j) + var1 * C3(econ, 1, j);
  }
  for (int k = 0; k < s; k++) {
    RW(u, s, k, 2) -= a1[2] * RW(u, s, k, 1);
    RW(u, s, k, jee - 1) -= a3[jee - 1] * RW(u, s, k, jee);
  }
  a1[2] = 0.0;
  a3[jee - 1] = 0.0;
  for (int j = 3; j <= jee - 1; j++) {
    double d = 1.0 / a2[j - 1];
    a2[j - 1] = d;
    a1[j] = -a1[j] * d;
    a2[j] = a2[j] + a1[j] * a3[j - 1];
  }
  a2[jee - 1] = 1.0 / a2[jee - 1];
  for (int k = 0; k < s; k++) {
    for (int j = 3; j <= jee - 1; j++) RW(u, s, k, j) += a1[j] * RW(u, s, k, j - 1);
    RW(u, s, k, jee - 1) *= a2[jee - 1];
    for (int j = jee - 2; j >= 2; j--) RW(u, s, k, j) = (RW(u, s, k, j) - a3[j] * RW(u, s, k, j + 1)) * a2[j];
  }
  free(a1);
  free(a2);
  free(a3);
}

/* ------------------------------------------------------------------------
 * Stats-only compact schemes (coeft.f90:56-285), (8, n) arrays: rows 1-5
 * the explicit right-hand side, rows 6-8 the tridiagonal (alpha, diagonal,
 * beta), LU-factored in place as coeft.f90 does (row 6 the multiplier,
 * row 7 the inverse pivot).
 * ---------------------------------------------------------------------- */
#define F8(a, r, i) (a)[8 * (size_t)(i) + (r)-1]
static double sq(double x) { return x * x; }
static double cu(double x) { return x * x * x; }

/* in-place LU of rows 6-8 from row j0 (coeft.f90:135-139, 191-195, 325-336) */
static void f8_factor(double *c, int n, int j0) {
  for (int j = j0; j <= n; j++) {
    F8(c, 6, j) = F8(c, 6, j) / F8(c, 7, j - 1);
    F8(c, 7, j) = F8(c, 7, j) - F8(c, 6, j) * F8(c, 8, j - 1);
  }
  for (int j = j0 - 1; j <= n; j++) F8(c, 7, j) = 1.0 / F8(c, 7, j);
}

static void stats_coef(oc_box *b) {
  const int nx = b->nx, ny = b->ny;
  const double *y = b->y;
  double a, bb, c, d, e, f, g;
  b->fd_dvdy = zalloc(8 * (size_t)(ny + 1));
  b->fd_iy = zalloc(8 * (size_t)(ny + 1));
  b->fd_iyp = zalloc(8 * (size_t)(ny + 1));
  b->fd_dx = zalloc(8 * (size_t)(nx + 1));
  b->fd_ix = zalloc(8 * (size_t)(nx + 1));
  double *h = zalloc(ny + 2), *hi = zalloc(ny + 2), *yp = zalloc(ny + 2);

  /* d/dy at the v rows (coeft.f90:56-139) */
  double *D = b->fd_dvdy;
  for (int j = 1; j <= ny + 1; j++) h[j] = y[j] - y[j - 1];
  a = h[2]; bb = h[2] + h[3]; c = h[2] + h[3] + h[4];
  F8(D, 1, 1) = -(a * bb + a * c + 2 * bb * c) / bb / a / c;
  F8(D, 3, 1) = -sq(a) * c / (-c + bb) / bb / sq(-bb + a);
  F8(D, 4, 1) = sq(a) * bb / c / sq(a - c) / (-c + bb);
  F8(D, 2, 1) = -(F8(D, 1, 1) + F8(D, 3, 1) + F8(D, 4, 1));
  F8(D, 7, 1) = 1.0;
  F8(D, 8, 1) = bb * c / (a - c) / (-bb + a);
  F8(D, 1, 2) = sq(-bb + a) * (a - c) * (a * bb + 2 * a * c + 2 * bb * c) / a / cu(bb) / sq(c);
  F8(D, 2, 2) = 1 / a * (-3 * a * bb + 2 * bb * c + 5 * sq(a) - 4 * a * c) / (a - c) / (-bb + a);
  F8(D, 3, 2) = sq(a) * (a - c) * (-2 * a * c + 3 * a * bb + 4 * bb * c - 5 * sq(bb)) / cu(bb) / sq(-c + bb) / (-bb + a);
  F8(D, 4, 2) = -(F8(D, 1, 2) + F8(D, 2, 2) + F8(D, 3, 2));
  F8(D, 6, 2) = -sq(-bb + a) * (a - c) / c / sq(bb);
  F8(D, 7, 2) = 1.0;
  F8(D, 8, 2) = sq(a) * (a - c) / (-c + bb) / sq(bb);
  for (int i = 3; i <= ny - 2; i++) {
    a = h[i - 1]; bb = h[i]; c = h[i + 1]; d = h[i + 2]; f = c + d; e = a + bb;
    F8(D, 1, i) = -sq(c) * sq(bb) * f / (f + e) / e / (sq(c) + 2 * c * e + sq(e)) / (sq(e) - 2 * e * bb + sq(bb));
    F8(D, 2, i) = f / bb * e * sq(c) *
                  (3 * c * bb * f - 2 * c * f * e + 4 * sq(bb) * c - 3 * bb * e * c + 5 * sq(bb) * f -
                   4 * bb * e * f + 6 * cu(bb) - 5 * e * sq(bb)) /
                  (sq(f) + 2 * bb * f + sq(bb)) / (cu(bb) + 3 * bb * sq(c) + cu(c) + 3 * sq(bb) * c) /
                  (sq(e) - 2 * e * bb + sq(bb));
    F8(D, 3, i) = (c * bb * f - bb * e * c - 2 * bb * e * f + 2 * c * f * e) / f / bb / e / c;
    F8(D, 4, i) = -(-3 * c * bb * f - 2 * bb * e * f + 3 * bb * e * c + 4 * bb * sq(c) - 5 * sq(c) * f -
                    4 * c * f * e + 6 * cu(c) + 5 * sq(c) * e) *
                  sq(bb) / c * e * f / (sq(f) - 2 * f * c + sq(c)) /
                  (3 * sq(e) * bb * sq(c) + 6 * e * sq(bb) * sq(c) + 6 * e * bb * cu(c) + sq(e) * cu(bb) +
                   sq(c) * cu(c) + 2 * sq(sq(c)) * e + cu(c) * sq(e) + 3 * sq(bb) * cu(c) + cu(bb) * sq(c) +
                   3 * bb * sq(sq(c)) + 2 * e * cu(bb) * c + 3 * sq(e) * sq(bb) * c);
    F8(D, 5, i) = -(F8(D, 1, i) + F8(D, 2, i) + F8(D, 3, i) + F8(D, 4, i));
    F8(D, 6, i) = -e * sq(c) * f / (f + bb) / (sq(bb) + 2 * bb * c + sq(c)) / (bb - e);
    F8(D, 7, i) = 1.0;
    F8(D, 8, i) = -e * sq(bb) * f / (-f + c) / (sq(bb) + 2 * bb * c + sq(c)) / (c + e);
  }
  a = h[ny - 2]; bb = h[ny - 1]; c = h[ny]; f = a + bb + c; e = bb + c;
  F8(D, 1, ny - 1) = sq(c) * sq(c - e) / sq(f) / sq(f - e) / (-f + c);
  F8(D, 3, ny - 1) = -(5 * sq(c) - 4 * f * c - 3 * c * e + 2 * f * e) / c / (-f + c) / (c - e);
  F8(D, 4, ny - 1) = -sq(c - e) * (-f + c) * (2 * f * c + c * e + 2 * f * e) / c / sq(f) / cu(e);
  F8(D, 2, ny - 1) = -(F8(D, 1, ny - 1) + F8(D, 3, ny - 1) + F8(D, 4, ny - 1));
  F8(D, 6, ny - 1) = -(-f + c) * sq(c) / sq(e) / (f - e);
  F8(D, 7, ny - 1) = 1.0;
  F8(D, 8, ny - 1) = -sq(c - e) * (-f + c) / f / sq(e);
  F8(D, 1, ny) = sq(c) * e / sq(-f + c) / (f - e) / f;
  F8(D, 2, ny) = -sq(c) * f / e / sq(c - e) / (f - e);
  F8(D, 4, ny) = (f * c + c * e + 2 * f * e) / c / f / e;
  F8(D, 3, ny) = -(F8(D, 1, ny) + F8(D, 2, ny) + F8(D, 4, ny));
  F8(D, 6, ny) = f * e / (-f + c) / (c - e);
  F8(D, 7, ny) = 1.0;
  f8_factor(D, ny, 2);

  /* d/dx on the uniform x grid (coeft.f90:146-195) */
  double *X = b->fd_dx;
  const double r1[5] = {-37.0 / 12.0, 2.0 / 3.0, 3.0, -2.0 / 3.0, 1.0 / 12.0};
  const double r2[5] = {-43.0 / 96.0, -5.0 / 6.0, 9.0 / 8.0, 1.0 / 6.0, -1.0 / 96.0};
  const double ri[5] = {-1.0 / 36.0, -7.0 / 9.0, 0.0, 7.0 / 9.0, 1.0 / 36.0};
  for (int r = 1; r <= 5; r++) {
    F8(X, r, 1) = r1[r - 1];
    F8(X, r, 2) = r2[r - 1];
    F8(X, r, nx - 1) = -r2[5 - r];
    F8(X, r, nx) = -r1[5 - r];
    for (int i = 3; i <= nx - 2; i++) F8(X, r, i) = ri[r - 1];
  }
  F8(X, 7, 1) = 1.0; F8(X, 8, 1) = 4.0;
  F8(X, 6, 2) = 1.0 / 8.0; F8(X, 7, 2) = 1.0; F8(X, 8, 2) = 3.0 / 4.0;
  for (int i = 3; i <= nx - 2; i++) {
    F8(X, 6, i) = 1.0 / 3.0; F8(X, 7, i) = 1.0; F8(X, 8, i) = 1.0 / 3.0;
  }
  F8(X, 6, nx - 1) = 3.0 / 4.0; F8(X, 7, nx - 1) = 1.0; F8(X, 8, nx - 1) = 1.0 / 8.0;
  F8(X, 6, nx) = 4.0; F8(X, 7, nx) = 1.0;
  const double point0 = b->ax * b->pi / (nx + 1);
  for (int i = 1; i <= nx; i++)
    for (int r = 1; r <= 5; r++) F8(X, r, i) /= point0;
  f8_factor(X, nx, 2);

  /* midpoint interpolation along x (coeft.f90:203-243) */
  double *I = b->fd_ix;
  const double q1[4] = {5.0 / 24.0, 15.0 / 8.0, 5.0 / 8.0, -1.0 / 24.0};
  const double qm[4] = {-1.0 / 84.0, 1.0 / 4.0, 5.0 / 4.0, 5.0 / 12.0};
  const double qn[4] = {1.0 / 8.0, -7.0 / 8.0, 35.0 / 8.0, 35.0 / 8.0};
  const double qi[4] = {1.0 / 20.0, 3.0 / 4.0, 3.0 / 4.0, 1.0 / 20.0};
  for (int r = 1; r <= 4; r++) {
    F8(I, r, 1) = q1[r - 1];
    F8(I, r, nx - 1) = qm[r - 1];
    F8(I, r, nx) = qn[r - 1];
    for (int i = 2; i <= nx - 2; i++) F8(I, r, i) = qi[r - 1];
  }
  F8(I, 7, 1) = 1.0; F8(I, 8, 1) = 5.0 / 3.0;
  for (int i = 2; i <= nx - 2; i++) {
    F8(I, 6, i) = 3.0 / 10.0; F8(I, 7, i) = 1.0; F8(I, 8, i) = 3.0 / 10.0;
  }
  F8(I, 6, nx - 1) = 5.0 / 6.0; F8(I, 7, nx - 1) = 1.0; F8(I, 8, nx - 1) = 1.0 / 14.0;
  F8(I, 6, nx) = 7.0; F8(I, 7, nx) = 1.0;
  f8_factor(I, nx, 2);

  /* midpoint interpolation along y from the u rows to the v rows
   * (coeft.f90:249-336): fd_iy for u, w (ny+1 rows in), fd_iyp for p */
  double *Y = b->fd_iy, *P = b->fd_iyp;
  for (int j = 1; j <= ny + 1; j++) {
    yp[j] = (y[j - 1] + y[j]) * 0.5;
    hi[j] = y[j] - y[j - 1];
  }
  for (int j = 2; j <= ny + 1; j++) h[j] = yp[j] - yp[j - 1];
  a = hi[2] / 2; bb = a + h[3]; c = bb + h[4]; d = c + h[5];
  F8(Y, 1, 1) = -c * bb * d / (-d + a) / (-c + a) / (-bb + a);
  F8(Y, 2, 1) = c * a * d / (-d + bb) / (-c + bb) / (-bb + a);
  F8(Y, 3, 1) = -bb * a * d / (-d + c) / (sq(c) - a * c - bb * c + bb * a);
  F8(Y, 4, 1) = 1.0 - (F8(Y, 1, 1) + F8(Y, 2, 1) + F8(Y, 3, 1));
  a = hi[3] / 2; bb = a + h[4]; c = bb + h[5]; d = c + h[6]; e = hi[2] / 2;
  F8(Y, 1, 2) = 2 * c * bb * sq(a) * d / (e + d) / (e + c) / (bb + e) / (2 * sq(a) + sq(e) + 3 * e * a);
  F8(Y, 2, 2) = -2 * c * bb * e * d / (a - d) / (a - c) / (a - bb) / (a + e);
  F8(Y, 3, 2) = 2 * c * e * sq(a) * d / (-d + bb) / (-c + bb) /
                (2 * sq(a) * e + 2 * sq(a) * bb - 3 * a * sq(bb) - 3 * a * e * bb + cu(bb) + e * sq(bb));
  F8(Y, 4, 2) = -2 * sq(a) * bb * e * d / (-d + c) /
                (3 * e * a * sq(c) - 2 * e * sq(a) * c - 3 * bb * e * a * c - 2 * sq(a) * sq(c) +
                 2 * sq(a) * e * bb + 3 * a * cu(c) + bb * cu(c) - e * cu(c) + 2 * bb * sq(a) * c -
                 3 * bb * a * sq(c) + bb * e * sq(c) - sq(sq(c)));
  F8(Y, 7, 2) = 1.0;
  F8(Y, 8, 2) = -c * bb * e * d / (2 * a + e) / (2 * a - d) / (2 * a - c) / (2 * a - bb);
  F8(Y, 5, 2) = (F8(Y, 7, 2) + F8(Y, 8, 2)) - (F8(Y, 1, 2) + F8(Y, 2, 2) + F8(Y, 3, 2) + F8(Y, 4, 2));
  for (int j = 3; j <= ny - 1; j++) {
    a = hi[j + 1] / 2; bb = a + h[j + 2]; e = hi[j] / 2; f = e + h[j];
    F8(Y, 1, j) = 4 * sq(e) * sq(a) * bb / (bb + f) /
                  (2 * sq(a) * sq(f) + sq(sq(f)) + 4 * sq(e) * sq(a) - 6 * sq(a) * e * f - 9 * e * a * sq(f) +
                   6 * sq(e) * a * f + 2 * sq(e) * sq(f) + 3 * a * cu(f) - 3 * e * cu(f));
    F8(Y, 2, j) = -4 * f * sq(a) * bb / (bb + e) / (a + e) / (e - f) / (2 * a + e);
    F8(Y, 3, j) = -4 * f * sq(e) * bb / (a - bb) / (a + f) / (2 * sq(e) + sq(a) + 3 * e * a);
    F8(Y, 6, j) = -f * bb * sq(a) / (a + e) / (-f + 2 * e) / (bb + 2 * e) / (a + 2 * e);
    F8(Y, 7, j) = 1.0;
    F8(Y, 8, j) = -f * sq(e) * bb / (2 * a - bb) / (2 * a + f) / (2 * sq(a) + sq(e) + 3 * e * a);
    F8(Y, 4, j) = (F8(Y, 6, j) + F8(Y, 7, j) + F8(Y, 8, j)) - (F8(Y, 1, j) + F8(Y, 2, j) + F8(Y, 3, j));
  }
  a = hi[ny + 1] / 2; e = hi[ny] / 2; f = e + h[ny]; g = f + h[ny - 1];
  F8(Y, 1, ny) = 2 * (4 * sq(e) + 2 * e * a - a * f - 2 * f * e) * sq(e) / (-g + e) / (f - g) / g / (g + a);
  F8(Y, 2, ny) = -2 * (2 * e * a + 4 * sq(e) - a * g - 2 * g * e) * sq(e) / (e - f) / f / (sq(f) + a * f - g * f - a * g);
  F8(Y, 3, ny) = 2 * (8 * cu(e) - 4 * g * sq(e) - 2 * f * a * e - 2 * g * e * a - 4 * f * sq(e) + g * a * f +
                      4 * a * sq(e) + 2 * f * g * e) /
                 (cu(e) + a * sq(e) - f * sq(e) - g * sq(e) + g * a * f - g * e * a + f * g * e - f * a * e);
  F8(Y, 4, ny) = 2 * (-2 * f * e - 2 * g * e + g * f + 4 * sq(e)) * sq(e) / (g + a) / (a + f) / (a + e) / a;
  F8(Y, 6, ny) = 1.0;
  F8(Y, 7, ny) = (8 * cu(e) - 4 * g * sq(e) - 2 * f * a * e - 2 * g * e * a - 4 * f * sq(e) + g * a * f +
                  4 * a * sq(e) + 2 * f * g * e) / g / a / f;
  memcpy(P, Y, 8 * (size_t)(ny + 1) * sizeof(double));
  {
    int j = ny - 1;
    a = hi[j + 1] / 2; e = hi[j] / 2; f = e + h[j];
    F8(P, 1, j) = 4 * sq(e) * sq(a) / (2 * a + f) / (a + f) / (2 * sq(e) + sq(f) - 3 * e * f);
    F8(P, 2, j) = -4 * f * sq(a) / (3 * e * a + sq(e) + 2 * sq(a)) / (e - f);
    F8(P, 3, j) = 4 * sq(e) * f / (cu(a) + 2 * sq(e) * a + 3 * e * sq(a) + 2 * sq(e) * f + f * sq(a) + 3 * f * e * a);
    F8(P, 6, j) = -f * sq(a) / (a + e) / (2 * e - f) / (2 * e + a);
    F8(P, 7, j) = 1.0;
    F8(P, 8, j) = sq(e) * f / (4 * cu(a) + 2 * sq(e) * a + 6 * e * sq(a) + 3 * f * e * a + sq(e) * f + 2 * f * sq(a));
    j = ny;
    e = hi[j] / 2; f = e + h[j]; g = f + h[j - 1];
    F8(P, 1, j) = 2 * sq(e) * (2 * e - f) / (e - g) / (f - g) / g;
    F8(P, 2, j) = -2 * (-g + 2 * e) * sq(e) / (e - f) / f / (f - g);
    F8(P, 3, j) = 2 * (4 * sq(e) + g * f - 2 * e * f - 2 * g * e) / (sq(e) - e * f + g * f - g * e);
    F8(P, 6, j) = 1.0;
    F8(P, 7, j) = (4 * sq(e) + g * f - 2 * e * f - 2 * g * e) / g / f;
  }
  f8_factor(P, ny, 3);
  f8_factor(Y, ny, 3);
  free(h);
  free(hi);
  free(yp);
}

/* ------------------------------------------------------------------------
 * Box
 * ---------------------------------------------------------------------- */
#define PL(b, a, i, R) ((a) + (size_t)((i)-1) * (size_t)(b)->nz1 * (R))

oc_box *oc_create(int nx, int ny, int nz, int xout, double re, double ax,
                  double az, double cfl, const double *y, int fp32_transposes,
                  int freeze_dt) {
  oc_box *b = (oc_box *)calloc(1, sizeof(oc_box));
  b->nx = nx; b->ny = ny; b->nz = nz;
  b->nz1 = 2 * (nz / 3); b->nz2 = b->nz1 / 2 - 1; b->ngz = nz / 2;
  b->xout = xout; b->re = re; b->ax = ax; b->az = az; b->cfl = cfl;
  b->pi = PI_D;
  b->fp32_transposes = fp32_transposes; b->freeze_dt = freeze_dt;
  int nz2 = b->nz2, nz1 = b->nz1;
  b->y = zalloc(ny + 2);
  memcpy(b->y, y, sizeof(double) * (ny + 2));
  b->dy = zalloc(ny + 1); b->dyi = zalloc(ny + 1); b->idy = zalloc(ny + 1);
  b->idyy = zalloc(ny + 1); b->dymin = zalloc(ny + 1); b->hy = zalloc(ny + 1);
  b->x = zalloc(nx + 2); b->dx = zalloc(nx + 1); b->dxi = zalloc(nx + 1);
  b->kaz = zalloc(nz2 + 1); b->kaz2 = zalloc(nz2 + 1); b->kvis = zalloc(nz2 + 1);
  b->kmod = zalloc(nx);
  /* wavenumbers (bl_1.F90:357-366): cmplx(0d0,k/az) is default (single)
   * precision complex, so the wavenumber is rounded to FP32. */
  for (int k = 0; k <= nz2; k++) {
    b->kaz[k] = f32(k / az);
    b->kaz2[k] = b->kaz[k] * b->kaz[k];
  }
  size_t n3x = 3 * (size_t)(nx + 2), n3y = 3 * (size_t)(ny + 3);
  size_t n2x = 2 * (size_t)(nx + 2), n2y = 2 * (size_t)(ny + 3);
  b->inxu = zalloc(n3x); b->inxv = zalloc(n3x); b->inyu = zalloc(n3y); b->inyv = zalloc(n3y);
  b->cofiux = zalloc(n2x); b->cofivx = zalloc(n2x); b->cofiuy = zalloc(n2y); b->cofivy = zalloc(n2y);
  b->dcxu = zalloc(n3x); b->dcxv = zalloc(n3x); b->dcyu = zalloc(n3y); b->dcyv = zalloc(n3y);
  b->cofcxu = zalloc(n2x); b->cofcxv = zalloc(n2x); b->cofcyu = zalloc(n2y); b->cofcyv = zalloc(n2y);
  b->vixu = zalloc(n3x); b->vixv = zalloc(n3x); b->viyu = zalloc(n3y); b->viyv = zalloc(n3y);
  b->vyui = zalloc(n3y); b->vyvi = zalloc(n3y);
  b->cofvxu = zalloc(n3x); b->cofvxv = zalloc(n3x); b->cofvyu = zalloc(n3y); b->cofvyv = zalloc(n3y);
  b->ayp = zalloc(n3y); b->vmagic = zalloc(nx + 1);
  b->um = zalloc(ny + 2); b->u0 = zalloc(ny + 2); b->v0 = zalloc(ny + 1);
  coef(b);
  stats_coef(b);

  b->nu = (size_t)nz1 * (ny + 1) * nx;
  b->nv = (size_t)nz1 * ny * nx;
  for (int f = 0; f < OC_NFIELDS; f++)
    b->fld[f] = zalloc((f == OC_V || f == OC_P || f == OC_RHSVPA) ? b->nv : b->nu);
  b->pen_u = zalloc(b->nu); b->pen_v = zalloc(b->nv); b->pen_w = zalloc(b->nu);
  b->pen_a = zalloc(b->nu);
  b->resu = zalloc(b->nu); b->resv = zalloc(b->nv); b->resw = zalloc(b->nu);
  b->rhsu = zalloc(b->nu); b->rhsv = zalloc(b->nv); b->rhsw = zalloc(b->nu);
  b->wki1 = zalloc(b->nu); b->wki2 = zalloc(b->nv); b->wki3 = zalloc(b->nu);
  b->rest = zalloc(b->nv); b->tbuf = zalloc(b->nu);
  size_t prow = (size_t)(nz + 2) * (ny + 1);
  b->vm = zalloc(ny + 1);
  b->fft_tw = zalloc(2 * (size_t)nz);
  for (int t = 0; t < nz; t++) {
    b->fft_tw[2 * t] = cos(2.0 * PI_D * t / nz);
    b->fft_tw[2 * t + 1] = sin(2.0 * PI_D * t / nz);
  }
  b->nth = omp_get_max_threads();
  b->tls = (struct oc_tls *)calloc(b->nth, sizeof(struct oc_tls));
  for (int t = 0; t < b->nth; t++) {
    struct oc_tls *q = &b->tls[t];
    q->wkp = zalloc(prow); q->wkpo = zalloc(prow);
    q->wkf = zalloc((size_t)nz1 * (ny + 1)); q->wkfo = zalloc((size_t)nz1 * (ny + 1));
    q->wk2r = zalloc(nz + 2); q->vm = zalloc(ny + 1);
    q->fft_a = zalloc(2 * (size_t)nz); q->fft_b = zalloc(2 * (size_t)nz);
    q->fft_t = zalloc(2 * (size_t)nz);
  }
  b->setstep = 1;
  return b;
}

void oc_destroy(oc_box *b) {
  if (!b) return;
  double **ptrs[] = {&b->y, &b->dy, &b->dyi, &b->idy, &b->idyy, &b->dymin, &b->hy, &b->x,
                     &b->dx, &b->dxi, &b->kaz, &b->kaz2, &b->kvis, &b->kmod, &b->inxu,
                     &b->inxv, &b->inyu, &b->inyv, &b->cofiux, &b->cofivx, &b->cofiuy,
                     &b->cofivy, &b->dcxu, &b->dcxv, &b->dcyu, &b->dcyv, &b->cofcxu,
                     &b->cofcxv, &b->cofcyu, &b->cofcyv, &b->vixu, &b->vixv, &b->viyu,
                     &b->viyv, &b->vyui, &b->vyvi, &b->cofvxu, &b->cofvxv, &b->cofvyu,
                     &b->cofvyv, &b->ayp, &b->vmagic, &b->um, &b->u0, &b->v0, &b->pen_u,
                     &b->pen_v, &b->pen_w, &b->pen_a, &b->resu, &b->resv, &b->resw,
                     &b->rhsu, &b->rhsv, &b->rhsw, &b->wki1, &b->wki2, &b->wki3, &b->rest,
                     &b->tbuf, &b->vm,
                     &b->fft_tw};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(*ptrs[i]);
  for (int t = 0; t < b->nth; t++) {
    struct oc_tls *q = &b->tls[t];
    double *tp[] = {q->wkp, q->wkpo, q->wkf, q->wkfo, q->wk2r, q->vm, q->fft_a, q->fft_b, q->fft_t};
    for (size_t i = 0; i < sizeof tp / sizeof tp[0]; i++) free(tp[i]);
  }
  free(b->tls);
  for (int f = 0; f < OC_NFIELDS; f++) free(b->fld[f]);
  free(b->fd_dvdy); free(b->fd_dx); free(b->fd_ix); free(b->fd_iy); free(b->fd_iyp);
  free(b->st); free(b->ens);
  free(b);
}

size_t oc_field_len(const oc_box *b, int id) {
  return (id == OC_V || id == OC_P || id == OC_RHSVPA) ? b->nv : b->nu;
}
double *oc_field(oc_box *b, int id) { return b->fld[id]; }

void oc_set_inflow(oc_box *b, const double *u0, const double *v0, const double *um) {
  for (int j = 1; j <= b->ny + 1; j++) {
    b->u0[j] = u0[j - 1];
    b->um[j] = um[j - 1];
  }
  for (int j = 1; j <= b->ny; j++) b->v0[j] = v0[j - 1];
  inlet_retheta(b);
  magicnumber(b);
}
void oc_get_um(const oc_box *b, double *um) {
  for (int j = 1; j <= b->ny + 1; j++) um[j - 1] = b->um[j];
}
void oc_set_time(oc_box *b, double tiempo, double timeinit, double dt) {
  b->tiempo = tiempo;
  b->timeinit = timeinit;
  b->dt = dt;
  b->dt_valid = dt > 0.0;
}
void oc_get_time(const oc_box *b, double *tiempo, double *dt) {
  *tiempo = b->tiempo;
  *dt = b->dt;
}
void oc_get_genflu(const oc_box *b, double *o) {
  o[0] = b->rthin; o[1] = b->din; o[2] = b->rthout; o[3] = b->dout;
  o[4] = b->gam[0]; o[5] = b->gam[1];
}
void oc_get_ener(const oc_box *b, double *e) {
  for (int i = 0; i < 15; i++) e[i] = b->ener[i + 1];
}

const double *oc_coef(const oc_box *b, const char *nm, int *rows, int *cols) {
  struct { const char *n; const double *p; int r, c; } t[] = {
      {"inxu", b->inxu + 3, 3, b->nx}, {"inxv", b->inxv + 3, 3, b->nx},
      {"inyu", b->inyu + 3, 3, b->ny}, {"inyv", b->inyv + 3, 3, b->ny},
      {"inbx", b->inbx, 4, 4}, {"inby", b->inby, 4, 4},
      {"cofiux", b->cofiux + 2, 2, b->nx}, {"cofivx", b->cofivx + 2, 2, b->nx},
      {"cofiuy", b->cofiuy + 2, 2, b->ny}, {"cofivy", b->cofivy + 2, 2, b->ny},
      {"dcxu", b->dcxu + 3, 3, b->nx}, {"dcxv", b->dcxv + 3, 3, b->nx},
      {"dcyu", b->dcyu + 3, 3, b->ny + 1}, {"dcyv", b->dcyv + 3, 3, b->ny},
      {"dcbx", b->dcbx, 5, 4}, {"dcby", b->dcby, 4, 4}, {"ccon", b->ccon, 2, 4},
      {"cofcxu", b->cofcxu + 2, 2, b->nx}, {"cofcxv", b->cofcxv + 2, 2, b->nx},
      {"cofcyu", b->cofcyu + 2, 2, b->ny}, {"cofcyv", b->cofcyv + 2, 2, b->ny},
      {"vixu", b->vixu + 3, 3, b->nx}, {"vixv", b->vixv + 3, 3, b->nx},
      {"viyu", b->viyu + 3, 3, b->ny + 1}, {"viyv", b->viyv + 3, 3, b->ny},
      {"vyui", b->vyui + 3, 3, b->ny + 1}, {"vyvi", b->vyvi + 3, 3, b->ny},
      {"cofvbx", b->cofvbx, 4, 4}, {"cofvby", b->cofvby, 4, 4},
      {"cofvxu", b->cofvxu + 3, 3, b->nx}, {"cofvxv", b->cofvxv + 3, 3, b->nx},
      {"cofvyu", b->cofvyu + 3, 3, b->ny}, {"cofvyv", b->cofvyv + 3, 3, b->ny},
      {"ayp", b->ayp + 3, 3, b->ny - 1}, {"kmod", b->kmod + 1, 1, b->nx - 1},
      {"kaz", b->kaz, 1, b->nz2 + 1}, {"kaz2", b->kaz2, 1, b->nz2 + 1},
      {"kvis", b->kvis, 1, b->nz2 + 1}, {"vmagic", b->vmagic + 1, 1, b->nx},
      {"dy", b->dy, 1, b->ny + 1}, {"dyi", b->dyi + 1, 1, b->ny},
      {"idy", b->idy, 1, b->ny + 1}, {"idyy", b->idyy + 1, 1, b->ny},
      {"dymin", b->dymin + 1, 1, b->ny}, {"dx", b->dx, 1, b->nx + 1},
      {"x", b->x, 1, b->nx + 2}, {"hy", b->hy, 1, b->ny + 1},
  };
  for (size_t i = 0; i < sizeof t / sizeof t[0]; i++)
    if (!strcmp(t[i].n, nm)) {
      *rows = t[i].r;
      *cols = t[i].c;
      return t[i].p;
    }
  *rows = *cols = 0;
  return NULL;
}

double oc_scalar(const oc_box *b, const char *nm) {
  if (!strcmp(nm, "idx")) return b->idx;
  if (!strcmp(nm, "idxx")) return b->idxx;
  if (!strcmp(nm, "dxmin")) return b->dxmin;
  if (!strcmp(nm, "dzmin")) return b->dzmin;
  if (!strcmp(nm, "dtret")) return b->dtret;
  if (!strcmp(nm, "rex")) return b->rex;
  if (!strcmp(nm, "rthin")) return b->rthin;
  if (!strcmp(nm, "din")) return b->din;
  if (!strcmp(nm, "dt")) return b->dt;
  if (!strcmp(nm, "tiempo")) return b->tiempo;
  return NAN;
}

/* ------------------------------------------------------------------------
 * energies / ministats (rhsp.F90:493-509, stotf.F90:395-440)
 * ---------------------------------------------------------------------- */
static void ministats(oc_box *b, const double *u, const double *v, const double *w,
                      double *uner, int i) {
  int ny = b->ny, nz2 = b->nz2, s = b->nz1;
  const double *hy = b->hy;
  const double *arr[3] = {u, v, w};
  int rows[3] = {ny + 1, ny, ny + 1};
  if (i == 1) {
    for (int c = 0; c < 3; c++) {
      double tot = 0.0;
      for (int j = 1; j <= rows[c]; j++) {
        double aux = 0.0;
        for (int k = 0; k <= nz2; k++) {
          double re = RW(arr[c], s, 2 * k, j), im = RW(arr[c], s, 2 * k + 1, j);
          aux += re * re + im * im;
        }
        tot += aux * hy[j - 1];
      }
      uner[1 + c] = tot;
    }
  }
  if (i == b->nx / 2) {
    for (int c = 0; c < 3; c++) {
      double aux1 = 0, aux2 = 0;
      double r01 = RW(arr[c], s, 0, 1);
      double tot = 0.0;
      for (int j = 1; j <= rows[c]; j++) {
        double aux = 0.0;
        for (int k = 1; k <= nz2; k++) {
          double re = RW(arr[c], s, 2 * k, j), im = RW(arr[c], s, 2 * k + 1, j);
          aux += re * re + im * im;
        }
        aux *= 2.0;
        if (j == 1) aux1 = aux;
        if (j == 2) aux2 = aux;
        tot += (aux + r01 * r01) * hy[j - 1];
      }
      uner[4 + 4 * c] = RW(arr[c], s, 0, 1);
      uner[5 + 4 * c] = RW(arr[c], s, 0, 2);
      uner[6 + 4 * c] = aux1;
      uner[7 + 4 * c] = aux2;
      uner[1 + c] = tot;
    }
  }
}

/* ------------------------------------------------------------------------
 * genflu (genflu.F90:3-167, NEWGENFLU), dobleu (:173-199), profile (:204-242)
 * Real view of complex fields: index 0 = Re mode 0, 1 = Im mode 0.
 * ---------------------------------------------------------------------- */
static void dobleu(const double *eta, double *www, double *www1, int n) {
  const double ag = 4.0, bg = 0.2, w1default = (double)0.7895f;
  double bg2 = 1.0 - 2.0 * bg, tag = 1.0 / tanh(ag), yyy;
  *www1 = w1default;
  www[1] = 0.0;
  for (int i = 2; i <= n; i++) {
    double xxx = 0.5 * (eta[i] + eta[i - 1]);
    if (xxx >= 1.0)
      yyy = 1.0;
    else
      yyy = 0.5 * (1.0 + tag * tanh(ag * (xxx - bg) / (bg2 * xxx + bg)));
    www[i] = www[i - 1] + yyy * (eta[i] - eta[i - 1]);
    if (eta[i] >= 1.0 && eta[i - 1] < 1.0) *www1 = www[i - 1] + (1.0 - eta[i - 1]) * yyy;
  }
}

static void profile(const oc_box *b, double *ut, double *wk, int ntop, const int *kl,
                    const double *fsca, int n) {
  int s = b->nz1;
  double ggg = 1.0 / b->gam[0];
  /* the copy of row ntop-3 below (genflu.F90:232-234) needs ntop >= 4; the
   * reference indexes out of bounds otherwise.  Such a profile is left
   * unchanged. */
  if (ntop < 4) return;
  if (n == b->ny + 1) {
    for (int i = 1; i <= ntop; i++) {
      int kll = kl[i];
      for (int k = 2; k <= s - 1; k++)
        RW(wk, s, k, i + 1) = RW(ut, s, k, kll + 2) * fsca[i] + RW(ut, s, k, kll + 1) * (1 - fsca[i]);
    }
    for (int k = 2; k <= s - 1; k++) RW(wk, s, k, 1) = -RW(wk, s, k, 2);
  } else {
    for (int i = 1; i <= ntop; i++) {
      int kll = kl[i];
      for (int k = 2; k <= s - 1; k++)
        RW(wk, s, k, i) = RW(ut, s, k, kll + 1) * fsca[i] + RW(ut, s, k, kll) * (1 - fsca[i]);
    }
  }
  for (int i = ntop - 2; i <= n; i++)
    for (int k = 2; k <= s - 1; k++) RW(wk, s, k, i) = RW(wk, s, k, ntop - 3);
  for (int i = 1; i <= n; i++)
    for (int k = 2; k <= s - 1; k++) RW(ut, s, k, i) = RW(wk, s, k, i) * ggg;
}

static void genflu(oc_box *b) {
  int ny = b->ny, s = b->nz1, ru = ny + 1;
  double *u = b->fld[OC_U], *v = b->fld[OC_V], *w = b->fld[OC_W];
  double *u1 = PL(b, u, 1, ru), *v1 = PL(b, v, 1, ny), *w1 = PL(b, w, 1, ru);
  double tloc = b->tiempo - b->timeinit;
  /* bring the reference plane xout to the inlet */
  memcpy(u1, PL(b, u, b->xout, ru), sizeof(double) * s * ru);
  memcpy(v1, PL(b, v, b->xout, ny), sizeof(double) * s * ny);
  memcpy(w1, PL(b, w, b->xout, ru), sizeof(double) * s * ru);
  int jy = 119 < ny ? 119 : ny; /* y(119) clamped to the grid */
  for (int j = 1; j <= ny + 1; j++) {
    double fblend = 0.5 * (1 - tanh(b->y[j] - b->y[jy]));
    for (int k = 2; k < s; k++) {
      RW(u1, s, k, j) *= fblend;
      RW(w1, s, k, j) *= fblend;
    }
  }
  if (tloc < 20)
    b->timec = 20;
  else if (tloc < 100)
    b->timec = 100;
  else
    b->timec = tloc;
  double dtc = b->dt / b->timec;
  for (int j = 1; j <= ny + 1; j++) b->um[j] = dtc * RW(u1, s, 0, j) + b->um[j] * (1.0 - dtc);
  b->utauout = sqrt(fabs((b->um[2] - b->um[1]) / (b->y[2] - b->y[1]) / b->re));
  double uinf = 0.0;
  for (int j = ny + 1 - 20; j <= ny + 1 - 5; j++) uinf += b->um[j];
  uinf /= 16.0;
  double u99 = (double)0.99f * uinf;
  int jtop = 0;
  for (int j = 1; j <= ny - 2; j++) {
    if (b->um[j] >= u99) break;
    jtop = j + 1;
  }
  b->dout = b->y[jtop];
  double sum = 0.0;
  for (int j = 2; j <= jtop + 3; j++) sum += (uinf - b->um[j]) * b->um[j] * b->dy[j - 1];
  b->rthout = sum * b->re / uinf;
  b->gam[1] = b->rthin / b->rthout;
  b->gam[0] = pow(b->gam[1], 0.125);

  double *ysca = zalloc(ny + 1), *wlwa = zalloc(ny + 1), www1;
  double *fsca[2] = {zalloc(ny + 1), zalloc(ny + 1)};
  int *kl[2] = {(int *)calloc(ny + 1, sizeof(int)), (int *)calloc(ny + 1, sizeof(int))};
  int ntop[2] = {0, 0};
  for (int kk = 0; kk < 2; kk++) {
    for (int j = 1; j <= ny; j++)
      ysca[j] = (kk == 0) ? 0.5 * (b->y[j] + b->y[j + 1]) / b->din : b->y[j] / b->din;
    dobleu(ysca, wlwa, &www1, ny);
    for (int j = 1; j <= ny; j++)
      wlwa[j] = ysca[j] / b->gam[0] + (1.0 / b->gam[1] - 1.0 / b->gam[0]) / www1 * wlwa[j];
    int kll = 1, khh;
    for (int i = 1; i <= ny; i++) {
      if (wlwa[i] <= ysca[ny]) {
        ntop[kk] = i;
        khh = ny;
        while (khh - kll > 1) {
          int km = (khh + kll) / 2;
          if (ysca[km] >= wlwa[i])
            khh = km;
          else
            kll = km;
        }
        kl[kk][i] = kll;
        fsca[kk][i] = (wlwa[i] - ysca[kll]) / (ysca[khh] - ysca[kll]);
      }
    }
  }
  /* the zero mode */
  for (int j = 1; j <= ny + 1; j++) {
    RW(u1, s, 0, j) = b->u0[j];
    RW(u1, s, 1, j) = 0.0;
    RW(w1, s, 0, j) = 0.0;
    RW(w1, s, 1, j) = 0.0;
  }
  for (int j = 1; j <= ny; j++) {
    RW(v1, s, 0, j) = b->v0[j];
    RW(v1, s, 1, j) = 0.0;
  }
  double *wk = b->tbuf;
  profile(b, u1, wk, ntop[0], kl[0], fsca[0], ny + 1);
  profile(b, w1, wk, ntop[0], kl[0], fsca[0], ny + 1);
  profile(b, v1, wk, ntop[1], kl[1], fsca[1], ny);
  free(ysca); free(wlwa); free(fsca[0]); free(fsca[1]); free(kl[0]); free(kl[1]);
}

/* ------------------------------------------------------------------------
 * boun (boun.f90:7-56)
 * ---------------------------------------------------------------------- */
static void boun(oc_box *b) {
  int ny = b->ny, s = b->nz1, ru = ny + 1;
  double *u = b->fld[OC_U], *v = b->fld[OC_V], *w = b->fld[OC_W];
  const double *inby = b->inby, *ccon = b->ccon;
  for (int i = 1; i <= b->nx; i++) {
    double *ui = PL(b, u, i, ru), *vi = PL(b, v, i, ny), *wi = PL(b, w, i, ru);
    for (int k = 0; k < s; k++) {
      RW(vi, s, k, 1) = 0.0;
      RW(vi, s, k, ny) = 0.0;
    }
    RW(vi, s, 0, ny) = b->vmagic[i];
    double *fs[2] = {ui, wi};
    for (int f = 0; f < 2; f++)
      for (int k = 0; k < s; k++) {
        double *a = fs[f];
        RW(a, s, k, 1) = -1.0 / M4(inby, 2, 1) *
                         (M4(inby, 2, 2) * RW(a, s, k, 2) + M4(inby, 2, 3) * RW(a, s, k, 3) +
                          M4(inby, 2, 4) * RW(a, s, k, 4));
        RW(a, s, k, ny + 1) = -1.0 / M2(ccon, 1, 1) *
                              (M2(ccon, 1, 2) * RW(a, s, k, ny) + M2(ccon, 1, 3) * RW(a, s, k, ny - 1) +
                               M2(ccon, 1, 4) * RW(a, s, k, ny - 2));
      }
  }
}

/* ------------------------------------------------------------------------
 * rhsp (rhsp.F90:22-489)
 * ---------------------------------------------------------------------- */
static double cmaxabs(const double *a, int n, double m) {
  for (int i = 0; i < n; i++)
    if (fabs(a[i]) > m) m = fabs(a[i]);
  return m;
}

void oc_rhsp(oc_box *b, int m) {
  const int nx = b->nx, ny = b->ny, nz = b->nz, s = b->nz1, ru = ny + 1, nz2 = b->nz2;
  const int ps = nz + 2; /* physical row stride (wkp) */
  double *u = b->fld[OC_U], *v = b->fld[OC_V], *w = b->fld[OC_W], *p = b->fld[OC_P];
  double *rhsupa = b->fld[OC_RHSUPA], *rhsvpa = b->fld[OC_RHSVPA], *rhswpa = b->fld[OC_RHSWPA];
  double *rhsu = b->rhsu, *rhsv = b->rhsv, *rhsw = b->rhsw;
  double *wki1 = b->wki1, *wki2 = b->wki2, *wki3 = b->wki3;
  int mpu = s * ru, mpv = s * ny;
  const double *kaz = b->kaz;
  int setstep = b->setstep && !(b->freeze_dt && b->dt_valid);

  /* velocities to pencils (rhsp.F90:67-69) */
  chp2x(b, b->pen_u, u, ru);
  chp2x(b, b->pen_v, v, ny);
  chp2x(b, b->pen_w, w, ru);
  /* interpolation in x (rhsp.F90:105-112) */
  interpxx(b, b->pen_u, b->pen_a, b->inxu, b->cofiux, b->inbx, 2, mpu, 1);
  chx2p(b, wki1, b->pen_a, ru);
  interpxx(b, b->pen_v, b->pen_a, b->inxv, b->cofivx, b->inbx, 1, mpv, 0);
  chx2p(b, wki2, b->pen_a, ny);
  interpxx(b, b->pen_w, b->pen_a, b->inxv, b->cofivx, b->inbx, 1, mpu, 0);
  chx2p(b, wki3, b->pen_a, ru);

  double um_ = -1e14, wm = -1e14, poco = 1e-7;
  for (int j = 1; j <= ny; j++) b->vm[j] = -1e14;
  if (m == 1) {
    double uner[16] = {0};
    for (int i = 1; i <= nx; i++)
      ministats(b, PL(b, u, i, ru), PL(b, v, i, ny), PL(b, w, i, ru), uner, i);
    for (int q = 1; q <= 15; q++) b->ener[q] = uner[q];
  }

  /* plane loop 1 (rhsp.F90:120-171), planes over threads */
  for (int t = 0; t < b->nth; t++)
    for (int j = 1; j <= ny; j++) b->tls[t].vm[j] = -1e14;
#pragma omp parallel num_threads(b->nth) reduction(max : um_, wm)
  {
  struct oc_tls *tl = TLS(b);
  double *wkp = tl->wkp, *wkpo = tl->wkpo, *wkf = tl->wkf, *wk2r = tl->wk2r, *vm = tl->vm;
#pragma omp for schedule(static)
  for (int i = 1; i <= nx; i++) {
    double *ui = PL(b, u, i, ru), *w1i = PL(b, wki1, i, ru), *w2i = PL(b, wki2, i, ny),
           *w3i = PL(b, wki3, i, ru), *rui = PL(b, rhsu, i, ru);
    for (int j = 1; j <= ny + 1; j++) oc_c2r(b, &RW(ui, s, 0, j), &RW(wkp, ps, 0, j));
    if (setstep)
      for (int j = 1; j <= ny + 1; j++) um_ = cmaxabs(&RW(wkp, ps, 0, j), nz, um_);
    for (int j = 1; j <= ny + 1; j++) {
      oc_c2r(b, &RW(w1i, s, 0, j), wk2r);
      for (int n = 0; n < nz; n++) wk2r[n] = wk2r[n] * wk2r[n];
      oc_r2c(b, wk2r, &RW(w1i, s, 0, j));
      oc_c2r(b, &RW(w3i, s, 0, j), wk2r);
      if (setstep) wm = cmaxabs(wk2r, nz, wm);
      for (int n = 0; n < nz; n++) wk2r[n] = wk2r[n] * RW(wkp, ps, n, j);
      oc_r2c(b, wk2r, &RW(w3i, s, 0, j));
    }
    interpyy(b, wkp, ps, wkpo, ps, b->inyu, b->cofiuy, b->inby, 0, nz);
    for (int j = 1; j <= ny; j++) {
      oc_c2r(b, &RW(w2i, s, 0, j), wk2r);
      if (setstep) vm[j] = cmaxabs(wk2r, nz, vm[j]);
      for (int n = 0; n < nz; n++) wk2r[n] = wk2r[n] * RW(wkpo, ps, n, j);
      oc_r2c(b, wk2r, &RW(w2i, s, 0, j));
    }
    differzy(b, w2i, s, wkf, s, s, b->dcyu, b->dcby, b->cofcyu, 3, -1, 0, ny + 1);
    for (int j = 1; j <= ny + 1; j++)
      for (int k = 0; k <= nz2; k++) {
        double wr = RW(w3i, s, 2 * k, j), wi = RW(w3i, s, 2 * k + 1, j);
        RW(rui, s, 2 * k, j) = RW(wkf, s, 2 * k, j) - wi * kaz[k];
        RW(rui, s, 2 * k + 1, j) = RW(wkf, s, 2 * k + 1, j) + wr * kaz[k];
      }
  }
  }
  for (int t = 0; t < b->nth; t++)
    for (int j = 1; j <= ny; j++)
      if (b->tls[t].vm[j] > b->vm[j]) b->vm[j] = b->tls[t].vm[j];

  /* x storage: convective + viscous x terms (rhsp.F90:181-193) */
  chp2x(b, b->pen_a, wki1, ru);
  difvisxx(b, b->pen_a, b->pen_u, b->pen_a, b->dcxu, b->vixu, b->dcbx, b->cofcxu, b->cofvxu,
           b->cofvbx, 1, mpu, 1, b->rex);
  chx2p(b, wki1, b->pen_a, ru);
  chp2x(b, b->pen_a, wki2, ny);
  difvisxx(b, b->pen_a, b->pen_v, b->pen_a, b->dcxv, b->vixv, b->dcbx, b->cofcxv, b->cofvxv,
           b->cofvbx, 2, mpv, 0, b->rex);
  chx2p(b, wki2, b->pen_a, ny);
  chp2x(b, b->pen_a, wki3, ru);
  difvisxx(b, b->pen_a, b->pen_w, b->pen_a, b->dcxv, b->vixv, b->dcbx, b->cofcxv, b->cofvxv,
           b->cofvbx, 2, mpu, 0, b->rex);
  chx2p(b, wki3, b->pen_a, ru);
  /* complete rhs (rhsp.F90:199-204) */
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < b->nu; q++) rhsu[q] += wki1[q];
  memcpy(rhsv, wki2, sizeof(double) * b->nv);
  memcpy(rhsw, wki3, sizeof(double) * b->nu);
  memcpy(PL(b, rhsu, nx, ru), PL(b, wki1, nx, ru), sizeof(double) * mpu);

  /* time step (rhsp.F90:214-262) */
  if (setstep) {
    double dt1 = b->dxmin / (poco > um_ ? poco : um_);
    double dt2 = b->dxmin / (poco > wm ? poco : wm);
    double dt3 = INFINITY;
    for (int j = 1; j <= ny; j++) {
      double t = b->dymin[j] / (poco > b->vm[j] ? poco : b->vm[j]);
      if (t < dt3) dt3 = t;
    }
    double dtl = dt1;
    if (dt2 < dtl) dtl = dt2;
    if (dt3 < dtl) dtl = dt3;
    if (b->dtret < dtl) dtl = b->dtret;
    b->dt = dtl;
    b->dt_valid = 1;
  }
  b->setstep = 0;
  double dt = b->dt;
  double var1 = dt * (b->rkcv[m] + b->rkdv[m]);
  const double var2_ = dt * b->rex * b->rkcv[m];
  double var3 = dt * b->rkc[m];
  double var4 = dt * b->rkd[m];

  /* save velocities (rhsp.F90:281-285) */
  memcpy(b->resu, u, sizeof(double) * b->nu);
  memcpy(b->resv, v, sizeof(double) * b->nv);
  memcpy(b->resw, w, sizeof(double) * b->nu);
  genflu(b);

  /* pressure-gradient predictor (rhsp.F90:291-339).  With one process the
   * last plane receives no p(nx+1) halo and is left unchanged. */
#pragma omp parallel for schedule(static)
  for (int i = 2; i <= nx - 1; i++) {
    double *ui = PL(b, u, i, ru), *pi = PL(b, p, i, ny), *pn = PL(b, p, i + 1, ny);
    for (int j = 2; j <= ny; j++)
      for (int k = 0; k < s; k++)
        RW(ui, s, k, j) -= var1 * b->idxx * (RW(pn, s, k, j) - RW(pi, s, k, j));
  }
#pragma omp parallel for schedule(static)
  for (int i = 2; i <= nx; i++) {
    double *vi = PL(b, v, i, ny), *wi = PL(b, w, i, ru), *pi = PL(b, p, i, ny);
    for (int j = 2; j <= ny - 1; j++)
      for (int k = 0; k < s; k++)
        RW(vi, s, k, j) -= var1 * b->idyy[j] * (RW(pi, s, k, j + 1) - RW(pi, s, k, j));
    for (int j = 2; j <= ny; j++)
      for (int k = 0; k <= nz2; k++) {
        double pr = RW(pi, s, 2 * k, j), pim = RW(pi, s, 2 * k + 1, j);
        RW(wi, s, 2 * k, j) -= var1 * (-kaz[k] * pim);
        RW(wi, s, 2 * k + 1, j) -= var1 * (kaz[k] * pr);
      }
  }

  /* plane loop 2 (rhsp.F90:351-418), planes over threads */
#pragma omp parallel num_threads(b->nth)
  {
  struct oc_tls *tl = TLS(b);
  double *wkp = tl->wkp, *wkpo = tl->wkpo, *wkf = tl->wkf, *wkfo = tl->wkfo, *wk2r = tl->wk2r;
#pragma omp for schedule(static)
  for (int i = 2; i <= nx - 1; i++) {
    double *rui = PL(b, rhsu, i, ru), *rvi = PL(b, rhsv, i, ny), *rwi = PL(b, rhsw, i, ru);
    double *sui = PL(b, b->resu, i, ru), *svi = PL(b, b->resv, i, ny), *swi = PL(b, b->resw, i, ru);
    for (int j = 2; j <= ny - 1; j++)
      for (int k = 0; k < s; k++) {
        double kv = b->kvis[k / 2];
        RW(rui, s, k, j) -= RW(sui, s, k, j) * kv;
        RW(rwi, s, k, j) -= RW(swi, s, k, j) * kv;
        RW(rvi, s, k, j) -= RW(svi, s, k, j) * kv;
      }
    for (int k = 0; k < s; k++) {
      double kv = b->kvis[k / 2];
      RW(rui, s, k, ny) -= RW(sui, s, k, ny) * kv;
      RW(rwi, s, k, ny) -= RW(swi, s, k, ny) * kv;
    }
    /* vw (w to v positions) */
    interpyy(b, swi, s, wkf, s, b->inyu, b->cofiuy, b->inby, 0, s);
    for (int j = 1; j <= ny + 1; j++) oc_c2r(b, &RW(wkf, s, 0, j), &RW(wkpo, ps, 0, j));
    for (int j = 1; j <= ny; j++) {
      oc_c2r(b, &RW(svi, s, 0, j), wk2r);
      for (int n = 0; n < nz; n++) RW(wkp, ps, n, j) = wk2r[n] * RW(wkpo, ps, n, j);
    }
    for (int j = 1; j <= ny + 1; j++) oc_r2c(b, &RW(wkp, ps, 0, j), &RW(wkfo, s, 0, j));
    for (int j = 1; j <= ny; j++)
      for (int k = 0; k <= nz2; k++) {
        double fr = RW(wkfo, s, 2 * k, j), fi = RW(wkfo, s, 2 * k + 1, j);
        RW(rvi, s, 2 * k, j) += -fi * kaz[k];
        RW(rvi, s, 2 * k + 1, j) += fr * kaz[k];
      }
    differzy(b, wkfo, s, wkf, s, s, b->dcyu, b->dcby, b->cofcyu, 3, -1, 0, ny + 1);
    for (int j = 1; j <= ny + 1; j++)
      for (int k = 0; k < s; k++) RW(rwi, s, k, j) += RW(wkf, s, k, j);
    /* vv (v to u positions) */
    interpyy(b, svi, s, wkf, s, b->inyv, b->cofivy, b->inby, 1, s);
    for (int j = 1; j <= ny; j++) oc_c2r(b, &RW(wkf, s, 0, j), &RW(wkpo, ps, 0, j));
    for (int j = 1; j <= ny; j++)
      for (int n = 0; n < nz; n++) RW(wkp, ps, n, j) = RW(wkpo, ps, n, j) * RW(wkpo, ps, n, j);
    for (int j = 1; j <= ny; j++) oc_r2c(b, &RW(wkp, ps, 0, j), &RW(wkfo, s, 0, j));
    differzy(b, wkfo, s, wkf, s, s, b->dcyv, b->dcby, b->cofcyv, 1, 0, 1, ny);
    for (int j = 1; j <= ny; j++) {
      for (int k = 0; k < s; k++) RW(rvi, s, k, j) += RW(wkf, s, k, j);
      oc_c2r(b, &RW(swi, s, 0, j), wk2r);
      for (int n = 0; n < nz; n++) wk2r[n] = wk2r[n] * wk2r[n];
      oc_r2c(b, wk2r, &RW(wkf, s, 0, j));
      for (int k = 0; k <= nz2; k++) {
        double fr = RW(wkf, s, 2 * k, j), fi = RW(wkf, s, 2 * k + 1, j);
        RW(rwi, s, 2 * k, j) += -fi * kaz[k];
        RW(rwi, s, 2 * k + 1, j) += fr * kaz[k];
      }
    }
  }
  }

  boun(b);
  /* RHS boundary rows (rhsp.F90:423-445) */
  memset(PL(b, rhsu, 1, ru), 0, sizeof(double) * mpu);
  memset(PL(b, rhsv, 1, ny), 0, sizeof(double) * mpv);
  memset(PL(b, rhsw, 1, ru), 0, sizeof(double) * mpu);
  for (int i = 1; i <= nx; i++) {
    double *rvi = PL(b, rhsv, i, ny);
    for (int k = 0; k < s; k++) {
      RW(rvi, s, k, 1) = 0.0;
      RW(rvi, s, k, ny) = 0.0;
    }
  }
  for (int i = 1; i <= nx; i++) {
    double *fs[2] = {PL(b, rhsu, i, ru), PL(b, rhsw, i, ru)};
    for (int f = 0; f < 2; f++) {
      double *a = fs[f];
      for (int k = 0; k < s; k++) {
        RW(a, s, k, 1) = -1.0 / M4(b->inby, 2, 1) *
                         (M4(b->inby, 2, 2) * RW(a, s, k, 2) + M4(b->inby, 2, 3) * RW(a, s, k, 3) +
                          M4(b->inby, 2, 4) * RW(a, s, k, 4));
        RW(a, s, k, ny + 1) = 0.0;
      }
    }
  }
  /* explicit RK3 combination (rhsp.F90:448-475) */
#pragma omp parallel num_threads(b->nth)
  {
  double *wkf = TLS(b)->wkf;
#pragma omp for schedule(static)
  for (int i = 2; i <= nx; i++) {
    const double var2 = i == nx ? 0.0 : var2_;
    vistyy(b, PL(b, b->resv, i, ny), wkf, b->viyv, b->cofvyv, b->cofvby, ny);
    for (size_t q = 0; q < (size_t)mpv; q++)
      PL(b, wki2, i, ny)[q] = PL(b, v, i, ny)[q] -
                              (var3 * PL(b, rhsv, i, ny)[q] + var2 * wkf[q] + var4 * PL(b, rhsvpa, i, ny)[q]);
    vistyy(b, PL(b, b->resu, i, ru), wkf, b->viyu, b->cofvyu, b->cofvby, ny + 1);
    for (size_t q = 0; q < (size_t)mpu; q++)
      PL(b, wki1, i, ru)[q] = PL(b, u, i, ru)[q] -
                              (var3 * PL(b, rhsu, i, ru)[q] + var2 * wkf[q] + var4 * PL(b, rhsupa, i, ru)[q]);
    vistyy(b, PL(b, b->resw, i, ru), wkf, b->viyu, b->cofvyu, b->cofvby, ny + 1);
    for (size_t q = 0; q < (size_t)mpu; q++)
      PL(b, wki3, i, ru)[q] = PL(b, w, i, ru)[q] -
                              (var3 * PL(b, rhsw, i, ru)[q] + var2 * wkf[q] + var4 * PL(b, rhswpa, i, ru)[q]);
  }
  }
  memcpy(rhsupa, rhsu, sizeof(double) * b->nu);
  memcpy(rhsvpa, rhsv, sizeof(double) * b->nv);
  memcpy(rhswpa, rhsw, sizeof(double) * b->nu);
  /* implicit y viscous step (rhsp.F90:479-484) */
  double rkk = b->rex * dt * b->rkdv[m];
#pragma omp parallel for schedule(static)
  for (int i = 2; i <= nx; i++) {
    implzy(b, PL(b, u, i, ru), PL(b, wki1, i, ru), b->vyui, b->cofvyu, ny + 1, rkk);
    implzy(b, PL(b, v, i, ny), PL(b, wki2, i, ny), b->vyvi, b->cofvyv, ny, rkk);
    implzy(b, PL(b, w, i, ru), PL(b, wki3, i, ru), b->vyui, b->cofvyu, ny + 1, rkk);
  }
  b->ener[13] = b->ener[14] = b->ener[15] = 0.0;
}

/* outflow_correction (boun.f90:63-93) */
void oc_outflow_correction(oc_box *b) {
  int nx = b->nx, ny = b->ny, s = b->nz1, ru = ny + 1;
  double *u = b->fld[OC_U], *v = b->fld[OC_V], *ra = b->fld[OC_RHSUPA];
  double sumax = 0.0;
  for (int i = 2; i <= nx; i++) sumax += RW(PL(b, v, i, ny), s, 0, ny);
  sumax *= b->dx[1];
  double s1 = 0.0, s2 = 0.0;
  for (int j = 2; j <= ny; j++) s1 += RW(PL(b, u, 1, ru), s, 0, j) * b->dy[j - 1];
  sumax = sumax - s1;
  for (int j = 2; j <= ny; j++) s2 += RW(PL(b, u, nx, ru), s, 0, j) * b->dy[j - 1];
  sumax = sumax + s2;
  double alpha = -sumax / (b->y[ny] - b->y[1]);
  for (int j = 2; j <= ny; j++) {
    RW(PL(b, u, nx, ru), s, 0, j) += alpha;
    RW(PL(b, ra, nx, ru), s, 0, j) += alpha * b->idx;
  }
}

/* soltriy (pois.f90:211-239) for one real column of a plane */
static void soltriy(double *col, int s, int ny1, double (*a)[3]) {
  /* a[j][0..2] for j = 1..ny1 (1-based rows) */
  for (int j = 2; j <= ny1; j++) {
    double d = 1.0 / a[j - 1][1];
    a[j - 1][1] = d;
    a[j][0] = -a[j][0] * d;
    a[j][1] = a[j][1] + a[j][0] * a[j - 1][2];
  }
  a[ny1][1] = 1.0 / a[ny1][1];
  for (int j = 2; j <= ny1; j++) col[(size_t)(j - 1) * s] += a[j][0] * col[(size_t)(j - 2) * s];
  col[(size_t)(ny1 - 1) * s] *= a[ny1][1];
  for (int j = ny1 - 1; j >= 1; j--)
    col[(size_t)(j - 1) * s] = (col[(size_t)(j - 1) * s] - a[j][2] * col[(size_t)j * s]) * a[j][1];
}

/* pois (pois.f90:11-189) */
void oc_pois(oc_box *b, double varstep) {
  int nx = b->nx, ny = b->ny, s = b->nz1, ru = ny + 1, nz2 = b->nz2;
  double *u = b->fld[OC_U], *v = b->fld[OC_V], *w = b->fld[OC_W], *p = b->fld[OC_P];
  double *rest = b->rest;
  const double *kaz = b->kaz;
  /* divergence (pois.f90:41-54) */
#pragma omp parallel for schedule(static)
  for (int i = 2; i <= nx; i++) {
    double *ri = PL(b, rest, i, ny), *wi = PL(b, w, i, ru), *vi = PL(b, v, i, ny);
    for (int j = 1; j <= ny - 1; j++)
      for (int k = 0; k <= nz2; k++) {
        double wr = RW(wi, s, 2 * k, j + 1), wim = RW(wi, s, 2 * k + 1, j + 1);
        RW(ri, s, 2 * k, j) = -wim * kaz[k] + (RW(vi, s, 2 * k, j + 1) - RW(vi, s, 2 * k, j)) * b->idy[j];
        RW(ri, s, 2 * k + 1, j) = wr * kaz[k] + (RW(vi, s, 2 * k + 1, j + 1) - RW(vi, s, 2 * k + 1, j)) * b->idy[j];
      }
  }
#pragma omp parallel for schedule(static)
  for (int i = 2; i <= nx; i++) {
    double *ri = PL(b, rest, i, ny), *ui = PL(b, u, i, ru), *um1 = PL(b, u, i - 1, ru);
    for (int j = 1; j <= ny - 1; j++)
      for (int k = 0; k < s; k++) RW(ri, s, k, j) += b->idx * (RW(ui, s, k, j + 1) - RW(um1, s, k, j + 1));
  }
  /* DCT-II along x over planes 2..nx (pois.f90:82-100) */
  double *pen = b->pen_a;
  chp2x(b, pen, rest, ny);
  cosft_lines(nx - 1, s * ny, pen + 1, (size_t)nx, -1);
  chx2p(b, rest, pen, ny);
  for (int i = 1; i <= nx; i++)
    for (int j = 1; j <= ny; j++) RW(PL(b, rest, i, ny), s, 1, j) = 0.0;
  /* per-mode tridiagonal in y (pois.f90:106-126) */
#pragma omp parallel
  {
  double(*a)[3] = (double(*)[3])malloc(sizeof(double) * 3 * (ny + 1));
#pragma omp for schedule(static)
  for (int i = 2; i <= nx; i++) {
    double *ri = PL(b, rest, i, ny);
    for (int k = 0; k <= nz2; k++) {
      for (int part = 0; part < 2; part++) {
        for (int j = 1; j <= ny - 1; j++) {
          a[j][0] = C3(b->ayp, 1, j);
          a[j][1] = C3(b->ayp, 2, j) - b->kaz2[k] - b->kmod[i - 1];
          a[j][2] = C3(b->ayp, 3, j);
        }
        if (i == 2 && k == 0) {
          a[ny - 1][1] = 2.0 * C3(b->ayp, 2, ny - 1);
          RW(ri, s, 2 * k + part, ny - 1) = 0.0;
        }
        soltriy(&RW(ri, s, 2 * k + part, 1), s, ny - 1, a);
      }
    }
  }
  free(a);
  }
  /* inverse DCT (pois.f90:130-142) */
  chp2x(b, pen, rest, ny);
  cosft_lines(nx - 1, s * ny, pen + 1, (size_t)nx, 1);
  chx2p(b, rest, pen, ny);
  /* projection (pois.f90:144-162); with one process the last plane of u
   * gets no correction (no phi(nx+1) halo) */
#pragma omp parallel for schedule(static)
  for (int i = 2; i <= nx; i++) {
    double *ri = PL(b, rest, i, ny), *pi = PL(b, p, i, ny), *wi = PL(b, w, i, ru), *vi = PL(b, v, i, ny);
    for (int j = 2; j <= ny; j++)
      for (int k = 0; k <= nz2; k++) {
        double fr = RW(ri, s, 2 * k, j - 1), fi = RW(ri, s, 2 * k + 1, j - 1);
        RW(pi, s, 2 * k, j) = varstep * fr + RW(pi, s, 2 * k, j);
        RW(pi, s, 2 * k + 1, j) = varstep * fi + RW(pi, s, 2 * k + 1, j);
        RW(wi, s, 2 * k, j) -= -kaz[k] * fi;
        RW(wi, s, 2 * k + 1, j) -= kaz[k] * fr;
      }
    for (int j = 2; j <= ny - 1; j++)
      for (int k = 0; k < s; k++) RW(vi, s, k, j) -= b->idyy[j] * (RW(ri, s, k, j) - RW(ri, s, k, j - 1));
  }
#pragma omp parallel for schedule(static)
  for (int i = 2; i <= nx - 1; i++) {
    double *ui = PL(b, u, i, ru), *ri = PL(b, rest, i, ny), *rn = PL(b, rest, i + 1, ny);
    for (int j = 1; j <= ny - 1; j++)
      for (int k = 0; k < s; k++) RW(ui, s, k, j + 1) -= b->idxx * (RW(rn, s, k, j) - RW(ri, s, k, j));
  }
}

void oc_substep(oc_box *b, int m) {
  oc_rhsp(b, m);
  oc_outflow_correction(b);
  oc_pois(b, 0.5 / b->dt / b->rkdv[m]);
}

void oc_step(oc_box *b) {
  b->setstep = 1;
  for (int m = 1; m <= 3; m++) oc_substep(b, m);
  b->tiempo += b->dt;
}

/* ------------------------------------------------------------------------
 * Operator-level entry points on plane-layout fields
 * ---------------------------------------------------------------------- */
void oc_op_interpxx(oc_box *b, const double *gen, double *geni, int R, int which_u) {
  int nlin = b->nz1 * R;
  chp2x(b, b->pen_u, gen, R);
  if (which_u)
    interpxx(b, b->pen_u, b->pen_a, b->inxu, b->cofiux, b->inbx, 2, nlin, 1);
  else
    interpxx(b, b->pen_u, b->pen_a, b->inxv, b->cofivx, b->inbx, 1, nlin, 0);
  chx2p(b, geni, b->pen_a, R);
}

void oc_op_difvisxx(oc_box *b, const double *prod, const double *vel, double *out, int R,
                    int which_u) {
  int nlin = b->nz1 * R;
  chp2x(b, b->pen_a, prod, R);
  chp2x(b, b->pen_u, vel, R);
  if (which_u)
    difvisxx(b, b->pen_a, b->pen_u, b->pen_a, b->dcxu, b->vixu, b->dcbx, b->cofcxu, b->cofvxu,
             b->cofvbx, 1, nlin, 1, b->rex);
  else
    difvisxx(b, b->pen_a, b->pen_u, b->pen_a, b->dcxv, b->vixv, b->dcbx, b->cofcxv, b->cofvxv,
             b->cofvbx, 2, nlin, 0, b->rex);
  chx2p(b, out, b->pen_a, R);
}

/* loguv 0: u-type (ny+1 rows) in -> ny rows out; 1: v-type in -> ny rows out
 * (output stored with ny+1 rows per plane, row ny+1 untouched). */
void oc_op_interpyy(oc_box *b, const double *gen, double *geni, int loguv, int nplanes) {
  int s = b->nz1, ny = b->ny, rin = loguv == 0 ? ny + 1 : ny;
  for (int i = 0; i < nplanes; i++)
    interpyy(b, gen + (size_t)i * s * rin, s, geni + (size_t)i * s * (ny + 1), s,
             loguv == 0 ? b->inyu : b->inyv, loguv == 0 ? b->cofiuy : b->cofivy, b->inby, loguv, s);
}

/* which 0: d/dy from v rows to u rows (dcyu set, jee=ny+1);
 *       1: d/dy from u rows (1..ny used) to v rows (dcyv set, jee=ny).
 * Input and output are stored with ny+1 rows per plane. */
void oc_op_differzy(oc_box *b, const double *geni, double *res, int which, int nplanes) {
  int s = b->nz1, ny = b->ny;
  for (int i = 0; i < nplanes; i++) {
    const double *g = geni + (size_t)i * s * (ny + 1);
    double *r = res + (size_t)i * s * (ny + 1);
    if (which == 0)
      differzy(b, g, s, r, s, s, b->dcyu, b->dcby, b->cofcyu, 3, -1, 0, ny + 1);
    else
      differzy(b, g, s, r, s, s, b->dcyv, b->dcby, b->cofcyv, 1, 0, 1, ny);
  }
}

void oc_op_vistyy(oc_box *b, const double *u, double *res, int R, int nplanes) {
  int s = b->nz1, ny = b->ny;
  for (int i = 0; i < nplanes; i++) {
    if (R == ny + 1)
      vistyy(b, u + (size_t)i * s * R, res + (size_t)i * s * R, b->viyu, b->cofvyu, b->cofvby, ny + 1);
    else
      vistyy(b, u + (size_t)i * s * R, res + (size_t)i * s * R, b->viyv, b->cofvyv, b->cofvby, ny);
  }
}

void oc_op_implzy(oc_box *b, double *u, const double *res, int R, double rkk, int nplanes) {
  int s = b->nz1, ny = b->ny;
  for (int i = 0; i < nplanes; i++) {
    if (R == ny + 1)
      implzy(b, u + (size_t)i * s * R, res + (size_t)i * s * R, b->vyui, b->cofvyu, ny + 1, rkk);
    else
      implzy(b, u + (size_t)i * s * R, res + (size_t)i * s * R, b->vyvi, b->cofvyv, ny, rkk);
  }
}

/* Re_theta(x) with the genflu.F90:79-87 sums, and Cf(x) = 2 u_tau^2 with
 * u_tau^2 from genflu.F90:76, on the instantaneous kz=0 profile. */
void oc_retheta_cf(const oc_box *b, double *rth, double *cf) {
  int ny = b->ny, s = b->nz1, ru = ny + 1;
  const double *u = b->fld[OC_U];
  double *U = zalloc(ny + 2);
  for (int i = 1; i <= b->nx; i++) {
    const double *ui = u + (size_t)(i - 1) * s * ru;
    for (int j = 1; j <= ny + 1; j++) U[j] = RW(ui, s, 0, j);
    double uinf = 0.0;
    for (int j = ny + 1 - 20; j <= ny + 1 - 5; j++) uinf += U[j];
    uinf /= 16.0;
    double u99 = (double)0.99f * uinf;
    int jtop = 0;
    for (int j = 1; j <= ny - 2; j++) {
      if (U[j] >= u99) break;
      jtop = j + 1;
    }
    double sum = 0.0;
    for (int j = 2; j <= jtop + 3; j++) sum += (uinf - U[j]) * U[j] * b->dy[j - 1];
    rth[i - 1] = sum * b->re / uinf;
    cf[i - 1] = 2.0 * fabs((U[2] - U[1]) / (b->y[2] - b->y[1]) / b->re);
  }
  free(U);
}

int oc_max_threads(void) { return omp_get_max_threads(); }

/* ========================================================================
 * Statistics: statsp (stotf.F90:51-291) as the production build compiles it
 * (makefile-OMP-vul:11: -DNOCORR -DNOSPECTRA, dissipation and triple
 * products on), with the prologue of rhsp.F90:72-102 that forms its x
 * interpolation and x derivatives.
 * ====================================================================== */
static const char *const kStatNames[OC_NSTAT] = {
    "us", "vs", "ws", "uv", "ua", "va", "wa", "pp", "pm", "vortx", "vorty", "vortz", "vortxa", "vortya",
    "vortza", "uw", "vw", "dispu", "dispv", "dispw", "dispuv", "pup", "pvp", "pdudx", "pdvdy", "pdudy",
    "pdvdx", "pdwdz", "v3", "u2v", "v2u", "w2v", "dudx0", "dudy0", "dvdx0", "dvdy0", "dwdx0", "dwdy0",
    "u3", "w2u"};

const char *oc_stat_name(int id) { return (id >= 0 && id < OC_NSTAT) ? kStatNames[id] : NULL; }

/* the tridiagonal back end shared by the four operators on the line u
 * (row j at u[(j-1) s]): forward from row j0 (multipliers in row 6), last
 * row times its inverse pivot, backward to row j1 (lip.f90:690-697,
 * 725-731, 759-765, 801-807) */
static void f8_solve(const double *c, double *u, size_t s, int n, int j0, int j1) {
  for (int j = j0; j <= n; j++) u[(size_t)(j - 1) * s] -= F8(c, 6, j) * u[(size_t)(j - 2) * s];
  u[(size_t)(n - 1) * s] *= F8(c, 7, n);
  for (int j = n - 1; j >= j1; j--)
    u[(size_t)(j - 1) * s] = (u[(size_t)(j - 1) * s] - F8(c, 8, j) * u[(size_t)j * s]) * F8(c, 7, j);
}

/* diffx_inplace (lip.f90:704-732) on pencils (nx, nlin) */
static void diffx_pen(const oc_box *b, const double *in, double *out, int nlin) {
  const int nx = b->nx;
  const double *c = b->fd_dx;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nlin; k++) {
    const double *g = in + (size_t)k * nx - 1; /* g[1..nx] */
    double *o = out + (size_t)k * nx - 1;
    for (int r = 0; r < 2; r++) { /* rows 1, 2 and nx-1, nx */
      int i = 1 + r, ie = nx - 1 + r;
      o[i] = o[ie] = 0.0;
      for (int q = 1; q <= 5; q++) {
        o[i] += F8(c, q, i) * g[q];
        o[ie] += F8(c, q, ie) * g[nx - 5 + q];
      }
    }
    for (int i = 3; i <= nx - 2; i++) {
      double t = 0.0;
      for (int q = 1; q <= 5; q++) t += F8(c, q, 3) * g[i - 3 + q];
      o[i] = t;
    }
    f8_solve(c, o + 1, 1, nx, 2, 1);
  }
}

/* interpx_new (lip.f90:737-766) on pencils */
static void interpx_pen(const oc_box *b, const double *in, double *out, int nlin) {
  const int nx = b->nx;
  const double *c = b->fd_ix;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nlin; k++) {
    const double *g = in + (size_t)k * nx - 1;
    double *o = out + (size_t)k * nx - 1;
    o[1] = o[nx - 1] = o[nx] = 0.0;
    for (int q = 1; q <= 4; q++) {
      o[1] += F8(c, q, 1) * g[q];
      o[nx - 1] += F8(c, q, nx - 1) * g[nx - 4 + q];
      o[nx] += F8(c, q, nx) * g[nx - 4 + q];
    }
    for (int i = 2; i <= nx - 2; i++) {
      double t = 0.0;
      for (int q = 1; q <= 4; q++) t += F8(c, q, 3) * g[i - 2 + q];
      o[i] = t;
    }
    f8_solve(c, o + 1, 1, nx, 2, 1);
  }
}

/* diffy_inplace (lip.f90:666-698) on the ny rows of one plane (stride s) */
static void diffy_plane(const oc_box *b, const double *in, double *out, int s) {
  const int ny = b->ny;
  const double *c = b->fd_dvdy;
  for (int k = 0; k < s; k++) {
    RW(out, s, k, 1) = RW(out, s, k, 2) = RW(out, s, k, ny - 1) = RW(out, s, k, ny) = 0.0;
    for (int q = 1; q <= 4; q++) {
      RW(out, s, k, 1) += F8(c, q, 1) * RW(in, s, k, q);
      RW(out, s, k, 2) += F8(c, q, 2) * RW(in, s, k, q);
      RW(out, s, k, ny - 1) += F8(c, q, ny - 1) * RW(in, s, k, ny - 4 + q);
      RW(out, s, k, ny) += F8(c, q, ny) * RW(in, s, k, ny - 4 + q);
    }
    for (int j = 3; j <= ny - 2; j++) {
      double t = 0.0;
      for (int q = 1; q <= 5; q++) t += F8(c, q, j) * RW(in, s, k, j - 3 + q);
      RW(out, s, k, j) = t;
    }
    f8_solve(c, &RW(out, s, k, 1), s, ny, 2, 1);
  }
}

/* interpy_new (lip.f90:771-809) on one plane: tipo 0 takes ny+1 rows (u, w
 * and their x derivatives; the wall value is 0), tipo 1 the ny rows of p */
static void interpy_plane(const oc_box *b, const double *in, double *out, int s, int tipo) {
  const int ny = b->ny, ch = tipo == 1 ? 0 : 1;
  const double *c = tipo == 1 ? b->fd_iyp : b->fd_iy;
  for (int k = 0; k < s; k++) {
    double r1 = 0.0, r2 = 0.0, rm = 0.0, rn = 0.0;
    for (int q = 1; q <= 4; q++) r1 += F8(c, q, 1) * RW(in, s, k, 1 + q);
    for (int q = 1; q <= 5; q++) r2 += F8(c, q, 2) * RW(in, s, k, 1 + q);
    RW(out, s, k, 1) = r1;
    RW(out, s, k, 2) = r2;
    for (int j = 3; j <= ny - 2; j++) {
      double t = 0.0;
      for (int q = 1; q <= 4; q++) t += F8(c, q, j) * RW(in, s, k, j - 2 + q);
      RW(out, s, k, j) = t;
    }
    if (tipo == 0) RW(out, s, k, 1) = 0.0;
    for (int q = 1; q <= 3 + ch; q++) {
      rm += F8(c, q, ny - 1) * RW(in, s, k, ny - 3 + q);
      rn += F8(c, q, ny) * RW(in, s, k, ny - 3 + q);
    }
    RW(out, s, k, ny - 1) = rm;
    RW(out, s, k, ny) = rn;
    f8_solve(c, &RW(out, s, k, 1), s, ny, 3, 2); /* rows 2..ny (lip.f90:801-807) */
  }
}

#define ST(b, id, i, j) (b)->st[((size_t)(id) * (b)->nx + (size_t)((i)-1)) * ((b)->ny + 1) + (size_t)((j)-1)]

void oc_statsp(oc_box *b) {
  const int nx = b->nx, ny = b->ny, nz = b->nz, s = b->nz1, ru = ny + 1, nz2 = b->nz2;
  const int mpu = s * ru, mpv = s * ny;
  const double *u = b->fld[OC_U], *v = b->fld[OC_V], *w = b->fld[OC_W], *p = b->fld[OC_P];
  const double *kaz = b->kaz;
  if (!b->st) b->st = zalloc((size_t)OC_NSTAT * nx * (ny + 1));
  if (!b->ens) {
    b->ens = (float *)calloc((size_t)nz * ny * nx, sizeof(float));
    if (!b->ens) abort();
  }
  /* prologue (rhsp.F90:67-69, 72-90): u_x = interpx_new(u) moved one plane
   * up (u_x(1) keeps the plane-1 value), dudx = diffx(u_x), dvdx = diffx(v),
   * dwdx = diffx(w), each back to planes */
  double *ux = zalloc(b->nu), *dudx = zalloc(b->nu), *dvdx = zalloc(b->nv), *dwdx = zalloc(b->nu);
  double *pa = zalloc(b->nu), *pb = zalloc(b->nu);
  chp2x(b, b->pen_u, u, ru);
  chp2x(b, b->pen_v, v, ny);
  chp2x(b, b->pen_w, w, ru);
  interpx_pen(b, b->pen_u, pa, mpu);
  for (int l = 0; l < mpu; l++) {
    double *q = pa + (size_t)l * nx - 1;
    for (int i = nx - 1; i >= 1; i--) q[i + 1] = q[i];
  }
  diffx_pen(b, pa, pb, mpu);
  chx2p(b, ux, pa, ru);
  chx2p(b, dudx, pb, ru);
  diffx_pen(b, b->pen_v, pa, mpv);
  chx2p(b, dvdx, pa, ny);
  diffx_pen(b, b->pen_w, pa, mpu);
  chx2p(b, dwdx, pa, ru);
  free(pa);
  free(pb);

#pragma omp parallel
  {
    /* buf1 = (u_x)_y, buf2 = w_y, buf3 = (dwdx)_y, buf4 = (dudx)_y, buf5 = p_y,
     * buf6 = dudy, buf7 = dwdy, buf8 = dvdy, all at the v rows (stotf.F90:105-116) */
    double *bf[9];
    for (int q = 1; q <= 8; q++) bf[q] = zalloc((size_t)s * ny);
    double *fo = zalloc(s), *ph[9];
    for (int q = 0; q < 9; q++) ph[q] = zalloc(nz);
#pragma omp for schedule(static)
    for (int i = 1; i <= nx; i++) {
      const double *vi = PL(b, v, i, ny), *dvi = PL(b, dvdx, i, ny);
      interpy_plane(b, PL(b, ux, i, ru), bf[1], s, 0);
      interpy_plane(b, PL(b, w, i, ru), bf[2], s, 0);
      interpy_plane(b, PL(b, dwdx, i, ru), bf[3], s, 0);
      interpy_plane(b, PL(b, dudx, i, ru), bf[4], s, 0);
      interpy_plane(b, PL(b, p, i, ny), bf[5], s, 1);
      diffy_plane(b, bf[1], bf[6], s);
      diffy_plane(b, bf[2], bf[7], s);
      diffy_plane(b, vi, bf[8], s);
      for (int j = 1; j <= ny; j++) {
#define CR(a, k) RW(a, s, 2 * (k), j)
#define CI(a, k) RW(a, s, 2 * (k) + 1, j)
        const double *B1 = bf[1], *B2 = bf[2], *B3 = bf[3], *B4 = bf[4], *B5 = bf[5], *B6 = bf[6], *B7 = bf[7],
                     *B8 = bf[8];
        /* k = 0: means (stotf.F90:128-144) */
        ST(b, 4, i, j) += CR(B1, 0);
        ST(b, 6, i, j) += CR(B2, 0);
        ST(b, 5, i, j) += CR(vi, 0);
        ST(b, 8, i, j) += CR(B5, 0);
        ST(b, 32, i, j) += CR(B4, 0);
        ST(b, 34, i, j) += CR(dvi, 0);
        ST(b, 36, i, j) += CR(B3, 0);
        ST(b, 33, i, j) += CR(B6, 0);
        ST(b, 35, i, j) += CR(B8, 0);
        ST(b, 37, i, j) += CR(B7, 0);
        ST(b, 13, i, j) -= CR(B3, 0);
        ST(b, 12, i, j) += CR(B7, 0);
        ST(b, 14, i, j) += CR(dvi, 0) - CR(B6, 0);
        /* the nine velocity gradients in physical space, enstrophy
         * (stotf.F90:148-154); d/dz multiplies by kaz = i k/az */
        const double *grad[9] = {B4, B6, B1, dvi, B8, vi, B3, B7, B2}; /* dudx dudy dudz dvdx dvdy dvdz dwdx dwdy dwdz */
        for (int q = 0; q < 9; q++) {
          const int dz = (q % 3) == 2;
          for (int k = 0; k <= nz2; k++) {
            double re = CR(grad[q], k), im = CI(grad[q], k);
            fo[2 * k] = dz ? -kaz[k] * im : re;
            fo[2 * k + 1] = dz ? kaz[k] * re : im;
          }
          oc_c2r(b, fo, ph[q]);
        }
        float *ens = b->ens + ((size_t)(i - 1) * ny + (j - 1)) * nz;
        for (int z = 0; z < nz; z++)
          ens[z] = (float)sqrt(sq(ph[7][z] - ph[5][z]) + sq(ph[2][z] - ph[6][z]) + sq(ph[3][z] - ph[1][z]));
        /* sums over the modes (stotf.F90:156-218) */
        for (int k = 0; k <= nz2; k++) {
          const double cte = k == 0 ? 1.0 : 2.0, kz = kaz[k];
          const double b1r = CR(B1, k), b1i = CI(B1, k), b2r = CR(B2, k), b2i = CI(B2, k);
          const double b3r = CR(B3, k), b3i = CI(B3, k), b4r = CR(B4, k), b4i = CI(B4, k);
          const double b5r = CR(B5, k), b5i = CI(B5, k), b6r = CR(B6, k), b6i = CI(B6, k);
          const double b7r = CR(B7, k), b7i = CI(B7, k), b8r = CR(B8, k), b8i = CI(B8, k);
          const double vr = CR(vi, k), vim = CI(vi, k), dr = CR(dvi, k), di = CI(dvi, k);
          /* z derivatives: a kaz = (-kz a_i, kz a_r) */
          const double uzr = -kz * b1i, uzi = kz * b1r, vzr = -kz * vim, vzi = kz * vr;
          const double wzr = -kz * b2i, wzi = kz * b2r;
#define RE(ar, ai, br, bi) ((ar) * (br) + (ai) * (bi)) /* Re(a conj(b)) */
          ST(b, 0, i, j) += cte * RE(b1r, b1i, b1r, b1i);
          ST(b, 2, i, j) += cte * RE(b2r, b2i, b2r, b2i);
          ST(b, 1, i, j) += cte * RE(vr, vim, vr, vim);
          ST(b, 3, i, j) += cte * RE(b1r, b1i, vr, vim);
          ST(b, 15, i, j) += cte * RE(b1r, b1i, b2r, b2i);
          ST(b, 16, i, j) += cte * RE(vr, vim, b2r, b2i);
          ST(b, 21, i, j) += cte * RE(b5r, b5i, b1r, b1i);
          ST(b, 22, i, j) += cte * RE(b5r, b5i, vr, vim);
          ST(b, 7, i, j) += cte * RE(b5r, b5i, b5r, b5i);
          ST(b, 27, i, j) += cte * RE(b5r, b5i, wzr, wzi);
          ST(b, 23, i, j) += cte * RE(b5r, b5i, b4r, b4i);
          ST(b, 26, i, j) += cte * RE(b5r, b5i, dr, di);
          ST(b, 24, i, j) += cte * RE(b5r, b5i, b8r, b8i);
          ST(b, 25, i, j) += cte * RE(b5r, b5i, b6r, b6i);
          double ar = b7r - vzr, ai = b7i - vzi;
          ST(b, 9, i, j) += cte * (ar * ar + ai * ai);
          ar = uzr - b3r; ai = uzi - b3i;
          ST(b, 10, i, j) += cte * (ar * ar + ai * ai);
          ar = dr - b6r; ai = di - b6i;
          ST(b, 11, i, j) += cte * (ar * ar + ai * ai);
          ST(b, 17, i, j) += cte * (RE(b4r, b4i, b4r, b4i) + RE(uzr, uzi, uzr, uzi) + RE(b6r, b6i, b6r, b6i));
          ST(b, 18, i, j) += cte * (RE(dr, di, dr, di) + RE(vzr, vzi, vzr, vzi) + RE(b8r, b8i, b8r, b8i));
          ST(b, 19, i, j) += cte * (RE(wzr, wzi, wzr, wzi) + RE(b3r, b3i, b3r, b3i) + RE(b7r, b7i, b7r, b7i));
          ST(b, 20, i, j) += cte * (RE(uzr, uzi, vzr, vzi) + RE(b4r, b4i, dr, di) + RE(b6r, b6i, b8r, b8i));
#undef RE
        }
        /* triple products of u_y, v, w_y in physical space (stotf.F90:256-283) */
        oc_c2r(b, &RW(B1, s, 0, j), ph[0]);
        oc_c2r(b, &RW(B2, s, 0, j), ph[1]);
        oc_c2r(b, &RW(vi, s, 0, j), ph[2]);
        for (int z = 0; z < nz; z++) {
          const double U = ph[0][z], W = ph[1][z], V = ph[2][z];
          ST(b, 38, i, j) += U * U * U;
          ST(b, 39, i, j) += U * W * W;
          ST(b, 28, i, j) += V * V * V;
          ST(b, 31, i, j) += V * W * W;
          ST(b, 30, i, j) += U * V * V;
          ST(b, 29, i, j) += V * U * U;
        }
#undef CR
#undef CI
      }
    }
    for (int q = 1; q <= 8; q++) free(bf[q]);
    for (int q = 0; q < 9; q++) free(ph[q]);
    free(fo);
  }
  b->ical += 1;
  free(ux);
  free(dudx);
  free(dvdx);
  free(dwdx);
}

void oc_stats_get(const oc_box *b, int id, double *out) {
  const size_t n = (size_t)b->nx * (b->ny + 1);
  if (b->st)
    memcpy(out, b->st + (size_t)id * n, n * sizeof(double));
  else
    memset(out, 0, n * sizeof(double));
}
void oc_stats_reset(oc_box *b) {
  if (b->st) memset(b->st, 0, (size_t)OC_NSTAT * b->nx * (b->ny + 1) * sizeof(double));
  b->ical = 0;
}
int oc_stats_ical(const oc_box *b) { return b->ical; }
const float *oc_stats_ens(const oc_box *b) { return b->ens; }
const double *oc_stats_coef(const oc_box *b, const char *name) {
  if (!strcmp(name, "fd_dvdy")) return b->fd_dvdy + 8;
  if (!strcmp(name, "fd_dx")) return b->fd_dx + 8;
  if (!strcmp(name, "fd_ix")) return b->fd_ix + 8;
  if (!strcmp(name, "fd_iy")) return b->fd_iy + 8;
  if (!strcmp(name, "fd_iyp")) return b->fd_iyp + 8;
  return NULL;
}
/* the four operators on one plane / pencil set, for the tests:
 * op 0 diffy (ny rows), 1 interpy u-type (ny+1 rows in), 2 interpy p (ny
 * rows in) on nplanes planes of s = nz1 columns; op 3 diffx, 4 interpx on
 * nlin pencils of nx points */
void oc_stats_op(const oc_box *b, int op, const double *in, double *out, int n) {
  const int s = b->nz1, ny = b->ny;
  if (op <= 2) {
    for (int i = 0; i < n; i++) {
      if (op == 0) diffy_plane(b, in + (size_t)i * s * ny, out + (size_t)i * s * ny, s);
      else interpy_plane(b, in + (size_t)i * s * (op == 1 ? ny + 1 : ny), out + (size_t)i * s * ny, s, op == 1 ? 0 : 1);
    }
  } else if (op == 3) {
    diffx_pen(b, in, out, n);
  } else {
    interpx_pen(b, in, out, n);
  }
}
