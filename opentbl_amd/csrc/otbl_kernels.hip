// HIP kernels of the RK3 substep (FP64 path) for gfx950.
//
// Data layout (the reference's planes): a complex field (0:nz2, R, 1:nx) as
// doubles, real index kr = 0..nz1-1 fastest, then row j = 1..R, then plane i.
// Element (kr, j, i) sits at (i-1)*P + (j-1)*nz1 + kr with P = R*nz1.
//
// Line solves map one thread to one line:
//  * x lines (lip.f90 interpxx/difvisxx): a line is a fixed (kr, j); the
//    thread marches over planes, so a wavefront reads 64 consecutive doubles
//    of one plane per step (coalesced), no plane<->pencil transpose.
//  * y lines (interpyy, differzy, vistyy, implzy, soltriy): a line is a fixed
//    (kr, i); the thread marches over rows, again reading consecutive kr.
// The z transforms work on one (j, i) row per workgroup in LDS (mixed-radix
// Stockham, radices 8/4/2/3/5/7), two real signals per complex transform.
#include "otbl_kcommon.h"
#include "otbl_fft.h"
#include "otbl_ycol.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <type_traits>
#include <stdexcept>

namespace otbl {


// ---------------------------------------------------------------------------
// coefficient factorisation (coefinv, coeft.f90:1338-1353) in FP64
// ---------------------------------------------------------------------------
struct FactorJob {
  double *a[12];
  int n[12];
};

__global__ void factor_kernel(FactorJob job) {
  int t = threadIdx.x;
  if (t >= 12) return;
  double *a = job.a[t];
  int neq = job.n[t];
  for (int j = 2; j <= neq; ++j) {
    double d = 1.0 / C3(a, 2, j - 1);
    C3(a, 2, j - 1) = d;
    C3(a, 1, j) = -C3(a, 1, j) * d;
    C3(a, 2, j) = C3(a, 2, j) + C3(a, 1, j) * C3(a, 3, j - 1);
  }
  C3(a, 2, neq) = 1.0 / C3(a, 2, neq);
}

void k_factor_coef(Box &b, double *) {
  FactorJob job;
  const DevCoef &c = b.dc;
  double *arr[12] = {(double *)c.inxu, (double *)c.inxv, (double *)c.inyu, (double *)c.inyv,
                     (double *)c.dcxu, (double *)c.dcxv, (double *)c.dcyu, (double *)c.dcyv,
                     (double *)c.vixu, (double *)c.vixv, (double *)c.viyu, (double *)c.viyv};
  int n[12] = {b.nx, b.nx, b.ny, b.ny, b.nx, b.nx, b.ny + 1, b.ny, b.nx, b.nx, b.ny + 1, b.ny};
  for (int q = 0; q < 12; ++q) {
    job.a[q] = arr[q];
    job.n[q] = n[q];
  }
  LAUNCH(b, "factor_coef", 1, 64, 0, factor_kernel, job);
}

// coefinv on one caller array (the drop-in of coeft.f90:1338-1353): the
// same in-place LU on a (3, npter1:neq) array
__global__ void coefinv_kernel(double *a, int nrow) {
  if (threadIdx.x != 0) return;
  for (int j = 2; j <= nrow; ++j) {
    double d = 1.0 / C3(a, 2, j - 1);
    C3(a, 2, j - 1) = d;
    C3(a, 1, j) = -C3(a, 1, j) * d;
    C3(a, 2, j) = C3(a, 2, j) + C3(a, 1, j) * C3(a, 3, j - 1);
  }
  C3(a, 2, nrow) = 1.0 / C3(a, 2, nrow);
}

void k_coefinv(Box &b, double *a, int nrow) { LAUNCH(b, "coefinv", 1, 64, 0, coefinv_kernel, a, nrow); }

// soltrix's sweeps (coeft.f90:1952-1964) over nlen columns of nstr doubles,
// one thread per column: rows 0..nrow-1 of each column of u solved into the
// same rows of us with the factored a (row i at a[3 i]); in place allowed
// (each thread reads and writes only its own column)
__global__ void soltrix_kernel(const double *u, double *us, const double *__restrict__ a, int nstr, int nlen,
                               int nrow) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlen) return;
  const double *uc = u + (size_t)j * nstr;
  double *sc = us + (size_t)j * nstr;
  double y = uc[0];
  sc[0] = y;
  for (int i = 1; i < nrow; ++i) {
    y = uc[i] + a[3 * i] * y;
    sc[i] = y;
  }
  double x = y * a[3 * (nrow - 1) + 1];
  sc[nrow - 1] = x;
  for (int i = nrow - 2; i >= 0; --i) {
    x = (sc[i] - a[3 * i + 2] * x) * a[3 * i + 1];
    sc[i] = x;
  }
}

void k_soltrix(Box &b, const double *u, double *us, const double *a, int nstr, int nlen, int nrow) {
  LAUNCH(b, "soltrix", nblk(nlen, 64), 64, 0, soltrix_kernel, u, us, a, nstr, nlen, nrow);
}

// soltriy (pois.f90:211-239): LU of the tridiagonal a(3, ny1) in place, then
// the Re and Im parts of complex column ii of u(0:nmode-1, ny1) solved into
// us (us may be u).  Lane 0 factors, lanes 0 and 1 sweep one part each.
__global__ void soltriy_kernel(const double *u, double *us, int ii, double *a, int ny1, int nmode) {
  const int t = threadIdx.x;
  if (t == 0) {
    for (int j = 2; j <= ny1; ++j) {
      const double d = 1.0 / C3(a, 2, j - 1);
      C3(a, 2, j - 1) = d;
      C3(a, 1, j) = -C3(a, 1, j) * d;
      C3(a, 2, j) = C3(a, 2, j) + C3(a, 1, j) * C3(a, 3, j - 1);
    }
    C3(a, 2, ny1) = 1.0 / C3(a, 2, ny1);
  }
  __syncthreads();
  if (t >= 2) return;
  const size_t s = 2 * (size_t)nmode, c = 2 * (size_t)ii + t;
  double y = u[c];
  us[c] = y;
  for (int j = 2; j <= ny1; ++j) {
    y = u[(size_t)(j - 1) * s + c] + C3(a, 1, j) * y;
    us[(size_t)(j - 1) * s + c] = y;
  }
  double x = y * C3(a, 2, ny1);
  us[(size_t)(ny1 - 1) * s + c] = x;
  for (int j = ny1 - 1; j >= 1; --j) {
    x = (us[(size_t)(j - 1) * s + c] - C3(a, 3, j) * x) * C3(a, 2, j);
    us[(size_t)(j - 1) * s + c] = x;
  }
}

void k_soltriy(Box &b, const double *u, double *us, int ii, double *a, int ny1, int nmode) {
  LAUNCH(b, "soltriy", 1, 64, 0, soltriy_kernel, u, us, ii, a, ny1, nmode);
}

// ---------------------------------------------------------------------------
// x-line operators
// ---------------------------------------------------------------------------
// interpxx (lip.f90:248-300): midpoint interpolation along x.
__global__ void interpxx_kernel(const double *__restrict__ gen, double *__restrict__ geni,
                                const double *__restrict__ icon, const double *__restrict__ econ,
                                const double *__restrict__ inbx, int chi, int chi2, int nx,
                                size_t P) {
  size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= P) return;
  const double *g = gen + l;
  double *o = geni + l;
  double g1 = g[0], g2 = g[P], g3 = g[2 * P], g4 = g[3 * P];
  const double gl0 = g[(size_t)(nx - 1) * P], gl1 = g[(size_t)(nx - 2) * P],
               gl2 = g[(size_t)(nx - 3) * P], gl3 = g[(size_t)(nx - 4) * P];
  double y = M4(inbx, chi2, 1) * g1 + M4(inbx, chi2, 2) * g2 + M4(inbx, chi2, 3) * g3 +
             M4(inbx, chi2, 4) * g4;
  o[0] = y;
  double gi = g2; // g(i)
  for (int i = 2; i <= nx - 1; ++i) {
    double gn = g[(size_t)i * P]; // g(i+1)
    double rhs = C2(econ, 1, i) * gi + C2(econ, 2, i) * gn;
    y = rhs + C3(icon, 1, i) * y;
    o[(size_t)(i - 1) * P] = y;
    gi = gn;
  }
  {
    double rhs = M4(inbx, chi, 1) * gl0 + M4(inbx, chi, 2) * gl1 + M4(inbx, chi, 3) * gl2 +
                 M4(inbx, chi, 4) * gl3;
    y = rhs + C3(icon, 1, nx) * y;
  }
  double x = y * C3(icon, 2, nx);
  o[(size_t)(nx - 1) * P] = x;
  for (int i = nx - 1; i >= 1; --i) {
    double yi = o[(size_t)(i - 1) * P];
    x = (yi - C3(icon, 3, i) * x) * C3(icon, 2, i);
    o[(size_t)(i - 1) * P] = x;
  }
}

void k_interpxx(Box &b, const double *gen, double *geni, int R, int which_u) {
  size_t P = (size_t)b.nz1 * R;
  const double *icon = which_u ? b.dc.inxu : b.dc.inxv;
  const double *econ = which_u ? b.dc.cofiux : b.dc.cofivx;
  int chi = which_u ? 2 : 1;
  int chi2 = which_u ? chi + 1 : chi + 3;
  if (b.xwin) {
    kx_interpxx(b, gen, geni, icon, econ, chi, chi2, P);
    return;
  }
  LAUNCH(b, "interpxx", nblk(P), TPB, 0, interpxx_kernel, gen, geni, icon, econ, b.dc.inbx, chi,
         chi2, b.nx, P);
}

// difvisxx (lip.f90:52-142): convective d/dx of the product plus rex times
// the compact d2/dx2 of the velocity.  mode 0: out = result; mode 1:
// out += result except on the last plane, where out = result (the outflow
// handling of rhsp.F90:200,204).  sc, sv: scratch fields for the two
// forward sweeps.
__global__ void difvisxx_kernel(const double *__restrict__ gp, const double *__restrict__ q,
                                double *__restrict__ out, double *__restrict__ sc,
                                double *__restrict__ sv, const double *__restrict__ icon,
                                const double *__restrict__ iconv, const double *__restrict__ econ,
                                const double *__restrict__ econv, const double *__restrict__ dcbx,
                                const double *__restrict__ cofvbx, int chv, int ic, int nx,
                                size_t P, double rex, double idx, int mode) {
  size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= P) return;
  const double *g = gp + l;
  const double *v = q + l;
  double *c = sc + l, *s = sv + l, *o = out + l;
  const int chv1 = chv + 2;
  double g1 = g[0], g2 = g[P], g3 = g[2 * P], g4 = g[3 * P];
  double q1 = v[0], q2 = v[P], q3 = v[2 * P], q4 = v[3 * P];
  double yc = M5(dcbx, chv, 1) * g1 + M5(dcbx, chv, 2) * g2 + M5(dcbx, chv, 3) * g3 +
              M5(dcbx, chv, 4) * g4;
  double yv = M4(cofvbx, ic, 1) * q1 + M4(cofvbx, ic, 2) * q2 + M4(cofvbx, ic, 3) * q3 +
              M4(cofvbx, ic, 4) * q4;
  c[0] = yc;
  s[0] = yv;
  double gim1 = g1, qim1 = q1, qi = q2;
  for (int i = 2; i <= nx - 1; ++i) {
    double gi = g[(size_t)(i - 1) * P];
    double qn = v[(size_t)i * P];
    double rc = C2(econ, 1, i) * gim1 + C2(econ, 2, i) * gi;
    double rv = C3(econv, 1, i) * qn + C3(econv, 2, i) * qi + C3(econv, 3, i) * qim1;
    yc = rc + C3(icon, 1, i) * yc;
    yv = rv + C3(iconv, 1, i) * yv;
    c[(size_t)(i - 1) * P] = yc;
    s[(size_t)(i - 1) * P] = yv;
    gim1 = gi;
    qim1 = qi;
    qi = qn;
  }
  double gn = g[(size_t)(nx - 1) * P], gn1 = g[(size_t)(nx - 2) * P], gn2 = g[(size_t)(nx - 3) * P],
         gn3 = g[(size_t)(nx - 4) * P];
  double qn = qi, qn1 = qim1, qn2 = v[(size_t)(nx - 3) * P], qn3 = v[(size_t)(nx - 4) * P];
  double rc = M5(dcbx, chv1, 1) * gn + M5(dcbx, chv1, 2) * gn1 + M5(dcbx, chv1, 3) * gn2 +
              M5(dcbx, chv1, 4) * gn3;
  double rv = M4(cofvbx, ic + 1, 1) * qn + M4(cofvbx, ic + 1, 2) * qn1 +
              M4(cofvbx, ic + 1, 3) * qn2 + M4(cofvbx, ic + 1, 4) * qn3;
  yc = rc + C3(icon, 1, nx) * yc;
  yv = rv + C3(iconv, 1, nx) * yv;
  double xc = yc * C3(icon, 2, nx);
  double xv = yv * C3(iconv, 2, nx);
  // outflow rows: U_inf du/dx convective, no viscous term (lip.f90:105,132)
  double last = 1.0 * (qn - qn1) * idx + rex * 0.0;
  size_t on = (size_t)(nx - 1) * P;
  o[on] = last;
  for (int i = nx - 1; i >= 1; --i) {
    size_t oi = (size_t)(i - 1) * P;
    xc = (c[oi] - C3(icon, 3, i) * xc) * C3(icon, 2, i);
    xv = (s[oi] - C3(iconv, 3, i) * xv) * C3(iconv, 2, i);
    double r = xc + rex * xv;
    if (mode == 1) r = o[oi] + r;
    o[oi] = r;
  }
}

void k_difvisxx(Box &b, const double *prod, const double *vel, double *out, int R, int which_u,
                int mode) {
  size_t P = (size_t)b.nz1 * R;
  const DevCoef &c = b.dc;
  int vel_u = which_u; // viscous coefficient set: u (ic=3) or v/w (ic=1)
  if (b.xwin) {
    kx_difvisxx(b, prod, vel, out, which_u ? c.dcxu : c.dcxv, which_u ? c.vixu : c.vixv,
                which_u ? c.cofcxu : c.cofcxv, which_u ? c.cofvxu : c.cofvxv, which_u ? 1 : 2, vel_u ? 3 : 1, P,
                mode);
    return;
  }
  LAUNCH(b, "difvisxx", nblk(P), TPB, 0, difvisxx_kernel, prod, vel, out, b.ya, b.yb,
         which_u ? c.dcxu : c.dcxv, which_u ? c.vixu : c.vixv, which_u ? c.cofcxu : c.cofcxv,
         which_u ? c.cofvxu : c.cofvxv, c.dcbx, c.cofvbx, which_u ? 1 : 2, vel_u ? 3 : 1, b.nx,
         P, b.hc.rex, b.hc.idx, mode);
}

// ---------------------------------------------------------------------------
// y-line operators.  Column c = (kr, plane); rows at stride nz1.
// ---------------------------------------------------------------------------
// Tridiagonal sweep helpers on a column already holding right-hand sides
// rows 1..n in `r` (stride s).
__device__ inline void tri_solve_col(double *r, size_t s, const double *icon, int n) {
  double y = r[0];
  for (int j = 2; j <= n; ++j) {
    y = r[(size_t)(j - 1) * s] + C3(icon, 1, j) * y;
    r[(size_t)(j - 1) * s] = y;
  }
  double x = y * C3(icon, 2, n);
  r[(size_t)(n - 1) * s] = x;
  for (int j = n - 1; j >= 1; --j) {
    x = (r[(size_t)(j - 1) * s] - C3(icon, 3, j) * x) * C3(icon, 2, j);
    r[(size_t)(j - 1) * s] = x;
  }
}

// interpyy (lip.f90:359-440) on spectral rows.
__global__ void interpyy_kernel(const double *__restrict__ gen, double *__restrict__ geni,
                                const double *__restrict__ icon, const double *__restrict__ econ,
                                const double *__restrict__ inby, int loguv, int ny, int nz1,
                                int i0, int nplanes, size_t ips, size_t ops) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * nplanes) return;
  int kr = (int)(cidx % nz1);
  size_t ip = (size_t)(i0 - 1) + cidx / nz1;
  const double *g = gen + ip * ips + kr;
  double *o = geni + ip * ops + kr;
  const size_t s = nz1;
#define G(j) g[(size_t)((j)-1) * s]
  if (loguv == 0) {
    double top = M4(inby, 1, 1) * G(ny + 1) + M4(inby, 1, 2) * G(ny) + M4(inby, 1, 3) * G(ny - 1) +
                 M4(inby, 1, 4) * G(ny - 2);
    o[0] = 0.0;
    for (int j = 2; j <= ny - 1; ++j) o[(size_t)(j - 1) * s] = C2(econ, 1, j) * G(j) + C2(econ, 2, j) * G(j + 1);
    o[(size_t)(ny - 1) * s] = top;
  } else {
    double bot = M4(inby, 3, 1) * G(1) + M4(inby, 3, 2) * G(2) + M4(inby, 3, 3) * G(3) +
                 M4(inby, 3, 4) * G(4);
    double top = M4(inby, 4, 1) * G(ny) + M4(inby, 4, 2) * G(ny - 1) + M4(inby, 4, 3) * G(ny - 2) +
                 M4(inby, 4, 4) * G(ny - 3);
    for (int j = ny - 1; j >= 2; --j) o[(size_t)(j - 1) * s] = C2(econ, 2, j) * G(j) + C2(econ, 1, j) * G(j - 1);
    o[0] = bot;
    o[(size_t)(ny - 1) * s] = top;
  }
#undef G
  tri_solve_col(o, s, icon, ny);
}

void k_interpyy(Box &b, const double *gen, double *geni, int loguv, int i0, int nplanes,
                size_t ips, size_t ops) {
  if (b.yL) {
    ky_interpyy(b, gen, geni, loguv, i0, nplanes, ips, ops);
    return;
  }
  size_t n = (size_t)b.nz1 * nplanes;
  LAUNCH(b, "interpyy", nblk(n), TPB, 0, interpyy_kernel, gen, geni,
         loguv == 0 ? b.dc.inyu : b.dc.inyv, loguv == 0 ? b.dc.cofiuy : b.dc.cofivy, b.dc.inby,
         loguv, b.ny, b.nz1, i0, nplanes, ips, ops);
}

// differzy (lip.f90:152-232).  which 0: v rows -> u rows (dcyu, cofcyu,
// chv=3, fwb1=-1, fwb2=0, jee=ny+1); which 1: u rows -> v rows (dcyv,
// cofcyv, chv=1, fwb1=0, fwb2=1, jee=ny).
__device__ inline void differzy_col(const double *g, double *o, size_t s, int which, int ny,
                                    const DevCoef &c) {
#define G(j) g[(size_t)((j)-1) * s]
  if (which == 0) {
    const int jee = ny + 1;
    double bot = M4(c.dcby, 3, 1) * G(1) + M4(c.dcby, 3, 2) * G(2) + M4(c.dcby, 3, 3) * G(3) +
                 M4(c.dcby, 3, 4) * G(4);
    double top = M4(c.dcby, 4, 1) * G(ny) + M4(c.dcby, 4, 2) * G(ny - 1) +
                 M4(c.dcby, 4, 3) * G(ny - 2) + M4(c.dcby, 4, 4) * G(ny - 3);
    for (int j = jee - 1; j >= 2; --j)
      o[(size_t)(j - 1) * s] = C2(c.cofcyu, 2, j) * G(j) + C2(c.cofcyu, 1, j) * G(j - 1);
    o[0] = bot;
    o[(size_t)(jee - 1) * s] = top;
    tri_solve_col(o, s, c.dcyu, jee);
  } else {
    const int jee = ny;
    double bot = M4(c.dcby, 1, 1) * G(1) + M4(c.dcby, 1, 2) * G(2) + M4(c.dcby, 1, 3) * G(3) +
                 M4(c.dcby, 1, 4) * G(4);
    double top = M4(c.dcby, 2, 1) * G(ny) + M4(c.dcby, 2, 2) * G(ny - 1) +
                 M4(c.dcby, 2, 3) * G(ny - 2) + M4(c.dcby, 2, 4) * G(ny - 3);
    for (int j = 2; j <= jee - 1; ++j)
      o[(size_t)(j - 1) * s] = C2(c.cofcyv, 2, j) * G(j + 1) + C2(c.cofcyv, 1, j) * G(j);
    o[0] = bot;
    o[(size_t)(jee - 1) * s] = top;
    tri_solve_col(o, s, c.dcyv, jee);
  }
#undef G
}

__global__ void differzy_kernel(const double *__restrict__ gen, double *__restrict__ res, DevCoef c,
                                int which, int ny, int nz1, int nplanes, size_t ips, size_t ops) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * nplanes) return;
  int kr = (int)(cidx % nz1);
  size_t ip = cidx / nz1;
  differzy_col(gen + ip * ips + kr, res + ip * ops + kr, nz1, which, ny, c);
}

void k_differzy(Box &b, const double *geni, double *res, int which, int nplanes, size_t ips,
                size_t ops) {
  if (b.yL) {
    ky_differzy(b, geni, res, which, nplanes, ips, ops);
    return;
  }
  size_t n = (size_t)b.nz1 * nplanes;
  LAUNCH(b, "differzy", nblk(n), TPB, 0, differzy_kernel, geni, res, b.dc, which, b.ny, b.nz1,
         nplanes, ips, ops);
}

// vistyy (lip.f90:447-513): compact d2/dy2 of one column into o.
__device__ inline void vistyy_col(const double *u, double *o, size_t s, int jee, int ny,
                                  const DevCoef &c) {
  const double *icon = (jee == ny) ? c.viyv : c.viyu;
  const double *econ = (jee == ny) ? c.cofvyv : c.cofvyu;
  const int ic = (jee == ny) ? 1 : 3;
#define U(j) u[(size_t)((j)-1) * s]
  o[0] = M4(c.cofvby, ic, 1) * U(1) + M4(c.cofvby, ic, 2) * U(2) + M4(c.cofvby, ic, 3) * U(3) +
         M4(c.cofvby, ic, 4) * U(4);
  for (int j = 2; j <= jee - 1; ++j)
    o[(size_t)(j - 1) * s] = C3(econ, 1, j) * U(j + 1) + C3(econ, 2, j) * U(j) + C3(econ, 3, j) * U(j - 1);
  o[(size_t)(jee - 1) * s] = M4(c.cofvby, ic + 1, 1) * U(jee) + M4(c.cofvby, ic + 1, 2) * U(jee - 1) +
                             M4(c.cofvby, ic + 1, 3) * U(jee - 2) + M4(c.cofvby, ic + 1, 4) * U(jee - 3);
#undef U
  tri_solve_col(o, s, icon, jee);
}

__global__ void vistyy_kernel(const double *__restrict__ u, double *__restrict__ res, DevCoef c,
                              int jee, int ny, int nz1, int nplanes) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * nplanes) return;
  int kr = (int)(cidx % nz1);
  size_t ip = cidx / nz1;
  size_t ps = (size_t)nz1 * jee;
  vistyy_col(u + ip * ps + kr, res + ip * ps + kr, nz1, jee, ny, c);
}

void k_vistyy(Box &b, const double *u, double *res, int R, int nplanes) {
  if (b.yL) {
    ky_vistyy(b, u, res, R, nplanes);
    return;
  }
  size_t n = (size_t)b.nz1 * nplanes;
  LAUNCH(b, "vistyy", nblk(n), TPB, 0, vistyy_kernel, u, res, b.dc, R, b.ny, b.nz1, nplanes);
}

// Implicit y-viscous factors for the current dt (implzy's a1, a2, a3 and
// their LU, lip.f90:545-568) in FP64, one workgroup per set (blockIdx.x 0:
// u-type, jee = ny+1; 1: v-type, jee = ny).  The rows are formed by all
// threads, the (sequential) factorisation runs in LDS, then the chunked
// solver's table of the factored system is built.
// layout per set (n = jee): f[0..n] a1 factored, f[n+1..2n+1] a2 (inverse
// diagonal), f[2n+2..3n+2] a3, f[3n+3] = a1(2) and f[3n+4] = a3(jee-1)
// before they are zeroed.  Dynamic LDS: 3*(ny+2) + 4*YNCH*L doubles.
__global__ void impl_factor_kernel(double *fu, double *fv, DevCoef c, int ny, const DevScalars *ds,
                                   double rkk_explicit, int use_explicit, double *tu, double *tv, int L, float *tuf,
                                   float *tvf, float *thu, float *thv) {
  extern __shared__ double shf[];
  const int jee = blockIdx.x == 0 ? ny + 1 : ny;
  const double *icon = blockIdx.x == 0 ? c.vyui : c.vyvi;
  const double *econ = blockIdx.x == 0 ? c.cofvyu : c.cofvyv;
  double *f = blockIdx.x == 0 ? fu : fv;
  double *tab = blockIdx.x == 0 ? tu : tv;
  const double v1 = use_explicit ? rkk_explicit : ds->rkk;
  double *a1 = shf, *a2 = shf + (ny + 2), *a3 = shf + 2 * (ny + 2);
  for (int j = threadIdx.x + 2; j <= jee - 1; j += blockDim.x) {
    a1[j] = C3(icon, 1, j) + v1 * C3(econ, 3, j);
    a2[j] = C3(icon, 2, j) + v1 * C3(econ, 2, j);
    a3[j] = C3(icon, 3, j) + v1 * C3(econ, 1, j);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    f[3 * (jee + 1)] = a1[2];
    f[3 * (jee + 1) + 1] = a3[jee - 1];
    a1[2] = 0.0;
    a3[jee - 1] = 0.0;
    for (int j = 3; j <= jee - 1; ++j) {
      double d = 1.0 / a2[j - 1];
      a2[j - 1] = d;
      a1[j] = -a1[j] * d;
      a2[j] = a2[j] + a1[j] * a3[j - 1];
    }
    a2[jee - 1] = 1.0 / a2[jee - 1];
  }
  __syncthreads();
  for (int j = threadIdx.x + 2; j <= jee - 1; j += blockDim.x) {
    f[j] = a1[j];
    f[(jee + 1) + j] = a2[j];
    f[2 * (jee + 1) + j] = a3[j];
  }
  if (L)
    ytab_build_block(
        tab, jee, L,
        [=](int j) {
          if (j == 1 || j == jee) return TriRow{0.0, 1.0, 0.0};
          return TriRow{a1[j], a2[j], a3[j]};
        },
        [=](int j) { return Sten3{C3(icon, 1, j), C3(icon, 2, j), C3(icon, 3, j)}; }, shf + 3 * (ny + 2));
  // FP32 copy of the table for the mixed-precision substep
  float *tabf = blockIdx.x == 0 ? tuf : tvf;
  if (L && tabf) {
    __syncthreads();
    for (int q = threadIdx.x; q < ytab_size(L); q += blockDim.x) tabf[q] = (float)tab[q];
  }
  // and of its half-wave chunking
  float *tabh = blockIdx.x == 0 ? thu : thv;
  if (L && tabh) {
    __syncthreads();
    ytab_halve_block(tabh, tab, L);
  }
}

static void launch_impl_factor(Box &b, double rkk, int use_explicit) {
  double *tu = b.yL ? b.ytab + YT_IMPU * ytab_size(b.yL) : nullptr;
  double *tv = b.yL ? b.ytab + YT_IMPV * ytab_size(b.yL) : nullptr;
  float *tuf = b.ytab_f ? b.ytab_f + YT_IMPU * ytab_size(b.yL) : nullptr;
  float *tvf = b.ytab_f ? b.ytab_f + YT_IMPV * ytab_size(b.yL) : nullptr;
  float *thu = b.ytabh_f ? b.ytabh_f + YT_IMPU * ytabh_size(b.yL / 2) : nullptr;
  float *thv = b.ytabh_f ? b.ytabh_f + YT_IMPV * ytabh_size(b.yL / 2) : nullptr;
  const size_t shm = (3 * (size_t)(b.ny + 2) + 4 * (size_t)YNCH * b.yL) * sizeof(double);
  LAUNCH(b, "impl_factor", 2, TPB, shm, impl_factor_kernel, b.implu, b.implv, b.dc, b.ny, b.ds, rkk,
         use_explicit, tu, tv, b.yL, tuf, tvf, thu, thv);
}

// implzy (lip.f90:528-597): interior rows 2..jee-1 of u from res; rows 1 and
// jee of u are the boundary values already in place.
__device__ inline void implzy_col(double *u, const double *res, size_t s, int jee,
                                  const double *icon, const double *f) {
  const double *a1 = f, *a2 = f + (jee + 1), *a3 = f + 2 * (jee + 1);
  const double a1b = f[3 * (jee + 1)], a3b = f[3 * (jee + 1) + 1];
#define R(j) res[(size_t)((j)-1) * s]
#define UU(j) u[(size_t)((j)-1) * s]
  double rm = R(1), r0 = R(2);
  double u1 = UU(1), ujee = UU(jee);
  double y = 0.0;
  for (int j = 2; j <= jee - 1; ++j) {
    double rp = R(j + 1);
    double rhs = C3(icon, 1, j) * rm + C3(icon, 2, j) * r0 + C3(icon, 3, j) * rp;
    if (j == 2) rhs -= a1b * u1;
    if (j == jee - 1) rhs -= a3b * ujee;
    y = (j == 2) ? rhs : rhs + a1[j] * y;
    UU(j) = y;
    rm = r0;
    r0 = rp;
  }
  double x = y * a2[jee - 1];
  UU(jee - 1) = x;
  for (int j = jee - 2; j >= 2; --j) {
    x = (UU(j) - a3[j] * x) * a2[j];
    UU(j) = x;
  }
#undef R
#undef UU
}

__global__ void implzy_kernel(double *u, const double *res, DevCoef c, const double *f, int jee,
                              int ny, int nz1, int nplanes) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * nplanes) return;
  int kr = (int)(cidx % nz1);
  size_t ip = cidx / nz1;
  size_t ps = (size_t)nz1 * jee;
  implzy_col(u + ip * ps + kr, res + ip * ps + kr, nz1, jee, jee == ny + 1 ? c.vyui : c.vyvi, f);
}

void k_implzy_op(Box &b, double *u, const double *res, int R, double rkk, int nplanes) {
  launch_impl_factor(b, rkk, 1);
  if (b.yL) {
    ky_implzy(b, u, res, R, nplanes);
    return;
  }
  size_t n = (size_t)b.nz1 * nplanes;
  LAUNCH(b, "implzy", nblk(n), TPB, 0, implzy_kernel, u, res, b.dc, R == b.ny + 1 ? b.implu : b.implv,
         R, b.ny, b.nz1, nplanes);
}

// ---------------------------------------------------------------------------
// z transforms in LDS (FFT core in otbl_fft.h)
// ---------------------------------------------------------------------------
// LDS layout of the transform kernels: lo[64], hi[128] twiddle tables, then
// the signal buffers.
static constexpr int TW_SLOTS = 64 + 128;
// rows per block of the LDS products kernels: at most ZROWS_MAX, about 512
// points per block (Box::zrb)
static constexpr int ZROWS_MAX = 8;

// Hermitian-extended spectrum of two real signals packed as z = x + i y,
// x = c2r(X), y = c2r(Y).  X, Y rows hold modes 0..nz2 (re, im pairs); Im of
// mode 0 does not enter a real inverse transform.  Every z entry is written.
template <typename R>
__device__ inline void pack2(typename CT<R>::c *z, const R *X, const R *Y, int n, int nz2) {
  using C = typename CT<R>::c;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    int kk = k <= nz2 ? k : (k >= n - nz2 ? n - k : -1);
    C v = mkc<C>(0.0, 0.0);
    if (kk >= 0) {
      R xr = X ? X[2 * kk] : R(0), xi = X ? X[2 * kk + 1] : R(0);
      R yr = Y ? Y[2 * kk] : R(0), yi = Y ? Y[2 * kk + 1] : R(0);
      if (k == 0) v = mkc<C>(xr, yr);
      else if (k == kk) v = mkc<C>(xr - yi, xi + yr);
      else v = mkc<C>(xr + yi, -xi + yr);
    }
    z[k] = v;
  }
}

// Split the transform Z of z = a + i b into the spectra of a and b, modes
// 0..nz2, scaled by 1/n (the r2c scaling of ffts.f90:51).  With XA, A also
// receives the mean-flow terms of a product whose mean was left out of the
// transform: A_k += ck XA_k for k >= 1, A_0 += c0 XA_0, XA read before A is
// written (XA may be A itself: each k is read and written by one thread).
template <typename R>
__device__ inline void unpack2(const typename CT<R>::c *Z, R *A, R *B, int n, int nz2, const R *XA = nullptr,
                               R c0 = R(0), R ck = R(0)) {
  using C = typename CT<R>::c;
  const R sc = R(1.0 / n);
  for (int k = threadIdx.x; k <= nz2; k += blockDim.x) {
    C z = Z[k];
    C w = Z[k == 0 ? 0 : n - k];
    if (A) {
      R ar = R(0.5) * (z.x + w.x) * sc, ai = (k == 0) ? R(0) : R(0.5) * (z.y - w.y) * sc;
      if (XA) {
        const R xr = XA[2 * k], xi = XA[2 * k + 1];
        ar += (k == 0 ? c0 : ck) * xr;
        ai += (k == 0) ? R(0) : ck * xi;
      }
      A[2 * k] = ar;
      A[2 * k + 1] = ai;
    }
    if (B) {
      B[2 * k] = R(0.5) * (z.y + w.y) * sc;
      B[2 * k + 1] = (k == 0) ? R(0) : R(-0.5) * (z.x - w.x) * sc;
    }
  }
}

// The block's maxima of three non-negative values, written for row `row`
// of the per-row maxima (three arrays of nrows; rowmax_kernel reduces them,
// so no block contends for one address with every other).
__device__ inline void block_rowmax3(double a, double b, double c, double *rowmax, size_t nrows, size_t row) {
  __shared__ double red[3][TPB / 64];
  for (int o = 32; o > 0; o >>= 1) {
    a = fmax(a, __shfl_down(a, o, 64));
    b = fmax(b, __shfl_down(b, o, 64));
    c = fmax(c, __shfl_down(c, o, 64));
  }
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x / 64); ++q) {
      a = fmax(a, red[0][q]);
      b = fmax(b, red[1][q]);
      c = fmax(c, red[2][q]);
    }
    rowmax[row] = a;
    rowmax[nrows + row] = b;
    rowmax[2 * nrows + row] = c;
  }
  __syncthreads();
}

// Plane loop 1 of rhsp (rhsp.F90:120-153) for rb consecutive rows (j, i) of a
// plane (their signals transformed together, sharing the block's twiddles
// and barriers; rb rows of nz points fill the block better than one when nz
// is short): with u, u_x,
// w_x, v_x and u interpolated to v rows (uy) in spectral space, form
// r2c(u_x^2) -> wki1, r2c(u w_x) -> wki3, r2c(v_x uy) -> wki2 (rows <= ny),
// and the maxima |u|, |w_x|, |v_x|(j) for the CFL time step.  In FP32 the
// mean flow (mode 0) of u_x and uy stays out of the transforms and its
// products are added in spectral space (u_x^2 = U^2 + 2 U u_x' + u_x'^2,
// v_x uy = Uy v_x + v_x uy'), as in the register-resident products1_z.
template <typename R>
__global__ void __launch_bounds__(TPB)
products1_kernel(const R *__restrict__ u, R *wki1, R *wki2, R *wki3, const R *__restrict__ uy, FFTPlan plan, int ny,
                 int nz1, int nz2, double *rowmax, int nx, int rb) {
  using C = typename CT<R>::c;
  constexpr bool SPLIT = sizeof(R) == 4;
  extern __shared__ double2 sh_raw[];
  __shared__ R sU[2][ZROWS_MAX]; // u_x and uy of mode 0 of each row, read before any row is written
  C *sh = reinterpret_cast<C *>(sh_raw);
  const int n = plan.n;
  // rows j0 .. j0+nr-1 of plane i; row r's two signals at Q0 + 2 n r
  C *lo = sh, *hi = sh + 64, *Q0 = sh + TW_SLOTS, *Q1 = Q0 + 2 * n * rb;
  const int j0 = blockIdx.x * rb + 1, i = blockIdx.y + 1;
  const int nr = min(rb, ny + 2 - j0);
  auto ou = [&](int r) { return ((size_t)(i - 1) * (ny + 1) + (j0 + r - 1)) * nz1; };
  auto ov = [&](int r) { return ((size_t)(i - 1) * ny + (j0 + r - 1)) * nz1; };
  auto vrow = [&](int r) { return j0 + r <= ny; };
  if (SPLIT && threadIdx.x < nr) {
    sU[0][threadIdx.x] = wki1[ou(threadIdx.x)];
    sU[1][threadIdx.x] = vrow(threadIdx.x) ? uy[ou(threadIdx.x)] : R(0);
  }
  fill_twiddles(lo, hi, n);
  for (int r = 0; r < rb; ++r) {
    C *q = Q0 + 2 * n * r;
    const bool ok = r < nr;
    pack2(q, ok ? u + ou(r) : (const R *)nullptr, ok ? wki1 + ou(r) : (const R *)nullptr, n, nz2);
    pack2(q + n, ok ? wki3 + ou(r) : (const R *)nullptr, (ok && vrow(r)) ? wki2 + ov(r) : (const R *)nullptr, n,
          nz2);
    if (SPLIT && threadIdx.x == 0) q[0].y = R(0); // the mean of u_x (slot 0 was packed by this thread)
  }
  __syncthreads();
  const bool inA = fft_multi<+1>(Q0, Q1, n, 2 * rb, plan, lo, hi);
  C *Rb = inA ? Q0 : Q1, *F = inA ? Q1 : Q0;
  const C *UY = nullptr;
  if (vrow(0)) {
    for (int r = 0; r < rb; ++r) {
      pack2(F + n * r, (r < nr && vrow(r)) ? uy + ou(r) : (const R *)nullptr, (const R *)nullptr, n, nz2);
      if (SPLIT && threadIdx.x == 0) F[n * r].x = R(0); // the mean of uy (slot 0 was packed by this thread)
    }
    __syncthreads();
    UY = fft_multi<+1>(F, F + n * rb, n, rb, plan, lo, hi) ? F : F + n * rb;
  }
  for (int r = 0; r < nr; ++r) {
    double um = 0.0, wm = 0.0, vm = 0.0;
    const bool vr = vrow(r);
    C *q = Rb + 2 * n * r;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      C a = q[k], b2 = q[n + k];
      R uu = a.x, ux = a.y, wx = b2.x, vx = b2.y;
      um = fmax(um, (double)fabs(uu));
      wm = fmax(wm, (double)fabs(wx));
      q[k] = mkc<C>(ux * ux, uu * wx);
      if (vr) {
        vm = fmax(vm, (double)fabs(vx));
        q[n + k] = mkc<C>(vx * UY[n * r + k].x, R(0));
      }
    }
    // rows numbered (i-1)(ny+1) + (j-1), as rowmax_kernel reads them
    block_rowmax3(um, wm, vr ? vm : 0.0, rowmax, (size_t)(ny + 1) * nx, (size_t)(i - 1) * (ny + 1) + (j0 + r - 1));
  }
  const C *O = fft_multi<-1>(Rb, F, n, 2 * rb, plan, lo, hi) ? Rb : F;
  for (int r = 0; r < nr; ++r) {
    const C *o = O + 2 * n * r;
    if (SPLIT) {
      const R Ux = sU[0][r], Uy = sU[1][r];
      unpack2(o, wki1 + ou(r), wki3 + ou(r), n, nz2, (const R *)(wki1 + ou(r)), Ux, R(2) * Ux);
      if (vrow(r)) unpack2(o + n, wki2 + ov(r), (R *)nullptr, n, nz2, (const R *)(wki2 + ov(r)), Uy, Uy);
    } else {
      unpack2(o, wki1 + ou(r), wki3 + ou(r), n, nz2);
      if (vrow(r)) unpack2(o + n, wki2 + ov(r), (R *)nullptr, n, nz2);
    }
  }
}

// Plane loop 2 of rhsp (rhsp.F90:367-415) for rb consecutive rows (j, i), j <= ny:
// vw = r2c(v * w_v) -> ya (w interpolated to v rows is read from ya),
// vv = r2c(v_u^2) -> yb (v interpolated to u rows is read from yb),
// ww = r2c(w^2) -> yc.
template <typename R>
__global__ void __launch_bounds__(TPB)
products2_kernel(const R *__restrict__ resv, const R *__restrict__ resw, R *ya, R *yb, R *yc, FFTPlan plan, int ny,
                 int nz1, int nz2, int rb) {
  using C = typename CT<R>::c;
  extern __shared__ double2 sh_raw[];
  C *sh = reinterpret_cast<C *>(sh_raw);
  const int n = plan.n;
  C *lo = sh, *hi = sh + 64, *Q0 = sh + TW_SLOTS, *Q1 = Q0 + 2 * n * rb;
  const int j0 = blockIdx.x * rb + 1, i = blockIdx.y + 2;
  const int nr = min(rb, ny + 1 - j0);
  auto ou = [&](int r) { return ((size_t)(i - 1) * (ny + 1) + (j0 + r - 1)) * nz1; };
  auto ov = [&](int r) { return ((size_t)(i - 1) * ny + (j0 + r - 1)) * nz1; };
  fill_twiddles(lo, hi, n);
  for (int r = 0; r < rb; ++r) {
    C *q = Q0 + 2 * n * r;
    const bool ok = r < nr;
    pack2(q, ok ? resv + ov(r) : (const R *)nullptr, ok ? ya + ou(r) : (const R *)nullptr, n, nz2);
    pack2(q + n, ok ? yb + ou(r) : (const R *)nullptr, ok ? resw + ou(r) : (const R *)nullptr, n, nz2);
  }
  __syncthreads();
  const bool inA = fft_multi<+1>(Q0, Q1, n, 2 * rb, plan, lo, hi);
  C *Rb = inA ? Q0 : Q1, *F = inA ? Q1 : Q0;
  for (int k = threadIdx.x; k < 2 * n * rb; k += blockDim.x) {
    const int r = k / (2 * n), q = k - r * 2 * n;
    if (q < n) {
      C a = Rb[k], b2 = Rb[k + n];
      Rb[k] = mkc<C>(a.x * a.y, b2.x * b2.x);
      Rb[k + n] = mkc<C>(b2.y * b2.y, R(0));
    }
  }
  __syncthreads();
  const C *O = fft_multi<-1>(Rb, F, n, 2 * rb, plan, lo, hi) ? Rb : F;
  for (int r = 0; r < nr; ++r) {
    const C *o = O + 2 * n * r;
    unpack2(o, ya + ou(r), yb + ou(r), n, nz2);
    unpack2(o + n, yc + ou(r), (R *)nullptr, n, nz2);
  }
}

// fourxz (ffts.f90:27-67) on independent rows, for the operator-level API.
__global__ void __launch_bounds__(TPB)
fourxz_kernel(double *fou, double *phys, int iopt, FFTPlan plan, int nz1, int nz2) {
  extern __shared__ double2 sh[];
  const int n = plan.n;
  double2 *lo = sh, *hi = sh + 64, *A = sh + TW_SLOTS, *T = A + n;
  size_t r = blockIdx.x;
  double *F = fou + r * nz1;
  double *Ph = phys + r * n;
  fill_twiddles(lo, hi, n);
  if (iopt >= 0) {
    pack2(A, F, (const double *)nullptr, n, nz2);
    __syncthreads();
    const double2 *R = fft_multi<+1>(A, T, n, 1, plan, lo, hi) ? A : T;
    for (int q = threadIdx.x; q < n; q += blockDim.x) Ph[q] = R[q].x;
  } else {
    for (int q = threadIdx.x; q < n; q += blockDim.x) A[q] = make_double2(Ph[q], 0.0);
    __syncthreads();
    const double2 *R = fft_multi<-1>(A, T, n, 1, plan, lo, hi) ? A : T;
    unpack2(R, F, (double *)nullptr, n, nz2);
  }
}

static size_t zfft_shm(const Box &b, int nbuf) { return (TW_SLOTS + (size_t)nbuf * b.nz) * sizeof(double2); }

void k_fourxz(Box &b, double *fou, double *phys, int iopt, int nrows) {
  if (b.zfast) {
    kz_fourxz(b, fou, phys, iopt, nrows);
    return;
  }
  LAUNCH(b, "fourxz", nrows, TPB, zfft_shm(b, 2), fourxz_kernel, fou, phys, iopt, b.zplan, b.nz1, b.nz2);
}

// ---------------------------------------------------------------------------
// Poisson x transforms.  FFTW's REDFT10 / REDFT01 of length N = nx-1
// (ffts.f90:124-181) by Makhoul's reordering into one complex DFT of length
// N per pair of real lines, evaluated with Bluestein's chirp-z algorithm as
// a circular convolution of length M >= 2N-1 (M with factors 2, 3, 5, 7).
// Lines are x-contiguous pencils: pencil l (l = (j-1)*nz1 + kr, rows
// j = 1..ny-1) holds planes 2..nx at P[l*N + x], x = i-2.
// ---------------------------------------------------------------------------
struct DctArgs {
  int N, M, ld; // points per line, Bluestein length, line stride
  FFTPlan fp;
  const double2 *chirp; // exp(-i pi n^2 / N), n < N
  const double2 *bf;    // FFT_M of the conjugate chirp, / M
  const double2 *tw4;   // exp(-i pi k / (2N)), k < N
};

// out of place allowed; in place when src == dst.  iopt < 0: DCT-II x 1/N,
// iopt > 0: DCT-III x 0.5.
__global__ void __launch_bounds__(TPB) dct_rows_kernel(const double *src, double *dst, int Lp, int iopt, DctArgs d) {
  extern __shared__ double2 sh[];
  const int N = d.N, M = d.M;
  double2 *lo = sh, *hi = sh + 64, *A = sh + TW_SLOTS, *B = A + M;
  fill_twiddles(lo, hi, M);
  const int l0 = 2 * blockIdx.x, l1 = l0 + 1;
  const bool has1 = l1 < Lp;
  const double *x0 = src + (size_t)l0 * d.ld, *x1 = src + (size_t)l1 * d.ld;
  const int half = (N + 1) / 2;
  for (int k = threadIdx.x; k < M; k += blockDim.x) {
    double2 a = make_double2(0.0, 0.0);
    if (k < N) {
      double2 z;
      if (iopt < 0) { // Makhoul reordering of the two real lines
        int s = k < half ? 2 * k : 2 * (N - 1 - k) + 1;
        z = make_double2(x0[s], has1 ? x1[s] : 0.0);
      } else { // V = (X_k - i X_{N-k}) exp(+i pi k / 2N), Z = V0 + i V1, use conj(Z)
        double2 t = d.tw4[k];
        t.y = -t.y;
        double a0 = x0[k], b0 = k ? x0[N - k] : 0.0;
        double a1 = has1 ? x1[k] : 0.0, b1 = (has1 && k) ? x1[N - k] : 0.0;
        double2 v0 = cmulc(make_double2(a0, -b0), t), v1 = cmulc(make_double2(a1, -b1), t);
        z = make_double2(v0.x - v1.y, -(v0.y + v1.x));
      }
      a = cmulc(z, d.chirp[k]);
    }
    A[k] = a;
  }
  __syncthreads();
  double2 *Y = fft_multi<-1>(A, B, M, 1, d.fp, lo, hi) ? A : B;
  double2 *F = (Y == A) ? B : A;
  for (int k = threadIdx.x; k < M; k += blockDim.x) Y[k] = cmulc(Y[k], d.bf[k]);
  __syncthreads();
  Y = fft_multi<+1>(Y, F, M, 1, d.fp, lo, hi) ? Y : F;
  F = (Y == A) ? B : A;
  for (int k = threadIdx.x; k < N; k += blockDim.x) F[k] = cmulc(d.chirp[k], Y[k]);
  __syncthreads();
  double *y0 = dst + (size_t)l0 * d.ld, *y1 = dst + (size_t)l1 * d.ld;
  if (iopt < 0) {
    const double sc = 2.0 / N;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      double2 z = F[k], w = F[k ? N - k : 0];
      double2 v0 = make_double2(0.5 * (z.x + w.x), 0.5 * (z.y - w.y));
      double2 v1 = make_double2(0.5 * (z.y + w.y), -0.5 * (z.x - w.x));
      double2 t = d.tw4[k];
      y0[k] = sc * (t.x * v0.x - t.y * v0.y);
      if (has1) y1[k] = sc * (t.x * v1.x - t.y * v1.y);
    }
  } else {
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      double2 zc = F[k]; // conj of sum_k Z_k exp(+2 pi i k n / N)
      int o = k < half ? 2 * k : 2 * (N - 1 - k) + 1;
      y0[o] = 0.5 * zc.x;
      if (has1) y1[o] = -0.5 * zc.y;
    }
  }
}

// threads of the 64 x 64 tile kernels between planes and pencils (the
// divergence and the transposes back): 16 rows of the tile per pass, so a
// thread issues the loads of its 4 rows at once
static constexpr int DIV_TPB = 1024;

// Tiled transposes between the planes layout (planes 2..nx, lines
// 0..Lp-1, plane stride ps) and x-contiguous pencils P[l*ld + x], x < N.
__global__ void __launch_bounds__(TPB) planes_to_pencils_kernel(const double *F, double *P, int Lp, int N, size_t ps,
                                                                int ld) {
  __shared__ double tile[64][65];
  unsigned bx, by;
  xcd_block2(bx, by); // XCD-contiguous tiles: neighbouring rows' tiles share lines
  const int l0 = (int)bx * 64, x0 = (int)by * 64, t = threadIdx.x;
  for (int xx = t / 64; xx < 64; xx += TPB / 64) {
    int l = l0 + (t & 63), x = x0 + xx;
    if (l < Lp && x < N) tile[xx][t & 63] = F[(size_t)(x + 1) * ps + l];
  }
  __syncthreads();
  for (int ll = t / 64; ll < 64; ll += TPB / 64) {
    int l = l0 + ll, x = x0 + (t & 63);
    if (l < Lp && x < N) P[(size_t)l * ld + x] = tile[t & 63][ll];
  }
}

// With P0, the lines of mode 0 (kr = 0, 1) are taken from the FP64 pencils
// P0 (line (j-1, kr) at P0[(2 (j-1) + kr) N]) instead of P.
template <typename T>
__global__ void __launch_bounds__(DIV_TPB)
pencils_to_planes_kernel(const T *__restrict__ P, T *__restrict__ F, int Lp, int N, int ld, size_t ps,
                         const double *__restrict__ P0, int nz1) {
  constexpr int XR = DIV_TPB / 64, NQ = 64 / XR; // lines per pass, passes
  __shared__ T tile[64][65];
  unsigned bx, by;
  xcd_block2(bx, by); // XCD-contiguous tiles: neighbouring rows' tiles share lines
  const int l0 = (int)bx * 64, x0 = (int)by * 64, t = threadIdx.x;
  const int lane = t & 63, xr = t / 64;
  // the NQ lines of this thread: every load issued before the first use
  T a[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int l = l0 + xr + XR * q, x = x0 + lane;
    if (l < Lp && x < N)
      a[q] = (P0 && l % nz1 < 2) ? (T)P0[(size_t)(2 * (l / nz1) + l % nz1) * N + x] : P[(size_t)l * ld + x];
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int l = l0 + xr + XR * q, x = x0 + lane;
    if (l < Lp && x < N) tile[xr + XR * q][lane] = a[q];
  }
  __syncthreads();
  for (int xx = xr; xx < 64; xx += XR) {
    const int l = l0 + lane, x = x0 + xx;
    if (l < Lp && x < N) F[(size_t)(x + 1) * ps + l] = tile[lane][xx];
  }
}

// Dense fallback for Bluestein lengths whose FP64 buffers exceed LDS:
// out[q][l] = sum_n M[q][n] in[n][l] over planes 2..nx.
__global__ void dct_gemm_kernel(const double *__restrict__ M, const double *__restrict__ in,
                                double *__restrict__ out, int N, int L, size_t ps) {
  __shared__ double As[16][64 + 1];
  __shared__ double Bs[16][64];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int q0 = blockIdx.y * 64, l0 = blockIdx.x * 64;
  double acc[4][4] = {{0}};
  for (int n0 = 0; n0 < N; n0 += 16) {
    for (int e = threadIdx.x; e < 64 * 16; e += 256) {
      int qq = e / 16, nn = e % 16;
      int q = q0 + qq, n = n0 + nn;
      As[nn][qq] = (q < N && n < N) ? M[(size_t)q * N + n] : 0.0;
      int ll = e % 64, n2 = e / 64;
      int l = l0 + ll, nb = n0 + n2;
      Bs[n2][ll] = (l < L && nb < N) ? in[(size_t)nb * ps + l] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int nn = 0; nn < 16; ++nn) {
      double a[4], bb[4];
      for (int r = 0; r < 4; ++r) a[r] = As[nn][ty * 4 + r];
      for (int c = 0; c < 4; ++c) bb[c] = Bs[nn][tx + 16 * c];
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * bb[c];
    }
    __syncthreads();
  }
  for (int r = 0; r < 4; ++r) {
    int q = q0 + ty * 4 + r;
    if (q >= N) continue;
    for (int c = 0; c < 4; ++c) {
      int l = l0 + tx + 16 * c;
      if (l < L) out[(size_t)q * ps + l] = acc[r][c];
    }
  }
}

static DctArgs dct_args(const Box &b, int ld) {
  DctArgs d;
  d.N = b.nx - 1;
  d.ld = ld;
  d.M = b.dct_M;
  d.fp = b.dct_plan;
  d.chirp = (const double2 *)b.dct_chirp;
  d.bf = (const double2 *)b.dct_bf;
  d.tw4 = (const double2 *)b.dct_tw4;
  return d;
}

static size_t dct_shm(const Box &b) { return (TW_SLOTS + 2 * (size_t)b.dct_M) * sizeof(double2); }

// Row transforms on pencils (in place, line stride Box::pld).
static void dct_pencils(Box &b, double *P, int iopt) {
  if (b.dct_zfast) {
    kz_dct_pencils<double>(b, P, iopt, (b.ny - 1) * b.nz1, b.pld);
    return;
  }
  const int Lp = (b.ny - 1) * b.nz1;
  LAUNCH(b, iopt < 0 ? "dct2_rows" : "dct3_rows", (Lp + 1) / 2, TPB, dct_shm(b), dct_rows_kernel, P, P, Lp,
         iopt, dct_args(b, b.pld));
}

static dim3 tgrid(const Box &b) {
  const int Lp = (b.ny - 1) * b.nz1, N = b.nx - 1;
  return dim3(nblk(Lp, 64), nblk(N, 64));
}

// cosftx on nlines contiguous lines of nx-1 doubles, in place
void k_cosft_lines(Box &b, double *P, int nlines, int iopt) {
  if (b.dct_zfast) {
    kz_dct_pencils<double>(b, P, iopt, nlines, b.nx - 1);
    return;
  }
  if (!b.dct_fft) throw std::invalid_argument("cosft on lines needs an FFT-based Poisson x transform");
  LAUNCH(b, iopt < 0 ? "dct2_rows" : "dct3_rows", (nlines + 1) / 2, TPB, dct_shm(b), dct_rows_kernel, P, P, nlines,
         iopt, dct_args(b, b.nx - 1));
}

// cosftx on a planes-layout v-type field (rows 1..ny-1 of planes 2..nx).
void k_cosftx(Box &b, const double *in, double *out, int iopt) {
  const int Lp = (b.ny - 1) * b.nz1, N = b.nx - 1;
  if (b.dct_fft) {
    // FP64 pencils: the box's own on an FP64 box, else a buffer of the
    // operator's own (a mixed box's pencil buffers hold FP32)
    double *P = b.rest2;
    if (b.esz != 8) {
      if (!b.op64) {
        OTBL_HIP(hipStreamSynchronize(b.s));
        const size_t pen = std::max(b.nv, (size_t)(b.ny - 1) * b.nz1 * b.pld);
        OTBL_HIP(hipMalloc((void **)&b.op64, pen * sizeof(double) + 64));
      }
      P = b.op64;
    }
    LAUNCH(b, "to_pencils", tgrid(b), TPB, 0, planes_to_pencils_kernel, in, P, Lp, N, b.pv, b.pld);
    dct_pencils(b, P, iopt);
    LAUNCH(b, "to_planes", tgrid(b), DIV_TPB, 0, pencils_to_planes_kernel<double>, P, out, Lp, N, b.pld, b.pv,
           (const double *)nullptr, b.nz1);
  } else {
    dim3 grid(nblk(Lp, 64), nblk(N, 64));
    LAUNCH(b, "dct_gemm", grid, 256, 0, dct_gemm_kernel, iopt < 0 ? b.dct2 : b.dct3, in + b.pv, out + b.pv, N,
           Lp, b.pv);
  }
}

// ---------------------------------------------------------------------------
// rhsp pieces
// ---------------------------------------------------------------------------
// Plane-marching element-wise kernels: a thread owns element (j, kr) of a
// run of pm planes (blockIdx.y), so a value of plane i+1 it reads is its own
// plane-i value of the next iteration (read from HBM once).  Run lengths as
// measured fastest on a 1536 x 316 x 682 box (runs of 1..32 planes timed):
// 4 planes for FP32 fields, 2 for FP64.
template <typename T> static inline int prun() { return sizeof(T) == 4 ? 4 : 2; }
// the box's plane run: the autotuner's choice, else the default
template <typename T> static inline int prun(const Box &b) { return b.prun > 0 ? b.prun : prun<T>(); }
__device__ __forceinline__ bool pcol(int nz1, int R, int &kr, int &j) {
  const unsigned el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= (unsigned)nz1 * (unsigned)R) return false;
  const unsigned row = el / (unsigned)nz1;
  kr = (int)(el - row * (unsigned)nz1);
  j = (int)row + 1;
  return true;
}
static inline dim3 pgrid(const Box &b, int pm) { return dim3(nblk(b.pu), (unsigned)((b.nx + pm - 1) / pm)); }

__global__ void reset_max_kernel(DevScalars *ds, unsigned long long *vmbits, int ny) {
  for (int j = threadIdx.x; j < ny; j += blockDim.x) vmbits[j] = 0ull;
  if (threadIdx.x == 0) {
    ds->umax_bits = 0ull;
    ds->wmax_bits = 0ull;
  }
}

// energies / ministats (rhsp.F90:116,493-509, stotf.F90:395-440).  The
// reference's plane-1 sums land in the slots the plane nx/2 sums overwrite,
// so only plane nx/2 is evaluated: block c sums hy_j (2 sum_{k>=1} |a_jk|^2
// + r01^2) over the rows of component c (r01 = Re of mode 0, row 1) and
// records rows 1-2 of mode 0 and their mode sums; slots 13-15 are cleared
// afterwards (rhsp.F90:487).
template <typename T>
__global__ void __launch_bounds__(1024)
energies_kernel(const T *u, const T *v, const T *w, const double *hy, DevScalars *ds, int nx, int ny, int nz1,
                int nz2) {
  __shared__ double red[16];
  const int c = blockIdx.x;
  const T *fl = c == 0 ? u : (c == 1 ? v : w);
  const int rows = c == 1 ? ny : ny + 1;
  const T *a = fl + (size_t)(nx / 2 - 1) * rows * nz1;
  const double r01 = a[0];
  double part = 0.0;
  for (size_t e = threadIdx.x; e < (size_t)rows * nz1; e += blockDim.x) {
    const int j = (int)(e / nz1), kr = (int)(e % nz1);
    const double x = a[e];
    if (kr >= 2 && kr <= 2 * nz2 + 1) part += 2.0 * x * x * hy[j];
  }
  for (int j = threadIdx.x; j < rows; j += blockDim.x) part += r01 * r01 * hy[j];
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x / 64); ++q) t += red[q];
    double a1 = 0.0, a2 = 0.0;
    for (int k = 1; k <= nz2; ++k) {
      const double x0 = a[2 * k], x1 = a[2 * k + 1], y0 = a[nz1 + 2 * k], y1 = a[nz1 + 2 * k + 1];
      a1 += x0 * x0 + x1 * x1;
      a2 += y0 * y0 + y1 * y1;
    }
    ds->ener[1 + c] = t;
    ds->ener[4 + 4 * c] = a[0];
    ds->ener[5 + 4 * c] = a[nz1];
    ds->ener[6 + 4 * c] = 2.0 * a1;
    ds->ener[7 + 4 * c] = 2.0 * a2;
    if (c == 2) ds->ener[13] = ds->ener[14] = ds->ener[15] = 0.0;
  }
}

// rhs = d(uv)/dy + i kz (uw) on all planes (rhsp.F90:164-169); the
// d/dy part is differzy from v rows to u rows.
__global__ void rhsu_loop1_kernel(const double *__restrict__ uv, const double *__restrict__ uw,
                                  double *__restrict__ rhsu, DevCoef c, int ny, int nz1, int nplanes) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * nplanes) return;
  int kr = (int)(cidx % nz1);
  size_t ip = cidx / nz1;
  size_t ps = (size_t)nz1 * (ny + 1);
  double *o = rhsu + ip * ps + kr;
  differzy_col(uv + ip * (size_t)nz1 * ny + kr, o, nz1, 0, ny, c);
  const double kz = c.kaz[kr / 2];
  const double *w = uw + ip * ps + (kr ^ 1);
  const double sg = (kr & 1) ? 1.0 : -1.0;
  for (int j = 1; j <= ny + 1; ++j) o[(size_t)(j - 1) * nz1] += sg * w[(size_t)(j - 1) * nz1] * kz;
}

// Save the velocities (rhsp.F90:281-285) and apply the pressure-gradient
// predictor (rhsp.F90:291-339) in one pass over u-type elements.  The
// velocities u, v, w are left in place as the saved copy; the predicted
// velocities (or the copies, outside the predictor's rows and planes) are
// written to un, vn, wn, and the caller swaps the roles of the two sets.
// With FP32 fields: p0, the FP64 mean-flow pressure [nx][ny], whose
// gradients replace those of p's mode-0 real part; um, the FP64 mean-flow u
// [ny+1][nx] (mode-0 real part), updated in FP64 and saved to rm (else
// all null).
template <typename T>
__device__ __forceinline__ void save_predict_col(const T *__restrict__ u, const T *__restrict__ v,
                                                 const T *__restrict__ w, const T *__restrict__ p,
                                                 T *__restrict__ un, T *__restrict__ vn, T *__restrict__ wn,
                                                 const T *idyy, const T *kaz, const DevScalars *ds, double idxx_d,
                                                 int nx, int ny, int nz1, const double *__restrict__ p0,
                                                 const double *idyy_d, double *um, double *rm, int pm, int kr,
                                                 int j) {
  const int i0 = (int)blockIdx.y * pm + 1, i1 = min(nx, i0 + pm - 1);
  const size_t pu = (size_t)nz1 * (ny + 1), pv = (size_t)nz1 * ny;
  const T var1 = (T)ds->var1, idxx = (T)idxx_d;
  const bool mean = p0 && kr == 0, vrow = j <= ny;
  size_t e = (size_t)(i0 - 1) * pu + (size_t)(j - 1) * nz1 + kr;
  size_t ev = (size_t)(i0 - 1) * pv + (size_t)(j - 1) * nz1 + kr;
  T pc = vrow ? p[ev] : T(0); // p(i, j), carried from plane to plane
  for (int i = i0; i <= i1; ++i, e += pu, ev += pv) {
    const T pn = (vrow && i < nx) ? p[ev + pv] : T(0); // p(i+1, j)
    const T uu = u[e], ww = w[e];
    const double *pm = mean ? p0 + (size_t)(i - 1) * ny + (j - 1) : nullptr; // mean pressure at (i, j)
    const bool upred = i >= 2 && i <= nx - 1 && j >= 2 && j <= ny;
    T uo = uu;
    if (um && kr == 0) {
      const size_t q = (size_t)(j - 1) * nx + (i - 1);
      double U = um[q];
      rm[q] = U;
      if (upred) {
        U -= ds->var1 * idxx_d * (pm[ny] - pm[0]);
        um[q] = U;
        uo = (T)U;
      }
    } else if (upred) {
      if (mean)
        uo = (T)((double)uu - ds->var1 * idxx_d * (pm[ny] - pm[0]));
      else
        uo = uu - var1 * idxx * (pn - pc);
    }
    un[e] = uo;
    T wo = ww;
    if (i >= 2 && j >= 2 && j <= ny) {
      // w -= var1 * (i kz p): Re gets -(-kz Im p), Im gets -(kz Re p)
      T kz = kaz[kr / 2];
      T pp = p[ev ^ 1];
      T t = (kr & 1) ? (kz * pp) : (-kz * pp);
      wo = ww - var1 * t;
    }
    wn[e] = wo;
    if (vrow) {
      const T vv = v[ev];
      T vo = vv;
      if (i >= 2 && j >= 2 && j <= ny - 1) {
        if (mean)
          vo = (T)((double)vv - ds->var1 * idyy_d[j] * (pm[1] - pm[0]));
        else
          vo = vv - var1 * idyy[j] * (p[ev + nz1] - pc);
      }
      vn[ev] = vo;
    }
    pc = pn;
  }
}

template <typename T>
__global__ void save_predict_kernel(const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
                                    const T *__restrict__ p, T *__restrict__ un, T *__restrict__ vn,
                                    T *__restrict__ wn, const T *idyy, const T *kaz, const DevScalars *ds,
                                    double idxx_d, int nx, int ny, int nz1, const double *__restrict__ p0,
                                    const double *idyy_d, double *um, double *rm, int pm) {
  int kr, j;
  if (!pcol(nz1, ny + 1, kr, j)) return;
  save_predict_col(u, v, w, p, un, vn, wn, idyy, kaz, ds, idxx_d, nx, ny, nz1, p0, idyy_d, um, rm, pm, kr, j);
}

// The FP32 save and predictor with one thread per mode (Re, Im as float2,
// as project_pair_kernel); mode 0 (its Re part carries the FP64 mean flow)
// runs save_predict_col per part.
__global__ void save_predict_pair_kernel(const float *__restrict__ u, const float *__restrict__ v,
                                         const float *__restrict__ w, const float *__restrict__ p,
                                         float *__restrict__ un, float *__restrict__ vn, float *__restrict__ wn,
                                         const float *idyy, const float *kaz, const DevScalars *ds, double idxx_d,
                                         int nx, int ny, int nz1, const double *__restrict__ p0,
                                         const double *idyy_d, double *um, double *rm, int pm) {
  int kp, j;
  if (!pcol(nz1 / 2, ny + 1, kp, j)) return;
  if (kp == 0) {
    for (int kr = 0; kr < 2; ++kr)
      save_predict_col(u, v, w, p, un, vn, wn, idyy, kaz, ds, idxx_d, nx, ny, nz1, p0, idyy_d, um, rm, pm, kr, j);
    return;
  }
  const int i0 = (int)blockIdx.y * pm + 1, i1 = min(nx, i0 + pm - 1);
  const size_t pu2 = (size_t)nz1 * (ny + 1) / 2, pv2 = (size_t)nz1 * ny / 2, s2 = (size_t)nz1 / 2;
  const float var1 = (float)ds->var1, idxx = (float)idxx_d;
  const bool vrow = j <= ny;
  const float kz = kaz[kp];
  const float2 *u2 = reinterpret_cast<const float2 *>(u), *v2 = reinterpret_cast<const float2 *>(v);
  const float2 *w2 = reinterpret_cast<const float2 *>(w), *p2 = reinterpret_cast<const float2 *>(p);
  float2 *un2 = reinterpret_cast<float2 *>(un), *vn2 = reinterpret_cast<float2 *>(vn);
  float2 *wn2 = reinterpret_cast<float2 *>(wn);
  size_t e = (size_t)(i0 - 1) * pu2 + (size_t)(j - 1) * s2 + kp;
  size_t ev = (size_t)(i0 - 1) * pv2 + (size_t)(j - 1) * s2 + kp;
  float2 pc = vrow ? p2[ev] : make_float2(0.0f, 0.0f); // p(i, j), carried from plane to plane
  for (int i = i0; i <= i1; ++i, e += pu2, ev += pv2) {
    const float2 pn = (vrow && i < nx) ? p2[ev + pv2] : make_float2(0.0f, 0.0f); // p(i+1, j)
    const float2 uu = u2[e], ww = w2[e];
    float2 uo = uu;
    if (i >= 2 && i <= nx - 1 && j >= 2 && j <= ny) {
      uo.x = uu.x - var1 * idxx * (pn.x - pc.x);
      uo.y = uu.y - var1 * idxx * (pn.y - pc.y);
    }
    un2[e] = uo;
    float2 wo = ww;
    if (i >= 2 && j >= 2 && j <= ny) {
      // w -= var1 * (i kz p): Re gets -(-kz Im p), Im gets -(kz Re p)
      const float tx = -kz * pc.y, ty = kz * pc.x;
      wo.x = ww.x - var1 * tx;
      wo.y = ww.y - var1 * ty;
    }
    wn2[e] = wo;
    if (vrow) {
      const float2 vv = v2[ev];
      float2 vo = vv;
      if (i >= 2 && j >= 2 && j <= ny - 1) {
        const float2 pj = p2[ev + s2]; // p(i, j+1)
        vo.x = vv.x - var1 * idyy[j] * (pj.x - pc.x);
        vo.y = vv.y - var1 * idyy[j] * (pj.y - pc.y);
      }
      vn2[ev] = vo;
    }
    pc = pn;
  }
}

// dt from the CFL limits (rhsp.F90:214-262, coeft.f90:432-454) and the RK3
// constants of substep m (rhsp.F90:264-267, 479).
__global__ void dt_kernel(DevScalars *ds, const unsigned long long *vmbits, const double *dymin,
                          int ny, int m, int every_step, double dxmin, double dtret, double rex,
                          double rkc, double rkd, double rkcv, double rkdv) {
  if (threadIdx.x != 0) return;
  // setstep is set at every step start (bl_1.F90:91); a frozen dt is only
  // computed when none is known yet
  if (m == 1 && (every_step || !ds->dt_valid)) {
    const double poco = 1e-7;
    double um = __longlong_as_double((long long)ds->umax_bits);
    double wm = __longlong_as_double((long long)ds->wmax_bits);
    double dt1 = dxmin / fmax(poco, um);
    double dt2 = dxmin / fmax(poco, wm);
    double dt3 = 1e300;
    for (int j = 1; j <= ny; ++j) {
      double vm = __longlong_as_double((long long)vmbits[j - 1]);
      dt3 = fmin(dt3, dymin[j] / fmax(poco, vm));
    }
    ds->dt = fmin(fmin(dt1, dt2), fmin(dt3, dtret));
    ds->dt_valid = 1;
  }
  double dt = ds->dt;
  ds->var1 = dt * (rkcv + rkdv);
  ds->var2 = dt * rex * rkcv;
  ds->var3 = dt * rkc;
  ds->var4 = dt * rkd;
  ds->rkk = rex * dt * rkdv;
  ds->varstep = 0.5 / dt / rkdv;
  (void)m;
}

// genflu (genflu.F90:3-167, NEWGENFLU branch) in two kernels.  Every value
// genflu writes into plane 1 is a function of the saved velocities (res*,
// equal to u, v, w before the predictor) on the recycling plane xout, so:
//  * genflu_stats (one workgroup): running mean um from plane xout's mode 0,
//    utau, u_inf, the outlet momentum thickness, the rescaling gammas and
//    the interpolation tables kl / fsca / ntop of both profiles (dobleu);
//  * genflu_plane (a grid over plane 1): the copy with the outer damping
//    factor, the zero modes from u0 / v0, the profile remapping, the copy of
//    row ntop-3 above ntop-2 and the 1/gamma1 scaling, evaluated per element.
// ibuf: kl[2][ny+1], ntop[2]; dbuf: fsca[2][ny+1], ysca[ny+1], wlwa[ny+1].
template <typename T>
__global__ void genflu_stats_kernel(const T *resu, double *um, int *ibuf, double *dbuf, DevCoef c, DevScalars *ds,
                                    int ny, int nz1, int xout, double re) {
  const int s = nz1;
  const T *ux = resu + (size_t)(xout - 1) * s * (ny + 1);
  __shared__ double sh_dtc;
  if (threadIdx.x == 0) {
    double tloc = ds->tiempo - ds->timeinit;
    double timec = tloc < 20 ? 20.0 : (tloc < 100 ? 100.0 : tloc);
    ds->timec = timec;
    sh_dtc = ds->dt / timec;
  }
  __syncthreads();
  const double dtc = sh_dtc;
  // u1 mode 0 after the copy is ux mode 0 (the damping factor is 1 there)
  for (int j = threadIdx.x + 1; j <= ny + 1; j += blockDim.x)
    um[j - 1] = dtc * (double)ux[(size_t)(j - 1) * s] + um[j - 1] * (1.0 - dtc);
  __syncthreads();
  int *kl = ibuf;                  // kl[2][ny+1]
  int *ntop = ibuf + 2 * (ny + 1); // ntop[2]
  double *fsca = dbuf;             // fsca[2][ny+1]
  // the profile work runs in LDS: y, um and, per profile kk, ysca, wlwa and
  // the dobleu weights yyy ([2][ny+2] each)
  extern __shared__ double gsh[];
  const int S_ = ny + 2;
  double *y = gsh, *ums = gsh + S_, *ysca = gsh + 2 * S_, *wlwa = gsh + 4 * S_, *yyy = gsh + 6 * S_;
  __shared__ double sh_g[3], sh_www1[2];
  for (int j = threadIdx.x; j < ny + 2; j += blockDim.x) {
    y[j] = c.y[j];
    ums[j] = j <= ny ? um[j] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double *dy = c.dy;
    double utau = sqrt(fabs((ums[1] - ums[0]) / (y[2] - y[1]) / re));
    double uinf = 0.0;
    for (int j = ny + 1 - 20; j <= ny + 1 - 5; ++j) uinf += ums[j - 1];
    uinf /= 16.0;
    double u99 = (double)0.99f * uinf;
    int jtop = 0;
    for (int j = 1; j <= ny - 2; ++j) {
      if (ums[j - 1] >= u99) break;
      jtop = j + 1;
    }
    double sum = 0.0;
    for (int j = 2; j <= jtop + 3; ++j) sum += (uinf - ums[j - 1]) * ums[j - 1] * dy[j - 1];
    double rthout = sum * re / uinf;
    double g2 = ds->rthin / rthout, g1 = pow(g2, 0.125);
    ds->utauout = utau;
    ds->jtopout = jtop;
    ds->dout = y[jtop];
    ds->rthout = rthout;
    ds->gamma1 = g1;
    ds->gamma2 = g2;
    sh_g[0] = g1;
    sh_g[1] = g2;
    sh_g[2] = ds->din;
  }
  __syncthreads();
  const double g1 = sh_g[0], g2 = sh_g[1], din = sh_g[2];
  // both profiles' scaled rows
  for (int e = threadIdx.x; e < 2 * ny; e += blockDim.x) {
    const int kk = e / ny, j = e % ny + 1;
    ysca[kk * S_ + j] = (kk == 0) ? 0.5 * (y[j] + y[j + 1]) / din : y[j] / din;
  }
  __syncthreads();
  // dobleu (genflu.F90:173-199): the blending weight of every interval,
  // then its running integral and w1 per profile in row order
  const double ag = 4.0, bg = 0.2;
  const double bg2 = 1.0 - 2.0 * bg, tag = 1.0 / tanh(ag);
  for (int e = threadIdx.x; e < 2 * (ny - 1); e += blockDim.x) {
    const int kk = e / (ny - 1), q = e % (ny - 1) + 2;
    const double *ys = ysca + kk * S_;
    const double xxx = 0.5 * (ys[q] + ys[q - 1]);
    yyy[kk * S_ + q] = (xxx >= 1.0) ? 1.0 : 0.5 * (1.0 + tag * tanh(ag * (xxx - bg) / (bg2 * xxx + bg)));
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int kk = threadIdx.x;
    const double *ys = ysca + kk * S_, *yy = yyy + kk * S_;
    double *wl = wlwa + kk * S_;
    double www1 = (double)0.7895f;
    wl[1] = 0.0;
    for (int q = 2; q <= ny; ++q) {
      wl[q] = wl[q - 1] + yy[q] * (ys[q] - ys[q - 1]);
      if (ys[q] >= 1.0 && ys[q - 1] < 1.0) www1 = wl[q - 1] + (1.0 - ys[q - 1]) * yy[q];
    }
    sh_www1[kk] = www1;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * ny; e += blockDim.x) {
    const int kk = e / ny, q = e % ny + 1;
    double *wl = wlwa + kk * S_;
    wl[q] = ysca[kk * S_ + q] / g1 + (1.0 / g2 - 1.0 / g1) / sh_www1[kk] * wl[q];
    kl[kk * (ny + 1) + q] = 0;
    fsca[kk * (ny + 1) + q] = 0.0;
  }
  __syncthreads();
  // the interpolation tables.  The reference bisects the rows in order,
  // carrying the lower bound from row to row; where wl is non-decreasing
  // (the rescaled coordinate of a monotone profile) a bisection from row 1
  // finds the same interval, so the rows are searched in parallel, and in
  // order (one thread per profile) otherwise.
  __shared__ int sh_mono, sh_ntop[2];
  if (threadIdx.x == 0) {
    sh_mono = 1;
    sh_ntop[0] = sh_ntop[1] = 0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * (ny - 1); e += blockDim.x) {
    const int kk = e / (ny - 1), q = e % (ny - 1) + 2;
    if (wlwa[kk * S_ + q] < wlwa[kk * S_ + q - 1]) atomicAnd(&sh_mono, 0);
  }
  __syncthreads();
  if (sh_mono) {
    for (int e = threadIdx.x; e < 2 * ny; e += blockDim.x) {
      const int kk = e / ny, q = e % ny + 1;
      const double *ys = ysca + kk * S_;
      const double w = wlwa[kk * S_ + q];
      if (w <= ys[ny]) {
        int kll = 1, khh = ny;
        while (khh - kll > 1) {
          const int km = (khh + kll) / 2;
          if (ys[km] >= w) khh = km; else kll = km;
        }
        kl[kk * (ny + 1) + q] = kll;
        fsca[kk * (ny + 1) + q] = (w - ys[kll]) / (ys[khh] - ys[kll]);
        atomicMax(&sh_ntop[kk], q);
      }
    }
    __syncthreads();
    if (threadIdx.x < 2) ntop[threadIdx.x] = sh_ntop[threadIdx.x];
  } else if (threadIdx.x < 2) {
    const int kk = threadIdx.x;
    const double *ys = ysca + kk * S_, *wl = wlwa + kk * S_;
    int kll = 1, khh;
    ntop[kk] = 0;
    for (int q = 1; q <= ny; ++q) {
      if (wl[q] <= ys[ny]) {
        ntop[kk] = q;
        khh = ny;
        while (khh - kll > 1) {
          int km = (khh + kll) / 2;
          if (ys[km] >= wl[q]) khh = km; else kll = km;
        }
        kl[kk * (ny + 1) + q] = kll;
        fsca[kk * (ny + 1) + q] = (wl[q] - ys[kll]) / (ys[khh] - ys[kll]);
      }
    }
  }
}

// info_per_step (genflu.F90:247-295): one record per substep into the
// diagnostics ring, in the column order of the reference's unit-62 line:
// tiempo, rthout, utauout, gamma(1:2), dout, jtop, ntop(1:2) of the recycling
// statistics (genflu_stats of this substep), then per sample plane x_plane,
// u_inf, u_tau, d_out, r_thout (8 values each).  The reference samples the
// first plane of ranks ((nummpi-1) i)/8, i = 1..8; with one plane per rank
// that is plane ((nx-1) i)/8 + 1, the planes sampled here.  The profile is
// mode 0 (real part) of u saved before the predictor (the reference's ut at
// the call inside genflu, planes >= 2), or, with FP32 fields, the FP64
// mean-flow copy rm ([ny+1][nx]).  One wavefront per sample plane: its lanes
// stage the profile in LDS, lane 0 forms the sums in the reference's order.
template <typename T>
__global__ void __launch_bounds__(512)
diag_kernel(const T *us, const double *rm, const int *ntop, const DevScalars *ds, DevCoef c, double *ring,
            unsigned long long *count, int cap, int nx, int ny, int nz1, double re) {
  extern __shared__ double prof[]; // [8][ny+1]
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int ip = ((nx - 1) * (w + 1)) / 8 + 1; // 1-based sample plane
  double *um = prof + (size_t)w * (ny + 1);
  for (int j = lane; j <= ny; j += 64)
    um[j] = rm ? rm[(size_t)j * nx + (ip - 1)] : (double)us[((size_t)(ip - 1) * (ny + 1) + j) * nz1];
  __syncthreads();
  const unsigned long long n = *count;
  double *rec = ring + (size_t)(n % (unsigned long long)cap) * OTBL_DIAG_LEN;
  if (lane == 0) {
    const double *y = c.y, *dy = c.dy;
    // um(j) of the reference is um[j-1]
    double uinf = 0.0;
    for (int j = ny + 1 - 20; j <= ny + 1 - 5; ++j) uinf += um[j - 1];
    uinf /= 16.0;
    const double u99 = (double)0.99f * uinf; // default-real literal
    const double utau = sqrt(fabs((um[1] - um[0]) / (y[2] - y[1]) / re));
    int jtop = 0;
    for (int j = 1; j <= ny - 2; ++j) {
      if (um[j - 1] >= u99) break;
      jtop = j + 1;
    }
    double sum = 0.0;
    for (int j = 2; j <= jtop + 3; ++j) sum += (uinf - um[j - 1]) * um[j - 1] * dy[j - 1];
    rec[9 + w] = ip;
    rec[17 + w] = uinf;
    rec[25 + w] = utau;
    rec[33 + w] = y[jtop];
    rec[41 + w] = sum * re / uinf;
    if (w == 0) {
      rec[0] = ds->tiempo;
      rec[1] = ds->rthout;
      rec[2] = ds->utauout;
      rec[3] = ds->gamma1;
      rec[4] = ds->gamma2;
      rec[5] = ds->dout;
      rec[6] = ds->jtopout;
      rec[7] = ntop[0];
      rec[8] = ntop[1];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *count = n + 1;
}

// Plane 1 of u (blockIdx.y 0), w (1) and v (2).  For modes k >= 1 the value
// of row i is the remapped profile (genflu.F90:203-240, u/w rows shifted by
// one against v rows, row 1 of u/w = -row 2) with rows >= ntop-2 copied from
// row ntop-3, times 1/gamma1; with ntop < 4 the copied plane stays (the
// reference's copy of row ntop-3 would read out of bounds).
template <typename T>
__global__ void genflu_plane_kernel(T *u, T *v, T *w, const T *resu, const T *resv, const T *resw, const double *u0,
                                    const double *v0, const int *ibuf, const double *dbuf, DevCoef c,
                                    const DevScalars *ds, int ny, int nz1, int xout, double *um, int nx1) {
  const int f = blockIdx.y;
  const int n = f == 2 ? ny : ny + 1, s = nz1;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)s * n) return;
  const int i = (int)(e / s) + 1, k = (int)(e % s);
  T *out = f == 0 ? u : (f == 1 ? w : v);
  if (k < 2) {
    double z = 0.0;
    if (k == 0 && f == 0) {
      z = u0[i - 1];
      if (um) um[(size_t)(i - 1) * nx1] = z; // plane 1 of the FP64 mean-flow u [ny+1][nx]
    }
    if (k == 0 && f == 2) z = v0[i - 1];
    out[e] = (T)z;
    return;
  }
  const T *rx = (f == 0 ? resu : (f == 1 ? resw : resv)) + (size_t)(xout - 1) * s * n;
  const double yb = c.y[119 < ny ? 119 : ny]; // y(119) clamped to the grid
  auto cval = [&](int row) { // plane 1 after the copy, 1-based row
    double r = (double)rx[(size_t)(row - 1) * s + k];
    if (f != 2) r *= 0.5 * (1 - tanh(c.y[row] - yb));
    return r;
  };
  const int *kl = ibuf + (f == 2 ? ny + 1 : 0);
  const double *fsca = dbuf + (f == 2 ? ny + 1 : 0);
  const int nt = ibuf[2 * (ny + 1) + (f == 2 ? 1 : 0)];
  if (nt < 4) {
    out[e] = (T)cval(i);
    return;
  }
  const int isrc = i >= nt - 2 ? nt - 3 : i;
  double W;
  if (f != 2) {
    const int ii = isrc == 1 ? 2 : isrc;
    const int kll = kl[ii - 1];
    const double fs = fsca[ii - 1];
    W = cval(kll + 2) * fs + cval(kll + 1) * (1 - fs);
    if (isrc == 1) W = -W;
  } else {
    const int kll = kl[isrc];
    const double fs = fsca[isrc];
    W = cval(kll + 1) * fs + cval(kll) * (1 - fs);
  }
  out[e] = (T)(W * (1.0 / ds->gamma1));
}

// z-viscous terms and the rest of rhsv/rhsw (rhsp.F90:355-415) for planes
// 2..nx-1.  Columns (kr, i).
__global__ void rhs_loop2_kernel(double *rhsu, double *rhsv, double *rhsw, const double *resu,
                                 const double *resv, const double *resw, const double *vw,
                                 const double *vv, const double *ww, double *scr, DevCoef c, int ny,
                                 int nz1, int nplanes) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * nplanes) return;
  int kr = (int)(cidx % nz1);
  size_t ip = cidx / nz1 + 1; // plane index 0-based (planes 2..nx-1)
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  const double kv = c.kvis[kr / 2], kz = c.kaz[kr / 2];
  const double sg = (kr & 1) ? 1.0 : -1.0;
  const int kp = kr ^ 1;
  double *ru = rhsu + ip * pu + kr, *rv = rhsv + ip * pv + kr, *rw = rhsw + ip * pu + kr;
  const double *su = resu + ip * pu + kr, *sv = resv + ip * pv + kr, *sw = resw + ip * pu + kr;
  for (int j = 2; j <= ny; ++j) {
    size_t o = (size_t)(j - 1) * s;
    ru[o] -= su[o] * kv;
    rw[o] -= sw[o] * kv;
    if (j <= ny - 1) rv[o] -= sv[o] * kv;
  }
  // rhsv += i kz vw (rows 1..ny), then += d(vv)/dy
  const double *vwc = vw + ip * pu;
  for (int j = 1; j <= ny; ++j) rv[(size_t)(j - 1) * s] += sg * vwc[(size_t)(j - 1) * s + kp] * kz;
  double *t = scr + ip * pu + kr;
  differzy_col(vv + ip * pu + kr, t, s, 1, ny, c);
  for (int j = 1; j <= ny; ++j) rv[(size_t)(j - 1) * s] += t[(size_t)(j - 1) * s];
  // rhsw += d(vw)/dy, then += i kz ww (rows 1..ny)
  differzy_col(vw + ip * pu + kr, t, s, 0, ny, c);
  for (int j = 1; j <= ny + 1; ++j) rw[(size_t)(j - 1) * s] += t[(size_t)(j - 1) * s];
  const double *wwc = ww + ip * pu;
  for (int j = 1; j <= ny; ++j) rw[(size_t)(j - 1) * s] += sg * wwc[(size_t)(j - 1) * s + kp] * kz;
}

// boun (boun.f90:7-56) and the right-hand-side boundary rows
// (rhsp.F90:421-445) for every plane.  Columns (kr, i).
// um: with FP32 fields, the FP64 mean-flow u [ny+1][nx], whose boundary
// rows are extrapolated in FP64 from its own interior (else null).
template <typename T>
__global__ void boun_kernel(T *u, T *v, T *w, T *rhsu, T *rhsv, T *rhsw, DevCoefT<T> c, int ny, int nz1, int nx,
                            double *um, DevCoef cd) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * nx) return;
  int kr = (int)(cidx % nz1);
  size_t ip = cidx / nz1;
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  T *vi = v + ip * pv + kr;
  vi[0] = T(0);
  vi[(size_t)(ny - 1) * s] = (kr == 0) ? c.vmagic[ip + 1] : T(0);
  const T i21 = T(-1) / M4(c.inby, 2, 1), c11 = T(-1) / M2(c.ccon, 1, 1);
  T *f[2] = {u + ip * pu + kr, w + ip * pu + kr};
  for (int q = 0; q < 2; ++q) {
    T *a = f[q];
    if (q == 0 && um && kr == 0) {
      double *U = um + ip;
      auto UR = [&](int r) -> double & { return U[(size_t)r * nx]; }; // row r+1
      UR(0) = -1.0 / M4(cd.inby, 2, 1) * (M4(cd.inby, 2, 2) * UR(1) + M4(cd.inby, 2, 3) * UR(2) + M4(cd.inby, 2, 4) * UR(3));
      UR(ny) = -1.0 / M2(cd.ccon, 1, 1) *
               (M2(cd.ccon, 1, 2) * UR(ny - 1) + M2(cd.ccon, 1, 3) * UR(ny - 2) + M2(cd.ccon, 1, 4) * UR(ny - 3));
      a[0] = (T)UR(0);
      a[(size_t)ny * s] = (T)UR(ny);
      continue;
    }
    a[0] = i21 * (M4(c.inby, 2, 2) * a[s] + M4(c.inby, 2, 3) * a[2 * s] + M4(c.inby, 2, 4) * a[3 * s]);
    a[(size_t)ny * s] = c11 * (M2(c.ccon, 1, 2) * a[(size_t)(ny - 1) * s] + M2(c.ccon, 1, 3) * a[(size_t)(ny - 2) * s] +
                               M2(c.ccon, 1, 4) * a[(size_t)(ny - 3) * s]);
  }
  // rhs: plane 1 zero, v rows 1 and ny zero, u/w row 1 extrapolated, row ny+1 zero
  T *rvi = rhsv + ip * pv + kr;
  T *rf[2] = {rhsu + ip * pu + kr, rhsw + ip * pu + kr};
  if (ip == 0) {
    for (int j = 0; j < ny; ++j) rvi[(size_t)j * s] = T(0);
    for (int q = 0; q < 2; ++q)
      for (int j = 0; j <= ny; ++j) rf[q][(size_t)j * s] = T(0);
  }
  rvi[0] = T(0);
  rvi[(size_t)(ny - 1) * s] = T(0);
  for (int q = 0; q < 2; ++q) {
    T *a = rf[q];
    a[0] = i21 * (M4(c.inby, 2, 2) * a[s] + M4(c.inby, 2, 3) * a[2 * s] + M4(c.inby, 2, 4) * a[3 * s]);
    a[(size_t)ny * s] = T(0);
  }
}

// Explicit RK3 combination with vistyy (rhsp.F90:448-475) and the implicit
// y-viscous solve (rhsp.F90:479-484) for one component, columns (kr, i) over
// planes 2..nx.  The reference then copies rhs into rhs_prev; here the two
// buffers swap roles instead (k_swap_saved).  wk: scratch of field size.
__global__ void rk_final_kernel(double *u, const double *res, const double *rhs, const double *rhsp,
                                double *wk, DevCoef c, const double *f, const DevScalars *ds, int jee,
                                int ny, int nz1, int nx) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * nx) return;
  int kr = (int)(cidx % nz1);
  size_t ip = cidx / nz1;
  const size_t s = nz1, ps = s * jee;
  const size_t off = ip * ps + kr;
  const double *rh = rhs + off, *rp = rhsp + off;
  if (ip == 0) return;
  double *uu = u + off, *t = wk + off;
  const double var2 = (ip == (size_t)(nx - 1)) ? 0.0 : ds->var2;
  const double var3 = ds->var3, var4 = ds->var4;
  vistyy_col(res + off, t, s, jee, ny, c);
  for (int j = 0; j < jee; ++j) {
    size_t o = (size_t)j * s;
    double r = rh[o];
    t[o] = uu[o] - (var3 * r + var2 * t[o] + var4 * rp[o]);
  }
  implzy_col(uu, t, s, jee, jee == ny + 1 ? c.vyui : c.vyvi, f);
}

// The FP64 mean-flow u ([ny+1][nx]) rounded into the kz = 0 real part of
// the FP32 field after its rk_final.
template <typename T>
__global__ void mean_to_field_kernel(T *u, const double *um, int nx, int jee, int nz1) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nx * jee) return;
  const int i = q % nx, j = q / nx;
  u[((size_t)i * jee + j) * nz1] = (T)um[q];
}

// ---------------------------------------------------------------------------
// outflow_correction (boun.f90:63-93), FP64, one workgroup
// ---------------------------------------------------------------------------
template <typename T>
__global__ void outflow_kernel(T *u, const T *v, T *rhsupa, DevCoef c, DevScalars *ds, int nx, int ny, int nz1,
                               double dx1, double idx, double *um) {
  __shared__ double red[3][TPB];
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  double a = 0.0, b1 = 0.0, b2 = 0.0;
  // With FP32 fields: u from the FP64 mean-flow u um [ny+1][nx]; a top-row
  // v that is the FP32 image of its boundary value (vmagic, set by boun)
  // enters as that FP64 value (the rounding of this smooth profile is
  // correlated along the boundary and would bias the flux).
  constexpr bool F32 = sizeof(T) == 4;
  for (int i = threadIdx.x + 2; i <= nx; i += blockDim.x) {
    const T vt = v[(size_t)(i - 1) * pv + (size_t)(ny - 1) * s];
    a += (F32 && vt == (T)c.vmagic[i]) ? c.vmagic[i] : (double)vt;
  }
  for (int j = threadIdx.x + 2; j <= ny; j += blockDim.x) {
    const double u1 = um ? um[(size_t)(j - 1) * nx] : (double)u[(size_t)(j - 1) * s];
    const double un = um ? um[(size_t)(j - 1) * nx + (nx - 1)] : (double)u[(size_t)(nx - 1) * pu + (size_t)(j - 1) * s];
    b1 += u1 * c.dy[j - 1];
    b2 += un * c.dy[j - 1];
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b1;
  red[2][threadIdx.x] = b2;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
    __syncthreads();
  }
  double sumax = red[0][0] * dx1;
  sumax = sumax - red[1][0];
  sumax = sumax + red[2][0];
  double alpha = -sumax / (c.y[ny] - c.y[1]);
  if (threadIdx.x == 0) ds->sumax = sumax;
  for (int j = threadIdx.x + 2; j <= ny; j += blockDim.x) {
    size_t e = (size_t)(nx - 1) * pu + (size_t)(j - 1) * s;
    if (um) {
      double &U = um[(size_t)(j - 1) * nx + (nx - 1)];
      U += alpha;
      u[e] = (T)U;
    } else {
      u[e] = (T)((double)u[e] + alpha);
    }
    rhsupa[e] = (T)((double)rhsupa[e] + alpha * idx);
  }
}

template <typename T> static void outflow_t(Box &b) {
  auto kf = &outflow_kernel<T>;
  LAUNCH(b, "outflow", 1, TPB, 0, kf, fld<T>(b.f[OTBL_U]), fld<T>(b.f[OTBL_V]), fld<T>(b.f[OTBL_RHSUPA]), b.dc,
         b.ds, b.nx, b.ny, b.nz1, b.hc.dx[1], b.hc.idx, std::is_same<T, double>::value ? nullptr : b.um64);
}

void k_outflow(Box &b) {
  if (b.precision == OTBL_MIXED)
    outflow_t<float>(b);
  else
    outflow_t<double>(b);
}

// ---------------------------------------------------------------------------
// pois (pois.f90:11-189)
// ---------------------------------------------------------------------------
// divergence (pois.f90:41-54) into rest, planes 2..nx, rows 1..ny-1
__global__ void div_kernel(const double *u, const double *v, const double *w, double *rest, DevCoef c,
                           double idx, int nx, int ny, int nz1) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  size_t per = (size_t)(ny - 1) * s;
  if (e >= per * (nx - 1)) return;
  int i = (int)(e / per) + 2;
  size_t r = e % per;
  int j = (int)(r / s) + 1;
  int kr = (int)(r % s);
  double kz = c.kaz[kr / 2];
  size_t ow = (size_t)(i - 1) * pu + (size_t)j * s;       // row j+1
  size_t ov = (size_t)(i - 1) * pv + (size_t)(j - 1) * s; // row j
  double wp = w[ow + (kr ^ 1)];
  double t = (kr & 1) ? wp * kz : -wp * kz;
  double d = t + (v[ov + s + kr] - v[ov + kr]) * c.idy[j];
  d = d + idx * (u[ow + kr] - u[ow - pu + kr]);
  rest[ov + kr] = d;
}

// per-mode y tridiagonal (pois.f90:102-117, soltriy pois.f90:211-239) on
// the DCT coefficients in rr (planes 2..nx), columns (kr, i); a2inv: scratch
__global__ void poisson_y_kernel(double *rr, double *a2s, DevCoef c, int nx, int ny, int nz1) {
  size_t cidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= (size_t)nz1 * (nx - 1)) return;
  int kr = (int)(cidx % nz1);
  int i = (int)(cidx / nz1) + 2;
  const size_t s = nz1, pv = s * ny;
  double *col = rr + (size_t)(i - 1) * pv + kr;
  double *d2 = a2s + (size_t)(i - 1) * pv + kr;
  const int ny1 = ny - 1;
  if (kr == 1) { // the zero mode is real (pois.f90:102)
    for (int j = 0; j < ny1; ++j) col[(size_t)j * s] = 0.0;
    return;
  }
  const double kz2 = c.kaz2[kr / 2], km = c.kmod[i - 1];
  const bool pin = (i == 2 && kr == 0);
  if (pin) col[(size_t)(ny1 - 1) * s] = 0.0;
  // LU as soltriy does it, forward elimination fused
  double a2prev = C3(c.ayp, 2, 1) - kz2 - km;
  if (pin && ny1 == 1) a2prev = 2.0 * C3(c.ayp, 2, 1);
  double a3prev = C3(c.ayp, 3, 1);
  double y = col[0];
  for (int j = 2; j <= ny1; ++j) {
    double d = 1.0 / a2prev;
    d2[(size_t)(j - 2) * s] = d;
    double a1 = -C3(c.ayp, 1, j) * d;
    double a2 = (pin && j == ny1) ? 2.0 * C3(c.ayp, 2, j) : C3(c.ayp, 2, j) - kz2 - km;
    a2 = a2 + a1 * a3prev;
    y = col[(size_t)(j - 1) * s] + a1 * y;
    col[(size_t)(j - 1) * s] = y;
    a2prev = a2;
    a3prev = C3(c.ayp, 3, j);
  }
  double x = y * (1.0 / a2prev);
  col[(size_t)(ny1 - 1) * s] = x;
  for (int j = ny1 - 1; j >= 1; --j) {
    x = (col[(size_t)(j - 1) * s] - C3(c.ayp, 3, j) * x) * d2[(size_t)(j - 1) * s];
    col[(size_t)(j - 1) * s] = x;
  }
}

// divergence (pois.f90:41-54) written straight into x-contiguous pencils
// (line stride ld) through an LDS tile (64 lines x 64 planes)
template <typename T>
__global__ void __launch_bounds__(DIV_TPB)
div_pencils_kernel(const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w, T *__restrict__ P,
                   int ld, DevCoefT<T> c, double idx_d, int nx, int ny, int nz1, double *__restrict__ P0, DevCoef cd,
                   const double *__restrict__ um) {
  constexpr int XR = DIV_TPB / 64, NQ = 64 / XR; // planes per pass, passes
  __shared__ T tile[64][65];
  const T idx = (T)idx_d;
  const int Lp = (ny - 1) * nz1, N = nx - 1;
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  unsigned bx, by;
  xcd_block2(bx, by); // XCD-contiguous tiles: neighbouring rows' tiles share lines
  const int l0 = (int)bx * 64, x0 = (int)by * 64, t = threadIdx.x;
  const int lane = t & 63, xr = t / 64;
  const int l = l0 + lane;
  const bool lok = l < Lp;
  const int j = lok ? l / nz1 + 1 : 1, kr = lok ? l % nz1 : 0;
  const T kz = c.kaz[kr / 2], iy = c.idy[j];
  // the NQ planes of this thread: every load issued before the first use
  T wp[NQ], v1[NQ], v0[NQ], u1[NQ], u0[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int x = x0 + xr + XR * q;
    if (lok && x < N) {
      const int i = x + 2;
      const size_t ow = (size_t)(i - 1) * pu + (size_t)j * s;       // row j+1
      const size_t ov = (size_t)(i - 1) * pv + (size_t)(j - 1) * s; // row j
      wp[q] = w[ow + (kr ^ 1)];
      v1[q] = v[ov + s + kr];
      v0[q] = v[ov + kr];
      u1[q] = u[ow + kr];
      u0[q] = u[ow - pu + kr];
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int x = x0 + xr + XR * q;
    if (lok && x < N) {
      const T tz = (kr & 1) ? wp[q] * kz : -wp[q] * kz;
      const T d = tz + (v1[q] - v0[q]) * iy;
      tile[xr + XR * q][lane] = d + idx * (u1[q] - u0[q]);
    }
  }
  if (P0 && lok && kr < 2) {
    // the mean-flow (kz = 0) lines of FP32 fields, formed in FP64 (with the
    // FP64 mean-flow u um); line (j-1, kr) at P0[(2 (j-1) + kr) N]
    for (int q = 0; q < NQ; ++q) {
      const int x = x0 + xr + XR * q;
      if (x >= N) break;
      const int i = x + 2;
      const size_t ow = (size_t)(i - 1) * pu + (size_t)j * s;
      const size_t ov = (size_t)(i - 1) * pv + (size_t)(j - 1) * s;
      const double wq = w[ow + (kr ^ 1)], kz0 = cd.kaz[0];
      const double d = (kr ? wq * kz0 : -wq * kz0) + ((double)v[ov + s + kr] - (double)v[ov + kr]) * cd.idy[j];
      const double du = (um && kr == 0) ? um[(size_t)j * nx + (i - 1)] - um[(size_t)j * nx + (i - 2)]
                                        : (double)u[ow + kr] - (double)u[ow - pu + kr];
      P0[(size_t)(2 * (j - 1) + kr) * N + x] = d + idx_d * du;
    }
  }
  __syncthreads();
  for (int ll = xr; ll < 64; ll += XR) {
    const int lw = l0 + ll, x = x0 + lane;
    if (lw < Lp && x < N) P[(size_t)lw * ld + x] = tile[lane][ll];
  }
}

// per-mode y tridiagonal on pencils (pois.f90:102-117, soltriy
// pois.f90:211-239): one thread per (kr, kx-mode x), line (j-1, kr) at
// P + ((j-1) nz1 + kr) ld, so rows at stride nz1 ld.
__global__ void __launch_bounds__(TPB)
poisson_y_pencil_kernel(double *__restrict__ P, double *__restrict__ S, DevCoef c, int nx, int ny, int nz1,
                        int kr_end, int ld) {
  constexpr int RB = 8; // rows whose loads are issued together
  const int N = nx - 1;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)kr_end * N) return; // columns kr < kr_end
  const int x = (int)(t % N), kr = (int)(t / N), i = x + 2;
  const size_t rs = (size_t)nz1 * ld;
  double *col = P + (size_t)kr * ld + x;
  double *d2 = S + (size_t)kr * ld + x;
  const int ny1 = ny - 1;
  if (kr == 1) { // the zero mode is real (pois.f90:102)
    for (int j = 0; j < ny1; ++j) col[(size_t)j * rs] = 0.0;
    return;
  }
  const double kz2 = c.kaz2[kr / 2], km = c.kmod[i - 1];
  const bool pin = (i == 2 && kr == 0); // its last row: phi = 0 with a doubled diagonal
  double a2prev = C3(c.ayp, 2, 1) - kz2 - km;
  double a3prev = C3(c.ayp, 3, 1);
  double y = (pin && ny1 == 1) ? 0.0 : col[0];
  if (pin && ny1 == 1) col[0] = 0.0;
  // forward elimination as soltriy, rows 2..ny-1
  auto fwd = [&](int j, double b) {
    if (pin && j == ny1) b = 0.0;
    const double d = 1.0 / a2prev;
    d2[(size_t)(j - 2) * rs] = d;
    const double a1 = -C3(c.ayp, 1, j) * d;
    double a2 = (pin && j == ny1) ? 2.0 * C3(c.ayp, 2, j) : C3(c.ayp, 2, j) - kz2 - km;
    a2 = a2 + a1 * a3prev;
    y = b + a1 * y;
    col[(size_t)(j - 1) * rs] = y;
    a2prev = a2;
    a3prev = C3(c.ayp, 3, j);
  };
  int j = 2;
  for (; j + RB - 1 <= ny1; j += RB) {
    double bv[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) bv[q] = col[(size_t)(j + q - 1) * rs];
#pragma unroll
    for (int q = 0; q < RB; ++q) fwd(j + q, bv[q]);
  }
  for (; j <= ny1; ++j) fwd(j, col[(size_t)(j - 1) * rs]);
  double x_ = y * (1.0 / a2prev);
  col[(size_t)(ny1 - 1) * rs] = x_;
  // back substitution, rows ny-2..1
  auto bwd = [&](int j, double yv, double d) {
    x_ = (yv - C3(c.ayp, 3, j) * x_) * d;
    col[(size_t)(j - 1) * rs] = x_;
  };
  j = ny1 - 1;
  for (; j - RB + 1 >= 1; j -= RB) {
    double yv[RB], dv[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      yv[q] = col[(size_t)(j - q - 1) * rs];
      dv[q] = d2[(size_t)(j - q - 1) * rs];
    }
#pragma unroll
    for (int q = 0; q < RB; ++q) bwd(j - q, yv[q], dv[q]);
  }
  for (; j >= 1; --j) bwd(j, col[(size_t)(j - 1) * rs], d2[(size_t)(j - 1) * rs]);
}

// projection (pois.f90:144-162): p, w, v, u from phi = rest
// With FP32 fields, P0 holds phi of the mean-flow lines in FP64 (line
// (j-1, kr) at P0[(2 (j-1) + kr)(nx-1)], planes 2..nx), p0 the FP64
// mean-flow pressure and um the FP64 mean-flow u: their differences are
// taken in FP64 (else all null).
template <typename T>
__device__ __forceinline__ void project_col(T *u, T *v, T *w, T *p, const T *rest, size_t phs, DevCoefT<T> c,
                                            const DevScalars *ds, double vstep, int use_m, double idxx_d, int nx,
                                            int ny, int nz1, const double *__restrict__ P0, double *p0,
                                            const double *idyy_d, double *um, int pm, int kr, int j) {
  const int i0 = max(2, (int)blockIdx.y * pm + 1), i1 = min(nx, (int)blockIdx.y * pm + pm);
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  const double vs = use_m ? ds->varstep : vstep;
  size_t e = (size_t)(i0 - 1) * pu + (size_t)(j - 1) * s + kr;
  size_t ov = (size_t)(i0 - 1) * pv + (size_t)(j - 1) * s + kr;
  if (P0 && kr == 0) {
    // phi(row r, plane q) = P0[2 (r-1) N + q-2]; kaz(0) = 0 leaves w
    const int N = nx - 1;
    const size_t rN = 2 * (size_t)N;
    for (int i = i0; i <= i1; ++i, e += pu, ov += pv) {
      const double phm = P0[(size_t)(j - 2) * rN + (i - 2)];
      double &pm = p0[(size_t)(i - 1) * ny + (j - 1)];
      pm = vs * phm + pm;
      p[ov] = (T)pm;
      if (j <= ny - 1) v[ov] = (T)((double)v[ov] - idyy_d[j] * (P0[(size_t)(j - 1) * rN + (i - 2)] - phm));
      if (i <= nx - 1) {
        const double du = idxx_d * (P0[(size_t)(j - 2) * rN + (i - 1)] - phm);
        if (um) {
          double &U = um[(size_t)(j - 1) * nx + (i - 1)];
          U -= du;
          u[e] = (T)U;
        } else {
          u[e] = (T)((double)u[e] - du);
        }
      }
    }
    return;
  }
  const T varstep = (T)vs, idxx = (T)idxx_d;
  const T kz = c.kaz[kr / 2];
  const T iy = j <= ny - 1 ? c.idyy[j] : T(0);
  const size_t o = (size_t)(j - 2) * s + kr; // phi row j-1 (v-type row j-2)
  const T *ph = rest + (size_t)(i0 - 1) * phs;
  T phc = i0 <= i1 ? ph[o] : T(0); // phi(i, row j-1), carried from plane to plane
  for (int i = i0; i <= i1; ++i, e += pu, ov += pv, ph += phs) {
    const T phn = i <= nx - 1 ? ph[phs + o] : T(0); // phi(i+1, row j-1)
    p[ov] = varstep * phc + p[ov];
    const T pp = ph[o - kr + (kr ^ 1)];
    const T t = (kr & 1) ? kz * pp : -kz * pp;
    w[e] -= t;
    if (j <= ny - 1) v[ov] -= iy * (ph[o + s] - phc);
    if (i <= nx - 1) u[e] -= idxx * (phn - phc);
    phc = phn;
  }
}

template <typename T>
__global__ void project_kernel(T *u, T *v, T *w, T *p, const T *rest, size_t phs, DevCoefT<T> c, const DevScalars *ds,
                               double vstep, int use_m, double idxx_d, int nx, int ny, int nz1,
                               const double *__restrict__ P0, double *p0, const double *idyy_d, double *um,
                               int pm) {
  int kr, j;
  // rows 2..ny of planes 2..nx change
  if (!pcol(nz1, ny + 1, kr, j) || j < 2 || j > ny) return;
  project_col(u, v, w, p, rest, phs, c, ds, vstep, use_m, idxx_d, nx, ny, nz1, P0, p0, idyy_d, um, pm, kr, j);
}

// The FP32 projection with one thread per mode (Re and Im, kr = 2 kz and
// 2 kz + 1 adjacent in memory): 8-byte accesses, and the Im (Re) part of
// phi that the w update of the Re (Im) part needs is the thread's own.
// Mode 0 (its Re part takes the FP64 mean-flow lines) runs project_col per
// part.  nz1 is even, so every pair is 8-byte aligned.
__global__ void project_pair_kernel(float *u, float *v, float *w, float *p, const float *rest, size_t phs,
                                    DevCoefT<float> c,
                                    const DevScalars *ds, double vstep, int use_m, double idxx_d, int nx, int ny,
                                    int nz1, const double *__restrict__ P0, double *p0, const double *idyy_d,
                                    double *um, int pm) {
  int kp, j;
  if (!pcol(nz1 / 2, ny + 1, kp, j) || j < 2 || j > ny) return;
  if (kp == 0) {
    for (int kr = 0; kr < 2; ++kr)
      project_col(u, v, w, p, rest, phs, c, ds, vstep, use_m, idxx_d, nx, ny, nz1, P0, p0, idyy_d, um, pm, kr, j);
    return;
  }
  const int i0 = max(2, (int)blockIdx.y * pm + 1), i1 = min(nx, (int)blockIdx.y * pm + pm);
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  const float varstep = (float)(use_m ? ds->varstep : vstep), idxx = (float)idxx_d;
  const float kz = c.kaz[kp];
  const float iy = j <= ny - 1 ? c.idyy[j] : 0.0f;
  float2 *u2 = reinterpret_cast<float2 *>(u), *v2 = reinterpret_cast<float2 *>(v);
  float2 *w2 = reinterpret_cast<float2 *>(w), *p2 = reinterpret_cast<float2 *>(p);
  const float2 *ph = reinterpret_cast<const float2 *>(rest + (size_t)(i0 - 1) * phs);
  // float2 indices: element offsets / 2
  size_t e = ((size_t)(i0 - 1) * pu + (size_t)(j - 1) * s) / 2 + kp;
  size_t ov = ((size_t)(i0 - 1) * pv + (size_t)(j - 1) * s) / 2 + kp;
  const size_t o = (size_t)(j - 2) * s / 2 + kp, pv2 = pv / 2, pu2 = pu / 2, s2 = s / 2, ph2 = phs / 2;
  float2 phc = i0 <= i1 ? ph[o] : make_float2(0.0f, 0.0f);
  for (int i = i0; i <= i1; ++i, e += pu2, ov += pv2, ph += ph2) {
    const float2 phn = i <= nx - 1 ? ph[ph2 + o] : make_float2(0.0f, 0.0f);
    float2 pp = p2[ov];
    pp.x = varstep * phc.x + pp.x;
    pp.y = varstep * phc.y + pp.y;
    p2[ov] = pp;
    // w -= i kz phi: Re gets +kz Im phi, Im gets -kz Re phi
    float2 ww = w2[e];
    ww.x += kz * phc.y;
    ww.y -= kz * phc.x;
    w2[e] = ww;
    if (j <= ny - 1) {
      const float2 pn = ph[o + s2];
      float2 vv = v2[ov];
      vv.x -= iy * (pn.x - phc.x);
      vv.y -= iy * (pn.y - phc.y);
      v2[ov] = vv;
    }
    if (i <= nx - 1) {
      float2 uu = u2[e];
      uu.x -= idxx * (phn.x - phc.x);
      uu.y -= idxx * (phn.y - phc.y);
      u2[e] = uu;
    }
    phc = phn;
  }
}

// The save and pressure-gradient predictor launch (rhsp.F90:281-339): u, v,
// w, p read, the predicted velocities written to un, vn, wn; in the mixed
// substep the FP64 mean-flow u um is saved to rm and predicted in place.
template <typename T>
static void save_predict_launch(Box &b, const T *u, const T *v, const T *w, const T *p, T *un, T *vn, T *wn,
                                double *um, double *rm, int pm) {
  const DevCoefT<T> &c = coefs<T>(b);
  constexpr bool F64 = sizeof(T) == 8;
  if constexpr (!F64) {
    const dim3 pg(nblk(b.pu / 2), (unsigned)((b.nx + pm - 1) / pm)); // one thread per mode (Re, Im)
    LAUNCH(b, "save_predict", pg, TPB, 0, save_predict_pair_kernel, u, v, w, p, un, vn, wn, c.idyy, c.kaz, b.ds,
           b.hc.idxx, b.nx, b.ny, b.nz1, (const double *)b.p0, b.dc.idyy, um, rm, pm);
  } else {
    LAUNCH(b, "save_predict", pgrid(b, pm), TPB, 0, save_predict_kernel<T>, u, v, w, p, un, vn, wn, c.idyy, c.kaz,
           b.ds, b.hc.idxx, b.nx, b.ny, b.nz1, (const double *)nullptr, b.dc.idyy, (double *)nullptr,
           (double *)nullptr, pm);
  }
}

// The projection launch of pois (pois.f90:144-181): phi (planes layout,
// plane stride phs) applied to u, v, w and p in place; in the mixed substep
// the FP64 mean-flow phi pencils P0 update the FP64 mean pressure p0 and
// mean-flow u um.
static void project_launch(Box &b, double *u, double *v, double *w, double *p, const double *phi, size_t phs,
                           double varstep, bool use_m, double *p0, double *um, int pm) {
  if (b.precision == OTBL_MIXED) {
    const dim3 pg(nblk(b.pu / 2), (unsigned)((b.nx + pm - 1) / pm)); // one thread per mode (Re, Im)
    LAUNCH(b, "project", pg, TPB, 0, project_pair_kernel, fld<float>(u), fld<float>(v), fld<float>(w), fld<float>(p),
           (const float *)phi, phs, b.dcf, b.ds, varstep, use_m ? 1 : 0, b.hc.idxx, b.nx, b.ny, b.nz1,
           (const double *)b.P0, p0, b.dc.idyy, um, pm);
  } else {
    LAUNCH(b, "project", pgrid(b, pm), TPB, 0, project_kernel<double>, u, v, w, p, phi, phs, b.dc, b.ds, varstep,
           use_m ? 1 : 0, b.hc.idxx, b.nx, b.ny, b.nz1, (const double *)nullptr, (double *)nullptr, b.dc.idyy,
           (double *)nullptr, pm);
  }
}

// FP32 pressure solve of the mixed path: the kz = 0 lines run in FP64 on
// their own pencils (divergence, both x transforms and the pinned y solve),
// the other modes in FP32 with one FP64-residual refinement of the y solve.
static void pois_mixed(Box &b, double varstep, bool use_m) {
  const int nx = b.nx, ny = b.ny, nz1 = b.nz1;
  const int Lp = (ny - 1) * nz1, N = nx - 1;
  float *P = fld<float>(b.rest2), *F = fld<float>(b.rest);
  LAUNCH(b, "div", tgrid(b), DIV_TPB, 0, div_pencils_kernel<float>, fld<float>(b.f[OTBL_U]), fld<float>(b.f[OTBL_V]),
         fld<float>(b.f[OTBL_W]), P, b.pld, b.dcf, b.hc.idx, nx, ny, nz1, b.P0, b.dc, (const double *)b.um64);
  kz_dct_pencils<float>(b, P, -1, Lp, b.pld);
  kz_dct_pencils<double>(b, b.P0, -1, 2 * (ny - 1), N);
  kp_poisson_y<float, true>(b, P);
  LAUNCH(b, "poisson_y0", nblk(2 * (size_t)N), TPB, 0, poisson_y_pencil_kernel, b.P0, b.S0, b.dc, nx, ny, 2, 2, N);
  kz_dct_pencils<float>(b, P, 1, Lp, b.pld);
  kz_dct_pencils<double>(b, b.P0, 1, 2 * (ny - 1), N);
  LAUNCH(b, "to_planes", tgrid(b), DIV_TPB, 0, pencils_to_planes_kernel<float>, P, F, Lp, N, b.pld, b.phs, b.P0, nz1);
  project_launch(b, b.f[OTBL_U], b.f[OTBL_V], b.f[OTBL_W], b.f[OTBL_P], (const double *)F, b.phs, varstep, use_m, b.p0,
                 b.um64, prun<float>(b));
}

void k_pois(Box &b, int m, double varstep, bool use_m) {
  const int nx = b.nx, ny = b.ny, nz1 = b.nz1;
  const int Lp = (ny - 1) * nz1, N = nx - 1;
  (void)m;
  if (b.precision == OTBL_MIXED) {
    pois_mixed(b, varstep, use_m);
    return;
  }
  if (b.dct_fft) {
    LAUNCH(b, "div", tgrid(b), DIV_TPB, 0, div_pencils_kernel<double>, b.f[OTBL_U], b.f[OTBL_V], b.f[OTBL_W], b.rest2,
           b.pld, b.dc, b.hc.idx, nx, ny, nz1, (double *)nullptr, b.dc, (const double *)nullptr);
    dct_pencils(b, b.rest2, -1);
    if (b.pchunk) {
      // modes kz >= 1 partitioned in registers; the pinned mean-flow mode
      // (kr = 0, 1) marching, one thread per column
      kp_poisson_y<double, false>(b, b.rest2);
      LAUNCH(b, "poisson_y0", nblk(2 * (size_t)N), TPB, 0, poisson_y_pencil_kernel, b.rest2, b.rest, b.dc, nx, ny,
             nz1, 2, b.pld);
    } else {
      LAUNCH(b, "poisson_y", nblk((size_t)nz1 * N), TPB, 0, poisson_y_pencil_kernel, b.rest2, b.rest, b.dc, nx, ny,
             nz1, nz1, b.pld);
    }
    dct_pencils(b, b.rest2, 1);
    LAUNCH(b, "to_planes", tgrid(b), DIV_TPB, 0, pencils_to_planes_kernel<double>, b.rest2, b.rest, Lp, N, b.pld,
           b.phs, (const double *)nullptr, nz1);
  } else {
    size_t ndiv = (size_t)(ny - 1) * nz1 * (nx - 1);
    LAUNCH(b, "div", nblk(ndiv), TPB, 0, div_kernel, b.f[OTBL_U], b.f[OTBL_V], b.f[OTBL_W], b.rest, b.dc,
           b.hc.idx, nx, ny, nz1);
    k_cosftx(b, b.rest, b.rest2, -1);
    size_t ncol = (size_t)nz1 * (nx - 1);
    LAUNCH(b, "poisson_y", nblk(ncol), TPB, 0, poisson_y_kernel, b.rest2, b.rest, b.dc, nx, ny, nz1);
    k_cosftx(b, b.rest2, b.rest, 1);
  }
  // phi: padded planes from the pencils, the planes layout from the dense x transforms
  project_launch(b, b.f[OTBL_U], b.f[OTBL_V], b.f[OTBL_W], b.f[OTBL_P], b.rest, b.dct_fft ? b.phs : b.pv, varstep,
                 use_m, nullptr, nullptr, prun<double>(b));
}

// ---------------------------------------------------------------------------
// rhsp (rhsp.F90:22-489)
// ---------------------------------------------------------------------------
// Swap the roles of the velocity fields and their saved copies (after
// save_predict_kernel); a captured substep applies the same swap on replay.
// Role swap of a substep: the velocities with their saved copies (the
// predictor wrote the new velocities into the save buffers) and the rhs work
// fields with the previous-substep rhs fields (the rhs just formed is the
// next substep's rhs_prev, rhsp.F90:475, and the old rhs_prev buffer is
// rewritten whole by the next substep).
void k_swap_saved(Box &b) {
  std::swap(b.f[OTBL_U], b.resu);
  std::swap(b.f[OTBL_V], b.resv);
  std::swap(b.f[OTBL_W], b.resw);
  std::swap(b.f[OTBL_RHSUPA], b.rhsu);
  std::swap(b.f[OTBL_RHSVPA], b.rhsv);
  std::swap(b.f[OTBL_RHSWPA], b.rhsw);
  b.flip ^= 1;
}

// The substep in the element type T of the fields: double (OTBL_FP64, with
// the one-thread-per-line fallbacks) or float (OTBL_MIXED, fast paths only;
// statistics, CFL maxima and the recycling profiles stay FP64).
template <typename T> static void rhsp_t(Box &b, int m) {
  constexpr bool F64 = std::is_same<T, double>::value;
  const int nx = b.nx, ny = b.ny, nz1 = b.nz1;
  const int ru = ny + 1;
  T *u = fld<T>(b.f[OTBL_U]), *v = fld<T>(b.f[OTBL_V]), *w = fld<T>(b.f[OTBL_W]), *p = fld<T>(b.f[OTBL_P]);
  T *resu = fld<T>(b.resu), *resv = fld<T>(b.resv), *resw = fld<T>(b.resw);
  T *rhsu = fld<T>(b.rhsu), *rhsv = fld<T>(b.rhsv), *rhsw = fld<T>(b.rhsw);
  // the previous substep's rhs (read before the role swap below)
  const T *rp[3] = {fld<T>(b.f[OTBL_RHSUPA]), fld<T>(b.f[OTBL_RHSVPA]), fld<T>(b.f[OTBL_RHSWPA])};
  T *wki1 = fld<T>(b.wki1), *wki2 = fld<T>(b.wki2), *wki3 = fld<T>(b.wki3);
  T *ya = fld<T>(b.ya), *yb = fld<T>(b.yb), *yc = fld<T>(b.yc);
  const DevCoefT<T> &c = coefs<T>(b);
  const FFTPlan &plan = b.zplan;
  const size_t shm = zfft_shm(b, 4 * b.zrb);

  LAUNCH(b, "reset_max", 1, TPB, 0, reset_max_kernel, b.ds, b.vmbits, ny);
  // plane loop 2 beside plane loop 1 (per-kernel timing runs them in order):
  // its y interpolations (rhsp.F90:367,393) read the substep's initial v, w
  // on planes 2..nx-1, which nothing before the loop changes (the predictor
  // writes the other buffer of each pair, genflu plane 1 only); ya, yb are
  // untouched until then
  const bool side = b.s2 && !b.kprof && nx > 2;
  const bool tail = b.tail(); // rhs_loop2, boun and rk_final fused per component
  auto on_side = [&](auto &&f) {
    std::swap(b.s, b.s2);
    try {
      f();
    } catch (...) {
      std::swap(b.s, b.s2);
      throw;
    }
    std::swap(b.s, b.s2);
  };
  if (side) {
    OTBL_HIP(hipEventRecord(b.ev_fork, b.s));
    OTBL_HIP(hipStreamWaitEvent(b.s2, b.ev_fork, 0));
    on_side([&] {
      if constexpr (F64) {
        k_interpyy(b, w, ya, 0, 2, nx - 2, b.pu, b.pu);
        k_interpyy(b, v, yb, 1, 2, nx - 2, b.pv, b.pu);
      } else {
        ky_interpyy<T>(b, w, ya, 0, 2, nx - 2, b.pu, b.pu);
        ky_interpyy<T>(b, v, yb, 1, 2, nx - 2, b.pv, b.pu);
      }
    });
  }
  // x interpolation (rhsp.F90:105-107)
  if constexpr (F64) {
    k_interpxx(b, u, wki1, ru, 1);
    k_interpxx(b, v, wki2, ny, 0);
    k_interpxx(b, w, wki3, ru, 0);
  } else {
    kx_interpxx<T>(b, u, wki1, c.inxu, c.cofiux, 2, 3, b.pu);
    kx_interpxx<T>(b, v, wki2, c.inxv, c.cofivx, 1, 4, b.pv);
    kx_interpxx<T>(b, w, wki3, c.inxv, c.cofivx, 1, 4, b.pu);
  }
  if (m == 1)
    LAUNCH(b, "energies", 3, 1024, 0, energies_kernel<T>, u, v, w, b.dc.hy, b.ds, nx, ny, nz1, b.nz2);
  // u interpolated to v rows in spectral space (interpyy is linear along y and
  // commutes with the z transform of rhsp.F90:122,146)
  if constexpr (F64)
    k_interpyy(b, u, yc, 0, 1, nx, b.pu, b.pu);
  else
    ky_interpyy<T>(b, u, yc, 0, 1, nx, b.pu, b.pu);
  // plane loop 1 products
  if (b.zfast) {
    kz_products1<T>(b);
  } else {
    LAUNCH(b, "products1", dim3((ny + b.zrb) / b.zrb, nx), TPB, shm, products1_kernel<T>, u, wki1, wki2, wki3, yc,
           plan, ny, nz1, b.nz2, b.rowmax, nx, b.zrb);
    kz_rowmax(b);
  }
  if (side) OTBL_HIP(hipEventRecord(b.ev_p1, b.s)); // yc read: plane loop 2's products may overwrite it
  if (b.yL)
    ky_rhsu_loop1<T>(b, wki2, wki3, rhsu);
  else if constexpr (F64)
    LAUNCH(b, "rhsu_loop1", nblk((size_t)nz1 * nx), TPB, 0, rhsu_loop1_kernel, wki2, wki3, rhsu, b.dc, ny, nz1,
           nx);
  // x derivatives (rhsp.F90:181-204)
  if constexpr (F64) {
    k_difvisxx(b, wki1, u, rhsu, ru, 1, 1);
    k_difvisxx(b, wki2, v, rhsv, ny, 0, 0);
    k_difvisxx(b, wki3, w, rhsw, ru, 0, 0);
  } else {
    kx_difvisxx<T>(b, wki1, u, rhsu, c.dcxu, c.vixu, c.cofcxu, c.cofvxu, 1, 3, b.pu, 1);
    kx_difvisxx<T>(b, wki2, v, rhsv, c.dcxv, c.vixv, c.cofcxv, c.cofvxv, 2, 1, b.pv, 0);
    kx_difvisxx<T>(b, wki3, w, rhsw, c.dcxv, c.vixv, c.cofcxv, c.cofvxv, 2, 1, b.pu, 0);
  }
  // time step and RK constants
  LAUNCH(b, "dt", 1, 64, 0, dt_kernel, b.ds, b.vmbits, b.dc.dymin, ny, m, b.dt_mode == OTBL_DT_EVERY_STEP ? 1 : 0,
         b.hc.dxmin, b.hc.dtret, b.hc.rex, b.hc.rkc[m], b.hc.rkd[m], b.hc.rkcv[m], b.hc.rkdv[m]);
  // save velocities, pressure-gradient predictor
  double *um = F64 ? nullptr : b.um64; // the FP64 mean-flow u of the mixed substep
  save_predict_launch<T>(b, u, v, w, p, resu, resv, resw, um, b.rm64, prun<T>(b));
  // the predicted velocities now live in the former save buffers, and this
  // substep's rhs buffers become the rhs_prev fields
  k_swap_saved(b);
  std::swap(u, resu);
  std::swap(v, resv);
  std::swap(w, resw);
  // recycling inflow (rhsp.F90:287)
  {
    int *ibuf = (int *)(b.gwork + (size_t)nz1 * (ny + 1));
    double *dbuf = b.gwork + (size_t)nz1 * (ny + 1) + 4 * (ny + 2);
    LAUNCH(b, "genflu", 1, TPB, 8 * (size_t)(ny + 2) * sizeof(double), genflu_stats_kernel<T>, resu, b.um, ibuf,
           dbuf, b.dc, b.ds, ny, nz1, b.xout, b.re);
    LAUNCH(b, "genflu", dim3(nblk((size_t)nz1 * (ny + 1)), 3), TPB, 0, genflu_plane_kernel<T>, u, v, w, resu, resv,
           resw, b.u0, b.v0, ibuf, dbuf, b.dc, b.ds, ny, nz1, b.xout, um, nx);
    if (b.diag)
      LAUNCH(b, "diag", 1, 512, 8 * (size_t)(ny + 1) * sizeof(double), diag_kernel<T>, resu,
             F64 ? (const double *)nullptr : (const double *)b.rm64, ibuf + 2 * (ny + 1), b.ds, b.dc, b.diag,
             b.diag_count, b.diag_cap, nx, ny, nz1, b.re);
  }
  // plane loop 2 (planes 2..nx-1)
  if (side) {
    // its products on the side stream once plane loop 1's have read yc,
    // joined back before rhs_loop2 (which also needs this loop's rhs)
    OTBL_HIP(hipStreamWaitEvent(b.s2, b.ev_p1, 0));
    on_side([&] { kz_products2<T>(b); });
    OTBL_HIP(hipEventRecord(b.ev_join, b.s2));
    OTBL_HIP(hipStreamWaitEvent(b.s, b.ev_join, 0));
    if (!tail) ky_rhs_loop2<T>(b, rhsu, rhsv, rhsw, resu, resv, resw, ya, yb, yc);
  } else if (nx > 2) {
    if constexpr (F64) {
      k_interpyy(b, resw, ya, 0, 2, nx - 2, b.pu, b.pu);
      k_interpyy(b, resv, yb, 1, 2, nx - 2, b.pv, b.pu);
    } else {
      ky_interpyy<T>(b, resw, ya, 0, 2, nx - 2, b.pu, b.pu);
      ky_interpyy<T>(b, resv, yb, 1, 2, nx - 2, b.pv, b.pu);
    }
    if (b.zfast)
      kz_products2<T>(b);
    else
      LAUNCH(b, "products2", dim3((ny + b.zrb - 1) / b.zrb, nx - 2), TPB, shm, products2_kernel<T>, resv, resw, ya,
             yb, yc, plan, ny, nz1, b.nz2, b.zrb);
    if (tail) {
      // plane loop 2's y terms are formed in the fused tail below
    } else if (b.yL)
      ky_rhs_loop2<T>(b, rhsu, rhsv, rhsw, resu, resv, resw, ya, yb, yc);
    else if constexpr (F64)
      LAUNCH(b, "rhs_loop2", nblk((size_t)nz1 * (nx - 2)), TPB, 0, rhs_loop2_kernel, rhsu, rhsv, rhsw, resu, resv,
             resw, ya, yb, yc, wki1, b.dc, ny, nz1, nx - 2);
  }
  if (tail) {
    // plane loop 2's y terms, boun's rows and the final update per component
    // in one pass each; the implicit factors for this dt first
    launch_impl_factor(b, 0.0, 0);
    ky_rk_tail<T>(b, 1, v, rhsv, resv, rp[1], ya, yb);
    if (um) k_boun_mean(b);
    ky_rk_tail<T>(b, 3, u, rhsu, resu, rp[0], (const T *)nullptr, (const T *)nullptr);
    if constexpr (!F64) {
      // the FP64 mean-flow u (kz = 0 real part)
      ky_mean_rk_final<T>(b, um, b.rm64, rhsu, rp[0]);
      LAUNCH(b, "rk_mean", nblk((size_t)nx * (ny + 1)), TPB, 0, mean_to_field_kernel<T>, u, (const double *)um, nx,
             ny + 1, nz1);
    }
    ky_rk_tail<T>(b, 2, w, rhsw, resw, rp[2], yc, ya);
    return;
  }
  // boundary conditions (rhsp.F90:421-445)
  LAUNCH(b, "boun", nblk((size_t)nz1 * nx), TPB, 0, boun_kernel<T>, u, v, w, rhsu, rhsv, rhsw, c, ny, nz1, nx, um,
         b.dc);
  // implicit factors for this dt, then the final update per component
  launch_impl_factor(b, 0.0, 0);
  if (b.yL) {
    ky_rk_final<T>(b, v, resv, rhsv, rp[1], ny);
    if constexpr (!F64) // the FP64 mean-flow u (kz = 0 real part)
      ky_mean_rk_final<T>(b, um, b.rm64, rhsu, rp[0]);
    ky_rk_final<T>(b, u, resu, rhsu, rp[0], ny + 1);
    if (um)
      LAUNCH(b, "rk_mean", nblk((size_t)nx * (ny + 1)), TPB, 0, mean_to_field_kernel<T>, u, (const double *)um, nx,
             ny + 1, nz1);
    ky_rk_final<T>(b, w, resw, rhsw, rp[2], ny + 1);
  } else if constexpr (F64) {
    size_t ncol = (size_t)nz1 * nx;
    LAUNCH(b, "rk_final", nblk(ncol), TPB, 0, rk_final_kernel, v, resv, rhsv, rp[1], wki2, b.dc, b.implv, b.ds, ny,
           ny, nz1, nx);
    LAUNCH(b, "rk_final", nblk(ncol), TPB, 0, rk_final_kernel, u, resu, rhsu, rp[0], wki1, b.dc, b.implu, b.ds,
           ny + 1, ny, nz1, nx);
    LAUNCH(b, "rk_final", nblk(ncol), TPB, 0, rk_final_kernel, w, resw, rhsw, rp[2], wki3, b.dc, b.implu, b.ds,
           ny + 1, ny, nz1, nx);
  }
}

void k_substep_rhsp(Box &b, int m) {
  if (b.precision == OTBL_MIXED)
    rhsp_t<float>(b, m);
  else
    rhsp_t<double>(b, m);
}

void k_setup_attrs(Box &b) {
  // rows per block of the LDS products kernels
  b.zrb = std::max(1, std::min(ZROWS_MAX, 512 / b.nz));
  const int shm = (int)zfft_shm(b, 4 * b.zrb);
  OTBL_HIP(hipFuncSetAttribute((const void *)products1_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
  OTBL_HIP(hipFuncSetAttribute((const void *)products2_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
  OTBL_HIP(hipFuncSetAttribute((const void *)products1_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
  OTBL_HIP(hipFuncSetAttribute((const void *)products2_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
  OTBL_HIP(hipFuncSetAttribute((const void *)fourxz_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)zfft_shm(b, 2)));
  if (b.dct_fft && !b.dct_zfast)
    OTBL_HIP(hipFuncSetAttribute((const void *)dct_rows_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)dct_shm(b)));
}

size_t k_dct_shm_bytes(int M) { return (TW_SLOTS + 2 * (size_t)M) * sizeof(double2); }
size_t k_zfft_shm_bytes(int nz, int nbuf) { return (TW_SLOTS + (size_t)nbuf * nz) * sizeof(double2); }

__global__ void d2f_kernel(const double *a, float *o, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    o[e] = (float)a[e];
}

__global__ void f2d_kernel(const float *a, double *o, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    o[e] = (double)a[e];
}

static unsigned cgrid(size_t n) { return (unsigned)std::min<size_t>(nblk(n), 8192); }
// (a null buffer is a setup error caught here, on the host)
void k_d2f(Box &b, const double *a, float *o, size_t n) {
  if (n && (!a || !o)) throw std::logic_error("k_d2f: null buffer");
  if (n) LAUNCH(b, "d2f", cgrid(n), TPB, 0, d2f_kernel, a, o, n);
}
void k_f2d(Box &b, const float *a, double *o, size_t n) {
  if (n && (!a || !o)) throw std::logic_error("k_f2d: null buffer");
  if (n) LAUNCH(b, "f2d", cgrid(n), TPB, 0, f2d_kernel, a, o, n);
}

__global__ void end_step_kernel(DevScalars *ds) {
  if (threadIdx.x == 0) ds->tiempo += ds->dt;
}

// Re_theta(x) and Cf(x) from the kz = 0 mode of u at every plane: the
// momentum-thickness Reynolds number of genflu.F90:77-87 (the same sums as
// compute-Re_theta.f90:34-49: U_inf the mean of rows ny-19..ny-4, jtop the
// first row past which U >= 0.99 U_inf) and Cf = 2 u_tau^2 with
// u_tau^2 = |U(2) - U(1)| / ((y(2) - y(1)) Re) (genflu.F90:76).  One thread
// per plane, sums in the reference's row order.  rm: the FP64 mean-flow u
// ([ny+1][nx]) of the mixed substep, else mode 0 of the field u.
template <typename T>
__global__ void retheta_cf_kernel(const T *u, const double *rm, DevCoef c, double *out, int nx, int ny, int nz1,
                                  double re) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x; // 0-based plane
  if (i >= nx) return;
  auto U = [&](int j) -> double { // 1-based row
    return rm ? rm[(size_t)(j - 1) * nx + i] : (double)u[((size_t)i * (ny + 1) + (j - 1)) * nz1];
  };
  double uinf = 0.0;
  for (int j = ny + 1 - 20; j <= ny + 1 - 5; ++j) uinf += U(j);
  uinf /= 16.0;
  const double u99 = (double)0.99f * uinf;
  int jtop = 0;
  for (int j = 1; j <= ny - 2; ++j) {
    if (U(j) >= u99) break;
    jtop = j + 1;
  }
  double sum = 0.0;
  for (int j = 2; j <= jtop + 3; ++j) sum += (uinf - U(j)) * U(j) * c.dy[j - 1];
  out[i] = sum * re / uinf;
  out[nx + i] = 2.0 * fabs((U(2) - U(1)) / (c.y[2] - c.y[1]) / re);
}

void k_retheta_cf(Box &b, double *out) {
  if (b.precision == OTBL_MIXED)
    LAUNCH(b, "retheta_cf", nblk((size_t)b.nx), TPB, 0, retheta_cf_kernel<float>, fld<float>(b.f[OTBL_U]),
           (const double *)b.um64, b.dc, out, b.nx, b.ny, b.nz1, b.re);
  else
    LAUNCH(b, "retheta_cf", nblk((size_t)b.nx), TPB, 0, retheta_cf_kernel<double>, b.f[OTBL_U],
           (const double *)nullptr, b.dc, out, b.nx, b.ny, b.nz1, b.re);
}

void k_end_step(Box &b) { LAUNCH(b, "end_step", 1, 64, 0, end_step_kernel, b.ds); }

// ---------------------------------------------------------------------------
// Autotuner: the launch shapes chosen per box by timing candidates on the
// box's own sizes.  Every timed launch reads state fields and writes only
// work fields (dead between substeps) or scratch, so the state is unchanged.
// ---------------------------------------------------------------------------
template <typename F> static double time_launches(Box &b, int reps, F &&launch) {
  hipEvent_t e0, e1;
  OTBL_HIP(hipEventCreate(&e0));
  OTBL_HIP(hipEventCreate(&e1));
  launch(); // warm
  OTBL_HIP(hipEventRecord(e0, b.s));
  for (int r = 0; r < reps; ++r) launch();
  OTBL_HIP(hipEventRecord(e1, b.s));
  OTBL_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  OTBL_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

template <typename T> static void x_lines_once(Box &b) {
  const DevCoefT<T> &c = coefs<T>(b);
  T *u = fld<T>(b.f[OTBL_U]), *wki1 = fld<T>(b.wki1), *rhsu = fld<T>(b.rhsu);
  if constexpr (sizeof(T) == 8) {
    k_interpxx(b, u, wki1, b.ny + 1, 1);
    k_difvisxx(b, wki1, u, rhsu, b.ny + 1, 1, 0);
  } else {
    kx_interpxx<T>(b, u, wki1, c.inxu, c.cofiux, 2, 3, b.pu);
    kx_difvisxx<T>(b, wki1, u, rhsu, c.dcxu, c.vixu, c.cofcxu, c.cofvxu, 1, 3, b.pu, 0);
  }
}

template <typename T> static std::string autotune_t(Box &b, int reps) {
  char buf[256];
  std::string rep;
  // x lines: segments per line
  {
    double best = 1e30;
    int pick = b.xseg;
    for (int n = 1; n <= b.xseg_max; ++n) {
      b.xseg = n;
      const double ms = time_launches(b, reps, [&] { x_lines_once<T>(b); });
      std::snprintf(buf, sizeof buf, "xseg%d=%.4fms ", n, ms);
      rep += buf;
      if (ms < best) best = ms, pick = n;
    }
    b.xseg = pick;
  }
  // element-wise predictor and projection: planes per thread, on work fields
  // (the mean-flow arrays of the mixed substep on scratch copies)
  {
    double *scr = nullptr;
    const size_t nm = (size_t)b.nx * (b.ny + 1);
    OTBL_HIP(hipMalloc((void **)&scr, 3 * nm * sizeof(double)));
    double *um = b.um64 ? scr : nullptr, *rm = b.um64 ? scr + nm : nullptr, *p0 = b.p0 ? scr + 2 * nm : nullptr;
    T *u = fld<T>(b.f[OTBL_U]), *v = fld<T>(b.f[OTBL_V]), *w = fld<T>(b.f[OTBL_W]), *p = fld<T>(b.f[OTBL_P]);
    double best = 1e30;
    int pick = prun<T>(b);
    try {
      for (int pm : {1, 2, 4, 8}) {
        const double ms = time_launches(b, reps, [&] {
          save_predict_launch<T>(b, u, v, w, p, fld<T>(b.resu), fld<T>(b.resv), fld<T>(b.resw), um, rm, pm);
          project_launch(b, b.resu, b.resv, b.resw, b.rhsv, b.rest, b.phs, 0.0, false, p0, um, pm);
        });
        std::snprintf(buf, sizeof buf, "prun%d=%.4fms ", pm, ms);
        rep += buf;
        if (ms < best) best = ms, pick = pm;
      }
    } catch (...) {
      (void)hipFree(scr);
      throw;
    }
    OTBL_HIP(hipStreamSynchronize(b.s));
    OTBL_HIP(hipFree(scr));
    b.prun = pick;
  }
  // FP32 rk_final: half-wave or full-wave chunks (u and v rows, on work fields)
  if constexpr (sizeof(T) == 4) {
    if (b.ytabh_f) {
      double best = 1e30;
      bool pick = b.yhalf;
      for (bool half : {true, false}) {
        b.yhalf = half;
        const double ms = time_launches(b, reps, [&] {
          ky_rk_final<T>(b, fld<T>(b.wki2), fld<T>(b.resv), fld<T>(b.rhsv), fld<T>(b.f[OTBL_RHSVPA]), b.ny);
          ky_rk_final<T>(b, fld<T>(b.wki1), fld<T>(b.resu), fld<T>(b.rhsu), fld<T>(b.f[OTBL_RHSUPA]), b.ny + 1);
        });
        std::snprintf(buf, sizeof buf, "yhalf%d=%.4fms ", half ? 1 : 0, ms);
        rep += buf;
        if (ms < best) best = ms, pick = half;
      }
      b.yhalf = pick;
    }
  }
  b.tuned = true;
  return rep;
}

std::string k_autotune(Box &b, int reps) {
  return b.precision == OTBL_MIXED ? autotune_t<float>(b, reps) : autotune_t<double>(b, reps);
}

} // namespace otbl
