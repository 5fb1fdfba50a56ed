// Internal declarations of libopentbl_amd: host coefficient set, device
// coefficient pointers, the box context and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "opentbl.h"
#include "otbl_fft.h"

namespace otbl {

// (r, n) Fortran arrays are stored with column i (1-based) at offset r*i, so
// a (3, 1:n) array has 3*(n+1) doubles and element (q, i) is a[3*i + q - 1].
struct HostCoef {
  int nx = 0, ny = 0, nz = 0, nz1 = 0, nz2 = 0;
  double re = 0, ax = 0, az = 0, cfl = 0, pi = 0;
  std::vector<double> y, dy, dyi, idy, idyy, dymin, x, dx, dxi, hy;
  std::vector<double> kaz, kaz2, kvis, kmod, vmagic;
  double idx = 0, idxx = 0, dxmin = 0, dzmin = 0, dtret = 0, rex = 0;
  double rkc[4] = {0}, rkd[4] = {0}, rkcv[4] = {0}, rkdv[4] = {0};
  // tridiagonal left-hand sides (unfactored here; factorised on the GPU)
  std::vector<double> inxu, inxv, inyu, inyv, dcxu, dcxv, dcyu, dcyv;
  std::vector<double> vixu, vixv, viyu, viyv, vyui, vyvi;
  // explicit right-hand sides
  std::vector<double> cofiux, cofivx, cofiuy, cofivy, cofcxu, cofcxv, cofcyu, cofcyv;
  std::vector<double> cofvxu, cofvxv, cofvyu, cofvyv, ayp;
  double inbx[16] = {0}, inby[16] = {0}, dcbx[20] = {0}, dcby[16] = {0}, ccon[8] = {0};
  double cofvbx[16] = {0}, cofvby[16] = {0};
  // stats-only schemes (8, 1:n), LU-factored: d/dy and midpoint y
  // interpolation (u rows, p rows) at the v rows, d/dx and midpoint x
  // interpolation on the planes
  std::vector<double> fd_dvdy, fd_iy, fd_iyp, fd_dx, fd_ix;
};

// Builds every coefficient of the RK3 step (coeft.f90:10-1234).
void build_coef(HostCoef &hc, int nx, int ny, int nz, double re, double ax, double az,
                double cfl, const double *y);
// the stats-only schemes (coeft.f90:56-285); called by build_coef
void build_stats_coef(HostCoef &hc);
// coefb (coeft.f90:1360-1789): boundary-closure weights for flag flagv
// (0, 1, 3..7), n of them; throws std::invalid_argument otherwise
void coefb_weights(int n, int flagv, double a, double b, double c, double d, double bcc[5]);
// inlet_retheta (bl_1.F90:372-396) and magicnumber (boun.f90:102-143)
void inlet_setup(HostCoef &hc, const double *u0, double *rthin, double *din);

// Device coefficient pointers (same (r, n+1) layout as HostCoef), FP64 or
// an FP32 copy of the factored FP64 set.
template <typename T> struct DevCoefT {
  const T *inxu, *inxv, *inyu, *inyv, *dcxu, *dcxv, *dcyu, *dcyv;
  const T *vixu, *vixv, *viyu, *viyv, *vyui, *vyvi;
  const T *cofiux, *cofivx, *cofiuy, *cofivy, *cofcxu, *cofcxv, *cofcyu, *cofcyv;
  const T *cofvxu, *cofvxv, *cofvyu, *cofvyv, *ayp;
  const T *inbx, *inby, *dcbx, *dcby, *ccon, *cofvbx, *cofvby;
  const T *y, *dy, *idy, *idyy, *dymin, *hy, *kaz, *kaz2, *kvis, *kmod, *vmagic;
};
using DevCoef = DevCoefT<double>;

// Scalars that live on the device so a substep never waits for the host.
struct DevScalars {
  double dt, tiempo, timeinit;
  double var1, var2, var3, var4, rkk, varstep;
  double rthin, din, rthout, dout, utauout, gamma1, gamma2, timec;
  double jtopout; // top row of the outlet profile (genflu.F90:82-85)
  double sumax;
  unsigned long long umax_bits, wmax_bits; // |u|, |w_x| maxima (IEEE order)
  int dt_valid, pad;
  double ener[16];
};

struct Box {
  int nx, ny, nz, nz1, nz2, mz, xout, precision, dt_mode;
  int dev = 0; // the HIP device the box lives on (current at otbl_create)
  // bytes per field element: 8 (OTBL_FP64) or 4 (OTBL_MIXED, FP32 fields);
  // fields and work fields are allocated at that size
  int esz = 8;
  // FP64 pencils of the cosftx operator entry point on a mixed box (whose
  // own pencil buffers hold FP32), allocated on first use
  double *op64 = nullptr;
  double re, ax, az, cfl;
  size_t nu, nv;     // elements per u-type / v-type field
  size_t pu, pv;     // elements per u-type / v-type plane
  // line stride of the Poisson pencils (lines of nx-1 points): nx-1 rounded
  // up to 32 values, so every line starts on a 128-byte boundary; the
  // values past nx-1 are never read or written
  int pld = 0;
  // plane stride of the pressure correction phi the pencils are transposed
  // back into (Box::rest): pv rounded up to 32 values
  size_t phs = 0;
  hipStream_t s = nullptr;
  // plane loop 2 of rhsp on a stream of its own beside plane loop 1 (its y
  // interpolations need only the substep's initial v, w; its products wait
  // for plane loop 1's products, whose input yc they overwrite); joined
  // back into s before rhs_loop2.  s2 null (OTBL_SERIAL_LOOPS, or paths
  // that use ya / yb as scratch): everything on s.
  hipStream_t s2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_p1 = nullptr, ev_join = nullptr;
  HostCoef hc;
  DevCoef dc{};
  double *coefbuf = nullptr;
  // FP32 copies for the mixed-precision substep (made after the FP64
  // factorisation): coefficients, y-line tables, twiddles, Bluestein tables
  DevCoefT<float> dcf{};
  float *coefbuf_f = nullptr, *ytab_f = nullptr, *ztw_f = nullptr;
  // FP32 tables of the half-wave chunked y kernels of the mixed substep,
  // one per system in the order of ytab; null when those kernels are off
  // (odd row count per chunk, OTBL_Y_HALF=0)
  float *ytabh_f = nullptr;
  float *dct_chirp_f = nullptr, *dct_bf_f = nullptr, *dct_tw4_f = nullptr, *dct_ztw_f = nullptr;
  // Poisson y solves of the modes kz >= 1 partitioned in registers: rows
  // per chunk (0: one thread marches each column)
  int pchunk = 0;
  // mixed Poisson: the FP64 pencils of the mean-flow (kz = 0: kr = 0, 1)
  // lines with their scratch (2(ny-1)(nx-1) each)
  double *P0 = nullptr, *S0 = nullptr;
  // mixed: the mean-flow pressure (real part of mode 0 of p) in FP64,
  // [nx][ny]; its gradients feed the predictor and its phi the projection
  double *p0 = nullptr;
  // mixed: the mean-flow u (real part of mode 0 of u) in FP64, [ny+1][nx],
  // updated in FP64 wherever the substep updates u (the FP32 field holds its
  // rounding) and its saved copy
  double *um64 = nullptr, *rm64 = nullptr;
  // persistent state; every field and work-field pointer below holds
  // elements of esz bytes (read through fld<T>)
  double *f[OTBL_NFIELDS] = {nullptr};
  // work fields
  double *resu = nullptr, *resv = nullptr, *resw = nullptr;
  double *rhsu = nullptr, *rhsv = nullptr, *rhsw = nullptr;
  double *wki1 = nullptr, *wki2 = nullptr, *wki3 = nullptr;
  double *ya = nullptr, *yb = nullptr, *yc = nullptr; // y-interpolated / product work
  double *rest = nullptr, *rest2 = nullptr;
  // Poisson x transforms (DCT matrices, (nx-1)^2 each)
  double *dct2 = nullptr, *dct3 = nullptr;
  // implicit y LU (dt dependent), u-type and v-type: a1, a2, a3 rows
  double *implu = nullptr, *implv = nullptr;
  // x lines with the windowed backward sweep (enabled when every window
  // product of the factored x matrices is below round-off)
  bool xwin = false;
  // segments per x line (threads per line) of the windowed kernels, > 1
  // when every forward multiplier product over XLINE_LEAD planes is below
  // round-off as well
  int xseg = 1;
  // largest segment count the x systems allow (1: segments off)
  int xseg_max = 1;
  // plane run of the element-wise predictor / projection kernels (0: the
  // default for the element size)
  int prun = 0;
  // the FP32 rk_final on half-wave chunks when its tables exist
  bool yhalf = true;
  // rhs_loop2, boun and rk_final fused per component (ky_rk_tail) where the
  // y solves run on full-wave chunks (Box::tail()); OTBL_Y_NOTAIL at create
  // keeps the three passes
  bool ytail = true;
  bool tail() const { return yL && ytail && !(esz == 4 && ytabh_f && yhalf); }
  // set by otbl_autotune: the choices above were timed on this box
  bool tuned = false;
  // chunked y-line solver: rows per chunk (0: one thread marches each
  // column) and the tables of its eight systems
  int yL = 0;
  double *ytab = nullptr;
  // genflu / inflow
  double *um = nullptr, *u0 = nullptr, *v0 = nullptr, *gwork = nullptr;
  DevScalars *ds = nullptr;
  unsigned long long *vmbits = nullptr; // per-row max |v_x|
  // z transform plan (length nz); zfast: register-resident engine with the
  // twiddle table ztw (nz complex), else the LDS Stockham kernels
  FFTPlan zplan{};
  bool zfast = false;
  // FP32 z products with two rows per team (0: one row, OTBL_Z_NOPAIR; 1:
  // plane loop 1's pair kernel held to three waves per SIMD; 2: to two,
  // OTBL_ZP_OCC2)
  int zpair = 1;
  double *ztw = nullptr;
  double *zrow = nullptr; // nz1 zeros: the absent partner of a packed transform
  double *rowmax = nullptr; // per-row CFL maxima of products1, 3 (ny+1) nx
  // Poisson x transforms: Bluestein length M and tables (complex as
  // double pairs), or the dense matrices when FP64 buffers of length M do
  // not fit in LDS
  bool dct_fft = false;
  bool dct_zfast = false; // M a power of two run on the register engine
  double *dct_ztw = nullptr; // exp(2 pi i m / M)
  int dct_M = 0;
  // rows per workgroup of the LDS products kernels (short z rows batched)
  int zrb = 1;
  FFTPlan dct_plan{};
  double *dct_chirp = nullptr, *dct_bf = nullptr, *dct_tw4 = nullptr;
  // per-substep diagnostics ring (info_per_step records): diag_cap records
  // of OTBL_DIAG_LEN doubles, the device count of records written, and the
  // host's read cursor
  double *diag = nullptr;
  unsigned long long *diag_count = nullptr;
  // Re_theta(x) and Cf(x) of the last otbl_retheta_cf call (2 nx)
  double *rthcf = nullptr;
  int diag_cap = 0;
  unsigned long long diag_read = 0;
  // statistics (otbl_stats_*): the accumulators [OTBL_NSTAT][nx][ny+1], the
  // enstrophy field of the last sample [nx][ny][nz] (FP32), the stats-only
  // schemes on the device (fd_dvdy, fd_iy, fd_iyp, fd_dx, fd_ix), the
  // samples taken (ical), the step counter and the sampling rule of
  // bl_1.F90:92 (every avcrt steps after step stest), and the request for
  // a sample at the next substep 1
  double *st = nullptr, *stcoef = nullptr;
  // mixed box: FP64 mean-flow lines of the statistics (u_x, du/dx as
  // [ny+1][nx]; B1, B4, B6 as [ny][nx] in slots of [ny+1][nx])
  double *stmean = nullptr;
  float *ens = nullptr;
  int ical = 0, avcrt = 0, stest = 0;
  long long istep = 0;
  bool dostat = false;
  // graphs
  bool use_graphs = false;
  // one graph per substep m and buffer role assignment (flip: the velocity
  // fields and their saved copies swap roles at every substep); slot 0 is
  // substep 1 with a statistics sample
  hipGraphExec_t gexec[4][2] = {};
  int flip = 0;
  // bumped whenever the captured substep would change (diagnostics on or
  // off), so that graphs captured elsewhere (an ensemble's) are captured again
  unsigned graph_gen = 0;
  // measurement
  bool kprof = false;
  struct KStat { int n = 0; double ms = 0.0; };
  std::map<std::string, KStat> kstat;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

// a field buffer as elements of the precision the substep runs in
template <typename T> inline T *fld(double *p) { return reinterpret_cast<T *>(p); }
template <typename T> inline const DevCoefT<T> &coefs(const Box &b);
template <> inline const DevCoefT<double> &coefs<double>(const Box &b) { return b.dc; }
template <> inline const DevCoefT<float> &coefs<float>(const Box &b) { return b.dcf; }

// kernel launchers (otbl_kernels.hip); each enqueues on b.s
void k_factor_coef(Box &b, double *coefbuf_rw);
void k_interpxx(Box &b, const double *gen, double *geni, int R, int which_u);
void k_difvisxx(Box &b, const double *prod, const double *vel, double *out, int R,
                int which_u, int mode);
void k_interpyy(Box &b, const double *gen, double *geni, int loguv, int i0, int nplanes,
                size_t in_pstride, size_t out_pstride);
void k_differzy(Box &b, const double *geni, double *res, int which, int nplanes,
                size_t in_pstride, size_t out_pstride);
void k_vistyy(Box &b, const double *u, double *res, int R, int nplanes);
void k_implzy_op(Box &b, double *u, const double *res, int R, double rkk, int nplanes);
void k_fourxz(Box &b, double *fou, double *phys, int iopt, int nrows);
void k_cosftx(Box &b, const double *in, double *out, int iopt);
void k_cosft_lines(Box &b, double *P, int nlines, int iopt);
void k_coefinv(Box &b, double *a, int nrow);
void k_soltrix(Box &b, const double *u, double *us, const double *a, int nstr, int nlen, int nrow);
void k_soltriy(Box &b, const double *u, double *us, int ii, double *a, int ny1, int nmode);
void k_substep_rhsp(Box &b, int m);
void k_outflow(Box &b);
void k_pois(Box &b, int m, double varstep, bool use_m);
void k_end_step(Box &b);
// Re_theta(x), Cf(x) of the kz = 0 mode of u into out[0..nx), out[nx..2nx)
void k_retheta_cf(Box &b, double *out);
void k_swap_saved(Box &b);
// times the candidates of xseg, prun and yhalf on the box's work fields and
// keeps the fastest (otbl_autotune); returns a one-line report
std::string k_autotune(Box &b, int reps);
// partitioned Poisson y solves of the modes kz >= 1 on pencils
// (otbl_poisson.hip); REF: one FP64-residual refinement step
int kp_chunk_rows(int ny);
template <typename T, bool REF> void kp_poisson_y(Box &b, T *P);
// element conversions between FP64 and FP32 buffers on b.s
void k_d2f(Box &b, const double *a, float *o, size_t n);
void k_f2d(Box &b, const float *a, double *o, size_t n);
void k_setup_attrs(Box &b);
size_t k_dct_shm_bytes(int M);
// register-resident z transforms (otbl_zfft.hip), valid when Box::zfast
void kz_twiddles(Box &b, double *tw, int n);
// windowed x lines (otbl_xline.hip), valid when Box::xwin
constexpr int XLINE_WINDOW = 24; // planes of the backward window
constexpr int XLINE_EXACT = 32;  // lines up to this length are solved exactly
constexpr int XLINE_LEAD = 32;   // forward lead-in of a line segment, planes
template <typename T>
void kx_interpxx(Box &b, const T *gen, T *geni, const T *icon, const T *econ, int chi, int chi2, size_t P);
template <typename T>
void kx_difvisxx(Box &b, const T *prod, const T *vel, T *out, const T *icon, const T *iconv, const T *econ,
                 const T *econv, int chv, int ic, size_t P, int mode);
// register-resident z transforms (otbl_zfft.hip), valid when Box::zfast /
// Box::dct_zfast
// Lp lines of the Poisson x transforms on pencils (all (ny-1)*nz1 lines,
// or the 2(ny-1) mean-flow lines of the mixed path), line l at P + l ld
template <typename R> void kz_dct_pencils(Box &b, R *P, int iopt, int Lp, int ld);
template <typename R> void kz_products1(Box &b);
// the CFL maxima from the per-row maxima products1 left in Box::rowmax
void kz_rowmax(Box &b);
template <typename R> void kz_products2(Box &b);
void kz_fourxz(Box &b, double *fou, double *phys, int iopt, int nrows);
size_t k_zfft_shm_bytes(int nz, int nbuf);

// statistics (otbl_stats.hip): one sample of the current state added to
// Box::st (statsp, stotf.F90:51-291); clearing the accumulators; launch
// attributes; LDS bytes of the row kernel for length nz
void k_statsp(Box &b);
void k_stats_zero(Box &b);
void k_stats_attrs(Box &b);
size_t k_stats_shm_bytes(int nz);

// launch bookkeeping (otbl_api.cpp)
void kbegin(Box &b, const char *name);
void kend(Box &b);
[[noreturn]] void fail_hip(hipError_t e, const char *what);

} // namespace otbl

#define OTBL_HIP(call)                                                         \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) otbl::fail_hip(e_, #call);                           \
  } while (0)
