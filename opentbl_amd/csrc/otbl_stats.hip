// Statistics of the boundary layer (the reference's statsp, stotf.F90:51-291,
// as its production build compiles it: -DNOCORR -DNOSPECTRA,
// makefile-OMP-vul:11), with the x prologue of rhsp.F90:72-102.  One sample
// is taken from the current u, v, w, p and added to 40 accumulators of
// (ny+1, nx) doubles; the enstrophy field of the sample is written in FP32.
//
// Three passes over the box, on work fields that hold nothing at the start
// of a substep:
//   stats_x     x lines: u_x = midpoint interpolation of u (interpx_new)
//               moved one plane up, and d/dx (diffx_inplace) of u_x, v, w;
//               one thread per line marching over the planes
//   stats_y     y lines of every (kr, plane) column: the midpoint
//               interpolations to the v rows (interpy_new) of u_x, w, dw/dx,
//               du/dx and p, and d/dy (diffy_inplace) of (u_x)_y, w_y, v
//   stats_z     one (j, i) row per workgroup: the sums over the spanwise
//               modes, the k = 0 means, and through three packed c2r
//               transforms (LDS Stockham) the enstrophy and the triple
//               products in physical space; the row's 40 sums are added to
//               the accumulators by one thread (no atomics: the sums are
//               reproducible).
// Arithmetic is FP64; the mixed box's FP32 fields are widened on load and
// its intermediate fields are held in FP32.
#include "otbl_kcommon.h"
#include "otbl_fft.h"
#include "otbl_zfft.h"

#include <type_traits>

namespace otbl {

// element (r, i) of an (8, 1:n) scheme array
#define F8(a, r, i) (a)[8 * (size_t)(i) + (r)-1]

// ---------------------------------------------------------------------------
// x lines (lip.f90:704-766).  KIND 0: d/dx (5-point right-hand side, rows 1,
// 2, nx-1, nx their own weights, the interior the weights of row 3); KIND 1:
// midpoint interpolation (4 points; rows 1, nx-1, nx own weights, interior
// row 3's), written one plane up (u_x(i+1) = I(i), u_x(1) = I(1),
// rhsp.F90:75-80).  The forward sweep leaves y in out; the backward sweep
// reads it back.
// ---------------------------------------------------------------------------
// Line l of P starts at l * ls, plane i of a line at (i-1) * ps: the planes
// layout (ls = 1, ps = P) or the FP64 mean-flow u of a mixed box, [ny+1][nx]
// (ls = nx, ps = 1).
template <typename T, int KIND>
__global__ void __launch_bounds__(TPB)
stats_x_kernel(const T *__restrict__ in, T *__restrict__ out, const double *__restrict__ c, int nx, size_t P,
               size_t ls, size_t ps) {
  const size_t l = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (l >= P) return;
  const T *g = in + l * ls; // g[(i-1) ps] = plane i
  T *o = out + l * ls;
  auto G = [&](int i) -> double { return (double)g[(size_t)(i - 1) * ps]; };
  double y = 0.0;
  if (KIND == 0) {
    double w[5] = {G(1), G(2), G(3), G(4), G(5)}; // planes i-2..i+2 once i >= 3
    for (int i = 1; i <= nx; ++i) {
      double r = 0.0;
      if (i <= 2 || i >= nx - 1) {
        const int f = i <= 2 ? 1 : nx - 4;
#pragma unroll
        for (int q = 1; q <= 5; ++q) r += F8(c, q, i) * G(f + q - 1);
      } else {
        if (i > 3) { // slide the window to planes i-2..i+2
#pragma unroll
          for (int q = 0; q < 4; ++q) w[q] = w[q + 1];
          w[4] = G(i + 2);
        }
#pragma unroll
        for (int q = 1; q <= 5; ++q) r += F8(c, q, 3) * w[q - 1];
      }
      y = i == 1 ? r : r - F8(c, 6, i) * y;
      o[(size_t)(i - 1) * ps] = (T)y;
    }
    double x = y * F8(c, 7, nx);
    o[(size_t)(nx - 1) * ps] = (T)x;
    for (int i = nx - 1; i >= 1; --i) {
      x = ((double)o[(size_t)(i - 1) * ps] - F8(c, 8, i) * x) * F8(c, 7, i);
      o[(size_t)(i - 1) * ps] = (T)x;
    }
  } else {
    double w[4] = {G(1), G(2), G(3), G(4)}; // planes i-1..i+2 once i >= 2
    for (int i = 1; i <= nx; ++i) {
      double r = 0.0;
      if (i == 1 || i >= nx - 1) {
        const int f = i == 1 ? 1 : nx - 3;
#pragma unroll
        for (int q = 1; q <= 4; ++q) r += F8(c, q, i) * G(f + q - 1);
      } else {
        if (i > 2) {
#pragma unroll
          for (int q = 0; q < 3; ++q) w[q] = w[q + 1];
          w[3] = G(i + 2);
        }
#pragma unroll
        for (int q = 1; q <= 4; ++q) r += F8(c, q, 3) * w[q - 1];
      }
      y = i == 1 ? r : r - F8(c, 6, i) * y;
      o[(size_t)(i - 1) * ps] = (T)y;
    }
    // I(nx) falls off the end of the shifted line
    double x = y * F8(c, 7, nx);
    for (int i = nx - 1; i >= 1; --i) {
      x = ((double)o[(size_t)(i - 1) * ps] - F8(c, 8, i) * x) * F8(c, 7, i);
      o[(size_t)i * ps] = (T)x;
    }
    o[0] = (T)x;
  }
}

// ---------------------------------------------------------------------------
// y lines of one (kr, plane) column (lip.f90:666-698, 771-809).  in/out
// address row r at [(r-1) nz1]; the forward sweep leaves y in out.
// ---------------------------------------------------------------------------
// interpy_new: tipo 0 reads ny+1 rows (u, w, their x derivatives) and puts
// 0 at the wall row; tipo 1 reads the ny rows of p and keeps its
// extrapolated wall row
template <typename T>
__device__ void col_interpy(const T *__restrict__ in, T *__restrict__ out, const double *__restrict__ c, int ny,
                            int s, int tipo) {
  auto I = [&](int r) -> double { return (double)in[(size_t)(r - 1) * s]; };
  double r1 = 0.0;
  if (tipo == 1) {
#pragma unroll
    for (int q = 1; q <= 4; ++q) r1 += F8(c, q, 1) * I(1 + q);
  }
  out[0] = (T)r1;
  double w[5] = {I(2), I(3), I(4), I(5), I(6)};
  double y = 0.0;
#pragma unroll
  for (int q = 1; q <= 5; ++q) y += F8(c, q, 2) * w[q - 1]; // row 2 on rows 2..6
  out[(size_t)s] = (T)y;
  // interior rows j = 3..ny-2 on rows j-1..j+2 (window w[0..3])
  w[0] = I(2); w[1] = I(3); w[2] = I(4); w[3] = I(5);
  for (int j = 3; j <= ny - 2; ++j) {
    if (j > 3) {
      w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = I(j + 2);
    }
    double r = 0.0;
#pragma unroll
    for (int q = 1; q <= 4; ++q) r += F8(c, q, j) * w[q - 1];
    y = r - F8(c, 6, j) * y;
    out[(size_t)(j - 1) * s] = (T)y;
  }
  const int np = tipo == 1 ? 3 : 4; // rows ny-2..ny+1 (u rows) or ..ny (p)
  for (int j = ny - 1; j <= ny; ++j) {
    double r = 0.0;
    for (int q = 1; q <= np; ++q) r += F8(c, q, j) * I(ny - 3 + q);
    y = r - F8(c, 6, j) * y;
    out[(size_t)(j - 1) * s] = (T)y;
  }
  double x = y * F8(c, 7, ny);
  out[(size_t)(ny - 1) * s] = (T)x;
  for (int j = ny - 1; j >= 2; --j) {
    x = ((double)out[(size_t)(j - 1) * s] - F8(c, 8, j) * x) * F8(c, 7, j);
    out[(size_t)(j - 1) * s] = (T)x;
  }
}

// diffy_inplace on ny rows
template <typename T>
__device__ void col_diffy(const T *__restrict__ in, T *__restrict__ out, const double *__restrict__ c, int ny,
                          int s) {
  auto I = [&](int r) -> double { return (double)in[(size_t)(r - 1) * s]; };
  double w[5] = {I(1), I(2), I(3), I(4), I(5)};
  double y = 0.0;
  for (int j = 1; j <= 2; ++j) {
    double r = 0.0;
#pragma unroll
    for (int q = 1; q <= 4; ++q) r += F8(c, q, j) * w[q - 1];
    y = j == 1 ? r : r - F8(c, 6, j) * y;
    out[(size_t)(j - 1) * s] = (T)y;
  }
  for (int j = 3; j <= ny - 2; ++j) {
    if (j > 3) {
      w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4]; w[4] = I(j + 2);
    }
    double r = 0.0;
#pragma unroll
    for (int q = 1; q <= 5; ++q) r += F8(c, q, j) * w[q - 1];
    y = r - F8(c, 6, j) * y;
    out[(size_t)(j - 1) * s] = (T)y;
  }
  for (int j = ny - 1; j <= ny; ++j) {
    double r = 0.0;
#pragma unroll
    for (int q = 1; q <= 4; ++q) r += F8(c, q, j) * I(ny - 4 + q);
    y = r - F8(c, 6, j) * y;
    out[(size_t)(j - 1) * s] = (T)y;
  }
  double x = y * F8(c, 7, ny);
  out[(size_t)(ny - 1) * s] = (T)x;
  for (int j = ny - 1; j >= 1; --j) {
    x = ((double)out[(size_t)(j - 1) * s] - F8(c, 8, j) * x) * F8(c, 7, j);
    out[(size_t)(j - 1) * s] = (T)x;
  }
}

struct StatsY {
  const void *ux, *w, *dwdx, *dudx, *p, *v;
  void *B[8];
  const double *fd_iy, *fd_iyp, *fd_dvdy;
};

// B1 = (u_x)_y, B2 = w_y, B3 = (dw/dx)_y, B4 = (du/dx)_y, B5 = p_y,
// B6 = d(B1)/dy, B7 = d(B2)/dy, B8 = dv/dy (stotf.F90:105-116), all at the
// v rows; one thread per real column (kr, i)
template <typename T>
__global__ void __launch_bounds__(TPB) stats_y_kernel(StatsY a, int nx, int ny, int nz1) {
  const size_t col = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (col >= (size_t)nz1 * nx) return;
  const size_t i = col / nz1, kr = col - i * nz1;
  const size_t ou = i * (size_t)(ny + 1) * nz1 + kr, ov = i * (size_t)ny * nz1 + kr;
  T *B[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) B[q] = (T *)a.B[q] + ov;
  col_interpy<T>((const T *)a.ux + ou, B[0], a.fd_iy, ny, nz1, 0);
  col_interpy<T>((const T *)a.w + ou, B[1], a.fd_iy, ny, nz1, 0);
  col_interpy<T>((const T *)a.dwdx + ou, B[2], a.fd_iy, ny, nz1, 0);
  col_interpy<T>((const T *)a.dudx + ou, B[3], a.fd_iy, ny, nz1, 0);
  col_interpy<T>((const T *)a.p + ov, B[4], a.fd_iyp, ny, nz1, 1);
  col_diffy<T>(B[0], B[5], a.fd_dvdy, ny, nz1);
  col_diffy<T>(B[1], B[6], a.fd_dvdy, ny, nz1);
  col_diffy<T>((const T *)a.v + ov, B[7], a.fd_dvdy, ny, nz1);
}

// The FP64 mean-flow columns of a mixed box (the real part of mode 0 of u_x
// and du/dx, [ny+1][nx]): B1, B4 and B6 of that column in FP64 ([ny][nx]);
// one thread per plane.
__global__ void __launch_bounds__(TPB)
stats_y_mean_kernel(const double *__restrict__ ux0, const double *__restrict__ dudx0, double *m1, double *m4,
                    double *m6, const double *fd_iy, const double *fd_dvdy, int nx, int ny) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= nx) return;
  col_interpy<double>(ux0 + i, m1 + i, fd_iy, ny, nx, 0);
  col_interpy<double>(dudx0 + i, m4 + i, fd_iy, ny, nx, 0);
  col_diffy<double>(m1 + i, m6 + i, fd_dvdy, ny, nx);
}

// ---------------------------------------------------------------------------
// Row sums of one (j, i) row
// ---------------------------------------------------------------------------
// accumulator ids (the order of the reference's statistics files:
// escru.F90:318-341 then 490-516)
enum {
  S_US, S_VS, S_WS, S_UV, S_UA, S_VA, S_WA, S_PP, S_PM, S_VORTX, S_VORTY, S_VORTZ, S_VORTXA, S_VORTYA,
  S_VORTZA, S_UW, S_VW, S_DISPU, S_DISPV, S_DISPW, S_DISPUV, S_PUP, S_PVP, S_PDUDX, S_PDVDY, S_PDUDY,
  S_PDVDX, S_PDWDZ, S_V3, S_U2V, S_V2U, S_W2V, S_DUDX0, S_DUDY0, S_DVDX0, S_DVDY0, S_DWDX0, S_DWDY0, S_U3,
  S_W2U, S_COUNT
};
static_assert(S_COUNT == OTBL_NSTAT, "statistic count");

// the sums over modes and over z reduced across the workgroup
constexpr int NSUM = 21 + 6;
constexpr int ZSLOTS = 4096 / TPB; // physical points per thread (nz <= 4096)

struct StatsZ {
  const void *B[8], *v, *dvdx;
  // mixed box: the real part of mode 0 of B1, B4, B6 from the FP64 mean
  // flow, [ny][nx] (null: from the fields)
  const double *m1, *m4, *m6;
  const double *kaz;
  double *acc; // [S_COUNT][nx][ny+1]
  float *ens;  // [nx][ny][nz]
};

template <typename T>
__global__ void __launch_bounds__(TPB) stats_z_kernel(StatsZ a, FFTPlan plan, int nx, int ny, int nz1, int nz2) {
  extern __shared__ double2 sh[];
  __shared__ double red[TPB / 64][NSUM];
  const int n = plan.n;
  double2 *lo = sh, *hi = sh + 64, *A = sh + 192, *Bf = A + n;
  const int j = blockIdx.x + 1, i = blockIdx.y + 1, t = threadIdx.x;
  const size_t ov = ((size_t)(i - 1) * ny + (j - 1)) * nz1;
  const T *Bq[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) Bq[q] = (const T *)a.B[q] + ov;
  const T *V = (const T *)a.v + ov, *DV = (const T *)a.dvdx + ov;
  const size_t om = (size_t)(j - 1) * nx + (i - 1); // (i, j) of the [ny][nx] mean-flow rows
  auto ld = [&](const T *p, int k) {
    double r = (double)p[2 * k];
    if (k == 0 && a.m1) r = p == Bq[0] ? a.m1[om] : p == Bq[3] ? a.m4[om] : p == Bq[5] ? a.m6[om] : r;
    return make_double2(r, (double)p[2 * k + 1]);
  };
  fill_twiddles(lo, hi, n);
  // ---- sums over the modes (stotf.F90:156-218) ----
  double sum[NSUM];
#pragma unroll
  for (int q = 0; q < NSUM; ++q) sum[q] = 0.0;
  auto re = [](double2 x, double2 y) { return x.x * y.x + x.y * y.y; }; // Re(x conj(y))
  for (int k = t; k <= nz2; k += TPB) {
    const double cte = k == 0 ? 1.0 : 2.0, kz = a.kaz[k];
    const double2 b1 = ld(Bq[0], k), b2 = ld(Bq[1], k), b3 = ld(Bq[2], k), b4 = ld(Bq[3], k);
    const double2 b5 = ld(Bq[4], k), b6 = ld(Bq[5], k), b7 = ld(Bq[6], k), b8 = ld(Bq[7], k);
    const double2 v = ld(V, k), dv = ld(DV, k);
    // z derivatives: x kaz = (-kz x.y, kz x.x)
    const double2 uz = make_double2(-kz * b1.y, kz * b1.x), vz = make_double2(-kz * v.y, kz * v.x);
    const double2 wz = make_double2(-kz * b2.y, kz * b2.x);
    const double2 ox = make_double2(b7.x - vz.x, b7.y - vz.y), oy = make_double2(uz.x - b3.x, uz.y - b3.y);
    const double2 oz = make_double2(dv.x - b6.x, dv.y - b6.y);
    sum[0] += cte * re(b1, b1);  // us
    sum[1] += cte * re(v, v);    // vs
    sum[2] += cte * re(b2, b2);  // ws
    sum[3] += cte * re(b1, v);   // uv
    sum[4] += cte * re(b5, b5);  // pp
    sum[5] += cte * re(ox, ox);  // vortx
    sum[6] += cte * re(oy, oy);  // vorty
    sum[7] += cte * re(oz, oz);  // vortz
    sum[8] += cte * re(b1, b2);  // uw
    sum[9] += cte * re(v, b2);   // vw
    sum[10] += cte * (re(b4, b4) + re(uz, uz) + re(b6, b6)); // dispu
    sum[11] += cte * (re(dv, dv) + re(vz, vz) + re(b8, b8)); // dispv
    sum[12] += cte * (re(wz, wz) + re(b3, b3) + re(b7, b7)); // dispw
    sum[13] += cte * (re(uz, vz) + re(b4, dv) + re(b6, b8)); // dispuv
    sum[14] += cte * re(b5, b1); // pup
    sum[15] += cte * re(b5, v);  // pvp
    sum[16] += cte * re(b5, b4); // pdudx
    sum[17] += cte * re(b5, b8); // pdvdy
    sum[18] += cte * re(b5, b6); // pdudy
    sum[19] += cte * re(b5, dv); // pdvdx
    sum[20] += cte * re(b5, wz); // pdwdz
  }
  // ---- physical space (stotf.F90:148-154, 256-283) ----
  // spectrum of signal q at mode k <= nz2: 0 omega_x, 1 omega_y, 2 omega_z,
  // 3 (u_x)_y, 4 v, 5 w_y
  auto spec = [&](int q, int k) -> double2 {
    const double kz = a.kaz[k];
    switch (q) {
    case 0: { const double2 b7 = ld(Bq[6], k), v = ld(V, k); return make_double2(b7.x + kz * v.y, b7.y - kz * v.x); }
    case 1: { const double2 b1 = ld(Bq[0], k), b3 = ld(Bq[2], k); return make_double2(-kz * b1.y - b3.x, kz * b1.x - b3.y); }
    case 2: { const double2 dv = ld(DV, k), b6 = ld(Bq[5], k); return make_double2(dv.x - b6.x, dv.y - b6.y); }
    case 3: return ld(Bq[0], k);
    case 4: return ld(V, k);
    default: return ld(Bq[1], k);
    }
  };
  // x = c2r(X), y = c2r(Y) packed as the transform of the Hermitian
  // extension of X + i Y (Im of mode 0 does not enter a real signal)
  auto transform = [&](int qx, int qy) -> const double2 * {
    for (int k = t; k < n; k += TPB) {
      const int kk = k <= nz2 ? k : (k >= n - nz2 ? n - k : -1);
      double2 z = make_double2(0.0, 0.0);
      if (kk >= 0) {
        const double2 X = spec(qx, kk), Y = spec(qy, kk);
        if (k == 0) z = make_double2(X.x, Y.x);
        else if (k == kk) z = make_double2(X.x - Y.y, X.y + Y.x);
        else z = make_double2(X.x + Y.y, -X.y + Y.x);
      }
      A[k] = z;
    }
    __syncthreads();
    return fft_multi<+1>(A, Bf, n, 1, plan, lo, hi) ? A : Bf;
  };
  double s2[ZSLOTS], U[ZSLOTS];
  const double2 *Z = transform(0, 1);
#pragma unroll
  for (int q = 0; q < ZSLOTS; ++q) {
    const int z = t + q * TPB;
    if (z < n) s2[q] = Z[z].x * Z[z].x + Z[z].y * Z[z].y;
  }
  __syncthreads();
  Z = transform(2, 3);
  float *ens = a.ens + ((size_t)(i - 1) * ny + (j - 1)) * n;
#pragma unroll
  for (int q = 0; q < ZSLOTS; ++q) {
    const int z = t + q * TPB;
    if (z < n) {
      ens[z] = (float)sqrt(s2[q] + Z[z].x * Z[z].x);
      U[q] = Z[z].y;
    }
  }
  __syncthreads();
  Z = transform(4, 5);
#pragma unroll
  for (int q = 0; q < ZSLOTS; ++q) {
    const int z = t + q * TPB;
    if (z < n) {
      const double u = U[q], v = Z[z].x, w = Z[z].y;
      sum[21] += u * u * u;  // u3
      sum[22] += u * w * w;  // w2u
      sum[23] += v * v * v;  // v3
      sum[24] += v * w * w;  // w2v
      sum[25] += u * v * v;  // v2u
      sum[26] += v * u * u;  // u2v
    }
  }
  // ---- reduce across the workgroup ----
#pragma unroll
  for (int q = 0; q < NSUM; ++q) {
    double x = sum[q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if ((t & 63) == 0) red[t / 64][q] = x;
  }
  __syncthreads();
  if (t != 0) return;
#pragma unroll
  for (int q = 0; q < NSUM; ++q)
    for (int w = 1; w < TPB / 64; ++w) red[0][q] += red[w][q];
  const double *r = red[0];
  const size_t NR = (size_t)(ny + 1), at0 = (size_t)(i - 1) * NR + (j - 1), S = (size_t)nx * NR;
  double *acc = a.acc + at0;
  acc[S_US * S] += r[0];
  acc[S_VS * S] += r[1];
  acc[S_WS * S] += r[2];
  acc[S_UV * S] += r[3];
  acc[S_PP * S] += r[4];
  acc[S_VORTX * S] += r[5];
  acc[S_VORTY * S] += r[6];
  acc[S_VORTZ * S] += r[7];
  acc[S_UW * S] += r[8];
  acc[S_VW * S] += r[9];
  acc[S_DISPU * S] += r[10];
  acc[S_DISPV * S] += r[11];
  acc[S_DISPW * S] += r[12];
  acc[S_DISPUV * S] += r[13];
  acc[S_PUP * S] += r[14];
  acc[S_PVP * S] += r[15];
  acc[S_PDUDX * S] += r[16];
  acc[S_PDVDY * S] += r[17];
  acc[S_PDUDY * S] += r[18];
  acc[S_PDVDX * S] += r[19];
  acc[S_PDWDZ * S] += r[20];
  acc[S_U3 * S] += r[21];
  acc[S_W2U * S] += r[22];
  acc[S_V3 * S] += r[23];
  acc[S_W2V * S] += r[24];
  acc[S_V2U * S] += r[25];
  acc[S_U2V * S] += r[26];
  // k = 0 means (stotf.F90:128-144): real parts
  const double b1 = ld(Bq[0], 0).x, b2 = (double)Bq[1][0], b3 = (double)Bq[2][0], b4 = ld(Bq[3], 0).x;
  const double b5 = (double)Bq[4][0], b6 = ld(Bq[5], 0).x, b7 = (double)Bq[6][0], b8 = (double)Bq[7][0];
  const double v0 = (double)V[0], dv0 = (double)DV[0];
  acc[S_UA * S] += b1;
  acc[S_WA * S] += b2;
  acc[S_VA * S] += v0;
  acc[S_PM * S] += b5;
  acc[S_DUDX0 * S] += b4;
  acc[S_DVDX0 * S] += dv0;
  acc[S_DWDX0 * S] += b3;
  acc[S_DUDY0 * S] += b6;
  acc[S_DVDY0 * S] += b8;
  acc[S_DWDY0 * S] += b7;
  acc[S_VORTYA * S] -= b3;
  acc[S_VORTXA * S] += b7;
  acc[S_VORTZA * S] += dv0 - b6;
}

// ---------------------------------------------------------------------------
// The same row sums on the register-resident z engine (nz of the engine,
// nz2 = nz/3 - 1): one (j, i) row per team of W wavefronts, rows numbered
// (i-1) ny + (j-1).  The three packed c2r transforms run in the fields'
// precision in registers (their inputs formed from the spectral rows as
// they are loaded), the physical products and every sum in FP64.
// ---------------------------------------------------------------------------
template <typename C> struct CPair {
  C x, y;
};

// Complex input of the c2r transform of two real signals whose spectra
// spec(k) gives for 0 <= k <= nz2: the Hermitian-extended spectrum of
// x + i y in the register slots of the engine's first pass (Im of mode 0
// does not enter a real signal; zero between nz2 and N - nz2).
template <int N, typename C, class F>
__device__ __forceinline__ void zload_pair(C (&v)[N / (64 * ZPlan<N>::W)], int t, F spec) {
  constexpr int G = 64 * ZPlan<N>::W, RD = ZPlan<N>::rad(0), NB = N / (RD * G), nz2 = N / 3 - 1;
#pragma unroll
  for (int m = 0; m < NB; ++m)
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int base = G * m + r * (N / RD), k = base + t;
      C z = mkc<C>(0.0, 0.0);
      if (base + G - 1 <= nz2) { // all lanes below the band
        const CPair<C> s = spec(k);
        z = k == 0 ? mkc<C>(s.x.x, s.y.x) : mkc<C>(s.x.x - s.y.y, s.x.y + s.y.x);
      } else if (base >= N - nz2) { // all above: conj(X[N-k]) + i conj(Y[N-k])
        const CPair<C> s = spec(N - k);
        z = mkc<C>(s.x.x + s.y.y, -s.x.y + s.y.x);
      } else if (base <= nz2 || base + G - 1 >= N - nz2) { // straddles an edge
        const bool lo = k <= nz2, hi = k >= N - nz2;
        const CPair<C> s = spec(lo ? k : (hi ? N - k : 0));
        C r2 = lo ? mkc<C>(s.x.x - s.y.y, s.x.y + s.y.x) : mkc<C>(s.x.x + s.y.y, -s.x.y + s.y.x);
        r2 = k == 0 ? mkc<C>(s.x.x, s.y.x) : r2;
        z = (lo || hi) ? r2 : mkc<C>(0.0, 0.0);
      }
      v[m * RD + r] = z;
    }
}

template <int N, typename T>
__global__ void __launch_bounds__(ZTeam<N>::TPB_)
stats_zr(StatsZ a, const typename CT<T>::c *__restrict__ tw, int nx, int ny, int nz1, int nrows) {
  using C = typename CT<T>::c;
  using Z = ZTeam<N>;
  constexpr int nz2 = N / 3 - 1, NK = (nz2 + 1 + Z::G - 1) / Z::G;
  __shared__ C lds_all[Z::TEAMS * zlds_elems<N>()];
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / Z::G), t = threadIdx.x % Z::G;
  C *lds = lds_all + team * zlds_elems<N>();
  const int row = blockIdx.x * Z::TEAMS + team;
  if (row >= nrows) return; // a whole team
  const int j = row % ny + 1, i = row / ny + 1;
  const size_t ov = (size_t)row * nz1;
  const C *Bq[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) Bq[q] = (const C *)((const T *)a.B[q] + ov);
  const C *V = (const C *)((const T *)a.v + ov), *DV = (const C *)((const T *)a.dvdx + ov);
  const size_t om = (size_t)(j - 1) * nx + (i - 1); // (i, j) of the [ny][nx] mean-flow rows
  // mode 0 of B1 (u_x at the v rows) with the FP64 mean flow of a mixed box
  const T m1 = a.m1 ? (T)a.m1[om] : T(0);
  auto b1at = [&](int k) {
    C b = Bq[0][k];
    if (a.m1 && k == 0) b.x = m1;
    return b;
  };
  double sum[NSUM];
#pragma unroll
  for (int q = 0; q < NSUM; ++q) sum[q] = 0.0;
  // ---- physical space (stotf.F90:148-154, 256-283) ----
  T s2[Z::E], U[Z::E];
  {
    C v[Z::E]; // omega_x + i omega_y
    zload_pair<N>(v, t, [&](int k) {
      const T kz = (T)a.kaz[k];
      const C b7 = Bq[6][k], vv = V[k], b1 = b1at(k), b3 = Bq[2][k];
      return CPair<C>{mkc<C>(b7.x + kz * vv.y, b7.y - kz * vv.x), mkc<C>(-kz * b1.y - b3.x, kz * b1.x - b3.y)};
    });
    zrun<N, +1, false>(v, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) s2[e] = v[e].x * v[e].x + v[e].y * v[e].y;
  }
  {
    C v[Z::E]; // omega_z + i u
    zload_pair<N>(v, t, [&](int k) {
      const C dv = DV[k], b6 = Bq[5][k];
      return CPair<C>{mkc<C>(dv.x - b6.x, dv.y - b6.y), b1at(k)};
    });
    zrun<N, +1, false>(v, t, tw, lds);
    float *ens = a.ens + (size_t)row * N;
#pragma unroll
    for (int e = 0; e < Z::E; ++e) {
      ens[zphys_index<N>(t, e)] = (float)sqrt((double)s2[e] + (double)v[e].x * (double)v[e].x);
      U[e] = v[e].y;
    }
  }
  {
    C v[Z::E]; // v + i w
    zload_pair<N>(v, t, [&](int k) { return CPair<C>{V[k], Bq[1][k]}; });
    zrun<N, +1, false>(v, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) {
      const double u = U[e], vv = v[e].x, w = v[e].y;
      sum[21] += u * u * u;   // u3
      sum[22] += u * w * w;   // w2u
      sum[23] += vv * vv * vv; // v3
      sum[24] += vv * w * w;  // w2v
      sum[25] += u * vv * vv; // v2u
      sum[26] += vv * u * u;  // u2v
    }
  }
  // ---- sums over the modes (stotf.F90:156-218), modes t + G it ----
  auto re = [](double2 x, double2 y) { return x.x * y.x + x.y * y.y; }; // Re(x conj(y))
  auto wd = [](C x) { return make_double2((double)x.x, (double)x.y); };
  // one mode per lane at a time: the 10 loads and 21 FP64 sums of a mode
  // fill the registers the transforms no longer need
#pragma unroll 1
  for (int it = 0; it < NK; ++it) {
    const int k0 = t + Z::G * it, k = min(k0, nz2);
    const double cte = k0 > nz2 ? 0.0 : (k0 == 0 ? 1.0 : 2.0), kz = a.kaz[k];
    double2 b1 = wd(Bq[0][k]), b4 = wd(Bq[3][k]), b6 = wd(Bq[5][k]);
    if (a.m1 && k0 == 0) {
      b1.x = a.m1[om];
      b4.x = a.m4[om];
      b6.x = a.m6[om];
    }
    const double2 b2 = wd(Bq[1][k]), b3 = wd(Bq[2][k]), b5 = wd(Bq[4][k]), b7 = wd(Bq[6][k]), b8 = wd(Bq[7][k]);
    const double2 v = wd(V[k]), dv = wd(DV[k]);
    const double2 uz = make_double2(-kz * b1.y, kz * b1.x), vz = make_double2(-kz * v.y, kz * v.x);
    const double2 wz = make_double2(-kz * b2.y, kz * b2.x);
    const double2 ox = make_double2(b7.x - vz.x, b7.y - vz.y), oy = make_double2(uz.x - b3.x, uz.y - b3.y);
    const double2 oz = make_double2(dv.x - b6.x, dv.y - b6.y);
    sum[0] += cte * re(b1, b1);
    sum[1] += cte * re(v, v);
    sum[2] += cte * re(b2, b2);
    sum[3] += cte * re(b1, v);
    sum[4] += cte * re(b5, b5);
    sum[5] += cte * re(ox, ox);
    sum[6] += cte * re(oy, oy);
    sum[7] += cte * re(oz, oz);
    sum[8] += cte * re(b1, b2);
    sum[9] += cte * re(v, b2);
    sum[10] += cte * (re(b4, b4) + re(uz, uz) + re(b6, b6));
    sum[11] += cte * (re(dv, dv) + re(vz, vz) + re(b8, b8));
    sum[12] += cte * (re(wz, wz) + re(b3, b3) + re(b7, b7));
    sum[13] += cte * (re(uz, vz) + re(b4, dv) + re(b6, b8));
    sum[14] += cte * re(b5, b1);
    sum[15] += cte * re(b5, v);
    sum[16] += cte * re(b5, b4);
    sum[17] += cte * re(b5, b8);
    sum[18] += cte * re(b5, b6);
    sum[19] += cte * re(b5, dv);
    sum[20] += cte * re(b5, wz);
  }
  // ---- reduce across the team ----
#pragma unroll
  for (int q = 0; q < NSUM; ++q)
    for (int o = 32; o > 0; o >>= 1) sum[q] += __shfl_down(sum[q], o, 64);
  if constexpr (Z::W > 1) { // combine the team's wavefronts through its LDS buffer (8 bytes an element... as doubles)
    double *red = reinterpret_cast<double *>(lds);
    if ((t & 63) == 0)
      for (int q = 0; q < NSUM; ++q) red[q * Z::W + t / 64] = sum[q];
    zsync<Z::W>();
    if (t == 0)
      for (int q = 0; q < NSUM; ++q)
        for (int w = 1; w < Z::W; ++w) sum[q] += red[q * Z::W + w];
  }
  if (t != 0) return;
  const size_t NR = (size_t)(ny + 1), at0 = (size_t)(i - 1) * NR + (j - 1), S = (size_t)nx * NR;
  double *acc = a.acc + at0;
  constexpr int ID[NSUM] = {S_US,   S_VS,    S_WS,    S_UV,    S_PP,    S_VORTX, S_VORTY, S_VORTZ, S_UW,
                            S_VW,   S_DISPU, S_DISPV, S_DISPW, S_DISPUV, S_PUP,  S_PVP,   S_PDUDX, S_PDVDY,
                            S_PDUDY, S_PDVDX, S_PDWDZ, S_U3,   S_W2U,   S_V3,    S_W2V,   S_V2U,   S_U2V};
#pragma unroll
  for (int q = 0; q < NSUM; ++q) acc[ID[q] * S] += sum[q];
  // k = 0 means (stotf.F90:128-144): real parts
  const double b1 = a.m1 ? a.m1[om] : (double)Bq[0][0].x, b4 = a.m1 ? a.m4[om] : (double)Bq[3][0].x;
  const double b6 = a.m1 ? a.m6[om] : (double)Bq[5][0].x;
  const double b2 = Bq[1][0].x, b3 = Bq[2][0].x, b5 = Bq[4][0].x, b7 = Bq[6][0].x, b8 = Bq[7][0].x;
  const double v0 = V[0].x, dv0 = DV[0].x;
  acc[S_UA * S] += b1;
  acc[S_WA * S] += b2;
  acc[S_VA * S] += v0;
  acc[S_PM * S] += b5;
  acc[S_DUDX0 * S] += b4;
  acc[S_DVDX0 * S] += dv0;
  acc[S_DWDX0 * S] += b3;
  acc[S_DUDY0 * S] += b6;
  acc[S_DVDY0 * S] += b8;
  acc[S_DWDY0 * S] += b7;
  acc[S_VORTYA * S] -= b3;
  acc[S_VORTXA * S] += b7;
  acc[S_VORTZA * S] += dv0 - b6;
}

__global__ void stats_zero_kernel(double *a, size_t n) {
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < n; e += (size_t)gridDim.x * TPB) a[e] = 0.0;
}

size_t k_stats_shm_bytes(int nz) { return (192 + 2 * (size_t)nz) * sizeof(double2); }

template <typename T> static void statsp_t(Box &b) {
  const int nx = b.nx, ny = b.ny, nz1 = b.nz1;
  const T *u = fld<T>(b.f[OTBL_U]), *v = fld<T>(b.f[OTBL_V]), *w = fld<T>(b.f[OTBL_W]), *p = fld<T>(b.f[OTBL_P]);
  // work fields (free at the start of a substep)
  T *ux = fld<T>(b.wki1), *dudx = fld<T>(b.wki3), *dwdx = fld<T>(b.rhsw), *dvdx = fld<T>(b.rhsv);
  T *B[8] = {fld<T>(b.resu), fld<T>(b.resw), fld<T>(b.rhsu), fld<T>(b.ya),
             fld<T>(b.yb),   fld<T>(b.yc),   fld<T>(b.resv), fld<T>(b.wki2)};
  const double *fd_dvdy = b.stcoef, *fd_iy = fd_dvdy + 8 * (size_t)(ny + 1), *fd_iyp = fd_iy + 8 * (size_t)(ny + 1);
  const double *fd_dx = fd_iyp + 8 * (size_t)(ny + 1), *fd_ix = fd_dx + 8 * (size_t)(nx + 1);
  LAUNCH(b, "stats_x", nblk(b.pu), TPB, 0, (stats_x_kernel<T, 1>), u, ux, fd_ix, nx, b.pu, 1, b.pu);
  LAUNCH(b, "stats_x", nblk(b.pu), TPB, 0, (stats_x_kernel<T, 0>), ux, dudx, fd_dx, nx, b.pu, 1, b.pu);
  LAUNCH(b, "stats_x", nblk(b.pv), TPB, 0, (stats_x_kernel<T, 0>), v, dvdx, fd_dx, nx, b.pv, 1, b.pv);
  LAUNCH(b, "stats_x", nblk(b.pu), TPB, 0, (stats_x_kernel<T, 0>), w, dwdx, fd_dx, nx, b.pu, 1, b.pu);
  // mixed box: the mean-flow u (mode 0, real part) is held in FP64 (its
  // FP32 rounding, amplified by d/dx, would swamp du/dx of the mean flow);
  // its lines run in FP64 beside the FP32 fields
  double *ux0 = nullptr, *dudx0 = nullptr, *m1 = nullptr, *m4 = nullptr, *m6 = nullptr;
  if (!std::is_same<T, double>::value) {
    const size_t nm = (size_t)nx * (ny + 1);
    ux0 = b.stmean;
    dudx0 = ux0 + nm;
    m1 = dudx0 + nm;
    m4 = m1 + nm;
    m6 = m4 + nm;
    LAUNCH(b, "stats_x", nblk(ny + 1), TPB, 0, (stats_x_kernel<double, 1>), (const double *)b.um64, ux0, fd_ix, nx,
           (size_t)(ny + 1), (size_t)nx, (size_t)1);
    LAUNCH(b, "stats_x", nblk(ny + 1), TPB, 0, (stats_x_kernel<double, 0>), (const double *)ux0, dudx0, fd_dx, nx,
           (size_t)(ny + 1), (size_t)nx, (size_t)1);
  }
  StatsY ay{ux, w, dwdx, dudx, p, v, {}, fd_iy, fd_iyp, fd_dvdy};
  for (int q = 0; q < 8; ++q) ay.B[q] = B[q];
  LAUNCH(b, "stats_y", nblk((size_t)nz1 * nx), TPB, 0, stats_y_kernel<T>, ay, nx, ny, nz1);
  if (m1)
    LAUNCH(b, "stats_y", nblk(nx), TPB, 0, stats_y_mean_kernel, (const double *)ux0, (const double *)dudx0, m1, m4,
           m6, fd_iy, fd_dvdy, nx, ny);
  StatsZ az{};
  az.m1 = m1;
  az.m4 = m4;
  az.m6 = m6;
  for (int q = 0; q < 8; ++q) az.B[q] = B[q];
  az.v = v;
  az.dvdx = dvdx;
  az.kaz = b.dc.kaz;
  az.acc = b.st;
  az.ens = b.ens;
  if (b.zfast) {
    const int nrows = ny * nx;
    const typename CT<T>::c *tw = std::is_same<T, double>::value ? (const typename CT<T>::c *)b.ztw
                                                                  : (const typename CT<T>::c *)b.ztw_f;
#define SZR(NN)                                                                                                    \
  case NN:                                                                                                         \
    LAUNCH(b, "stats_z", (unsigned)((nrows + ZTeam<NN>::TEAMS - 1) / ZTeam<NN>::TEAMS), ZTeam<NN>::TPB_, 0,       \
           (stats_zr<NN, T>), az, tw, nx, ny, nz1, nrows);                                                      \
    break;
    switch (b.nz) {
      SZR(256)
      SZR(512)
      SZR(1024)
      SZR(2048)
      SZR(4096)
#undef SZR
    default: throw std::invalid_argument("stats: z length not on the register engine");
    }
    return;
  }
  LAUNCH(b, "stats_z", dim3(ny, nx), TPB, k_stats_shm_bytes(b.nz), stats_z_kernel<T>, az, b.zplan, nx, ny, nz1,
         b.nz2);
}

void k_statsp(Box &b) {
  if (b.precision == OTBL_MIXED)
    statsp_t<float>(b);
  else
    statsp_t<double>(b);
}

void k_stats_zero(Box &b) {
  const size_t n = (size_t)OTBL_NSTAT * b.nx * (b.ny + 1);
  LAUNCH(b, "stats_zero", (unsigned)std::min<size_t>(nblk(n), 4096), TPB, 0, stats_zero_kernel, b.st, n);
}

void k_stats_attrs(Box &b) {
  const int shm = (int)k_stats_shm_bytes(b.nz);
  OTBL_HIP(hipFuncSetAttribute((const void *)stats_z_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
  OTBL_HIP(hipFuncSetAttribute((const void *)stats_z_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
}

} // namespace otbl
