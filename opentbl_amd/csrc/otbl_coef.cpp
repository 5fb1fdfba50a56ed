// FP64 coefficient setup of the RK3 step (the reference's coef,
// coeft.f90:10-1234).  The staggered compact schemes keep the reference's
// formulas: interior weights from the non-uniform Pade formulas, boundary
// closures from small Taylor-moment systems (coefb, coeft.f90:1360-1789),
// the Poisson y operator from mtr (coeft.f90:1244-1332).  Left-hand sides are
// left unfactored; the LU pre-factoring (coefinv, coeft.f90:1338-1353) runs
// on the GPU in FP64 (k_factor_coef).
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "otbl_internal.h"

namespace otbl {
namespace {

// element (q, i) of an (r, 1:n) array stored with column i at offset r*i
inline double &at(std::vector<double> &a, int r, int q, int i) { return a[(size_t)r * i + q - 1]; }

// Gaussian elimination with partial pivoting on an n x n system (n <= 5).
void solve_small(int n, double m[5][5], double rhs[5]) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
    if (m[p][c] == 0.0) throw std::runtime_error("singular boundary closure");
    if (p != c) {
      for (int j = 0; j < n; ++j) std::swap(m[p][j], m[c][j]);
      std::swap(rhs[p], rhs[c]);
    }
    for (int r = c + 1; r < n; ++r) {
      double f = m[r][c] / m[c][c];
      for (int j = c; j < n; ++j) m[r][j] -= f * m[c][j];
      rhs[r] -= f * rhs[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double acc = rhs[r];
    for (int j = r + 1; j < n; ++j) acc -= m[r][j] * rhs[j];
    rhs[r] = acc / m[r][r];
  }
}

enum ClosureKind {
  MIDPOINT3 = 0,   // coefb flag 0: 3-point interpolation to a half-cell point
  SECOND5 = 1,     // flag 1: compact second derivative, 4 explicit points
  FIRST_OFF = 3,   // flag 3: compact first derivative, offsets s
  MIDPOINT4 = 4,   // flag 4: 4-point interpolation
  FIRST_MIX = 5,   // flag 5
  ZEROVORT6 = 6,   // flag 6
  ZEROVORT7 = 7,   // flag 7 (its second-moment row carries s^2, not s^2/2, on columns 3,4)
};

// The Taylor-moment systems of coefb.  Returns the weights in w[0..n-1].
void closure(ClosureKind kind, double a, double b, double c, double d, double w[5]) {
  double m[5][5] = {{0}};
  double r[5] = {0, 0, 0, 0, 0};
  int n = 4;
  auto moments = [&](const double s[4], int row0, double sc2[4]) {
    // rows row0..row0+2: s, s^2*sc2, s^3/6
    for (int j = 0; j < 4; ++j) {
      m[row0][j] = s[j];
      m[row0 + 1][j] = s[j] * s[j] * sc2[j];
      m[row0 + 2][j] = s[j] * s[j] * s[j] / 6.0;
    }
  };
  double half[4] = {0.5, 0.5, 0.5, 0.5};
  switch (kind) {
  case MIDPOINT3: {
    n = 3;
    double s[3] = {-0.5 * a, -(0.5 * a + b), -(0.5 * a + b + c)};
    for (int j = 0; j < 3; ++j) {
      m[0][j] = 1.0;
      m[1][j] = s[j];
      m[2][j] = 0.5 * s[j] * s[j];
    }
    r[0] = 1.0;
    break;
  }
  case SECOND5: {
    n = 5;
    double t[3] = {a, a + b, a + b + c};
    m[0][0] = -1.0; m[3][0] = -a; m[4][0] = -0.5 * a * a;
    m[1][1] = 1.0;
    for (int j = 0; j < 3; ++j) {
      m[0][2 + j] = 0.5 * t[j] * t[j];
      m[1][2 + j] = 1.0;
      m[2][2 + j] = t[j];
      m[3][2 + j] = t[j] * t[j] * t[j] / 6.0;
      m[4][2 + j] = t[j] * t[j] * t[j] * t[j] / 24.0;
    }
    r[0] = 1.0;
    break;
  }
  case FIRST_OFF: {
    double s[4] = {0.5 * a, 0.5 * a + b, 0.5 * a + b + c, 0.5 * a + b + c + d};
    for (int j = 0; j < 4; ++j) {
      m[0][j] = s[j];
      m[1][j] = 1.0;
      m[2][j] = 0.5 * s[j] * s[j];
      m[3][j] = s[j] * s[j] * s[j] / 6.0;
    }
    r[0] = 1.0;
    break;
  }
  case FIRST_MIX: {
    double s[4] = {0.5 * a, -0.5 * b, -(0.5 * b + c), -(0.5 * b + c + d)};
    for (int j = 0; j < 4; ++j) {
      m[0][j] = s[j];
      m[1][j] = 1.0;
      m[2][j] = 0.5 * s[j] * s[j];
      m[3][j] = s[j] * s[j] * s[j] / 6.0;
    }
    r[0] = 1.0;
    break;
  }
  case MIDPOINT4: {
    double s[4] = {-0.5 * a, 0.5 * b, 0.5 * b + c, 0.5 * b + c + d};
    for (int j = 0; j < 4; ++j) m[0][j] = 1.0;
    moments(s, 1, half);
    r[0] = 1.0;
    break;
  }
  case ZEROVORT6: {
    double s[4] = {0.0, -0.5 * a, -(0.5 * a + b), -(0.5 * a + b + c)};
    for (int j = 0; j < 4; ++j) m[0][j] = 1.0;
    moments(s, 1, half);
    r[1] = 1.0;
    break;
  }
  case ZEROVORT7: {
    double s[4] = {0.5 * a, -0.5 * b, -(0.5 * b + c), -(0.5 * b + c + d)};
    double sc[4] = {0.5, 0.5, 1.0, 1.0};
    for (int j = 0; j < 4; ++j) m[0][j] = 1.0;
    moments(s, 1, sc);
    r[1] = 1.0;
    break;
  }
  }
  solve_small(n, m, r);
  for (int j = 0; j < 5; ++j) w[j] = j < n ? r[j] : 0.0;
}

inline double &m4(double *a, int r, int c) { return a[4 * (c - 1) + r - 1]; }
inline double &m5(double *a, int r, int c) { return a[5 * (c - 1) + r - 1]; }
inline double &m2(double *a, int r, int c) { return a[2 * (c - 1) + r - 1]; }

// Non-uniform compact-scheme rows.  A, B are the two neighbouring spacings
// as the reference names them (a, b), C the third where used.
struct Rows {
  // midpoint interpolation, u positions (inxu / inyv form)
  static void interp_a(double a, double b, double c, double &l, double &r) {
    l = c * b * b / ((2 * a + b) * (-b + 2 * a) * (c + a));
    r = -a * b * b / ((b - 2 * c) * (b + 2 * c) * (c + a));
  }
  // midpoint interpolation, v positions (inyu / inxv form)
  static void interp_b(double a, double b, double &l, double &r) {
    l = a * a / ((a + b) * (a + 2 * b));
    r = b * b / (b * b + 2 * a * a + 3 * a * b);
  }
  // first derivative, dcxu / dcyv form
  static void d1_a(double a, double b, double &l, double &r) {
    double den = 2.0 * b * b * b + 2.0 * a * a * a + 9.0 * b * b * a + 9.0 * b * a * a;
    l = -a * (-2.0 * a * b - b * b + 2.0 * a * a) / den;
    r = ((a * a + 2.0 * a * b - 2.0 * b * b) * b) / den;
  }
  // first derivative, dcyu / dcxv form
  static void d1_b(double a, double b, double c, double &l, double &r) {
    double den = -c * c * a - c * c * b + 12.0 * b * b * a + 12.0 * b * a * a;
    l = (b * c * c) / den;
    r = (c * c * a) / den;
  }
  // second derivative, implicit part
  static void d2(double a, double b, double &l, double &r) {
    double den = b * b * b + a * a * a + 4 * a * a * b + 4 * a * b * b;
    l = b * (a * a + a * b - b * b) / den;
    r = -(a * a - a * b - b * b) * a / den;
  }
  // second derivative, explicit part
  static void d2e(double a, double b, double &e1, double &e2, double &e3) {
    double den = b * b * b + a * a * a + 4 * a * a * b + 4 * a * b * b;
    e1 = 12 * b / den;
    e2 = -12 / (a * a + 3 * a * b + b * b);
    e3 = 12 * a / den;
  }
};

void set_tri(std::vector<double> &t, int i, double l, double r) {
  at(t, 3, 1, i) = l;
  at(t, 3, 2, i) = 1.0;
  at(t, 3, 3, i) = r;
}
void set_identity_ends(std::vector<double> &t, int n) {
  set_tri(t, 1, 0.0, 0.0);
  set_tri(t, n, 0.0, 0.0);
}

} // namespace

// coefb (coeft.f90:1360-1789) for the flags the reference calls it with
// (coeft.f90:589-1023): the n closure weights; another flag or another n
// than the flag's is rejected.
void coefb_weights(int n, int flagv, double a, double b, double c, double d, double bcc[5]) {
  ClosureKind k;
  int want;
  switch (flagv) {
  case 0: k = MIDPOINT3; want = 3; break;
  case 1: k = SECOND5; want = 5; break;
  case 3: k = FIRST_OFF; want = 4; break;
  case 4: k = MIDPOINT4; want = 4; break;
  case 5: k = FIRST_MIX; want = 4; break;
  case 6: k = ZEROVORT6; want = 4; break;
  case 7: k = ZEROVORT7; want = 4; break;
  default: throw std::invalid_argument("coefb: flagv not one the reference calls (0, 1, 3..7)");
  }
  if (n != want) throw std::invalid_argument("coefb: n does not match the flag's closure (0: 3, 1: 5, 3..7: 4)");
  try {
    closure(k, a, b, c, d, bcc);
  } catch (const std::runtime_error &e) {
    throw std::invalid_argument(e.what()); // coincident points
  }
}


void build_coef(HostCoef &hc, int nx, int ny, int nz, double re, double ax, double az,
                double cfl, const double *yin) {
  hc.nx = nx; hc.ny = ny; hc.nz = nz;
  hc.nz1 = 2 * (nz / 3);
  hc.nz2 = hc.nz1 / 2 - 1;
  hc.re = re; hc.ax = ax; hc.az = az; hc.cfl = cfl;
  const double pi = 4.0 * std::atan(1.0);
  hc.pi = pi;
  hc.y.assign(yin, yin + ny + 2);
  const std::vector<double> &y = hc.y;
  auto X3 = [&](std::vector<double> &v, int n) { v.assign(3 * (size_t)(n + 2), 0.0); };
  auto X2 = [&](std::vector<double> &v, int n) { v.assign(2 * (size_t)(n + 2), 0.0); };

  // ---- metrics (coeft.f90:32-37, 397-454) ----
  hc.hy.assign(ny + 1, 0.0);
  for (int j = 1; j <= ny - 1; ++j) hc.hy[j] = (y[j + 1] - y[j - 1]) / 2.5 / 2.0;
  hc.hy[0] = (y[1] - y[0]) / 2.5;
  hc.hy[ny] = (y[ny] - y[ny - 1]) / 2.5;
  hc.dy.assign(ny + 1, 0.0);
  hc.dyi.assign(ny + 1, 0.0);
  hc.idy.assign(ny + 1, 0.0);
  hc.idyy.assign(ny + 1, 0.0);
  hc.dymin.assign(ny + 1, 0.0);
  for (int j = 0; j <= ny; ++j) hc.dy[j] = y[j + 1] - y[j];
  for (int j = 1; j <= ny; ++j) hc.dyi[j] = 0.5 * (hc.dy[j] + hc.dy[j - 1]);
  const double h0 = ax * pi / (nx + 1);
  hc.x.assign(nx + 2, 0.0);
  hc.dx.assign(nx + 1, 0.0);
  hc.dxi.assign(nx + 1, 0.0);
  for (int i = 0; i <= nx + 1; ++i) hc.x[i] = (double)i * h0;
  double dmin = 1000.0;
  for (int i = 0; i <= nx; ++i) {
    hc.dx[i] = hc.x[i + 1] - hc.x[i];
    dmin = std::min(dmin, hc.dx[i]);
  }
  for (int i = 1; i <= nx; ++i) hc.dxi[i] = 0.5 * (hc.dx[i] + hc.dx[i - 1]);
  const std::vector<double> &dy = hc.dy, &dyi = hc.dyi, &dx = hc.dx, &dxi = hc.dxi;
  hc.idx = 1.0 / dx[1];
  hc.idxx = 1.0 / dx[1];
  for (int j = 0; j <= ny; ++j) hc.idy[j] = 1.0 / dy[j];
  for (int j = 1; j <= ny; ++j) hc.idyy[j] = 2.0 / (dy[j - 1] + dy[j]);
  double dz = az * 2 * pi / nz;
  double dtrex = re * (cfl * dmin * dmin) / 6.0;
  double dtrez = re * (cfl * dz * dz) / (pi * pi);
  hc.dtret = std::min(dtrex, dtrez);
  for (int j = 1; j <= ny; ++j) hc.dymin[j] = dy[j] * cfl / std::sqrt(3.0);
  hc.dxmin = dmin * cfl / std::sqrt(3.0);
  hc.dzmin = dz * cfl / pi;
  hc.rex = -1.0 / re;

  // spanwise wavenumbers: the reference's cmplx(0d0,k/az) is single precision
  // (bl_1.F90:360), so k/az is rounded to FP32 before use
  hc.kaz.assign(hc.nz2 + 1, 0.0);
  hc.kaz2.assign(hc.nz2 + 1, 0.0);
  hc.kvis.assign(hc.nz2 + 1, 0.0);
  for (int k = 0; k <= hc.nz2; ++k) {
    hc.kaz[k] = (double)(float)(k / az);
    hc.kaz2[k] = hc.kaz[k] * hc.kaz[k];
    hc.kvis[k] = hc.rex * hc.kaz2[k];
  }
  // RK3 (coeft.f90:494-508)
  const double gam[4] = {0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
  const double zet[4] = {0, 0.0, -17.0 / 60.0, -5.0 / 12.0};
  const double alp[4] = {0, 4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0};
  for (int m = 1; m <= 3; ++m) {
    hc.rkc[m] = gam[m];
    hc.rkd[m] = zet[m];
    hc.rkcv[m] = alp[m];
    hc.rkdv[m] = alp[m];
  }

  // ---- midpoint interpolation (coeft.f90:514-713) ----
  X3(hc.inxu, nx); X3(hc.inxv, nx); X3(hc.inyu, ny); X3(hc.inyv, ny);
  double l, r;
  for (int i = 2; i <= nx - 1; ++i) {
    Rows::interp_a(dxi[i], dx[i], dxi[i + 1], l, r);
    set_tri(hc.inxu, i, l, r);
    Rows::interp_b(dx[i], dx[i - 1], l, r);
    set_tri(hc.inxv, i, l, r);
  }
  for (int i = 2; i <= ny - 1; ++i) {
    Rows::interp_b(dy[i], dy[i - 1], l, r);
    set_tri(hc.inyu, i, l, r);
    Rows::interp_a(dyi[i], dy[i - 1], dyi[i - 1], l, r);
    set_tri(hc.inyv, i, l, r);
  }
  set_identity_ends(hc.inxu, nx);
  set_identity_ends(hc.inxv, nx);
  set_identity_ends(hc.inyu, ny);
  set_identity_ends(hc.inyv, ny);
  double w[5];
  closure(MIDPOINT3, dx[nx - 1], dxi[nx - 1], dxi[nx - 2], dxi[nx - 3], w);
  for (int q = 1; q <= 3; ++q) m4(hc.inbx, 1, q) = w[q - 1];
  closure(MIDPOINT3, dx[nx], dx[nx - 1], dx[nx - 2], dx[nx - 3], w);
  for (int q = 1; q <= 3; ++q) m4(hc.inbx, 2, q) = w[q - 1];
  closure(MIDPOINT4, dx[1], dx[1], dx[2], dx[3], w);
  for (int q = 1; q <= 4; ++q) m4(hc.inbx, 3, q) = w[q - 1];
  closure(MIDPOINT4, dx[0], dx[1], dxi[2], dxi[3], w);
  for (int q = 1; q <= 4; ++q) m4(hc.inbx, 4, q) = w[q - 1];
  closure(MIDPOINT4, dy[ny], dy[ny - 1], dyi[ny - 1], dyi[ny - 2], w);
  for (int q = 1; q <= 4; ++q) m4(hc.inby, 1, q) = w[q - 1];
  closure(MIDPOINT3, -dy[0], -dy[1], -dy[2], dy[4], w);
  for (int q = 1; q <= 3; ++q) m4(hc.inby, 3, q) = w[q - 1];
  closure(MIDPOINT4, dy[0], dy[1], dyi[2], dyi[3], w);
  for (int q = 1; q <= 4; ++q) m4(hc.inby, 2, q) = w[q - 1];
  closure(MIDPOINT4, -dy[ny - 1], -dy[ny - 1], -dy[ny - 2], -dy[ny - 3], w);
  for (int q = 1; q <= 4; ++q) m4(hc.inby, 4, q) = w[q - 1];
  X2(hc.cofivy, ny); X2(hc.cofiuy, ny); X2(hc.cofivx, nx); X2(hc.cofiux, nx);
  for (int i = 2; i <= ny - 1; ++i) {
    double a = dyi[i], b = dy[i - 1], c = dyi[i - 1];
    at(hc.cofivy, 2, 2, i) = (-2.0 * c * a) / (2 * a * b + b * b - 4 * c * a - 2 * c * b);
    at(hc.cofivy, 2, 1, i) = (2.0 * c * a) / (-b * b + 2 * a * b - 2 * c * b + 4 * c * a);
    a = dy[i]; b = dy[i - 1];
    at(hc.cofiuy, 2, 2, i) = 4 * b * b / (2 * b * b + a * a + 3 * a * b);
    at(hc.cofiuy, 2, 1, i) = 4 * a * a / ((a + b) * (2 * a + b));
  }
  for (int i = 2; i <= nx - 1; ++i) {
    double a = dx[i], b = dx[i - 1];
    at(hc.cofivx, 2, 2, i) = (4.0 * b * b) / (2.0 * b * b + a * a + 3.0 * a * b);
    at(hc.cofivx, 2, 1, i) = (4.0 * a * a) / ((b + a) * (2.0 * a + b));
    double c = dxi[i + 1];
    a = dxi[i]; b = dx[i];
    at(hc.cofiux, 2, 2, i) = (-2.0 * c * a) / (2 * a * b + b * b - 4 * c * a - 2 * c * b);
    at(hc.cofiux, 2, 1, i) = (2.0 * c * a) / (-b * b + 2 * a * b - 2 * c * b + 4 * c * a);
  }

  // ---- first derivative (coeft.f90:722-965) ----
  X3(hc.dcxu, nx); X3(hc.dcxv, nx); X3(hc.dcyu, ny + 1); X3(hc.dcyv, ny);
  for (int i = 2; i <= nx - 1; ++i) {
    Rows::d1_a(dx[i], dx[i - 1], l, r);
    set_tri(hc.dcxu, i, l, r);
    Rows::d1_b(dxi[i], dxi[i - 1], dx[i - 1], l, r);
    set_tri(hc.dcxv, i, l, r);
  }
  for (int i = 2; i <= ny; ++i) {
    Rows::d1_b(dyi[i], dyi[i - 1], dy[i - 1], l, r);
    set_tri(hc.dcyu, i, l, r);
  }
  for (int i = 2; i <= ny - 1; ++i) {
    Rows::d1_a(dy[i], dy[i - 1], l, r);
    set_tri(hc.dcyv, i, l, r);
  }
  set_identity_ends(hc.dcxu, nx);
  set_identity_ends(hc.dcxv, nx);
  set_identity_ends(hc.dcyu, ny + 1);
  set_identity_ends(hc.dcyv, ny);
  closure(FIRST_OFF, dx[1], dxi[2], dxi[3], dxi[4], w);
  for (int q = 1; q <= 4; ++q) m5(hc.dcbx, 1, q) = w[q - 1];
  closure(FIRST_OFF, dx[0], dx[1], dx[2], dx[3], w);
  for (int q = 1; q <= 4; ++q) m5(hc.dcbx, 2, q) = w[q - 1];
  closure(FIRST_MIX, dx[nx], dx[nx - 1], dxi[nx - 1], dxi[nx - 2], w);
  for (int q = 1; q <= 4; ++q) m5(hc.dcbx, 3, q) = w[q - 1];
  closure(FIRST_MIX, dx[nx - 1], dx[nx - 1], dx[nx - 2], dx[nx - 3], w);
  for (int q = 1; q <= 4; ++q) m5(hc.dcbx, 4, q) = w[q - 1];
  closure(FIRST_MIX, -dx[nx - 1], -dxi[nx - 1], -dxi[nx - 2], -dxi[nx - 3], w);
  for (int q = 1; q <= 4; ++q) m5(hc.dcbx, 5, q) = w[q - 1];
  closure(FIRST_MIX, -dy[0], -dy[1], -dyi[2], -dyi[3], w);
  for (int q = 1; q <= 4; ++q) m4(hc.dcby, 1, q) = w[q - 1];
  closure(FIRST_OFF, -dy[ny - 1], -dyi[ny - 1], -dyi[ny - 2], -dyi[ny - 3], w);
  for (int q = 1; q <= 4; ++q) m4(hc.dcby, 2, q) = w[q - 1];
  closure(FIRST_OFF, dy[0], dy[1], dy[2], dy[3], w);
  for (int q = 1; q <= 4; ++q) m4(hc.dcby, 3, q) = w[q - 1];
  closure(FIRST_OFF, -dy[ny], -dy[ny - 1], -dy[ny - 2], -dy[ny - 3], w);
  for (int q = 1; q <= 4; ++q) m4(hc.dcby, 4, q) = w[q - 1];
  closure(ZEROVORT7, dy[ny], dy[ny - 1], dyi[ny - 1], dyi[ny - 2], w);
  for (int q = 1; q <= 4; ++q) m2(hc.ccon, 1, q) = w[q - 1];
  closure(ZEROVORT6, -dy[0], -dyi[1], -dyi[2], 0.0, w);
  for (int q = 1; q <= 4; ++q) m2(hc.ccon, 2, q) = w[q - 1];
  X2(hc.cofcyv, ny); X2(hc.cofcyu, ny); X2(hc.cofcxv, nx); X2(hc.cofcxu, nx);
  auto c24 = [](double a, double b) {
    return 24 * a * b / (2.0 * b * b * b + 2.0 * a * a * a + 9.0 * b * b * a + 9.0 * b * a * a);
  };
  auto c12 = [](double a, double b, double c) { return 12.0 * a * b / ((12.0 * a * b - c * c) * c); };
  for (int i = 2; i <= ny - 1; ++i) {
    double v = c24(dy[i], dy[i - 1]);
    at(hc.cofcyv, 2, 1, i) = -v;
    at(hc.cofcyv, 2, 2, i) = v;
  }
  for (int i = 2; i <= ny; ++i) {
    double v = c12(dyi[i], dyi[i - 1], dy[i - 1]);
    at(hc.cofcyu, 2, 1, i) = -v;
    at(hc.cofcyu, 2, 2, i) = v;
  }
  for (int i = 2; i <= nx - 1; ++i) {
    double v = c12(dxi[i], dxi[i - 1], dx[i - 1]);
    at(hc.cofcxv, 2, 1, i) = -v;
    at(hc.cofcxv, 2, 2, i) = v;
    v = c24(dx[i], dx[i - 1]);
    at(hc.cofcxu, 2, 1, i) = -v;
    at(hc.cofcxu, 2, 2, i) = v;
  }

  // ---- second derivative (coeft.f90:971-1168) ----
  X3(hc.vixu, nx); X3(hc.vixv, nx); X3(hc.viyu, ny + 1); X3(hc.viyv, ny);
  for (int i = 2; i <= nx - 1; ++i) {
    Rows::d2(dx[i], dx[i - 1], l, r);
    set_tri(hc.vixu, i, l, r);
    Rows::d2(dxi[i], dxi[i - 1], l, r);
    set_tri(hc.vixv, i, l, r);
  }
  for (int i = 2; i <= ny; ++i) {
    Rows::d2(dyi[i], dyi[i - 1], l, r);
    set_tri(hc.viyu, i, l, r);
  }
  for (int i = 2; i <= ny - 1; ++i) {
    Rows::d2(dy[i], dy[i - 1], l, r);
    set_tri(hc.viyv, i, l, r);
  }
  // first and last rows: unit diagonal, off-diagonal from the closure
  struct End { std::vector<double> *t; int i, q; double a, b, c; double *out; int row; };
  End ends[8] = {
      {&hc.vixv, 1, 3, dxi[1], dxi[2], dxi[3], hc.cofvbx, 1},
      {&hc.vixv, nx, 1, dxi[nx - 1], dxi[nx - 2], dxi[nx - 3], hc.cofvbx, 2},
      {&hc.vixu, 1, 3, dx[1], dx[2], dx[3], hc.cofvbx, 3},
      {&hc.vixu, nx, 1, dx[nx - 1], dx[nx - 2], dx[nx - 3], hc.cofvbx, 4},
      {&hc.viyv, 1, 3, dy[1], dy[2], dy[3], hc.cofvby, 1},
      {&hc.viyv, ny, 1, dy[ny - 1], dy[ny - 2], dy[ny - 3], hc.cofvby, 2},
      {&hc.viyu, 1, 3, dyi[1], dyi[2], dyi[3], hc.cofvby, 3},
      {&hc.viyu, ny + 1, 1, dyi[ny], dyi[ny - 1], dyi[ny - 2], hc.cofvby, 4},
  };
  for (auto &e : ends) {
    at(*e.t, 3, 2, e.i) = 1.0;
    at(*e.t, 3, 4 - e.q, e.i) = 0.0; // the other off-diagonal of an end row
    closure(SECOND5, e.a, e.b, e.c, 0.0, w);
    at(*e.t, 3, e.q, e.i) = w[0];
    for (int q = 1; q <= 4; ++q) m4(e.out, e.row, q) = w[q];
  }
  // unfactored copies for the implicit y solve (coeft.f90:1116-1127)
  hc.vyui = hc.viyu;
  hc.vyvi = hc.viyv;
  X3(hc.cofvyv, ny); X3(hc.cofvyu, ny); X3(hc.cofvxv, nx); X3(hc.cofvxu, nx);
  double e1, e2, e3;
  for (int i = 2; i <= ny - 1; ++i) {
    Rows::d2e(dy[i], dy[i - 1], e1, e2, e3);
    at(hc.cofvyv, 3, 1, i) = e1; at(hc.cofvyv, 3, 2, i) = e2; at(hc.cofvyv, 3, 3, i) = e3;
  }
  for (int i = 2; i <= ny; ++i) {
    Rows::d2e(dyi[i], dyi[i - 1], e1, e2, e3);
    at(hc.cofvyu, 3, 1, i) = e1; at(hc.cofvyu, 3, 2, i) = e2; at(hc.cofvyu, 3, 3, i) = e3;
  }
  for (int i = 2; i <= nx - 1; ++i) {
    Rows::d2e(dxi[i], dxi[i - 1], e1, e2, e3);
    at(hc.cofvxv, 3, 1, i) = e1; at(hc.cofvxv, 3, 2, i) = e2; at(hc.cofvxv, 3, 3, i) = e3;
    Rows::d2e(dx[i], dx[i - 1], e1, e2, e3);
    at(hc.cofvxu, 3, 1, i) = e1; at(hc.cofvxu, 3, 2, i) = e2; at(hc.cofvxu, 3, 3, i) = e3;
  }

  // ---- Poisson: x eigenvalues and the y operator (coeft.f90:1181-1186) ----
  hc.kmod.assign(nx, 0.0);
  for (int i = 1; i <= nx - 1; ++i)
    hc.kmod[i] = 2.0 / (dx[1] * dx[1]) * (1 - std::cos((pi * (i - 1)) / (double)(nx - 1)));
  // mtr: row j couples phi(j-1), phi(j), phi(j+1) through fluxes between
  // v rows; rows 1 and ny-1 carry zero-flux ends
  X3(hc.ayp, ny);
  const int lt = ny - 1;
  double acp = 0.0, bcp = 0.0; // flux weights of the face below
  double dy1 = 0, dy2 = 0;
  for (int j = 1; j <= lt - 1; ++j) {
    dy1 = y[j + 1] - y[j];
    dy2 = y[j + 2] - y[j + 1];
    double a3 = 1.0 / dy1, b2 = -1.0 / dy1;
    double ac2 = 2.0 / (dy1 + dy2), bc2 = -2.0 / (dy1 + dy2);
    at(hc.ayp, 3, 1, j) = b2 * bcp;
    at(hc.ayp, 3, 3, j) = a3 * ac2;
    at(hc.ayp, 3, 2, j) = a3 * bc2 + b2 * acp;
    acp = ac2;
    bcp = bc2;
  }
  {
    double b2 = -1.0 / dy2;
    at(hc.ayp, 3, 1, lt) = b2 * bcp;
    at(hc.ayp, 3, 3, lt) = 0.0;
    at(hc.ayp, 3, 2, lt) = b2 * acp;
  }
  hc.vmagic.assign(nx + 1, 0.0);
  build_stats_coef(hc);
}

void inlet_setup(HostCoef &hc, const double *u0, double *rthin, double *din) {
  const int ny = hc.ny, nx = hc.nx;
  // inlet_retheta (bl_1.F90:372-396); 0.99 is a default-real literal there
  double uinf = 0.0;
  for (int j = ny + 1 - 20; j <= ny + 1 - 5; ++j) uinf += u0[j - 1];
  uinf /= 16.0;
  const double u99 = (double)0.99f * uinf;
  int jtop = 0;
  for (int j = 1; j <= ny - 2; ++j) {
    if (u0[j - 1] >= u99) break;
    jtop = j + 1;
  }
  *din = hc.y[jtop];
  double s = 0.0;
  for (int j = 2; j <= jtop + 3; ++j) s += (uinf - u0[j - 1]) * u0[j - 1] * hc.dy[j - 1];
  *rthin = s * hc.re / uinf;
  // magicnumber (boun.f90:102-143): d(delta*)/dx from a Re_theta(x) fit,
  // with its own dx = ax*pi/(nx-1)
  const double e1 = 0.8659, c1 = 0.01277, ei1 = 1 / e1, kap = 0.384, ckap = 4.127, cprim = 7.135;
  const double dxm = hc.ax * hc.pi / (nx - 1);
  std::vector<double> xx(nx + 1), h(nx + 1);
  for (int i = 1; i <= nx; ++i) {
    xx[i] = dxm * (i - 1);
    double reth = std::pow(std::pow(*rthin, ei1) + std::pow(c1, ei1) * hc.re * xx[i], e1);
    double uep = std::log(reth) / kap + ckap;
    h[i] = reth / ((1 - cprim / uep) * hc.re);
  }
  hc.vmagic[1] = (h[2] - h[1]) / (xx[2] - xx[1]);
  for (int i = 2; i <= nx - 1; ++i) hc.vmagic[i] = (h[i + 1] - h[i - 1]) / (xx[i + 1] - xx[i - 1]);
  hc.vmagic[nx] = (h[nx] - h[nx - 1]) / (xx[nx] - xx[nx - 1]);
}


// ---------------------------------------------------------------------------
// Stats-only compact schemes (the reference's coeft.f90:56-285: fd_dvdy,
// fd_dx, fd_ix, fd_iy, fd_iyp).  Every row is the scheme of maximal order on
// the reference's stencil: with the left-hand points L (weights a_l; the
// pinned point has weight 1) and the right-hand points P (weights c_q), the
// weights make
//     sum_l a_l g(L_l) = sum_q c_q f(P_q),   g = f' (derivative) or f,
// exact for every polynomial of degree below the number of free weights.
// That condition fixes the reference's closed-form weights (the same
// schemes, derived independently here).  Arrays are (8, 1:n) with column i
// at offset 8 i: rows 1-5 explicit weights, 6-8 (alpha, diagonal, beta),
// then LU-factored in place (row 6 the multiplier, row 7 the inverse pivot).
// ---------------------------------------------------------------------------
namespace {

// Gaussian elimination with partial pivoting, n <= 8 unknowns.
void solve_dense(int n, double m[8][8], double r[8]) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int q = c + 1; q < n; ++q)
      if (std::fabs(m[q][c]) > std::fabs(m[p][c])) p = q;
    if (m[p][c] == 0.0) throw std::runtime_error("singular compact-scheme moment system");
    if (p != c) {
      for (int j = 0; j < n; ++j) std::swap(m[p][j], m[c][j]);
      std::swap(r[p], r[c]);
    }
    for (int q = c + 1; q < n; ++q) {
      const double f = m[q][c] / m[c][c];
      for (int j = c; j < n; ++j) m[q][j] -= f * m[c][j];
      r[q] -= f * r[c];
    }
  }
  for (int q = n - 1; q >= 0; --q) {
    double acc = r[q];
    for (int j = q + 1; j < n; ++j) acc -= m[q][j] * r[j];
    r[q] = acc / m[q][q];
  }
}

// One row of a compact scheme about t0.  L[0..nl) are the left-hand points,
// L[pin] carries weight 1 (its g value moves to the right of the moment
// equations), the other left weights go to a (a[pin] = 1); P[0..np) the
// right-hand points, weights to c.  Distances are scaled by the stencil's
// half-width h for conditioning; a derivative weight scales back by 1/h.
void compact_row(double t0, const double *L, int nl, int pin, const double *P, int np, bool deriv, double *a,
                 double *c) {
  double h = 0.0;
  for (int q = 0; q < np; ++q) h = std::max(h, std::fabs(P[q] - t0));
  for (int l = 0; l < nl; ++l) h = std::max(h, std::fabs(L[l] - t0));
  const int n = nl - 1 + np;
  double m[8][8] = {{0}}, r[8] = {0};
  auto g = [&](double tau, int k) { // g of the monomial tau^k (scaled units)
    if (!deriv) return std::pow(tau, k);
    return k == 0 ? 0.0 : k * std::pow(tau, k - 1);
  };
  for (int k = 0; k < n; ++k) {
    int col = 0;
    for (int l = 0; l < nl; ++l) {
      const double tau = (L[l] - t0) / h;
      if (l == pin)
        r[k] = -g(tau, k);
      else
        m[k][col++] = g(tau, k);
    }
    for (int q = 0; q < np; ++q) m[k][col++] = -std::pow((P[q] - t0) / h, k);
  }
  solve_dense(n, m, r);
  int col = 0;
  for (int l = 0; l < nl; ++l) a[l] = l == pin ? 1.0 : r[col++];
  for (int q = 0; q < np; ++q) c[q] = deriv ? r[col++] / h : r[col++];
}

// Fills column i of an (8, n) array from one row: the left points are the
// rows i-1, i, i+1 present in lrow (offsets -1, 0, +1), the diagonal is the
// weight of offset 0.
void put_row(std::vector<double> &A, int i, const int *loff, int nl, const double *a, const double *c, int np) {
  for (int q = 0; q < np; ++q) A[8 * (size_t)i + q] = c[q];
  for (int l = 0; l < nl; ++l) A[8 * (size_t)i + 6 + loff[l]] = a[l]; // row 6 + off: 5 alpha, 6 diag, 7 beta
}

void factor8(std::vector<double> &A, int n, int j0) {
  auto at8 = [&](int r, int i) -> double & { return A[8 * (size_t)i + r - 1]; };
  for (int j = j0; j <= n; ++j) {
    at8(6, j) = at8(6, j) / at8(7, j - 1);
    at8(7, j) = at8(7, j) - at8(6, j) * at8(8, j - 1);
  }
  for (int j = j0 - 1; j <= n; ++j) at8(7, j) = 1.0 / at8(7, j);
}

// One row: points given by index into a coordinate array.
struct RowSpec {
  int nl, pin, np;
  int loff[3];
  int rhs[5];
};

void build_rows(std::vector<double> &A, int n, const std::vector<double> &to, const std::vector<double> &from,
                bool deriv, const RowSpec *spec, int i) {
  double L[3], P[5], a[3], c[5];
  for (int l = 0; l < spec->nl; ++l) L[l] = to[i + spec->loff[l]];
  for (int q = 0; q < spec->np; ++q) P[q] = from[spec->rhs[q]];
  compact_row(to[i], L, spec->nl, spec->pin, P, spec->np, deriv, a, c);
  put_row(A, i, spec->loff, spec->nl, a, c, spec->np);
  (void)n;
}

} // namespace

void build_stats_coef(HostCoef &hc) {
  const int nx = hc.nx, ny = hc.ny;
  const std::vector<double> &y = hc.y; // y(0:ny+1)
  hc.fd_dvdy.assign(8 * (size_t)(ny + 1), 0.0);
  hc.fd_iy.assign(8 * (size_t)(ny + 1), 0.0);
  hc.fd_iyp.assign(8 * (size_t)(ny + 1), 0.0);
  hc.fd_dx.assign(8 * (size_t)(nx + 1), 0.0);
  hc.fd_ix.assign(8 * (size_t)(nx + 1), 0.0);
  auto span = [](RowSpec &s, int first, int count) {
    s.np = count;
    for (int q = 0; q < count; ++q) s.rhs[q] = first + q;
  };
  // d/dy at the v rows y(1..ny) from the v rows (diffy_inplace,
  // lip.f90:666-698): rows 1, 2 on y(1..4), interior on y(j-2..j+2), rows
  // ny-1, ny on y(ny-3..ny)
  for (int j = 1; j <= ny; ++j) {
    RowSpec s{};
    if (j == 1) { s.nl = 2; s.loff[0] = 0; s.loff[1] = 1; s.pin = 0; span(s, 1, 4); }
    else if (j == 2) { s.nl = 3; s.loff[0] = -1; s.loff[1] = 0; s.loff[2] = 1; s.pin = 1; span(s, 1, 4); }
    else if (j <= ny - 2) { s.nl = 3; s.loff[0] = -1; s.loff[1] = 0; s.loff[2] = 1; s.pin = 1; span(s, j - 2, 5); }
    else if (j == ny - 1) { s.nl = 3; s.loff[0] = -1; s.loff[1] = 0; s.loff[2] = 1; s.pin = 1; span(s, ny - 3, 4); }
    else { s.nl = 2; s.loff[0] = -1; s.loff[1] = 0; s.pin = 1; span(s, ny - 3, 4); }
    build_rows(hc.fd_dvdy, ny, y, y, true, &s, j);
  }
  factor8(hc.fd_dvdy, ny, 2);
  // midpoint interpolation from the u rows yp(j) = (y(j-1) + y(j))/2 to the
  // v rows (interpy_new, lip.f90:771-809); u, w take ny+1 input rows, p
  // ny.  Row 1 (the wall) is explicit; the tridiagonal starts at row 2,
  // which has no alpha.  The last row pins its alpha and frees the
  // diagonal, as the reference's row does (coeft.f90:311-312).
  std::vector<double> yp(ny + 2, 0.0);
  for (int j = 1; j <= ny + 1; ++j) yp[j] = 0.5 * (y[j - 1] + y[j]);
  for (int tipo = 0; tipo < 2; ++tipo) {
    std::vector<double> &A = tipo ? hc.fd_iyp : hc.fd_iy;
    const int last = ny + 1 - tipo; // last input row
    for (int j = 1; j <= ny; ++j) {
      RowSpec s{};
      if (j == 1) { s.nl = 1; s.loff[0] = 0; s.pin = 0; span(s, 2, 4); }
      else if (j == 2) { s.nl = 2; s.loff[0] = 0; s.loff[1] = 1; s.pin = 0; span(s, 2, 5); }
      else if (j <= ny - 2 || (j == ny - 1 && tipo == 0)) {
        s.nl = 3; s.loff[0] = -1; s.loff[1] = 0; s.loff[2] = 1; s.pin = 1; span(s, j - 1, 4);
      } else if (j == ny - 1) { s.nl = 3; s.loff[0] = -1; s.loff[1] = 0; s.loff[2] = 1; s.pin = 1; span(s, ny - 2, 3); }
      else { s.nl = 2; s.loff[0] = -1; s.loff[1] = 0; s.pin = 0; span(s, ny - 2, last - (ny - 2) + 1); }
      build_rows(A, ny, y, yp, false, &s, j);
    }
    factor8(A, ny, 3);
  }
  // the uniform x grid, spacing ax pi / (nx + 1) (coeft.f90:181): d/dx at
  // the planes (diffx_inplace) with the reference's left-hand weights --
  // boundary rows f'1 + 4 f'2, (1/8) f'1 + f'2 + (3/4) f'3 and their mirror
  // images, interior (1/3, 1, 1/3) -- and explicit weights of maximal order
  const double dx = hc.ax * hc.pi / (nx + 1);
  std::vector<double> xg(nx + 2);
  for (int i = 0; i <= nx + 1; ++i) xg[i] = i * dx;
  {
    auto row_fixed = [&](int i, const double *lw, const int *loff, int nl, int first, int np) {
      // explicit weights for fixed left weights: moment conditions of order
      // 0..np-1 (a square system in the explicit weights only)
      double m[8][8] = {{0}}, r[8] = {0};
      for (int k = 0; k < np; ++k) {
        double gk = 0.0;
        for (int l = 0; l < nl; ++l) {
          const double tau = (xg[i + loff[l]] - xg[i]) / dx;
          gk += lw[l] * (k == 0 ? 0.0 : k * std::pow(tau, k - 1));
        }
        for (int q = 0; q < np; ++q) m[k][q] = std::pow((xg[first + q] - xg[i]) / dx, k);
        r[k] = gk;
      }
      solve_dense(np, m, r);
      for (int q = 0; q < np; ++q) hc.fd_dx[8 * (size_t)i + q] = r[q] / dx;
      for (int l = 0; l < nl; ++l) hc.fd_dx[8 * (size_t)i + 6 + loff[l]] = lw[l];
    };
    const int o2[2] = {0, 1}, o2m[2] = {-1, 0}, o3[3] = {-1, 0, 1};
    const double w1[2] = {1.0, 4.0}, w2[3] = {1.0 / 8.0, 1.0, 3.0 / 4.0}, wi[3] = {1.0 / 3.0, 1.0, 1.0 / 3.0};
    const double w2m[3] = {3.0 / 4.0, 1.0, 1.0 / 8.0}, w1m[2] = {4.0, 1.0};
    row_fixed(1, w1, o2, 2, 1, 5);
    row_fixed(2, w2, o3, 3, 1, 5);
    for (int i = 3; i <= nx - 2; ++i) row_fixed(i, wi, o3, 3, i - 2, 5);
    row_fixed(nx - 1, w2m, o3, 3, nx - 4, 5);
    row_fixed(nx, w1m, o2m, 2, nx - 4, 5);
    factor8(hc.fd_dx, nx, 2);
  }
  // midpoint interpolation along x from plane i to i + 1/2 (interpx_new):
  // rows 1 (beta 5/3), 2..nx-2 (3/10, 1, 3/10), nx-1 (5/6, 1, 1/14), nx
  // (7, 1; the point beyond the last plane), explicit weights of maximal
  // order on 4 points
  {
    auto row_fixed = [&](int i, const double *lw, const int *loff, int nl, int first) {
      double m[8][8] = {{0}}, r[8] = {0};
      const double t0 = xg[i] + 0.5 * dx;
      for (int k = 0; k < 4; ++k) {
        double gk = 0.0;
        for (int l = 0; l < nl; ++l) gk += lw[l] * std::pow((xg[i + loff[l]] + 0.5 * dx - t0) / dx, k);
        for (int q = 0; q < 4; ++q) m[k][q] = std::pow((xg[first + q] - t0) / dx, k);
        r[k] = gk;
      }
      solve_dense(4, m, r);
      for (int q = 0; q < 4; ++q) hc.fd_ix[8 * (size_t)i + q] = r[q];
      for (int l = 0; l < nl; ++l) hc.fd_ix[8 * (size_t)i + 6 + loff[l]] = lw[l];
    };
    const int o2[2] = {0, 1}, o2m[2] = {-1, 0}, o3[3] = {-1, 0, 1};
    const double w1[2] = {1.0, 5.0 / 3.0}, wi[3] = {3.0 / 10.0, 1.0, 3.0 / 10.0};
    const double wm[3] = {5.0 / 6.0, 1.0, 1.0 / 14.0}, wn[2] = {7.0, 1.0};
    row_fixed(1, w1, o2, 2, 1);
    for (int i = 2; i <= nx - 2; ++i) row_fixed(i, wi, o3, 3, i - 1);
    row_fixed(nx - 1, wm, o3, 3, nx - 3);
    row_fixed(nx, wn, o2m, 2, nx - 3);
    factor8(hc.fd_ix, nx, 2);
  }
}

} // namespace otbl
