// x-line compact operators (interpxx, difvisxx) with a windowed backward
// sweep.  A thread owns one line (kr, j) and marches over the planes once:
// the forward elimination y_i = r_i + m_i y_{i-1} is exact; the backward
// substitution x_i = a2_i (y_i - c_i x_{i+1}) of a block of XB planes is run
// from XW = (XK-1) XB planes further on, starting there from x = 0.  The
// x schemes live on the uniform x grid, where the backward recurrence damps
// a start error by a factor |c a2| < 0.2 per plane, so after XW planes the
// error is below round-off (the host checks every window product of the
// factored matrices before this path is enabled).  The last XK blocks of the
// line are solved exactly from plane nx.  The y values of the last XK blocks
// stay in registers (a ring of XK x XB), so every input is read once and
// every output written once, with no line scratch in HBM.
//
// Segments: with nseg > 1 a line is split into nseg runs of groups of XK
// blocks, one thread each (more threads in flight).  A segment's forward
// elimination starts one group (XLINE_LEAD planes) before its first block,
// from y = 0; the forward recurrence damps that start error by |m_i| per
// plane (the host checks every product over XLINE_LEAD planes), so the
// segment's own planes see the exact y.  Its backward windows reach XK - 1
// blocks past its last block.  Segments start on a group boundary, so the
// ring slot of a block is its index mod XK in every segment.
#include "otbl_kcommon.h"

namespace otbl {

static constexpr int XB = 8, XK = 4; // block of planes, ring length in blocks
static_assert((XK - 1) * XB == XLINE_WINDOW, "window length");
static_assert(XK * XB == XLINE_EXACT, "lines of at most XK blocks take the exact sweep alone");
static_assert(XK * XB == XLINE_LEAD, "a segment's forward lead-in is one group of XK blocks");

// block range of segment sg of nseg over nb blocks (whole groups of XK
// blocks): emitted blocks [bs0, bs1), swept blocks [qs, qe], qs a multiple of
// XK
struct XSeg {
  int bs0, bs1, qs, qe;
  __device__ XSeg(int sg, int nseg, int nb) {
    const int ng = (nb + XK - 1) / XK;
    bs0 = (int)((long)sg * ng / nseg) * XK;
    bs1 = min((int)((long)(sg + 1) * ng / nseg) * XK, nb);
    qs = bs0 > XK ? bs0 - XK : 0;
    qe = min(bs1 + XK - 2, nb - 1);
  }
  __device__ bool emits(int e) const { return e >= bs0 && e < bs1; }
};

// interpxx (lip.f90:248-300): midpoint interpolation along x
template <typename T>
__global__ void __launch_bounds__(TPB)
interpxx_w(const T *__restrict__ gen, T *__restrict__ geni, const T *__restrict__ icon,
           const T *__restrict__ econ, const T *__restrict__ inbx, int chi, int chi2, int nx, size_t P, int nseg) {
  // the segment is a function of the block (uniform): the plane loop and
  // its coefficient indices stay scalar
  const int nbp = (int)((P + TPB - 1) / TPB);
  const int bk = (int)xcd_block(); // XCD-contiguous: neighbouring blocks share lines
  const int sg = bk / nbp;
  const size_t l = (size_t)(bk - sg * nbp) * TPB + threadIdx.x;
  if (l >= P) return;
  const int nb = (nx + XB - 1) / XB;
  const XSeg sgm(sg, nseg, nb);
  const T *g = gen + l;
  T *o = geni + l;
  T ring[XK][XB];
  T y = 0;
  // g(i) of the interior plane i being processed: from i = 2, or from the
  // first plane of the segment's lead-in
  T gi = sgm.qs > 0 ? g[(size_t)(sgm.qs * XB) * P] : g[P];
  // the loads of the next block when it is interior, issued one block ahead
  T pre[XB];
  bool hp = false;
  for (int it = sgm.qs; it <= sgm.qe; it += XK) {
#pragma unroll
    for (int kk = 0; kk < XK; ++kk) {
      const int qq = it + kk;
      if (qq > sgm.qe) break;
      // forward elimination over block qq; a block of interior planes
      // issues its XB loads before any use (one basic block)
      const int i0 = qq * XB + 1;
      if (i0 >= 2 && i0 + XB - 1 <= nx - 1) {
        T gn[XB];
        if (hp) {
#pragma unroll
          for (int b = 0; b < XB; ++b) gn[b] = pre[b];
        } else {
#pragma unroll
          for (int b = 0; b < XB; ++b) gn[b] = g[(size_t)(i0 + b) * P]; // g(i+1)
        }
        // (FP32 only: FP64 holds no spare registers for it)
        hp = sizeof(T) == 4 && qq + 1 <= sgm.qe && i0 + 2 * XB - 1 <= nx - 1;
        if (hp) {
#pragma unroll
          for (int b = 0; b < XB; ++b) pre[b] = g[(size_t)(i0 + XB + b) * P];
        }
#pragma unroll
        for (int b = 0; b < XB; ++b) {
          const int i = i0 + b;
          const T rhs = C2(econ, 1, i) * gi + C2(econ, 2, i) * gn[b];
          gi = gn[b];
          y = rhs + C3(icon, 1, i) * y;
          ring[kk][b] = y;
        }
      } else
#pragma unroll
      for (int b = 0; b < XB; ++b) {
        const int i = qq * XB + b + 1;
        if (i <= nx) {
          T rhs;
          if (i == 1) {
            rhs = M4(inbx, chi2, 1) * g[0] + M4(inbx, chi2, 2) * g[P] + M4(inbx, chi2, 3) * g[2 * P] +
                  M4(inbx, chi2, 4) * g[3 * P];
            y = rhs;
          } else if (i < nx) {
            const T gn = g[(size_t)i * P]; // g(i+1)
            rhs = C2(econ, 1, i) * gi + C2(econ, 2, i) * gn;
            gi = gn;
            y = rhs + C3(icon, 1, i) * y;
          } else {
            rhs = M4(inbx, chi, 1) * g[(size_t)(nx - 1) * P] + M4(inbx, chi, 2) * g[(size_t)(nx - 2) * P] +
                  M4(inbx, chi, 3) * g[(size_t)(nx - 3) * P] + M4(inbx, chi, 4) * g[(size_t)(nx - 4) * P];
            y = rhs + C3(icon, 1, nx) * y;
          }
          ring[kk][b] = y;
        }
      }
      if (qq < nb - 1) {
        if (qq >= XK - 1 && sgm.emits(qq - (XK - 1))) {
          // windowed backward sweep: blocks qq .. qq-XK+1, emit the oldest
          T x = 0;
#pragma unroll
          for (int d = 0; d < XK; ++d) {
            const int slot = (kk - d + XK) % XK;
#pragma unroll
            for (int b = XB - 1; b >= 0; --b) {
              const int i = (qq - d) * XB + b + 1;
              x = (ring[slot][b] - C3(icon, 3, i) * x) * C3(icon, 2, i);
              if (d == XK - 1) o[(size_t)(i - 1) * P] = x;
            }
          }
        }
      } else {
        // last block: exact backward sweep from plane nx over the blocks
        // not emitted yet (those of this segment)
        T x = 0;
        bool started = false;
#pragma unroll
        for (int d = 0; d < XK; ++d) {
          const int slot = (kk - d + XK) % XK;
          if (qq - d >= sgm.qs) {
#pragma unroll
            for (int b = XB - 1; b >= 0; --b) {
              const int i = (qq - d) * XB + b + 1;
              if (i <= nx) {
                const T yy = ring[slot][b];
                x = started ? (yy - C3(icon, 3, i) * x) * C3(icon, 2, i) : yy * C3(icon, 2, nx);
                started = true;
                if (sgm.emits(qq - d)) o[(size_t)(i - 1) * P] = x;
              }
            }
          }
        }
      }
    }
  }
}

// difvisxx (lip.f90:52-142): convective d/dx of the product plus rex times
// the compact d2/dx2 of the velocity; mode 1 adds to out except on plane nx,
// which gets the outflow derivative U_inf du/dx (lip.f90:105,132).
template <typename T>
__global__ void __launch_bounds__(TPB)
difvisxx_w(const T *__restrict__ gp, const T *__restrict__ q, T *out,
           const T *__restrict__ icon, const T *__restrict__ iconv, const T *__restrict__ econ,
           const T *__restrict__ econv, const T *__restrict__ dcbx, const T *__restrict__ cofvbx,
           int chv, int ic, int nx, size_t P, double rex_d, double idx_d, int mode, int nseg) {
  // the segment is a function of the block (uniform): the plane loop and
  // its coefficient indices stay scalar
  const int nbp = (int)((P + TPB - 1) / TPB);
  const int bk = (int)xcd_block(); // XCD-contiguous: neighbouring blocks share lines
  const int sg = bk / nbp;
  const size_t l = (size_t)(bk - sg * nbp) * TPB + threadIdx.x;
  if (l >= P) return;
  const int nb = (nx + XB - 1) / XB;
  const XSeg sgm(sg, nseg, nb);
  const T *g = gp + l, *v = q + l;
  T *o = out + l;
  const int chv1 = chv + 2;
  const T rex = (T)rex_d, idx = (T)idx_d;
  T rc_[XK][XB], rv_[XK][XB];
  T yc = 0, yv = 0;
  // g(i-1), q(i-1), q(i) before the first interior plane i processed (2, or
  // the first plane of the segment's lead-in)
  const int i0s = sgm.qs > 0 ? sgm.qs * XB + 1 : 2;
  T gim1 = g[(size_t)(i0s - 2) * P], qim1 = v[(size_t)(i0s - 2) * P], qi = v[(size_t)(i0s - 1) * P];
  for (int it = sgm.qs; it <= sgm.qe; it += XK) {
#pragma unroll
    for (int kk = 0; kk < XK; ++kk) {
      const int qq = it + kk;
      if (qq > sgm.qe) break;
      const int i0 = qq * XB + 1;
      if (i0 >= 2 && i0 + XB - 1 <= nx - 1) {
        // interior block: its 2 XB loads issued before any use
        T gb[XB], qb[XB];
#pragma unroll
        for (int b = 0; b < XB; ++b) {
          gb[b] = g[(size_t)(i0 + b - 1) * P]; // g(i)
          qb[b] = v[(size_t)(i0 + b) * P];     // q(i+1)
        }
#pragma unroll
        for (int b = 0; b < XB; ++b) {
          const int i = i0 + b;
          const T rc = C2(econ, 1, i) * gim1 + C2(econ, 2, i) * gb[b];
          const T rv = C3(econv, 1, i) * qb[b] + C3(econv, 2, i) * qi + C3(econv, 3, i) * qim1;
          yc = rc + C3(icon, 1, i) * yc;
          yv = rv + C3(iconv, 1, i) * yv;
          gim1 = gb[b];
          qim1 = qi;
          qi = qb[b];
          rc_[kk][b] = yc;
          rv_[kk][b] = yv;
        }
      } else
#pragma unroll
      for (int b = 0; b < XB; ++b) {
        const int i = qq * XB + b + 1;
        if (i <= nx) {
          if (i == 1) {
            yc = M5(dcbx, chv, 1) * g[0] + M5(dcbx, chv, 2) * g[P] + M5(dcbx, chv, 3) * g[2 * P] +
                 M5(dcbx, chv, 4) * g[3 * P];
            yv = M4(cofvbx, ic, 1) * v[0] + M4(cofvbx, ic, 2) * v[P] + M4(cofvbx, ic, 3) * v[2 * P] +
                 M4(cofvbx, ic, 4) * v[3 * P];
          } else if (i < nx) {
            const T gi = g[(size_t)(i - 1) * P];
            const T qn = v[(size_t)i * P];
            const T rc = C2(econ, 1, i) * gim1 + C2(econ, 2, i) * gi;
            const T rv = C3(econv, 1, i) * qn + C3(econv, 2, i) * qi + C3(econv, 3, i) * qim1;
            yc = rc + C3(icon, 1, i) * yc;
            yv = rv + C3(iconv, 1, i) * yv;
            gim1 = gi;
            qim1 = qi;
            qi = qn;
          } else {
            const T gn = g[(size_t)(nx - 1) * P], gn1 = g[(size_t)(nx - 2) * P],
                         gn2 = g[(size_t)(nx - 3) * P], gn3 = g[(size_t)(nx - 4) * P];
            const T qn2 = v[(size_t)(nx - 3) * P], qn3 = v[(size_t)(nx - 4) * P];
            const T rc = M5(dcbx, chv1, 1) * gn + M5(dcbx, chv1, 2) * gn1 + M5(dcbx, chv1, 3) * gn2 +
                              M5(dcbx, chv1, 4) * gn3;
            const T rv = M4(cofvbx, ic + 1, 1) * qi + M4(cofvbx, ic + 1, 2) * qim1 +
                              M4(cofvbx, ic + 1, 3) * qn2 + M4(cofvbx, ic + 1, 4) * qn3;
            yc = rc + C3(icon, 1, nx) * yc;
            yv = rv + C3(iconv, 1, nx) * yv;
          }
          rc_[kk][b] = yc;
          rv_[kk][b] = yv;
        }
      }
      if (qq < nb - 1) {
        if (qq >= XK - 1 && sgm.emits(qq - (XK - 1))) {
          // mode 1: the emitted block's previous values, loaded before the
          // sweep (a load after a store to the same array would wait for it)
          T prev[XB];
          const int ie = (qq - (XK - 1)) * XB; // 0-based first plane emitted
#pragma unroll
          for (int b = 0; b < XB; ++b) prev[b] = mode == 1 ? o[(size_t)(ie + b) * P] : T(0);
          T xc = 0, xv = 0;
#pragma unroll
          for (int d = 0; d < XK; ++d) {
            const int slot = (kk - d + XK) % XK;
#pragma unroll
            for (int b = XB - 1; b >= 0; --b) {
              const int i = (qq - d) * XB + b + 1;
              xc = (rc_[slot][b] - C3(icon, 3, i) * xc) * C3(icon, 2, i);
              xv = (rv_[slot][b] - C3(iconv, 3, i) * xv) * C3(iconv, 2, i);
              if (d == XK - 1) o[(size_t)(i - 1) * P] = prev[b] + (xc + rex * xv);
            }
          }
        }
      } else {
        T xc = 0, xv = 0;
        bool started = false;
#pragma unroll
        for (int d = 0; d < XK; ++d) {
          const int slot = (kk - d + XK) % XK;
          if (qq - d >= sgm.qs) {
            const bool em = sgm.emits(qq - d);
            T prev[XB];
#pragma unroll
            for (int b = 0; b < XB; ++b) {
              const int i = (qq - d) * XB + b + 1;
              prev[b] = (mode == 1 && em && i <= nx) ? o[(size_t)(i - 1) * P] : T(0);
            }
#pragma unroll
            for (int b = XB - 1; b >= 0; --b) {
              const int i = (qq - d) * XB + b + 1;
              if (i <= nx) {
                const size_t oi = (size_t)(i - 1) * P;
                if (!started) {
                  xc = rc_[slot][b] * C3(icon, 2, nx);
                  xv = rv_[slot][b] * C3(iconv, 2, nx);
                  // outflow row: the velocity difference qi - qim1 = q(nx) - q(nx-1)
                  if (em) o[oi] = T(1) * (qi - qim1) * idx + rex * T(0);
                  started = true;
                } else {
                  xc = (rc_[slot][b] - C3(icon, 3, i) * xc) * C3(icon, 2, i);
                  xv = (rv_[slot][b] - C3(iconv, 3, i) * xv) * C3(iconv, 2, i);
                  if (em) o[oi] = prev[b] + (xc + rex * xv);
                }
              }
            }
          }
        }
      }
    }
  }
}

template <typename T>
void kx_interpxx(Box &b, const T *gen, T *geni, const T *icon, const T *econ, int chi, int chi2, size_t P) {
  auto kf = &interpxx_w<T>;
  LAUNCH(b, "interpxx", (unsigned)(nblk(P) * b.xseg), TPB, 0, kf, gen, geni, icon, econ, coefs<T>(b).inbx, chi, chi2, b.nx, P,
         b.xseg);
}

template <typename T>
void kx_difvisxx(Box &b, const T *prod, const T *vel, T *out, const T *icon, const T *iconv, const T *econ,
                 const T *econv, int chv, int ic, size_t P, int mode) {
  auto kf = &difvisxx_w<T>;
  LAUNCH(b, "difvisxx", (unsigned)(nblk(P) * b.xseg), TPB, 0, kf, prod, vel, out, icon, iconv, econ, econv, coefs<T>(b).dcbx,
         coefs<T>(b).cofvbx, chv, ic, b.nx, P, b.hc.rex, b.hc.idx, mode, b.xseg);
}

template void kx_interpxx<double>(Box &, const double *, double *, const double *, const double *, int, int, size_t);
template void kx_interpxx<float>(Box &, const float *, float *, const float *, const float *, int, int, size_t);
template void kx_difvisxx<double>(Box &, const double *, const double *, double *, const double *, const double *,
                                  const double *, const double *, int, int, size_t, int);
template void kx_difvisxx<float>(Box &, const float *, const float *, float *, const float *, const float *,
                                 const float *, const float *, int, int, size_t, int);

} // namespace otbl
