// Chunked y-line solver: constants, the table builder shared by the kernel
// translation units, and the launchers of otbl_ycol.hip.
#pragma once
#include "otbl_kcommon.h"

namespace otbl {

static constexpr int YNCH = 16;           // chunks (waves) per column
static constexpr int YTPB = 64 * YNCH;    // threads per workgroup

// Table of one chunked y system, YNCH*L rows (NL) in chunk order:
//   M, A2, B   factored tridiagonal: y_j = r_j + M_j y_{j-1},
//              x_j = A2_j y_j + B_j x_{j+1}
//   G, S       responses of a chunk to unit forward / backward carries
//   E1, E2, E3 the explicit 3-point stencil whose result the system solves
//              for: rhs_j = E1 g(j-1) + E2 g(j) + E3 g(j+1) (interior rows)
//   PE, G0, S0 per chunk: carry gain (product of M), G and S at the start
__host__ __device__ constexpr size_t ytab_size(int L) { return (size_t)8 * YNCH * L + 3 * YNCH; }

// rows per chunk for a box with ny (u columns have ny+1 rows); 0 when the
// columns are too long for the instantiated chunk lengths
inline int ytab_rows(int ny) {
  const int need = (ny + 1 + YNCH - 1) / YNCH;
  for (int L : {4, 8, 12, 16, 20, 24})
    if (L >= need) return L;
  return 0;
}

// tables in Box::ytab, in this order
enum { YT_INYU = 0, YT_INYV, YT_DCYU, YT_DCYV, YT_VIYU, YT_VIYV, YT_IMPU, YT_IMPV, YT_COUNT };

struct TriRow {
  double m, a2, c;
};
struct Sten3 {
  double e1, e2, e3;
};

// Builds the table of a system of n rows from 1-based accessors: row(j) the
// factored tridiagonal row, sten(j) the stencil of interior row j
// (2 <= j <= n-1; zero elsewhere).  Rows beyond n are zero, so a partly
// empty last chunk needs no special case.
template <class F, class E>
__device__ inline void ytab_build(double *tab, int n, int L, F row, E sten) {
  const int NL = YNCH * L;
  double *M = tab, *A2 = tab + NL, *B = tab + 2 * NL, *G = tab + 3 * NL, *S = tab + 4 * NL;
  double *E1 = tab + 5 * NL, *E2 = tab + 6 * NL, *E3 = tab + 7 * NL;
  double *PE = tab + 8 * NL, *G0 = PE + YNCH, *S0 = G0 + YNCH;
  for (int j = 0; j < NL; ++j) {
    if (j < n) {
      const TriRow t = row(j + 1);
      M[j] = j == 0 ? 0.0 : t.m;
      A2[j] = t.a2;
      B[j] = j == n - 1 ? 0.0 : -t.c * t.a2;
    } else {
      M[j] = A2[j] = B[j] = 0.0;
    }
    const Sten3 e = (j >= 1 && j <= n - 2) ? sten(j + 1) : Sten3{0.0, 0.0, 0.0};
    E1[j] = e.e1;
    E2[j] = e.e2;
    E3[j] = e.e3;
  }
  for (int q = 0; q < YNCH; ++q) {
    double P = 1.0;
    for (int r = 0; r < L; ++r) {
      P *= M[q * L + r];
      G[q * L + r] = P; // forward response to a unit carry-in
    }
    PE[q] = P;
    double g = 0.0, sa = 1.0;
    for (int r = L - 1; r >= 0; --r) {
      const int j = q * L + r;
      g = A2[j] * G[j] + B[j] * g;
      G[j] = g;
      sa = B[j] * sa;
      S[j] = sa;
    }
    G0[q] = G[q * L];
    S0[q] = S[q * L];
  }
}

// The same table built by a whole workgroup: rows by all threads, the
// chunk recurrences by one thread per chunk.  sh: LDS scratch of 4*YNCH*L
// doubles.  Ends with the workgroup synchronised.
template <class F, class E>
__device__ inline void ytab_build_block(double *tab, int n, int L, F row, E sten, double *sh) {
  const int NL = YNCH * L;
  double *M = sh, *A2 = sh + NL, *B = sh + 2 * NL, *P = sh + 3 * NL;
  for (int j = threadIdx.x; j < NL; j += blockDim.x) {
    double m = 0.0, a2 = 0.0, b = 0.0;
    if (j < n) {
      const TriRow t = row(j + 1);
      m = j == 0 ? 0.0 : t.m;
      a2 = t.a2;
      b = j == n - 1 ? 0.0 : -t.c * t.a2;
    }
    const Sten3 e = (j >= 1 && j <= n - 2) ? sten(j + 1) : Sten3{0.0, 0.0, 0.0};
    M[j] = tab[j] = m;
    A2[j] = tab[NL + j] = a2;
    B[j] = tab[2 * NL + j] = b;
    tab[5 * NL + j] = e.e1;
    tab[6 * NL + j] = e.e2;
    tab[7 * NL + j] = e.e3;
  }
  __syncthreads();
  if (threadIdx.x < YNCH) {
    const int q = threadIdx.x;
    double *G = tab + 3 * NL, *S = tab + 4 * NL, *PE = tab + 8 * NL, *G0 = PE + YNCH, *S0 = G0 + YNCH;
    double pr = 1.0;
    for (int r = 0; r < L; ++r) {
      pr *= M[q * L + r];
      P[q * L + r] = pr;
    }
    PE[q] = pr;
    double g = 0.0, sa = 1.0;
    for (int r = L - 1; r >= 0; --r) {
      const int j = q * L + r;
      g = A2[j] * P[j] + B[j] * g;
      sa = B[j] * sa;
      G[j] = g;
      S[j] = sa;
      if (r == 0) {
        G0[q] = g;
        S0[q] = sa;
      }
    }
  }
  __syncthreads();
}

// Half-wave chunking of the same systems: YHCH chunks of L/2 rows (the
// same YNCH*L rows), one per half-wavefront of a workgroup of YTPB threads
// that takes 32 columns.
static constexpr int YHCH = 2 * YNCH;
__host__ __device__ constexpr size_t ytabh_size(int L2) { return (size_t)8 * YHCH * L2 + 3 * YHCH; }

// FP32 table of the half-wave chunking from the FP64 table src of the
// YNCH-chunk layout with L rows per chunk (L even, at most 24): the row
// arrays M, A2, B, E1..E3 are the same rows; the chunk responses G, S and
// PE, G0, S0 are recomputed (in FP64) for chunks of L/2 rows.  Run by a
// whole workgroup after src is complete; no synchronisation at the end.
__device__ inline void ytab_halve_block(float *dst, const double *src, int L) {
  const int NL = YNCH * L, L2 = L / 2;
  for (int j = threadIdx.x; j < NL; j += blockDim.x) {
    dst[j] = (float)src[j];
    dst[NL + j] = (float)src[NL + j];
    dst[2 * NL + j] = (float)src[2 * NL + j];
    dst[5 * NL + j] = (float)src[5 * NL + j];
    dst[6 * NL + j] = (float)src[6 * NL + j];
    dst[7 * NL + j] = (float)src[7 * NL + j];
  }
  if (threadIdx.x < YHCH) {
    const int q = threadIdx.x;
    const double *M = src, *A2 = src + NL, *B = src + 2 * NL;
    float *G = dst + 3 * NL, *S = dst + 4 * NL, *PE = dst + 8 * NL, *G0 = PE + YHCH, *S0 = G0 + YHCH;
    double pr[12], acc = 1.0; // L2 <= 12: unrolled, in registers
#pragma unroll
    for (int r = 0; r < 12; ++r)
      if (r < L2) {
        acc *= M[q * L2 + r];
        pr[r] = acc;
      }
    PE[q] = (float)acc;
    double g = 0.0, sa = 1.0;
#pragma unroll
    for (int r = 11; r >= 0; --r)
      if (r < L2) {
        const int j = q * L2 + r;
        g = A2[j] * pr[r] + B[j] * g;
        sa = B[j] * sa;
        G[j] = (float)g;
        S[j] = (float)sa;
      }
    G0[q] = (float)g;
    S0[q] = (float)sa;
  }
}

// matrices factored by coefinv (coeft.f90:1338-1353): a(1,j) multiplier,
// a(2,j) inverse pivot, a(3,j) super-diagonal
__device__ inline TriRow coefinv_row(const double *a, int j) {
  return TriRow{C3(a, 1, j), C3(a, 2, j), C3(a, 3, j)};
}

// Table of the implicit y-viscous system (implzy, lip.f90:528-597) factored
// by impl_factor: rows 2..jee-1 as factored, rows 1 and jee identity rows
// carrying the boundary values; stencil: the icon rows (vyui / vyvi).
__device__ inline void ytab_impl(double *tab, const double *f, const double *icon, int jee, int L) {
  const double *a1 = f, *a2 = f + (jee + 1), *a3 = f + 2 * (jee + 1);
  ytab_build(
      tab, jee, L,
      [=](int j) {
        if (j == 1 || j == jee) return TriRow{0.0, 1.0, 0.0};
        return TriRow{a1[j], a2[j], a3[j]};
      },
      [=](int j) { return Sten3{C3(icon, 1, j), C3(icon, 2, j), C3(icon, 3, j)}; });
}

// launchers (otbl_ycol.hip); valid when Box::yL > 0
void k_ytab_fixed(Box &b);
template <typename T>
void ky_interpyy(Box &b, const T *gen, T *geni, int loguv, int i0, int nplanes, size_t ips, size_t ops);
template <typename T> void ky_differzy(Box &b, const T *geni, T *res, int which, int nplanes, size_t ips, size_t ops);
template <typename T> void ky_vistyy(Box &b, const T *u, T *res, int R, int nplanes);
// the implicit factors (implu / implv) stay FP64; the table of the
// factored system comes in the precision of the fields
template <typename T> void ky_implzy(Box &b, T *u, const T *res, int R, int nplanes);
template <typename T> void ky_rhsu_loop1(Box &b, const T *uv, const T *uw, T *rhsu);
template <typename T>
void ky_rhs_loop2(Box &b, T *rhsu, T *rhsv, T *rhsw, const T *resu, const T *resv, const T *resw, const T *vw,
                  const T *vv, const T *ww);
template <typename T> void ky_rk_final(Box &b, T *u, const T *res, const T *rhs, const T *rhsp, int jee);
// plane loop 2's y terms, boun's rows and rk_final of one component in one
// pass (k2 = 1 v, 2 w, 3 u); pa, pb: the loop's products (v: vw, vv; w: ww,
// vw; u: none)
template <typename T>
void ky_rk_tail(Box &b, int k2, T *u, T *rhs, const T *res, const T *rhsp, const T *pa, const T *pb);
// boun's boundary rows of the FP64 mean-flow u of the mixed substep
void k_boun_mean(Box &b);
// FP32 tables of the half-wave chunking of the six fixed systems
// (Box::ytabh_f; the implicit systems' are written with each factorisation)
void k_ytab_halve_fixed(Box &b);
// rk_final of the FP64 mean-flow u (kz = 0 real part) of the mixed substep
template <typename T> void ky_mean_rk_final(Box &b, double *um, const double *rm, const T *rhs, const T *rhsp);

} // namespace otbl
