// y-line kernels with each column split over the waves of a workgroup.
//
// A workgroup handles 64 columns (consecutive flattened (kr, plane) indices)
// with YNCH = 16 waves; wave p keeps rows p*L+1 .. p*L+L of each of its
// columns in registers (L rows per chunk, a template parameter).  Every wave
// access reads or writes 64 consecutive doubles of one row and every input
// row is read once: no column scratch goes through HBM.
//
// A factored tridiagonal system along the column (forward
// y_j = r_j + m_j y_{j-1}, backward x_j = a2_j y_j + b_j x_{j+1}) is solved
// chunk by chunk with zero carries, then corrected by a partition (spike)
// step:  x = x_local + C g + D s,  where g is the chunk's response to a unit
// forward carry-in (the y of the row above the chunk) and s its response to a
// unit backward carry-in (the x of the row below the chunk).  g and s depend
// only on the factored matrix and on the chunking, so they are tabulated once
// per matrix (once per substep for the dt-dependent implicit operator).  The
// carries C and D of each chunk follow from the chunk end values by two
// recurrences over the 16 chunks, exchanged through LDS.  The explicit stencil
// coefficients of each operator are stored in its system's table.  The
// kernels read the tables in place: a wave's table indices are uniform, so
// they are scalar loads into scalar registers (FP32 rk_final, which measured
// faster that way, stages its two tables in LDS).
#include "otbl_kcommon.h"
#include "otbl_ycol.h"

namespace otbl {

// the six fixed y systems with their stencils: inyu/cofiuy and inyv/cofivy
// (interpyy), dcyu/cofcyu and dcyv/cofcyv (differzy), viyu/cofvyu and
// viyv/cofvyv (vistyy)
__global__ void ytab_fixed_kernel(double *tabs, DevCoef c, int ny, int L) {
  const int t = threadIdx.x;
  if (t >= 6) return;
  const double *a[6] = {c.inyu, c.inyv, c.dcyu, c.dcyv, c.viyu, c.viyv};
  const int n[6] = {ny, ny, ny + 1, ny, ny + 1, ny};
  const double *m = a[t];
  auto row = [m](int j) { return coefinv_row(m, j); };
  double *tab = tabs + (size_t)t * ytab_size(L);
  switch (t) {
  case 0: // G(j), G(j+1)
    ytab_build(tab, n[t], L, row, [&](int j) { return Sten3{0.0, C2(c.cofiuy, 1, j), C2(c.cofiuy, 2, j)}; });
    break;
  case 1: // G(j-1), G(j)
    ytab_build(tab, n[t], L, row, [&](int j) { return Sten3{C2(c.cofivy, 1, j), C2(c.cofivy, 2, j), 0.0}; });
    break;
  case 2: // G(j-1), G(j)
    ytab_build(tab, n[t], L, row, [&](int j) { return Sten3{C2(c.cofcyu, 1, j), C2(c.cofcyu, 2, j), 0.0}; });
    break;
  case 3: // G(j), G(j+1)
    ytab_build(tab, n[t], L, row, [&](int j) { return Sten3{0.0, C2(c.cofcyv, 1, j), C2(c.cofcyv, 2, j)}; });
    break;
  case 4:
    ytab_build(tab, n[t], L, row, [&](int j) { return Sten3{C3(c.cofvyu, 3, j), C3(c.cofvyu, 2, j), C3(c.cofvyu, 1, j)}; });
    break;
  default:
    ytab_build(tab, n[t], L, row, [&](int j) { return Sten3{C3(c.cofvyv, 3, j), C3(c.cofvyv, 2, j), C3(c.cofvyv, 1, j)}; });
    break;
  }
}

void k_ytab_fixed(Box &b) {
  LAUNCH(b, "ytab_fixed", 1, 64, 0, ytab_fixed_kernel, b.ytab, b.dc, b.ny, b.yL);
}

// half-wave chunked FP32 tables of the fixed systems, one per block
__global__ void ytab_halve_kernel(float *dst, const double *src, int L) {
  ytab_halve_block(dst + blockIdx.x * ytabh_size(L / 2), src + blockIdx.x * ytab_size(L), L);
}

void k_ytab_halve_fixed(Box &b) {
  static_assert(YT_INYU == 0 && YT_VIYV == 5, "the fixed systems lead the table order");
  LAUNCH(b, "ytab_halve", YT_VIYV + 1, 256, 0, ytab_halve_kernel, b.ytabh_f, b.ytab, b.yL);
}


// ---------------------------------------------------------------------------
// per-wave column helpers
// ---------------------------------------------------------------------------
struct YCol {
  int p, lane, kr;
  bool ok;
  size_t ip; // 0-based plane
};

__device__ inline YCol ycol(size_t ncols, int nz1, size_t ip0) {
  YCol y;
  y.lane = threadIdx.x & 63;
  y.p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  size_t c = (size_t)xcd_block() * 64 + y.lane;
  y.ok = c < ncols;
  if (!y.ok) c = 0; // a dead lane addresses column 0 of the first plane
  y.kr = (int)(c % nz1);
  y.ip = ip0 + c / nz1;
  return y;
}

__device__ __forceinline__ int yclamp(int j, int n) { return j < 0 ? 0 : (j >= n ? n - 1 : j); }

// An opaque copy of a wave-uniform value: row offsets built from it in one
// phase of a kernel are not shared with (and kept live from) another phase.
__device__ __forceinline__ size_t yopaque(size_t x) {
  asm volatile("" : "+s"(x));
  return x;
}

// Scheduling fence: the unrolled sweeps are split into groups of YGRP rows
// so the compiler cannot hoist every load and coefficient read of a sweep
// ahead of its dependency chain (which runs out of registers and spills).
#define YFENCE() __builtin_amdgcn_sched_barrier(0)
static constexpr int YGRP = 4;

// the largest divisor of L that is at most g: rows per load group
__host__ __device__ constexpr int yrows_group(int L, int g) {
  for (int k = g; k > 1; --k)
    if (L % k == 0) return k;
  return 1;
}

// Copy one system's table into LDS (all threads); the caller synchronises.
template <int L, typename T>
__device__ __forceinline__ void ytab_stage(T *dst, const T *__restrict__ src) {
  for (int k = threadIdx.x; k < (int)ytab_size(L); k += YTPB) dst[k] = src[k];
}

// sum_k cf(row, k) * G(j1 + dir*(k-1)), k = 1..4, over 1-based rows of col:
// the one-sided closures of the compact y operators
template <typename T>
__device__ __forceinline__ T bnd4(const T *cf, int ld, int row, const T *col, size_t s, int j1, int dir) {
  T r = 0;
#pragma unroll
  for (int k = 1; k <= 4; ++k) r += cf[ld * (k - 1) + row - 1] * col[(size_t)(j1 + dir * (k - 1) - 1) * s];
  return r;
}

// Right-hand side of one chunk fused with the forward local sweep.  fetch(k)
// returns input row j0-1+k (k = 0..L+1, 0-based rows, j0 = p*L); interior
// rows take the table's 3-point stencil, the system's first row takes `bot`
// and its last row (n-1) `top`; add1 and addm are added to rows 1 and n-2
// (the boundary columns of implzy).  Rows past n have a zero stencil.
// Returns y at the chunk's last row (its forward carry-out).
// G: rows per fenced group (a group's input loads are issued together);
// the FP32 interpolation and rhsu kernels gain from whole-chunk groups, the
// kernels with more live state lose.
template <int L, int G = YGRP, int NCH = YNCH, typename T, class Fetch>
__device__ __forceinline__ T yfwd(T (&v)[L], Fetch fetch, const T *tab, int p, int n, T bot, T top, T add1, T addm) {
  constexpr int NL = NCH * L;
  const T *M = tab + p * L, *E1 = tab + 5 * NL + p * L, *E2 = tab + 6 * NL + p * L, *E3 = tab + 7 * NL + p * L;
  const int j0 = p * L;
  T gm = fetch(0), g0 = fetch(1), y = 0;
#pragma unroll
  for (int r = 0; r < L; ++r) {
    if (r % G == 0) YFENCE();
    const int j = j0 + r;
    const T gp = fetch(r + 2);
    T x = E1[r] * gm + E2[r] * g0 + E3[r] * gp;
    x = j == 0 ? bot : x;
    x = j == n - 1 ? top : x;
    x = j == 1 ? x + add1 : x;
    x = j == n - 2 ? x + addm : x;
    y = x + M[r] * y;
    v[r] = y;
    gm = g0;
    g0 = gp;
  }
  YFENCE();
  return y;
}

// Backward local sweep, carry exchange between the chunks of each column
// and the spike correction: on return v holds the solution of the chunk.
// xch: LDS, 2*NCH*CW values (NCH chunks of CW columns).
template <int L, int NCH = YNCH, int CW = 64, typename T>
__device__ __forceinline__ void ybwd(T (&v)[L], T yend, const T *tab, int p, int lane, T *xch) {
  constexpr int NL = NCH * L;
  const T *A2 = tab + NL + p * L, *B = tab + 2 * NL + p * L;
  const T *G = tab + 3 * NL + p * L, *S = tab + 4 * NL + p * L;
  const T *PE = tab + 8 * NL, *G0 = PE + NCH, *S0 = G0 + NCH;
  v[L - 1] = A2[L - 1] * v[L - 1];
#pragma unroll
  for (int r = L - 2; r >= 0; --r) {
    if (r % YGRP == 0) YFENCE();
    v[r] = A2[r] * v[r] + B[r] * v[r + 1];
  }
  YFENCE();
  __syncthreads(); // earlier readers of xch are done
  xch[p * CW + lane] = yend;
  xch[(NCH + p) * CW + lane] = v[0];
  __syncthreads();
  // forward carry C_p = y(row above chunk p): C_{q+1} = yend_q + PE_q C_q.
  // backward carry D_p = x(row below chunk p) = sum over q > p of
  // (xstart_q + G0_q C_q) * S0_{p+1} ... S0_{q-1}, accumulated forwards.
  T C = 0;
  for (int q = 0; q < p; ++q) C = xch[q * CW + lane] + PE[q] * C;
  T Cq = C, D = 0, prod = 1;
  for (int q = p + 1; q < NCH; ++q) {
    Cq = xch[(q - 1) * CW + lane] + PE[q - 1] * Cq;
    D += prod * (xch[(NCH + q) * CW + lane] + G0[q] * Cq);
    prod *= S0[q];
  }
#pragma unroll
  for (int r = 0; r < L; ++r) {
    if (r % YGRP == 0) YFENCE();
    v[r] = v[r] + C * G[r] + D * S[r];
  }
  YFENCE();
}

// solve for a chunk whose input rows are read from a global column
template <int L, int G = YGRP, int NCH = YNCH, int CW = 64, typename T>
__device__ __forceinline__ void ysolve_col(T (&v)[L], const T *col, size_t s, int nin, const T *tab, int p, int lane,
                                           int n, T bot, T top, T *xch) {
  const int j0 = p * L;
  const T yend = yfwd<L, G, NCH>(
      v, [&](int k) { return col[(size_t)yclamp(j0 - 1 + k, nin) * s]; }, tab, p, n, bot, top, T(0), T(0));
  ybwd<L, NCH, CW>(v, yend, tab, p, lane, xch);
}

// one-sided closures of the operators (lip.f90), evaluated by the wave that
// holds the row they belong to
template <typename T> struct YBnd {
  T bot, top;
};

// interpyy (lip.f90:359-440): loguv 0 u rows -> v rows, 1 v rows -> u rows;
// n = ny
template <typename T>
__device__ __forceinline__ YBnd<T> interpyy_bnd(int loguv, int j0, int L, int ny, const T *inby, const T *g,
                                                size_t s) {
  YBnd<T> b{0, 0};
  if (j0 == 0 && loguv != 0) b.bot = bnd4(inby, 4, 3, g, s, 1, 1);
  if (ny - 1 >= j0 && ny - 1 < j0 + L)
    b.top = loguv == 0 ? bnd4(inby, 4, 1, g, s, ny + 1, -1) : bnd4(inby, 4, 4, g, s, ny, -1);
  return b;
}

// differzy (lip.f90:152-232): which 0 v rows -> u rows (n = ny+1), which 1
// u rows -> v rows (n = ny)
template <typename T>
__device__ __forceinline__ YBnd<T> differzy_bnd(int which, int j0, int L, int ny, const T *dcby, const T *g,
                                                size_t s) {
  YBnd<T> b{0, 0};
  const int n = which == 0 ? ny + 1 : ny;
  if (j0 == 0) b.bot = bnd4(dcby, 4, which == 0 ? 3 : 1, g, s, 1, 1);
  if (n - 1 >= j0 && n - 1 < j0 + L) b.top = bnd4(dcby, 4, which == 0 ? 4 : 2, g, s, ny, -1);
  return b;
}

// vistyy (lip.f90:447-513), n = jee
template <typename T>
__device__ __forceinline__ YBnd<T> vistyy_bnd(int jee, int j0, int L, int ny, const T *cofvby, const T *u,
                                              size_t s) {
  YBnd<T> b{0, 0};
  const int ic = (jee == ny) ? 1 : 3;
  if (j0 == 0) b.bot = bnd4(cofvby, 4, ic, u, s, 1, 1);
  if (jee - 1 >= j0 && jee - 1 < j0 + L) b.top = bnd4(cofvby, 4, ic + 1, u, s, jee, -1);
  return b;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int L, typename T>
__global__ void __launch_bounds__(YTPB, sizeof(T) == 8 ? 8 : 1)
interpyy_y(const T *__restrict__ gen, T *__restrict__ geni, const T *__restrict__ tab,
           const T *__restrict__ inby, int loguv, int ny, int nz1, int i0, int nplanes, size_t ips,
           size_t ops) {
  __shared__ T xch[2 * YNCH * 64];
  const T *tl = tab; // read in place (wave-uniform indices: scalar loads)
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, (size_t)(i0 - 1));
  const T *g = gen + q.ip * ips + q.kr;
  T *o = geni + q.ip * ops + q.kr;
  const size_t s = nz1;
  const int j0 = q.p * L;
  const YBnd<T> bd = interpyy_bnd(loguv, j0, L, ny, inby, g, s);
  __syncthreads(); // table staged
  T v[L];
  ysolve_col<L, sizeof(T) == 4 ? L : YGRP>(v, g, s, loguv == 0 ? ny + 1 : ny, tl, q.p, q.lane, ny, bd.bot, bd.top,
                                           xch);
#pragma unroll
  for (int r = 0; r < L; ++r)
    if (q.ok && j0 + r < ny) o[(size_t)(j0 + r) * s] = v[r];
}

template <int L, typename T>
__global__ void __launch_bounds__(YTPB)
differzy_y(const T *__restrict__ gen, T *__restrict__ res, const T *__restrict__ tab, DevCoefT<T> c,
           int which, int ny, int nz1, int nplanes, size_t ips, size_t ops) {
  __shared__ T xch[2 * YNCH * 64];
  const T *tl = tab; // read in place (wave-uniform indices: scalar loads)
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 0);
  const T *g = gen + q.ip * ips + q.kr;
  T *o = res + q.ip * ops + q.kr;
  const size_t s = nz1;
  const int j0 = q.p * L, n = which == 0 ? ny + 1 : ny;
  const YBnd<T> bd = differzy_bnd(which, j0, L, ny, c.dcby, g, s);
  __syncthreads();
  T v[L];
  // differzy reads input rows 1..ny in both directions
  ysolve_col<L>(v, g, s, ny, tl, q.p, q.lane, n, bd.bot, bd.top, xch);
#pragma unroll
  for (int r = 0; r < L; ++r)
    if (q.ok && j0 + r < n) o[(size_t)(j0 + r) * s] = v[r];
}

template <int L, typename T>
__global__ void __launch_bounds__(YTPB)
vistyy_y(const T *__restrict__ u, T *__restrict__ res, const T *__restrict__ tab, DevCoefT<T> c,
         int jee, int ny, int nz1, int nplanes) {
  __shared__ T xch[2 * YNCH * 64];
  const T *tl = tab; // read in place (wave-uniform indices: scalar loads)
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 0);
  const size_t s = nz1, off = q.ip * s * jee + q.kr;
  const int j0 = q.p * L;
  const YBnd<T> bd = vistyy_bnd(jee, j0, L, ny, c.cofvby, u + off, s);
  __syncthreads();
  T v[L];
  ysolve_col<L>(v, u + off, s, jee, tl, q.p, q.lane, jee, bd.bot, bd.top, xch);
#pragma unroll
  for (int r = 0; r < L; ++r)
    if (q.ok && j0 + r < jee) res[off + (size_t)(j0 + r) * s] = v[r];
}

// implzy (lip.f90:528-597): rows 2..jee-1 of u from res; rows 1 and jee of
// u are the boundary values already in place
template <int L, typename T>
__global__ void __launch_bounds__(YTPB)
implzy_y(T *u, const T *__restrict__ res, const T *__restrict__ tab, const double *__restrict__ f,
         int jee, int nz1, int nplanes) {
  __shared__ T xch[2 * YNCH * 64];
  const T *tl = tab; // read in place (wave-uniform indices: scalar loads)
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 0);
  const size_t s = nz1, off = q.ip * s * jee + q.kr;
  const int j0 = q.p * L;
  const T u1 = u[off], uj = u[off + (size_t)(jee - 1) * s];
  const T a1b = (T)f[3 * (jee + 1)], a3b = (T)f[3 * (jee + 1) + 1];
  __syncthreads();
  T w[L];
  const T *rc = res + off;
  const T yend = yfwd<L>(
      w, [&](int k) { return rc[(size_t)yclamp(j0 - 1 + k, jee) * s]; }, tl, q.p, jee, u1, uj, -a1b * u1,
      -a3b * uj);
  ybwd<L>(w, yend, tl, q.p, q.lane, xch);
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int j = j0 + r;
    if (q.ok && j >= 1 && j < jee - 1) u[off + (size_t)j * s] = w[r];
  }
}

// rhsu = d(uv)/dy + i kz (uw) on all planes (rhsp.F90:164-169)
template <int L, typename T>
__global__ void __launch_bounds__(YTPB, sizeof(T) == 8 ? 8 : 1)
rhsu_loop1_y(const T *__restrict__ uv, const T *__restrict__ uw, T *__restrict__ rhsu,
             const T *__restrict__ tab, DevCoefT<T> c, int ny, int nz1, int nplanes) {
  __shared__ T xch[2 * YNCH * 64];
  const T *tl = tab; // read in place (wave-uniform indices: scalar loads)
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 0);
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  const int j0 = q.p * L;
  const T *g = uv + q.ip * pv + q.kr;
  const YBnd<T> bd = differzy_bnd(0, j0, L, ny, c.dcby, g, s);
  __syncthreads();
  T v[L];
  ysolve_col<L, sizeof(T) == 4 ? L : YGRP>(v, g, s, ny, tl, q.p, q.lane, ny + 1, bd.bot, bd.top, xch);
  const T kz = c.kaz[q.kr / 2];
  const T sg = (q.kr & 1) ? T(1) : T(-1);
  const T *w = uw + q.ip * pu + (q.kr ^ 1);
  T *o = rhsu + q.ip * pu + q.kr;
  const size_t s1 = yopaque(s);
  // a group's loads are all issued before its first store
  constexpr int GG = sizeof(T) == 4 ? yrows_group(L, 10) : 1; // FP64: no spare registers
#pragma unroll
  for (int g0 = 0; g0 < L; g0 += GG) {
    T wp[GG];
#pragma unroll
    for (int r = 0; r < GG; ++r) wp[r] = w[(size_t)yclamp(j0 + g0 + r, ny + 1) * s1];
#pragma unroll
    for (int r = 0; r < GG; ++r) {
      const int j = j0 + g0 + r;
      if (q.ok && j <= ny) o[(size_t)j * s1] = v[g0 + r] + sg * wp[r] * kz;
    }
    YFENCE();
  }
}

// z-viscous terms and the rest of rhsv/rhsw (rhsp.F90:355-415), planes
// 2..nx-1
template <int L, typename T>
__global__ void __launch_bounds__(YTPB)
rhs_loop2_y(T *rhsu, T *rhsv, T *rhsw, const T *__restrict__ resu,
            const T *__restrict__ resv, const T *__restrict__ resw, const T *__restrict__ vw,
            const T *__restrict__ vv, const T *__restrict__ ww, const T *__restrict__ tabu,
            const T *__restrict__ tabv, DevCoefT<T> c, int ny, int nz1, int nplanes) {
  __shared__ T xch[2 * YNCH * 64];
  const T *tlu = tabu, *tlv = tabv; // read in place (scalar loads)
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 1);
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  const int j0 = q.p * L;
  const T kv = c.kvis[q.kr / 2], kz = c.kaz[q.kr / 2];
  const T sg = (q.kr & 1) ? T(1) : T(-1);
  const size_t ou = q.ip * pu + q.kr, ov = q.ip * pv + q.kr, op = q.ip * pu + (q.kr ^ 1);
  const YBnd<T> bv = differzy_bnd(1, j0, L, ny, c.dcby, vv + ou, s);
  const YBnd<T> bw = differzy_bnd(0, j0, L, ny, c.dcby, vw + ou, s);
  // rows per group of the explicit terms: a group's loads are all issued
  // before its first store (the stores go to arrays that are also read, at
  // clamped rows the compiler cannot tell apart)
  constexpr int GG = sizeof(T) == 4 ? yrows_group(L, 10) : 1; // FP64: no spare registers
  __syncthreads();
  T d[L];
  // rhsv: -kv resv (rows 2..ny-1), + i kz vw (rows 1..ny), + d(vv)/dy
  ysolve_col<L>(d, vv + ou, s, ny, tlv, q.p, q.lane, ny, bv.bot, bv.top, xch);
  {
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int g0 = 0; g0 < L; g0 += GG) {
      T x[GG], sv[GG], pw[GG];
#pragma unroll
      for (int r = 0; r < GG; ++r) {
        const size_t jc = (size_t)yclamp(j0 + g0 + r, ny) * s1;
        x[r] = rhsv[ov + jc];
        sv[r] = resv[ov + jc];
        pw[r] = vw[op + jc];
      }
#pragma unroll
      for (int r = 0; r < GG; ++r) {
        const int j = j0 + g0 + r;
        T e = x[r];
        if (j >= 1 && j <= ny - 2) e -= sv[r] * kv;
        e += sg * pw[r] * kz;
        if (q.ok && j < ny) rhsv[ov + (size_t)j * s1] = e + d[g0 + r];
      }
      YFENCE();
    }
  }
  // rhsw: -kv resw (rows 2..ny), + d(vw)/dy, + i kz ww (rows 1..ny)
  ysolve_col<L>(d, vw + ou, s, ny, tlu, q.p, q.lane, ny + 1, bw.bot, bw.top, xch);
  {
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int g0 = 0; g0 < L; g0 += GG) {
      T x[GG], sw[GG], pw[GG];
#pragma unroll
      for (int r = 0; r < GG; ++r) {
        const int j = j0 + g0 + r;
        x[r] = rhsw[ou + (size_t)yclamp(j, ny + 1) * s1];
        sw[r] = resw[ou + (size_t)yclamp(j, ny + 1) * s1];
        pw[r] = ww[op + (size_t)yclamp(j, ny) * s1]; // ww holds rows 1..ny
      }
#pragma unroll
      for (int r = 0; r < GG; ++r) {
        const int j = j0 + g0 + r;
        T e = x[r];
        if (j >= 1 && j <= ny - 1) e -= sw[r] * kv;
        e += d[g0 + r];
        if (j <= ny - 1) e += sg * pw[r] * kz;
        if (q.ok && j <= ny) rhsw[ou + (size_t)j * s1] = e;
      }
      YFENCE();
    }
  }
  // rhsu: -kv resu (rows 2..ny)
  {
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int g0 = 0; g0 < L; g0 += GG) {
      T x[GG], su[GG];
#pragma unroll
      for (int r = 0; r < GG; ++r) {
        const size_t jc = (size_t)yclamp(j0 + g0 + r, ny + 1) * s1;
        x[r] = rhsu[ou + jc];
        su[r] = resu[ou + jc];
      }
#pragma unroll
      for (int r = 0; r < GG; ++r) {
        const int j = j0 + g0 + r;
        if (q.ok && j >= 1 && j <= ny - 1) rhsu[ou + (size_t)j * s1] = x[r] - su[r] * kv;
      }
      YFENCE();
    }
  }
}

// Explicit RK3 combination with vistyy (rhsp.F90:448-475) and the implicit
// y-viscous solve (rhsp.F90:479-484) for one component, planes 2..nx
// (plane nx: no viscous term).  The reference then copies rhs into rhs_prev
// (rhsp.F90:475); here the buffers swap roles instead (k_swap_saved).
template <int L, typename T>
__global__ void __launch_bounds__(YTPB)
rk_final_y(T *u, const T *__restrict__ res, const T *__restrict__ rhs, const T *__restrict__ rhsp,
           const T *__restrict__ tabv, const T *__restrict__ tabi, DevCoefT<T> c,
           const double *__restrict__ f, const DevScalars *__restrict__ ds, int jee, int ny, int nz1, int nx) {
  __shared__ T xch[2 * YNCH * 64];
  __shared__ T hal[2 * YNCH * 64];
  // FP64: the tables are read in place (wave-uniform indices: scalar loads,
  // which frees the vector registers the LDS copies took); FP32 measured
  // faster with the tables staged in LDS
  constexpr bool STAGE = sizeof(T) == 4;
  __shared__ T tlv_s[STAGE ? ytab_size(L) : 1], tli_s[STAGE ? ytab_size(L) : 1];
  const T *tlv = tabv, *tli = tabi;
  if constexpr (STAGE) {
    ytab_stage<L>(tlv_s, tabv);
    ytab_stage<L>(tli_s, tabi);
    tlv = tlv_s;
    tli = tli_s;
  }
  const YCol q = ycol((size_t)nz1 * nx, nz1, 0);
  const size_t s = nz1, off = q.ip * s * jee + q.kr;
  const bool act = q.ok && q.ip > 0;
  const int j0 = q.p * L;
  const YBnd<T> bd = vistyy_bnd(jee, j0, L, ny, c.cofvby, res + off, s);
  const T u1 = u[off], uj = u[off + (size_t)(jee - 1) * s];
  const T a1b = (T)f[3 * (jee + 1)], a3b = (T)f[3 * (jee + 1) + 1];
  const T var2 = (q.ip == (size_t)(nx - 1)) ? T(0) : (T)ds->var2;
  const T var3 = (T)ds->var3, var4 = (T)ds->var4;
  T t[L];
  if constexpr (sizeof(T) == 4) {
    // FP32: the explicit part w = u - (var3 rhs + var4 rhs_prev) is formed
    // as soon as its 3L rows arrive (all issued at once), so only w stays
    // live through the vistyy solve
    T w[L];
    {
      T rh[L], rp[L], uu[L];
      const size_t s1 = yopaque(s);
#pragma unroll
      for (int r = 0; r < L; ++r) {
        const size_t o = off + (size_t)yclamp(j0 + r, jee) * s1;
        rh[r] = rhs[o];
        rp[r] = rhsp[o];
        uu[r] = u[o];
      }
#pragma unroll
      for (int r = 0; r < L; ++r) w[r] = uu[r] - (var3 * rh[r] + var4 * rp[r]);
    }
    YFENCE();
    __syncthreads();
    ysolve_col<L>(t, res + off, s, jee, tlv, q.p, q.lane, jee, bd.bot, bd.top, xch);
#pragma unroll
    for (int r = 0; r < L; ++r)
      if (act && j0 + r < jee) t[r] = w[r] - var2 * t[r];
  } else {
    __syncthreads();
    ysolve_col<L>(t, res + off, s, jee, tlv, q.p, q.lane, jee, bd.bot, bd.top, xch);
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int r = 0; r < L; ++r) {
      if (r % YGRP == 0) YFENCE();
      const int j = j0 + r;
      const size_t o = off + (size_t)yclamp(j, jee) * s1;
      const T rh = rhs[o], rp = rhsp[o], uu = u[o];
      if (act && j < jee) t[r] = uu - (var3 * rh + var2 * t[r] + var4 * rp);
    }
  }
  YFENCE();
  // neighbour rows of the combined right-hand side
  hal[q.p * 64 + q.lane] = t[0];
  hal[(YNCH + q.p) * 64 + q.lane] = t[L - 1];
  __syncthreads();
  const T tm = q.p > 0 ? hal[(YNCH + q.p - 1) * 64 + q.lane] : T(0);
  const T tp = q.p < YNCH - 1 ? hal[(q.p + 1) * 64 + q.lane] : T(0);
  // implzy on t, in place (row r is overwritten after rows r+1 is read)
  const T yend = yfwd<L>(
      t, [&](int k) { return k == 0 ? tm : (k == L + 1 ? tp : t[k - 1]); }, tli, q.p, jee, u1, uj, -a1b * u1,
      -a3b * uj);
  ybwd<L>(t, yend, tli, q.p, q.lane, xch);
  const size_t s2 = yopaque(s);
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int j = j0 + r;
    if (act && j >= 1 && j < jee - 1) u[off + (size_t)j * s2] = t[r];
  }
}

// ---------------------------------------------------------------------------
// Half-wave chunked FP32 rk_final of the mixed substep.  A workgroup takes
// YHC = 32 columns and each of its YHCH = 32 half-wavefronts L rows of them
// (half the rows per chunk of the kernels above), so a lane holds half the
// rows, two workgroups share a CU (64 VGPRs) and one workgroup's loads
// overlap the other's two serial solves.  The two half-waves of a wave hold
// different chunks, so the tables are staged in LDS (two addresses per
// wave: broadcast reads).  Field elements are addressed by 32-bit offsets
// from the uniform field base (one VGPR per row address; the launcher
// checks that every field stays below 2^30 values).  The one-solve FP32
// kernels (interpyy, rhsu_loop1) measured slower in this layout, and
// rhs_loop2's live state does not fit 64 VGPRs: they keep the full-wave
// chunks.
// ---------------------------------------------------------------------------
static constexpr int YHC = 32;

struct YColH {
  int p, lane;
  bool ok;
  unsigned kr, ip; // ip: 0-based plane
};

__device__ inline YColH ycolh(unsigned ncols, int nz1, unsigned ip0) {
  YColH y;
  y.lane = threadIdx.x & (YHC - 1);
  y.p = threadIdx.x / YHC;
  unsigned c = xcd_block() * YHC + y.lane;
  y.ok = c < ncols;
  if (!y.ok) c = 0; // a dead lane addresses column 0 of the first plane
  y.kr = c % (unsigned)nz1;
  y.ip = ip0 + c / (unsigned)nz1;
  return y;
}

__device__ __forceinline__ float yat(const float *base, unsigned e) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + (size_t)(e * 4u));
}
__device__ __forceinline__ float &yref(float *base, unsigned e) {
  return *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + (size_t)(e * 4u));
}

template <int L> __device__ __forceinline__ void ytabh_stage(float *dst, const float *__restrict__ src) {
  for (int k = threadIdx.x; k < (int)ytabh_size(L); k += YTPB) dst[k] = src[k];
}

// rk_final (rhsp.F90:448-484), as rk_final_y
template <int L>
__global__ void __launch_bounds__(YTPB, 8)
rk_final_yh(float *u, const float *__restrict__ res, const float *__restrict__ rhs, const float *__restrict__ rhsp,
            const float *__restrict__ tabv, const float *__restrict__ tabi, DevCoefT<float> c,
            const double *__restrict__ f, const DevScalars *__restrict__ ds, int jee, int ny, int nz1, int nx) {
  using T = float;
  __shared__ T xch[2 * YHCH * YHC];
  __shared__ T hal[2 * YHCH * YHC];
  __shared__ T tlv[ytabh_size(L)], tli[ytabh_size(L)];
  ytabh_stage<L>(tlv, tabv);
  ytabh_stage<L>(tli, tabi);
  const YColH q = ycolh((unsigned)nz1 * nx, nz1, 0);
  const unsigned s = nz1, off = q.ip * s * (unsigned)jee + q.kr;
  const bool act = q.ok && q.ip > 0;
  const int p = q.p, lane = q.lane, j0 = p * L;
  const YBnd<T> bd = vistyy_bnd(jee, j0, L, ny, c.cofvby, res + off, (size_t)s);
  const T u1 = yat(u, off), uj = yat(u, off + (unsigned)(jee - 1) * s);
  const T a1b = (T)f[3 * (jee + 1)], a3b = (T)f[3 * (jee + 1) + 1];
  const T var2 = (q.ip == (unsigned)(nx - 1)) ? T(0) : (T)ds->var2;
  const T var3 = (T)ds->var3, var4 = (T)ds->var4;
  T t[L], w[L];
  {
    // the explicit part w = u - (var3 rhs + var4 rhs_prev), its 3L rows
    // issued at once
    T rh[L], rp[L], uu[L];
#pragma unroll
    for (int r = 0; r < L; ++r) {
      const unsigned o = off + (unsigned)yclamp(j0 + r, jee) * s;
      rh[r] = yat(rhs, o);
      rp[r] = yat(rhsp, o);
      uu[r] = yat(u, o);
    }
#pragma unroll
    for (int r = 0; r < L; ++r) {
      w[r] = uu[r] - (var3 * rh[r] + var4 * rp[r]);
      asm volatile("" : "+v"(w[r])); // formed here, not sunk to its use after the solve
    }
  }
  YFENCE();
  __syncthreads(); // tables staged
  {
    const T yend = yfwd<L, YGRP, YHCH>(
        t, [&](int k) { return yat(res, off + (unsigned)yclamp(j0 - 1 + k, jee) * s); }, tlv, p, jee, bd.bot,
        bd.top, T(0), T(0));
    ybwd<L, YHCH, YHC>(t, yend, tlv, p, lane, xch);
  }
#pragma unroll
  for (int r = 0; r < L; ++r)
    if (act && j0 + r < jee) t[r] = w[r] - var2 * t[r];
  YFENCE();
  // neighbour rows of the combined right-hand side
  hal[p * YHC + lane] = t[0];
  hal[(YHCH + p) * YHC + lane] = t[L - 1];
  __syncthreads();
  const T tm = p > 0 ? hal[(YHCH + p - 1) * YHC + lane] : T(0);
  const T tp = p < YHCH - 1 ? hal[(p + 1) * YHC + lane] : T(0);
  // implzy on t, in place
  const T yend = yfwd<L, YGRP, YHCH>(
      t, [&](int k) { return k == 0 ? tm : (k == L + 1 ? tp : t[k - 1]); }, tli, p, jee, u1, uj, -a1b * u1,
      -a3b * uj);
  ybwd<L, YHCH, YHC>(t, yend, tli, p, lane, xch);
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int j = j0 + r;
    if (act && j >= 1 && j < jee - 1) yref(u, off + (unsigned)j * s) = t[r];
  }
}

// ---------------------------------------------------------------------------
// The tail of rhsp fused per component: plane loop 2's y terms
// (rhsp.F90:355-415), the boundary rows of boun (rhsp.F90:421-445) and
// rk_final (rhsp.F90:448-484) in one pass over the component's columns.  The
// final rhs is written once (it is the next substep's rhs_prev) and read back
// at once by the RK combination, from cache, instead of being written by
// rhs_loop2 and read back from HBM by rk_final.  K2 selects the component:
//   1 v: rhsv  = rhsv - kv resv (rows 2..ny-1) + i kz vw (rows 1..ny)
//                + d(vv)/dy;
//   2 w: rhsw  = rhsw - kv resw (rows 2..ny) + d(vw)/dy + i kz ww (rows
//                1..ny-1);
//   3 u: rhsu  = rhsu - kv resu (rows 2..ny),
// each on planes 2..nx-1 (vw, ww taken at the paired column kr ^ 1, pa the
// explicit product, pb the one differentiated).  boun then zeroes the rhs of
// plane 1, v's rows 1 and ny, u's and w's row ny+1, and extrapolates u's and
// w's row 1 from rows 2..4; the field's boundary rows are set (v: 0 and the
// inflow vmagic, u and w: the one-sided closures of inby and ccon; mode 0 of
// a mixed box's u: its FP64 mean flow um, whose boundary rows
// boun_mean_kernel set).  Same chunks and arithmetic as rhs_loop2_y,
// boun_kernel and rk_final_y, in the same order: the same results bit for bit.
// ---------------------------------------------------------------------------

// u / w field rows 1 and ny+1 of boun from the column's interior rows
template <typename T, class At>
__device__ __forceinline__ void tail_field_bnd(const DevCoefT<T> &c, int ny, At at, T &b0, T &bn) {
  const T i21 = T(-1) / M4(c.inby, 2, 1), c11 = T(-1) / M2(c.ccon, 1, 1);
  b0 = i21 * (M4(c.inby, 2, 2) * at(1) + M4(c.inby, 2, 3) * at(2) + M4(c.inby, 2, 4) * at(3));
  bn = c11 * (M2(c.ccon, 1, 2) * at(ny - 1) + M2(c.ccon, 1, 3) * at(ny - 2) + M2(c.ccon, 1, 4) * at(ny - 3));
}

// The fused tail runs on the full-wave chunks (FP64, and FP32 fields of 2^30
// values or more).  Each rhs row is written as soon as it is formed, with
// rows 2..4 also to LDS for row 1, and the RK update reads it back from cache.
// On the half-wave FP32 chunks, whose rk_final alone fills its 64 registers,
// the fused kernel spilled and measured slower than the three passes (9.0
// against 7.7 ms per substep at 1536 x 315 x 1024), so those keep them; on
// the full-wave chunks FP64 took 12.7 ms against 14.6.
template <int L, typename T, int K2>
__global__ void __launch_bounds__(YTPB)
rk_tail_y(T *u, T *rhs, const T *__restrict__ res, const T *__restrict__ rhsp, const T *__restrict__ pa,
          const T *__restrict__ pb, const T *__restrict__ tabv, const T *__restrict__ tabi,
          const T *__restrict__ tabd, DevCoefT<T> c, const double *__restrict__ f,
          const DevScalars *__restrict__ ds, const double *__restrict__ um, int jee, int ny, int nz1, int nx) {
  constexpr bool SOLVE = K2 != 3;
  __shared__ T xch[2 * YNCH * 64];
  __shared__ T hal[2 * YNCH * 64];
  __shared__ T bx[3 * 64];
  constexpr bool STAGE = sizeof(T) == 4;
  __shared__ T tlv_s[STAGE ? ytab_size(L) : 1], tli_s[STAGE ? ytab_size(L) : 1];
  __shared__ T tld_s[STAGE && SOLVE ? ytab_size(L) : 1];
  const T *tlv = tabv, *tli = tabi, *tld = tabd;
  if constexpr (STAGE) {
    ytab_stage<L>(tlv_s, tabv);
    ytab_stage<L>(tli_s, tabi);
    tlv = tlv_s;
    tli = tli_s;
    if constexpr (SOLVE) {
      ytab_stage<L>(tld_s, tabd);
      tld = tld_s;
    }
  }
  const YCol q = ycol((size_t)nz1 * nx, nz1, 0);
  const size_t s = nz1, off = q.ip * s * jee + q.kr;
  const size_t opu = q.ip * s * (ny + 1) + q.kr, opp = opu - q.kr + (q.kr ^ 1);
  const bool act = q.ok && q.ip > 0;
  const bool loop2 = q.ip >= 1 && q.ip + 2 <= (size_t)nx; // planes 2..nx-1
  const int j0 = q.p * L;
  const T kv = c.kvis[q.kr / 2], kz = c.kaz[q.kr / 2];
  const T sg = (q.kr & 1) ? T(1) : T(-1);
  constexpr int GG = yrows_group(L, sizeof(T) == 4 ? 5 : 4);
  __syncthreads(); // tables staged
  // the final rhs, row by row: written at once (the RK update reads it back
  // from cache), rows 2..4 also to LDS for row 1
  T d[SOLVE ? L : 1];
  if constexpr (SOLVE) {
    const YBnd<T> bd = differzy_bnd(K2 == 1 ? 1 : 0, j0, L, ny, c.dcby, pb + opu, s);
    ysolve_col<L>(d, pb + opu, s, ny, tld, q.p, q.lane, jee, bd.bot, bd.top, xch);
  }
  {
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int g0 = 0; g0 < L; g0 += GG) {
      T x[GG], sr[GG], pw[GG];
#pragma unroll
      for (int r = 0; r < GG; ++r) {
        const size_t jc = (size_t)yclamp(j0 + g0 + r, jee) * s1;
        x[r] = rhs[off + jc];
        sr[r] = res[off + jc];
        if constexpr (SOLVE) pw[r] = pa[opp + (size_t)yclamp(j0 + g0 + r, ny) * s1];
      }
#pragma unroll
      for (int r = 0; r < GG; ++r) {
        const int j = j0 + g0 + r;
        T e = x[r];
        if constexpr (K2 == 1) {
          if (j >= 1 && j <= ny - 2) e -= sr[r] * kv;
          e += sg * pw[r] * kz;
          e = e + d[g0 + r];
        } else if constexpr (K2 == 2) {
          if (j >= 1 && j <= ny - 1) e -= sr[r] * kv;
          e += d[g0 + r];
          if (j <= ny - 1) e += sg * pw[r] * kz;
        } else {
          if (j >= 1 && j <= ny - 1) e = x[r] - sr[r] * kv;
        }
        e = loop2 ? e : x[r];
        if (q.ip == 0) e = T(0);
        if constexpr (K2 == 1) {
          if (j == 0 || j == ny - 1) e = T(0);
        } else {
          if (j == ny) e = T(0);
          if (j >= 1 && j <= 3) bx[(j - 1) * 64 + q.lane] = e;
        }
        if (q.ok && j < jee) rhs[off + (size_t)j * s1] = e;
      }
      YFENCE();
    }
  }
  if constexpr (K2 != 1) {
    // row 1 from rows 2..4 (another chunk may hold them)
    __syncthreads();
    if (q.p == 0 && q.ok) {
      const T i21 = T(-1) / M4(c.inby, 2, 1);
      rhs[off] = i21 * (M4(c.inby, 2, 2) * bx[q.lane] + M4(c.inby, 2, 3) * bx[64 + q.lane] +
                        M4(c.inby, 2, 4) * bx[128 + q.lane]);
    }
  }
  YFENCE();
  T u1, uj;
  if constexpr (K2 == 1) {
    u1 = T(0);
    uj = q.kr == 0 ? c.vmagic[q.ip + 1] : T(0);
  } else if (K2 == 3 && um && q.kr == 0) {
    u1 = (T)um[q.ip];
    uj = (T)um[(size_t)ny * nx + q.ip];
  } else {
    tail_field_bnd(c, ny, [&](int j) { return u[off + (size_t)j * s]; }, u1, uj);
  }
  // boun's boundary rows into the field now, by the lanes that hold them:
  // the RK update below reads them back
  if (q.ok && j0 == 0) u[off] = u1;
  if (q.ok && jee - 1 >= j0 && jee - 1 < j0 + L) u[off + (size_t)(jee - 1) * s] = uj;
  YFENCE();
  const T a1b = (T)f[3 * (jee + 1)], a3b = (T)f[3 * (jee + 1) + 1];
  const T var2 = (q.ip == (size_t)(nx - 1)) ? T(0) : (T)ds->var2;
  const T var3 = (T)ds->var3, var4 = (T)ds->var4;
  const YBnd<T> bd = vistyy_bnd(jee, j0, L, ny, c.cofvby, res + off, s);
  T t[L];
  if constexpr (sizeof(T) == 4) {
    T w[L];
    {
      T rh[L], rp[L], uu[L];
      const size_t s1 = yopaque(s);
#pragma unroll
      for (int r = 0; r < L; ++r) {
        const size_t o = off + (size_t)yclamp(j0 + r, jee) * s1;
        rh[r] = rhs[o];
        rp[r] = rhsp[o];
        uu[r] = u[o];
      }
#pragma unroll
      for (int r = 0; r < L; ++r) w[r] = uu[r] - (var3 * rh[r] + var4 * rp[r]);
    }
    YFENCE();
    ysolve_col<L>(t, res + off, s, jee, tlv, q.p, q.lane, jee, bd.bot, bd.top, xch);
#pragma unroll
    for (int r = 0; r < L; ++r)
      if (act && j0 + r < jee) t[r] = w[r] - var2 * t[r];
  } else {
    YFENCE();
    ysolve_col<L>(t, res + off, s, jee, tlv, q.p, q.lane, jee, bd.bot, bd.top, xch);
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int r = 0; r < L; ++r) {
      if (r % YGRP == 0) YFENCE();
      const int j = j0 + r;
      const size_t o = off + (size_t)yclamp(j, jee) * s1;
      const T rr = rhs[o], rp = rhsp[o], uu = u[o];
      if (act && j < jee) t[r] = uu - (var3 * rr + var2 * t[r] + var4 * rp);
    }
  }
  YFENCE();
  hal[q.p * 64 + q.lane] = t[0];
  hal[(YNCH + q.p) * 64 + q.lane] = t[L - 1];
  __syncthreads();
  const T tm = q.p > 0 ? hal[(YNCH + q.p - 1) * 64 + q.lane] : T(0);
  const T tp = q.p < YNCH - 1 ? hal[(q.p + 1) * 64 + q.lane] : T(0);
  const T yend = yfwd<L>(
      t, [&](int k) { return k == 0 ? tm : (k == L + 1 ? tp : t[k - 1]); }, tli, q.p, jee, u1, uj, -a1b * u1,
      -a3b * uj);
  ybwd<L>(t, yend, tli, q.p, q.lane, xch);
  const size_t s2 = yopaque(s);
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int j = j0 + r;
    if (act && j >= 1 && j < jee - 1) u[off + (size_t)j * s2] = t[r];
  }
}

// boun's rows 1 and ny+1 of the FP64 mean-flow u of the mixed substep
// (um [ny+1][nx], plane fastest), from its own interior
__global__ void boun_mean_kernel(double *um, DevCoef cd, int ny, int nx) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= nx) return;
  double *U = um + ip;
  auto UR = [&](int r) -> double & { return U[(size_t)r * nx]; }; // row r+1
  UR(0) = -1.0 / M4(cd.inby, 2, 1) * (M4(cd.inby, 2, 2) * UR(1) + M4(cd.inby, 2, 3) * UR(2) + M4(cd.inby, 2, 4) * UR(3));
  UR(ny) = -1.0 / M2(cd.ccon, 1, 1) *
           (M2(cd.ccon, 1, 2) * UR(ny - 1) + M2(cd.ccon, 1, 3) * UR(ny - 2) + M2(cd.ccon, 1, 4) * UR(ny - 3));
}

// rk_final of the FP64 mean-flow u of the mixed substep (FP32 fields): the
// explicit RK3 combination with vistyy (rhsp.F90:448-475) and the implicit y
// solve (rhsp.F90:479-484) of the kz = 0 real part of u, every plane a
// column.  um and its saved copy rm are [ny+1][nx] (plane fastest: column =
// plane, row stride nx); rhs and rhs_prev are the FP32 fields (mode 0 of
// plane ip at ip (ny+1) nz1, row stride nz1).  FP64 tables and arithmetic.
template <int L, typename T>
__global__ void __launch_bounds__(YTPB)
mean_rk_final_y(double *um, const double *__restrict__ rm, const T *__restrict__ rhs, const T *__restrict__ rhsp,
                const double *__restrict__ tabv, const double *__restrict__ tabi, DevCoef c,
                const double *__restrict__ f, const DevScalars *__restrict__ ds, int ny, int nz1, int nx) {
  __shared__ double xch[2 * YNCH * 64];
  const double *tlv = tabv, *tli = tabi; // read in place (scalar loads)
  __shared__ double hal[2 * YNCH * 64];

  const int jee = ny + 1;
  const YCol q = ycol((size_t)nx, nx, 0); // column kr = plane index
  const int ip = q.kr;
  const size_t s = nx, off = ip;
  const T *rh = rhs + (size_t)ip * nz1 * jee, *rp = rhsp + (size_t)ip * nz1 * jee;
  const bool act = q.ok && ip > 0; // plane 1 keeps u
  const int j0 = q.p * L;
  const YBnd<double> bd = vistyy_bnd(jee, j0, L, ny, c.cofvby, rm + off, s);
  const double u1 = um[off], uj = um[off + (size_t)(jee - 1) * s];
  const double a1b = f[3 * (jee + 1)], a3b = f[3 * (jee + 1) + 1];
  const double var2 = ip == nx - 1 ? 0.0 : ds->var2, var3 = ds->var3, var4 = ds->var4;
  __syncthreads();
  double t[L];
  ysolve_col<L>(t, rm + off, s, jee, tlv, q.p, q.lane, jee, bd.bot, bd.top, xch);
#pragma unroll
  for (int r = 0; r < L; ++r) {
    if (r % YGRP == 0) YFENCE();
    const int j = j0 + r, jc = yclamp(j, jee);
    const double rr = (double)rh[(size_t)jc * nz1], rq = (double)rp[(size_t)jc * nz1];
    const double uu = um[off + (size_t)jc * s];
    if (act && j < jee) t[r] = uu - (var3 * rr + var2 * t[r] + var4 * rq);
  }
  YFENCE();
  hal[q.p * 64 + q.lane] = t[0];
  hal[(YNCH + q.p) * 64 + q.lane] = t[L - 1];
  __syncthreads();
  const double tm = q.p > 0 ? hal[(YNCH + q.p - 1) * 64 + q.lane] : 0.0;
  const double tp = q.p < YNCH - 1 ? hal[(q.p + 1) * 64 + q.lane] : 0.0;
  const double yend = yfwd<L>(
      t, [&](int k) { return k == 0 ? tm : (k == L + 1 ? tp : t[k - 1]); }, tli, q.p, jee, u1, uj, -a1b * u1,
      -a3b * uj);
  ybwd<L>(t, yend, tli, q.p, q.lane, xch);
#pragma unroll
  for (int r = 0; r < L; ++r) {
    const int j = j0 + r;
    if (act && j >= 1 && j < jee - 1) um[off + (size_t)j * s] = t[r];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define YDISPATCH(Lv, ...)                                                     \
  switch (Lv) {                                                                \
  case 4: { constexpr int L = 4; __VA_ARGS__; } break;                         \
  case 8: { constexpr int L = 8; __VA_ARGS__; } break;                         \
  case 12: { constexpr int L = 12; __VA_ARGS__; } break;                       \
  case 16: { constexpr int L = 16; __VA_ARGS__; } break;                       \
  case 20: { constexpr int L = 20; __VA_ARGS__; } break;                       \
  case 24: { constexpr int L = 24; __VA_ARGS__; } break;                       \
  default: throw std::invalid_argument("y chunk length not instantiated");    \
  }

static inline unsigned ygrid(size_t ncols) { return (unsigned)((ncols + 63) / 64); }
template <typename T> static inline const T *ytab(const Box &b, int k);
template <> inline const double *ytab<double>(const Box &b, int k) {
  return b.ytab + (size_t)k * ytab_size(b.yL);
}
template <> inline const float *ytab<float>(const Box &b, int k) { return b.ytab_f + (size_t)k * ytab_size(b.yL); }

// the half-wave rk_final (FP32 with half-wave tables): workgroups for
// ncols columns, table of a system, and the 32-bit offset bound
static inline unsigned ygridh(size_t ncols) { return (unsigned)((ncols + YHC - 1) / YHC); }
static inline const float *ytabh(const Box &b, int k) { return b.ytabh_f + (size_t)k * ytabh_size(b.yL / 2); }
static void yh_check(size_t last) {
  if (last >= ((size_t)1 << 30))
    throw std::invalid_argument("half-wave y kernels: a field of more than 2^30 values");
}

template <typename T>
void ky_interpyy(Box &b, const T *gen, T *geni, int loguv, int i0, int nplanes, size_t ips, size_t ops) {
  const unsigned g = ygrid((size_t)b.nz1 * nplanes);
  YDISPATCH(b.yL, auto kf = &interpyy_y<L, T>;
            LAUNCH(b, "interpyy", g, YTPB, 0, kf, gen, geni, ytab<T>(b, loguv == 0 ? YT_INYU : YT_INYV),
                   coefs<T>(b).inby, loguv, b.ny, b.nz1, i0, nplanes, ips, ops));
}

template <typename T> void ky_differzy(Box &b, const T *geni, T *res, int which, int nplanes, size_t ips, size_t ops) {
  const unsigned g = ygrid((size_t)b.nz1 * nplanes);
  YDISPATCH(b.yL, auto kf = &differzy_y<L, T>;
            LAUNCH(b, "differzy", g, YTPB, 0, kf, geni, res, ytab<T>(b, which == 0 ? YT_DCYU : YT_DCYV), coefs<T>(b),
                   which, b.ny, b.nz1, nplanes, ips, ops));
}

template <typename T> void ky_vistyy(Box &b, const T *u, T *res, int R, int nplanes) {
  const unsigned g = ygrid((size_t)b.nz1 * nplanes);
  YDISPATCH(b.yL, auto kf = &vistyy_y<L, T>;
            LAUNCH(b, "vistyy", g, YTPB, 0, kf, u, res, ytab<T>(b, R == b.ny + 1 ? YT_VIYU : YT_VIYV), coefs<T>(b), R,
                   b.ny, b.nz1, nplanes));
}

template <typename T> void ky_implzy(Box &b, T *u, const T *res, int R, int nplanes) {
  const unsigned g = ygrid((size_t)b.nz1 * nplanes);
  const bool uu = R == b.ny + 1;
  YDISPATCH(b.yL, auto kf = &implzy_y<L, T>;
            LAUNCH(b, "implzy", g, YTPB, 0, kf, u, res, ytab<T>(b, uu ? YT_IMPU : YT_IMPV), uu ? b.implu : b.implv,
                   R, b.nz1, nplanes));
}

template <typename T> void ky_rhsu_loop1(Box &b, const T *uv, const T *uw, T *rhsu) {
  const unsigned g = ygrid((size_t)b.nz1 * b.nx);
  YDISPATCH(b.yL, auto kf = &rhsu_loop1_y<L, T>;
            LAUNCH(b, "rhsu_loop1", g, YTPB, 0, kf, uv, uw, rhsu, ytab<T>(b, YT_DCYU), coefs<T>(b), b.ny, b.nz1, b.nx));
}

template <typename T>
void ky_rhs_loop2(Box &b, T *rhsu, T *rhsv, T *rhsw, const T *resu, const T *resv, const T *resw, const T *vw,
                  const T *vv, const T *ww) {
  const unsigned g = ygrid((size_t)b.nz1 * (b.nx - 2));
  YDISPATCH(b.yL, auto kf = &rhs_loop2_y<L, T>;
            LAUNCH(b, "rhs_loop2", g, YTPB, 0, kf, rhsu, rhsv, rhsw, resu, resv, resw, vw, vv, ww,
                   ytab<T>(b, YT_DCYU), ytab<T>(b, YT_DCYV), coefs<T>(b), b.ny, b.nz1, b.nx - 2));
}

template <typename T> void ky_rk_final(Box &b, T *u, const T *res, const T *rhs, const T *rhsp, int jee) {
  const unsigned g = ygrid((size_t)b.nz1 * b.nx);
  const bool uu = jee == b.ny + 1;
  if constexpr (sizeof(T) == 4) {
    if (b.ytabh_f && b.yhalf) {
      yh_check((size_t)b.nz1 * b.nx * (b.ny + 1));
      YDISPATCH(b.yL, auto kf = &rk_final_yh<L / 2>;
                LAUNCH(b, "rk_final", ygridh((size_t)b.nz1 * b.nx), YTPB, 0, kf, u, res, rhs, rhsp,
                       ytabh(b, uu ? YT_VIYU : YT_VIYV), ytabh(b, uu ? YT_IMPU : YT_IMPV), coefs<T>(b),
                       uu ? b.implu : b.implv, b.ds, jee, b.ny, b.nz1, b.nx));
      return;
    }
  }
  YDISPATCH(b.yL, auto kf = &rk_final_y<L, T>;
            LAUNCH(b, "rk_final", g, YTPB, 0, kf, u, res, rhs, rhsp, ytab<T>(b, uu ? YT_VIYU : YT_VIYV),
                   ytab<T>(b, uu ? YT_IMPU : YT_IMPV), coefs<T>(b), uu ? b.implu : b.implv, b.ds, jee, b.ny, b.nz1,
                   b.nx));
}

// the fused tail of one component (k2 = 1 v, 2 w, 3 u) on the full-wave
// chunks; pa, pb the products of plane loop 2 (v: vw, vv; w: ww, vw; u:
// unused)
template <typename T>
void ky_rk_tail(Box &b, int k2, T *u, T *rhs, const T *res, const T *rhsp, const T *pa, const T *pb) {
  const int jee = k2 == 1 ? b.ny : b.ny + 1;
  const bool uu = jee == b.ny + 1;
  const int kv = uu ? YT_VIYU : YT_VIYV, ki = uu ? YT_IMPU : YT_IMPV, kd = k2 == 1 ? YT_DCYV : YT_DCYU;
  const double *um = (k2 == 3 && b.precision == OTBL_MIXED) ? b.um64 : nullptr;
  if (k2 == 3) pa = pb = nullptr;
  const unsigned g = ygrid((size_t)b.nz1 * b.nx);
  YDISPATCH(b.yL, auto kf = k2 == 1 ? &rk_tail_y<L, T, 1> : (k2 == 2 ? &rk_tail_y<L, T, 2> : &rk_tail_y<L, T, 3>);
            LAUNCH(b, "rk_tail", g, YTPB, 0, kf, u, rhs, res, rhsp, pa, pb, ytab<T>(b, kv), ytab<T>(b, ki),
                   ytab<T>(b, kd), coefs<T>(b), uu ? b.implu : b.implv, b.ds, um, jee, b.ny, b.nz1, b.nx));
}
template void ky_rk_tail<double>(Box &, int, double *, double *, const double *, const double *, const double *,
                                 const double *);
template void ky_rk_tail<float>(Box &, int, float *, float *, const float *, const float *, const float *,
                                const float *);

void k_boun_mean(Box &b) {
  LAUNCH(b, "boun", (b.nx + 255) / 256, 256, 0, boun_mean_kernel, b.um64, b.dc, b.ny, b.nx);
}

template <typename T> void ky_mean_rk_final(Box &b, double *um, const double *rm, const T *rhs, const T *rhsp) {
  const unsigned g = ygrid((size_t)b.nx);
  YDISPATCH(b.yL, auto kf = &mean_rk_final_y<L, T>;
            LAUNCH(b, "rk_mean", g, YTPB, 0, kf, um, rm, rhs, rhsp, ytab<double>(b, YT_VIYU),
                   ytab<double>(b, YT_IMPU), b.dc, b.implu, b.ds, b.ny, b.nz1, b.nx));
}
template void ky_mean_rk_final<float>(Box &, double *, const double *, const float *, const float *);

#define YINST(T)                                                                                               \
  template void ky_interpyy<T>(Box &, const T *, T *, int, int, int, size_t, size_t);                          \
  template void ky_differzy<T>(Box &, const T *, T *, int, int, size_t, size_t);                               \
  template void ky_vistyy<T>(Box &, const T *, T *, int, int);                                                 \
  template void ky_implzy<T>(Box &, T *, const T *, int, int);                                                 \
  template void ky_rhsu_loop1<T>(Box &, const T *, const T *, T *);                                            \
  template void ky_rhs_loop2<T>(Box &, T *, T *, T *, const T *, const T *, const T *, const T *, const T *,     \
                                const T *);                                                                    \
  template void ky_rk_final<T>(Box &, T *, const T *, const T *, const T *, int);
YINST(double)
YINST(float)

} // namespace otbl
