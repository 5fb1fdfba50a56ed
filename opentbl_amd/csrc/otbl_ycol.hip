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
// coefficients of each operator are stored in its system's table, which a
// workgroup stages in LDS.
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
  size_t c = (size_t)blockIdx.x * 64 + y.lane;
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

// Copy one system's table into LDS (all threads); the caller synchronises.
template <int L>
__device__ __forceinline__ void ytab_stage(double *dst, const double *__restrict__ src) {
  for (int k = threadIdx.x; k < (int)ytab_size(L); k += YTPB) dst[k] = src[k];
}

// sum_k cf(row, k) * G(j1 + dir*(k-1)), k = 1..4, over 1-based rows of col:
// the one-sided closures of the compact y operators
__device__ __forceinline__ double bnd4(const double *cf, int ld, int row, const double *col, size_t s,
                                       int j1, int dir) {
  double r = 0.0;
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
template <int L, class Fetch>
__device__ __forceinline__ double yfwd(double (&v)[L], Fetch fetch, const double *tab, int p, int n, double bot,
                                       double top, double add1, double addm) {
  constexpr int NL = YNCH * L;
  const double *M = tab + p * L, *E1 = tab + 5 * NL + p * L, *E2 = tab + 6 * NL + p * L,
               *E3 = tab + 7 * NL + p * L;
  const int j0 = p * L;
  double gm = fetch(0), g0 = fetch(1), y = 0.0;
#pragma unroll
  for (int r = 0; r < L; ++r) {
    if (r % YGRP == 0) YFENCE();
    const int j = j0 + r;
    const double gp = fetch(r + 2);
    double x = E1[r] * gm + E2[r] * g0 + E3[r] * gp;
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
// xch: LDS, 2*YNCH*64 doubles.
template <int L>
__device__ __forceinline__ void ybwd(double (&v)[L], double yend, const double *tab, int p, int lane,
                                     double *xch) {
  constexpr int NL = YNCH * L;
  const double *A2 = tab + NL + p * L, *B = tab + 2 * NL + p * L;
  const double *G = tab + 3 * NL + p * L, *S = tab + 4 * NL + p * L;
  const double *PE = tab + 8 * NL, *G0 = PE + YNCH, *S0 = G0 + YNCH;
  v[L - 1] = A2[L - 1] * v[L - 1];
#pragma unroll
  for (int r = L - 2; r >= 0; --r) {
    if (r % YGRP == 0) YFENCE();
    v[r] = A2[r] * v[r] + B[r] * v[r + 1];
  }
  YFENCE();
  __syncthreads(); // earlier readers of xch are done
  xch[p * 64 + lane] = yend;
  xch[(YNCH + p) * 64 + lane] = v[0];
  __syncthreads();
  // forward carry C_p = y(row above chunk p): C_{q+1} = yend_q + PE_q C_q.
  // backward carry D_p = x(row below chunk p) = sum over q > p of
  // (xstart_q + G0_q C_q) * S0_{p+1} ... S0_{q-1}, accumulated forwards.
  double C = 0.0;
  for (int q = 0; q < p; ++q) C = xch[q * 64 + lane] + PE[q] * C;
  double Cq = C, D = 0.0, prod = 1.0;
  for (int q = p + 1; q < YNCH; ++q) {
    Cq = xch[(q - 1) * 64 + lane] + PE[q - 1] * Cq;
    D += prod * (xch[(YNCH + q) * 64 + lane] + G0[q] * Cq);
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
template <int L>
__device__ __forceinline__ void ysolve_col(double (&v)[L], const double *col, size_t s, int nin, const double *tab,
                                           int p, int lane, int n, double bot, double top, double *xch) {
  const int j0 = p * L;
  const double yend = yfwd<L>(
      v, [&](int k) { return col[(size_t)yclamp(j0 - 1 + k, nin) * s]; }, tab, p, n, bot, top, 0.0, 0.0);
  ybwd<L>(v, yend, tab, p, lane, xch);
}

// one-sided closures of the operators (lip.f90), evaluated by the wave that
// holds the row they belong to
struct YBnd {
  double bot, top;
};

// interpyy (lip.f90:359-440): loguv 0 u rows -> v rows, 1 v rows -> u rows;
// n = ny
__device__ __forceinline__ YBnd interpyy_bnd(int loguv, int j0, int L, int ny, const double *inby,
                                             const double *g, size_t s) {
  YBnd b{0.0, 0.0};
  if (j0 == 0 && loguv != 0) b.bot = bnd4(inby, 4, 3, g, s, 1, 1);
  if (ny - 1 >= j0 && ny - 1 < j0 + L)
    b.top = loguv == 0 ? bnd4(inby, 4, 1, g, s, ny + 1, -1) : bnd4(inby, 4, 4, g, s, ny, -1);
  return b;
}

// differzy (lip.f90:152-232): which 0 v rows -> u rows (n = ny+1), which 1
// u rows -> v rows (n = ny)
__device__ __forceinline__ YBnd differzy_bnd(int which, int j0, int L, int ny, const double *dcby,
                                             const double *g, size_t s) {
  YBnd b{0.0, 0.0};
  const int n = which == 0 ? ny + 1 : ny;
  if (j0 == 0) b.bot = bnd4(dcby, 4, which == 0 ? 3 : 1, g, s, 1, 1);
  if (n - 1 >= j0 && n - 1 < j0 + L) b.top = bnd4(dcby, 4, which == 0 ? 4 : 2, g, s, ny, -1);
  return b;
}

// vistyy (lip.f90:447-513), n = jee
__device__ __forceinline__ YBnd vistyy_bnd(int jee, int j0, int L, int ny, const double *cofvby, const double *u,
                                           size_t s) {
  YBnd b{0.0, 0.0};
  const int ic = (jee == ny) ? 1 : 3;
  if (j0 == 0) b.bot = bnd4(cofvby, 4, ic, u, s, 1, 1);
  if (jee - 1 >= j0 && jee - 1 < j0 + L) b.top = bnd4(cofvby, 4, ic + 1, u, s, jee, -1);
  return b;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(YTPB)
interpyy_y(const double *__restrict__ gen, double *__restrict__ geni, const double *__restrict__ tab,
           const double *__restrict__ inby, int loguv, int ny, int nz1, int i0, int nplanes, size_t ips,
           size_t ops) {
  __shared__ double xch[2 * YNCH * 64];
  __shared__ double tl[ytab_size(L)];
  ytab_stage<L>(tl, tab);
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, (size_t)(i0 - 1));
  const double *g = gen + q.ip * ips + q.kr;
  double *o = geni + q.ip * ops + q.kr;
  const size_t s = nz1;
  const int j0 = q.p * L;
  const YBnd bd = interpyy_bnd(loguv, j0, L, ny, inby, g, s);
  __syncthreads(); // table staged
  double v[L];
  ysolve_col<L>(v, g, s, loguv == 0 ? ny + 1 : ny, tl, q.p, q.lane, ny, bd.bot, bd.top, xch);
#pragma unroll
  for (int r = 0; r < L; ++r)
    if (q.ok && j0 + r < ny) o[(size_t)(j0 + r) * s] = v[r];
}

template <int L>
__global__ void __launch_bounds__(YTPB)
differzy_y(const double *__restrict__ gen, double *__restrict__ res, const double *__restrict__ tab, DevCoef c,
           int which, int ny, int nz1, int nplanes, size_t ips, size_t ops) {
  __shared__ double xch[2 * YNCH * 64];
  __shared__ double tl[ytab_size(L)];
  ytab_stage<L>(tl, tab);
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 0);
  const double *g = gen + q.ip * ips + q.kr;
  double *o = res + q.ip * ops + q.kr;
  const size_t s = nz1;
  const int j0 = q.p * L, n = which == 0 ? ny + 1 : ny;
  const YBnd bd = differzy_bnd(which, j0, L, ny, c.dcby, g, s);
  __syncthreads();
  double v[L];
  // differzy reads input rows 1..ny in both directions
  ysolve_col<L>(v, g, s, ny, tl, q.p, q.lane, n, bd.bot, bd.top, xch);
#pragma unroll
  for (int r = 0; r < L; ++r)
    if (q.ok && j0 + r < n) o[(size_t)(j0 + r) * s] = v[r];
}

template <int L>
__global__ void __launch_bounds__(YTPB)
vistyy_y(const double *__restrict__ u, double *__restrict__ res, const double *__restrict__ tab, DevCoef c,
         int jee, int ny, int nz1, int nplanes) {
  __shared__ double xch[2 * YNCH * 64];
  __shared__ double tl[ytab_size(L)];
  ytab_stage<L>(tl, tab);
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 0);
  const size_t s = nz1, off = q.ip * s * jee + q.kr;
  const int j0 = q.p * L;
  const YBnd bd = vistyy_bnd(jee, j0, L, ny, c.cofvby, u + off, s);
  __syncthreads();
  double v[L];
  ysolve_col<L>(v, u + off, s, jee, tl, q.p, q.lane, jee, bd.bot, bd.top, xch);
#pragma unroll
  for (int r = 0; r < L; ++r)
    if (q.ok && j0 + r < jee) res[off + (size_t)(j0 + r) * s] = v[r];
}

// implzy (lip.f90:528-597): rows 2..jee-1 of u from res; rows 1 and jee of
// u are the boundary values already in place
template <int L>
__global__ void __launch_bounds__(YTPB)
implzy_y(double *u, const double *__restrict__ res, const double *__restrict__ tab, const double *__restrict__ f,
         int jee, int nz1, int nplanes) {
  __shared__ double xch[2 * YNCH * 64];
  __shared__ double tl[ytab_size(L)];
  ytab_stage<L>(tl, tab);
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 0);
  const size_t s = nz1, off = q.ip * s * jee + q.kr;
  const int j0 = q.p * L;
  const double u1 = u[off], uj = u[off + (size_t)(jee - 1) * s];
  const double a1b = f[3 * (jee + 1)], a3b = f[3 * (jee + 1) + 1];
  __syncthreads();
  double w[L];
  const double *rc = res + off;
  const double yend = yfwd<L>(
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
template <int L>
__global__ void __launch_bounds__(YTPB)
rhsu_loop1_y(const double *__restrict__ uv, const double *__restrict__ uw, double *__restrict__ rhsu,
             const double *__restrict__ tab, DevCoef c, int ny, int nz1, int nplanes) {
  __shared__ double xch[2 * YNCH * 64];
  __shared__ double tl[ytab_size(L)];
  ytab_stage<L>(tl, tab);
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 0);
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  const int j0 = q.p * L;
  const double *g = uv + q.ip * pv + q.kr;
  const YBnd bd = differzy_bnd(0, j0, L, ny, c.dcby, g, s);
  __syncthreads();
  double v[L];
  ysolve_col<L>(v, g, s, ny, tl, q.p, q.lane, ny + 1, bd.bot, bd.top, xch);
  const double kz = c.kaz[q.kr / 2];
  const double sg = (q.kr & 1) ? 1.0 : -1.0;
  const double *w = uw + q.ip * pu + (q.kr ^ 1);
  double *o = rhsu + q.ip * pu + q.kr;
  const size_t s1 = yopaque(s);
#pragma unroll
  for (int r = 0; r < L; ++r) {
    if (r % YGRP == 0) YFENCE();
    const int j = j0 + r;
    const size_t jc = (size_t)yclamp(j, ny + 1) * s1;
    const double wp = w[jc];
    if (q.ok && j <= ny) o[jc] = v[r] + sg * wp * kz;
  }
}

// z-viscous terms and the rest of rhsv/rhsw (rhsp.F90:355-415), planes
// 2..nx-1
template <int L>
__global__ void __launch_bounds__(YTPB)
rhs_loop2_y(double *rhsu, double *rhsv, double *rhsw, const double *__restrict__ resu,
            const double *__restrict__ resv, const double *__restrict__ resw, const double *__restrict__ vw,
            const double *__restrict__ vv, const double *__restrict__ ww, const double *__restrict__ tabu,
            const double *__restrict__ tabv, DevCoef c, int ny, int nz1, int nplanes) {
  __shared__ double xch[2 * YNCH * 64];
  __shared__ double tlu[ytab_size(L)];
  __shared__ double tlv[ytab_size(L)];
  ytab_stage<L>(tlu, tabu);
  ytab_stage<L>(tlv, tabv);
  const YCol q = ycol((size_t)nz1 * nplanes, nz1, 1);
  const size_t s = nz1, pu = s * (ny + 1), pv = s * ny;
  const int j0 = q.p * L;
  const double kv = c.kvis[q.kr / 2], kz = c.kaz[q.kr / 2];
  const double sg = (q.kr & 1) ? 1.0 : -1.0;
  const size_t ou = q.ip * pu + q.kr, ov = q.ip * pv + q.kr, op = q.ip * pu + (q.kr ^ 1);
  const YBnd bv = differzy_bnd(1, j0, L, ny, c.dcby, vv + ou, s);
  const YBnd bw = differzy_bnd(0, j0, L, ny, c.dcby, vw + ou, s);
  __syncthreads();
  double d[L];
  // rhsv: -kv resv (rows 2..ny-1), + i kz vw (rows 1..ny), + d(vv)/dy
  ysolve_col<L>(d, vv + ou, s, ny, tlv, q.p, q.lane, ny, bv.bot, bv.top, xch);
  {
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int r = 0; r < L; ++r) {
      if (r % YGRP == 0) YFENCE();
      const int j = j0 + r;
      const size_t jc = (size_t)yclamp(j, ny) * s1;
      double x = rhsv[ov + jc];
      const double sv = resv[ov + jc], pw = vw[op + jc];
      if (j >= 1 && j <= ny - 2) x -= sv * kv;
      x += sg * pw * kz;
      if (q.ok && j < ny) rhsv[ov + jc] = x + d[r];
    }
  }
  YFENCE();
  // rhsw: -kv resw (rows 2..ny), + d(vw)/dy, + i kz ww (rows 1..ny)
  ysolve_col<L>(d, vw + ou, s, ny, tlu, q.p, q.lane, ny + 1, bw.bot, bw.top, xch);
  {
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int r = 0; r < L; ++r) {
      if (r % YGRP == 0) YFENCE();
      const int j = j0 + r;
      const size_t jc = (size_t)yclamp(j, ny + 1) * s1;
      double x = rhsw[ou + jc];
      const double sw = resw[ou + jc];
      const double pw = ww[op + (size_t)yclamp(j, ny) * s1]; // ww holds rows 1..ny
      if (j >= 1 && j <= ny - 1) x -= sw * kv;
      x += d[r];
      if (j <= ny - 1) x += sg * pw * kz;
      if (q.ok && j <= ny) rhsw[ou + jc] = x;
    }
  }
  YFENCE();
  // rhsu: -kv resu (rows 2..ny)
  {
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int r = 0; r < L; ++r) {
      if (r % YGRP == 0) YFENCE();
      const int j = j0 + r;
      const size_t jc = (size_t)yclamp(j, ny + 1) * s1;
      const double x = rhsu[ou + jc], su = resu[ou + jc];
      if (q.ok && j >= 1 && j <= ny - 1) rhsu[ou + jc] = x - su * kv;
    }
  }
}

// Explicit RK3 combination with vistyy (rhsp.F90:448-475) and the implicit
// y-viscous solve (rhsp.F90:479-484) for one component, planes 1..nx
// (plane 1: rhs_prev = rhs only; plane nx: no viscous term).
template <int L>
__global__ void __launch_bounds__(YTPB)
rk_final_y(double *u, const double *__restrict__ res, const double *__restrict__ rhs, double *rhsp,
           const double *__restrict__ tabv, const double *__restrict__ tabi, DevCoef c,
           const double *__restrict__ f, const DevScalars *__restrict__ ds, int jee, int ny, int nz1, int nx) {
  __shared__ double xch[2 * YNCH * 64];
  __shared__ double tlv[ytab_size(L)];
  __shared__ double tli[ytab_size(L)];
  __shared__ double hal[2 * YNCH * 64];
  ytab_stage<L>(tlv, tabv);
  ytab_stage<L>(tli, tabi);
  const YCol q = ycol((size_t)nz1 * nx, nz1, 0);
  const size_t s = nz1, off = q.ip * s * jee + q.kr;
  const bool act = q.ok && q.ip > 0;
  const int j0 = q.p * L;
  const YBnd bd = vistyy_bnd(jee, j0, L, ny, c.cofvby, res + off, s);
  const double u1 = u[off], uj = u[off + (size_t)(jee - 1) * s];
  const double a1b = f[3 * (jee + 1)], a3b = f[3 * (jee + 1) + 1];
  __syncthreads();
  double t[L];
  ysolve_col<L>(t, res + off, s, jee, tlv, q.p, q.lane, jee, bd.bot, bd.top, xch);
  const double var2 = (q.ip == (size_t)(nx - 1)) ? 0.0 : ds->var2;
  const double var3 = ds->var3, var4 = ds->var4;
  {
    const size_t s1 = yopaque(s);
#pragma unroll
    for (int r = 0; r < L; ++r) {
      if (r % YGRP == 0) YFENCE();
      const int j = j0 + r;
      const size_t o = off + (size_t)yclamp(j, jee) * s1;
      const double rh = rhs[o], rp = rhsp[o], uu = u[o];
      if (act && j < jee) t[r] = uu - (var3 * rh + var2 * t[r] + var4 * rp);
      if (q.ok && j < jee) rhsp[o] = rh;
    }
  }
  YFENCE();
  // neighbour rows of the combined right-hand side
  hal[q.p * 64 + q.lane] = t[0];
  hal[(YNCH + q.p) * 64 + q.lane] = t[L - 1];
  __syncthreads();
  const double tm = q.p > 0 ? hal[(YNCH + q.p - 1) * 64 + q.lane] : 0.0;
  const double tp = q.p < YNCH - 1 ? hal[(q.p + 1) * 64 + q.lane] : 0.0;
  // implzy on t, in place (row r is overwritten after rows r+1 is read)
  const double yend = yfwd<L>(
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
// launchers
// ---------------------------------------------------------------------------
#define YDISPATCH(Lv, CALL)                                                    \
  switch (Lv) {                                                                \
  case 4: { constexpr int L = 4; CALL; } break;                                \
  case 8: { constexpr int L = 8; CALL; } break;                                \
  case 12: { constexpr int L = 12; CALL; } break;                              \
  case 16: { constexpr int L = 16; CALL; } break;                              \
  case 20: { constexpr int L = 20; CALL; } break;                              \
  case 24: { constexpr int L = 24; CALL; } break;                              \
  default: throw std::invalid_argument("y chunk length not instantiated");    \
  }

static inline unsigned ygrid(size_t ncols) { return (unsigned)((ncols + 63) / 64); }
static inline const double *ytab(const Box &b, int k) { return b.ytab + (size_t)k * ytab_size(b.yL); }

void ky_interpyy(Box &b, const double *gen, double *geni, int loguv, int i0, int nplanes, size_t ips,
                 size_t ops) {
  const unsigned g = ygrid((size_t)b.nz1 * nplanes);
  YDISPATCH(b.yL, LAUNCH(b, "interpyy", g, YTPB, 0, interpyy_y<L>, gen, geni,
                         ytab(b, loguv == 0 ? YT_INYU : YT_INYV), b.dc.inby, loguv, b.ny, b.nz1, i0, nplanes,
                         ips, ops));
}

void ky_differzy(Box &b, const double *geni, double *res, int which, int nplanes, size_t ips, size_t ops) {
  const unsigned g = ygrid((size_t)b.nz1 * nplanes);
  YDISPATCH(b.yL, LAUNCH(b, "differzy", g, YTPB, 0, differzy_y<L>, geni, res,
                         ytab(b, which == 0 ? YT_DCYU : YT_DCYV), b.dc, which, b.ny, b.nz1, nplanes, ips, ops));
}

void ky_vistyy(Box &b, const double *u, double *res, int R, int nplanes) {
  const unsigned g = ygrid((size_t)b.nz1 * nplanes);
  YDISPATCH(b.yL, LAUNCH(b, "vistyy", g, YTPB, 0, vistyy_y<L>, u, res,
                         ytab(b, R == b.ny + 1 ? YT_VIYU : YT_VIYV), b.dc, R, b.ny, b.nz1, nplanes));
}

void ky_implzy(Box &b, double *u, const double *res, int R, int nplanes) {
  const unsigned g = ygrid((size_t)b.nz1 * nplanes);
  const bool uu = R == b.ny + 1;
  YDISPATCH(b.yL, LAUNCH(b, "implzy", g, YTPB, 0, implzy_y<L>, u, res, ytab(b, uu ? YT_IMPU : YT_IMPV),
                         uu ? b.implu : b.implv, R, b.nz1, nplanes));
}

void ky_rhsu_loop1(Box &b, const double *uv, const double *uw, double *rhsu) {
  const unsigned g = ygrid((size_t)b.nz1 * b.nx);
  YDISPATCH(b.yL, LAUNCH(b, "rhsu_loop1", g, YTPB, 0, rhsu_loop1_y<L>, uv, uw, rhsu, ytab(b, YT_DCYU), b.dc,
                         b.ny, b.nz1, b.nx));
}

void ky_rhs_loop2(Box &b, const double *vw, const double *vv, const double *ww) {
  const unsigned g = ygrid((size_t)b.nz1 * (b.nx - 2));
  YDISPATCH(b.yL, LAUNCH(b, "rhs_loop2", g, YTPB, 0, rhs_loop2_y<L>, b.rhsu, b.rhsv, b.rhsw, b.resu, b.resv,
                         b.resw, vw, vv, ww, ytab(b, YT_DCYU), ytab(b, YT_DCYV), b.dc, b.ny, b.nz1, b.nx - 2));
}

void ky_rk_final(Box &b, double *u, const double *res, const double *rhs, double *rhsp, int jee) {
  const unsigned g = ygrid((size_t)b.nz1 * b.nx);
  const bool uu = jee == b.ny + 1;
  YDISPATCH(b.yL, LAUNCH(b, "rk_final", g, YTPB, 0, rk_final_y<L>, u, res, rhs, rhsp,
                         ytab(b, uu ? YT_VIYU : YT_VIYV), ytab(b, uu ? YT_IMPU : YT_IMPV), b.dc,
                         uu ? b.implu : b.implv, b.ds, jee, b.ny, b.nz1, b.nx));
}

} // namespace otbl
