// Per-mode y Poisson solves of pois (pois.f90:102-117, soltriy
// pois.f90:211-239) for the modes kz >= 1, as partitioned tridiagonal solves
// held in registers.
//
// The pencils hold the x-transformed divergence: row j of line (kz, kx) (Re
// and Im parts, kr = 2 kz, 2 kz + 1) at P[((j-1) nz1 + kr) ld + kx], kx <
// N = nx - 1, ld the line stride.  Row j of the operator of mode (kz, kx):
//   a(1,j) x(j-1) + (a(2,j) - kz^2 - kmod(kx)) x(j) + a(3,j) x(j+1) = b(j),
// a = ayp, rows 1..ny-1 (no a(1,1), no a(3,ny-1) term).  For kz >= 1 the
// operator is strictly diagonally dominant, so chunks factor without
// pivoting.
//
// A workgroup takes 16 consecutive (kz, kx) modes, i.e. 32 real columns
// (Re and Im, one per lane of a half-wavefront); each of its 32 half-
// wavefronts takes L consecutive rows of them (chunk k: rows kL+1 .. kL+L;
// rows past ny-1 are identity rows).  A chunk factors its rows, solves them
// for the right-hand side y and for the two coupling columns V (the
// sub-diagonal entry of its first row times e_1) and W (the super-diagonal
// entry of its last row times e_L), so that x = y - V X(k-1) - W X'(k+1)
// with X(k-1) the last unknown of chunk k-1 and X'(k+1) the first of chunk
// k+1.  The first and last rows of every chunk form a 2 x 32 interface
// system, eliminated by one lane per column through LDS.  Inputs are read
// once and outputs written once.
//
// REF (FP32 pencils): after the solve, the residual b - A x of every row is
// formed in FP64 (rows of the neighbouring chunks from the interface
// values) and the FP32 solution of A d = r, through the same factors and
// spikes, is added: one step of iterative refinement.
#include "otbl_kcommon.h"

namespace otbl {

static constexpr int PK = 32;     // chunks per column = columns per workgroup
static constexpr int PROWS = 512; // rows the LDS coefficient copy holds (ny - 1 <= PROWS)

// 1 / x: the FP32 solves (followed by a refinement step whose FP64
// residual absorbs a 1-ulp pivot error) take the hardware reciprocal, off
// the serial interface chain's critical path; FP64 divides
template <typename T> __device__ __forceinline__ T prcp(T x) {
  if constexpr (sizeof(T) == 4)
    return __builtin_amdgcn_rcpf(x);
  else
    return T(1) / x;
}

template <int L, typename T, bool REF>
__global__ void __launch_bounds__(1024, sizeof(T) == 4 ? 8 : 1)
poisson_y_chunk(T *__restrict__ P, DevCoefT<T> c, DevCoef cd, int nx, int ny, int nz1, int ld) {
  // interface system per chunk and lane column: yF, yL, VF, VL, WF, WL; then
  // the interface unknowns F, L
  __shared__ T red[PK][6][PK];
  __shared__ T itf[PK][2][PK];
  // REF: the interface elimination's matrix factors of each step (VF den,
  // den, VL, the be entering the step, de), kept for the refinement's
  // second solve, whose matrix is the same
  __shared__ T fac[REF ? PK : 1][5][PK];
  // row coefficients of the operator, rows 0 .. PK L: a(1,r) (0 where absent),
  // a(2,r) (1 past ny-1), a(3,r) (0 where absent); FP64 copies for REF
  __shared__ T ca1[PK * L + 2], ca2[PK * L + 2], ca3[PK * L + 2];
  __shared__ double da1[REF ? PK * L + 2 : 1], da2[REF ? PK * L + 2 : 1], da3[REF ? PK * L + 2 : 1];
  const int N = nx - 1, nzh = nz1 / 2, ny1 = ny - 1;
  for (int r = threadIdx.x; r <= PK * L; r += blockDim.x) {
    const bool in = r >= 1 && r <= ny1;
    ca1[r] = (in && r >= 2) ? C3(c.ayp, 1, r) : T(0);
    ca2[r] = in ? C3(c.ayp, 2, r) : T(1);
    ca3[r] = (in && r < ny1) ? C3(c.ayp, 3, r) : T(0);
    if constexpr (REF) {
      da1[r] = (in && r >= 2) ? C3(cd.ayp, 1, r) : 0.0;
      da2[r] = in ? C3(cd.ayp, 2, r) : 1.0;
      da3[r] = (in && r < ny1) ? C3(cd.ayp, 3, r) : 0.0;
    }
  }
  __syncthreads();
  // lane column: the Re (even) or Im (odd) part of one of 16 (kz, kx) modes
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & (PK - 1), k = wave * 2 + (lane >> 5);
  // XCD-contiguous block order: neighbouring workgroups share the cache
  // lines at the ends of their 16-mode runs
  const size_t t = (size_t)N + (size_t)xcd_block() * (PK / 2) + (col >> 1); // kz >= 1 only
  const bool live = t < (size_t)nzh * N;
  const int kz = live ? (int)(t / N) : 1, x = live ? (int)(t % N) : 0;
  // 32-bit byte offsets from a wave-uniform base (the wave's first row, a
  // scalar 64-bit address): the two chunks of a wave span 2 L rows, which
  // stay below 4 GiB for any pencil set the launcher accepts
  const unsigned rs = (unsigned)(nz1 * ld * sizeof(T));
  const unsigned base = (unsigned)(((2 * kz + (col & 1)) * ld + x) * sizeof(T));
  const int r0 = k * L + 1;
  // rows q < nq of the chunk are operator rows (r0 + q <= ny - 1), the rest
  // identity rows: one count instead of L row indices kept live
  const int nq = ny1 - r0 + 1;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  char *Pc = reinterpret_cast<char *>(P) + (size_t)(2 * L * wave_u) * rs;
  // the chunk's first row relative to the wave's; row q adds the uniform q rs
  const unsigned b0 = base + (unsigned)((lane >> 5) * L) * rs;
  auto AT = [&](int q) -> T & { return *reinterpret_cast<T *>(Pc + (b0 + (unsigned)q * rs)); };
  const T sh = c.kaz2[kz] + c.kmod[x + 1];
  auto A1 = [&](int r) -> T { return ca1[r]; };
  auto A3 = [&](int r) -> T { return ca3[r]; };
  auto DG = [&](int q) -> T { return q < nq ? ca2[r0 + q] - sh : T(1); };

  // chunk factorisation: inverse pivots
  T ip[L];
#pragma unroll
  for (int q = 0; q < L; ++q) {
    const int r = r0 + q;
    T piv = DG(q);
    if (q > 0) piv -= A1(r) * ip[q - 1] * A3(r - 1);
    ip[q] = prcp(piv);
  }
  // right-hand side, read once
  T xv[L];
#pragma unroll
  for (int q = 0; q < L; ++q) xv[q] = (live && q < nq) ? AT(q) : T(0);
  // local solve in place
  auto local_solve = [&](T(&v)[L]) {
#pragma unroll
    for (int q = 1; q < L; ++q) v[q] -= A1(r0 + q) * ip[q - 1] * v[q - 1];
    v[L - 1] *= ip[L - 1];
#pragma unroll
    for (int q = L - 2; q >= 0; --q) v[q] = (v[q] - A3(r0 + q) * v[q + 1]) * ip[q];
  };
  local_solve(xv);
  __builtin_amdgcn_sched_barrier(0);
  // spikes V (left coupling) and W (right coupling): only their end values
  // are kept; the arrays are formed again where they are applied
  auto spike_v = [&](T(&v)[L]) {
    v[0] = A1(r0);
#pragma unroll
    for (int q = 1; q < L; ++q) v[q] = -A1(r0 + q) * ip[q - 1] * v[q - 1];
    v[L - 1] *= ip[L - 1];
#pragma unroll
    for (int q = L - 2; q >= 0; --q) v[q] = (v[q] - A3(r0 + q) * v[q + 1]) * ip[q];
  };
  auto spike_w = [&](T(&w)[L]) {
    w[L - 1] = A3(r0 + L - 1) * ip[L - 1];
#pragma unroll
    for (int q = L - 2; q >= 0; --q) w[q] = -A3(r0 + q) * w[q + 1] * ip[q];
  };
  T VF, VL, WF, WL;
  {
    T v[L];
    spike_v(v);
    VF = v[0];
    VL = v[L - 1];
  }
  {
    T w[L];
    spike_w(w);
    WF = w[0];
    WL = w[L - 1];
  }
  // x -= V XL + W XF, one spike array alive at a time
  auto combine = [&](T(&v_)[L], T XL, T XF) {
    {
      T v[L];
      spike_v(v);
#pragma unroll
      for (int q = 0; q < L; ++q) v_[q] -= v[q] * XL;
    }
    {
      T w[L];
      spike_w(w);
#pragma unroll
      for (int q = 0; q < L; ++q) v_[q] -= w[q] * XF;
    }
  };
  // interface system: one lane per lane column (wave 0, lanes 0..31)
  // eliminates over the chunks; every chunk then reads the unknowns it
  // couples to
  auto interface_solve = [&](T yF, T yL, T &XL, T &XF) {
    red[k][0][col] = yF;
    red[k][1][col] = yL;
    red[k][2][col] = VF;
    red[k][3][col] = VL;
    red[k][4][col] = WF;
    red[k][5][col] = WL;
    __syncthreads();
    if (wave == 0 && lane < PK) {
      // L(q) = al(q) + be(q) F(q+1), F(q) = ga(q) + de(q) F(q+1); the LDS
      // reads of step q+1 are issued before step q's arithmetic (the sweep
      // is a chain of dependent steps run by one half-wave)
      T al = 0, be = 0;
      T n0 = red[0][0][col], n1 = red[0][1][col], n2 = red[0][2][col], n3 = red[0][3][col];
      T n4 = red[0][4][col], n5 = red[0][5][col];
#pragma unroll 2
      for (int q = 0; q < PK; ++q) {
        const T qyF = n0, qyL = n1, qVF = n2, qVL = n3, qWF = n4, qWL = n5;
        if (q + 1 < PK) {
          n0 = red[q + 1][0][col];
          n1 = red[q + 1][1][col];
          n2 = red[q + 1][2][col];
          n3 = red[q + 1][3][col];
          n4 = red[q + 1][4][col];
          n5 = red[q + 1][5][col];
        }
        const T den = prcp(T(1) + qVF * be);
        const T ga = (qyF - qVF * al) * den, de = -qWF * den;
        const T nal = qyL - qVL * (al + be * ga), nbe = -qVL * be * de - qWL;
        if constexpr (REF) {
          fac[q][0][col] = qVF * den;
          fac[q][1][col] = den;
          fac[q][2][col] = qVL;
          fac[q][3][col] = be;
          fac[q][4][col] = de;
        }
        red[q][0][col] = ga;
        red[q][1][col] = de;
        red[q][2][col] = nal;
        red[q][3][col] = nbe;
        al = nal;
        be = nbe;
      }
      T Fn = 0; // F(q+1)
      T m0 = red[PK - 1][0][col], m1 = red[PK - 1][1][col], m2 = red[PK - 1][2][col], m3 = red[PK - 1][3][col];
#pragma unroll 2
      for (int q = PK - 1; q >= 0; --q) {
        const T ga = m0, de = m1, nal = m2, nbe = m3;
        if (q > 0) {
          m0 = red[q - 1][0][col];
          m1 = red[q - 1][1][col];
          m2 = red[q - 1][2][col];
          m3 = red[q - 1][3][col];
        }
        itf[q][1][col] = nal + nbe * Fn; // L(q)
        Fn = ga + de * Fn;
        itf[q][0][col] = Fn; // F(q)
      }
    }
    __syncthreads();
    XL = k > 0 ? itf[k - 1][1][col] : T(0);
    XF = k < PK - 1 ? itf[k + 1][0][col] : T(0);
    __syncthreads(); // red / itf are reused by a second solve
  };
  // the same interface system with another right-hand side, through the
  // factors the first elimination kept (no division on the serial chain)
  auto interface_rhs = [&](T yF, T yL, T &XL, T &XF) {
    red[k][0][col] = yF;
    red[k][1][col] = yL;
    __syncthreads();
    if (wave == 0 && lane < PK) {
      T al = 0;
      T n0 = red[0][0][col], n1 = red[0][1][col];
      T f0 = fac[0][0][col], f1 = fac[0][1][col], f2 = fac[0][2][col], f3 = fac[0][3][col];
#pragma unroll 2
      for (int q = 0; q < PK; ++q) {
        const T qyF = n0, qyL = n1, vden = f0, den = f1, qVL = f2, be = f3;
        if (q + 1 < PK) {
          n0 = red[q + 1][0][col];
          n1 = red[q + 1][1][col];
          f0 = fac[q + 1][0][col];
          f1 = fac[q + 1][1][col];
          f2 = fac[q + 1][2][col];
          f3 = fac[q + 1][3][col];
        }
        const T ga = qyF * den - vden * al;
        const T nal = qyL - qVL * (al + be * ga);
        red[q][0][col] = ga;
        red[q][2][col] = nal;
        al = nal;
      }
      // the reads of step q-1 are issued before step q's arithmetic
      T Fn = 0; // F(q+1)
      T m0 = red[PK - 1][0][col], m2 = red[PK - 1][2][col], m4 = fac[PK - 1][4][col], m3 = T(0);
#pragma unroll 2
      for (int q = PK - 1; q >= 0; --q) {
        const T ga = m0, nal = m2, de = m4, nbe = m3; // nbe: the be leaving step q
        if (q > 0) {
          m0 = red[q - 1][0][col];
          m2 = red[q - 1][2][col];
          m4 = fac[q - 1][4][col];
          m3 = fac[q][3][col];
        }
        itf[q][1][col] = nal + nbe * Fn; // L(q)
        Fn = ga + de * Fn;
        itf[q][0][col] = Fn; // F(q)
      }
    }
    __syncthreads();
    XL = k > 0 ? itf[k - 1][1][col] : T(0);
    XF = k < PK - 1 ? itf[k + 1][0][col] : T(0);
  };
  T XL, XF;
  interface_solve(xv[0], xv[L - 1], XL, XF);
  combine(xv, XL, XF);
  if constexpr (REF) {
    // residual of every row in FP64, b read again
    const double shd = cd.kaz2[kz] + cd.kmod[x + 1];
    T d[L];
#pragma unroll
    for (int q = 0; q < L; ++q) {
      // one row at a time: its loads are not hoisted above the previous row
      __builtin_amdgcn_sched_barrier(0);
      const int r = r0 + q;
      const bool in = live && q < nq;
      const double b = in ? (double)AT(q) : 0.0;
      const double pm = q > 0 ? (double)xv[q - 1] : (double)XL, pp = q < L - 1 ? (double)xv[q + 1] : (double)XF;
      const double dg = q < nq ? da2[r] - shd : 1.0;
      d[q] = in ? (T)(b - (da1[r] * pm + dg * (double)xv[q] + da3[r] * pp)) : T(0);
    }
    __builtin_amdgcn_sched_barrier(0);
    local_solve(d);
    T dL, dF;
    interface_rhs(d[0], d[L - 1], dL, dF);
#pragma unroll
    for (int q = 0; q < L; ++q) xv[q] += d[q];
    combine(xv, dL, dF);
  }
#pragma unroll
  for (int q = 0; q < L; ++q)
    if (live && q < nq) AT(q) = xv[q];
}

// rows per chunk for ny-1 rows (0: not instantiated)
int kp_chunk_rows(int ny) {
  if (ny - 1 > PROWS) return 0;
  const int need = (ny - 1 + PK - 1) / PK;
  for (int L : {2, 4, 6, 8, 10, 12, 14, 16})
    if (L >= need) return L;
  return 0;
}

template <typename T, bool REF> void kp_poisson_y(Box &b, T *P) {
  const int N = b.nx - 1, nzh = b.nz1 / 2;
  const size_t cols = (size_t)(nzh - 1) * N;
  if (cols == 0) return;
  // a wave addresses 2 L rows plus one row of offsets in 32 bits
  if ((size_t)(2 * kp_chunk_rows(b.ny) + 1) * b.nz1 * b.pld * sizeof(T) >= (size_t(1) << 32))
    throw std::invalid_argument("Poisson pencil rows too long for 32-bit offsets");
  const unsigned grid = (unsigned)((cols + PK / 2 - 1) / (PK / 2)); // 16 modes (Re, Im) per workgroup
  switch (kp_chunk_rows(b.ny)) {
#define PCASE(LL)                                                                                          \
  case LL: {                                                                                               \
    auto kf = &poisson_y_chunk<LL, T, REF>;                                                                \
    LAUNCH(b, "poisson_y", grid, 1024, 0, kf, P, coefs<T>(b), b.dc, b.nx, b.ny, b.nz1, b.pld);             \
  } break;
    PCASE(2)
    PCASE(4)
    PCASE(6)
    PCASE(8)
    PCASE(10)
    PCASE(12)
    PCASE(14)
    PCASE(16)
#undef PCASE
  default: throw std::invalid_argument("Poisson chunk length not instantiated");
  }
}

template void kp_poisson_y<double, false>(Box &, double *);
template void kp_poisson_y<float, true>(Box &, float *);

} // namespace otbl
