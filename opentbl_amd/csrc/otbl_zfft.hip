// z-direction product kernels of rhsp on the register-resident FFT engine
// (otbl_zfft.h): one (j, i) row per team of W wavefronts, every transform of
// the row held in registers, products formed in physical space without data
// movement, one LDS exchange buffer per team.
#include "otbl_kcommon.h"
#include "otbl_zfft.h"

#include <type_traits>

namespace otbl {

// exp(2 pi i m / n), m = 0..n-1
__global__ void ztw_kernel(double2 *tw, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(2.0 * m / n, &s, &c);
  tw[m] = make_double2(c, s);
}

void kz_twiddles(Box &b, double *tw, int n) { LAUNCH(b, "ztw", nblk(n), TPB, 0, ztw_kernel, (double2 *)tw, n); }

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Plane loop 1 of rhsp (rhsp.F90:120-153) for one (j, i) row: from u, u_x
// (wki1), w_x (wki3), v_x (wki2, rows <= ny) and u interpolated to v rows
// (uy) in spectral space form r2c(u_x^2) -> wki1, r2c(u w_x) -> wki3,
// r2c(v_x uy) -> wki2 (rows <= ny), and the maxima |u|, |w_x|, |v_x|(j)
// of the CFL time step.  Rows are numbered (i-1)(ny+1) + (j-1).
template <int N, typename R>
__global__ void __launch_bounds__(ZTeam<N>::TPB_)
products1_z(const R *__restrict__ u, R *wki1, R *wki2, R *wki3, const R *__restrict__ uy,
            const typename CT<R>::c *__restrict__ tw, const R *__restrict__ zrow, int ny, int nz1, int nz2, int nrows,
            double *rowmax) {
  using C = typename CT<R>::c;
  using Z = ZTeam<N>;
  __shared__ C lds_all[Z::TEAMS * zlds_elems<N>()];
  // the team (and so the row and every row pointer) is uniform per wavefront
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / Z::G), t = threadIdx.x % Z::G;
  C *lds = lds_all + team * zlds_elems<N>();
  const int row = blockIdx.x * Z::TEAMS + team;
  if (row >= nrows) return; // a whole team (one wavefront when TEAMS > 1)
  const int j = row % (ny + 1) + 1, i = row / (ny + 1) + 1;
  const size_t ou = (size_t)row * nz1, ov = ((size_t)(i - 1) * ny + (j - 1)) * nz1;
  const bool vrow = j <= ny;
  // Pairs chosen so that almost nothing is held across a transform:
  // (v_x, u_y) gives v_x u_y, inverted at once; (u, w_x) gives u w_x, kept
  // through the transform of u_x, which gives u_x^2.
  // In FP32 the mean flow (mode 0) of u_x and u_y, O(1) against O(1e-3)
  // fluctuations, stays out of the transforms: their round-off scales with
  // the signal, and in u_x^2 it would reach the mean mode and, through
  // d/dx, the x-momentum balance.  With U the mean, u = U + u':
  // u_x^2 = U^2 + 2 U u_x' + u_x'^2 and v_x u_y = U_y v_x + v_x u_y', the
  // U terms added in spectral space.
  constexpr bool SPLIT = sizeof(R) == 4;
  R um = 0.0, wm = 0.0, vm = 0.0;
  if (vrow) {
    C Q[Z::E];
    zload_spectral<N>(Q, t, wki2 + ov, uy + ou);
    const R Uy = SPLIT ? uy[ou] : R(0);
    if (SPLIT && t == 0) Q[0].y = R(0); // k = 0 sits in lane 0, slot 0
    zrun<N, +1, false>(Q, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) {
      vm = fmax(vm, fabs(Q[e].x));
      Q[e] = mkc<C>(Q[e].x * Q[e].y, R(0));
    }
    zrun<N, -1, true>(Q, t, tw, lds);
    if constexpr (SPLIT)
      zstore_spectral<N, R>(Q, t, lds, wki2 + ov, nullptr, nz2, wki2 + ov, Uy, Uy);
    else
      zstore_spectral<N, R>(Q, t, lds, wki2 + ov, nullptr, nz2);
  }
  R s1[Z::E];
  {
    C A[Z::E];
    zload_spectral<N>(A, t, u + ou, wki3 + ou);
    zrun<N, +1, false>(A, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) {
      um = fmax(um, fabs(A[e].x));
      wm = fmax(wm, fabs(A[e].y));
      s1[e] = A[e].x * A[e].y;
    }
  }
  {
    C B[Z::E];
    zload_spectral<N>(B, t, wki1 + ou, zrow);
    const R Ux = SPLIT ? wki1[ou] : R(0);
    if (SPLIT && t == 0) B[0].x = R(0);
    zrun<N, +1, false>(B, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) B[e] = mkc<C>(B[e].x * B[e].x, s1[e]);
    zrun<N, -1, true>(B, t, tw, lds);
    if constexpr (SPLIT)
      zstore_spectral<N, R>(B, t, lds, wki1 + ou, wki3 + ou, nz2, wki1 + ou, Ux, R(2) * Ux);
    else
      zstore_spectral<N, R>(B, t, lds, wki1 + ou, wki3 + ou, nz2);
  }
  // per-row maxima; rowmax_kernel reduces them (no atomics on shared
  // addresses from every row)
  um = wave_max(um);
  wm = wave_max(wm);
  vm = wave_max(vm);
  if constexpr (Z::W > 1) { // combine the team's wavefronts through its LDS buffer
    if ((t & 63) == 0) lds[t / 64] = mkc<C>(um, wm), lds[Z::W + t / 64] = mkc<C>(vm, R(0));
    zsync<Z::W>();
    if (t == 0)
      for (int q = 1; q < Z::W; ++q) {
        um = fmax(um, lds[q].x);
        wm = fmax(wm, lds[q].y);
        vm = fmax(vm, lds[Z::W + q].x);
      }
  }
  if (t == 0) {
    rowmax[row] = um;
    rowmax[nrows + row] = wm;
    rowmax[2 * (size_t)nrows + row] = vrow ? (double)vm : 0.0;
  }
}

// Plane loop 1 in FP32 for two rows at once (rows a = 2 pr, b = 2 pr + 1 of
// the numbering above): the FFTs that products1_z fills with one real
// signal (v_x u_y inverted alone, u_x forward with an empty partner) carry
// the second row's signal instead, 8 transforms per two rows against 10.
// Each product's two rows share the inverse transform: (u w_x of a, of b),
// (u_x^2 of a, of b), (v_x u_y of a, of b), so at most one row's product is
// held across a transform.  The mean flow of u_x and u_y stays out of the
// transforms as in products1_z.
template <int N, int MINB>
__global__ void __launch_bounds__(ZTeam<N>::TPB_, MINB)
products1_zp(const float *__restrict__ u, float *wki1, float *wki2, float *wki3, const float *__restrict__ uy,
             const float2 *__restrict__ tw, const float *__restrict__ zrow, int ny, int nz1, int nz2, int nrows,
             double *rowmax) {
  using C = float2;
  using Z = ZTeam<N>;
  __shared__ C lds_all[Z::TEAMS * zlds_elems<N>()];
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / Z::G), t = threadIdx.x % Z::G;
  C *lds = lds_all + team * zlds_elems<N>();
  const int pr = blockIdx.x * Z::TEAMS + team;
  const int ra = 2 * pr, rb = ra + 1;
  if (ra >= nrows) return; // a whole team
  const bool hasb = rb < nrows;
  const int ja = ra % (ny + 1) + 1, ia = ra / (ny + 1) + 1;
  const int jb = hasb ? rb % (ny + 1) + 1 : ny + 1, ib = hasb ? rb / (ny + 1) + 1 : ia;
  const size_t oua = (size_t)ra * nz1, oub = (size_t)(hasb ? rb : ra) * nz1;
  const size_t ova = ((size_t)(ia - 1) * ny + (ja - 1)) * nz1, ovb = ((size_t)(ib - 1) * ny + (jb - 1)) * nz1;
  const bool va = ja <= ny, vb = hasb && jb <= ny;
  float uma = 0.f, wma = 0.f, vma = 0.f, umb = 0.f, wmb = 0.f, vmb = 0.f;
  // ---- u w_x of both rows ----
  {
    float sa[Z::E];
    {
      C A[Z::E];
      zload_spectral<N>(A, t, u + oua, wki3 + oua);
      zrun<N, +1, false>(A, t, tw, lds);
#pragma unroll
      for (int e = 0; e < Z::E; ++e) {
        uma = fmaxf(uma, fabsf(A[e].x));
        wma = fmaxf(wma, fabsf(A[e].y));
        sa[e] = A[e].x * A[e].y;
      }
    }
    C S[Z::E];
    zload_spectral<N>(S, t, u + oub, hasb ? wki3 + oub : zrow);
    zrun<N, +1, false>(S, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) {
      umb = fmaxf(umb, fabsf(S[e].x));
      wmb = fmaxf(wmb, fabsf(S[e].y));
      S[e] = make_float2(sa[e], S[e].x * S[e].y);
    }
    zrun<N, -1, true>(S, t, tw, lds);
    zstore_spectral2<N, float>(S, t, lds, wki3 + oua, hasb ? wki3 + oub : nullptr, nz2, nullptr, 0.f, 0.f, nullptr,
                               0.f, 0.f);
  }
  // ---- u_x^2 of both rows: one forward transform of (u_x a, u_x b) ----
  {
    C B[Z::E];
    zload_spectral<N>(B, t, wki1 + oua, hasb ? wki1 + oub : zrow);
    const float Uxa = wki1[oua], Uxb = hasb ? wki1[oub] : 0.f;
    if (t == 0) B[0] = make_float2(0.f, 0.f); // the means of both u_x (k = 0: lane 0, slot 0)
    zrun<N, +1, false>(B, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) B[e] = make_float2(B[e].x * B[e].x, B[e].y * B[e].y);
    zrun<N, -1, true>(B, t, tw, lds);
    zstore_spectral2<N, float>(B, t, lds, wki1 + oua, hasb ? wki1 + oub : nullptr, nz2, wki1 + oua, Uxa, 2.f * Uxa,
                               hasb ? wki1 + oub : nullptr, Uxb, 2.f * Uxb);
  }
  // ---- v_x u_y of both rows ----
  if (va || vb) {
    float pa[Z::E];
#pragma unroll
    for (int e = 0; e < Z::E; ++e) pa[e] = 0.f;
    if (va) {
      C Q[Z::E];
      zload_spectral<N>(Q, t, wki2 + ova, uy + oua);
      if (t == 0) Q[0].y = 0.f; // the mean of u_y
      zrun<N, +1, false>(Q, t, tw, lds);
#pragma unroll
      for (int e = 0; e < Z::E; ++e) {
        vma = fmaxf(vma, fabsf(Q[e].x));
        pa[e] = Q[e].x * Q[e].y;
      }
    }
    C P[Z::E];
    if (vb) {
      zload_spectral<N>(P, t, wki2 + ovb, uy + oub);
      if (t == 0) P[0].y = 0.f;
      zrun<N, +1, false>(P, t, tw, lds);
#pragma unroll
      for (int e = 0; e < Z::E; ++e) {
        vmb = fmaxf(vmb, fabsf(P[e].x));
        P[e] = make_float2(pa[e], P[e].x * P[e].y);
      }
    } else {
#pragma unroll
      for (int e = 0; e < Z::E; ++e) P[e] = make_float2(pa[e], 0.f);
    }
    zrun<N, -1, true>(P, t, tw, lds);
    const float Uya = va ? uy[oua] : 0.f, Uyb = vb ? uy[oub] : 0.f;
    zstore_spectral2<N, float>(P, t, lds, va ? wki2 + ova : nullptr, vb ? wki2 + ovb : nullptr, nz2,
                               va ? wki2 + ova : nullptr, Uya, Uya, vb ? wki2 + ovb : nullptr, Uyb, Uyb);
  }
  // per-row maxima (reduced by rowmax_kernel)
  float mx[6] = {uma, wma, vma, umb, wmb, vmb};
#pragma unroll
  for (int q = 0; q < 6; ++q)
    for (int o = 32; o > 0; o >>= 1) mx[q] = fmaxf(mx[q], __shfl_down(mx[q], o, 64));
  if constexpr (Z::W > 1) {
    if ((t & 63) == 0)
      for (int q = 0; q < 6; ++q) lds[q * Z::W + t / 64] = make_float2(mx[q], 0.f);
    zsync<Z::W>();
    if (t == 0)
      for (int q = 0; q < 6; ++q)
        for (int w = 1; w < Z::W; ++w) mx[q] = fmaxf(mx[q], lds[q * Z::W + w].x);
  }
  if (t == 0) {
    rowmax[ra] = mx[0];
    rowmax[nrows + ra] = mx[1];
    rowmax[2 * (size_t)nrows + ra] = va ? (double)mx[2] : 0.0;
    if (hasb) {
      rowmax[rb] = mx[3];
      rowmax[nrows + rb] = mx[4];
      rowmax[2 * (size_t)nrows + rb] = vb ? (double)mx[5] : 0.0;
    }
  }
}

// max |u|, max |w_x| over the box and max |v_x| per row j from the per-row
// maxima of products1_z: block j-1 reduces rows (i, j) over the planes.
__global__ void __launch_bounds__(256)
rowmax_kernel(const double *__restrict__ rowmax, DevScalars *ds, unsigned long long *vmbits, int ny, int nx) {
  const int j = blockIdx.x + 1, nrows = (ny + 1) * nx;
  double um = 0.0, wm = 0.0, vm = 0.0;
  for (int i = threadIdx.x + 1; i <= nx; i += blockDim.x) {
    const int row = (i - 1) * (ny + 1) + (j - 1);
    um = fmax(um, rowmax[row]);
    wm = fmax(wm, rowmax[nrows + row]);
    vm = fmax(vm, rowmax[2 * (size_t)nrows + row]);
  }
  __shared__ double red[3][4];
  um = wave_max(um);
  wm = wave_max(wm);
  vm = wave_max(vm);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x / 64] = um;
    red[1][threadIdx.x / 64] = wm;
    red[2][threadIdx.x / 64] = vm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; ++q) {
      um = fmax(um, red[0][q]);
      wm = fmax(wm, red[1][q]);
      vm = fmax(vm, red[2][q]);
    }
    atomicMax(&ds->umax_bits, (unsigned long long)__double_as_longlong(um));
    atomicMax(&ds->wmax_bits, (unsigned long long)__double_as_longlong(wm));
    if (j <= ny) atomicMax(&vmbits[j - 1], (unsigned long long)__double_as_longlong(vm));
  }
}

// Plane loop 2 of rhsp (rhsp.F90:367-415) for one (j, i) row, j <= ny,
// planes 2..nx-1 (rows numbered (i-2) ny + (j-1)): vw = r2c(v w_v) -> ya
// (w interpolated to v rows is read from ya), vv = r2c(v_u^2) -> yb (v
// interpolated to u rows is read from yb), ww = r2c(w^2) -> yc.
template <int N, typename R>
__global__ void __launch_bounds__(ZTeam<N>::TPB_, sizeof(R) == 4 ? 4 : 1)
products2_z(const R *__restrict__ resv, const R *__restrict__ resw, R *ya, R *yb, R *yc,
            const typename CT<R>::c *__restrict__ tw, int ny, int nz1, int nz2, int nrows) {
  using C = typename CT<R>::c;
  using Z = ZTeam<N>;
  __shared__ C lds_all[Z::TEAMS * zlds_elems<N>()];
  // the team (and so the row and every row pointer) is uniform per wavefront
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / Z::G), t = threadIdx.x % Z::G;
  C *lds = lds_all + team * zlds_elems<N>();
  const int row = blockIdx.x * Z::TEAMS + team;
  if (row >= nrows) return;
  const int j = row % ny + 1, i = row / ny + 2;
  const size_t ou = ((size_t)(i - 1) * (ny + 1) + (j - 1)) * nz1;
  const size_t ov = ((size_t)(i - 1) * ny + (j - 1)) * nz1;
  // (v_u, w) gives (v_u^2, w^2), inverted at once; (v, w_v) gives v w_v
  {
    C B[Z::E];
    zload_spectral<N>(B, t, yb + ou, resw + ou);
    zrun<N, +1, false>(B, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) B[e] = mkc<C>(B[e].x * B[e].x, B[e].y * B[e].y);
    zrun<N, -1, true>(B, t, tw, lds);
    zstore_spectral<N, R>(B, t, lds, yb + ou, yc + ou, nz2);
  }
  {
    C A[Z::E];
    zload_spectral<N>(A, t, resv + ov, ya + ou);
    zrun<N, +1, false>(A, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) A[e] = mkc<C>(A[e].x * A[e].y, R(0));
    zrun<N, -1, true>(A, t, tw, lds);
    zstore_spectral<N, R>(A, t, lds, ya + ou, nullptr, nz2);
  }
}

// fourxz (ffts.f90:27-67) on independent rows: iopt >= 0 c2r (spectral ->
// physical, nz doubles per row), iopt < 0 r2c.
template <int N>
__global__ void __launch_bounds__(ZTeam<N>::TPB_)
fourxz_z(double *fou, double *phys, int iopt, const double2 *__restrict__ tw, const double *__restrict__ zrow,
         int nz1, int nz2, int nrows) {
  using Z = ZTeam<N>;
  __shared__ double2 lds_all[Z::TEAMS * zlds_elems<N>()];
  // the team (and so the row and every row pointer) is uniform per wavefront
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / Z::G), t = threadIdx.x % Z::G;
  double2 *lds = lds_all + team * zlds_elems<N>();
  const int row = blockIdx.x * Z::TEAMS + team;
  if (row >= nrows) return;
  double *F = fou + (size_t)row * nz1;
  double *Ph = phys + (size_t)row * N;
  double2 A[Z::E];
  if (iopt >= 0) {
    zload_spectral<N>(A, t, F, zrow);
    zrun<N, +1, false>(A, t, tw, lds);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) Ph[zphys_index<N>(t, e)] = A[e].x;
  } else {
#pragma unroll
    for (int e = 0; e < Z::E; ++e) A[e] = make_double2(Ph[zphys_index<N>(t, e)], 0.0);
    zrun<N, -1, true>(A, t, tw, lds);
    zstore_spectral<N, double>(A, t, lds, F, nullptr, nz2);
  }
}

// base[idx] through a 32-bit byte offset (idx sizeof(T) < 2^32): the load
// takes a scalar base and one offset register instead of a 64-bit address
template <typename T> __device__ __forceinline__ T ld32(const T *base, unsigned idx) {
  return *(const T *)((const char *)base + idx * (unsigned)sizeof(T));
}

// Poisson x transforms (ffts.f90:124-181: FFTW REDFT10 / REDFT01 of length
// N = nx-1) on x-contiguous pencils, two real lines per complex transform:
// Makhoul's reordering turns each DCT into a complex DFT of length N, and
// Bluestein's chirp-z evaluates that DFT as a circular convolution of
// length M (>= 2N-1, so N <= M/2) on the register engine: forward
// transform (sign -1), product with the transformed conjugate chirp bf at
// each register's physical index, inverse transform (sign +1), then the
// chirp and the Makhoul untwisting through the team's LDS buffer.
// IOPT < 0: DCT-II x 1/N, IOPT > 0: DCT-III x 0.5; in place (src == dst)
// allowed, each team reads its two lines before writing them.
// Register slot m R0 + r of the first pass holds index t + G m + r M/R0;
// slots with G m + r M/R0 >= M/2 are zero on input (the forward pass skips
// them) and unused on output (the inverse's last pass never forms them).
// Every global load of a phase is issued before its arithmetic, with
// clamped indices and selects in place of branches.
template <int M, typename R, int IOPT>
__global__ void __launch_bounds__(ZTeam<M>::TPB_, sizeof(R) == 4 ? 6 : 1)
dct_rows_z(const R *src, R *dst, int Lp, int N, int ld, const typename CT<R>::c *__restrict__ chirp,
           const typename CT<R>::c *__restrict__ bf, const typename CT<R>::c *__restrict__ tw4,
           const typename CT<R>::c *__restrict__ tw) {
  using C = typename CT<R>::c;
  using Z = ZTeam<M>;
  constexpr int R0 = ZPlan<M>::rad(0), NB0 = M / (R0 * Z::G), NO = (M / 2 + Z::G - 1) / Z::G;
  __shared__ C lds_all[Z::TEAMS * zlds_elems<M>()];
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / Z::G);
  int t = threadIdx.x % Z::G;
  C *lds = lds_all + team * zlds_elems<M>();
  const int pair = blockIdx.x * Z::TEAMS + team;
  const int l0 = 2 * pair, l1 = l0 + 1;
  if (l0 >= Lp) return; // a whole team
  const bool has1 = l1 < Lp;
  const R *x0 = src + (size_t)l0 * ld, *x1 = src + (size_t)(has1 ? l1 : l0) * ld;
  const int half = (N + 1) / 2;
  // live slots of butterfly m: r < RH (G m + r M/R0 < M/2 exactly then,
  // as G NB0 = M / R0)
  constexpr int RH = R0 / 2, NL = NB0 * RH;
  auto kof = [&](int s) { return t + Z::G * (s / RH) + (s % RH) * (M / R0); }; // live slot s -> index
  C v[Z::E];
  {
    R a0[NL], a1[NL], b0[NL], b1[NL];
    C ch[NL], tt[NL];
#pragma unroll
    for (int s = 0; s < NL; ++s) {
      const unsigned k = min(kof(s), N - 1);
      if constexpr (IOPT < 0) { // Makhoul reordering of the two real lines
        const unsigned sidx = (int)k < half ? 2 * k : 2 * (N - 1 - k) + 1;
        a0[s] = ld32(x0, sidx);
        a1[s] = ld32(x1, sidx);
      } else {
        const unsigned nk = k ? N - k : 0;
        a0[s] = ld32(x0, k);
        a1[s] = ld32(x1, k);
        b0[s] = ld32(x0, nk);
        b1[s] = ld32(x1, nk);
        tt[s] = ld32(tw4, k);
      }
      ch[s] = ld32(chirp, k);
    }
#pragma unroll
    for (int s = 0; s < NL; ++s) {
      const int k = kof(s);
      C z;
      if constexpr (IOPT < 0) {
        z = mkc<C>(a0[s], has1 ? a1[s] : R(0));
      } else { // V = (X_k - i X_{N-k}) exp(+i pi k / 2N), Z = V0 + i V1, conj(Z)
        C w = tt[s];
        w.y = -w.y;
        const R c0 = k ? b0[s] : R(0), c1 = (has1 && k) ? b1[s] : R(0), d1 = has1 ? a1[s] : R(0);
        const C v0 = cmulc(mkc<C>(a0[s], -c0), w), v1 = cmulc(mkc<C>(d1, -c1), w);
        z = mkc<C>(v0.x - v1.y, -(v0.y + v1.x));
      }
      const C a = cmulc(z, ch[s]);
      v[(s / RH) * R0 + s % RH] = k < N ? a : mkc<C>(R(0), R(0));
    }
  }
  zrun<M, -1, false, true>(v, t, tw, lds);
  {
    C w[Z::E];
#pragma unroll
    for (int e = 0; e < Z::E; ++e) w[e] = ld32(bf, (unsigned)zphys_index<M>(t, e));
#pragma unroll
    for (int e = 0; e < Z::E; ++e) v[e] = cmulc(v[e], w[e]);
  }
  zrun<M, +1, true>(v, t, tw, lds);
  // natural-order outputs t + G m + q M / R0 (q < R0 / 2 hold every k < N);
  // keep k < N times the chirp.  The load offsets are formed again here
  // rather than kept from the input phase (kept, they spill).
  asm volatile("" : "+v"(t));
  {
    C ch[NL];
#pragma unroll
    for (int s = 0; s < NL; ++s) ch[s] = ld32(chirp, (unsigned)min(kof(s), N - 1));
#pragma unroll
    for (int s = 0; s < NL; ++s) {
      const int k = kof(s);
      if (k < N) lds[zpad(k)] = cmulc(ch[s], v[(s / RH) * R0 + s % RH]);
    }
  }
  zsync<Z::W>();
  R *y0 = dst + (size_t)l0 * ld, *y1 = dst + (size_t)l1 * ld;
  if constexpr (IOPT < 0) {
    const R sc = R(2.0 / N);
    C tt[NO];
#pragma unroll
    for (int it = 0; it < NO; ++it) tt[it] = ld32(tw4, (unsigned)min(t + Z::G * it, N - 1));
#pragma unroll
    for (int it = 0; it < NO; ++it) {
      const int k = t + Z::G * it;
      if (k < N) {
        const C z = lds[zpad(k)], w = lds[zpad(k ? N - k : 0)];
        const C v0 = mkc<C>(R(0.5) * (z.x + w.x), R(0.5) * (z.y - w.y));
        const C v1 = mkc<C>(R(0.5) * (z.y + w.y), R(-0.5) * (z.x - w.x));
        y0[k] = sc * (tt[it].x * v0.x - tt[it].y * v0.y);
        if (has1) y1[k] = sc * (tt[it].x * v1.x - tt[it].y * v1.y);
      }
    }
  } else {
#pragma unroll
    for (int it = 0; it < NO; ++it) {
      const int k = t + Z::G * it;
      if (k < N) {
        const C zc = lds[zpad(k)]; // conj of sum_k Z_k exp(+2 pi i k n / N)
        const int o = k < half ? 2 * k : 2 * (N - 1 - k) + 1;
        y0[o] = R(0.5) * zc.x;
        if (has1) y1[o] = R(-0.5) * zc.y;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define ZDISPATCH(n, ...)                                                      \
  switch (n) {                                                                 \
  case 256: { constexpr int N = 256; __VA_ARGS__; } break;                     \
  case 512: { constexpr int N = 512; __VA_ARGS__; } break;                     \
  case 1024: { constexpr int N = 1024; __VA_ARGS__; } break;                   \
  case 2048: { constexpr int N = 2048; __VA_ARGS__; } break;                   \
  case 4096: { constexpr int N = 4096; __VA_ARGS__; } break;                   \
  default: throw std::invalid_argument("z transform length not instantiated"); \
  }

template <int N> static unsigned zgrid(int nrows) {
  return (unsigned)((nrows + ZTeam<N>::TEAMS - 1) / ZTeam<N>::TEAMS);
}

// the engine fixes the dealiased mode count by the length
static void zcheck(const Box &b) {
  if (b.nz2 != b.nz / 3 - 1 || b.nz1 != 2 * (b.nz / 3))
    throw std::invalid_argument("z transform: nz1/nz2 do not follow nz");
}

template <typename R> static const typename CT<R>::c *ztw(const Box &b);
template <> const double2 *ztw<double>(const Box &b) { return (const double2 *)b.ztw; }
template <> const float2 *ztw<float>(const Box &b) { return (const float2 *)b.ztw_f; }

template <typename R> void kz_products1(Box &b) {
  zcheck(b);
  const int nrows = (b.ny + 1) * b.nx;
  const R *zrow = fld<R>(b.zrow);
  if (std::is_same<R, float>::value && b.zpair) {
    // FP32: two rows per team (the registers two rows need are there in FP32)
    const float *zr = (const float *)b.zrow;
    if (b.zpair == 1) {
      ZDISPATCH(b.nz, auto kf = &products1_zp<N, 3>;
                LAUNCH(b, "products1", zgrid<N>((nrows + 1) / 2), ZTeam<N>::TPB_, 0, kf, fld<float>(b.f[OTBL_U]),
                       fld<float>(b.wki1), fld<float>(b.wki2), fld<float>(b.wki3), fld<float>(b.yc),
                       (const float2 *)b.ztw_f, zr, b.ny, b.nz1, b.nz2, nrows, b.rowmax));
    } else {
      ZDISPATCH(b.nz, auto kf = &products1_zp<N, 1>;
                LAUNCH(b, "products1", zgrid<N>((nrows + 1) / 2), ZTeam<N>::TPB_, 0, kf, fld<float>(b.f[OTBL_U]),
                       fld<float>(b.wki1), fld<float>(b.wki2), fld<float>(b.wki3), fld<float>(b.yc),
                       (const float2 *)b.ztw_f, zr, b.ny, b.nz1, b.nz2, nrows, b.rowmax));
    }
  } else {
    ZDISPATCH(b.nz, auto kf = &products1_z<N, R>;
              LAUNCH(b, "products1", zgrid<N>(nrows), ZTeam<N>::TPB_, 0, kf, fld<R>(b.f[OTBL_U]), fld<R>(b.wki1),
                     fld<R>(b.wki2), fld<R>(b.wki3), fld<R>(b.yc), ztw<R>(b), zrow, b.ny, b.nz1, b.nz2, nrows,
                     b.rowmax));
  }
  kz_rowmax(b);
}

void kz_rowmax(Box &b) {
  LAUNCH(b, "rowmax", b.ny + 1, 256, 0, rowmax_kernel, b.rowmax, b.ds, b.vmbits, b.ny, b.nx);
}

template <typename R> void kz_products2(Box &b) {
  zcheck(b);
  const int nrows = b.ny * (b.nx - 2);
  ZDISPATCH(b.nz, auto kf = &products2_z<N, R>;
            LAUNCH(b, "products2", zgrid<N>(nrows), ZTeam<N>::TPB_, 0, kf, fld<R>(b.resv), fld<R>(b.resw),
                   fld<R>(b.ya), fld<R>(b.yb), fld<R>(b.yc), ztw<R>(b), b.ny, b.nz1, b.nz2, nrows));
}

// Lp lines of N points from P (in place), line l at P + l ld; the lines of
// the mean-flow mode alone (Lp = 2 (ny-1)) are timed under their own names
template <typename R> static void dct_launch(Box &b, R *P, int iopt, int Lp, int ld, const R *chirp, const R *bf,
                                             const R *tw4, const R *tw) {
  using C = typename CT<R>::c;
  const int N = b.nx - 1, pairs = (Lp + 1) / 2;
  const bool all = Lp == (b.ny - 1) * b.nz1;
  const char *name = iopt < 0 ? (all ? "dct2_rows" : "dct2_mean") : (all ? "dct3_rows" : "dct3_mean");
  switch (b.dct_M) {
#define ZDCT(MM)                                                                                              \
  case MM: {                                                                                                   \
    auto kf = iopt < 0 ? &dct_rows_z<MM, R, -1> : &dct_rows_z<MM, R, 1>;                                       \
    LAUNCH(b, name, zgrid<MM>(pairs), ZTeam<MM>::TPB_, 0, kf, P, P, Lp, N, ld, (const C *)chirp,              \
           (const C *)bf, (const C *)tw4, (const C *)tw);                                                      \
  } break;
    ZDCT(512)
    ZDCT(1024)
    ZDCT(2048)
    ZDCT(3072)
    ZDCT(4096)
    ZDCT(8192)
#undef ZDCT
  default: throw std::invalid_argument("Bluestein length not instantiated");
  }
}

template <> void kz_dct_pencils<double>(Box &b, double *P, int iopt, int Lp, int ld) {
  dct_launch<double>(b, P, iopt, Lp, ld, b.dct_chirp, b.dct_bf, b.dct_tw4, b.dct_ztw);
}
template <> void kz_dct_pencils<float>(Box &b, float *P, int iopt, int Lp, int ld) {
  dct_launch<float>(b, P, iopt, Lp, ld, b.dct_chirp_f, b.dct_bf_f, b.dct_tw4_f, b.dct_ztw_f);
}

template void kz_products1<double>(Box &);
template void kz_products1<float>(Box &);
template void kz_products2<double>(Box &);
template void kz_products2<float>(Box &);

void kz_fourxz(Box &b, double *fou, double *phys, int iopt, int nrows) {
  zcheck(b);
  ZDISPATCH(b.nz, auto kf = &fourxz_z<N>;
            LAUNCH(b, "fourxz", zgrid<N>(nrows), ZTeam<N>::TPB_, 0, kf, fou, phys, iopt, (const double2 *)b.ztw,
                   b.zrow, b.nz1, b.nz2, nrows));
}

} // namespace otbl
