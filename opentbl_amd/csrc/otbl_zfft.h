// Register-resident FFTs of the z direction for one row per team of W
// wavefronts (G = 64 W lanes, E = N / G elements per lane).
//
// A transform of length N runs as a sequence of Stockham passes with the
// radices of ZPlan<N>.  In pass p (radix R, Ns = product of the earlier
// radices) lane t owns the butterflies j = t + G m (m < N / (R G)); its
// registers hold the inputs x[j + r N / R] and, after the pass, the outputs
// y[(j / Ns) Ns R + (j mod Ns) + q Ns].  Between passes the outputs go
// through one LDS buffer per team, indexed a + a / 16, which spreads the
// strided writes of the early passes over the banks (the bank-conflict
// counters read about 0.2 conflict cycles per LDS instruction for the FP32
// products and 2-3 for FP64 elements).
//
// The last pass of a transform leaves output index j + q N / R_last in the
// registers of the lane owning butterfly j; an inverse transform whose first
// radix is R_last expects exactly that index set, so point-wise products in
// physical space need no data movement between a forward and an inverse
// transform.  Twiddles W_M^k = exp(2 pi i k / M): the powers 1, 2, 4, 8 from
// a table of exp(2 pi i m / N), the others as products of those; the sign S = +1 transform is the complex-to-real
// direction of fourxz (ffts.f90:27-67), S = -1 the real-to-complex one.
#pragma once
#include "otbl_fft.h"

namespace otbl {

// radix plans (the inverse runs the radices in reverse order)
template <int N> struct ZPlan;
template <> struct ZPlan<256> {
  static constexpr int W = 1, NP = 4;
  static constexpr int rad(int p) { return 4; }
};
template <> struct ZPlan<512> {
  static constexpr int W = 1, NP = 3;
  static constexpr int rad(int p) { return 8; }
};
template <> struct ZPlan<1024> {
  static constexpr int W = 1, NP = 3;
  static constexpr int rad(int p) { return p < 2 ? 16 : 4; }
};
template <> struct ZPlan<2048> {
  static constexpr int W = 2, NP = 3;
  static constexpr int rad(int p) { return p < 2 ? 16 : 8; }
};
// 3 x 1024: the Bluestein length of the x transforms for N up to 1536
template <> struct ZPlan<3072> {
  static constexpr int W = 2, NP = 4;
  static constexpr int rad(int p) { return p == 0 ? 12 : (p < 3 ? 8 : 4); }
};
template <> struct ZPlan<4096> {
  static constexpr int W = 4, NP = 3;
  static constexpr int rad(int p) { return 16; }
};

template <> struct ZPlan<8192> {
  static constexpr int W = 8, NP = 4;
  static constexpr int rad(int p) { return p < 3 ? 16 : 2; }
};

// lengths of the z products and of the Poisson x transforms (Bluestein
// convolution length) that the engine is instantiated for
inline bool zfft_supported(int n) { return n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096; }
inline bool zdct_supported(int m) {
  return m == 512 || m == 1024 || m == 2048 || m == 3072 || m == 4096 || m == 8192;
}
// the shortest engine length that holds a Bluestein convolution of length
// n (0: none; convolutions of at most 256 points stay on the LDS kernels)
inline int zdct_length(int n) {
  if (n <= 256) return 0;
  for (int m : {512, 1024, 2048, 3072, 4096, 8192})
    if (m >= n) return m;
  return 0;
}
inline int zfft_waves(int n) { return n <= 1024 ? 1 : n / 1024; }

// Teams of a launch: one row per team of W wavefronts (G lanes, E elements
// per lane); single-wavefront teams run four to a workgroup.
template <int N> struct ZTeam {
  static constexpr int W = ZPlan<N>::W, G = 64 * W, E = N / G;
  static constexpr int TEAMS = W == 1 ? 4 : 1; // teams per workgroup
  static constexpr int TPB_ = TEAMS * G;
};

// padded LDS index of element a
__device__ __forceinline__ int zpad(int a) { return a + (a >> 4); }
template <int N> constexpr int zlds_elems() { return N + N / 16; }

// synchronisation of a team: one wavefront needs only ordering of its own
// LDS accesses, several need the workgroup barrier (one team per workgroup)
template <int W> __device__ __forceinline__ void zsync() {
  if constexpr (W == 1) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// 4-point DFT of (t0, t1, 0, 0)
template <int S, typename C> __device__ __forceinline__ void bfly4h(C *t) {
  const C a = t[0], b = t[1], ib = mul_si<S>(b);
  t[0] = cadd(a, b);
  t[2] = csub(a, b);
  t[1] = cadd(a, ib);
  t[3] = csub(a, ib);
}

// 16-point DFT as 4 x 4: inner DFTs over stride-4 inputs, twiddles W16^{bq},
// outer DFTs; X[q + 4p].  H: inputs 8..15 are zero (not read).
template <int S, bool H = false, typename C> __device__ __forceinline__ void bfly16(C *v) {
  using T = decltype(v[0].x);
  C y[4][4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    C t[4] = {v[b], v[b + 4], H ? v[b] : v[b + 8], H ? v[b] : v[b + 12]};
    if constexpr (H) bfly4h<S>(t);
    else bfly4<S>(t);
#pragma unroll
    for (int q = 0; q < 4; ++q) y[b][q] = t[q];
  }
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
  // W16^k = cos(2 pi k / 16) + S i sin(2 pi k / 16), k = b q
  auto tw = [&](C a, int k) -> C {
    T cr, si;
    switch (k) {
    case 1: cr = c1; si = s1; break;
    case 2: cr = h; si = h; break;
    case 3: cr = s1; si = c1; break;
    case 4: cr = 0; si = 1; break;
    case 6: cr = -h; si = h; break;
    default: cr = -c1; si = -s1; break; // k = 9
    }
    si = S > 0 ? si : -si;
    if (k == 4) return mul_si<S>(a);
    return mkc<C>(a.x * cr - a.y * si, a.x * si + a.y * cr);
  };
#pragma unroll
  for (int b = 1; b < 4; ++b)
#pragma unroll
    for (int q = 1; q < 4; ++q) y[b][q] = tw(y[b][q], b * q);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    C t[4] = {y[0][q], y[1][q], y[2][q], y[3][q]};
    bfly4<S>(t);
#pragma unroll
    for (int p = 0; p < 4; ++p) v[q + 4 * p] = t[p];
  }
}

// 12-point DFT as 3 x 4: inner 4-point DFTs over stride-3 inputs
// v[b + 3a], twiddles W12^{bq}, outer 3-point DFTs over b; X[q + 4p].
// H: inputs 6..11 are zero (not read).
template <int S, bool H = false, typename C> __device__ __forceinline__ void bfly12(C *v) {
  using T = decltype(v[0].x);
  C y[3][4];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    C t[4] = {v[b], v[b + 3], H ? v[b] : v[b + 6], H ? v[b] : v[b + 9]};
    if constexpr (H) bfly4h<S>(t);
    else bfly4<S>(t);
#pragma unroll
    for (int q = 0; q < 4; ++q) y[b][q] = t[q];
  }
  // W12^k = cos(pi k / 6) + S i sin(pi k / 6), k = b q in {1, 2, 3, 4, 6}
  const T hr = (T)0.86602540378443864676;
  auto tw = [&](C a, int k) -> C {
    if (k == 3) return mul_si<S>(a);
    if (k == 6) return mkc<C>(-a.x, -a.y);
    const T cr = k == 1 ? hr : (k == 2 ? T(0.5) : T(-0.5));
    const T si = S * (k == 1 ? T(0.5) : hr);
    return mkc<C>(a.x * cr - a.y * si, a.x * si + a.y * cr);
  };
#pragma unroll
  for (int b = 1; b < 3; ++b)
#pragma unroll
    for (int q = 1; q < 4; ++q) y[b][q] = tw(y[b][q], b * q);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    C t[3] = {y[0][q], y[1][q], y[2][q]};
    bfly3<S>(t);
#pragma unroll
    for (int p = 0; p < 3; ++p) v[q + 4 * p] = t[p];
  }
}

// 8-point DFT of (v0..v3, 0, 0, 0, 0): bfly8's even/odd split with half
// inputs
template <int S, typename C> __device__ __forceinline__ void bfly8h(C *v) {
  using T = decltype(v[0].x);
  const T h = (T)0.70710678118654752440;
  C e[4] = {v[0], v[2], v[0], v[0]}, o[4] = {v[1], v[3], v[1], v[1]};
  bfly4h<S>(e);
  bfly4h<S>(o);
  C o1 = cadd(o[1], mul_si<S>(o[1]));
  o1.x *= h;
  o1.y *= h;
  C o2 = mul_si<S>(o[2]);
  C o3 = csub(mul_si<S>(o[3]), o[3]);
  o3.x *= h;
  o3.y *= h;
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// radix-R butterfly; H: the upper half of the inputs is zero
template <int R, int S, bool H = false, typename C> __device__ __forceinline__ void zbfly(C *v) {
  if constexpr (R == 16) bfly16<S, H>(v);
  else if constexpr (R == 12) bfly12<S, H>(v);
  else if constexpr (H && R == 8) bfly8h<S>(v);
  else if constexpr (H && R == 4) bfly4h<S>(v);
  else {
    static_assert(!H, "no half-input butterfly of this radix");
    bfly<R, S>(v);
  }
}

// Twiddle table entries of one pass: for butterfly m of the lane, W^(k 2^q)
// (k = (t + G m) mod Ns, q with 2^q < R) in units of N / (Ns R) of the
// table.  Loaded before the LDS exchange that precedes the pass, so their
// latency overlaps the exchange.
template <int N, int G, int R, int Ns, typename C>
__device__ __forceinline__ void zpass_twload(C (&base)[N / (R * G)][4], int t, const C *__restrict__ tw) {
  if constexpr (Ns > 1) {
    constexpr int NB = N / (R * G), st = N / (Ns * R);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int k = (t + G * m) % Ns;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if ((1 << q) < R) base[m][q] = tw[k * (1 << q) * st];
    }
  }
}

// Twiddles and butterflies of one pass on the lane's registers (H: the
// upper half of every butterfly's inputs is zero, first pass only).
template <int N, int G, int R, int Ns, int S, bool H, typename C>
__device__ __forceinline__ void zpass_compute(C (&v)[N / G], const C (&base_in)[N / (R * G)][4]) {
  constexpr int NB = N / (R * G);
  static_assert(NB * R * G == N, "radix does not tile the lanes");
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    C *b = v + m * R;
    if constexpr (Ns > 1) {
      // W^r: the powers 1, 2, 4, 8 from the table, every other power as the
      // product of a smaller power already formed and the table power of
      // its top bit (one multiplication each, at most three deep); r and
      // r + HB are taken together so that few powers stay live
      constexpr int HB = R > 8 ? 8 : (R > 4 ? 4 : (R > 2 ? 2 : 1));
      constexpr int QB = HB == 8 ? 3 : (HB == 4 ? 2 : (HB == 2 ? 1 : 0));
      C base[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if ((1 << q) < R) {
          base[q] = base_in[m][q];
          if (S < 0) base[q].y = -base[q].y;
        }
      }
      C w[HB];
#pragma unroll
      for (int r = 1; r < HB; ++r) {
        const int hb = r >= 4 ? 4 : (r >= 2 ? 2 : 1), qb = r >= 4 ? 2 : (r >= 2 ? 1 : 0);
        w[r] = r == hb ? base[qb] : cmulc(w[r - hb], base[qb]);
        b[r] = cmulc(b[r], w[r]);
        if (r + HB < R) b[r + HB] = cmulc(b[r + HB], cmulc(w[r], base[QB]));
      }
      if (HB < R) b[HB] = cmulc(b[HB], base[QB]);
    }
    zbfly<R, S, H>(b);
  }
}

// padded index of a + d for a compile-time offset d that keeps the 16-element
// group structure (d a multiple of 16, or a below a group boundary that a + d
// does not cross): zpad(a) + d + d / 16
__device__ __forceinline__ int zpad_off(int pa, int d) { return pa + d + (d >> 4); }

// The store index base + q Ns of pass (R, Ns), base = (j / Ns) Ns R + j mod Ns,
// is zpad(base) + q Ns + (q Ns) / 16 for every lane when Ns is a multiple of
// 16, or when Ns divides 16 and Ns R is a multiple of 16 (the group
// boundary carry is then (q Ns) / 16), or when Ns R divides 16 (no lane's
// q range crosses a group boundary); one address per butterfly then serves
// all R stores through the instruction offset.
template <int R, int Ns> constexpr bool zstore_linear() {
  return Ns % 16 == 0 || (16 % Ns == 0 && (Ns * R) % 16 == 0) || 16 % (Ns * R) == 0;
}

// outputs of a pass (radix R, Ns) to LDS
template <int N, int G, int R, int Ns, typename C>
__device__ __forceinline__ void zpass_store(const C (&v)[N / G], int t, C *lds) {
  constexpr int NB = N / (R * G);
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int j = t + G * m;
    const int base = (j / Ns) * Ns * R + (j % Ns);
    if constexpr (zstore_linear<R, Ns>()) {
      const int pb = zpad(base);
#pragma unroll
      for (int q = 0; q < R; ++q) lds[zpad_off(pb, q * Ns)] = v[m * R + q];
    } else {
#pragma unroll
      for (int q = 0; q < R; ++q) lds[zpad(base + q * Ns)] = v[m * R + q];
    }
  }
}

// inputs of a pass of radix R from LDS: x[j + r N / R] (N / R and G are
// multiples of 16, so every index is zpad(t) plus a compile-time offset)
template <int N, int G, int R, typename C>
__device__ __forceinline__ void zpass_load(C (&v)[N / G], int t, const C *lds) {
  constexpr int NB = N / (R * G);
  static_assert((N / R) % 16 == 0 && G % 16 == 0, "load offsets break the padding groups");
  const int pt = zpad(t);
#pragma unroll
  for (int m = 0; m < NB; ++m)
#pragma unroll
    for (int r = 0; r < R; ++r) v[m * R + r] = lds[zpad_off(pt, G * m + r * (N / R))];
}

// Passes P..NP-1 of the forward (Inv = false: radices in plan order) or the
// inverse (Inv = true: reversed order) transform; v holds the inputs of pass
// P and base its twiddle entries.  On return v holds the outputs of the last
// pass.
template <int N, int S, bool Inv, int P, int Ns, bool H, typename C, int NBP>
__device__ __forceinline__ void zrun_from(C (&v)[N / (64 * ZPlan<N>::W)], int t, const C *__restrict__ tw, C *lds,
                                          const C (&base)[NBP][4]) {
  constexpr int W = ZPlan<N>::W, G = 64 * W, NP = ZPlan<N>::NP;
  constexpr int R = ZPlan<N>::rad(Inv ? NP - 1 - P : P);
  static_assert(NBP == N / (R * G), "twiddle entries of another pass");
  static_assert(!H || (P == 0 && Ns == 1), "half inputs in the first pass only");
  zpass_compute<N, G, R, Ns, S, H>(v, base);
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (P + 1 < NP) {
    constexpr int R2 = ZPlan<N>::rad(Inv ? NP - 2 - P : P + 1);
    C base2[N / (R2 * G)][4];
    zpass_twload<N, G, R2, Ns * R>(base2, t, tw);
    zpass_store<N, G, R, Ns>(v, t, lds);
    zsync<W>();
    zpass_load<N, G, R2>(v, t, lds);
    zsync<W>();
    zrun_from<N, S, Inv, P + 1, Ns * R, false>(v, t, tw, lds, base2);
  }
}

// The whole forward (Inv = false) or inverse (Inv = true) transform of the
// team's row in v, sign S.  H: the inputs of the first pass's upper half
// (register slots m R + r, r >= R / 2) are zero and not read.
template <int N, int S, bool Inv, bool H = false, typename C>
__device__ __forceinline__ void zrun(C (&v)[N / (64 * ZPlan<N>::W)], int t, const C *__restrict__ tw, C *lds) {
  constexpr int W = ZPlan<N>::W, G = 64 * W, NP = ZPlan<N>::NP;
  constexpr int R = ZPlan<N>::rad(Inv ? NP - 1 : 0);
  C base[N / (R * G)][4]; // the first pass has no twiddles (Ns = 1)
  zrun_from<N, S, Inv, 0, 1, H>(v, t, tw, lds, base);
}

// The physical index held in register slot e after the last pass of a
// forward transform (and expected there by the first pass of an inverse).
template <int N> __device__ __forceinline__ int zphys_index(int t, int e) {
  constexpr int W = ZPlan<N>::W, G = 64 * W;
  constexpr int R = ZPlan<N>::rad(ZPlan<N>::NP - 1);
  return t + G * (e / R) + (e % R) * (N / R);
}

// Complex input of the c2r transform of two real signals X, Y (spectral rows
// of nz1 = 2 (nz2 + 1) doubles, modes 0..nz2; a signal that is absent is a
// row of zeros, never a null pointer, so the loads carry no branches): the
// Hermitian-extended spectrum of x + i y at index k (pack of two real
// inverse transforms into one complex one).
template <typename R>
__device__ __forceinline__ typename CT<R>::c zpack_one(const R *X, const R *Y, int k, int n, int nz2) {
  using C = typename CT<R>::c;
  const bool lo = k <= nz2, hi = k >= n - nz2;
  const int kk = lo ? k : (hi ? n - k : 0);
  const C x = ((const C *)X)[kk];
  const C y = ((const C *)Y)[kk];
  // k <= nz2: X[k] + i Y[k]; mirrored: conj(X[n-k]) + i conj(Y[n-k]);
  // Im of mode 0 does not enter; zero between nz2 and n - nz2
  const R sx = lo ? R(-1) : R(1);
  C r = mkc<C>(x.x + sx * y.y, -sx * x.y + y.x);
  r = k == 0 ? mkc<C>(x.x, y.x) : r;
  return (lo || hi) ? r : mkc<C>(0.0, 0.0);
}

// The first forward pass's inputs x[j + r N / R0] straight from the spectral
// rows in global memory.  The dealiased mode count is fixed by the length
// (nz1 = 2 (N / 3), nz2 = N / 3 - 1), so register slots whose indices all
// fall in the zero band nz2 < k < N - nz2 are known at compile time and
// load nothing, and slots entirely below or above the band need no select.
template <int N, typename R>
__device__ __forceinline__ void zload_spectral(typename CT<R>::c (&v)[N / (64 * ZPlan<N>::W)], int t, const R *X,
                                               const R *Y) {
  using C = typename CT<R>::c;
  constexpr int G = 64 * ZPlan<N>::W, RD = ZPlan<N>::rad(0), NB = N / (RD * G);
  constexpr int nz2 = N / 3 - 1;
  const C *X2 = (const C *)X, *Y2 = (const C *)Y;
#pragma unroll
  for (int m = 0; m < NB; ++m)
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int base = G * m + r * (N / RD); // k = base + t, t < G
      const int k = base + t;
      C z = mkc<C>(0.0, 0.0);
      if (base + G - 1 <= nz2) { // all lanes below the band: X[k] + i Y[k]
        const C x = X2[k], y = Y2[k];
        z = k == 0 ? mkc<C>(x.x, y.x) : mkc<C>(x.x - y.y, x.y + y.x);
      } else if (base >= N - nz2) { // all above: conj(X[N-k]) + i conj(Y[N-k])
        const C x = X2[N - k], y = Y2[N - k];
        z = mkc<C>(x.x + y.y, -x.y + y.x);
      } else if (base <= nz2 || base + G - 1 >= N - nz2) { // straddles an edge
        z = zpack_one(X, Y, k, N, nz2);
      }
      v[m * RD + r] = z;
    }
  __builtin_amdgcn_sched_barrier(0);
}

// After the last inverse pass (natural order, R = first forward radix,
// outputs y[j + q N / R]): through LDS, split the transform Z of a + i b
// into the spectra of a and b, modes 0..nz2, scaled by 1/N (ffts.f90:51),
// and add the mean-flow terms of products whose mean was left out of the
// transform: A_k += cka XA_k for k >= 1, A_0 += c0a XA_0, the same for B
// with XB, c0b, ckb.  XA and XB are read before A and B are written (either
// may be the output itself); a null XA or XB adds no term, a null A or B is
// not written.  nz2 = N / 3 - 1, the engine's dealiasing: the mode loop is
// unrolled to those N / 3 modes, and the mean-flow rows are loaded in one
// batch ahead of the arithmetic.
template <int N, typename R>
__device__ __forceinline__ void zstore_spectral2(const typename CT<R>::c (&v)[N / (64 * ZPlan<N>::W)], int t,
                                                 typename CT<R>::c *lds, R *A, R *B, int nz2, const R *XA, R c0a,
                                                 R cka, const R *XB, R c0b, R ckb) {
  using C = typename CT<R>::c;
  constexpr int W = ZPlan<N>::W, G = 64 * W, RD = ZPlan<N>::rad(0), NB = N / (RD * G);
  constexpr int NI = (N / 3 + G - 1) / G; // mode loop trips (nz2 + 1 = N / 3 modes)
  const int pt = zpad(t);
  // slots whose indices all fall in the zero band nz2 < k < N - nz2 are never
  // read back, and the last pass does not form them
  constexpr int NZ2 = N / 3 - 1;
#pragma unroll
  for (int m = 0; m < NB; ++m)
#pragma unroll
    for (int q = 0; q < RD; ++q) {
      const int b = G * m + q * (N / RD);
      if (!(b > NZ2 && b + G - 1 < N - NZ2)) lds[zpad_off(pt, b)] = v[m * RD + q];
    }
  C xa[NI], xb[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int k = min(t + G * it, nz2);
    xa[it] = XA ? ((const C *)XA)[k] : mkc<C>(0.0, 0.0);
    xb[it] = XB ? ((const C *)XB)[k] : mkc<C>(0.0, 0.0);
  }
  zsync<W>();
  const R sc = R(1.0 / N);
  // N - k for k = t + G it is zpad(N - t) - G it - G it / 16 (t > 0)
  const int pw = zpad(N - t);
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int k = t + G * it;
    if (k <= nz2) {
      const C z = lds[zpad_off(pt, G * it)];
      const C w = k == 0 ? lds[0] : lds[pw - G * it - ((G * it) >> 4)];
      if (A) {
        R ar = R(0.5) * (z.x + w.x) * sc, ai = (k == 0) ? R(0) : R(0.5) * (z.y - w.y) * sc;
        ar += (k == 0 ? c0a : cka) * xa[it].x;
        ai += (k == 0) ? R(0) : cka * xa[it].y;
        ((C *)A)[k] = mkc<C>(ar, ai);
      }
      if (B) {
        R br = R(0.5) * (z.y + w.y) * sc, bi = (k == 0) ? R(0) : R(-0.5) * (z.x - w.x) * sc;
        br += (k == 0 ? c0b : ckb) * xb[it].x;
        bi += (k == 0) ? R(0) : ckb * xb[it].y;
        ((C *)B)[k] = mkc<C>(br, bi);
      }
    }
  }
  zsync<W>();
}

// zstore_spectral2 with at most one mean-flow term (on A)
template <int N, typename R>
__device__ __forceinline__ void zstore_spectral(const typename CT<R>::c (&v)[N / (64 * ZPlan<N>::W)], int t,
                                                typename CT<R>::c *lds, R *A, R *B, int nz2,
                                                const R *XA = nullptr, R c0 = R(0), R ck = R(0)) {
  zstore_spectral2<N, R>(v, t, lds, A, B, nz2, XA, c0, ck, nullptr, R(0), R(0));
}

} // namespace otbl
