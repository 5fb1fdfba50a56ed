// Register-resident FFTs of the z direction for one row per team of W
// wavefronts (G = 64 W lanes, E = N / G elements per lane).
//
// A transform of length N runs as a sequence of Stockham passes with the
// radices of ZPlan<N>.  In pass p (radix R, Ns = product of the earlier
// radices) lane t owns the butterflies j = t + G m (m < N / (R G)); its
// registers hold the inputs x[j + r N / R] and, after the pass, the outputs
// y[(j / Ns) Ns R + (j mod Ns) + q Ns].  Between passes the outputs go
// through one LDS buffer per team, indexed a + a / 16 (16-byte elements), so
// the strided writes of the early passes and the unit-stride reads hit
// distinct banks.
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
  static constexpr int W = 4, NP = 5;
  static constexpr int rad(int p) { return p == 0 ? 12 : 4; }
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

// 16-point DFT as 4 x 4: inner DFTs over stride-4 inputs, twiddles W16^{bq},
// outer DFTs; X[q + 4p]
template <int S, typename C> __device__ __forceinline__ void bfly16(C *v) {
  using T = decltype(v[0].x);
  C y[4][4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    C t[4] = {v[b], v[b + 4], v[b + 8], v[b + 12]};
    bfly4<S>(t);
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
// v[b + 3a], twiddles W12^{bq}, outer 3-point DFTs over b; X[q + 4p]
template <int S, typename C> __device__ __forceinline__ void bfly12(C *v) {
  using T = decltype(v[0].x);
  C y[3][4];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    C t[4] = {v[b], v[b + 3], v[b + 6], v[b + 9]};
    bfly4<S>(t);
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

template <int R, int S, typename C> __device__ __forceinline__ void zbfly(C *v) {
  if constexpr (R == 16) bfly16<S>(v);
  else if constexpr (R == 12) bfly12<S>(v);
  else bfly<R, S>(v);
}

// Twiddles and butterflies of one pass on the lane's registers.
template <int N, int G, int R, int Ns, int S, typename C>
__device__ __forceinline__ void zpass_compute(C (&v)[N / G], int t, const C *__restrict__ tw) {
  constexpr int NB = N / (R * G);
  static_assert(NB * R * G == N, "radix does not tile the lanes");
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    C *b = v + m * R;
    if constexpr (Ns > 1) {
      // W^r: the powers 1, 2, 4, 8 from the table, every other power as the
      // product of the table powers of its set bits (at most three
      // multiplications deep), formed where it is applied
      const int k = (t + G * m) % Ns;
      constexpr int st = N / (Ns * R);
      C base[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if ((1 << q) < R) {
          base[q] = tw[k * (1 << q) * st];
          if (S < 0) base[q].y = -base[q].y;
        }
      }
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C w = mkc<C>(1.0, 0.0);
        bool first = true;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (r & (1 << q)) {
            w = first ? base[q] : cmulc(w, base[q]);
            first = false;
          }
        b[r] = cmulc(b[r], w);
      }
    }
    zbfly<R, S>(b);
  }
}

// outputs of a pass (radix R, Ns) to LDS
template <int N, int G, int R, int Ns, typename C>
__device__ __forceinline__ void zpass_store(const C (&v)[N / G], int t, C *lds) {
  constexpr int NB = N / (R * G);
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int j = t + G * m;
    const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
    for (int q = 0; q < R; ++q) lds[zpad(base + q * Ns)] = v[m * R + q];
  }
}

// inputs of a pass of radix R from LDS: x[j + r N / R]
template <int N, int G, int R, typename C>
__device__ __forceinline__ void zpass_load(C (&v)[N / G], int t, const C *lds) {
  constexpr int NB = N / (R * G);
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int j = t + G * m;
#pragma unroll
    for (int r = 0; r < R; ++r) v[m * R + r] = lds[zpad(j + r * (N / R))];
  }
}

// Passes from..NP-1 of the forward (Dir = +1: radices in plan order) or the
// inverse (Dir = -1: reversed order) transform; v holds the inputs of pass
// `from`.  On return v holds the outputs of the last pass.
template <int N, int S, bool Inv, int P = 0, int Ns = 1, typename C>
__device__ __forceinline__ void zrun(C (&v)[N / (64 * ZPlan<N>::W)], int t, const C *__restrict__ tw, C *lds) {
  constexpr int W = ZPlan<N>::W, G = 64 * W, NP = ZPlan<N>::NP;
  if constexpr (P < NP) {
    constexpr int R = ZPlan<N>::rad(Inv ? NP - 1 - P : P);
    zpass_compute<N, G, R, Ns, S>(v, t, tw);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (P + 1 < NP) {
      constexpr int R2 = ZPlan<N>::rad(Inv ? NP - 2 - P : P + 1);
      zpass_store<N, G, R, Ns>(v, t, lds);
      zsync<W>();
      zpass_load<N, G, R2>(v, t, lds);
      zsync<W>();
      zrun<N, S, Inv, P + 1, Ns * R>(v, t, tw, lds);
    }
  }
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
// into the spectra of a and b, modes 0..nz2, scaled by 1/N (ffts.f90:51).
// With XA, A also receives the mean-flow terms of a product whose mean was
// left out of the transform: A_k += ck XA_k for k >= 1 and A_0 += c0 XA_0,
// XA read before A is written (XA may be A itself).
template <int N, typename R>
__device__ __forceinline__ void zstore_spectral(const typename CT<R>::c (&v)[N / (64 * ZPlan<N>::W)], int t,
                                                typename CT<R>::c *lds, R *A, R *B, int nz2,
                                                const R *XA = nullptr, R c0 = R(0), R ck = R(0)) {
  using C = typename CT<R>::c;
  constexpr int W = ZPlan<N>::W, G = 64 * W, RD = ZPlan<N>::rad(0), NB = N / (RD * G);
#pragma unroll
  for (int m = 0; m < NB; ++m)
#pragma unroll
    for (int q = 0; q < RD; ++q) lds[zpad(t + G * m + q * (N / RD))] = v[m * RD + q];
  zsync<W>();
  const R sc = R(1.0 / N);
  for (int k = t; k <= nz2; k += G) {
    const C z = lds[zpad(k)];
    const C w = lds[zpad(k == 0 ? 0 : N - k)];
    if (A) {
      R ar = R(0.5) * (z.x + w.x) * sc, ai = (k == 0) ? R(0) : R(0.5) * (z.y - w.y) * sc;
      if (XA) {
        const R xr = XA[2 * k], xi = XA[2 * k + 1];
        ar += (k == 0 ? c0 : ck) * xr;
        ai += (k == 0) ? R(0) : ck * xi;
      }
      A[2 * k] = ar;
      A[2 * k + 1] = ai;
    }
    if (B) {
      B[2 * k] = R(0.5) * (z.y + w.y) * sc;
      B[2 * k + 1] = (k == 0) ? R(0) : R(-0.5) * (z.x - w.x) * sc;
    }
  }
  zsync<W>();
}

// zstore_spectral with the mean-flow terms on both outputs: A_k += cka XA_k
// (k >= 1), A_0 += c0a XA_0, and the same for B with XB, c0b, ckb (XA and XB
// read before A and B are written; either may be the output itself; null:
// no such term).  A or B null: that output is not written.
template <int N, typename R>
__device__ __forceinline__ void zstore_spectral2(const typename CT<R>::c (&v)[N / (64 * ZPlan<N>::W)], int t,
                                                 typename CT<R>::c *lds, R *A, R *B, int nz2, const R *XA, R c0a,
                                                 R cka, const R *XB, R c0b, R ckb) {
  using C = typename CT<R>::c;
  constexpr int W = ZPlan<N>::W, G = 64 * W, RD = ZPlan<N>::rad(0), NB = N / (RD * G);
#pragma unroll
  for (int m = 0; m < NB; ++m)
#pragma unroll
    for (int q = 0; q < RD; ++q) lds[zpad(t + G * m + q * (N / RD))] = v[m * RD + q];
  zsync<W>();
  const R sc = R(1.0 / N);
  for (int k = t; k <= nz2; k += G) {
    const C z = lds[zpad(k)];
    const C w = lds[zpad(k == 0 ? 0 : N - k)];
    if (A) {
      R ar = R(0.5) * (z.x + w.x) * sc, ai = (k == 0) ? R(0) : R(0.5) * (z.y - w.y) * sc;
      if (XA) {
        ar += (k == 0 ? c0a : cka) * XA[2 * k];
        ai += (k == 0) ? R(0) : cka * XA[2 * k + 1];
      }
      A[2 * k] = ar;
      A[2 * k + 1] = ai;
    }
    if (B) {
      R br = R(0.5) * (z.y + w.y) * sc, bi = (k == 0) ? R(0) : R(-0.5) * (z.x - w.x) * sc;
      if (XB) {
        br += (k == 0 ? c0b : ckb) * XB[2 * k];
        bi += (k == 0) ? R(0) : ckb * XB[2 * k + 1];
      }
      B[2 * k] = br;
      B[2 * k + 1] = bi;
    }
  }
  zsync<W>();
}

} // namespace otbl
