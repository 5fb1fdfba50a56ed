// Complex FFTs in LDS for one workgroup: self-sorting Stockham stages with
// radix-8/4/2/3/5/7 butterflies in registers, several signals of the same
// length advanced together so that every thread has work.  Twiddles come
// from a two-level table in LDS: w(t) = lo[t % 64] * hi[t / 64] with
// lo[a] = exp(2 pi i a / n), hi[b] = exp(2 pi i 64 b / n).
#pragma once
#include <hip/hip_runtime.h>

namespace otbl {

template <typename T> struct CT;
template <> struct CT<double> { using c = double2; };
template <> struct CT<float> { using c = float2; };

template <typename C> __device__ __forceinline__ C mkc(double x, double y) {
  C r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mkc<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mkc<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
// a * (S i): S = +1 -> i a, S = -1 -> -i a
template <int S, typename C> __device__ __forceinline__ C mul_si(C a) {
  C r;
  r.x = -S * a.y;
  r.y = S * a.x;
  return r;
}

struct FFTPlan {
  int n, nfac;
  int fac[14];
};

// ---- butterflies: X_q = sum_r a_r exp(S 2 pi i r q / R) ----
template <int S, typename C> __device__ __forceinline__ void bfly2(C *v) {
  C a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int S, typename C> __device__ __forceinline__ void bfly4(C *v) {
  C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  C t2 = cadd(v[1], v[3]), t3 = mul_si<S>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <int S, typename C> __device__ __forceinline__ void bfly8(C *v) {
  using T = decltype(v[0].x);
  const T h = (T)0.70710678118654752440;
  C e[4] = {v[0], v[2], v[4], v[6]};
  C o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<S>(e);
  bfly4<S>(o);
  // W8^1 = (1 + S i)/sqrt2, W8^2 = S i, W8^3 = (-1 + S i)/sqrt2
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

template <int S, typename C> __device__ __forceinline__ void bfly3(C *v) {
  using T = decltype(v[0].x);
  const T c = (T)-0.5, s = (T)0.86602540378443864676;
  C t = cadd(v[1], v[2]);
  C d = csub(v[1], v[2]);
  C m = mkc<C>(v[0].x + c * t.x, v[0].y + c * t.y);
  C r = mul_si<S>(d);
  r.x *= s;
  r.y *= s;
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <int R, int S, typename C> __device__ __forceinline__ void bfly_odd(C *v) {
  // direct DFT with constant roots, R = 5 or 7
  using T = decltype(v[0].x);
  constexpr double c5[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                            0.30901699437494742410};
  constexpr double s5[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                            -0.95105651629515357212};
  constexpr double c7[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429, -0.90096886790241912624,
                            -0.90096886790241912624, -0.22252093395631440429, 0.62348980185873353053};
  constexpr double s7[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702, 0.43388373911755812048,
                            -0.43388373911755812048, -0.97492791218182360702, -0.78183148246802980871};
  C o[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    T ax = v[0].x, ay = v[0].y;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const int t = (r * q) % R;
      const T cr = (T)(R == 5 ? c5[t] : c7[t]);
      const T sr = (T)(S * (R == 5 ? s5[t] : s7[t]));
      ax += v[r].x * cr - v[r].y * sr;
      ay += v[r].x * sr + v[r].y * cr;
    }
    o[q] = mkc<C>(ax, ay);
  }
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = o[q];
}

template <int R, int S, typename C> __device__ __forceinline__ void bfly(C *v) {
  if constexpr (R == 2) bfly2<S>(v);
  else if constexpr (R == 4) bfly4<S>(v);
  else if constexpr (R == 8) bfly8<S>(v);
  else if constexpr (R == 3) bfly3<S>(v);
  else bfly_odd<R, S>(v);
}

// twiddle exp(S 2 pi i t / n) from the two-level table
template <int S, typename C> __device__ __forceinline__ C twid(const C *lo, const C *hi, int t) {
  C w = cmulc(lo[t & 63], hi[t >> 6]);
  if (S < 0) w.y = -w.y;
  return w;
}

// Fill the two-level table for length n (n <= 64 * 256).  Called by the
// whole block; followed by __syncthreads() by the caller.
template <typename C> __device__ void fill_twiddles(C *lo, C *hi, int n) {
  for (int a = threadIdx.x; a < 64; a += blockDim.x) {
    double s, c;
    sincospi(2.0 * a / n, &s, &c);
    lo[a] = mkc<C>(c, s);
  }
  for (int b = threadIdx.x; b <= (n - 1) / 64; b += blockDim.x) {
    double s, c;
    sincospi(2.0 * (64.0 * b) / n, &s, &c);
    hi[b] = mkc<C>(c, s);
  }
}

template <int R, int S, typename C>
__device__ __forceinline__ void fft_stage_one(const C *__restrict__ a, C *__restrict__ b, int n, int Ns, int j,
                                              const C *lo, const C *hi) {
  const int nb = n / R;
  const int k = j % Ns;
  C v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = a[j + r * nb];
  if (Ns > 1) {
    const int tstep = n / (Ns * R);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmulc(v[r], twid<S>(lo, hi, k * r * tstep));
  }
  bfly<R, S>(v);
  const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int q = 0; q < R; ++q) b[idxD + q * Ns] = v[q];
}

// Transform nsig signals of length p.n together.  Signal s starts at
// A + s*sstride and ping-pongs with B + s*sstride; returns true if the
// results are back in A.
template <int S, typename C>
__device__ bool fft_multi(C *A, C *B, int sstride, int nsig, const FFTPlan &p, const C *lo, const C *hi) {
  int Ns = 1;
  bool inA = true;
  const int n = p.n;
  for (int s = 0; s < p.nfac; ++s) {
    const int R = p.fac[s];
    const int nb = n / R;
    const int total = nsig * nb;
    for (int g = threadIdx.x; g < total; g += blockDim.x) {
      const int sig = g / nb, j = g - sig * nb;
      const C *src = (inA ? A : B) + sig * sstride;
      C *dst = (inA ? B : A) + sig * sstride;
      switch (R) {
      case 8: fft_stage_one<8, S>(src, dst, n, Ns, j, lo, hi); break;
      case 4: fft_stage_one<4, S>(src, dst, n, Ns, j, lo, hi); break;
      case 2: fft_stage_one<2, S>(src, dst, n, Ns, j, lo, hi); break;
      case 3: fft_stage_one<3, S>(src, dst, n, Ns, j, lo, hi); break;
      case 5: fft_stage_one<5, S>(src, dst, n, Ns, j, lo, hi); break;
      default: fft_stage_one<7, S>(src, dst, n, Ns, j, lo, hi); break;
      }
    }
    __syncthreads();
    inA = !inA;
    Ns *= R;
  }
  return inA;
}

} // namespace otbl
