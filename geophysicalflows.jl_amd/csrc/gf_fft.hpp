// Workgroup-cooperative Stockham FFT for CDNA4 (gfx950).
//
// One transform of length N (a power of two) is held by TPT = N/P threads,
// P complex points per thread in registers.  Thread t holds x[t + TPT*q]
// (q = 0..P-1) before the first pass and y[t + TPT*q] after the last, so a
// transform reads and writes global memory with unit stride across lanes.
// Passes are radix-P Stockham autosort passes (plus one smaller remainder
// radix); between passes the points are exchanged through LDS.  Several
// independent transforms per thread (C) share the twiddles of a pass; their
// LDS exchanges go through one buffer, one transform at a time, so a
// workgroup needs only N complex values of LDS per thread group.
//
// The forward transform uses exp(-2 pi i jk/N) (FFTW/numpy sign, unnormalised).
// Inverse transforms are computed as conj(F(conj(x))) by the callers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gf {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};

template <typename T> __device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> z; z.re = a; z.im = b; return z; }
template <typename T> __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T> __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.re - b.re, a.im - b.im); }
// complex product in the (ac - bd, ad + bc) form
template <typename T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T> __device__ __forceinline__ cplx<T> rmul(T s, cplx<T> a) { return mk(s * a.re, s * a.im); }
template <typename T> __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return mk(a.re, -a.im); }
// multiply by -i
template <typename T> __device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) { return mk(a.im, -a.re); }

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

template <int N, int P>
struct Plan {
  static constexpr int LN = ilog2(N), LP = ilog2(P);
  static constexpr int NFULL = LN / LP;
  static constexpr int REM = 1 << (LN % LP);
  static constexpr int NPASS = NFULL + (REM > 1 ? 1 : 0);
  static constexpr int TPT = N / P;
  static constexpr int radix(int p) { return p < NFULL ? P : REM; }
  static_assert(N % P == 0, "P must divide N");
};

// ---- small DFT butterflies (forward sign), natural-order in/out ----------
template <typename T> __device__ __forceinline__ void dft2(cplx<T>& a, cplx<T>& b) {
  cplx<T> t = a; a = t + b; b = t - b;
}

template <typename T> __device__ __forceinline__ void dft4(cplx<T>* v) {
  cplx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
  cplx<T> t2 = v[1] + v[3], t3 = mul_mi(v[1] - v[3]);
  v[0] = t0 + t2; v[2] = t0 - t2; v[1] = t1 + t3; v[3] = t1 - t3;
}

template <typename T> __device__ __forceinline__ void dft8(cplx<T>* v) {
  const T h = (T)0.70710678118654752440084436210484903928;
  cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e); dft4(o);
  // o[k] *= w8^k
  cplx<T> o1 = mk(h * (o[1].re + o[1].im), h * (o[1].im - o[1].re));
  cplx<T> o2 = mul_mi(o[2]);
  cplx<T> o3 = mk(h * (o[3].im - o[3].re), -h * (o[3].re + o[3].im));
  v[0] = e[0] + o[0]; v[4] = e[0] - o[0];
  v[1] = e[1] + o1;   v[5] = e[1] - o1;
  v[2] = e[2] + o2;   v[6] = e[2] - o2;
  v[3] = e[3] + o3;   v[7] = e[3] - o3;
}

template <typename T> __device__ __forceinline__ void dft16(cplx<T>* v) {
  const T c1 = (T)0.92387953251128675612818318939678828682;  // cos(pi/8)
  const T s1 = (T)0.38268343236508977172845998403039886676;  // sin(pi/8)
  const T h = (T)0.70710678118654752440084436210484903928;
  cplx<T> e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
  dft8(e); dft8(o);
  // w16^k = cos(k pi/8) - i sin(k pi/8)
  const T wr[8] = {(T)1, c1, h, s1, (T)0, -s1, -h, -c1};
  const T wi[8] = {(T)0, -s1, -h, -c1, (T)-1, -c1, -h, -s1};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    cplx<T> ok = k == 0 ? o[0] : (k == 4 ? mul_mi(o[4]) : cmul(o[k], mk(wr[k], wi[k])));
    v[k] = e[k] + ok;
    v[k + 8] = e[k] - ok;
  }
}

template <int R, typename T> __device__ __forceinline__ void dft(cplx<T>* v) {
  if constexpr (R == 2) dft2(v[0], v[1]);
  else if constexpr (R == 4) dft4(v);
  else if constexpr (R == 8) dft8(v);
  else if constexpr (R == 16) dft16(v);
}

// XOR swizzle of LDS slots: spreads the strided Stockham writes over the banks
// without padding (the set of slots stays 0..N-1).
template <typename T> __device__ __forceinline__ int swz(int i) {
  if constexpr (sizeof(T) == 8) return i ^ ((i >> 3) & 7);   // 16-byte elements
  else return i ^ ((i >> 4) & 15);                            // 8-byte elements
}

// One Stockham pass (index PASS, stride NS) on C transforms, then recurse.
template <typename T, int N, int P, int C, int PASS, int NS>
__device__ __forceinline__ void stockham(cplx<T> (&v)[C][P], cplx<T>* __restrict__ lds, int t,
                                         const cplx<T>* __restrict__ tw) {
  using PL = Plan<N, P>;
  constexpr int R = PL::radix(PASS);
  constexpr int M = P / R;
  constexpr int TPT = PL::TPT;
#pragma unroll
  for (int b = 0; b < M; ++b) {
    const int j = t + b * TPT;
    const int k = j % NS;
    cplx<T> w[R];
    if constexpr (NS > 1) {
      constexpr int STEP = N / (NS * R);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        // 32-bit byte offset from a uniform base: SGPR-base + VGPR-offset loads
        const uint32_t off = (uint32_t)(r * k * STEP) * (uint32_t)sizeof(cplx<T>);
        w[r] = *reinterpret_cast<const cplx<T>*>(reinterpret_cast<const char*>(tw) + off);
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      cplx<T> u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        u[r] = v[c][b + r * M];
        if constexpr (NS > 1) { if (r > 0) u[r] = cmul(u[r], w[r]); }
      }
      dft<R>(u);
#pragma unroll
      for (int r = 0; r < R; ++r) v[c][b + r * M] = u[r];
    }
  }
  if constexpr (PASS + 1 < PL::NPASS) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
#pragma unroll
      for (int b = 0; b < M; ++b) {
        const int j = t + b * TPT;
        const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
        for (int r = 0; r < R; ++r) lds[swz<T>(base + r * NS)] = v[c][b + r * M];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < P; ++q) v[c][q] = lds[swz<T>(t + TPT * q)];
      __syncthreads();
    }
    stockham<T, N, P, C, PASS + 1, NS * R>(v, lds, t, tw);
  }
}

// Forward FFT of C transforms held as v[c][q] = x_c[t + TPT*q].
// lds: N complex slots private to this thread group; tw[i] = exp(-2 pi i i/N).
template <typename T, int N, int P, int C>
__device__ __forceinline__ void fft_forward(cplx<T> (&v)[C][P], cplx<T>* __restrict__ lds, int t,
                                            const cplx<T>* __restrict__ tw) {
  stockham<T, N, P, C, 0, 1>(v, lds, t, tw);
}

}  // namespace gf
