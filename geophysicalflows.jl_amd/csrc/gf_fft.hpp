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
// workgroup needs one padded N-point buffer per thread group.
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
template <typename T> __device__ __forceinline__ bool finite(cplx<T> a) { return __builtin_isfinite(a.re) && __builtin_isfinite(a.im); }
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

// Twiddle tables.  Pass p of the Stockham plan (stride NS_p, radix R_p) needs
// w^r with w = exp(-2 pi i k/(NS_p R_p)), k = j mod NS_p, r = 1..R_p-1.  The
// workgroup keeps, for every pass with NS_p > 1, the values w(k) contiguous in
// LDS so that the lanes of a wave read consecutive slots; the powers w^r are
// formed by squaring and multiplication (at most three roundings deep for
// R <= 16).  A pass with NS_p > TW_SPLIT stores w(k) as two tables,
// w(l) for l < TW_SPLIT and w(TW_SPLIT h) for h < NS_p/TW_SPLIT, and forms
// w(k) = w(TW_SPLIT h) w(l) with one more rounding: at N = 2048 the tables take
// 144 slots instead of 584, which lets two 512-thread workgroups share a CU.
constexpr int TW_SPLIT = 64;

template <int N, int P>
struct TwPlan {
  using PL = Plan<N, P>;
  static constexpr int ns(int p) { return p == 0 ? 1 : ns(p - 1) * PL::radix(p - 1); }
  static constexpr bool split(int p) { return ns(p) > TW_SPLIT; }
  static constexpr int slots(int p) { return split(p) ? TW_SPLIT + ns(p) / TW_SPLIT : ns(p); }
  // offset of pass p's table (passes 1..NPASS-1)
  static constexpr int off(int p) { return p <= 1 ? 0 : off(p - 1) + slots(p - 1); }
  static constexpr int size() { return PL::NPASS <= 1 ? 1 : off(PL::NPASS - 1) + slots(PL::NPASS - 1); }
};

// Fill the per-pass tables from the full table twg[i] = exp(-2 pi i i/N) in global memory.
template <typename T, int N, int P, int PASS = 1>
__device__ __forceinline__ void load_twiddles(cplx<T>* __restrict__ dst, const cplx<T>* __restrict__ twg) {
  using TP = TwPlan<N, P>;
  using PL = Plan<N, P>;
  if constexpr (PASS < PL::NPASS) {
    constexpr int ns = TP::ns(PASS), step = N / (ns * PL::radix(PASS)), o = TP::off(PASS);
    if constexpr (TP::split(PASS)) {
      for (int k = threadIdx.x; k < TW_SPLIT; k += blockDim.x) dst[o + k] = twg[k * step];
      for (int h = threadIdx.x; h < ns / TW_SPLIT; h += blockDim.x) dst[o + TW_SPLIT + h] = twg[h * TW_SPLIT * step];
    } else {
      for (int k = threadIdx.x; k < ns; k += blockDim.x) dst[o + k] = twg[k * step];
    }
    load_twiddles<T, N, P, PASS + 1>(dst, twg);
  }
}

// w^1..w^(R-1) from w^1
template <typename T, int R>
__device__ __forceinline__ void powers(cplx<T> (&w)[R]) {
  if constexpr (R >= 3) w[2] = cmul(w[1], w[1]);
  if constexpr (R >= 4) w[3] = cmul(w[2], w[1]);
  if constexpr (R >= 5) w[4] = cmul(w[2], w[2]);
  if constexpr (R >= 6) w[5] = cmul(w[4], w[1]);
  if constexpr (R >= 7) w[6] = cmul(w[3], w[3]);
  if constexpr (R >= 8) w[7] = cmul(w[4], w[3]);
  if constexpr (R >= 9) w[8] = cmul(w[4], w[4]);
  if constexpr (R >= 10) w[9] = cmul(w[8], w[1]);
  if constexpr (R >= 11) w[10] = cmul(w[5], w[5]);
  if constexpr (R >= 12) w[11] = cmul(w[8], w[3]);
  if constexpr (R >= 13) w[12] = cmul(w[6], w[6]);
  if constexpr (R >= 14) w[13] = cmul(w[8], w[5]);
  if constexpr (R >= 15) w[14] = cmul(w[7], w[7]);
  if constexpr (R >= 16) w[15] = cmul(w[8], w[7]);
}

// LDS slot of point i: one pad slot after every 8 points.  Stockham writes at
// stride R or NS (multiples of 8 or the first pass's 8j + r) and the unit-stride
// reads then fall on distinct banks.  pad(a + b) = pad(a) + pad(b) whenever
// a mod 8 + b mod 8 < 8, which holds for a thread's base point and the power-of-two
// multiples added to it (base + r NS with base mod NS < NS, t + q TPT with t < TPT), so
// every slot of a thread's accesses is its base slot plus a compile-time offset: one
// address register per pass and no per-access shift and mask.
__host__ __device__ __forceinline__ constexpr int pad(int i) { return i + (i >> 3); }
template <int N> struct LdsN { static constexpr int value = N + N / 8; };
// One pad slot after every G points (G a power of two, G >= 8: within LdsN), additive as pad is
// for sums whose parts mod G add up to less than G.
template <int G> __host__ __device__ __forceinline__ constexpr int padg(int i) { return i + i / G; }
// Padding of the Stockham exchange after a pass of radix R at stride NS, for points of S bytes,
// from LDS timings on gfx950 (64 banks of 4 bytes). The exchange writes point base + r NS and
// reads consecutive points on consecutive lanes.
// - FP64 (16-byte slots, ds_read_b128 / ds_write_b128): the reads run at full rate with one pad
//   per 32 points and at half rate with one per 8; writes at stride 16 run as fast with one per
//   32 as with one per 8, writes at stride 8 or 4 (a first pass of radix 8 or 4) only with one
//   per 8.
// - FP32 (8-byte slots, 64-bit accesses): one pad per 16 points keeps the reads and the writes
//   at stride 16 or 8 at full rate (one per 8 halves the reads); only the radix-4 passes at
//   stride 1 or 4 keep one per 8.
template <int S, int NS, int R>
constexpr int xpad_g() {
  if constexpr (S == 16) return (NS == 1 && R <= 8) ? 8 : 32;
  else return (R <= 4 && NS <= 4) ? 8 : 16;
}

// Slots of kr = t + i TPT and of its mirror (N - kr) mod N in an N-point buffer, as used by
// the split of a packed spectrum W = F(a + i b) of two real sequences (W[k] with W[N - k]).
// For t > 0 the mirror is tm + (N - (i + 1) TPT) with tm = TPT - t; for t = 0 it is
// (N - i TPT) mod N.  Both are a per-thread base plus a compile-time offset when TPT is a
// multiple of the padding period G (padg<G> is then additive over multiples of TPT); smaller TPT
// computes each slot.
template <int N, int TPT, int G = 8>
struct MirrorSlots {
  int lo, m0, mb;
  __host__ __device__ explicit MirrorSlots(int t) {
    lo = padg<G>(t);
    if constexpr (TPT % G == 0) {
      m0 = t != 0 ? padg<G>(TPT - t) + padg<G>(N - TPT) : 0;
      mb = t != 0 ? padg<G>(TPT - t) - padg<G>(TPT) : 0;
    } else {
      m0 = t;   // the thread index, for the computed form
      mb = 0;
    }
  }
  __host__ __device__ int k(int i) const { return lo + padg<G>(i * TPT); }
  __host__ __device__ int m(int i) const {
    if constexpr (TPT % G == 0) return i == 0 ? m0 : mb + padg<G>(N - i * TPT);
    else return padg<G>((N - m0 - i * TPT) & (N - 1));
  }
};

// The y-direction kernels interleave the two columns of a pair across adjacent lanes, each
// column with its own N-point buffer.  Buffers N + N/8 slots apart start on the same LDS bank,
// so the two lanes of a pair would collide on every exchange; the second column's buffer is
// shifted by 64 bytes (half of the 128 bytes the banks cover) to put them on disjoint banks.
template <typename T> struct ColSkew { static constexpr int value = 64 / (int)sizeof(cplx<T>); };
// Slots one column pair occupies (two buffers and the skew; pairs never overlap), and the
// buffer of column c of pair slot in a workgroup's LDS of G PairLds<T, N> slots.
template <typename T, int N> struct PairLds { static constexpr int value = 2 * LdsN<N>::value + ColSkew<T>::value; };
template <typename T, int N>
__device__ __forceinline__ cplx<T>* col_buffer(cplx<T>* lds, int slot, int c) {
  return lds + slot * PairLds<T, N>::value + c * (LdsN<N>::value + ColSkew<T>::value);
}

// Workgroup barrier for an exchange through LDS alone: this wave's LDS accesses complete, then the
// barrier. Unlike __syncthreads() it carries no memory fence, so it does not wait for the wave's
// outstanding global accesses, among them loads into LDS (global_load_lds) into another buffer
// that are meant to stay in flight across the exchange. The "memory" clobber keeps the compiler
// from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
template <bool LIGHT> __device__ __forceinline__ void xbarrier() {
  if constexpr (LIGHT) lds_barrier();
  else __syncthreads();
}

// One Stockham pass (index PASS, stride NS) on C transforms, then recurse.  LIGHT: exchange
// barriers without a memory fence (lds_barrier).
template <typename T, int N, int P, int C, int PASS, int NS, bool LIGHT = false>
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
      constexpr int TOFF = TwPlan<N, P>::off(PASS);
      if constexpr (TwPlan<N, P>::split(PASS))
        w[1] = cmul(tw[TOFF + TW_SPLIT + k / TW_SPLIT], tw[TOFF + k % TW_SPLIT]);
      else
        w[1] = tw[TOFF + k];
      powers<T, R>(w);
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
    constexpr int G = xpad_g<(int)sizeof(cplx<T>), NS, R>();
#pragma unroll
    for (int c = 0; c < C; ++c) {
#pragma unroll
      for (int b = 0; b < M; ++b) {
        const int j = t + b * TPT;
        const int base = (j / NS) * NS * R + (j % NS);
        cplx<T>* wp = lds + padg<G>(base);   // slot base + r NS is padg(base) + padg(r NS)
#pragma unroll
        for (int r = 0; r < R; ++r) wp[padg<G>(r * NS)] = v[c][b + r * M];
      }
      xbarrier<LIGHT>();
      const cplx<T>* rp = lds + padg<G>(t);
#pragma unroll
      for (int q = 0; q < P; ++q) v[c][q] = rp[padg<G>(TPT * q)];
      xbarrier<LIGHT>();
    }
    stockham<T, N, P, C, PASS + 1, NS * R, LIGHT>(v, lds, t, tw);
  }
}

// Forward FFT of C transforms held as v[c][q] = x_c[t + TPT*q].
// lds: LdsN<N> complex slots private to this thread group; tw: the per-pass
// twiddle tables in LDS (TwPlan<N, P>::size() slots, filled by load_twiddles and
// followed by a barrier).
//
//
// FRESH makes t opaque to the optimiser at the call (an empty asm).  A kernel that runs several
// transforms otherwise has the LDS and twiddle addresses of every pass computed once and kept
// live across all of them: at N = 2048, P = 8 that holds 25 more VGPRs in the advection x-pass
// (147 instead of 122), one wave per SIMD fewer.  Recomputing them costs a few integer
// operations per pass, 1-5% of a kernel's time where the registers it frees do not let more
// workgroups run at once, so kernels ask for it only at sizes where they do.
template <typename T, int N, int P, int C, bool FRESH = false, bool LIGHT = false>
__device__ __forceinline__ void fft_forward(cplx<T> (&v)[C][P], cplx<T>* __restrict__ lds, int t,
                                            const cplx<T>* __restrict__ tw) {
  if constexpr (FRESH) {
    asm volatile("" : "+v"(t));
    __builtin_assume(t >= 0 && t < N / P);   // the asm hides t's range from the multiplies below
  }
  stockham<T, N, P, C, 0, 1, LIGHT>(v, lds, t, tw);
}

// Direct DFT of a length n < N that is not a power of two, in the data layout of fft_forward:
// thread t holds x[t + TPT q] (zero where t + TPT q >= n) and receives y[k] = sum_j x[j] w^(jk),
// w = exp(-2 pi i/n), at k = t + TPT q < n (the other slots come back zero).  roots[j] = w^j,
// j < n, in LDS.  Each thread forms its P outputs from all n inputs (P n complex products), the
// index j k mod n advanced by k per term; the inputs are read from lds (broadcast across lanes).
template <typename T, int N, int P>
__device__ __forceinline__ void dft_direct(cplx<T> (&v)[1][P], cplx<T>* __restrict__ lds, int t, int n,
                                           const cplx<T>* __restrict__ roots) {
  constexpr int TPT = N / P;
  __syncthreads();   // lds is free: every earlier reader of it is done
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const int j = t + TPT * q;
    if (j < n) lds[j] = v[0][q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const int k = t + TPT * q;
    cplx<T> acc = mk((T)0, (T)0);
    if (k < n) {
      int idx = 0;
      for (int j = 0; j < n; ++j) {
        acc = acc + cmul(lds[j], roots[idx]);
        idx += k;
        if (idx >= n) idx -= n;
      }
    }
    v[0][q] = acc;
  }
  __syncthreads();   // lds is free again for the caller
}

}  // namespace gf
