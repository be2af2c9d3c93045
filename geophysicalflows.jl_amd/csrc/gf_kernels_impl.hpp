// Fused pseudospectral kernels for the nonlinear term and the stepper stages.
//
// One evaluation of calcN! (src/twodnavierstokes.jl:286-294,
// src/singlelayerqg.jl:443-451, src/multilayerqg.jl:651-663) is three launches:
//
//   K1  spectral -> mixed : one thread group per kr column: masked read of the
//       state (dealias!, P1), inversion to velocity (and S^-1 q for layers),
//       inverse FFT along y.  Only the kr < kmax columns that survive the
//       dealias mask are transformed.
//   K2  mixed -> flux     : per row pair (y, y+1): Hermitian extension, c2r
//       along x (two real rows per complex FFT), physical-space products,
//       r2c along x (two real rows per complex FFT), split of the packed
//       spectra.  Nothing physical touches HBM.
//   K3  flux -> spectral  : one thread group per kr column: forward FFT along
//       y, assembly of N (-i kr A - i l B, forcing, topography and drag terms)
//       and the fused stepper stage (ETDRK4 / RK4 / ForwardEuler) as epilogue.
//
// This header holds the templates; gf_inst_*.hip instantiate them per size.
#pragma once
#include "gf_kernels.hpp"
#include <type_traits>

namespace gf {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ bool kept_l(const Geo<T>& g, int l) {
  return l < g.la0 || l >= g.la1;
}
template <typename T> __device__ __forceinline__ T lwave(const Geo<T>& g, int l) {
  return (T)(l < (g.ny + 1) / 2 ? l : l - g.ny) * g.ky0;
}
template <typename T> __device__ __forceinline__ T krwave(const Geo<T>& g, int kr) {
  return (T)kr * g.kx0;
}
__device__ __forceinline__ long boff(int kr, int j, int ny) {
  return ((long)(kr / BW) * ny + j) * BW + (kr % BW);
}
// 32-bit element offset in the column-blocked layout (fields stay below 4 GiB)
__device__ __forceinline__ uint32_t boff32(int kr, int j, int ny) {
  return ((uint32_t)(kr / BW) * (uint32_t)ny + (uint32_t)j) * (uint32_t)BW + (uint32_t)(kr % BW);
}
// Column pair of pair-workgroup b (one pair per workgroup).  With quads the two pairs of a block
// go to workgroups 16 (q / 8) + (q mod 8) and that + 8, which the dispatcher hands to one XCD
// (workgroup b runs on XCD b mod 8): the pair that reads second finds the block's lines in that
// XCD's L2.  With pairs, workgroup b takes pair b.
__device__ __forceinline__ int pair_of_block(int b) { return BW == 4 ? 16 * (b / 16) + 2 * (b % 8) + (b / 8) % 2 : b; }
// Element i of an array whose base pointer is uniform across the workgroup, addressed
// through a 32-bit byte offset (scalar base + vector offset memory instructions).
template <typename X> __device__ __forceinline__ const X& at(const X* base, uint32_t i) {
  return *reinterpret_cast<const X*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(X));
}
template <typename X> __device__ __forceinline__ X& at(X* base, uint32_t i) {
  return *reinterpret_cast<X*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(X));
}
template <typename T> __device__ __forceinline__ cplx<T> czero() { return mk((T)0, (T)0); }

// Philox-4x32-10 (Salmon et al. 2011); key (k0,k1), counter (c0..c3).
__device__ __forceinline__ void philox(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                       uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
}

// Complex normal, variance 1/2 per part, keyed on (seed, step, index, instance).
template <typename T>
__device__ __noinline__ cplx<T> complex_normal(uint64_t seed, int64_t step, uint32_t index, uint32_t inst) {
  uint32_t c0 = index, c1 = inst, c2 = (uint32_t)(uint64_t)step, c3 = (uint32_t)((uint64_t)step >> 32);
  philox(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = ((uint64_t)c0 << 32) | c1, b = ((uint64_t)c2 << 32) | c3;
  const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;
  const double u2 = ((double)(b >> 11) + 1.0) * 0x1.0p-53;
  const double r = sqrt(-log(u1));
  const double th = 2.0 * 3.141592653589793 * u2;
  double sn, cs;
  sincos(th, &sn, &cs);
  return mk((T)(r * cs), (T)(r * sn));
}

template <typename T> struct Coef {
  const void* p; int cplx_;
  // The same two loads for either kind of table: a complex table's element i, or a real table's
  // elements i and i + 1 (every table carries 16 bytes of padding), the second then unused. With
  // a branch on the kind between a load and its use, the compiler waited for each coefficient
  // before issuing the next load.
  __device__ __forceinline__ cplx<T> mul(uint32_t i, cplx<T> x) const {
    const T* e = reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) +
                                            i * (uint32_t)(cplx_ ? sizeof(cplx<T>) : sizeof(T)));
    const T re = e[0], im = e[1];
    return cplx_ ? cmul(mk(re, im), x) : rmul(re, x);
  }
};

// ---------------------------------------------------------------------------
// K1: spectral state -> mixed fields (inverse y-transform), one column / group
// ---------------------------------------------------------------------------
#ifndef GF_K1_WAVES
#define GF_K1_WAVES 3
#endif
#ifndef GF_K2_WAVES
#define GF_K2_WAVES 2
#endif
// Sizes from which each kernel's transforms recompute their addresses (fft_forward's FRESH),
// chosen per precision by whole-step timings at 2048² (tuned schedules, ring-forced ETDRK4):
// - FP64 K2 (P = 8) drops from 147 to 122 VGPRs and fits 4 workgroups per CU, all 1024 row
//   pairs at once; FP64 K3 at P = 8 spills less.  0.436 -> 0.406 ms/step.
// - FP32 K1 at P = 8 goes from 82 to 60 VGPRs, 5 to 8 waves per SIMD; FP32 K2 and K3 only pay
//   the extra integer work.
// Below 2048 every kernel already has the workgroups it needs in flight and keeps its addresses.
constexpr int kFreshNever = 1 << 30;
template <typename T> struct FreshFrom;
template <> struct FreshFrom<double> { static constexpr int k1 = kFreshNever, k2 = 2048, k3 = 2048; };
template <> struct FreshFrom<float> { static constexpr int k1 = 2048, k2 = kFreshNever, k3 = kFreshNever; };
#ifndef GF_K3_WAVES
#define GF_K3_WAVES 4   // measured faster than 3 despite a small spill (two 512-thread workgroups per CU)
#endif

template <int MODEL, int FS> struct Fields {
  static constexpr int N = FS == FS_CALCN ? 3 : (FS == FS_RAW ? 1 : (MODEL == M_TWODNS ? 3 : 4));
  static constexpr int code(int f) {
    return FS == FS_CALCN ? (f == 0 ? F_U : f == 1 ? F_V : F_Q)
         : FS == FS_RAW ? F_RAW
         : MODEL == M_TWODNS ? (f == 0 ? F_Q : f == 1 ? F_U : F_V)
                             : (f == 0 ? F_Q : f == 1 ? F_PSI : f == 2 ? F_U : F_V);
  }
};

// Thread layout of the y-direction kernels: a workgroup holds G column pairs;
// lane bit 0 selects the column of a pair and the remaining bits the FFT thread,
// so the two columns of a pair (interleaved in memory) are read and written by
// one wave as one contiguous run.
template <int TPT> struct YMap {
  int c, t, slot;
  __device__ __forceinline__ YMap(int x) : c(x & 1), t((x >> 1) % TPT), slot((x >> 1) / TPT) {}
};

// Non-power-of-two column lengths ny (even, up to GEN_MAX) run in separate instantiations
// (GEN = true) of the y-pass kernels of NY = pow2_ceil(ny): they index with the runtime ny, the
// rows ny..NY-1 of each thread's points stay unused, and a direct DFT (dft_direct) replaces the
// Stockham transform, with the ny roots of unity (the global table, which has ny entries) in the
// LDS twiddle array. The power-of-two kernels keep the constant NY (a runtime ny in them cost
// 6-7% per step at 64^2-256^2).
template <int NY, int P, bool GEN> struct YTw {
  static constexpr int size = (GEN && NY > TwPlan<NY, P>::size()) ? NY : TwPlan<NY, P>::size();
};
template <typename T, int NY, int P, bool GEN>
__device__ __forceinline__ void load_y_twiddles(cplx<T>* __restrict__ tw, const cplx<T>* __restrict__ twg, int ny) {
  if constexpr (GEN) {
    for (int i = threadIdx.x; i < ny; i += blockDim.x) tw[i] = twg[i];
  } else {
    load_twiddles<T, NY, P>(tw, twg);
  }
}
// the y-direction transform of a column of ny points held as fft_forward holds them
template <typename T, int NY, int P, bool FR, bool GEN>
__device__ __forceinline__ void y_transform(cplx<T> (&w)[1][P], cplx<T>* __restrict__ L, int t,
                                            const cplx<T>* __restrict__ tw, int ny) {
  if constexpr (GEN) dft_direct<T, NY, P>(w, L, t, ny, tw);
  else fft_forward<T, NY, P, 1, FR>(w, L, t, tw);
}
// the column length the y-pass kernels index with
template <int NY, bool GEN, typename T> __device__ __forceinline__ int ycol(const Geo<T>& g) {
  if constexpr (GEN) return g.ny; else return NY;
}
// row l lies in the column (always, where the column fills the kernel's NY points)
template <bool GEN> __device__ __forceinline__ bool in_col(int l, int ny) {
  if constexpr (GEN) return l < ny; else return true;
}

// Body of K1 for block blk of column pairs; lds holds 2G padded N-point buffers and
// tw the per-pass twiddle tables (already loaded).
template <typename T, int NY, int P, int G, int NL, int MODEL, int FS, bool GEN = false>
__device__ __forceinline__ void k1_body(const K1Args<T>& a, int blk, cplx<T>* __restrict__ lds,
                                        const cplx<T>* __restrict__ tw) {
  constexpr int TPT = NY / P;
  using FSET = Fields<MODEL, FS>;
  const YMap<TPT> m(threadIdx.x);
  const int t = m.t;
  const int kr = 2 * (G == 1 ? pair_of_block(blk) : blk * G + m.slot) + m.c;
  const Geo<T>& g = a.g;
  const bool active = kr < a.ncols && kr < g.nkr;
  // the linearised MultiLayerQG calcN! does not dealias its input (src/multilayerqg.jl:675-683)
  const bool masked = FS != FS_RAW && !((NL > 1 || MODEL == M_SLQG) && FS == FS_CALCN && a.nomask);
  const bool krk = kr < g.kmax || !masked;
  const T kv = krwave(g, kr);
  cplx<T>* L = col_buffer<T, NY>(lds, m.slot, m.c);
  const T nrm = g.inv_norm;
  const int ny = ycol<NY, GEN>(g);
  // masked state value of layer j at y-index l
  auto sval = [&](int j, int l) -> cplx<T> {
    const bool lk = kept_l(g, l) || !masked;
    cplx<T> v = czero<T>();
    if (active && lk && krk && in_col<GEN>(l, ny)) v = at(a.sol, (uint32_t)(j * g.lstride) + boff32(kr, l, ny));
    return v;
  };
  // multi-layer problems may give each layer its own workgroups (blockIdx.y): the inversion
  // reads every layer, the fields and transforms are the layer's own
  const int jl0 = (NL > 1 && a.lsplit) ? (int)blockIdx.y : 0;
  const int jl1 = (NL > 1 && a.lsplit) ? jl0 + 1 : NL;
  for (int jl = jl0; jl < jl1; ++jl) {
    // per-point inversion factor, computed once before the fields that use it:
    // TwoDNS 1/|k|^2 (invKrsq, [FF] TwoDGrid), QG psih (src/singlelayerqg.jl:484-490,
    // src/multilayerqg.jl:495-509)
    T iK[P];
    cplx<T> ps[P];
    // A point's state is read in a loop of its own, before any is used: read where it is used,
    // each read's branch on the dealias mask made the compiler wait for it before the next.
    cplx<T> svs[P];
    if constexpr (FS != FS_RAW && MODEL != M_TWODNS && NL == 1) {
#pragma unroll
      for (int q = 0; q < P; ++q) svs[q] = sval(jl, t + TPT * q);
    }
    if constexpr (FS != FS_RAW) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int l = t + TPT * q;
        const T lv = lwave(g, l);
        const T K2 = kv * kv + lv * lv;
        if constexpr (MODEL == M_TWODNS) {
          iK[q] = (kr == 0 && l == 0) ? (T)0 : (T)1 / K2;
        } else if constexpr (NL > 1) {
          cplx<T> acc = czero<T>();
          if (active && in_col<GEN>(l, ny)) {
            const T* mm = a.sinv + boff(kr, l, ny) * (NL * NL) + jl * NL;
#pragma unroll
            for (int k = 0; k < NL; ++k) acc = acc + rmul(mm[k], sval(k, l));
          }
          ps[q] = acc;
          // the bottom layer's psih, for K3's bottom drag of the same stage
          if (FS == FS_CALCN && jl == NL - 1 && a.psi_out != nullptr && active && in_col<GEN>(l, ny)) at(a.psi_out, boff32(kr, l, ny)) = acc;
        } else {
          const cplx<T> sv = svs[q];
          if (a.eb) ps[q] = mk(-sv.re / (K2 + a.ell2inv), -sv.im / (K2 + a.ell2inv));
          else ps[q] = rmul(-((kr == 0 && l == 0) ? (T)0 : (T)1 / K2), sv);
        }
      }
    }
#pragma unroll
    for (int f = 0; f < FSET::N; ++f) {
      // with fsplit each workgroup row transforms one field (a uniform branch: every wave of
      // the workgroup skips the same fields)
      if constexpr (NL == 1 && FS == FS_CALCN) {
        if (a.fsplit && f != (int)blockIdx.y) continue;
      }
      const int code = FSET::code(f);
      cplx<T> w[1][P];
      // TwoDNS reads the state again for each field (an L2 hit, mostly, and it keeps the registers
      // few), except at 16 points per thread: there LDS already limits the kernel to 2 waves per
      // SIMD, so the state stays in registers across the three fields for free. At 2048^2 FP64
      // that took K1 from 22.9 to 21.2 us and the step from 0.3434 to 0.3308 ms; at 8 points
      // per thread (4096^2) holding it cost a wave per SIMD (K1 89.8 -> 118 us).
      constexpr bool HOLD = MODEL == M_TWODNS && P >= 16;
      const bool fsp = NL == 1 && FS == FS_CALCN && a.fsplit;
      if (HOLD ? (f == 0 || fsp) : (code == F_RAW || code == F_Q || MODEL == M_TWODNS)) {
#pragma unroll
        for (int q = 0; q < P; ++q) svs[q] = sval(jl, t + TPT * q);
      }
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int l = t + TPT * q;
        const T lv = lwave(g, l);
        cplx<T> r;
        if (code == F_RAW || code == F_Q) {
          r = svs[q];
        } else if constexpr (MODEL == M_TWODNS) {
          // uh = i l invK sol, vh = -i kr invK sol (src/twodnavierstokes.jl:254-255)
          const cplx<T> sv = svs[q];
          if (code == F_U) { const T aa = lv * iK[q]; r = mk(-aa * sv.im, aa * sv.re); }
          else if (code == F_V) { const T bb = kv * iK[q]; r = mk(bb * sv.im, -bb * sv.re); }
          else { r = rmul(-iK[q], sv); }
        } else {
          // uh = -i l psih, vh = i kr psih (src/singlelayerqg.jl:360-361)
          const cplx<T> p = ps[q];
          if (code == F_U) r = mk(lv * p.im, -lv * p.re);
          else if (code == F_V) r = mk(-kv * p.im, kv * p.re);
          else r = p;
        }
        // normalised inverse = conj(F(conj(x)))/(nx ny); the scale is a power of two
        w[0][q] = mk(r.re * nrm, -(r.im * nrm));
      }
      y_transform<T, NY, P, (NY >= FreshFrom<T>::k1), GEN>(w, L, t, tw, ny);
      if (active) {
        cplx<T>* o = a.out + (long)(jl * FSET::N + f) * a.fstride;
#pragma unroll
        for (int q = 0; q < P; ++q)
          if (in_col<GEN>(t + TPT * q, ny)) at(o, boff32(kr, t + TPT * q, ny)) = conj(w[0][q]);
      }
    }
  }
}

template <typename T, int NY, int P, int G, int NL, int MODEL, int FS, bool GEN = false>
__global__ __launch_bounds__(2 * G * (NY / P)) __attribute__((amdgpu_waves_per_eu(GF_K1_WAVES)))
void k1_kernel(const K1Args<T> a) {
  __shared__ cplx<T> lds[G * PairLds<T, NY>::value];
  __shared__ cplx<T> tw[YTw<NY, P, GEN>::size];
  load_y_twiddles<T, NY, P, GEN>(tw, a.tw, a.g.ny);
  __syncthreads();
  k1_body<T, NY, P, G, NL, MODEL, FS, GEN>(a, blockIdx.x, lds, tw);
}

// ---------------------------------------------------------------------------
// K2: mixed -> flux (x-direction c2r, products, r2c) / forward / inverse-out
// ---------------------------------------------------------------------------
// Padding period of the x pass's own exchanges through LDS (mirrored points, spectrum splits),
// where consecutive lanes touch consecutive points: one pad per 32 FP64 or 16 FP32 points keeps
// gfx950's LDS reads at full rate (one per 8 halves them).
template <typename T> struct XPad { static constexpr int G = sizeof(T) == 8 ? 32 : 16; };
// Hermitian-extended entry k of a half spectrum (kr = k or NX-k).
template <typename T, int NX>
__device__ __forceinline__ cplx<T> herm_load(const cplx<T>* f, int k, int y, int ny, int kin, bool active) {
  // Branch-free: every lane loads and a select clears what it must not use. With an early return
  // here the compiler waited for each load before issuing the next one. A lane past the kept
  // columns re-reads its row's last kept column (a line other lanes fetch anyway; one shared
  // address for all of them would send every such load in the chip to one L2 channel), and a
  // lane outside the rows reads element 0, which every field has.
  const int kk = k <= NX / 2 ? k : NX - k;
  const bool ok = active && kk < kin;
  const uint32_t idx = (active && kin > 0) ? boff32(kk < kin ? kk : kin - 1, y, ny) : 0u;
  const cplx<T> v = at(f, idx);
  // c2r ignores the imaginary DC/Nyquist parts
  const T im = (kk == 0 || kk == NX / 2) ? (T)0 : (k > NX / 2 ? -v.im : v.im);
  return ok ? mk(v.re, im) : czero<T>();
}

// Split of a packed spectrum W = F(a + i b) of two real rows:
// row 0: (W[k] + conj W[N-k])/2, row 1: (W[k] - conj W[N-k])/(2i).
template <typename T>
__device__ __forceinline__ cplx<T> split_row(cplx<T> Wk, cplx<T> Wm, int row) {
  return row == 0 ? mk((Wk.re + Wm.re) / 2, (Wk.im - Wm.im) / 2)
                  : mk((Wk.im + Wm.im) / 2, (Wm.re - Wk.re) / 2);
}

// Workgroups are handed to the 8 XCDs in turn (workgroup b to XCD b mod 8), each XCD with its
// own L2.  Mapping workgroup b to block (b mod 8) nb/8 + b/8 gives every XCD a contiguous range
// of row pairs, so the row pairs that share the 128-byte lines of the blocked layout (four rows
// of a column pair) run together on one L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

// Body of K2 for block blk of row pairs; lds holds G padded N-point buffers.
template <typename T, int NX, int P, int G, int MODE, int NL>
__device__ __forceinline__ void k2_body(const K2Args<T>& a, int blk, cplx<T>* __restrict__ lds,
                                        const cplx<T>* __restrict__ tw) {
  constexpr int TPT = NX / P;
  constexpr int MAXI = (NX / 2 + 1 + TPT - 1) / TPT;
  const int grp = threadIdx.x / TPT, t = threadIdx.x % TPT;
  const Geo<T>& g = a.g;
  const int ny = g.ny;
  const int rp = blk * G + grp;
  const bool active = rp < ny / 2;
  const int y0 = 2 * rp;
  cplx<T>* L = lds + grp * LdsN<NX>::value;

  if constexpr (MODE == K2_FORWARD_IN) {
    // physical rows (y0, y0+1) of each field -> half spectra along x
    for (int f = 0; f < a.nfields; ++f) {
      cplx<T> w[1][P];
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int x = t + TPT * q;
        T r0 = 0, r1 = 0;
        if (active) {
          r0 = a.phys_in[f * a.pstride + (long)y0 * NX + x];
          r1 = a.phys_in[f * a.pstride + (long)(y0 + 1) * NX + x];
        }
        w[0][q] = mk(r0, r1);
      }
      fft_forward<T, NX, P, 1, (NX >= FreshFrom<T>::k2)>(w, L, t, tw);
#pragma unroll
      for (int q = 0; q < P; ++q) L[padg<XPad<T>::G>(t) + padg<XPad<T>::G>(TPT * q)] = w[0][q];
      __syncthreads();
      if (active) {
        cplx<T>* o = a.out + f * a.ostride;
        for (int it = t; it < 2 * a.kout; it += TPT) {
          const int row = it & 1, kr = it >> 1;
          o[boff(kr, y0 + row, ny)] = split_row(L[padg<XPad<T>::G>(kr)], L[padg<XPad<T>::G>((NX - kr) & (NX - 1))], row);
        }
      }
      __syncthreads();
    }
  } else if constexpr (MODE == K2_INVERSE_OUT) {
    // nfields mixed fields -> physical rows; fields paired (f, f+1) per row
    for (int f = 0; f < a.nfields; f += 2) {
      const bool two = f + 1 < a.nfields;
      const cplx<T>* fa = a.in + f * a.fstride;
      const cplx<T>* fb = a.in + (f + 1) * a.fstride;
      for (int row = 0; row < 2; ++row) {
        cplx<T> w[1][P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
          const int k = t + TPT * q;
          const cplx<T> A = herm_load<T, NX>(fa, k, y0 + row, ny, a.kin, active);
          const cplx<T> B = two ? herm_load<T, NX>(fb, k, y0 + row, ny, a.kin, active) : czero<T>();
          w[0][q] = mk(A.re - B.im, -(A.im + B.re));     // conj(A + i B)
        }
        fft_forward<T, NX, P, 1, (NX >= FreshFrom<T>::k2)>(w, L, t, tw);
        if (active) {
#pragma unroll
          for (int q = 0; q < P; ++q) {
            const long o = (long)(y0 + row) * NX + t + TPT * q;
            a.phys[f * a.pstride + o] = w[0][q].re;
            if (two) a.phys[(f + 1) * a.pstride + o] = -w[0][q].im;
          }
        }
      }
    }
  } else if constexpr (MODE == K2_ADVECT) {
    // ---- advection without background gradients: u, v, q -> u (q [+ eta]), v (q [+ eta]) ----
    // (src/twodnavierstokes.jl:262-270, src/singlelayerqg.jl:368-375).  Every transform packs
    // the two rows of one field: inverse (q0 + i q1), (u0 + i u1), (v0 + i v1) and forward
    // (u0 q0 + i u1 q1), (v0 q0 + i v1 q1).  A phase thus reads and writes whole 64-byte pieces
    // (two rows of a column pair; with the workgroup's other row pair, whole 128-byte lines), and
    // only q stays live across phases.
    static_assert(NL == 1, "K2_ADVECT serves single-layer models");
    const cplx<T>* fu = a.in;
    const cplx<T>* fv = a.in + a.fstride;
    const cplx<T>* fq = a.in + 2 * a.fstride;
    auto load_rows = [&](const cplx<T>* f, cplx<T> (&w)[1][P]) {
      // Only points k <= NX/2 are read from memory. The Hermitian half k > NX/2 is the packed
      // value of the mirror NX - k, conj(conj A + i conj B): the loading lane leaves it in LDS slot
      // NX - k for the lane that holds that point. Where each lane read its own mirror too, half
      // of the x pass's load instructions fetched data another lane had already loaded (at
      // 2048^2 these instructions each touch about 30 pages); the exchange costs two barriers.
      // 2048^2 step 0.3480 -> 0.3435 ms, 1024^2 0.1569 -> 0.1531.
      // All loads are issued before any is used, so their latencies overlap.
      constexpr int QH = P / 2 + 1;                     // the q with TPT q <= NX/2
      cplx<T> A[QH], B[QH];
#pragma unroll
      for (int q = 0; q < QH; ++q) {
        const int k = t + TPT * q;
        A[q] = herm_load<T, NX>(f, k, y0, ny, a.kin, active && k <= NX / 2);
        B[q] = herm_load<T, NX>(f, k, y0 + 1, ny, a.kin, active && k <= NX / 2);
      }
#pragma unroll
      for (int q = 0; q < QH; ++q) {
        const int k = t + TPT * q;
        if (k > NX / 2) continue;
        w[0][q] = mk(A[q].re - B[q].im, -(A[q].im + B[q].re));     // conj(A + i B)
        if (k > 0 && k < NX / 2) L[padg<XPad<T>::G>(NX - k)] = mk(A[q].re + B[q].im, A[q].im - B[q].re);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (TPT * q + TPT - 1 <= NX / 2) continue;      // every lane of this q holds k <= NX/2
        const int k = t + TPT * q;
        if (k <= NX / 2) continue;
        w[0][q] = L[padg<XPad<T>::G>(k)];
      }
      __syncthreads();
    };
    auto inverse_rows = [&](const cplx<T>* f, cplx<T> (&w)[1][P]) {
      load_rows(f, w);
      fft_forward<T, NX, P, 1, (NX >= FreshFrom<T>::k2)>(w, L, t, tw);             // row y0 = Re, row y0 + 1 = -Im
    };
    auto forward_rows = [&](cplx<T>* o, cplx<T> (&w)[1][P]) {
      fft_forward<T, NX, P, 1, (NX >= FreshFrom<T>::k2)>(w, L, t, tw);             // of (row y0 + i row y0 + 1)
      const MirrorSlots<NX, TPT, XPad<T>::G> ms(t);
#pragma unroll
      for (int q = 0; q < P; ++q) L[ms.k(q)] = w[0][q];
      __syncthreads();
      if (active) {
#pragma unroll
        for (int i = 0; i < MAXI; ++i) {
          const int kr = t + i * TPT;
          if (kr < a.kout) {
            const cplx<T> Wk = L[ms.k(i)], Wm = L[ms.m(i)];
            at(o, boff32(kr, y0, ny)) = split_row(Wk, Wm, 0);
            at(o, boff32(kr, y0 + 1, ny)) = split_row(Wk, Wm, 1);
          }
        }
      }
      __syncthreads();
    };
    T q0[P], q1[P];
    {
      cplx<T> w[1][P];
      inverse_rows(fq, w);
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int x = t + TPT * q;
        q0[q] = w[0][q].re;
        q1[q] = -w[0][q].im;
        if (a.eta != nullptr && active) {
          q0[q] = q0[q] + at(a.eta, (uint32_t)(y0 * NX + x));
          q1[q] = q1[q] + at(a.eta, (uint32_t)((y0 + 1) * NX + x));
        }
      }
    }
    // with hsplit, workgroup row 0 forms u q and row 1 v q (each transforms q itself)
    const int half = a.hsplit == 1 ? (int)blockIdx.y : 2;
    if (half != 1) {
      cplx<T> w[1][P];
      inverse_rows(fu, w);
#pragma unroll
      for (int q = 0; q < P; ++q) w[0][q] = mk(w[0][q].re * q0[q], -w[0][q].im * q1[q]);   // u q, both rows
      forward_rows(a.out, w);
    }
    if (half != 0) {
      cplx<T> w[1][P];
      inverse_rows(fv, w);
#pragma unroll
      for (int q = 0; q < P; ++q) w[0][q] = mk(w[0][q].re * q0[q], -w[0][q].im * q1[q]);   // v q, both rows
      forward_rows(a.out + a.ostride, w);
    }
  } else {
    // ---- advection with background gradients: per layer, u, v, q -> (u q, v q, u Qx + v Qy) ----
    constexpr bool FULL = MODE == K2_ADVECT_FULL;
    constexpr int NF = FULL ? 3 : 2;
    const int jl0 = (NL > 1 && a.lsplit) ? (int)blockIdx.y : 0;
    const int jl1 = (NL > 1 && a.lsplit) ? jl0 + 1 : NL;
    for (int jl = jl0; jl < jl1; ++jl) {
      const cplx<T>* fu = a.in + (long)(3 * jl + 0) * a.fstride;
      const cplx<T>* fv = a.in + (long)(3 * jl + 1) * a.fstride;
      const cplx<T>* fq = a.in + (long)(3 * jl + 2) * a.fstride;
      // inverse x-transforms, one packed pair at a time: (u0 + i v0), (u1 + i v1), (q0 + i q1)
      cplx<T> w[3][P];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        cplx<T> wr[1][P];
        const cplx<T>* fa = r == 2 ? fq : fu;
        const cplx<T>* fb = r == 2 ? fq : fv;
        const int ya = r == 1 ? y0 + 1 : y0;
        const int yb = r == 0 ? y0 : y0 + 1;
#pragma unroll
        for (int q = 0; q < P; ++q) {
          const int k = t + TPT * q;
          const cplx<T> A = herm_load<T, NX>(fa, k, ya, ny, a.kin, active);
          const cplx<T> B = herm_load<T, NX>(fb, k, yb, ny, a.kin, active);
          wr[0][q] = mk(A.re - B.im, -(A.im + B.re));     // conj(A + i B)
        }
        fft_forward<T, NX, P, 1, (NX >= FreshFrom<T>::k2)>(wr, L, t, tw);
#pragma unroll
        for (int q = 0; q < P; ++q) w[r][q] = wr[0][q];
      }
      // physical values: u_r = Re, v_r = -Im of w[r]; q_0 = Re w[2], q_1 = -Im w[2]
      cplx<T> z[NF][P];
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int x = t + TPT * q;
        T u0 = w[0][q].re, v0 = -w[0][q].im, u1 = w[1][q].re, v1 = -w[1][q].im;
        T q0 = w[2][q].re, q1 = -w[2][q].im;
        if constexpr (!FULL) {
          if (a.eta != nullptr && active) {
            // u (q + eta), v (q + eta) (src/singlelayerqg.jl:368-370)
            q0 = q0 + at(a.eta, (uint32_t)(y0 * NX + x));
            q1 = q1 + at(a.eta, (uint32_t)((y0 + 1) * NX + x));
          }
          z[0][q] = mk(u0 * q0, v0 * q0);
          z[1][q] = mk(u1 * q1, v1 * q1);
        } else {
          // u += U; N = -(u Qx)^ - (v Qy)^ - i kr (u q)^ - i l (v q)^ (src/multilayerqg.jl:695-731);
          // linearised: N = -(u Qx)^ - (v Qy)^ - i kr (U q)^ (src/multilayerqg.jl:744-779)
          T c0 = 0, c1 = 0, U0 = 0, U1 = 0;
          if (active && a.has_bg) {
            U0 = a.Uy[jl * ny + y0];
            U1 = a.Uy[jl * ny + y0 + 1];
            u0 = u0 + U0;
            u1 = u1 + U1;
            T qy0 = a.Qy1[jl * ny + y0], qy1 = a.Qy1[jl * ny + y0 + 1];
            T qx0 = 0, qx1 = 0;
            if (jl == NL - 1) {
              if (a.etax) { qx0 = a.etax[(long)y0 * NX + x]; qx1 = a.etax[(long)(y0 + 1) * NX + x]; }
              if (a.etay) { qy0 = qy0 + a.etay[(long)y0 * NX + x]; qy1 = qy1 + a.etay[(long)(y0 + 1) * NX + x]; }
            }
            c0 = u0 * qx0 + v0 * qy0;
            c1 = u1 * qx1 + v1 * qy1;
          }
          if (a.linear) {
            z[0][q] = mk(U0 * q0, (T)0);
            z[1][q] = mk(U1 * q1, (T)0);
          } else {
            z[0][q] = mk(u0 * q0, v0 * q0);
            z[1][q] = mk(u1 * q1, v1 * q1);
          }
          z[2][q] = mk(c0, c1);
        }
      }
#pragma unroll
      for (int r = 0; r < NF; ++r) {
        cplx<T> zr[1][P];
#pragma unroll
        for (int q = 0; q < P; ++q) zr[0][q] = z[r][q];
        fft_forward<T, NX, P, 1, (NX >= FreshFrom<T>::k2)>(zr, L, t, tw);
#pragma unroll
        for (int q = 0; q < P; ++q) z[r][q] = zr[0][q];
      }
      // split the packed spectra; rows handled one at a time through LDS
      cplx<T>* oA = a.out + (long)(2 * jl + 0) * a.ostride;
      cplx<T>* oB = a.out + (long)(2 * jl + 1) * a.ostride;
      const MirrorSlots<NX, TPT, XPad<T>::G> ms(t);
      for (int row = 0; row < 2; ++row) {
#pragma unroll
        for (int q = 0; q < P; ++q) L[ms.k(q)] = z[row][q];
        __syncthreads();
        cplx<T> Arow[MAXI], Brow[MAXI];
#pragma unroll
        for (int i = 0; i < MAXI; ++i) {
          const int kr = t + i * TPT;
          if (kr < a.kout) {
            const cplx<T> Wk = L[ms.k(i)], Wm = L[ms.m(i)];
            Arow[i] = split_row(Wk, Wm, 0);
            Brow[i] = split_row(Wk, Wm, 1);
          }
        }
        __syncthreads();
        if constexpr (FULL) {
          // this row's C = u Qx + v Qy spectrum from the packed (C0 + i C1) transform
#pragma unroll
          for (int q = 0; q < P; ++q) L[ms.k(q)] = z[2][q];
          __syncthreads();
#pragma unroll
          for (int i = 0; i < MAXI; ++i) {
            const int kr = t + i * TPT;
            if (kr < a.kout) {
              const cplx<T> Cv = split_row(L[ms.k(i)], L[ms.m(i)], row);
              const T kv = krwave(g, kr);
              // P' = -i kr (u q)^ - (u Qx + v Qy)^
              Arow[i] = mk(kv * Arow[i].im - Cv.re, -kv * Arow[i].re - Cv.im);
            }
          }
          __syncthreads();
        }
        if (active) {
#pragma unroll
          for (int i = 0; i < MAXI; ++i) {
            const int kr = t + i * TPT;
            if (kr < a.kout) {
              const uint32_t o = boff32(kr, y0 + row, ny);
              at(oA, o) = Arow[i];
              at(oB, o) = Brow[i];
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K2 for rows whose length nx is not a power of two (even, 18..GEN_MAX): one workgroup per row,
// direct sums over the row's modes and points (O(nx^2) per row) in place of the packed two-row
// transforms. The same fields, products and outputs as k2_body's modes; the row's half spectra,
// physical products and the nx roots of unity w^j = exp(-2 pi i j/nx) (the global table) sit in
// LDS.
// ---------------------------------------------------------------------------
constexpr int K2G_THREADS = 64;
template <typename T> struct K2GLds {
  cplx<T> hs[3][GEN_MAX / 2 + 1];
  T ph[3][GEN_MAX];
  cplx<T> roots[GEN_MAX];
};
// c2r of the half spectrum h (columns kr < kin present) at point x: Re h_0 + 2 sum Re(h_k w^-kx)
// + Re h_(nx/2) (-1)^x; the imaginary parts of the DC and Nyquist columns are ignored, as c2r does
template <typename T>
__device__ __forceinline__ T c2r_point(const cplx<T>* h, int kin, int nx, int x, const cplx<T>* roots) {
  const int kc = kin < nx / 2 + 1 ? kin : nx / 2 + 1;
  T acc = kc > 0 ? h[0].re : (T)0;
  int idx = 0;
  for (int k = 1; k < kc; ++k) {
    idx += x;
    if (idx >= nx) idx -= nx;                          // k x mod nx
    const cplx<T> w = roots[idx];
    if (2 * k == nx) acc += h[k].re * w.re;
    else acc += (T)2 * (h[k].re * w.re + h[k].im * w.im);   // 2 Re(h_k conj(w))
  }
  return acc;
}
// r2c of the physical row p at column kr: sum_x p_x w^(kr x)
template <typename T>
__device__ __forceinline__ cplx<T> r2c_point(const T* p, int nx, int kr, const cplx<T>* roots) {
  cplx<T> acc = czero<T>();
  int idx = 0;
  for (int x = 0; x < nx; ++x) {
    acc = acc + rmul(p[x], roots[idx]);
    idx += kr;
    if (idx >= nx) idx -= nx;
  }
  return acc;
}
template <typename T, int MODE>
__global__ __launch_bounds__(K2G_THREADS) void k2_gen_kernel(const K2Args<T> a) {
  __shared__ K2GLds<T> S;
  const Geo<T>& g = a.g;
  const int nx = g.nx, ny = g.ny, nkr = nx / 2 + 1, t = (int)threadIdx.x, y = (int)blockIdx.x;
  for (int i = t; i < nx; i += K2G_THREADS) S.roots[i] = a.tw[i];
  // half spectrum of row y of field f (kr < kin, zero beyond) into S.hs[s]
  auto load = [&](int s, const cplx<T>* f) {
    for (int k = t; k < nkr; k += K2G_THREADS) S.hs[s][k] = k < a.kin ? at(f, boff32(k, y, ny)) : czero<T>();
  };
  if constexpr (MODE == K2_INVERSE_OUT) {
    for (int f = 0; f < a.nfields; ++f) {
      __syncthreads();
      load(0, a.in + (long)f * a.fstride);
      __syncthreads();
      for (int x = t; x < nx; x += K2G_THREADS) a.phys[f * a.pstride + (long)y * nx + x] = c2r_point(S.hs[0], a.kin, nx, x, S.roots);
    }
  } else if constexpr (MODE == K2_FORWARD_IN) {
    for (int f = 0; f < a.nfields; ++f) {
      __syncthreads();
      for (int x = t; x < nx; x += K2G_THREADS) S.ph[0][x] = a.phys_in[f * a.pstride + (long)y * nx + x];
      __syncthreads();
      cplx<T>* o = a.out + (long)f * a.ostride;
      for (int kr = t; kr < a.kout; kr += K2G_THREADS) at(o, boff32(kr, y, ny)) = r2c_point(S.ph[0], nx, kr, S.roots);
    }
  } else {
    // K2_ADVECT: u (q [+ eta]), v (q [+ eta]) (src/twodnavierstokes.jl:262-270,
    // src/singlelayerqg.jl:368-375); K2_ADVECT_FULL, per layer: u += U,
    // (-i kr (u q)^ - (u Qx + v Qy)^, (v q)^) (src/multilayerqg.jl:695-731), or with `linear`
    // U q in place of u q and no v q (src/multilayerqg.jl:744-779)
    constexpr bool FULL = MODE == K2_ADVECT_FULL;
    const int nl = FULL ? g.nlayers : 1;
    for (int jl = 0; jl < nl; ++jl) {
      __syncthreads();
      for (int s = 0; s < 3; ++s) load(s, a.in + (long)(3 * jl + s) * a.fstride);
      __syncthreads();
      for (int x = t; x < nx; x += K2G_THREADS) {
        T u = c2r_point(S.hs[0], a.kin, nx, x, S.roots);
        const T v = c2r_point(S.hs[1], a.kin, nx, x, S.roots);
        T q = c2r_point(S.hs[2], a.kin, nx, x, S.roots);
        if constexpr (!FULL) {
          if (a.eta != nullptr) q = q + a.eta[(long)y * nx + x];
          S.ph[0][x] = u * q;
          S.ph[1][x] = v * q;
        } else {
          T c = 0, U = 0;
          if (a.has_bg) {
            U = a.Uy[jl * ny + y];
            u = u + U;
            T qy = a.Qy1[jl * ny + y], qx = 0;
            if (jl == nl - 1) {
              if (a.etax) qx = a.etax[(long)y * nx + x];
              if (a.etay) qy = qy + a.etay[(long)y * nx + x];
            }
            c = u * qx + v * qy;
          }
          S.ph[0][x] = a.linear ? U * q : u * q;
          S.ph[1][x] = a.linear ? (T)0 : v * q;
          S.ph[2][x] = c;
        }
      }
      __syncthreads();
      cplx<T>* oA = a.out + (long)(FULL ? 2 * jl : 0) * a.ostride;
      cplx<T>* oB = oA + a.ostride;
      for (int kr = t; kr < a.kout; kr += K2G_THREADS) {
        cplx<T> A = r2c_point(S.ph[0], nx, kr, S.roots);
        const cplx<T> B = r2c_point(S.ph[1], nx, kr, S.roots);
        if constexpr (FULL) {
          const cplx<T> C = r2c_point(S.ph[2], nx, kr, S.roots);
          const T kv = krwave(g, kr);
          A = mk(kv * A.im - C.re, -kv * A.re - C.im);   // -i kr (u q)^ - (u Qx + v Qy)^
        }
        at(oA, boff32(kr, y, ny)) = A;
        at(oB, boff32(kr, y, ny)) = B;
      }
    }
  }
}
template <typename T>
void k2_gen(const K2Args<T>& a, hipStream_t s) {
  const dim3 grid(a.g.ny), blk(K2G_THREADS);
  switch (a.mode) {
    case K2_ADVECT: GF_LAUNCH((k2_gen_kernel<T, K2_ADVECT>), grid, blk, 0, s, a); break;
    case K2_ADVECT_FULL: GF_LAUNCH((k2_gen_kernel<T, K2_ADVECT_FULL>), grid, blk, 0, s, a); break;
    case K2_INVERSE_OUT: GF_LAUNCH((k2_gen_kernel<T, K2_INVERSE_OUT>), grid, blk, 0, s, a); break;
    case K2_FORWARD_IN: GF_LAUNCH((k2_gen_kernel<T, K2_FORWARD_IN>), grid, blk, 0, s, a); break;
    default: break;
  }
}

#ifndef GF_K2_ADVECT_WAVES
#define GF_K2_ADVECT_WAVES 3   // a floor: at 2048, P = 8 the kernel needs 122 VGPRs and runs at 4
#endif
template <typename T, int NX, int P, int G, int MODE, int NL>
__global__ __launch_bounds__(G * (NX / P))
__attribute__((amdgpu_waves_per_eu(MODE == K2_ADVECT ? GF_K2_ADVECT_WAVES : GF_K2_WAVES)))
void k2_kernel(const K2Args<T> a) {
  __shared__ cplx<T> lds[G * LdsN<NX>::value];
  __shared__ cplx<T> tw[TwPlan<NX, P>::size()];
  load_twiddles<T, NX, P>(tw, a.tw);
  __syncthreads();
  k2_body<T, NX, P, G, MODE, NL>(a, xcd_block(blockIdx.x, gridDim.x), lds, tw);
}

// ---------------------------------------------------------------------------
// K2, staged: the x pass of advection without background gradients, loads one field ahead
// ---------------------------------------------------------------------------
// Each workgroup walks row pairs rp, rp + gridDim.x, ...  A field's two rows reach LDS through
// direct loads into a staging buffer (global_load_lds, 16 bytes per lane), issued one field
// ahead: the rows of u stream in while q is transformed, those of v while u q is formed and
// stored, and q of the next row pair while v q is.  The exchanges inside the transforms use
// lds_barrier, which does not wait for those loads; dma_wait waits for them.  The staging holds
// the kept columns k < kin of the two rows as 64-byte pieces (both rows of a column pair), so
// every four lanes load one contiguous piece, a pattern that moves K2's bytes about 1.25x
// faster than one row per load (22.7 -> 18.3 us for 111.8 MB at 2048^2 FP64, memory alone).
// The arithmetic is K2_ADVECT's, in the same order; the compiler may contract the products into
// the first transform pass differently in the two kernels, so FP64 results agree with k2_kernel's
// under the same P to rounding (about 1e-16 relative), not always bit for bit.
template <typename T, int NX, int P> struct K2Stage {
  static constexpr int EPL = 16 / (int)sizeof(cplx<T>);   // complex values per 16-byte lane load
  static constexpr int SPAN = (NX / P) * EPL;               // slots one load instruction of every wave fills
  // 4 slots per column pair (k <= NX/2 + 1), rounded up to whole spans
  static constexpr int SLOTS = (NX + 4 + SPAN - 1) / SPAN * SPAN;
};
typedef __attribute__((address_space(1))) void gvoid_t;
typedef __attribute__((address_space(3))) void lvoid_t;

// Both rows y0, y0 + 1 of the kept columns k < kin of field f into S: slot 4 p + 2 row + c holds
// column 2 p + c of row y0 + row, i.e. the 64-byte piece of column pair p at rows y0, y0 + 1.
// Lane L of wave w writes slots j SPAN + 64 EPL w + EPL L .. + EPL - 1 in round j.  Every lane
// issues every load of a round (lanes past the kept columns re-read the last piece into slots
// nobody reads): with a load under a lane-dependent branch the compiler waits for it at the
// branch's end.  The trip count is uniform.  Blocked pair layout only.
template <typename T, int NX, int P>
__device__ __forceinline__ void stage_rows(const cplx<T>* f, int y0, int ny, int kin, cplx<T>* S) {
  static_assert(BW == 2, "the staged x pass reads the pair layout");
  using KS = K2Stage<T, NX, P>;
  const int nslots = 4 * ((kin + 1) / 2);
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  const int rounds = (nslots + KS::SPAN - 1) / KS::SPAN;
  const char* fb = reinterpret_cast<const char*>(f);
  for (int j = 0; j < rounds; ++j) {
    const int base = j * KS::SPAN + wave * 64 * KS::EPL;
    const int e0 = base + lane * KS::EPL;
    const int e = e0 < nslots ? e0 : nslots - KS::EPL;
    const uint32_t off = (((uint32_t)(e >> 2) * (uint32_t)ny + (uint32_t)y0) * 2u + (uint32_t)(e & 3)) * (uint32_t)sizeof(cplx<T>);
    __builtin_amdgcn_global_load_lds((gvoid_t*)(fb + off), (lvoid_t*)(S + base), 16, 0, 0);
  }
}
// Every wave's staged loads (and earlier global accesses) complete, then a barrier.
__device__ __forceinline__ void dma_wait() { asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Hermitian-extended entry k of the staged half spectrum of row `row` (herm_load's rule).
template <typename T, int NX>
__device__ __forceinline__ cplx<T> staged(const cplx<T>* S, int k, int row, int kin) {
  const int kk = k <= NX / 2 ? k : NX - k;
  const cplx<T> v = S[4 * (kk >> 1) + 2 * row + (kk & 1)];
  const T im = (kk == 0 || kk == NX / 2) ? (T)0 : (k > NX / 2 ? -v.im : v.im);
  return kk < kin ? mk(v.re, im) : czero<T>();
}

// Staged rows -> packed inverse input conj(A + i B) for k <= NX/2, the mirrored half through L.
// (Plain functions rather than lambdas: with the lambdas' captures the compiler lost track of
// which LDS buffer the staged loads write and waited for them before every LDS read.)
template <typename T, int NX, int P>
__device__ __forceinline__ void staged_take(cplx<T> (&w)[1][P], const cplx<T>* __restrict__ S,
                                            cplx<T>* __restrict__ L, int t, int kin) {
  constexpr int TPT = NX / P;
  constexpr int QH = P / 2 + 1;                     // the q with TPT q <= NX/2
#pragma unroll
  for (int q = 0; q < QH; ++q) {
    const int k = t + TPT * q;
    if (k > NX / 2) continue;
    const cplx<T> A = staged<T, NX>(S, k, 0, kin), B = staged<T, NX>(S, k, 1, kin);
    w[0][q] = mk(A.re - B.im, -(A.im + B.re));
    if (k > 0 && k < NX / 2) L[padg<XPad<T>::G>(NX - k)] = mk(A.re + B.im, A.im - B.re);
  }
  lds_barrier();
#pragma unroll
  for (int q = 0; q < P; ++q) {
    if (TPT * q + TPT - 1 <= NX / 2) continue;
    const int k = t + TPT * q;
    if (k <= NX / 2) continue;
    w[0][q] = L[padg<XPad<T>::G>(k)];
  }
  lds_barrier();   // S and L are free again
}
// Forward transform of the packed product rows y0, y0 + 1 and their split into kept columns of o.
template <typename T, int NX, int P>
__device__ __forceinline__ void staged_store(cplx<T>* o, cplx<T> (&w)[1][P], int y0, int ny, int kout,
                                             cplx<T>* __restrict__ L, const cplx<T>* __restrict__ tw, int t) {
  constexpr int TPT = NX / P;
  constexpr int MAXI = (NX / 2 + 1 + TPT - 1) / TPT;
  constexpr bool FR = NX >= FreshFrom<T>::k2;
  fft_forward<T, NX, P, 1, FR, true>(w, L, t, tw);
  const MirrorSlots<NX, TPT, XPad<T>::G> ms(t);
#pragma unroll
  for (int q = 0; q < P; ++q) L[ms.k(q)] = w[0][q];
  lds_barrier();
#pragma unroll
  for (int i = 0; i < MAXI; ++i) {
    const int kr = t + i * TPT;
    if (kr < kout) {
      const cplx<T> Wk = L[ms.k(i)], Wm = L[ms.m(i)];
      at(o, boff32(kr, y0, ny)) = split_row(Wk, Wm, 0);
      at(o, boff32(kr, y0 + 1, ny)) = split_row(Wk, Wm, 1);
    }
  }
  lds_barrier();
}

template <typename T, int NX, int P>
__device__ __forceinline__ void k2_staged_body(const K2Args<T>& a, cplx<T>* __restrict__ L, cplx<T>* __restrict__ S,
                                               const cplx<T>* __restrict__ tw) {
  constexpr int TPT = NX / P;
  constexpr bool FR = NX >= FreshFrom<T>::k2;
  const int t = (int)threadIdx.x;
  const Geo<T>& g = a.g;
  const int ny = g.ny, npairs = ny / 2, kin = a.kin;
  const cplx<T>* fu = a.in;
  const cplx<T>* fv = a.in + a.fstride;
  const cplx<T>* fq = a.in + 2 * a.fstride;
  const int nwg = (int)gridDim.x;
  const int first = xcd_block((int)blockIdx.x, nwg);
  if (first < npairs) stage_rows<T, NX, P>(fq, 2 * first, ny, kin, S);
  for (int rp = first; rp < npairs; rp += nwg) {
    const int y0 = 2 * rp;
    T q0[P], q1[P];
    {
      cplx<T> w[1][P];
      dma_wait();
      staged_take<T, NX, P>(w, S, L, t, kin);
      stage_rows<T, NX, P>(fu, y0, ny, kin, S);
      fft_forward<T, NX, P, 1, FR, true>(w, L, t, tw);
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int x = t + TPT * q;
        q0[q] = w[0][q].re;
        q1[q] = -w[0][q].im;
        if (a.eta != nullptr) {
          q0[q] = q0[q] + at(a.eta, (uint32_t)(y0 * NX + x));
          q1[q] = q1[q] + at(a.eta, (uint32_t)((y0 + 1) * NX + x));
        }
      }
    }
    {
      cplx<T> w[1][P];
      dma_wait();
      staged_take<T, NX, P>(w, S, L, t, kin);
      stage_rows<T, NX, P>(fv, y0, ny, kin, S);
      fft_forward<T, NX, P, 1, FR, true>(w, L, t, tw);
#pragma unroll
      for (int q = 0; q < P; ++q) w[0][q] = mk(w[0][q].re * q0[q], -w[0][q].im * q1[q]);   // u q, both rows
      staged_store<T, NX, P>(a.out, w, y0, ny, a.kout, L, tw, t);
    }
    {
      cplx<T> w[1][P];
      dma_wait();
      staged_take<T, NX, P>(w, S, L, t, kin);
      if (rp + nwg < npairs) stage_rows<T, NX, P>(fq, 2 * (rp + nwg), ny, kin, S);
      fft_forward<T, NX, P, 1, FR, true>(w, L, t, tw);
#pragma unroll
      for (int q = 0; q < P; ++q) w[0][q] = mk(w[0][q].re * q0[q], -w[0][q].im * q1[q]);   // v q, both rows
      staged_store<T, NX, P>(a.out + a.ostride, w, y0, ny, a.kout, L, tw, t);
    }
  }
}

template <typename T, int NX, int P>
__global__ __launch_bounds__(NX / P) __attribute__((amdgpu_waves_per_eu(2)))
void k2_staged_kernel(const K2Args<T> a) {
  __shared__ cplx<T> L[LdsN<NX>::value];
  __shared__ cplx<T> S[K2Stage<T, NX, P>::SLOTS];
  __shared__ cplx<T> tw[TwPlan<NX, P>::size()];
  load_twiddles<T, NX, P>(tw, a.tw);
  __syncthreads();
  k2_staged_body<T, NX, P>(a, L, S, tw);
}

// Workgroups of the staged x pass: as many as the CUs hold at once (by LDS), at most one per
// row pair.
template <typename T, int NX, int P>
int k2_staged_blocks(int npairs) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  }
  constexpr int bytes = (int)sizeof(cplx<T>) * (LdsN<NX>::value + K2Stage<T, NX, P>::SLOTS + TwPlan<NX, P>::size());
  constexpr int per_cu = 160 * 1024 / bytes > 0 ? 160 * 1024 / bytes : 1;
  const int n = cus * per_cu;
  return n < npairs ? n : npairs;
}

// ---------------------------------------------------------------------------
// K3: flux -> spectral N + stepper stage epilogue, one column / group
// ---------------------------------------------------------------------------
// Column of single-column workgroup b: the BW columns of block q go to workgroups
// 8 BW (q / 8) + (q mod 8) + 8 c, c < BW, which the dispatcher hands to the same XCD (workgroup b
// runs on XCD b mod 8), so the later reads of each 16 BW-byte piece of the blocked layout (one
// row of the block) hit in that XCD's L2.
__device__ __forceinline__ int single_column(int b) { return BW * (8 * (b / (8 * BW)) + b % 8) + (b / 8) % BW; }

// Body of K3 for block blk of column pairs (CPW = 2, G pairs per workgroup) or, with CPW = 1,
// for the single column single_column(blk) (G = 1).  EPI >= 0 fixes the stage epilogue at
// compile time (it must equal a.epi); EPI = -1 switches on a.epi.  ML1: a one-layer MultiLayerQG
// problem (SingleLayerParams), whose bottom drag mu Krsq psih uses psih = -invKrsq qh of the calcN
// input (src/multilayerqg.jl:658, 547-551) and whose linearised equations step every mode.
template <typename T, int NY, int P, int G, int NL, int CPW = 2, int EPI = -1, bool ML1 = false, bool GEN = false>
__device__ __forceinline__ void k3_body(const K3Args<T>& a, int blk, cplx<T>* __restrict__ lds,
                                        const cplx<T>* __restrict__ tw) {
  constexpr int TPT = NY / P;
  static_assert(CPW == 2 || G == 1, "single-column workgroups hold one column");
  int t, kr;
  cplx<T>* L;
  if constexpr (CPW == 2) {
    const YMap<TPT> m(threadIdx.x);
    t = m.t;
    kr = 2 * (G == 1 ? pair_of_block(blk) : blk * G + m.slot) + m.c;
    L = col_buffer<T, NY>(lds, m.slot, m.c);
  } else {
    t = threadIdx.x;
    kr = single_column(blk);
    L = lds;
  }
  const Geo<T>& g = a.g;
  const bool active = kr < a.ncols && kr < g.nkr;
  const T kv = krwave(g, kr);
  const int epi = EPI >= 0 ? EPI : a.epi;
  const int ny = ycol<NY, GEN>(g);

  // Ring forcing (examples/twodnavierstokes_stochasticforcing.jl:78-86: N += sqrt(spectrum) xi
  // / sqrt(dt), xi complex standard normal per mode and step): this column's terms are drawn
  // now, while few registers are live, and read back in the epilogue.  Drawn there, the Philox
  // and Box-Muller sequence set the whole kernel's register count (144 VGPRs instead of 118 at
  // 2048, P = 8), although only the box's few columns execute it.  A workgroup row of its own
  // per layer (lsplit) has its own slice of the box, so every workgroup reads back only what it
  // wrote.
  cplx<T>* const noise = a.noise + ((NL > 1 && a.lsplit) ? (long)blockIdx.y * (a.fkx + 1) * a.nbl : 0);
  if (a.sqrt_spec != nullptr && epi != E_STORE && active && kr <= a.fkx) {
    const int64_t fstep = a.fstep_dev != nullptr ? a.fstep_dev[0] : a.fstep;
    for (int q = 0; q < P; ++q) {
      const int l = t + TPT * q;
      if (in_col<GEN>(l, ny) && (l <= a.fky || l >= ny - a.fky)) {
        const T sp = at(a.sqrt_spec, boff32(kr, l, ny));
        if (sp != (T)0) {
          const cplx<T> xi = complex_normal<T>(a.seed, fstep, (uint32_t)(l * g.nkr + kr), (uint32_t)a.instance);
          noise[kr * a.nbl + ring_box_row(l, ny, a.fky, a.nbl)] = mk((xi.re * sp) / a.sqrt_dt, (xi.im * sp) / a.sqrt_dt);
        }
      }
    }
  }

  if (epi == E_STORE) {
    // forward y-transform of mixed field(s) -> spectral state (set_q!/set_zeta!)
    for (int jl = 0; jl < g.nlayers; ++jl) {
      cplx<T> w[1][P];
      const cplx<T>* f = a.in + (long)jl * a.fstride;
#pragma unroll
      for (int q = 0; q < P; ++q) w[0][q] = (active && in_col<GEN>(t + TPT * q, ny)) ? f[boff(kr, t + TPT * q, ny)] : czero<T>();
      y_transform<T, NY, P, (NY >= FreshFrom<T>::k3), GEN>(w, L, t, tw, ny);
      if (active) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
          const int l = t + TPT * q;
          if (!in_col<GEN>(l, ny)) continue;
          cplx<T> v = w[0][q];
          if (a.zero_mean && kr == 0 && l == 0) v = czero<T>();
          a.sol[jl * g.lstride + boff(kr, l, ny)] = v;
        }
      }
    }
    return;
  }

  // Bottom drag mu Krsq psih_n (src/multilayerqg.jl:658) needs psi_n of the calcN
  // input; the stage epilogue below may overwrite that input in place (ETDRK4
  // stage 3, RK4 stages 2-3), so psi_n is formed before any layer is written.
  // With one layer per workgroup (lsplit), K1 of the same stage has stored it (psi_in): a
  // workgroup of another layer may already have overwritten that layer's input.
  cplx<T> psin[P];
  const int jl0 = (NL > 1 && a.lsplit) ? (int)blockIdx.y : 0;
  const int jl1 = (NL > 1 && a.lsplit) ? jl0 + 1 : NL;
  if constexpr (NL > 1) {
    if (jl1 == NL) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int l = t + TPT * q;
        cplx<T> ps = czero<T>();
        if (active && in_col<GEN>(l, ny) && a.drag_src != nullptr && ((kept_l(g, l) && kr < g.kmax) || a.nomask)) {
          const long idx = boff(kr, l, ny);
          if (a.psi_in != nullptr) {
            ps = a.psi_in[idx];
          } else {
            const T* m = a.sinv + idx * (NL * NL) + (NL - 1) * NL;
#pragma unroll
            for (int k = 0; k < NL; ++k) ps = ps + rmul(m[k], a.drag_src[k * g.lstride + idx]);
          }
        }
        psin[q] = ps;
      }
    }
  }
  for (int jl = jl0; jl < jl1; ++jl) {
    cplx<T> A[1][P], B[1][P];
    const cplx<T>* fa = a.in + (long)(2 * jl) * a.fstride;
    const cplx<T>* fb = a.in + (long)(2 * jl + 1) * a.fstride;
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int y = t + TPT * q;
      A[0][q] = (active && in_col<GEN>(y, ny)) ? at(fa, boff32(kr, y, ny)) : czero<T>();
      B[0][q] = (active && in_col<GEN>(y, ny)) ? at(fb, boff32(kr, y, ny)) : czero<T>();
    }
    y_transform<T, NY, P, (NY >= FreshFrom<T>::k3), GEN>(A, L, t, tw, ny);
    y_transform<T, NY, P, (NY >= FreshFrom<T>::k3), GEN>(B, L, t, tw, ny);
    if (!active) continue;
    const long lo = (long)jl * g.lstride;
    bool bad = false;   // a non-finite value in the new state (final stages)
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int l = t + TPT * q;
      if (!in_col<GEN>(l, ny)) continue;
      const bool kept = (kept_l(g, l) && kr < g.kmax) || ((NL > 1 || ML1) && a.nomask);
      if (!kept && !a.full) continue;
      const T lv = lwave(g, l);
      const uint32_t idx = boff32(kr, l, ny);
      const uint32_t ix = (uint32_t)lo + idx;
      cplx<T> N;
      const cplx<T> Av = A[0][q], Bv = B[0][q];
      if (a.apply_kr) {
        // N = -i kr (u zeta)^ - i l (v zeta)^ (src/twodnavierstokes.jl:272)
        N = mk(kv * Av.im + lv * Bv.im, -kv * Av.re - lv * Bv.re);
      } else {
        N = mk(Av.re + lv * Bv.im, Av.im - lv * Bv.re);
      }
      if (a.etah != nullptr) {
        // N -= i kr U etah (src/singlelayerqg.jl:379)
        const cplx<T> e = at(a.etah, idx);
        const T cu = kv * a.Ueta;
        N = mk(N.re + cu * e.im, N.im - cu * e.re);
      }
      if constexpr (NL > 1) {
        if (a.drag_src != nullptr && jl == NL - 1) {
          // bottom drag N_n += mu Krsq psih_n (src/multilayerqg.jl:658)
          const T K2 = kv * kv + lv * lv;
          N = N + rmul(a.mu_drag * K2, psin[q]);
        }
      }
      if constexpr (ML1) {
        if (a.drag_src != nullptr) {
          // one layer: psih = -invKrsq qh of the (masked) calcN input, read before this point of
          // the in-place stages is written; N += mu Krsq psih
          const T K2 = kv * kv + lv * lv;
          const T iK = (kr == 0 && l == 0) ? (T)0 : (T)1 / K2;
          const cplx<T> ps = kept ? rmul(-iK, at(a.drag_src, ix)) : czero<T>();
          N = N + rmul(a.mu_drag * K2, ps);
        }
      }
      if (a.Fh != nullptr) N = N + rmul(a.fscale, at(a.Fh, ix));
      if (a.sqrt_spec != nullptr) {
        if (kr <= a.fkx && (l <= a.fky || l >= ny - a.fky)) {
          if (at(a.sqrt_spec, idx) != (T)0) N = N + noise[kr * a.nbl + ring_box_row(l, ny, a.fky, a.nbl)];
        }
      }
      // The stepper's coefficient tables and the filter depend on kr and Krsq only, so rows l and
      // NY - l hold equal values: reading the lower row lets the mirrored point's read hit in
      // cache and halves the tables' HBM traffic.
      const uint32_t cf = boff32(kr, l <= ny / 2 ? l : ny - l, ny);
      const uint32_t ci = a.cl_layered ? ix : cf;
      const cplx<T> svr = at(a.sol, ix);     // every lane of an active column has this point
      const cplx<T> sv = kept ? svr : czero<T>();
      if (a.nsave != nullptr) at(a.nsave, ix) = N;
      switch (epi) {
        case E_CALCN:
          at(a.nout, ix) = N;
          break;
        // The stages compute for every lane and store only where kept: loads under a branch on
        // kept made the compiler wait for each before issuing the next.
        case E_ETD1: {  // sol1 = E2 sol + zeta N1
          const Coef<T> E2{a.c0, a.ccplx}, Z{a.c1, a.ccplx};
          const cplx<T> r = E2.mul(ci, sv) + Z.mul(ci, N);
          if (kept) at(a.s1, ix) = r;
          at(a.n1, ix) = N;
        } break;
        case E_ETD2: {  // sol2 = E2 sol + zeta N2
          const Coef<T> E2{a.c0, a.ccplx}, Z{a.c1, a.ccplx};
          const cplx<T> r = E2.mul(ci, sv) + Z.mul(ci, N);
          if (kept) at(a.s2, ix) = r;
          at(a.n23, ix) = N;
        } break;
        case E_ETD3: {  // sol2 = E2 sol1 + zeta (2 N3 - N1)
          const Coef<T> E2{a.c0, a.ccplx}, Z{a.c1, a.ccplx};
          const cplx<T> s1v = at(a.s1, ix);
          const cplx<T> n1v = at(a.n1, ix);
          const cplx<T> n23v = at(a.n23, ix);
          const cplx<T> r = E2.mul(ci, s1v) + Z.mul(ci, mk((T)2 * N.re - n1v.re, (T)2 * N.im - n1v.im));
          if (kept) at(a.s2, ix) = r;
          at(a.n23, ix) = n23v + N;
        } break;
        case E_ETD4: {  // sol = E sol + alpha N1 + 2 beta (N2 + N3) + Gamma N4
          const Coef<T> E{a.c2, a.ccplx}, Al{a.c3, a.ccplx}, Be{a.c4, a.ccplx}, Ga{a.c5, a.ccplx};
          const cplx<T> n23v = at(a.n23, ix);
          cplx<T> r = E.mul(ci, sv) + Al.mul(ci, at(a.n1, ix));
          r = r + Be.mul(ci, mk((T)2 * n23v.re, (T)2 * n23v.im));
          r = r + Ga.mul(ci, N);
          if (a.filter) r = rmul(at(a.filter, cf), r);
          at(a.sol_out != nullptr ? a.sol_out : a.sol, ix) = r;
          bad |= !finite(r);
        } break;
        case E_RK1:
        case E_RK2:
        case E_RK3:
        case E_RK4: {
          const Coef<T> Lc{a.c0, a.ccplx};
          const cplx<T> xin = (epi == E_RK1) ? sv : (kept ? at(a.s1, ix) : czero<T>());
          const cplx<T> R = N + Lc.mul(ci, xin);   // RHS = N + L sol_i
          if (epi == E_RK1) {
            if (kept) at(a.s1, ix) = sv + rmul(a.dt / (T)2, R);
            at(a.racc, ix) = mk(R.re / (T)6, R.im / (T)6);
          } else if (epi == E_RK2) {
            if (kept) at(a.s1, ix) = sv + rmul(a.dt / (T)2, R);
            const cplx<T> ra = at(a.racc, ix);
            at(a.racc, ix) = mk(ra.re + R.re / (T)3, ra.im + R.im / (T)3);
          } else if (epi == E_RK3) {
            if (kept) at(a.s1, ix) = sv + rmul(a.dt, R);
            const cplx<T> ra = at(a.racc, ix);
            at(a.racc, ix) = mk(ra.re + R.re / (T)3, ra.im + R.im / (T)3);
          } else {
            const cplx<T> ra = at(a.racc, ix);
            const cplx<T> acc = mk(ra.re + R.re / (T)6, ra.im + R.im / (T)6);
            cplx<T> r = sv + rmul(a.dt, acc);
            if (a.filter) r = rmul(at(a.filter, cf), r);
            at(a.sol, ix) = r;
            bad |= !finite(r);
          }
        } break;
        case E_FE: {
          const Coef<T> Lc{a.c0, a.ccplx};
          cplx<T> r = sv + rmul(a.dt, N + Lc.mul(ci, sv));
          if (a.filter) r = rmul(at(a.filter, cf), r);
          at(a.sol, ix) = r;
          bad |= !finite(r);
        } break;
        case E_AB3: {
          // RHS = N + L sol; forward Euler start-up, then 23/12 RHS - 16/12 RHS_1 + 5/12 RHS_2;
          // the history (racc = RHS_1, n1 = RHS_2) shifts every step ([FF] AB3TimeStepper)
          const Coef<T> Lc{a.c0, a.ccplx};
          const cplx<T> R = N + Lc.mul(ci, sv);
          const cplx<T> R1 = at(a.racc, ix);
          cplx<T> r;
          if (a.ab3_start) {
            r = sv + rmul(a.dt, R);
          } else {
            const cplx<T> R2 = at(a.n1, ix);
            const cplx<T> acc = (rmul((T)(23.0 / 12.0), R) - rmul((T)(16.0 / 12.0), R1)) + rmul((T)(5.0 / 12.0), R2);
            r = sv + rmul(a.dt, acc);
          }
          if (a.filter) r = rmul(at(a.filter, cf), r);
          at(a.sol, ix) = r;
          at(a.n1, ix) = R1;
          at(a.racc, ix) = R;
          bad |= !finite(r);
        } break;
        default: break;
      }
      // With 8 points per thread, one point's loads and stores at a time keeps the epilogue
      // within 128 VGPRs (two 512-thread workgroups per CU) with less spilling than eight
      // points in flight; with 4 the registers allow all points' loads to be in flight at once.
      if constexpr (P >= 8) __builtin_amdgcn_sched_barrier(0);
    }
    // one store per wave (the lowest lane that saw a non-finite value): with a non-finite state
    // every lane would otherwise write the flag in host memory
    const unsigned long long badm = __ballot(bad);
    if (badm != 0 && a.nonfinite != nullptr && (int)(threadIdx.x & 63) == __ffsll(badm) - 1) *a.nonfinite = 1;
  }
}

// Occupancy target of K3.  The multi-layer kernels need about 170 VGPRs (the bottom drag's psi
// per point, layered coefficient tables): held to 128 for 4 waves they spilled 30-50.  At 2
// waves they do not spill, and a CU still holds more of their workgroups than the layered grids
// supply (513 single-column workgroups at 512^2 x 3).  At 4096 points (16 per thread) the
// epilogue's registers double too, and one 147 KB column-pair workgroup fills a CU's LDS anyway.
template <int N, int NL, int CPW> struct K3Waves { static constexpr int value = (NL > 1 || N >= 4096) ? 2 : GF_K3_WAVES; };
template <typename T, int NY, int P, int G, int NL, int CPW = 2, int EPI = -1, bool ML1 = false, bool GEN = false>
__global__ __launch_bounds__(CPW * G * (NY / P)) __attribute__((amdgpu_waves_per_eu(K3Waves<NY, NL, CPW>::value)))
void k3_kernel(const K3Args<T> a) {
  __shared__ cplx<T> lds[CPW == 2 ? G * PairLds<T, NY>::value : LdsN<NY>::value];
  __shared__ cplx<T> tw[YTw<NY, P, GEN>::size];
  load_y_twiddles<T, NY, P, GEN>(tw, a.tw, a.g.ny);
  __syncthreads();
  k3_body<T, NY, P, G, NL, CPW, EPI, ML1, GEN>(a, blockIdx.x, lds, tw);
}

// ---------------------------------------------------------------------------
// size dispatch helpers
// ---------------------------------------------------------------------------
#ifndef GF_P2048
#define GF_P2048 8
#endif
template <int N> struct PY { static constexpr int P = N >= 4096 ? 16 : (N == 2048 ? GF_P2048 : (N >= 32 ? 8 : 4)); };
// K2: G thread groups of N/P threads (one row pair each), at least one wave per workgroup.  The
// two row pairs that share a 128-byte line of the blocked layout run in neighbouring workgroups
// on one XCD (xcd_block), so one row pair per workgroup (three 256-thread workgroups per CU at
// 2048) reads each line from HBM once.
#ifndef GF_K2_MIN_G
#define GF_K2_MIN_G 1
#endif
template <int N, int P> struct GW {
  static constexpr int G0 = (N / P) >= 64 ? 1 : 64 / (N / P);
  static constexpr int G = G0 < GF_K2_MIN_G ? GF_K2_MIN_G : G0;
};
// K1/K3: G column pairs of 2 N/P threads each, at least one wave per workgroup
template <int N, int P> struct GY { static constexpr int G = 2 * (N / P) >= 64 ? 1 : 64 / (2 * (N / P)); };

// Radix schedules the autotuner may choose for the single-layer calcN kernels: FFT points per
// thread P.  P = 4 doubles the threads per transform and halves each thread's butterfly chain
// and register footprint.  P = 16 (from 1024 points) needs one LDS exchange fewer per transform
// (2048 = 16 16 8) at twice the registers, which Float32 can afford.  Every other path uses the
// built-in choice.  Keep in step with sched_candidate() on the host.
template <int N> struct PCand {
  // (4 points per thread stops at 2048: a K1/K3 column pair of 4096 points would need 2048 threads)
  static constexpr bool has(int p) {
    return p == 4 ? (N >= 16 && N <= 2048) : (p == 8 ? N >= 32 : (p == 16 ? N >= 1024 : false));
  }
};
// Points per thread of the multi-layer y passes: 4 (more, shorter transforms) up to 2048, 8 at
// 4096 (a column pair of 4096 points at 4 per thread would need 2048 threads)
template <int N> struct PML { static constexpr int P = N >= 4096 ? 8 : (N >= 32 ? 4 : PY<N>::P); };

template <typename T, int N, int NL, int P, bool GEN = false>
void k1_np(const K1Args<T>& a, hipStream_t s) {
  constexpr int G = GY<N, P>::G;
  const int nblk = pair_blocks(a.ncols, G);
  if (nblk <= 0) return;
  const int gy = (NL > 1 && a.lsplit) ? NL : ((NL == 1 && a.fsplit && a.fieldset == FS_CALCN) ? 3 : 1);
  const dim3 grid(nblk, gy), blk(2 * G * (N / P));
  if constexpr (NL > 1) {
    if (a.fieldset == FS_CALCN) GF_LAUNCH((k1_kernel<T, N, P, G, NL, M_MLQG, FS_CALCN, GEN>), grid, blk, 0, s, a);
    else if (a.fieldset == FS_OUT) GF_LAUNCH((k1_kernel<T, N, P, G, NL, M_MLQG, FS_OUT, GEN>), grid, blk, 0, s, a);
    else GF_LAUNCH((k1_kernel<T, N, P, G, NL, M_MLQG, FS_RAW, GEN>), grid, blk, 0, s, a);
  } else if constexpr (P != PY<N>::P) {   // a tuned schedule: the calcN field set only
    if (a.model == M_TWODNS) GF_LAUNCH((k1_kernel<T, N, P, G, 1, M_TWODNS, FS_CALCN, GEN>), grid, blk, 0, s, a);
    else GF_LAUNCH((k1_kernel<T, N, P, G, 1, M_SLQG, FS_CALCN, GEN>), grid, blk, 0, s, a);
  } else {
    if (a.model == M_TWODNS) {
      if (a.fieldset == FS_CALCN) GF_LAUNCH((k1_kernel<T, N, P, G, 1, M_TWODNS, FS_CALCN, GEN>), grid, blk, 0, s, a);
      else if (a.fieldset == FS_OUT) GF_LAUNCH((k1_kernel<T, N, P, G, 1, M_TWODNS, FS_OUT, GEN>), grid, blk, 0, s, a);
      else GF_LAUNCH((k1_kernel<T, N, P, G, 1, M_TWODNS, FS_RAW, GEN>), grid, blk, 0, s, a);
    } else {
      if (a.fieldset == FS_CALCN) GF_LAUNCH((k1_kernel<T, N, P, G, 1, M_SLQG, FS_CALCN, GEN>), grid, blk, 0, s, a);
      else if (a.fieldset == FS_OUT) GF_LAUNCH((k1_kernel<T, N, P, G, 1, M_SLQG, FS_OUT, GEN>), grid, blk, 0, s, a);
      else GF_LAUNCH((k1_kernel<T, N, P, G, 1, M_SLQG, FS_RAW, GEN>), grid, blk, 0, s, a);
    }
  }
}
template <typename T, int N, int NL>
void k1_n(const K1Args<T>& a, hipStream_t s) {
  if constexpr (NL > 1) {
    if constexpr (N <= GEN_MAX) {
      if (a.g.ny != N) { k1_np<T, N, NL, PML<N>::P, true>(a, s); return; }   // a column length below N
    }
    k1_np<T, N, NL, PML<N>::P>(a, s);
  } else {
    if constexpr (N <= GEN_MAX) {
      if (a.g.ny != N) { k1_np<T, N, 1, PY<N>::P, true>(a, s); return; }
    }
    const bool tuned = a.fieldset == FS_CALCN && a.p != 0 && a.p != PY<N>::P && PCand<N>::has(a.p);
    if (tuned && a.p == 4) { if constexpr (PCand<N>::has(4) && PY<N>::P != 4) { k1_np<T, N, 1, 4>(a, s); return; } }
    if (tuned && a.p == 8) { if constexpr (PCand<N>::has(8) && PY<N>::P != 8) { k1_np<T, N, 1, 8>(a, s); return; } }
    if (tuned && a.p == 16) { if constexpr (PCand<N>::has(16) && PY<N>::P != 16) { k1_np<T, N, 1, 16>(a, s); return; } }
    k1_np<T, N, 1, PY<N>::P>(a, s);
  }
}
template <typename T, int N, int NL, int P>
void k2_np(const K2Args<T>& a, hipStream_t s) {
  constexpr int G = GW<N, P>::G;
  const int gy = (NL > 1 && a.lsplit) ? NL : ((NL == 1 && a.hsplit == 1 && a.mode == K2_ADVECT) ? 2 : 1);
  const dim3 grid((a.g.ny / 2 + G - 1) / G, gy), blk(G * (N / P));
  if constexpr (NL == 1 && BW == 2 && N / P >= 64) {   // whole waves: the staging is issued per wave
    if (a.mode == K2_ADVECT && a.hsplit == 2) {   // the staged x pass (a schedule option)
      GF_LAUNCH((k2_staged_kernel<T, N, P>), dim3(k2_staged_blocks<T, N, P>(a.g.ny / 2)), dim3(N / P), 0, s, a);
      return;
    }
  }
  if constexpr (P != PY<N>::P) {          // a tuned schedule: the calcN modes only
    if (a.mode == K2_ADVECT) { if constexpr (NL == 1) GF_LAUNCH((k2_kernel<T, N, P, G, K2_ADVECT, 1>), grid, blk, 0, s, a); }
    else GF_LAUNCH((k2_kernel<T, N, P, G, K2_ADVECT_FULL, NL>), grid, blk, 0, s, a);
  } else {
    switch (a.mode) {
      case K2_ADVECT: if constexpr (NL == 1) GF_LAUNCH((k2_kernel<T, N, P, G, K2_ADVECT, 1>), grid, blk, 0, s, a); break;
      case K2_ADVECT_FULL: GF_LAUNCH((k2_kernel<T, N, P, G, K2_ADVECT_FULL, NL>), grid, blk, 0, s, a); break;
      case K2_INVERSE_OUT: if constexpr (NL == 1) GF_LAUNCH((k2_kernel<T, N, P, G, K2_INVERSE_OUT, 1>), grid, blk, 0, s, a); break;
      case K2_FORWARD_IN: if constexpr (NL == 1) GF_LAUNCH((k2_kernel<T, N, P, G, K2_FORWARD_IN, 1>), grid, blk, 0, s, a); break;
      default: break;
    }
  }
}
template <typename T, int N, int NL>
void k2_n(const K2Args<T>& a, hipStream_t s) {
  if constexpr (N <= GEN_MAX) {
    if (a.g.nx != N) { k2_gen<T>(a, s); return; }   // a row length below N: direct sums
  }
  const bool tuned = NL == 1 && (a.mode == K2_ADVECT || a.mode == K2_ADVECT_FULL) && a.p != 0 &&
                     a.p != PY<N>::P && PCand<N>::has(a.p);
  if constexpr (NL == 1) {
    if (tuned && a.p == 4) { if constexpr (PCand<N>::has(4) && PY<N>::P != 4) { k2_np<T, N, 1, 4>(a, s); return; } }
    if (tuned && a.p == 8) { if constexpr (PCand<N>::has(8) && PY<N>::P != 8) { k2_np<T, N, 1, 8>(a, s); return; } }
    if (tuned && a.p == 16) { if constexpr (PCand<N>::has(16) && PY<N>::P != 16) { k2_np<T, N, 1, 16>(a, s); return; } }
  }
  (void)tuned;
  if constexpr (NL > 1) {
    // multi-layer x pass: 4 points per thread (128-thread transforms at 512) -- one layer per
    // workgroup row leaves the kernel latency-bound, and more threads per row pair shorten it
    if (a.mode == K2_ADVECT_FULL && N >= 64) { k2_np<T, N, NL, 4>(a, s); return; }
  }
  k2_np<T, N, NL, PY<N>::P>(a, s);
}
template <typename T, int N, int NL, int P, bool GEN = false>
void k3_np(const K3Args<T>& a, hipStream_t s) {
  constexpr int G = GY<N, P>::G;
  if constexpr (NL > 1 && N >= 512 && k3_single_column(N, P)) {
    // multi-layer problems (fixed schedule): one column per workgroup from 512^2, where it
    // measured 3.7% faster over a 3-layer 512^2 ETDRK4 step (0.1915 -> 0.1845 ms)
    const int nblk = column_blocks(a.ncols);
    if (a.ncols > 0) GF_LAUNCH((k3_kernel<T, N, P, 1, NL, 1, -1, false, GEN>), dim3(nblk, a.lsplit ? NL : 1), dim3(N / P), 0, s, a);
    return;
  }
  if constexpr (NL == 1 && P == PY<N>::P) {
    if (a.model == M_MLQG) {   // one-layer MultiLayerQG: fixed schedule, column pairs
      const int nblk = pair_blocks(a.ncols, G);
      if (nblk > 0) GF_LAUNCH((k3_kernel<T, N, P, G, 1, 2, -1, true, GEN>), dim3(nblk), dim3(2 * G * (N / P)), 0, s, a);
      return;
    }
  }
  if constexpr (NL == 1 && k3_single_column(N, P)) {
    static_assert(G == 1, "a single column fills a wave");
    if (a.csplit) {
      // one column per workgroup, a pair's columns 8 workgroups apart (single_column): at 2048^2
      // 683 workgroups at 4 per CU instead of 342 column pairs at 2 per CU on 256 CUs
      const int nblk = column_blocks(a.ncols);
      if (a.ncols <= 0) return;
      const dim3 grid(nblk), blk(N / P);
      if constexpr (std::is_same<T, double>::value && N == 2048 && P == 8) {
        // FP64 at 2048, P = 8: one kernel per ETDRK4 stage.  Each fits the 4-wave budget of
        // 128 VGPRs without spilling (120-128); the kernel that switches on the stage needs
        // 137 and spills 8.
        switch (a.epi) {
          case E_ETD1: GF_LAUNCH((k3_kernel<T, N, P, 1, 1, 1, E_ETD1, false, GEN>), grid, blk, 0, s, a); return;
          case E_ETD2: GF_LAUNCH((k3_kernel<T, N, P, 1, 1, 1, E_ETD2, false, GEN>), grid, blk, 0, s, a); return;
          case E_ETD3: GF_LAUNCH((k3_kernel<T, N, P, 1, 1, 1, E_ETD3, false, GEN>), grid, blk, 0, s, a); return;
          case E_ETD4: GF_LAUNCH((k3_kernel<T, N, P, 1, 1, 1, E_ETD4, false, GEN>), grid, blk, 0, s, a); return;
          default: break;
        }
      }
      GF_LAUNCH((k3_kernel<T, N, P, 1, 1, 1, -1, false, GEN>), grid, blk, 0, s, a);
      return;
    }
  }
  const int nblk = pair_blocks(a.ncols, G);
  if (nblk > 0)
    GF_LAUNCH((k3_kernel<T, N, P, G, NL, 2, -1, false, GEN>), dim3(nblk, (NL > 1 && a.lsplit) ? NL : 1), dim3(2 * G * (N / P)), 0, s, a);
}
template <typename T, int N, int NL>
void k3_n(const K3Args<T>& a, hipStream_t s) {
  if constexpr (NL > 1) {
    if constexpr (N <= GEN_MAX) {
      if (a.g.ny != N) { k3_np<T, N, NL, PML<N>::P, true>(a, s); return; }   // a column length below N
    }
    k3_np<T, N, NL, PML<N>::P>(a, s);
  } else {
    if constexpr (N <= GEN_MAX) {
      if (a.g.ny != N) { k3_np<T, N, 1, PY<N>::P, true>(a, s); return; }
    }
    const bool tuned = a.p != 0 && a.p != PY<N>::P && PCand<N>::has(a.p);
    if (tuned && a.p == 4) { if constexpr (PCand<N>::has(4) && PY<N>::P != 4) { k3_np<T, N, 1, 4>(a, s); return; } }
    if (tuned && a.p == 8) { if constexpr (PCand<N>::has(8) && PY<N>::P != 8) { k3_np<T, N, 1, 8>(a, s); return; } }
    if (tuned && a.p == 16) { if constexpr (PCand<N>::has(16) && PY<N>::P != 16) { k3_np<T, N, 1, 16>(a, s); return; } }
    k3_np<T, N, 1, PY<N>::P>(a, s);
  }
}

}  // namespace gf
