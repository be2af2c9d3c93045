// Discrete adjoint of the TwoDNavierStokes ETDRK4 step (no reference counterpart).
//
// The adjoint of calcN at a masked state s maps the co-state lambda of N to
//   lambda_s = w [ F(u pA + v pB) - i l/|k|^2 F(zeta pA) + i kr/|k|^2 F(zeta pB) ],
//   pA = F^-1(i kr lambda / w),  pB = F^-1(i l lambda / w),
// with u, v, zeta the fields calcN forms from s, F the unnormalised rfft2, F^-1 the
// normalised irfft2 and w the Hermitian weights of the half plane (2 for 0 < kr < nx/2,
// 1 on the kr = 0 and kr = nx/2 columns).  It runs as three kernels shaped like K1/K2/K3:
//   KA1 (columns): 5 inverse y-transforms -> u, v, zeta, pA, pB in mixed space;
//   KA2 (row pairs): 5 inverse x-transforms, the products zeta pA, zeta pB, u pA + v pB,
//       3 forward x-transforms;
//   KA3 (columns): 3 forward y-transforms, lambda_s, and the adjoint stage arithmetic.
// Every co-state is masked (the states calcN reads are), so only the kept columns and rows
// are transformed and written.
#include "gf_kernels_impl.hpp"

namespace gf {

template <typename T> __device__ __forceinline__ T herm_w(const Geo<T>& g, int kr) {
  return (kr == 0 || 2 * kr == g.nx) ? (T)1 : (T)2;
}

// ---------------------------------------------------------------------------
// KA1: state s and co-state lambda = lamc * X -> mixed u, v, zeta, pA, pB
// ---------------------------------------------------------------------------
template <typename T, int NY, int P, int G>
__device__ __forceinline__ void ka1_body(const AdjK1Args<T>& a, int blk, cplx<T>* __restrict__ lds,
                                         const cplx<T>* __restrict__ tw) {
  constexpr int TPT = NY / P;
  const YMap<TPT> m(threadIdx.x);
  const int t = m.t;
  const int kr = 2 * (G == 1 ? pair_of_block(blk) : blk * G + m.slot) + m.c;
  const Geo<T>& g = a.g;
  const bool active = kr < a.ncols && kr < g.nkr;
  const bool krk = kr < g.kmax;
  const T kv = krwave(g, kr);
  const T wi = (T)1 / herm_w(g, kr);
  cplx<T>* L = col_buffer<T, NY>(lds, m.slot, m.c);
  const T nrm = g.inv_norm;
  auto masked = [&](const cplx<T>* f, int l) -> cplx<T> {
    cplx<T> v = czero<T>();
    if (active && krk && kept_l(g, l)) v = at(f, boff32(kr, l, NY));
    return v;
  };
  // the stage's calcN input at (kr, l), rebuilt from the taped s0 and Ns in the forward stages'
  // operation order (the ETD stage epilogues of K3); every lane reads, a select masks
  auto stage_state = [&](int l) -> cplx<T> {
    const bool keep = active && krk && kept_l(g, l);
    const uint32_t ix = active ? boff32(kr, l, NY) : 0u;
    const cplx<T> s0 = at(a.s0, ix);
    cplx<T> v = s0;
    if (a.stage != 1) {
      const T E2 = at(a.E2, ix), Z = at(a.Z, ix);
      const cplx<T> nb = at(a.stage == 3 ? a.N2 : a.N1, ix);
      v = rmul(E2, s0) + rmul(Z, nb);                        // s1 (stage 2, 4) or s2 (stage 3)
      if (a.stage == 4) {
        const cplx<T> n3 = at(a.N3, ix);
        v = rmul(E2, v) + rmul(Z, mk((T)2 * n3.re - nb.re, (T)2 * n3.im - nb.im));   // s3
      }
    }
    return keep ? v : czero<T>();
  };
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    if (a.fsplit && f != (int)blockIdx.y) continue;   // one field per workgroup row (uniform)
    cplx<T> w[1][P];
    // the field's reads first, all of them, then their use: read where used, each read's branch
    // on the dealias mask made the compiler wait for it before issuing the next
    cplx<T> vs[P];
    T cs[P];
#pragma unroll
    for (int q = 0; q < P; ++q) vs[q] = f < 3 ? stage_state(t + TPT * q) : masked(a.x, t + TPT * q);
    if (f >= 3 && a.lamc != nullptr) {
#pragma unroll
      for (int q = 0; q < P; ++q) cs[q] = at(a.lamc, active ? boff32(kr, t + TPT * q, NY) : 0u);
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int l = t + TPT * q;
      const T lv = lwave(g, l);
      cplx<T> r;
      if (f < 3) {
        const cplx<T> sv = vs[q];
        const T iK = (kr == 0 && l == 0) ? (T)0 : (T)1 / (kv * kv + lv * lv);
        if (f == 0) { const T aa = lv * iK; r = mk(-aa * sv.im, aa * sv.re); }   // u: i l s/|k|^2
        else if (f == 1) { const T bb = kv * iK; r = mk(bb * sv.im, -bb * sv.re); }  // v: -i kr s/|k|^2
        else r = sv;                                                                // zeta
      } else {
        cplx<T> lam = vs[q];
        if (a.lamc != nullptr && active && krk && kept_l(g, l)) lam = rmul(cs[q], lam);
        const T fac = (f == 3 ? kv : lv) * wi;                                      // i kr/w, i l/w
        r = mk(-fac * lam.im, fac * lam.re);
      }
      w[0][q] = mk(r.re * nrm, -(r.im * nrm));
    }
    fft_forward<T, NY, P, 1>(w, L, t, tw);
    if (active) {
      cplx<T>* o = a.out + (long)f * a.fstride;
#pragma unroll
      for (int q = 0; q < P; ++q) at(o, boff32(kr, t + TPT * q, NY)) = conj(w[0][q]);
    }
  }
}

template <typename T, int NY, int P, int G>
__global__ __launch_bounds__(2 * G * (NY / P)) __attribute__((amdgpu_waves_per_eu(GF_K1_WAVES)))
void ka1_kernel(const AdjK1Args<T> a) {
  __shared__ cplx<T> lds[G * PairLds<T, NY>::value];
  __shared__ cplx<T> tw[TwPlan<NY, P>::size()];
  load_twiddles<T, NY, P>(tw, a.tw);
  __syncthreads();
  ka1_body<T, NY, P, G>(a, blockIdx.x, lds, tw);
}

// ---------------------------------------------------------------------------
// KA2: mixed u, v, zeta, pA, pB -> x-spectra of zeta pA, zeta pB, u pA + v pB
// ---------------------------------------------------------------------------
template <typename T, int NX, int P, int G>
__device__ __forceinline__ void ka2_body(const AdjK2Args<T>& a, int blk, cplx<T>* __restrict__ lds,
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
  // inverse x-transform of the two rows of mixed field f: row y0 = Re, row y0 + 1 = -Im
  auto inverse_rows = [&](int f, cplx<T> (&w)[1][P]) {
    const cplx<T>* fp = a.in + (long)f * a.fstride;
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int k = t + TPT * q;
      const cplx<T> A = herm_load<T, NX>(fp, k, y0, ny, a.kin, active);
      const cplx<T> B = herm_load<T, NX>(fp, k, y0 + 1, ny, a.kin, active);
      w[0][q] = mk(A.re - B.im, -(A.im + B.re));
    }
    fft_forward<T, NX, P, 1>(w, L, t, tw);
  };
  // forward x-transform of two real rows packed as r0 + i r1, split, stored as flux field f
  auto forward_rows = [&](int f, cplx<T> (&w)[1][P]) {
    fft_forward<T, NX, P, 1>(w, L, t, tw);
    const MirrorSlots<NX, TPT, XPad<T>::G> ms(t);
#pragma unroll
    for (int q = 0; q < P; ++q) L[ms.k(q)] = w[0][q];
    __syncthreads();
    if (active) {
      cplx<T>* o = a.out + (long)f * a.ostride;
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
  // with hsplit, workgroup row 0 forms zeta pA, row 1 zeta pB and row 2 u pA + v pB, each
  // transforming the fields it needs (uniform branches); otherwise one workgroup forms all three
  const int part = a.hsplit ? (int)blockIdx.y : 3;
  const bool all = part == 3;
  T z0[P], z1[P], pa0[P], pa1[P], pb0[P], pb1[P];
  if (all || part == 0 || part == 1) {
    cplx<T> w[1][P];
    inverse_rows(2, w);                                 // zeta
#pragma unroll
    for (int q = 0; q < P; ++q) { z0[q] = w[0][q].re; z1[q] = -w[0][q].im; }
  }
  if (all || part == 0 || part == 2) {
    cplx<T> w[1][P];
    inverse_rows(3, w);                                 // pA
#pragma unroll
    for (int q = 0; q < P; ++q) { pa0[q] = w[0][q].re; pa1[q] = -w[0][q].im; }
    if (all || part == 0) {
#pragma unroll
      for (int q = 0; q < P; ++q) w[0][q] = mk(z0[q] * pa0[q], z1[q] * pa1[q]);   // zeta pA, both rows
      forward_rows(0, w);
    }
  }
  if (all || part == 1 || part == 2) {
    cplx<T> w[1][P];
    inverse_rows(4, w);                                 // pB
#pragma unroll
    for (int q = 0; q < P; ++q) { pb0[q] = w[0][q].re; pb1[q] = -w[0][q].im; }
    if (all || part == 1) {
#pragma unroll
      for (int q = 0; q < P; ++q) w[0][q] = mk(z0[q] * pb0[q], z1[q] * pb1[q]);   // zeta pB
      forward_rows(1, w);
    }
  }
  if (all || part == 2) {
    T c0[P], c1[P];
    {
      cplx<T> w[1][P];
      inverse_rows(0, w);                               // u
#pragma unroll
      for (int q = 0; q < P; ++q) { c0[q] = w[0][q].re * pa0[q]; c1[q] = -w[0][q].im * pa1[q]; }
    }
    {
      cplx<T> w[1][P];
      inverse_rows(1, w);                               // v
#pragma unroll
      for (int q = 0; q < P; ++q) w[0][q] = mk(c0[q] + w[0][q].re * pb0[q], c1[q] + (-w[0][q].im) * pb1[q]);
      forward_rows(2, w);                               // u pA + v pB
    }
  }
}

template <typename T, int NX, int P, int G>
__global__ __launch_bounds__(G * (NX / P)) __attribute__((amdgpu_waves_per_eu(GF_K2_WAVES)))
void ka2_kernel(const AdjK2Args<T> a) {
  __shared__ cplx<T> lds[G * LdsN<NX>::value];
  __shared__ cplx<T> tw[TwPlan<NX, P>::size()];
  load_twiddles<T, NX, P>(tw, a.tw);
  __syncthreads();
  ka2_body<T, NX, P, G>(a, xcd_block(blockIdx.x, gridDim.x), lds, tw);
}

// ---------------------------------------------------------------------------
// KA3: y-forward of the three x-spectra, lambda_s, adjoint stage arithmetic
// ---------------------------------------------------------------------------
template <typename T, int NY, int P, int G>
__device__ __forceinline__ void ka3_body(const AdjK3Args<T>& a, int blk, cplx<T>* __restrict__ lds,
                                         const cplx<T>* __restrict__ tw) {
  constexpr int TPT = NY / P;
  const YMap<TPT> m(threadIdx.x);
  const int t = m.t;
  const int kr = 2 * (G == 1 ? pair_of_block(blk) : blk * G + m.slot) + m.c;
  const Geo<T>& g = a.g;
  const bool active = kr < a.ncols && kr < g.kmax;
  cplx<T>* L = col_buffer<T, NY>(lds, m.slot, m.c);
  const T kv = krwave(g, kr);
  const T wk = herm_w(g, kr);
  // acc = F(u pA + v pB) - i l/|k|^2 F(zeta pA) + i kr/|k|^2 F(zeta pB), built one transform at a time
  cplx<T> acc[P];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    cplx<T> w[1][P];
    const cplx<T>* fp = a.in + (long)f * a.fstride;
#pragma unroll
    for (int q = 0; q < P; ++q) w[0][q] = active ? at(fp, boff32(kr, t + TPT * q, NY)) : czero<T>();
    fft_forward<T, NY, P, 1>(w, L, t, tw);
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int l = t + TPT * q;
      const T lv = lwave(g, l);
      const T iK = (kr == 0 && l == 0) ? (T)0 : (T)1 / (kv * kv + lv * lv);
      const cplx<T> v = w[0][q];
      if (f == 0) { const T c = lv * iK; acc[q] = mk(c * v.im, -c * v.re); }                // -i l/|k|^2
      else if (f == 1) { const T c = kv * iK; acc[q] = acc[q] + mk(-c * v.im, c * v.re); }  // +i kr/|k|^2
      else acc[q] = v + acc[q];
    }
  }
  T nu = 0;
  if (active) {
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int l = t + TPT * q;
      if (!kept_l(g, l)) continue;
      const uint32_t ix = boff32(kr, l, NY);
      const cplx<T> tv = rmul(wk, acc[q]);             // lambda of the stage's calcN input
      const T E2 = at(a.E2, ix), Z = at(a.Z, ix);
      auto re_dot = [](cplx<T> x, cplx<T> y) { return x.re * y.re + x.im * y.im; };  // Re(conj(x) y)
      if (a.stage == 4) {
        const cplx<T> lam = at(a.lam, ix);
        const T E = at(a.E, ix), Al = at(a.Al, ix), Be = at(a.Be, ix);
        at(a.l1, ix) = rmul(E2, tv);
        at(a.lN3, ix) = rmul((T)2 * Be, lam) + rmul((T)2 * Z, tv);
        at(a.lN1, ix) = rmul(Al, lam) - rmul(Z, tv);
        (void)E;
        if (a.nu_part != nullptr) {
          const cplx<T> s0 = at(a.s0, ix);
          const cplx<T> N1 = at(a.N1, ix), N2 = at(a.N2, ix), N3 = at(a.N3, ix), N4 = at(a.N4, ix);
          const cplx<T> s1 = rmul(E2, s0) + rmul(Z, N1);   // stage 2's input, as the forward formed it
          const cplx<T> d1 = rmul(at(a.dE, ix), s0) + rmul(at(a.dA, ix), N1) +
                             rmul((T)2 * at(a.dB, ix), N2 + N3) + rmul(at(a.dG, ix), N4);
          const cplx<T> d3 = rmul(at(a.dE2, ix), s1) + rmul(at(a.dZ, ix), mk((T)2 * N3.re - N1.re, (T)2 * N3.im - N1.im));
          nu += re_dot(lam, d1) + re_dot(tv, d3);
        }
      } else if (a.stage == 3) {
        const cplx<T> lam = at(a.lam, ix);
        const T E = at(a.E, ix), Be = at(a.Be, ix);
        at(a.l0, ix) = rmul(E, lam) + rmul(E2, tv);
        at(a.lN2, ix) = rmul((T)2 * Be, lam) + rmul(Z, tv);
        if (a.nu_part != nullptr) {
          const cplx<T> d2 = rmul(at(a.dE2, ix), at(a.s0, ix)) + rmul(at(a.dZ, ix), at(a.N2, ix));
          nu += re_dot(tv, d2);
        }
      } else if (a.stage == 2) {
        const cplx<T> l1 = at(a.l1, ix) + tv;
        at(a.l0, ix) = at(a.l0, ix) + rmul(E2, l1);
        at(a.lN1, ix) = at(a.lN1, ix) + rmul(Z, l1);
        if (a.nu_part != nullptr) {
          const cplx<T> d1 = rmul(at(a.dE2, ix), at(a.s0, ix)) + rmul(at(a.dZ, ix), at(a.N1, ix));
          nu += re_dot(l1, d1);
        }
      } else {
        at(a.lam, ix) = at(a.l0, ix) + tv;
      }
    }
  }
  if (a.nu_part != nullptr && a.stage >= 2) {
    // workgroup sum of the nu terms; each workgroup owns one slot, so the total is
    // summed in a fixed order and no atomics are needed
    __syncthreads();
    T* red = reinterpret_cast<T*>(lds);
    red[threadIdx.x] = nu;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.nu_part[blockIdx.x] += red[0];
  }
}

template <typename T, int NY, int P, int G>
__global__ __launch_bounds__(2 * G * (NY / P)) __attribute__((amdgpu_waves_per_eu(2)))
void ka3_kernel(const AdjK3Args<T> a) {
  __shared__ cplx<T> lds[G * PairLds<T, NY>::value];
  __shared__ cplx<T> tw[TwPlan<NY, P>::size()];
  load_twiddles<T, NY, P>(tw, a.tw);
  __syncthreads();
  ka3_body<T, NY, P, G>(a, blockIdx.x, lds, tw);
}

template <typename T, int N>
void adj_n(int k, const void* args, hipStream_t s) {
  constexpr int P = PY<N>::P;
  if (k == 2) {
    const AdjK2Args<T>& a = *(const AdjK2Args<T>*)args;
    constexpr int G = GW<N, P>::G;
    hipLaunchKernelGGL((ka2_kernel<T, N, P, G>), dim3((a.g.ny / 2 + G - 1) / G, a.hsplit ? 3 : 1), dim3(G * (N / P)), 0, s, a);
    return;
  }
  constexpr int G = GY<N, P>::G;
  if (k == 1) {
    const AdjK1Args<T>& a = *(const AdjK1Args<T>*)args;
    const int nblk = pair_blocks(a.ncols, G);
    if (nblk > 0) hipLaunchKernelGGL((ka1_kernel<T, N, P, G>), dim3(nblk, a.fsplit ? 5 : 1), dim3(2 * G * (N / P)), 0, s, a);
  } else {
    const AdjK3Args<T>& a = *(const AdjK3Args<T>*)args;
    const int nblk = pair_blocks(a.ncols, G);
    if (nblk > 0) hipLaunchKernelGGL((ka3_kernel<T, N, P, G>), dim3(nblk), dim3(2 * G * (N / P)), 0, s, a);
  }
}

template <typename T>
int adj_blocks(int n, int ncols) {
  int G = 1;
  switch (n) {
    case 32: G = GY<32, PY<32>::P>::G; break;
    case 64: G = GY<64, PY<64>::P>::G; break;
    case 128: G = GY<128, PY<128>::P>::G; break;
    case 256: G = GY<256, PY<256>::P>::G; break;
    case 512: G = GY<512, PY<512>::P>::G; break;
    case 1024: G = GY<1024, PY<1024>::P>::G; break;
    default: G = GY<2048, PY<2048>::P>::G; break;
  }
  return pair_blocks(ncols, G);
}

template <typename T>
void launch_adj(int k, int n, const void* args, hipStream_t s) {
  switch (n) {
    case 32: adj_n<T, 32>(k, args, s); break;
    case 64: adj_n<T, 64>(k, args, s); break;
    case 128: adj_n<T, 128>(k, args, s); break;
    case 256: adj_n<T, 256>(k, args, s); break;
    case 512: adj_n<T, 512>(k, args, s); break;
    case 1024: adj_n<T, 1024>(k, args, s); break;
    case 2048: adj_n<T, 2048>(k, args, s); break;
    default: break;
  }
}

template void launch_adj<double>(int, int, const void*, hipStream_t);
template int adj_blocks<double>(int, int);

}  // namespace gf
