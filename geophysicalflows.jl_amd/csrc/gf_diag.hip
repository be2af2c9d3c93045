// Domain-averaged diagnostics as device reductions over the blocked spectral state
// (src/twodnavierstokes.jl:374-535, src/singlelayerqg.jl:546-715, src/multilayerqg.jl:918-951).
//
// Every diagnostic is a Parseval sum over the rfft half plane ([FF] parsevalsum: the kr = 0
// column once, every other column twice), times Lx Ly/(nx ny)^2 and the diagnostic's own
// factor.  Each workgroup sums its modes and writes one partial per output; the host adds
// the partials in a fixed order, so the result does not depend on scheduling.
#include "gf_kernels_impl.hpp"

namespace gf {

constexpr int DIAG_THREADS = 256;

template <typename T>
__device__ __forceinline__ double rdot(cplx<T> a, cplx<T> b) {   // Re(a conj(b))
  return (double)a.re * (double)b.re + (double)a.im * (double)b.im;
}
template <typename T> __device__ __forceinline__ double abs2(cplx<T> a) { return rdot(a, a); }

template <typename T>
__global__ __launch_bounds__(DIAG_THREADS) void diag_kernel(const DiagArgs<T> a, double* __restrict__ part) {
  const Geo<T>& g = a.g;
  const int ny = g.ny;
  const long fsz = g.lstride;
  double acc[DIAG_MAXOUT];
#pragma unroll
  for (int k = 0; k < DIAG_MAXOUT; ++k) acc[k] = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < fsz; i += (long)gridDim.x * blockDim.x) {
    const int kr = BW * (int)(i / (BW * ny)) + (int)(i % BW);
    const int l = (int)((i / BW) % ny);
    if (kr >= g.nkr) continue;
    const double w = kr == 0 ? 1.0 : 2.0;
    const double kv = (double)krwave(g, kr), lv = (double)lwave(g, l);
    const double K2 = kv * kv + lv * lv;
    const bool k0 = kr == 0 && l == 0;
    const double iK2 = k0 ? 0.0 : 1.0 / K2;
    const uint32_t ix = (uint32_t)i;
    if (a.model == M_MLQG) {
      // psi_j = sum_k S^-1[j,k] q_k; KE_j ~ H_j K^2 |psi_j|^2, PE_{j+1/2} ~ f0^2/g'_j |psi_j - psi_{j+1}|^2
      const int n = a.nl;
      cplx<T> ps[MAXL];
      for (int j = 0; j < n; ++j) {
        cplx<T> s = czero<T>();
        const T* m = a.sinv + (size_t)i * (n * n) + j * n;
        for (int k = 0; k < n; ++k) s = s + rmul(m[k], at(a.sol, (uint32_t)(k * fsz) + ix));
        ps[j] = s;
      }
      for (int j = 0; j < n; ++j) acc[j] += w * a.lay_ke[j] * K2 * abs2(ps[j]);
      for (int j = 0; j + 1 < n; ++j) acc[n + j] += w * a.lay_pe[j] * abs2(ps[j] - ps[j + 1]);
      continue;
    }
    const cplx<T> s = at(a.sol, ix);
    const double s2 = abs2(s);
    double v = 0;
    switch (a.kind) {
      case DIAG_ENERGY:
      case DIAG_KINETIC_ENERGY:
      case DIAG_POTENTIAL_ENERGY: {
        if (a.model == M_TWODNS) { v = 0.5 * iK2 * s2; break; }
        // SLQG psih = -q/(K^2 + 1/ell^2) or -q/K^2 (src/singlelayerqg.jl:484-490)
        const double den = a.eb ? K2 + (double)a.ell2inv : K2;
        const double p2 = (k0 && !a.eb) ? 0.0 : s2 / (den * den);
        const double ke = 0.5 * K2 * p2, pe = 0.5 * (double)a.ell2inv * p2;
        v = a.kind == DIAG_KINETIC_ENERGY ? ke : a.kind == DIAG_POTENTIAL_ENERGY ? pe : ke + pe;
      } break;
      case DIAG_ENSTROPHY: {
        if (a.model == M_SLQG && a.etah != nullptr) v = 0.5 * abs2(s + at(a.etah, ix));   // |q + eta|^2
        else v = 0.5 * s2;
      } break;
      case DIAG_PALINSTROPHY: v = 0.5 * K2 * s2; break;
      case DIAG_ENERGY_DISSIPATION:
      case DIAG_ENERGY_DRAG:
      case DIAG_ENSTROPHY_DISSIPATION:
      case DIAG_ENSTROPHY_DRAG: {
        if (k0) break;
        const bool en = a.kind == DIAG_ENERGY_DISSIPATION || a.kind == DIAG_ENERGY_DRAG;
        if (a.model == M_TWODNS) {
          // -xi Krsq^(n - 1) |zeta|^2 or -xi Krsq^n |zeta|^2 (src/twodnavierstokes.jl:425-467)
          const int e = en ? a.nxi - 1 : a.nxi;
          double p = 1;
          for (int k = 0; k < (e < 0 ? -e : e); ++k) p *= K2;
          v = -(double)a.xi * (e < 0 ? 1.0 / p : p) * s2;
        } else {
          // SLQG: nu Krsq^(nnu - 1) |q|^2, nu Krsq^nnu |q|^2, mu invKrsq |q|^2, mu |q|^2
          // (src/singlelayerqg.jl:617-715)
          if (a.kind == DIAG_ENERGY_DRAG) v = (double)a.xi * iK2 * s2;
          else if (a.kind == DIAG_ENSTROPHY_DRAG) v = (double)a.xi * s2;
          else {
            const int e = a.kind == DIAG_ENERGY_DISSIPATION ? a.nxi - 1 : a.nxi;
            double p = 1;
            for (int k = 0; k < (e < 0 ? -e : e); ++k) p *= K2;
            v = (double)a.xi * (e < 0 ? 1.0 / p : p) * s2;
          }
        }
      } break;
      case DIAG_ENERGY_WORK:
      case DIAG_ENSTROPHY_WORK: {
        // invKrsq sol conj(Fh) or sol conj(Fh); with stochastic forcing (prevsol + sol)/2
        // and Fh the step's realisation (src/twodnavierstokes.jl:494-535)
        cplx<T> F = czero<T>();
        if (a.Fh != nullptr) F = rmul(a.fscale, at(a.Fh, ix));
        if (a.sqrt_spec != nullptr && kr <= a.fkx && (l <= a.fky || l >= ny - a.fky)) {
          const T sp = at(a.sqrt_spec, ix);
          if (sp != (T)0) {
            const cplx<T> xi = complex_normal<T>(a.seed, a.fstep, (uint32_t)(l * g.nkr + kr), (uint32_t)a.instance);
            F = mk((xi.re * sp) / a.sqrt_dt, (xi.im * sp) / a.sqrt_dt);
          }
        }
        cplx<T> x = s;
        if (a.prevsol != nullptr) {
          const cplx<T> p = at(a.prevsol, ix);
          x = mk((p.re + s.re) / 2, (p.im + s.im) / 2);
        }
        v = (a.kind == DIAG_ENERGY_WORK ? iK2 : 1.0) * rdot(x, F);
      } break;
      default: break;
    }
    acc[0] += w * v;
  }
  __shared__ double red[DIAG_THREADS];
  for (int k = 0; k < a.nout; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = DIAG_THREADS / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.x * a.nout + k] = red[0];
    __syncthreads();
  }
}

template <typename T>
void launch_diag(const DiagArgs<T>& a, double* part, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(diag_kernel<T>, dim3(nblocks), dim3(DIAG_THREADS), 0, s, a, part);
}

// Adds the per-block partials in block order and scales them, exactly as the host does for
// gf_diagnostic, so a value recorded on the device equals the synchronous one bit for bit.
__global__ __launch_bounds__(64) void diag_finish_kernel(const double* __restrict__ part, int nblocks, int nout,
                                                         double scale, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= nout) return;
  double s = 0;
  for (int b = 0; b < nblocks; ++b) s += part[(long)b * nout + k];
  out[k] = s * scale;
}

void launch_diag_finish(const double* part, int nblocks, int nout, double scale, double* out, hipStream_t s) {
  hipLaunchKernelGGL(diag_finish_kernel, dim3(1), dim3(64), 0, s, part, nblocks, nout, scale, out);
}

// ---- MultiLayerQG fluxes (src/multilayerqg.jl:977-1022) ----

// Spectra of psi, v and du/dy per layer from the dealiased state (fluxes calls updatevars!,
// which dealiases sol): psih = S^-1 qh (one layer: -invKrsq qh), vh = i kr psih, uh = -i l psih,
// (du/dy)h = i l uh.
template <typename T>
__global__ __launch_bounds__(DIAG_THREADS) void flux_fields_kernel(const FluxArgs<T> a) {
  const Geo<T>& g = a.g;
  const int ny = g.ny, n = a.nl;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.fstride) return;
  const int kr = BW * (int)(i / (BW * ny)) + (int)(i % BW);
  const int l = (int)((i / BW) % ny);
  if (kr >= g.nkr) return;
  const T kv = krwave(g, kr), lv = lwave(g, l);
  const T K2 = kv * kv + lv * lv;
  const T iK = (kr == 0 && l == 0) ? (T)0 : (T)1 / K2;
  for (int j = 0; j < n; ++j) {
    cplx<T> ps = czero<T>();
    if (a.sinv == nullptr) {
      ps = rmul(-iK, a.sol[i]);
    } else {
      const T* m = a.sinv + i * (n * n) + j * n;
      for (int k = 0; k < n; ++k) ps = ps + rmul(m[k], a.sol[k * a.fstride + i]);
    }
    const cplx<T> vh = mk(-kv * ps.im, kv * ps.re);
    const cplx<T> uh = mk(lv * ps.im, -lv * ps.re);
    const cplx<T> uyh = mk(-lv * uh.im, lv * uh.re);
    a.f[0][(long)j * a.fstride + i] = ps;
    a.f[1][(long)j * a.fstride + i] = vh;
    a.f[2][(long)j * a.fstride + i] = uyh;
  }
}

// x sums by Parseval along x: for real rows f, g with mixed spectra (y transformed, x not)
// m_f(kr, y), sum_x f g = nx sum_kr w_kr Re(m_f conj(m_g)), w = 2 for 0 < kr < nx/2; the kr = 0 and
// kr = nx/2 columns enter through their real parts only, as the c2r transform reads them.
template <typename T>
__global__ __launch_bounds__(DIAG_THREADS) void flux_kernel(const FluxArgs<T> a, double* __restrict__ part) {
  const Geo<T>& g = a.g;
  const int ny = g.ny, n = a.nl, nout = 2 * n - 1;
  double acc[DIAG_MAXOUT];
#pragma unroll
  for (int k = 0; k < DIAG_MAXOUT; ++k) acc[k] = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.fstride; i += (long)gridDim.x * blockDim.x) {
    const int kr = BW * (int)(i / (BW * ny)) + (int)(i % BW);
    const int y = (int)((i / BW) % ny);
    if (kr >= g.nkr) continue;
    const bool edge = kr == 0 || 2 * kr == g.nx;
    const double w = edge ? 1.0 : 2.0;
    auto prod = [&](int fa, int ja, int fb, int jb) -> double {
      const cplx<T> p = a.f[fa][(long)ja * a.fstride + i], q = a.f[fb][(long)jb * a.fstride + i];
      return edge ? (double)p.re * (double)q.re : rdot(p, q);
    };
    for (int j = 0; j < n; ++j)   // lateral_j: U_j v_j du_j/dy
      acc[j] += w * a.fac[j] * (double)a.Uy[j * ny + y] * prod(1, j, 2, j);
    for (int j = 0; j + 1 < n; ++j)   // vertical_{j+1/2}: (U_j - U_{j+1}) v_{j+1} psi_j
      acc[n + j] += w * a.fac[n + j] * ((double)a.Uy[j * ny + y] - (double)a.Uy[(j + 1) * ny + y]) *
                    prod(1, j + 1, 0, j);
  }
  __shared__ double red[DIAG_THREADS];
  for (int k = 0; k < nout; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = DIAG_THREADS / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.x * nout + k] = red[0];
    __syncthreads();
  }
}

template <typename T>
void launch_flux_fields(const FluxArgs<T>& a, hipStream_t s) {
  hipLaunchKernelGGL(flux_fields_kernel<T>, dim3((unsigned)((a.fstride + DIAG_THREADS - 1) / DIAG_THREADS)),
                     dim3(DIAG_THREADS), 0, s, a);
}
template <typename T>
void launch_flux(const FluxArgs<T>& a, double* part, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(flux_kernel<T>, dim3(nblocks), dim3(DIAG_THREADS), 0, s, a, part);
}

// ---- pvfromstreamfunction! (src/multilayerqg.jl:520-552) ----
// q_j = sum_k S[j, k] psi_k, S = -Krsq I + F, summed in k order as the static-matrix product does;
// one layer: q = -Krsq psi.
template <typename T>
__global__ __launch_bounds__(DIAG_THREADS) void pv_from_psi_kernel(const PvArgs<T> a) {
  const Geo<T>& g = a.g;
  const int ny = g.ny, n = a.nl;
  const long fs = g.lstride;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fs) return;
  const int kr = BW * (int)(i / (BW * ny)) + (int)(i % BW);
  const int l = (int)((i / BW) % ny);
  if (kr >= g.nkr) return;
  const T kv = krwave(g, kr), lv = lwave(g, l);
  const T K2 = kv * kv + lv * lv;
  cplx<T> ps[MAXL];
  for (int k = 0; k < n; ++k) ps[k] = a.psi[k * fs + i];
  for (int j = 0; j < n; ++j) {
    cplx<T> q = czero<T>();
    for (int k = 0; k < n; ++k) {
      const T s = k == j ? -K2 + a.F[j * MAXL + k] : a.F[j * MAXL + k];
      q = q + rmul(s, ps[k]);
    }
    a.q[j * fs + i] = q;
  }
}
template <typename T>
void launch_pv_from_psi(const PvArgs<T>& a, hipStream_t s) {
  hipLaunchKernelGGL(pv_from_psi_kernel<T>, dim3((unsigned)((a.g.lstride + DIAG_THREADS - 1) / DIAG_THREADS)),
                     dim3(DIAG_THREADS), 0, s, a);
}

template void launch_diag<double>(const DiagArgs<double>&, double*, int, hipStream_t);
template void launch_diag<float>(const DiagArgs<float>&, double*, int, hipStream_t);
template void launch_flux_fields<double>(const FluxArgs<double>&, hipStream_t);
template void launch_flux_fields<float>(const FluxArgs<float>&, hipStream_t);
template void launch_flux<double>(const FluxArgs<double>&, double*, int, hipStream_t);
template void launch_flux<float>(const FluxArgs<float>&, double*, int, hipStream_t);
template void launch_pv_from_psi<double>(const PvArgs<double>&, hipStream_t);
template void launch_pv_from_psi<float>(const PvArgs<float>&, hipStream_t);

}  // namespace gf
