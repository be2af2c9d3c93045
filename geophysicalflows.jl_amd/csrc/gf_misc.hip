// Layout conversion (Julia column-major <-> column-blocked) and in-place dealias.
#include "gf_kernels_impl.hpp"

namespace gf {

template <typename T>
__global__ void to_julia_kernel(const Geo<T> g, const cplx<T>* src, cplx<T>* dst, int nl) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;   // Julia linear index
  const long per = (long)g.nkr * g.ny;
  if (i >= per * nl) return;
  const int layer = (int)(i / per);
  const long r = i - layer * per;
  const int kr = (int)(r % g.nkr), l = (int)(r / g.nkr);
  dst[i] = src[layer * g.lstride + boff(kr, l, g.ny)];
}
template <typename T>
__global__ void from_julia_kernel(const Geo<T> g, const cplx<T>* src, cplx<T>* dst, int nl) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)g.nkr * g.ny;
  if (i >= per * nl) return;
  const int layer = (int)(i / per);
  const long r = i - layer * per;
  const int kr = (int)(r % g.nkr), l = (int)(r / g.nkr);
  dst[layer * g.lstride + boff(kr, l, g.ny)] = src[i];
}
// dealias! in place (src/twodnavierstokes.jl:287 and updatevars!, 334)
template <typename T>
__global__ void mask_kernel(const Geo<T> g, cplx<T>* sol, int nl) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)g.nkr * g.ny;
  if (i >= per * nl) return;
  const int layer = (int)(i / per);
  const long r = i - layer * per;
  const int kr = (int)(r % g.nkr), l = (int)(r / g.nkr);
  if (kr >= g.kmax || !kept_l(g, l)) sol[layer * g.lstride + boff(kr, l, g.ny)] = czero<T>();
}

template <typename T> void launch_to_julia(const Geo<T>& g, const cplx<T>* src, cplx<T>* dst, int nl, hipStream_t s) {
  const long n = (long)g.nkr * g.ny * nl;
  hipLaunchKernelGGL(to_julia_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, src, dst, nl);
}
template <typename T> void launch_from_julia(const Geo<T>& g, const cplx<T>* src, cplx<T>* dst, int nl, hipStream_t s) {
  const long n = (long)g.nkr * g.ny * nl;
  hipLaunchKernelGGL(from_julia_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, src, dst, nl);
}
template <typename T> void launch_mask(const Geo<T>& g, cplx<T>* sol, int nl, hipStream_t s) {
  const long n = (long)g.nkr * g.ny * nl;
  hipLaunchKernelGGL(mask_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, sol, nl);
}

// Forcing step counter of a captured step: set once before the replays (outside the graph),
// advanced by one at the end of each replayed step (inside it). One lane stores (a vector store).
__global__ void counter_kernel(int64_t* c, int64_t v, int add) {
  if (threadIdx.x == 0) c[0] = add ? c[0] + v : v;
}
void launch_counter(int64_t* c, int64_t v, int add, hipStream_t s) {
  hipLaunchKernelGGL(counter_kernel, dim3(1), dim3(64), 0, s, c, v, add);
}

bool size_supported(int n) {
  return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096;
}

#define INST(T)                                                                                  \
  template void launch_to_julia<T>(const Geo<T>&, const cplx<T>*, cplx<T>*, int, hipStream_t);   \
  template void launch_from_julia<T>(const Geo<T>&, const cplx<T>*, cplx<T>*, int, hipStream_t); \
  template void launch_mask<T>(const Geo<T>&, cplx<T>*, int, hipStream_t);
INST(double)
INST(float)

}  // namespace gf
