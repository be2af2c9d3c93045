// Instantiation unit: K3 for float (sizes 16..2048, 1..4 layers).
#include "gf_kernels_impl.hpp"
namespace gf {
template <int NL>
static void k3_nl_f(int n, const K3Args<float>& a, hipStream_t s) {
  switch (n) {
    case 16: k3_n<float, 16, NL>(a, s); break;
    case 32: k3_n<float, 32, NL>(a, s); break;
    case 64: k3_n<float, 64, NL>(a, s); break;
    case 128: k3_n<float, 128, NL>(a, s); break;
    case 256: k3_n<float, 256, NL>(a, s); break;
    case 512: k3_n<float, 512, NL>(a, s); break;
    case 1024: k3_n<float, 1024, NL>(a, s); break;
    case 2048: k3_n<float, 2048, NL>(a, s); break;
    default: break;
  }
}
template <> void launch_k3<float>(int n, const K3Args<float>& a, hipStream_t s) {
  switch (a.model == M_MLQG ? a.g.nlayers : 1) {
    case 1: k3_nl_f<1>(n, a, s); break;
    case 2: k3_nl_f<2>(n, a, s); break;
    case 3: k3_nl_f<3>(n, a, s); break;
    case 4: k3_nl_f<4>(n, a, s); break;
    default: break;
  }
}
}  // namespace gf
