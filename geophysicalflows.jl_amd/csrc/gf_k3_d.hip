// Instantiation unit: K3 for double (sizes 16..2048, 1..4 layers).
#include "gf_kernels_impl.hpp"
namespace gf {
template <int NL>
static void k3_nl_d(int n, const K3Args<double>& a, hipStream_t s) {
  switch (n) {
    case 16: k3_n<double, 16, NL>(a, s); break;
    case 32: k3_n<double, 32, NL>(a, s); break;
    case 64: k3_n<double, 64, NL>(a, s); break;
    case 128: k3_n<double, 128, NL>(a, s); break;
    case 256: k3_n<double, 256, NL>(a, s); break;
    case 512: k3_n<double, 512, NL>(a, s); break;
    case 1024: k3_n<double, 1024, NL>(a, s); break;
    case 2048: k3_n<double, 2048, NL>(a, s); break;
    default: break;
  }
}
template <> void launch_k3<double>(int n, const K3Args<double>& a, hipStream_t s) {
  switch (a.model == M_MLQG ? a.g.nlayers : 1) {
    case 1: k3_nl_d<1>(n, a, s); break;
    case 2: k3_nl_d<2>(n, a, s); break;
    case 3: k3_nl_d<3>(n, a, s); break;
    case 4: k3_nl_d<4>(n, a, s); break;
    default: break;
  }
}
}  // namespace gf
