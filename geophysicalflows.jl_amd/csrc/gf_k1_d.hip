// Instantiation unit: K1 for double (sizes 16..2048, 1..4 layers).
#include "gf_kernels_impl.hpp"
namespace gf {
template <int NL>
static void k1_nl_d(int n, const K1Args<double>& a, hipStream_t s) {
  switch (n) {
    case 16: k1_n<double, 16, NL>(a, s); break;
    case 32: k1_n<double, 32, NL>(a, s); break;
    case 64: k1_n<double, 64, NL>(a, s); break;
    case 128: k1_n<double, 128, NL>(a, s); break;
    case 256: k1_n<double, 256, NL>(a, s); break;
    case 512: k1_n<double, 512, NL>(a, s); break;
    case 1024: k1_n<double, 1024, NL>(a, s); break;
    case 2048: k1_n<double, 2048, NL>(a, s); break;
    default: break;
  }
}
template <> void launch_k1<double>(int n, const K1Args<double>& a, hipStream_t s) {
  switch (a.model == M_MLQG ? a.g.nlayers : 1) {
    case 1: k1_nl_d<1>(n, a, s); break;
    case 2: k1_nl_d<2>(n, a, s); break;
    case 3: k1_nl_d<3>(n, a, s); break;
    case 4: k1_nl_d<4>(n, a, s); break;
    default: break;
  }
}
}  // namespace gf
