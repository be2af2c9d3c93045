// Instantiation unit: K1 for double (sizes 16..2048, 1..4 layers; 5..8 layers in gf_k1_dm.hip).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k1_deep_d(int nl, int n, const K1Args<double>& a, hipStream_t s);
template <> void launch_k1<double>(int n, const K1Args<double>& a, hipStream_t s) {
  const int nl = a.model == M_MLQG ? a.g.nlayers : 1;
  switch (nl) {
    case 1: GF_SIZES(k1_n, double, 1, n, a, s); break;
    case 2: GF_SIZES(k1_n, double, 2, n, a, s); break;
    case 3: GF_SIZES(k1_n, double, 3, n, a, s); break;
    case 4: GF_SIZES(k1_n, double, 4, n, a, s); break;
    default: k1_deep_d(nl, n, a, s); break;
  }
}
}  // namespace gf
