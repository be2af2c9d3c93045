// Instantiation unit: K2 for double (sizes 16..2048, 1..4 layers; 5..8 layers in gf_k2_dm.hip).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k2_deep_d(int nl, int n, const K2Args<double>& a, hipStream_t s);
// Inverse-out and forward-in treat every layer as one more field (NL = 1).
template <> void launch_k2<double>(int n, const K2Args<double>& a, hipStream_t s) {
  const int nl = (a.mode == K2_ADVECT || a.mode == K2_ADVECT_FULL) ? a.g.nlayers : 1;
  switch (nl) {
    case 1: GF_SIZES(k2_n, double, 1, n, a, s); break;
    case 2: GF_SIZES(k2_n, double, 2, n, a, s); break;
    case 3: GF_SIZES(k2_n, double, 3, n, a, s); break;
    case 4: GF_SIZES(k2_n, double, 4, n, a, s); break;
    default: k2_deep_d(nl, n, a, s); break;
  }
}
}  // namespace gf
