// Instantiation unit: K1 for double, MultiLayerQG with 5..8 layers (sizes 16..2048).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k1_deep_d(int nl, int n, const K1Args<double>& a, hipStream_t s) {
  switch (nl) {
    case 5: GF_SIZES(k1_n, double, 5, n, a, s); break;
    case 6: GF_SIZES(k1_n, double, 6, n, a, s); break;
    case 7: GF_SIZES(k1_n, double, 7, n, a, s); break;
    case 8: GF_SIZES(k1_n, double, 8, n, a, s); break;
    default: break;
  }
}
}  // namespace gf
