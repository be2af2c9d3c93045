// Instantiation unit: K2 for float, MultiLayerQG with 5..8 layers (sizes 16..2048).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k2_deep_f(int nl, int n, const K2Args<float>& a, hipStream_t s) {
  switch (nl) {
    case 5: GF_SIZES(k2_n, float, 5, n, a, s); break;
    case 6: GF_SIZES(k2_n, float, 6, n, a, s); break;
    case 7: GF_SIZES(k2_n, float, 7, n, a, s); break;
    case 8: GF_SIZES(k2_n, float, 8, n, a, s); break;
    default: break;
  }
}
}  // namespace gf
