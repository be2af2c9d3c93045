// Instantiation unit: K1 for float, MultiLayerQG with 5..8 layers (sizes 16..2048).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k1_deep_f(int nl, int n, const K1Args<float>& a, hipStream_t s) {
  switch (nl) {
    case 5: GF_SIZES(k1_n, float, 5, n, a, s); break;
    case 6: GF_SIZES(k1_n, float, 6, n, a, s); break;
    case 7: GF_SIZES(k1_n, float, 7, n, a, s); break;
    case 8: GF_SIZES(k1_n, float, 8, n, a, s); break;
    default: break;
  }
}
}  // namespace gf
