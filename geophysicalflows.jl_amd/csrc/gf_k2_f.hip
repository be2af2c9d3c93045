// Instantiation unit: K2 for float (sizes 16..2048, 1..4 layers; 5..8 layers in gf_k2_fm.hip).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k2_deep_f(int nl, int n, const K2Args<float>& a, hipStream_t s);
// Inverse-out and forward-in treat every layer as one more field (NL = 1).
template <> void launch_k2<float>(int n, const K2Args<float>& a, hipStream_t s) {
  const int nl = (a.mode == K2_ADVECT || a.mode == K2_ADVECT_FULL) ? a.g.nlayers : 1;
  switch (nl) {
    case 1: GF_SIZES(k2_n, float, 1, n, a, s); break;
    case 2: GF_SIZES(k2_n, float, 2, n, a, s); break;
    case 3: GF_SIZES(k2_n, float, 3, n, a, s); break;
    case 4: GF_SIZES(k2_n, float, 4, n, a, s); break;
    default: k2_deep_f(nl, n, a, s); break;
  }
}
}  // namespace gf
