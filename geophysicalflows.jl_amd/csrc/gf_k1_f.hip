// Instantiation unit: K1 for float (sizes 16..2048, 1..4 layers; 5..8 layers in gf_k1_fm.hip).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k1_deep_f(int nl, int n, const K1Args<float>& a, hipStream_t s);
template <> void launch_k1<float>(int n, const K1Args<float>& a, hipStream_t s) {
  const int nl = a.model == M_MLQG ? a.g.nlayers : 1;
  switch (nl) {
    case 1: GF_SIZES(k1_n, float, 1, n, a, s); break;
    case 2: GF_SIZES(k1_n, float, 2, n, a, s); break;
    case 3: GF_SIZES(k1_n, float, 3, n, a, s); break;
    case 4: GF_SIZES(k1_n, float, 4, n, a, s); break;
    default: k1_deep_f(nl, n, a, s); break;
  }
}
}  // namespace gf
