// Instantiation unit: K3 for float (sizes 16..2048, 1..4 layers; 5..8 layers in gf_k3_fm.hip).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k3_deep_f(int nl, int n, const K3Args<float>& a, hipStream_t s);
template <> void launch_k3<float>(int n, const K3Args<float>& a, hipStream_t s) {
  const int nl = a.model == M_MLQG ? a.g.nlayers : 1;
  switch (nl) {
    case 1: GF_SIZES(k3_n, float, 1, n, a, s); break;
    case 2: GF_SIZES(k3_n, float, 2, n, a, s); break;
    case 3: GF_SIZES(k3_n, float, 3, n, a, s); break;
    case 4: GF_SIZES(k3_n, float, 4, n, a, s); break;
    default: k3_deep_f(nl, n, a, s); break;
  }
}
}  // namespace gf
