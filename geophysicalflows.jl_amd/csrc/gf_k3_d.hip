// Instantiation unit: K3 for double (sizes 16..2048, 1..4 layers; 5..8 layers in gf_k3_dm.hip).
#include "gf_kernels_impl.hpp"
#include "gf_inst.hpp"
namespace gf {
void k3_deep_d(int nl, int n, const K3Args<double>& a, hipStream_t s);
template <> void launch_k3<double>(int n, const K3Args<double>& a, hipStream_t s) {
  const int nl = a.model == M_MLQG ? a.g.nlayers : 1;
  switch (nl) {
    case 1: GF_SIZES(k3_n, double, 1, n, a, s); break;
    case 2: GF_SIZES(k3_n, double, 2, n, a, s); break;
    case 3: GF_SIZES(k3_n, double, 3, n, a, s); break;
    case 4: GF_SIZES(k3_n, double, 4, n, a, s); break;
    default: k3_deep_d(nl, n, a, s); break;
  }
}
}  // namespace gf
