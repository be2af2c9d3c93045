// Exhaustive host check of ring_box_row (gf_kernels.hpp): for every ny the library supports and
// every forcing half-width fky, the spectral rows l <= fky and l >= ny - fky map one to one
// onto the rows 0 .. nbl - 1 of the ring-forcing box, nbl = min(ny, 2 fky + 1).  Built with
// hipcc for the host only; exit status 0 when all agree.
#include "gf_kernels.hpp"
#include <algorithm>
#include <cstdio>
#include <vector>
int main() {
  int bad = 0, cases = 0;
  for (int ny = 16; ny <= 4096; ny *= 2)
    for (int fky = 0; fky <= ny / 2 + 1; ++fky) {
      const int nbl = std::min(ny, 2 * fky + 1);
      std::vector<int> seen(nbl, 0);
      for (int l = 0; l < ny; ++l) {
        if (!(l <= fky || l >= ny - fky)) continue;
        const int r = gf::ring_box_row(l, ny, fky, nbl);
        if (r < 0 || r >= nbl) { ++bad; continue; }
        ++seen[r];
      }
      for (int c : seen) bad += c != 1;
      ++cases;
    }
  std::printf("cases %d mismatches %d\n", cases, bad);
  return bad != 0;
}
