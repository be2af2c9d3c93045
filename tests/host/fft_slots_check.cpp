// Exhaustive host check of the padded LDS slot arithmetic of the FFT core (gf_fft.hpp): for
// every (N, P) plan the library instantiates, each Stockham write slot pad(base + r NS) and
// read slot pad(t + q TPT) equals the split form pad(base) + pad(r NS) the kernels use, and
// MirrorSlots gives pad(kr) and pad((N - kr) mod N) for every kr = t + i TPT <= N/2 inside the
// N + N/8-slot buffer.  Built with hipcc for the host only; exit status 0 when all agree.
#include "gf_fft.hpp"
#include <cstdio>
using namespace gf;
template <int N, int P> int check() {
  constexpr int TPT = N / P;
  int bad = 0;
  using PL = Plan<N, P>;
  // Stockham writes and reads
  int NS = 1;
  for (int p = 0; p < PL::NPASS; ++p) {
    const int R = PL::radix(p), M = P / R;
    for (int t = 0; t < TPT; ++t)
      for (int b = 0; b < M; ++b) {
        const int j = t + b * TPT, base = (j / NS) * NS * R + (j % NS);
        for (int r = 0; r < R; ++r) bad += pad(base) + pad(r * NS) != pad(base + r * NS);
      }
    NS *= R;
  }
  for (int t = 0; t < TPT; ++t)
    for (int q = 0; q < P; ++q) bad += pad(t) + pad(TPT * q) != pad(t + TPT * q);
  // mirror slots
  for (int t = 0; t < TPT; ++t) {
    MirrorSlots<N, TPT> s(t);
    for (int i = 0; t + i * TPT <= N / 2; ++i) {
      const int kr = t + i * TPT;
      bad += s.k(i) != pad(kr);
      bad += s.m(i) != pad((N - kr) & (N - 1));
      bad += s.m(i) < 0 || s.m(i) >= LdsN<N>::value;
    }
  }
  printf("N=%5d P=%2d TPT=%4d mismatches %d\n", N, P, TPT, bad);
  return bad;
}
int main() {
  int bad = 0;
  bad += check<16, 4>() + check<32, 4>() + check<32, 8>() + check<64, 4>() + check<64, 8>();
  bad += check<128, 4>() + check<128, 8>() + check<256, 4>() + check<256, 8>() + check<512, 4>() + check<512, 8>();
  bad += check<1024, 4>() + check<1024, 8>() + check<1024, 16>() + check<2048, 4>() + check<2048, 8>() + check<2048, 16>();
  bad += check<4096, 16>();
  printf("total %d\n", bad);
  return bad != 0;
}
