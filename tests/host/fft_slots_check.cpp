// Exhaustive host check of the padded LDS slot arithmetic of the FFT core (gf_fft.hpp): for
// every (N, P) plan the library instantiates and both point sizes (FP64 16-byte and FP32 8-byte
// slots), each Stockham write slot padg<G>(base + r NS) and read slot padg<G>(t + q TPT), with
// the exchange's padding period G = xpad_g, equals the split form padg<G>(base) + padg<G>(r NS)
// the kernels use and lies inside the N + N/8-slot buffer; MirrorSlots<N, TPT, G> gives
// padg<G>(kr) and padg<G>((N - kr) mod N) for every kr = t + i TPT <= N/2 and each period the
// kernels use (8, 16, 32).  Built with hipcc for the host only; exit status 0 when all agree.
#include "gf_fft.hpp"
#include <cstdio>
using namespace gf;
template <int N, int TPT, int G> int check_mirror() {
  int bad = 0;
  for (int t = 0; t < TPT; ++t) {
    MirrorSlots<N, TPT, G> s(t);
    for (int i = 0; t + i * TPT <= N / 2; ++i) {
      const int kr = t + i * TPT;
      bad += s.k(i) != padg<G>(kr);
      bad += s.m(i) != padg<G>((N - kr) & (N - 1));
      bad += s.m(i) < 0 || s.m(i) >= LdsN<N>::value;
    }
  }
  return bad;
}
template <int N, int P, int S, int PASS = 0, int NS = 1> int check_passes() {
  using PL = Plan<N, P>;
  if constexpr (PASS + 1 >= PL::NPASS) {
    return 0;   // the last pass writes no exchange
  } else {
    constexpr int TPT = N / P, R = PL::radix(PASS), M = P / R, G = xpad_g<S, NS, R>();
    int bad = 0;
    for (int t = 0; t < TPT; ++t)
      for (int b = 0; b < M; ++b) {
        const int j = t + b * TPT, base = (j / NS) * NS * R + (j % NS);
        for (int r = 0; r < R; ++r) {
          const int s = padg<G>(base + r * NS);
          bad += padg<G>(base) + padg<G>(r * NS) != s;
          bad += s < 0 || s >= LdsN<N>::value;
        }
      }
    for (int t = 0; t < TPT; ++t)
      for (int q = 0; q < P; ++q) bad += padg<G>(t) + padg<G>(TPT * q) != padg<G>(t + TPT * q);
    return bad + check_passes<N, P, S, PASS + 1, NS * R>();
  }
}
template <int N, int P> int check() {
  constexpr int TPT = N / P;
  int bad = check_passes<N, P, 16>() + check_passes<N, P, 8>();
  for (int t = 0; t < TPT; ++t)
    for (int q = 0; q < P; ++q) bad += pad(t) + pad(TPT * q) != pad(t + TPT * q);
  bad += check_mirror<N, TPT, 8>() + check_mirror<N, TPT, 16>() + check_mirror<N, TPT, 32>();
  printf("N=%5d P=%2d TPT=%4d mismatches %d\n", N, P, TPT, bad);
  return bad;
}
int main() {
  int bad = 0;
  bad += check<16, 4>() + check<32, 4>() + check<32, 8>() + check<64, 4>() + check<64, 8>();
  bad += check<128, 4>() + check<128, 8>() + check<256, 4>() + check<256, 8>() + check<512, 4>() + check<512, 8>();
  bad += check<1024, 4>() + check<1024, 8>() + check<1024, 16>() + check<2048, 4>() + check<2048, 8>() + check<2048, 16>();
  bad += check<4096, 8>() + check<4096, 16>();
  printf("total %d\n", bad);
  return bad != 0;
}
