// Batched ensembles of independent small instances of uneven size (FP64 or FP32, one precision
// per ensemble).
//
// One launch per kernel class (K1, K2, K3) and stage covers every member of the
// ensemble: the grid is a flat table of (instance, block) work entries, ordered
// largest instance first, and each 256-thread workgroup runs the K1/K2/K3 body of
// its entry's size N on one block of column pairs (K1/K3) or row pairs (K2) of that
// instance.  With P = 8 points per thread every N in 32..512 fills the same
// 2304-slot LDS buffer, so the launch mixes sizes freely and fills all CUs even
// when single members are far too small to.
//
// GF_BATCH_K (1, 2 or 3) selects the kernel class this translation unit
// instantiates, so the three compile in parallel.
#include "gf_kernels_impl.hpp"

namespace gf {

constexpr int BATCH_THREADS = 256;
constexpr int BATCH_P = 8;
// G1 PairLds<T, N> (2 G1 LdsN<N> plus one column skew per pair) and G2 LdsN<N>: at most
// 2304 slots plus 32 skews (N = 32, 32 column pairs per workgroup) for every N.  The skew is 64
// bytes: 4 slots in FP64, 8 in FP32.
template <typename T> struct BatchLds {
  static constexpr int value = BATCH_THREADS * BATCH_P * 9 / 8 + 32 * ColSkew<T>::value;
};
constexpr int BATCH_TW = 128;

template <typename T, int N> struct BatchGeom {
  static constexpr int TPT = N / BATCH_P;
  static constexpr int G1 = BATCH_THREADS / (2 * TPT);   // column pairs per K1/K3 workgroup
  static constexpr int G2 = BATCH_THREADS / TPT;         // row pairs per K2 workgroup
  static_assert(PY<N>::P == BATCH_P, "batched sizes use 8 points per thread");
  static_assert(G1 * PairLds<T, N>::value <= BatchLds<T>::value && G2 * LdsN<N>::value <= BatchLds<T>::value, "LDS");
  static_assert(TwPlan<N, BATCH_P>::size() <= BATCH_TW, "twiddle table");
};

// PLAIN (K = 2): every member advects without background gradients, so the kernel holds only that
// x-pass body (122 VGPRs instead of 243 with the K2_ADVECT_FULL body beside it).
template <typename T, int N, int K, bool PLAIN>
__device__ __forceinline__ void batch_body(const BatchItem<T>& it, int blk, cplx<T>* lds, cplx<T>* tw) {
  using BG = BatchGeom<T, N>;
  if constexpr (K == 1) {
    load_twiddles<T, N, BATCH_P>(tw, it.k1.tw);
    __syncthreads();
    if (it.k1.model == M_TWODNS) k1_body<T, N, BATCH_P, BG::G1, 1, M_TWODNS, FS_CALCN>(it.k1, blk, lds, tw);
    else k1_body<T, N, BATCH_P, BG::G1, 1, M_SLQG, FS_CALCN>(it.k1, blk, lds, tw);
  } else if constexpr (K == 2) {
    load_twiddles<T, N, BATCH_P>(tw, it.k2.tw);
    __syncthreads();
    if (!PLAIN && it.k2.mode == K2_ADVECT_FULL) {
      if constexpr (!PLAIN) k2_body<T, N, BATCH_P, BG::G2, K2_ADVECT_FULL, 1>(it.k2, blk, lds, tw);
    } else {
      k2_body<T, N, BATCH_P, BG::G2, K2_ADVECT, 1>(it.k2, blk, lds, tw);
    }
  } else {
    load_twiddles<T, N, BATCH_P>(tw, it.k3.tw);
    __syncthreads();
    if (it.k3.model == M_MLQG) k3_body<T, N, BATCH_P, BG::G1, 1, 2, -1, true>(it.k3, blk, lds, tw);   // one layer
    else k3_body<T, N, BATCH_P, BG::G1, 1>(it.k3, blk, lds, tw);
  }
}

template <typename T, int K, bool PLAIN>
__global__ __launch_bounds__(BATCH_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
void batch_kernel(const BatchItem<T>* __restrict__ items, const BatchWork* __restrict__ work) {
  __shared__ cplx<T> lds[BatchLds<T>::value];
  __shared__ cplx<T> tw[BATCH_TW];
  const BatchWork w = work[blockIdx.x];
  const BatchItem<T>& it = items[w.item];
  switch (it.n) {
    case 32: batch_body<T, 32, K, PLAIN>(it, w.blk, lds, tw); break;
    case 64: batch_body<T, 64, K, PLAIN>(it, w.blk, lds, tw); break;
    case 128: batch_body<T, 128, K, PLAIN>(it, w.blk, lds, tw); break;
    case 256: batch_body<T, 256, K, PLAIN>(it, w.blk, lds, tw); break;
    case 512: batch_body<T, 512, K, PLAIN>(it, w.blk, lds, tw); break;
    default: break;
  }
}

#ifndef GF_BATCH_K
#define GF_BATCH_K 1
#endif

template <typename T, int K>
void launch_batch(const BatchItem<T>* items, const BatchWork* work, int nwork, int plain, hipStream_t s) {
  if (nwork <= 0) return;
  if (K == 2 && plain) GF_LAUNCH((batch_kernel<T, K, true>), dim3(nwork), dim3(BATCH_THREADS), 0, s, items, work);
  else GF_LAUNCH((batch_kernel<T, K, false>), dim3(nwork), dim3(BATCH_THREADS), 0, s, items, work);
}
template void launch_batch<double, GF_BATCH_K>(const BatchItem<double>*, const BatchWork*, int, int, hipStream_t);
template void launch_batch<float, GF_BATCH_K>(const BatchItem<float>*, const BatchWork*, int, int, hipStream_t);

#if GF_BATCH_K == 1
bool batch_size_supported(int n) { return n == 32 || n == 64 || n == 128 || n == 256 || n == 512; }

int batch_blocks(int k, int n, int count) {
  const int tpt = n / BATCH_P;
  if (k == 2) {   // count = ny: row pairs in groups of G2
    const int g2 = BATCH_THREADS / tpt;
    return (count / 2 + g2 - 1) / g2;
  }
  const int g1 = BATCH_THREADS / (2 * tpt);   // count = columns: column pairs in groups of G1
  return ((count + 1) / 2 + g1 - 1) / g1;
}
#endif

}  // namespace gf
