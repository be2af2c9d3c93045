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
// instantiates, so the three compile in parallel; GF_BATCH_K = 0 instantiates the persistent
// stage kernel, which runs all three classes of a stage in one cooperative launch.
#include "gf_kernels_impl.hpp"
#include <hip/hip_cooperative_groups.h>
#include <algorithm>

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

// kernel class K of work entry w, at the entry's instance size
template <typename T, int K, bool PLAIN>
__device__ __forceinline__ void batch_entry(const BatchItem<T>* __restrict__ items, const BatchWork w,
                                            cplx<T>* lds, cplx<T>* tw) {
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

template <typename T, int K, bool PLAIN>
__global__ __launch_bounds__(BATCH_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
void batch_kernel(const BatchItem<T>* __restrict__ items, const BatchWork* __restrict__ work) {
  __shared__ cplx<T> lds[BatchLds<T>::value];
  __shared__ cplx<T> tw[BATCH_TW];
  batch_entry<T, K, PLAIN>(items, work[blockIdx.x], lds, tw);
}

#ifndef GF_BATCH_K
#define GF_BATCH_K 1
#endif

#if GF_BATCH_K >= 1
template <typename T, int K>
void launch_batch(const BatchItem<T>* items, const BatchWork* work, int nwork, int plain, hipStream_t s) {
  if (nwork <= 0) return;
  if (K == 2 && plain) GF_LAUNCH((batch_kernel<T, K, true>), dim3(nwork), dim3(BATCH_THREADS), 0, s, items, work);
  else GF_LAUNCH((batch_kernel<T, K, false>), dim3(nwork), dim3(BATCH_THREADS), 0, s, items, work);
}
template void launch_batch<double, GF_BATCH_K>(const BatchItem<double>*, const BatchWork*, int, int, hipStream_t);
template void launch_batch<float, GF_BATCH_K>(const BatchItem<float>*, const BatchWork*, int, int, hipStream_t);
#else
// One stage of every member in one persistent launch: the workgroups the GPU holds at once walk
// the K1 work entries, then (after a grid-wide barrier, which also makes K1's stores visible)
// the K2 entries, then the K3 entries. Every workgroup reaches both barriers: the loops hold no
// early exit and the grid never exceeds the co-resident workgroups (cooperative launch).
// Each class's walk is a function of its own (not inlined), so that its registers are allocated
// for that class alone; the LDS is the launch's dynamic allocation, which each walk addresses
// directly (so its accesses stay LDS instructions).
template <typename T> struct BatchStageLds {
  static constexpr size_t bytes = sizeof(cplx<T>) * (BatchLds<T>::value + BATCH_TW);
};
template <typename T, int K, bool PLAIN>
__device__ __attribute__((noinline)) void batch_walk(const BatchItem<T>* __restrict__ items, const BatchWork* __restrict__ w, int n) {
  extern __shared__ __align__(16) char batch_dyn_lds[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(batch_dyn_lds);
  cplx<T>* tw = lds + BatchLds<T>::value;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    __syncthreads();   // the previous entry's LDS reads are done before this entry's twiddles land
    batch_entry<T, K, PLAIN>(items, w[i], lds, tw);
  }
}
// (3 waves per SIMD where the x pass is the plain one, as the separate kernels of each class run;
// with the background-gradient x pass beside them, 2)
template <typename T, bool PLAIN>
__global__ __launch_bounds__(BATCH_THREADS) __attribute__((amdgpu_waves_per_eu(PLAIN ? 3 : 2)))
void batch_stage_kernel(const BatchItem<T>* __restrict__ items, BatchStageWork sw) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  batch_walk<T, 1, PLAIN>(items, sw.w[0], sw.n[0]);
  grid.sync();
  batch_walk<T, 2, PLAIN>(items, sw.w[1], sw.n[1]);
  grid.sync();
  batch_walk<T, 3, PLAIN>(items, sw.w[2], sw.n[2]);
}
// Co-resident workgroups of the stage kernel (CUs times workgroups per CU), computed once per
// process on the current device (every device of a process is taken to be the same GPU).
template <typename T, bool PLAIN>
int batch_stage_grid() {
  static const int g = [] {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_stage_kernel<T, PLAIN>, BATCH_THREADS, BatchStageLds<T>::bytes) != hipSuccess)
      return -1;
    return ncu * per;
  }();
  return g;
}
template <typename T>
hipError_t launch_batch_stage(const BatchItem<T>* items, const BatchStageWork& sw, int plain, hipStream_t s) {
  const int g = plain ? batch_stage_grid<T, true>() : batch_stage_grid<T, false>();
  if (g <= 0) return hipErrorInvalidConfiguration;
  const int most = std::max(sw.n[0], std::max(sw.n[1], sw.n[2]));
  const dim3 grid(std::max(1, std::min(g, most))), blk(BATCH_THREADS);
  BatchStageWork a = sw;
  void* args[] = {(void*)&items, (void*)&a};
  const size_t sh = BatchStageLds<T>::bytes;
  return plain ? hipLaunchCooperativeKernel((const void*)batch_stage_kernel<T, true>, grid, blk, args, sh, s)
               : hipLaunchCooperativeKernel((const void*)batch_stage_kernel<T, false>, grid, blk, args, sh, s);
}
template hipError_t launch_batch_stage<double>(const BatchItem<double>*, const BatchStageWork&, int, hipStream_t);
template hipError_t launch_batch_stage<float>(const BatchItem<float>*, const BatchStageWork&, int, hipStream_t);
#endif

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
