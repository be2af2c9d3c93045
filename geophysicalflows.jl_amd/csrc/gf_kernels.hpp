// Kernel argument blocks and launch entry points shared by the host context
// (gf_context.cpp) and the kernels (gf_kernels.hip).
//
// Device data layout ("column-blocked"): a spectral or mixed field with
// nkr columns (kr) and ny rows (l or y) is stored as
//     off(kr, j) = ((kr / BW) * ny + j) * BW + (kr % BW)
// i.e. blocks of BW kr columns are interleaved and each block runs contiguously
// along y.  The y-direction passes (K1, K3) read and write column pairs of a
// block; the x-direction pass (K2) reads and writes 16 BW-byte pieces (one row
// of a block) in FP64.  Layers (MultiLayerQG) and extra fields follow at fixed
// strides.
#pragma once
#include <stdint.h>
#include <hip/hip_ext.h>
#include "gf_fft.hpp"

namespace gf {

// Events that the next K1, K2, K3 or batched launch on this thread carries (kernel timing):
// hipExtLaunchKernel stamps `stop` with the kernel's end (and `start`, if set, marks the stream
// just before the launch). Timing by the ends of consecutive kernels leaves no event packets
// between them, which would add the queue's processing to each interval (about 4-5 us per
// kernel at 2048^2).
struct LaunchEvents { hipEvent_t start = nullptr, stop = nullptr; };
inline thread_local LaunchEvents launch_events;
#define GF_LAUNCH(K, GRID, BLK, SHM, S, ...)                                                        \
  do {                                                                                              \
    if (::gf::launch_events.stop)                                                                   \
      hipExtLaunchKernelGGL(K, GRID, BLK, SHM, S, ::gf::launch_events.start, ::gf::launch_events.stop, \
                            0u, __VA_ARGS__);                                                       \
    else                                                                                            \
      hipLaunchKernelGGL(K, GRID, BLK, SHM, S, __VA_ARGS__);                                        \
  } while (0)

// Columns per block of the blocked layout: 2 (pairs) or 4 (quads).
#ifndef GF_BW
#define GF_BW 2
#endif
constexpr int BW = GF_BW;
static_assert(BW == 2 || BW == 4, "the blocked layout interleaves 2 or 4 columns");

enum Model { M_TWODNS = 0, M_SLQG = 1, M_MLQG = 2 };

// Field codes computed by K1 from a spectral state.
enum Field { F_U = 0, F_V = 1, F_Q = 2, F_PSI = 3, F_RAW = 4 };

// Field sets produced by K1: FS_CALCN = (u, v, q) for the advection path,
// FS_OUT = the physical outputs of updatevars! ((zeta, u, v) for TwoDNS,
// (q, psi, u, v) for the QG models), FS_RAW = the state itself, unmasked.
enum FieldSet { FS_CALCN = 0, FS_OUT = 1, FS_RAW = 2 };

// K2 modes
enum K2Mode { K2_ADVECT = 0, K2_INVERSE_OUT = 1, K2_FORWARD_IN = 2, K2_ADVECT_FULL = 3 };

// K3 epilogues
enum Epi {
  E_CALCN = 0, E_STORE = 1,
  E_ETD1 = 2, E_ETD2 = 3, E_ETD3 = 4, E_ETD4 = 5,
  E_RK1 = 6, E_RK2 = 7, E_RK3 = 8, E_RK4 = 9,
  E_FE = 10, E_AB3 = 11
};

constexpr int MAXL = 8;  // max layers

template <typename T>
struct Geo {
  int nx, ny, nkr, nkr2;    // nkr2: nkr rounded up to a multiple of the block width BW
  int kmax;                 // kr < kmax survive dealias
  int la0, la1;             // l-index range [la0, la1) is aliased
  int nlayers;
  T kx0, ky0;               // 2 pi/Lx, 2 pi/Ly
  T inv_norm;               // 1/(nx ny)
  long lstride;             // complex elements between layers of a spectral field (nkr2*ny)
};

template <typename T>
struct K1Args {
  Geo<T> g;
  int model;
  int fieldset;             // FS_CALCN (u, v, q), FS_OUT (updatevars! outputs), FS_RAW (state)
  int p;                    // FFT points per thread (radix schedule); 0: the built-in choice
  const cplx<T>* sol;       // input state (masked on read)
  cplx<T>* out;             // fields of every layer, stride fstride
  long fstride;
  int ncols;                // kr columns to process (kr < ncols)
  int eb;                   // SLQG equivalent-barotropic: psi = -q/(K^2 + ell2inv)
  T ell2inv;
  const T* sinv;            // MLQG: S^-1 per mode, n*n reals, blocked layout (n*n per mode)
  int nomask;               // MLQG linearised equations: the state is read without dealiasing
  int lsplit;               // multi-layer: one layer per workgroup row (blockIdx.y)
  int fsplit;               // single-layer calcN: one output field per workgroup row (blockIdx.y)
  cplx<T>* psi_out;         // multi-layer calcN, nullable: the bottom layer's psih (for K3's drag)
  const cplx<T>* tw;
};

template <typename T>
struct K2Args {
  Geo<T> g;
  int mode;
  int nfields;              // inverse-out / forward-in: number of fields per layer
  int p;                    // FFT points per thread (radix schedule); 0: the built-in choice
  const cplx<T>* in;        // mixed fields (u, v, q per layer), stride fstride
  long fstride;
  int kin;                  // kr < kin present in the input mixed fields
  const T* eta;             // SLQG: physical eta (x fastest), nullable
  // advect-full (MLQG / SLQG varying U): per-layer U(y), Qy(y); bottom-layer Qx/Qy fields
  const T* Uy;              // nlayers*ny
  const T* Qy1;             // nlayers*ny
  const T* etax;            // ny*nx, bottom layer, nullable
  const T* etay;            // ny*nx, bottom layer, nullable
  int linear;               // MLQG linearised equations: advect q by U only
  int lsplit;               // multi-layer: one layer per workgroup row (blockIdx.y)
  int hsplit;               // single-layer advection: row 0 forms u q, row 1 v q (blockIdx.y)
  int has_bg;               // any nonzero Qx/Qy/U term
  cplx<T>* out;             // flux (2 fields per layer) or mixed (forward-in)
  long ostride;
  int kout;                 // write kr < kout
  T* phys;                  // inverse-out: nfields physical fields (ny*nx each, x fastest)
  const T* phys_in;         // forward-in
  long pstride;
  const cplx<T>* tw;
};

template <typename T>
struct K3Args {
  Geo<T> g;
  int model;
  const cplx<T>* in;        // flux fields (2 per layer), stride fstride
  int p;                    // FFT points per thread (radix schedule); 0: the built-in choice
  long fstride;
  int ncols;                // kr columns to process (kr < ncols)
  int full;                 // also produce the aliased band (last step / diagnostics)
  int epi;
  int apply_kr;             // N = -i kr A - i l B (else A - i l B)
  const cplx<T>* etah; T Ueta;   // SLQG constant-U topography term
  const cplx<T>* Fh;        // fixed forcing, nullable
  T fscale;                 // factor on Fh: 1, or cos(omega t + phase) at the stage time
  const T* sqrt_spec;       // ring forcing amplitude (blocked real), nullable
  uint64_t seed; int64_t fstep; T sqrt_dt; int fkx, fky; int instance;
  // ring forcing: (fkx + 1) x nbl box of this step's forcing terms (one per layer row when
  // lsplit), written by K3 itself before its transforms and read back in the stage epilogue
  // (ring_box_row)
  cplx<T>* noise; int nbl;
  const int64_t* fstep_dev; // nullable: the forcing step is read here (steps replayed from a graph)
  // MLQG bottom drag mu Krsq psi_n
  T mu_drag; const cplx<T>* drag_src;  // calcN input state (for psi_n)
  const T* sinv;
  // stepper state
  cplx<T>* sol; cplx<T>* s1; cplx<T>* s2; cplx<T>* n1; cplx<T>* n23; cplx<T>* racc; cplx<T>* nout;
  const void* c0; const void* c1; const void* c2; const void* c3; const void* c4; const void* c5;
  int ccplx;                // coefficient tables are complex
  int cl_layered;           // coefficient tables have one slice per layer
  T dt;
  const T* filter;          // nullable
  int zero_mean;            // E_STORE: set mode (0,0) to zero
  int nomask;               // MLQG linearised equations: no dealiasing of the state
  int lsplit;               // multi-layer: one layer per workgroup row (blockIdx.y)
  int csplit;               // single-layer: one column per workgroup where k3_single_column allows
  const cplx<T>* psi_in;    // multi-layer, nullable: the bottom layer's psih written by K1
  int ab3_start;            // AB3: forward Euler start-up step
  // adjoint tape (ETDRK4): nsave receives each stage's N, sol_out the new state of
  // E_ETD4 (both nullable: no tape / update sol in place)
  cplx<T>* nsave;
  cplx<T>* sol_out;
  // nullable: the final stage of a step writes 1 here (a plain store from the lanes that saw one)
  // when a value of the new state is not finite; host memory, read by the context after its stream
  int* nonfinite;
  const cplx<T>* tw;
};

// Adjoint of the TwoDNavierStokes ETDRK4 step (gf_adjoint_impl.hpp).
template <typename T>
struct AdjK1Args {
  Geo<T> g;
  // the stage's calcN input, rebuilt from the step's input s0 and the taped N1, N2, N3 as the
  // forward stages formed it (stage 1: s0; 2: E2 s0 + Z N1; 3: E2 s0 + Z N2;
  // 4: E2 s1 + Z (2 N3 - N1)); masked on read
  const cplx<T>* s0; const cplx<T>* N1; const cplx<T>* N2; const cplx<T>* N3;
  const T* E2; const T* Z;  // ETDRK4 coefficients (real, blocked)
  int stage;                // 4, 3, 2, 1
  const cplx<T>* x;         // co-state source; lambda_N = lamc * x (masked on read)
  const T* lamc;            // real coefficient table (blocked), nullable = 1
  cplx<T>* out;             // 5 mixed fields: u, v, zeta, pA, pB
  long fstride;
  int ncols;                // kmax
  int fsplit;               // one field per workgroup row (blockIdx.y)
  const cplx<T>* tw;
};

template <typename T>
struct AdjK2Args {
  Geo<T> g;
  const cplx<T>* in;        // the 5 mixed fields of AdjK1Args
  long fstride;
  int kin;
  cplx<T>* out;             // 3 flux fields: x-spectra of zeta pA, zeta pB, u pA + v pB
  long ostride;
  int kout;
  int hsplit;               // one flux field per workgroup row (blockIdx.y)
  const cplx<T>* tw;
};

template <typename T>
struct AdjK3Args {
  Geo<T> g;
  const cplx<T>* in;        // the 3 flux fields of AdjK2Args
  long fstride;
  int ncols;                // kmax
  int stage;                // 4, 3, 2, 1: adjoint of the ETDRK4 stage with that calcN
  // co-state buffers (blocked): lambda of sol' (in; out at stage 1), partial sums
  cplx<T>* lam; cplx<T>* l0; cplx<T>* l1; cplx<T>* lN1; cplx<T>* lN2; cplx<T>* lN3;
  const T* E; const T* E2; const T* Z; const T* Al; const T* Be;   // ETDRK4 coefficients (real)
  // dJ/dnu: coefficient derivatives times dL/dnu, the step's tape (s1 = E2 s0 + Z N1 is rebuilt),
  // one slot per workgroup
  const T* dE; const T* dE2; const T* dZ; const T* dA; const T* dB; const T* dG;
  const cplx<T>* s0;
  const cplx<T>* N1; const cplx<T>* N2; const cplx<T>* N3; const cplx<T>* N4;
  T* nu_part;               // nullable: no nu gradient
  const cplx<T>* tw;
};

// k = 1, 2, 3: launch KA1/KA2/KA3 for grid size n (square grids) with args of that kernel
template <typename T> void launch_adj(int k, int n, const void* args, hipStream_t s);
// workgroups of KA3 (and KA1) for ncols columns at size n
template <typename T> int adj_blocks(int n, int ncols);

// Domain-averaged diagnostics (gf_diag.hip); kinds as GF_DIAG_* in include/gf.h.
enum DiagKind {
  DIAG_ENERGY = 0, DIAG_KINETIC_ENERGY = 1, DIAG_POTENTIAL_ENERGY = 2, DIAG_ENSTROPHY = 3,
  DIAG_PALINSTROPHY = 4, DIAG_ENERGY_DISSIPATION = 5, DIAG_ENSTROPHY_DISSIPATION = 6,
  DIAG_ENERGY_DRAG = 7, DIAG_ENSTROPHY_DRAG = 8, DIAG_ENERGY_WORK = 9, DIAG_ENSTROPHY_WORK = 10,
  DIAG_LAYER_ENERGIES = 11, DIAG_FLUXES = 12
};
constexpr int DIAG_MAXOUT = 2 * MAXL - 1;

template <typename T>
struct DiagArgs {
  Geo<T> g;
  int kind, model, nl, nout;
  const cplx<T>* sol;       // raw state (not dealiased: the reference's diagnostics read sol as is)
  const cplx<T>* prevsol;   // stochastic forcing: state at the start of the last step, nullable
  const cplx<T>* Fh;        // fixed forcing table, nullable
  T fscale;                 // factor on Fh at the last forcing evaluation
  const T* sqrt_spec;       // ring forcing, nullable
  uint64_t seed; int64_t fstep; T sqrt_dt; int fkx, fky, instance;
  const cplx<T>* etah;      // SLQG topographic PV spectrum, nullable
  int eb; T ell2inv;
  T xi; int nxi;            // dissipation coefficient and exponent
  const T* sinv;            // MLQG S^-1 per mode
  double lay_ke[MAXL], lay_pe[MAXL];   // MLQG: H_j/(2V), f0^2/(2 V g'_j)
};
template <typename T> void launch_diag(const DiagArgs<T>& a, double* part, int nblocks, hipStream_t s);
void launch_diag_finish(const double* part, int nblocks, int nout, double scale, double* out, hipStream_t s);

// MultiLayerQG fluxes (src/multilayerqg.jl:977-1022).  Step 1 (flux_fields): from the dealiased
// state, per layer j, the spectra psih_j (S^-1 q), vh_j = i kr psih_j and (du/dy)h_j = i l uh_j with
// uh_j = -i l psih_j, written to layer j of f[0], f[1], f[2].  Step 2: those fields are
// transformed along y (K1, raw).  Step 3 (flux_kernel): the x sums of the products by Parseval
// along x on the mixed fields, weighted by U(y): lateral_j ~ U_j v_j du_j/dy, vertical_{j+1/2} ~
// (U_j - U_{j+1}) v_{j+1} psi_j; fac[o] holds each output's constant factor.
template <typename T>
struct FluxArgs {
  Geo<T> g;
  int nl;
  const cplx<T>* sol;       // dealiased state
  const T* sinv;            // S^-1 per mode (n n reals), nullable for one layer (psih = -invKrsq q)
  cplx<T>* f[3];            // psi, v, du/dy, nl layers each at stride fstride: spectral fields
                            // (flux_fields) or their y transforms (flux_kernel)
  long fstride;
  const T* Uy;              // nl ny
  double fac[DIAG_MAXOUT];
};
template <typename T> void launch_flux_fields(const FluxArgs<T>& a, hipStream_t s);
template <typename T> void launch_flux(const FluxArgs<T>& a, double* part, int nblocks, hipStream_t s);

// pvfromstreamfunction! (src/multilayerqg.jl:520-552): q_j = sum_k S[j, k] psi_k per mode with
// S = -Krsq I + F (calcS!, 603-613); one layer: q = -Krsq psi.  Blocked spectral fields.
template <typename T>
struct PvArgs {
  Geo<T> g;
  int nl;
  const cplx<T>* psi;
  cplx<T>* q;
  T F[MAXL * MAXL];         // layer coupling, row-major
};
template <typename T> void launch_pv_from_psi(const PvArgs<T>& a, hipStream_t s);

template <typename T> void launch_k1(int ny, const K1Args<T>& a, hipStream_t s);
template <typename T> void launch_k2(int nx, const K2Args<T>& a, hipStream_t s);
template <typename T> void launch_k3(int ny, const K3Args<T>& a, hipStream_t s);
template <typename T> void launch_to_julia(const Geo<T>& g, const cplx<T>* src, cplx<T>* dst, int nl, hipStream_t s);
template <typename T> void launch_from_julia(const Geo<T>& g, const cplx<T>* src, cplx<T>* dst, int nl, hipStream_t s);
template <typename T> void launch_mask(const Geo<T>& g, cplx<T>* sol, int nl, hipStream_t s);
void launch_counter(int64_t* c, int64_t v, int add, hipStream_t s);
// Grid lengths that are not powers of two (even, 18..GEN_MAX): a column length ny runs in the
// y-pass kernels of the next power of two with a direct DFT (dft_direct), a row length nx in the
// x-pass kernel of direct sums (k2_gen_kernel).
constexpr int GEN_MAX = 256;
__host__ __device__ constexpr int pow2_ceil(int n) { int m = 1; while (m < n) m *= 2; return m; }
__host__ __device__ constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
inline bool generic_len_supported(int n) { return n % 2 == 0 && n >= 18 && n <= GEN_MAX && !is_pow2(n); }
// Radix schedules offered to the autotuner for an FFT of length n (the host side of PCand).
inline bool sched_candidate(int n, int p) {
  return p == 4 ? (n >= 16 && n <= 2048) : (p == 8 ? n >= 32 : (p == 16 ? n >= 1024 : false));
}
// The built-in choice (PY in the kernels): P = 4 below 32 points, 16 from 4096, else 8.
inline int sched_default(int n) { return n < 32 ? 4 : (n >= 4096 ? 16 : 8); }
// Row of spectral row l in a ring-forcing box of nbl = min(ny, 2 fky + 1) rows holding
// l <= fky and l >= ny - fky.
__host__ __device__ inline int ring_box_row(int l, int ny, int fky, int nbl) { return l <= fky ? l : l - ny + nbl; }
// K3 runs one column per workgroup (csplit) when a column's n/p threads fill at least one wave;
// otherwise csplit is ignored and the column pairs share workgroups.
constexpr bool k3_single_column(int n, int p) { return n / p >= 64; }
// Workgroups of a launch with G column pairs per workgroup over ncols columns; with quads and
// G = 1, a multiple of 16 (pair_of_block puts the two pairs of a block on one XCD).
constexpr int pair_blocks(int ncols, int G) {
  const int n = ((ncols + 1) / 2 + G - 1) / G;
  return (BW == 4 && G == 1) ? 16 * ((n + 15) / 16) : n;
}
// Workgroups of a one-column-per-workgroup launch over ncols columns (single_column): the BW
// columns of a block go to workgroups 8 apart, on one XCD.
constexpr int column_blocks(int ncols) { return 8 * BW * (((ncols + BW - 1) / BW + 7) / 8); }
bool size_supported(int n);

}  // namespace gf

namespace gf {

// One instance of a batched ensemble for one stage: the three kernels' argument
// blocks (with the stage's input buffer, epilogue and full flag filled in).
template <typename T>
struct BatchItem {
  int n;                    // grid size (nx == ny)
  K1Args<T> k1;
  K2Args<T> k2;
  K3Args<T> k3;
};

// One workgroup's share of a batched launch: block blk of instance item.
struct BatchWork {
  int item;
  int blk;
};

// Launch kernel class K (1, 2, 3) of one stage for every instance: workgroup b runs
// block work[b].blk of instance work[b].item (FP64 or FP32).
// plain (K = 2): every member's x pass is K2_ADVECT, which selects the kernel without the
// background-gradient body
template <typename T, int K> void launch_batch(const BatchItem<T>* items, const BatchWork* work, int nwork, int plain, hipStream_t s);
// The three work tables of one stage (K1, K2, K3 entries) for the persistent stage launch.
struct BatchStageWork {
  const BatchWork* w[3];
  int n[3];
};
// All three kernel classes of one stage for every instance in one cooperative launch of the
// workgroups the GPU holds at once (grid-wide barriers between the classes).
template <typename T> hipError_t launch_batch_stage(const BatchItem<T>* items, const BatchStageWork& sw, int plain, hipStream_t s);
// workgroups of kernel class k for an instance of size n: count = columns (k = 1, 3) or ny (k = 2)
int batch_blocks(int k, int n, int count);
bool batch_size_supported(int n);

}  // namespace gf
