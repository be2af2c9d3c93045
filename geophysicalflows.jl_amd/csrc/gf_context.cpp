// Host side of libgfhip: problem context, setup tables and the C ABI (include/gf.h).
//
// The context mirrors what FourierFlows.Problem builds for the reference
// (grid, Equation.L, time-stepper coefficients and buffers, clock) and drives
// the fused kernels of gf_kernels_impl.hpp on one HIP stream.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "gf_kernels.hpp"
#include "gf_schedule.hpp"
#include "../../include/gf.h"

using namespace gf;
using cd = std::complex<double>;

namespace {

struct GfError : std::runtime_error {
  int code;
  GfError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      throw GfError(GF_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));        \
  } while (0)

// FourierFlows' aliased index rule ([FF] getaliasedwavenumbers), 0-based [lo, hi).
void aliased_range(int nk, int nkr, double f, int& lo_full, int& hi_full, int& lo_r) {
  if (f <= 0) { lo_full = hi_full = nk; lo_r = nkr; return; }
  const double L = (1 - f) / 2, R = (1 + f) / 2;
  const int iL = (int)std::floor(L * nk) + 1, iR = (int)std::ceil(R * nk);
  lo_full = iL - 1; hi_full = iR; lo_r = iL - 1;
}

inline long boff_h(int kr, int j, int ny) { return ((long)(kr / BW) * ny + j) * BW + (kr % BW); }

template <typename F>
void parallel_for(long n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  const long chunk = (n + nt - 1) / nt;
  for (unsigned i = 0; i < nt; ++i) {
    const long a = i * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=] { for (long j = a; j < b; ++j) f(j); });
  }
  for (auto& t : th) t.join();
}

// Kassam--Trefethen contour means ([FF] getetdcoeffs, ncirc = 32, rcirc = 1).
void etd_coeffs(cd dtL, double dt, cd& zeta, cd& alpha, cd& beta, cd& gamma) {
  const int M = 32;
  cd sz = 0, sa = 0, sb = 0, sg = 0;
  for (int m = 0; m < M; ++m) {
    const double th = 2 * M_PI / M * (m + 0.5);
    const cd z = dtL + cd(std::cos(th), std::sin(th));
    const cd ez = std::exp(z), z3 = z * z * z;
    sz += (std::exp(z / 2.0) - 1.0) / z;
    sa += (-4.0 - z + ez * (4.0 - 3.0 * z + z * z)) / z3;
    sb += (2.0 + z + ez * (-2.0 + z)) / z3;
    sg += (-4.0 - 3.0 * z - z * z + ez * (4.0 - z)) / z3;
  }
  zeta = dt * (sz / (double)M);
  alpha = dt * (sa / (double)M);
  beta = dt * (sb / (double)M);
  gamma = dt * (sg / (double)M);
}

// d/dL of the contour-mean coefficients (zeta, alpha, beta, Gamma): with z = dt L + e^{i theta},
// d/dL <f(z)> = dt <f'(z)>, times the dt the coefficients carry.
void etd_coeff_derivs(cd dtL, double dt, cd& dzeta, cd& dalpha, cd& dbeta, cd& dgamma) {
  const int M = 32;
  cd sz = 0, sa = 0, sb = 0, sg = 0;
  for (int m = 0; m < M; ++m) {
    const double th = 2 * M_PI / M * (m + 0.5);
    const cd z = dtL + cd(std::cos(th), std::sin(th));
    const cd ez = std::exp(z), ez2 = std::exp(z / 2.0), z2 = z * z, z4 = z2 * z2;
    sz += (ez2 * (z / 2.0 - 1.0) + 1.0) / z2;
    sa += (12.0 + 2.0 * z + ez * (-12.0 + 10.0 * z - 4.0 * z2 + z2 * z)) / z4;
    sb += (-6.0 - 2.0 * z + ez * (6.0 - 4.0 * z + z2)) / z4;
    sg += (12.0 + 6.0 * z + z2 + ez * (-12.0 + 6.0 * z - z2)) / z4;
  }
  const double f = dt * dt / M;
  dzeta = f * sz; dalpha = f * sa; dbeta = f * sb; dgamma = f * sg;
}

// Device-wide runtime calls and stream capture.
//
// hipMalloc, hipFree, hipHostMalloc/hipHostFree, the synchronous hipMemcpy and the null-stream
// hipMemset/hipStreamSynchronize, stream creation/destruction and graph-exec destruction may
// synchronise the device or take runtime-wide locks. Contexts may be driven from several host
// threads, each capturing its own stream into a HIP graph, and such a call on one thread must not
// interleave with a capture on another. Every such call goes through the wrappers below, which
// hold g_devwide shared; a capture (begin, the recorded launches, end, instantiate) holds it
// exclusively (CaptureLock). Captures use hipStreamCaptureModeThreadLocal, so an unsafe call made
// by the capturing thread itself fails instead of being recorded or blocking.
// (The raw runtime calls appear only between the two DEVWIDE markers; a host test checks that.)
static std::shared_mutex g_devwide;
// DEVWIDE-BEGIN
namespace dw {
using SL = std::shared_lock<std::shared_mutex>;
hipError_t malloc(void** p, size_t n) { SL l(g_devwide); return hipMalloc(p, n); }
hipError_t free(void* p) { SL l(g_devwide); return hipFree(p); }
hipError_t host_malloc(void** p, size_t n) { SL l(g_devwide); return hipHostMalloc(p, n, hipHostMallocDefault); }
hipError_t host_free(void* p) { SL l(g_devwide); return hipHostFree(p); }
hipError_t copy(void* dst, const void* src, size_t n, hipMemcpyKind k) { SL l(g_devwide); return hipMemcpy(dst, src, n, k); }
// zero n bytes and wait: the contexts' streams are non-blocking, so a null-stream memset is not
// ordered before their work (a late memset would clear what a kernel wrote)
hipError_t zero_sync(void* p, size_t n) {
  SL l(g_devwide);
  hipError_t e = hipMemset(p, 0, n);
  return e == hipSuccess ? hipStreamSynchronize(nullptr) : e;
}
hipError_t stream_create(hipStream_t* s) { SL l(g_devwide); return hipStreamCreateWithFlags(s, hipStreamNonBlocking); }
hipError_t stream_destroy(hipStream_t s) { SL l(g_devwide); return hipStreamDestroy(s); }
hipError_t exec_destroy(hipGraphExec_t g) { SL l(g_devwide); return hipGraphExecDestroy(g); }
// stream -> graph exec, holding g_devwide exclusively; rec() records the launches
template <typename F>
hipGraphExec_t capture(hipStream_t s, F&& rec) {
  std::unique_lock<std::shared_mutex> l(g_devwide);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) throw GfError(GF_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  try {
    rec();
  } catch (...) {
    (void)hipStreamEndCapture(s, &gr);
    if (gr) (void)hipGraphDestroy(gr);
    throw;
  }
  e = hipStreamEndCapture(s, &gr);
  if (e != hipSuccess) throw GfError(GF_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (e != hipSuccess) throw GfError(GF_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return ex;
}
}  // namespace dw
// DEVWIDE-END

// Scratch of gf_diagnostic_async_noalloc: the diagnostics' per-workgroup partial sums (at most
// kScratchBlocks workgroups x DIAG_MAXOUT outputs, 120 KiB), in a static device array that every
// GPU the library runs on holds until the process ends.  In host code `scratch` names a host-side
// stand-in, not a device address: its address on a device comes from hipGetSymbolAddress with
// that device current (scratch_on), and only that address is handed to kernels.
constexpr int kScratchBlocks = 1024;
constexpr size_t kScratchDoubles = (size_t)kScratchBlocks * DIAG_MAXOUT;
static __device__ __align__(256) double scratch[kScratchDoubles];
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_addr;   // device -> that device's copy of scratch

// Resolved once per device, with that device current (at context set-up).
double* scratch_on(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(dev);
  if (it != g_scratch_addr.end()) return it->second;
  int cur = -1;
  HIPCHECK(hipGetDevice(&cur));
  if (cur != dev) HIPCHECK(hipSetDevice(dev));
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (cur != dev) HIPCHECK(hipSetDevice(cur));
  if (e != hipSuccess || ptr == nullptr)
    throw GfError(GF_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_addr[dev] = (double*)ptr;
  return (double*)ptr;
}

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  void alloc(size_t bytes) {
    if (p) (void)dw::free(p);
    p = nullptr; n = bytes;
    if (bytes) {
      hipError_t e = dw::malloc(&p, bytes);
      if (e != hipSuccess) throw GfError(GF_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
      e = dw::zero_sync(p, bytes);
      if (e != hipSuccess) throw GfError(GF_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
    }
  }
  ~DevBuf() { if (p) (void)dw::free(p); }
  template <typename X> X* as() const { return (X*)p; }
};

// Per-launch times of kernel classes 0..2 (K1, K2, K3), summed on read. Between rec(k, true) and
// rec(k, false) the class's launch carries an event (launch_events) that the runtime stamps with
// the kernel's end; a launch's interval runs from the end of the timed launch before it on the
// stream (or from a marker, mark(), where other work came in between) to its own end: the
// kernel's execution plus the short dispatch gap after its predecessor, with no marker packets
// between the timed kernels. Events recorded on the stream around each launch would add the
// queue's processing of those packets (about 4-5 us per kernel at 2048^2).
struct KernelTimer {
  struct Ivl { int k; size_t a, b; };   // class, indices of the events that open and close it
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<Ivl> ivl;
  long last = -1;                       // event closing the previous interval, -1: none yet
  int64_t launches[3] = {0, 0, 0};
  double ms[3] = {0, 0, 0};
  ~KernelTimer() { drop(); }
  void drop() {
    launch_events = LaunchEvents{};
    for (auto& e : ev) (void)hipEventDestroy(e);
    ev.clear();
    ivl.clear();
    last = -1;
  }
  hipEvent_t fresh() {
    hipEvent_t e = nullptr;
    // no system-scope release at the kernel's end: with it the end of a timed kernel waits for a
    // cache write-back that an untimed kernel does not (about 4.5 us per kernel at 2048^2)
    HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    ev.push_back(e);
    return e;
  }
  // work not timed (a copy, another kernel, the host) may precede the next timed launch
  void mark(hipStream_t s) {
    if (!on) return;
    HIPCHECK(hipEventRecord(fresh(), s));
    last = (long)ev.size() - 1;
  }
  void rec(int k, bool start, hipStream_t s) {
    if (!on) return;
    if (start) {
      if (last < 0) mark(s);
      hipEvent_t e = fresh();
      ivl.push_back(Ivl{k, (size_t)last, ev.size() - 1});
      launch_events = LaunchEvents{nullptr, e};
    } else {
      launch_events = LaunchEvents{};
      last = (long)ivl.back().b;
    }
  }
  void reset(int enable, hipStream_t s) {
    HIPCHECK(hipStreamSynchronize(s));
    drop();
    for (int k = 0; k < 3; ++k) { launches[k] = 0; ms[k] = 0; }
    on = enable != 0;
  }
  void read(int k, int64_t* n, double* t, hipStream_t s) {
    if (k < 0 || k > 2) throw GfError(GF_EINVAL, "timing class must be 0, 1 or 2");
    HIPCHECK(hipStreamSynchronize(s));
    for (auto& i : ivl) {
      float x = 0;
      // a class call that launched nothing (no columns to transform) left its event unrecorded
      if (hipEventElapsedTime(&x, ev[i.a], ev[i.b]) != hipSuccess) { (void)hipGetLastError(); continue; }
      launches[i.k] += 1; ms[i.k] += x;
    }
    drop();
    *n = launches[k]; *t = ms[k];
  }
};

struct CtxBase {
  virtual ~CtxBase() = default;
  virtual void set_state(const void* src, hipMemcpyKind kind) = 0;
  virtual void get_state(void* dst, hipMemcpyKind kind) = 0;
  virtual void set_physical(const void* f) = 0;
  virtual void set_psi(const void*) { throw GfError(GF_EUNSUPPORTED, "set_psi is a MultiLayerQG entry point"); }
  virtual void updatevars(void* out) = 0;
  virtual void calcN(void* N, void* sol, double t) = 0;
  virtual void set_forcing(const void* Fh) = 0;
  virtual void set_forcing_harmonic(const void*, double, double) { throw GfError(GF_EINVAL, "forcing is set per problem"); }
  virtual bool is_harmonic() const { return false; }
  virtual void step(int n) = 0;
  virtual int nvars() const = 0;
  virtual void timing(int enable) = 0;
  virtual void timing_read(int k, int64_t* n, double* ms) = 0;
  virtual void time_kernel(int, int, double*) {
    throw GfError(GF_EUNSUPPORTED, "kernel classes are timed per problem; address one member through gf_ctx_instance");
  }
  virtual double kernel_bytes(int k, int full) const = 0;
  virtual void sync() = 0;
  virtual int64_t snapshot(void*) { throw GfError(GF_EINVAL, "snapshots are taken per problem; address one member through gf_ctx_instance"); }
  virtual void snapshot_wait(int64_t) {}
  virtual void set_graphs(int) {}
  virtual void set_persistent(int) {
    throw GfError(GF_EINVAL, "persistent stage launches are for ensembles (gf_ctx_create_batch)");
  }
  virtual void set_schedule(const int*) { throw GfError(GF_EUNSUPPORTED, "radix schedules are set per problem"); }
  virtual void get_schedule(int*) const { throw GfError(GF_EUNSUPPORTED, "radix schedules are set per problem"); }
  virtual void set_splits(const int*) { throw GfError(GF_EUNSUPPORTED, "workgroup splits are set per problem"); }
  virtual void get_splits(int*) const { throw GfError(GF_EUNSUPPORTED, "workgroup splits are set per problem"); }
  virtual void autotune(int, double*) { throw GfError(GF_EUNSUPPORTED, "the autotuner times one problem; address a member through gf_ctx_instance"); }
  // ensembles (a plain context is an ensemble of one)
  virtual int ninstances() const { return 1; }
  virtual CtxBase* instance(int i) { return i == 0 ? this : nullptr; }
  virtual void set_clock(double tt, int64_t s) { t = tt; stepno = s; }
  virtual int diagnostic(int, double, int, double*, int) {
    throw GfError(GF_EINVAL, "diagnostics are computed per problem; address one member through gf_ctx_instance");
  }
  virtual int diagnostic_async(int, double, int, double*, int) {
    throw GfError(GF_EINVAL, "diagnostics are computed per problem; address one member through gf_ctx_instance");
  }
  virtual int diagnostic_noalloc(int, double, int, double*, int) {
    throw GfError(GF_EINVAL, "diagnostics are computed per problem; address one member through gf_ctx_instance");
  }
  virtual void set_tape_budget(size_t) { throw GfError(GF_EUNSUPPORTED, "the adjoint's tape is set per problem"); }
  virtual void adjoint_grad(const double*, int, int, double*, double*, double*) {
    throw GfError(GF_EUNSUPPORTED, "the adjoint is provided for TwoDNavierStokes FP64 ETDRK4 problems");
  }
  // a step since the state was last set produced a non-finite value (read without waiting:
  // reflects the steps that have finished)
  virtual bool nonfinite() const { return false; }
  double t = 0;
  int64_t stepno = 0;
};

// Gradients in flight in this process. The two-stream sweep of gf_adjoint_grad pays when one
// gradient alone leaves the GPU partly idle. With several at once the GPU is already full,
// and the extra streams (beyond the 4 hardware queues) wait on each other, so a gradient that
// starts while another is running keeps its sweep on one stream.
static std::atomic<int> g_adjoint_active{0};

template <typename T>
struct Ctx final : CtxBase {
  using C = cplx<T>;
  gf_problem_desc d;
  Geo<T> g;
  int nl = 1;
  hipStream_t stream = nullptr;
  long fsz = 0;        // complex elements of one blocked spectral/mixed field (one layer)
  bool eb = false;     // SLQG equivalent barotropic
  double ell2inv = 0;
  bool advect_full = false;   // K2 with background gradients (MLQG, SLQG varying U)
  bool ccplx = false;         // complex L (SLQG)
  bool has_eta = false;
  int fkx = -1, fky = -1;
  // device buffers
  DevBuf sol, s1, s2, n1, n23, racc, nout, mixed, flux, twx, twy, phys, jl, cinp;
  DevBuf coef[6], filter, eta, etah, etax, etay, Uy, Qy1, sinv, Fh, sqrt_spec;
  DevBuf noise;   // ring forcing: K3's box of this step's forcing terms
  int nbl = 0;    // rows of that box (ring_box_row)
  DevBuf prevsol, diag_part;
  DevBuf psib;               // multi-layer with drag: the bottom layer's psih, K1 -> K3
  bool prevsol_valid = false;
  // time-harmonic forcing F(t) = Fh cos(omega t + phase) (gf_set_forcing_harmonic): K3 scales
  // the table by cur_fscale, evaluated at each stage's time; last_fscale is the factor of the
  // last evaluation, which the work diagnostics use (0 before any: the reference's vars.Fh
  // starts at zero)
  bool harmonic = false;
  double f_omega = 0, f_phase = 0;
  T cur_fscale = 1;
  double last_fscale = 1;
  std::vector<double> Hlay, gprime;   // MLQG layer depths and reduced gravities
  std::vector<double> Fcpl;           // MLQG layer coupling F (n x n, row-major): S = -Krsq I + F
  KernelTimer tm;
  // non-finite flag of the state (pinned host memory the final stage of K3 writes 1 into); cleared
  // whenever the state is set
  int* nf_host = nullptr;
  bool nonfinite() const override { return *(volatile const int*)nf_host != 0; }
  void clear_nonfinite() { *(volatile int*)nf_host = 0; }

  bool own_stream = true;
  int device = 0;
  double* scratch_dev = nullptr;   // this device's copy of the static scratch (scratch_on)
  explicit Ctx(const gf_problem_desc& desc, hipStream_t shared = nullptr) : d(desc) {
    HIPCHECK(hipGetDevice(&device));
    scratch_dev = scratch_on(device);
    if (shared) { stream = shared; own_stream = false; }
    else HIPCHECK(dw::stream_create(&stream));
    nl = d.model == GF_MODEL_MLQG ? d.nlayers : 1;
    HIPCHECK(dw::host_malloc((void**)&nf_host, 64));
    clear_nonfinite();
    const int nx = d.nx, ny = d.ny, nkr = nx / 2 + 1;
    g.nx = nx; g.ny = ny; g.nkr = nkr; g.nkr2 = (nkr + BW - 1) / BW * BW; g.nlayers = nl;
    int lo_full, hi_full, lo_r, dum1, dum2;
    aliased_range(nx, nkr, d.aliased_fraction, dum1, dum2, lo_r);
    aliased_range(ny, ny / 2 + 1, d.aliased_fraction, lo_full, hi_full, dum1);
    g.kmax = std::min(lo_r, nkr);
    g.la0 = lo_full; g.la1 = hi_full;
    if (d.aliased_fraction <= 0) { g.kmax = nkr; g.la0 = g.la1 = ny; }
    // rfftfreq(nx, 2 pi/Lx * nx) multiplier, as [FF] TwoDGrid builds it
    g.kx0 = (T)((2 * M_PI / d.Lx * nx) / nx);
    g.ky0 = (T)((2 * M_PI / d.Ly * ny) / ny);
    g.inv_norm = (T)(1.0 / ((double)nx * ny));
    fsz = (long)g.nkr2 * ny;
    g.lstride = fsz;
    const size_t cb = sizeof(C) * fsz * nl;
    sol.alloc(cb); s1.alloc(cb); s2.alloc(cb); n1.alloc(cb); n23.alloc(cb); racc.alloc(cb); nout.alloc(cb);
    cinp.alloc(cb);
    mixed.alloc(sizeof(C) * fsz * 4 * nl);
    flux.alloc(sizeof(C) * fsz * 2 * nl);
    phys.alloc(sizeof(T) * (size_t)nx * ny * 4 * nl);
    jl.alloc(sizeof(C) * (size_t)nkr * ny * nl);
    twx.alloc(sizeof(C) * nx); twy.alloc(sizeof(C) * ny);
    upload_twiddles(twx, nx); upload_twiddles(twy, ny);
    build_model();
    build_stepper();
    build_forcing();
    if (d.forcing == GF_FORCING_FIXED) Fh.alloc(sizeof(C) * fsz * nl);   // zero until gf_set_forcing
    if (nl > 1 && d.mu != 0) psib.alloc(sizeof(C) * fsz);
    Sched tuned;
    if (tunable() && sched_lookup(sched_key(), tuned)) {
      for (int k = 0; k < 3; ++k) sched[k] = tuned.p[k];
      for (int k = 0; k < 3; ++k) split[k] = tuned.split[k];
    }
    HIPCHECK(hipStreamSynchronize(stream));
  }
  ~Ctx() override {
    if (stream) { (void)hipStreamSynchronize(stream); if (own_stream) (void)dw::stream_destroy(stream); }
    for (auto& ge : gexec) if (ge) (void)dw::exec_destroy(ge);
    if (rstream) {
      (void)hipStreamSynchronize(rstream);
      (void)dw::stream_destroy(rstream);
      (void)hipEventDestroy(fwd_done);
    }
    for (hipEvent_t e : adj_done) (void)hipEventDestroy(e);
    for (hipEvent_t e : rec_done) (void)hipEventDestroy(e);
    if (copy_stream) {
      (void)hipStreamSynchronize(copy_stream);
      for (int k = 0; k < 2; ++k) { (void)hipEventDestroy(snap_ready[k]); (void)hipEventDestroy(snap_done[k]); }
      (void)dw::stream_destroy(copy_stream);
    }
    if (nf_host) (void)dw::host_free(nf_host);
  }

  // ---------------- snapshots ----------------
  // The state is converted to the Julia layout on the compute stream into one of two staging
  // buffers; the device-to-host copy runs on a copy stream, so the steps queued after it
  // overlap the copy. Before a staging buffer is reused the compute stream waits for its
  // previous copy.
  hipStream_t copy_stream = nullptr;
  hipEvent_t snap_ready[2] = {}, snap_done[2] = {};
  DevBuf snap[2];
  bool snap_used[2] = {false, false};
  int64_t snap_count = 0;

  int64_t snapshot(void* dst) override {
    const size_t bytes = sizeof(C) * (size_t)g.nkr * d.ny * nl;
    if (!copy_stream) {
      HIPCHECK(dw::stream_create(&copy_stream));
      for (int k = 0; k < 2; ++k) {
        HIPCHECK(hipEventCreateWithFlags(&snap_ready[k], hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&snap_done[k], hipEventDisableTiming));
        snap[k].alloc(bytes);
      }
    }
    const int k = (int)(snap_count & 1);
    if (snap_used[k]) HIPCHECK(hipStreamWaitEvent(stream, snap_done[k], 0));
    launch_to_julia<T>(g, sol.as<C>(), snap[k].as<C>(), nl, stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(snap_ready[k], stream));
    HIPCHECK(hipStreamWaitEvent(copy_stream, snap_ready[k], 0));
    HIPCHECK(hipMemcpyAsync(dst, snap[k].p, bytes, hipMemcpyDeviceToHost, copy_stream));
    HIPCHECK(hipEventRecord(snap_done[k], copy_stream));
    snap_used[k] = true;
    return snap_count++;
  }
  // the copy stream runs in order: the event of the ticket's slot is this snapshot's or a later one's
  void snapshot_wait(int64_t ticket) override {
    if (!copy_stream) return;
    if (ticket < 0 || ticket >= snap_count) { HIPCHECK(hipStreamSynchronize(copy_stream)); return; }
    HIPCHECK(hipEventSynchronize(snap_done[ticket & 1]));
  }

  // ---------------- setup helpers ----------------
  void upload_twiddles(DevBuf& b, int n) {
    std::vector<C> h(n);
    for (int i = 0; i < n; ++i) {
      const long double a = -2.0L * 3.141592653589793238462643383279502884L * i / n;
      h[i].re = (T)std::cos(a); h[i].im = (T)std::sin(a);
    }
    HIPCHECK(dw::copy(b.p, h.data(), sizeof(C) * n, hipMemcpyHostToDevice));
  }
  double krw(int kr) const { return (double)kr * ((2 * M_PI / d.Lx * d.nx) / d.nx); }
  double lw(int l) const {
    const int ny = d.ny;
    return (double)(l < (ny + 1) / 2 ? l : l - ny) * ((2 * M_PI / d.Ly * ny) / ny);
  }
  // Krsq in the problem precision, as the reference's grid holds it
  double krsq(int kr, int l) const {
    const T a = (T)krw(kr), b = (T)lw(l);
    return (double)(T)(a * a + b * b);
  }
  // pad: zeroed bytes after the table (the stepper's coefficient tables carry 16, so that the
  // stage epilogue may read a real table's element i + 1 along with element i)
  template <typename X>
  void upload_blocked(DevBuf& b, const std::vector<X>& full /* [l][kr] */, int nlay = 1, size_t pad = 0) {
    std::vector<X> h((size_t)fsz * nlay, X{});
    for (int j = 0; j < nlay; ++j)
      for (int l = 0; l < d.ny; ++l)
        for (int kr = 0; kr < g.nkr; ++kr)
          h[(size_t)j * fsz + boff_h(kr, l, d.ny)] = full[((size_t)j * d.ny + l) * g.nkr + kr];
    b.alloc(sizeof(X) * h.size() + pad);
    HIPCHECK(dw::copy(b.p, h.data(), sizeof(X) * h.size(), hipMemcpyHostToDevice));
  }
  // device rfft2 of nf physical fields (x fastest) -> blocked spectral fields
  void fwd_device(const T* phys_dev, C* spec_dev, int nf, int zero_mean) {
    K2Args<T> a{};
    a.g = g; a.g.nlayers = 1; a.mode = K2_FORWARD_IN; a.nfields = nf; a.phys_in = phys_dev;
    a.pstride = (long)d.nx * d.ny; a.out = mixed.as<C>(); a.ostride = fsz; a.kout = g.nkr;
    a.tw = twx.as<C>();
    launch_k2<T>(d.nx, a, stream);
    K3Args<T> b{};
    b.g = g; b.g.nlayers = nf; b.model = GF_MODEL_TWODNS; b.in = mixed.as<C>(); b.fstride = fsz;
    b.ncols = g.nkr; b.full = 1; b.epi = E_STORE; b.sol = spec_dev; b.zero_mean = zero_mean;
    b.tw = twy.as<C>();
    launch_k3<T>(d.ny, b, stream);
    HIPCHECK(hipGetLastError());
  }
  // device irfft2 (normalised) of nf blocked spectral fields -> physical fields
  void inv_device(const C* spec_dev, T* phys_dev, int nf) {
    for (int f = 0; f < nf; ++f) {
      K1Args<T> a{};
      a.g = g; a.g.nlayers = 1; a.model = GF_MODEL_TWODNS; a.fieldset = FS_RAW;
      a.sol = spec_dev + (size_t)f * fsz; a.out = mixed.as<C>() + (size_t)f * fsz; a.fstride = fsz;
      a.ncols = g.nkr; a.tw = twy.as<C>();
      launch_k1<T>(d.ny, a, stream);
    }
    K2Args<T> b{};
    b.g = g; b.g.nlayers = 1; b.mode = K2_INVERSE_OUT; b.nfields = nf; b.in = mixed.as<C>();
    b.fstride = fsz; b.kin = g.nkr; b.phys = phys_dev; b.pstride = (long)d.nx * d.ny; b.tw = twx.as<C>();
    launch_k2<T>(d.nx, b, stream);
    HIPCHECK(hipGetLastError());
  }
  // spectral derivative / multiplier of physical fields, computed with the device transforms
  std::vector<double> spectral_op(const std::vector<double>& f /* ny*nx */, int op) {
    const size_t np = (size_t)d.nx * d.ny;
    std::vector<T> ft(np);
    for (size_t i = 0; i < np; ++i) ft[i] = (T)f[i];
    HIPCHECK(dw::copy(phys.p, ft.data(), sizeof(T) * np, hipMemcpyHostToDevice));
    fwd_device(phys.as<T>(), nout.as<C>(), 1, 0);
    std::vector<C> h(fsz);
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(dw::copy(h.data(), nout.p, sizeof(C) * fsz, hipMemcpyDeviceToHost));
    for (int l = 0; l < d.ny; ++l)
      for (int kr = 0; kr < g.nkr; ++kr) {
        C& z = h[boff_h(kr, l, d.ny)];
        const T k = (T)krw(kr), m = (T)lw(l);
        C r;
        if (op == 0) { r.re = -k * z.im; r.im = k * z.re; }          // i kr
        else if (op == 1) { r.re = -m * z.im; r.im = m * z.re; }     // i l
        else { const T m2 = -(m * m); r.re = m2 * z.re; r.im = m2 * z.im; }  // -l^2
        z = r;
      }
    HIPCHECK(dw::copy(nout.p, h.data(), sizeof(C) * fsz, hipMemcpyHostToDevice));
    inv_device(nout.as<C>(), phys.as<T>(), 1);
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(dw::copy(ft.data(), phys.p, sizeof(T) * np, hipMemcpyDeviceToHost));
    std::vector<double> out(np);
    for (size_t i = 0; i < np; ++i) out[i] = (double)ft[i];
    return out;
  }
  void upload_real(DevBuf& b, const std::vector<double>& v) {
    std::vector<T> h(v.size());
    for (size_t i = 0; i < v.size(); ++i) h[i] = (T)v[i];
    b.alloc(sizeof(T) * h.size());
    HIPCHECK(dw::copy(b.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  }

  std::vector<cd> Lop;   // [l][kr], one layer (MLQG L is the same in every layer)

  void build_model() {
    const int nx = d.nx, ny = d.ny, nkr = g.nkr;
    const size_t np = (size_t)nx * ny;
    Lop.assign((size_t)ny * nkr, cd(0));
    // Krsq^n with an integer exponent (pow, correctly rounded to within an ulp)
    auto ipow = [](double x, int n) { return n == 0 ? 1.0 : std::pow(x, (double)n); };
    if (d.model == GF_MODEL_TWODNS) {
      // L = -nu Krsq^nnu - mu Krsq^nmu, L[1,1] = 0 (src/twodnavierstokes.jl:150-151)
      for (int l = 0; l < ny; ++l)
        for (int kr = 0; kr < nkr; ++kr) {
          const double K2 = krsq(kr, l);
          const double pm = d.nmu >= 0 ? ipow(K2, d.nmu) : (K2 == 0 ? INFINITY : 1.0 / ipow(K2, -d.nmu));
          Lop[(size_t)l * nkr + kr] = cd(-d.nu * ipow(K2, d.nnu) - d.mu * pm, 0);
        }
      Lop[0] = 0;
      return;
    }
    if (d.eta) {
      has_eta = false;
      for (size_t i = 0; i < np; ++i) if (d.eta[i] != 0) { has_eta = true; break; }
    }
    std::vector<double> etav(np, 0.0);
    if (d.eta) for (size_t i = 0; i < np; ++i) etav[i] = d.eta[i];
    if (d.model == GF_MODEL_SLQG) {
      eb = d.deformation_radius > 0 && std::isfinite(d.deformation_radius);
      ell2inv = eb ? 1.0 / (d.deformation_radius * d.deformation_radius) : 0.0;
      const bool varyingU = d.Uy != nullptr;
      ccplx = true;
      // L = -mu - nu Krsq^nnu + i beta kr/(Krsq [+ 1/ell^2]) [- i U kr] (src/singlelayerqg.jl:230-252)
      for (int l = 0; l < ny; ++l)
        for (int kr = 0; kr < nkr; ++kr) {
          const double K2 = krsq(kr, l), k = (double)(T)krw(kr);
          const double iK = (kr == 0 && l == 0) ? 0.0 : 1.0 / K2;
          const double bterm = eb ? d.beta * k / (K2 + ell2inv) : d.beta * k * iK;
          cd L(-d.mu - d.nu * ipow(K2, d.nnu), bterm);
          if (!varyingU) L -= cd(0, d.U * k);
          Lop[(size_t)l * nkr + kr] = L;
        }
      Lop[0] = 0;
      if (has_eta) {
        // etah = rfft(eta): the -i kr U etah term and the enstrophy (src/singlelayerqg.jl:190)
        std::vector<T> ft(np);
        for (size_t i = 0; i < np; ++i) ft[i] = (T)etav[i];
        HIPCHECK(dw::copy(phys.p, ft.data(), sizeof(T) * np, hipMemcpyHostToDevice));
        etah.alloc(sizeof(C) * fsz);
        fwd_device(phys.as<T>(), etah.as<C>(), 1, 0);
        HIPCHECK(hipStreamSynchronize(stream));
      }
      if (!varyingU) {
        if (has_eta) upload_real(eta, etav);
      } else {
        // varying U(y): Qx = d eta/dx, Qy = d eta/dy - U_yy (src/singlelayerqg.jl:182-189)
        advect_full = true;
        std::vector<double> U2(np);
        for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x) U2[(size_t)y * nx + x] = d.Uy[y];
        std::vector<double> Uyy = spectral_op(U2, 2);
        std::vector<double> ex = spectral_op(etav, 0), ey = spectral_op(etav, 1);
        for (size_t i = 0; i < np; ++i) ey[i] -= Uyy[i];
        upload_real(etax, ex); upload_real(etay, ey);
        upload_real(Uy, std::vector<double>(d.Uy, d.Uy + ny));
        upload_real(Qy1, std::vector<double>(ny, 0.0));
      }
      return;
    }
    // MultiLayerQG (src/multilayerqg.jl:255-320, 338-347, 603-634)
    advect_full = true;
    const int n = nl;
    for (int l = 0; l < ny; ++l)
      for (int kr = 0; kr < nkr; ++kr) Lop[(size_t)l * nkr + kr] = cd(-d.nu * ipow(krsq(kr, l), d.nnu), 0);
    Lop[0] = 0;
    // one layer (SingleLayerParams, src/multilayerqg.jl:193-219): no H, b or coupling
    std::vector<double> H = n > 1 ? std::vector<double>(d.H, d.H + n) : std::vector<double>{1.0};
    std::vector<double> b = n > 1 ? std::vector<double>(d.b, d.b + n) : std::vector<double>{0.0};
    std::vector<double> gp(n - 1), Fm(n - 1), Fp(n - 1);
    Hlay = H;
    for (int j = 0; j < n - 1; ++j) {
      gp[j] = b[j] - b[j + 1];
      Fm[j] = d.f0 * d.f0 / (gp[j] * H[j + 1]);
      Fp[j] = d.f0 * d.f0 / (gp[j] * H[j]);
    }
    gprime = gp;
    std::vector<double> F((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
      F[j * n + j] = -((j < n - 1 ? Fp[j] : 0.0) + (j > 0 ? Fm[j - 1] : 0.0));
      if (j < n - 1) F[j * n + j + 1] = Fp[j];
      if (j > 0) F[j * n + j - 1] = Fm[j - 1];
    }
    Fcpl = F;
    // S^-1 = I / (-k^2 I + F), k^2 = 0 -> 1, S^-1[1,1] = 0
    std::vector<double> S((size_t)fsz * n * n, 0.0);
    for (int l = 0; l < ny; ++l)
      for (int kr = 0; kr < nkr; ++kr) {
        double k2 = krsq(kr, l);
        if (k2 == 0) k2 = 1;
        std::vector<double> A((size_t)n * n), Inv((size_t)n * n, 0.0);
        for (int i = 0; i < n * n; ++i) A[i] = F[i];
        for (int i = 0; i < n; ++i) A[i * n + i] -= k2, Inv[i * n + i] = 1.0;
        // Gauss-Jordan with partial pivoting
        for (int c = 0; c < n; ++c) {
          int p = c;
          for (int r = c + 1; r < n; ++r) if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
          if (p != c) for (int k = 0; k < n; ++k) { std::swap(A[c * n + k], A[p * n + k]); std::swap(Inv[c * n + k], Inv[p * n + k]); }
          const double piv = A[c * n + c];
          for (int k = 0; k < n; ++k) { A[c * n + k] /= piv; Inv[c * n + k] /= piv; }
          for (int r = 0; r < n; ++r) if (r != c) {
            const double fct = A[r * n + c];
            for (int k = 0; k < n; ++k) { A[r * n + k] -= fct * A[c * n + k]; Inv[r * n + k] -= fct * Inv[c * n + k]; }
          }
        }
        if (kr == 0 && l == 0) std::fill(Inv.begin(), Inv.end(), 0.0);
        const long o = boff_h(kr, l, ny);
        for (int i = 0; i < n * n; ++i) S[(size_t)o * n * n + i] = Inv[i];
      }
    upload_real(sinv, S);
    // U(y) per layer, U_yy, Qy (src/multilayerqg.jl:271-316)
    std::vector<double> U((size_t)n * ny, 0.0);
    if (d.Uy) for (int i = 0; i < n * ny; ++i) U[i] = d.Uy[i];
    std::vector<double> Qy((size_t)n * ny);
    std::vector<double> Uyy((size_t)n * ny);
    for (int j = 0; j < n; ++j) {
      std::vector<double> U2(np);
      for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x) U2[(size_t)y * nx + x] = U[(size_t)j * ny + y];
      std::vector<double> uyy = spectral_op(U2, 2);
      for (int y = 0; y < ny; ++y) Uyy[(size_t)j * ny + y] = uyy[(size_t)y * nx];
    }
    for (int j = 0; j < n; ++j)
      for (int y = 0; y < ny; ++y) {
        double q = d.beta - Uyy[(size_t)j * ny + y];
        auto Uj = [&](int jj) { return U[(size_t)jj * ny + y]; };
        if (n == 1) {}
        else if (j == 0) q -= Fp[0] * (Uj(1) - Uj(0));
        else if (j == n - 1) q -= Fm[n - 2] * (Uj(n - 2) - Uj(n - 1));
        else q = q - Fp[j] * (Uj(j + 1) - Uj(j)) - Fm[j - 1] * (Uj(j - 1) - Uj(j));
        Qy[(size_t)j * ny + y] = q;
      }
    upload_real(Uy, U);
    upload_real(Qy1, Qy);
    std::vector<double> ex = has_eta ? spectral_op(etav, 0) : std::vector<double>(np, 0.0);
    std::vector<double> ey = has_eta ? spectral_op(etav, 1) : std::vector<double>(np, 0.0);
    for (size_t i = 0; i < np; ++i) { ex[i] += d.topo_grad_x; ey[i] += d.topo_grad_y; }
    upload_real(etax, ex); upload_real(etay, ey);
  }

  bool filtered() const {
    return d.stepper == GF_FILTERED_ETDRK4 || d.stepper == GF_FILTERED_RK4 || d.stepper == GF_FILTERED_FORWARD_EULER ||
           d.stepper == GF_FILTERED_AB3;
  }
  bool ab3() const { return d.stepper == GF_AB3 || d.stepper == GF_FILTERED_AB3; }
  bool etd() const { return d.stepper == GF_ETDRK4 || d.stepper == GF_FILTERED_ETDRK4; }

  void build_stepper() {
    const int ny = d.ny, nkr = g.nkr;
    const size_t nm = (size_t)ny * nkr;
    auto up = [&](DevBuf& b, const std::vector<cd>& v) {
      if (ccplx) {
        std::vector<C> h(nm);
        for (size_t i = 0; i < nm; ++i) { h[i].re = (T)v[i].real(); h[i].im = (T)v[i].imag(); }
        upload_blocked(b, h, 1, 16);
      } else {
        std::vector<T> h(nm);
        for (size_t i = 0; i < nm; ++i) h[i] = (T)v[i].real();
        upload_blocked(b, h, 1, 16);
      }
    };
    if (etd()) {
      std::vector<cd> E(nm), E2(nm), Z(nm), A(nm), B(nm), G(nm);
      const double dt = d.dt;
      parallel_for((long)nm, [&](long i) {
        const cd L = Lop[i];
        E[i] = std::exp(dt * L);
        E2[i] = std::exp(dt * L / 2.0);
        etd_coeffs(dt * L, dt, Z[i], A[i], B[i], G[i]);
        if (!ccplx) { Z[i] = Z[i].real(); A[i] = A[i].real(); B[i] = B[i].real(); G[i] = G[i].real(); }
      });
      up(coef[0], E2); up(coef[1], Z); up(coef[2], E); up(coef[3], A); up(coef[4], B); up(coef[5], G);
    } else {
      up(coef[0], Lop);
    }
    if (filtered()) {
      // [FF] makefilter: K = sqrt((kr dx/pi)^2 + (l dy/pi)^2), exp(-decay (K - 0.65)^4) for K >= 0.65
      const double dx = d.Lx / d.nx, dy = d.Ly / d.ny;
      const double decay = -std::log(1e-15) / std::pow(1 - 0.65, 4);
      std::vector<T> h(nm);
      for (int l = 0; l < ny; ++l)
        for (int kr = 0; kr < nkr; ++kr) {
          const double a = (double)(T)krw(kr) * dx / M_PI, b = (double)(T)lw(l) * dy / M_PI;
          const double K = std::sqrt(a * a + b * b);
          h[(size_t)l * nkr + kr] = (T)(K < 0.65 ? 1.0 : std::exp(-decay * std::pow(K - 0.65, 4)));
        }
      upload_blocked(filter, h);
    }
  }

  void build_forcing() {
    if (d.forcing != GF_FORCING_RING) return;
    // ring spectrum (examples/twodnavierstokes_stochasticforcing.jl:53-66)
    const int ny = d.ny, nkr = g.nkr;
    std::vector<double> spec((size_t)ny * nkr);
    double U = 0;
    for (int l = 0; l < ny; ++l)
      for (int kr = 0; kr < nkr; ++kr) {
        const double K2 = krsq(kr, l), K = std::sqrt(K2);
        double s = std::exp(-(K - d.kf) * (K - d.kf) / (2 * d.dkf * d.dkf));
        if (K2 == 0) s = 0;
        spec[(size_t)l * nkr + kr] = s;
        const double iK = K2 == 0 ? 0 : 1.0 / K2;
        U += (kr == 0 ? 1.0 : 2.0) * s * iK / 2;
      }
    const double eps0 = U * d.Lx * d.Ly / ((double)d.nx * d.nx * (double)ny * ny) / (d.Lx * d.Ly);
    std::vector<T> ss((size_t)ny * nkr);
    fkx = -1; fky = -1;
    for (int l = 0; l < ny; ++l)
      for (int kr = 0; kr < nkr; ++kr) {
        const double s = spec[(size_t)l * nkr + kr] * (d.eps / eps0);
        ss[(size_t)l * nkr + kr] = (T)std::sqrt(s);
        if (s > 0) {
          fkx = std::max(fkx, kr);
          fky = std::max(fky, l < (ny + 1) / 2 ? l : ny - l);
        }
      }
    upload_blocked(sqrt_spec, ss);
    if (fkx >= 0) {
      nbl = std::min(ny, 2 * fky + 1);
      noise.alloc(sizeof(C) * (size_t)(fkx + 1) * nbl * nl);   // one slice per layer row (lsplit)
    }
  }

  // ---------------- hot path ----------------
  void rec(int k, bool start) { tm.rec(k, start, stream); }

  // argument blocks of the three kernels for one calcN! evaluation with stage epilogue epi;
  // ab3_start >= 0 overrides the AB3 start-up flag the clock gives (ensemble tables)
  void make_args(const C* in, int epi, bool full, K1Args<T>& a, K2Args<T>& b, K3Args<T>& c, int ab3_start = -1) const {
    a = K1Args<T>{};
    a.g = g; a.model = d.model; a.fieldset = FS_CALCN; a.sol = in; a.out = mixed.as<C>(); a.fstride = fsz;
    a.ncols = g.kmax; a.eb = eb; a.ell2inv = (T)ell2inv; a.sinv = sinv.as<T>(); a.tw = twy.as<C>();
    b = K2Args<T>{};
    b.g = g; b.mode = advect_full ? K2_ADVECT_FULL : K2_ADVECT; b.in = mixed.as<C>(); b.fstride = fsz;
    b.kin = g.kmax; b.eta = (d.model == GF_MODEL_SLQG && !advect_full && has_eta) ? eta.as<T>() : nullptr;
    b.Uy = Uy.as<T>(); b.Qy1 = Qy1.as<T>(); b.etax = etax.as<T>(); b.etay = etay.as<T>(); b.has_bg = advect_full;
    b.out = flux.as<C>(); b.ostride = fsz; b.kout = full ? g.nkr : g.kmax; b.tw = twx.as<C>();
    c = K3Args<T>{};
    c.g = g; c.model = d.model; c.in = flux.as<C>(); c.fstride = fsz; c.ncols = full ? g.nkr : g.kmax;
    c.full = full; c.epi = epi; c.apply_kr = !advect_full;
    c.etah = (d.model == GF_MODEL_SLQG && !advect_full && d.U != 0) ? etah.as<C>() : nullptr;
    c.Ueta = (T)d.U;
    c.Fh = (d.forcing == GF_FORCING_FIXED) ? Fh.as<C>() : nullptr;
    c.fscale = cur_fscale;
    c.sqrt_spec = (d.forcing == GF_FORCING_RING && fkx >= 0) ? sqrt_spec.as<T>() : nullptr;
    c.seed = d.seed; c.fstep = stepno; c.sqrt_dt = (T)std::sqrt(d.dt); c.fkx = fkx; c.fky = fky;
    c.noise = noise.as<C>(); c.nbl = nbl;
    c.instance = d.instance;
    if (d.model == GF_MODEL_MLQG && d.mu != 0) { c.mu_drag = (T)d.mu; c.drag_src = in; c.sinv = sinv.as<T>(); }
    if (nl > 1) {
      // one layer per workgroup row: three times the workgroups of the layer loop (at 512^2 the
      // x pass had one 64-thread workgroup per row pair); K1 hands the drag's psih to K3
      a.lsplit = b.lsplit = c.lsplit = 1;
      a.psi_out = psib.as<C>();
      c.psi_in = psib.as<C>();
    }
    c.sol = sol.as<C>(); c.s1 = s1.as<C>(); c.s2 = s2.as<C>(); c.n1 = n1.as<C>(); c.n23 = n23.as<C>();
    c.racc = racc.as<C>(); c.nout = nout.as<C>();
    c.c0 = coef[0].p; c.c1 = coef[1].p; c.c2 = coef[2].p; c.c3 = coef[3].p; c.c4 = coef[4].p; c.c5 = coef[5].p;
    c.ccplx = ccplx; c.cl_layered = 0; c.dt = (T)d.dt; c.filter = filtered() ? filter.as<T>() : nullptr;
    c.nonfinite = nf_host;
    c.tw = twy.as<C>();
    a.p = sched[0]; b.p = sched[1]; c.p = sched[2];
    if (tunable()) { a.fsplit = split[0]; b.hsplit = split[1]; c.csplit = split[2]; }
    if (ab3()) {
      // the RHS history of the aliased band enters later steps: every AB3 step is computed in full
      c.ab3_start = ab3_start >= 0 ? ab3_start : stepno < 3;
      b.kout = g.nkr; c.ncols = g.nkr; c.full = 1;
    }
    if (d.model == GF_MODEL_MLQG && d.linear) {
      // calcNlinear! works on the whole spectrum: no dealiasing of its input, every mode stepped
      a.ncols = g.nkr; a.nomask = 1;
      b.kin = g.nkr; b.kout = g.nkr; b.linear = 1;
      c.ncols = g.nkr; c.full = 1; c.nomask = 1;
    }
  }

  void calcN_launch(const C* in, int epi, bool full, const int64_t* fstep_dev = nullptr) {
    K1Args<T> a; K2Args<T> b; K3Args<T> c;
    make_args(in, epi, full, a, b, c);
    c.fstep_dev = fstep_dev;
    rec(0, true); launch_k1<T>(d.ny, a, stream); rec(0, false);
    rec(1, true); launch_k2<T>(d.nx, b, stream); rec(1, false);
    rec(2, true); launch_k3<T>(d.ny, c, stream); rec(2, false);
    HIPCHECK(hipGetLastError());
  }

  // the stages of one step: (calcN input, epilogue) pairs
  int stages(const C* ins[4], int epis[4]) const {
    if (etd()) {
      ins[0] = sol.as<C>(); ins[1] = s1.as<C>(); ins[2] = s2.as<C>(); ins[3] = s2.as<C>();
      epis[0] = E_ETD1; epis[1] = E_ETD2; epis[2] = E_ETD3; epis[3] = E_ETD4;
      return 4;
    }
    if (d.stepper == GF_RK4 || d.stepper == GF_FILTERED_RK4) {
      ins[0] = sol.as<C>(); ins[1] = s1.as<C>(); ins[2] = s1.as<C>(); ins[3] = s1.as<C>();
      epis[0] = E_RK1; epis[1] = E_RK2; epis[2] = E_RK3; epis[3] = E_RK4;
      return 4;
    }
    ins[0] = sol.as<C>(); epis[0] = ab3() ? E_AB3 : E_FE;
    return 1;
  }

  // ---------------- radix schedules and the autotuner ----------------
  // sched[k]: FFT points per thread of K1, K2, K3 (0: the built-in choice); single-layer models.
  int sched[3] = {0, 0, 0};
  // split[0]: K1 transforms one field per workgroup row; split[1]: K2 forms u q and v q in two
  // workgroup rows (1): more, shorter workgroups for grids too small to fill the GPU, or walks row
  // pairs with each field's rows staged through LDS one field ahead (2, k2_staged_kernel; where
  // k2_staged_offered does not hold the unsplit pass runs). split[2]: K3
  // takes one column per workgroup instead of a column pair, where a column fills a wave
  // (k3_single_column): 683 workgroups at 4 per CU balance 256 CUs better than 342 at 2.
  int split[3] = {0, 0, 0};
  // TwoDNavierStokes and SingleLayerQG take radix schedules and workgroup splits; MultiLayerQG
  // (any number of layers) has a fixed schedule
  // (a grid length that is not a power of two keeps the built-in schedule too)
  bool tunable() const { return d.model != GF_MODEL_MLQG && is_pow2(d.nx) && is_pow2(d.ny); }
  // the staged x pass: advection without background gradients, whole waves per workgroup
  bool k2_staged_offered(int p) const { return nl == 1 && BW == 2 && !advect_full && d.nx / p >= 64; }
  void set_splits(const int* v) override {
    if (!tunable() && (v[0] || v[1] || v[2]))
      throw GfError(GF_EINVAL, "workgroup splits are for TwoDNavierStokes and SingleLayerQG problems with nx and ny powers of two");
    if (v[1] < 0 || v[1] > 2) throw GfError(GF_EINVAL, "K2 split must be 0, 1 (halves) or 2 (staged loads)");
    bool changed = false;
    for (int k = 0; k < 3; ++k) {
      const int x = k == 1 ? v[k] : (v[k] != 0);
      changed |= split[k] != x;
      split[k] = x;
    }
    if (changed)
      for (auto& ge : gexec) if (ge) { HIPCHECK(hipStreamSynchronize(stream)); (void)dw::exec_destroy(ge); ge = nullptr; }
  }
  void get_splits(int* v) const override { for (int k = 0; k < 3; ++k) v[k] = split[k]; }
  SchedKey sched_key() const { return SchedKey{d.model, nl, d.nx, d.ny, d.dtype}; }
  // K2 transforms rows, K1/K3 columns (in the kernels of the next power of two where ny is not one;
  // a row length that is not one runs in the x pass of direct sums, which has no schedule)
  int sched_len(int k) const { return k == 1 ? pow2_ceil(d.nx) : pow2_ceil(d.ny); }
  void set_schedule(const int* p) override {
    for (int k = 0; k < 3; ++k)
      if (p[k] != 0 && !(tunable() && sched_candidate(sched_len(k), p[k])))
        throw GfError(GF_EINVAL, "radix schedule: P = " + std::to_string(p[k]) + " is not offered for kernel " +
                                 std::to_string(k + 1) + " of this problem");
    bool changed = false;
    for (int k = 0; k < 3; ++k) { changed |= sched[k] != p[k]; sched[k] = p[k]; }
    if (changed)   // the captured steps hold the old launches
      for (auto& ge : gexec) if (ge) { HIPCHECK(hipStreamSynchronize(stream)); (void)dw::exec_destroy(ge); ge = nullptr; }
  }
  void get_schedule(int* p) const override {
    for (int k = 0; k < 3; ++k) p[k] = sched[k] ? sched[k] : (nl == 1 ? sched_default(sched_len(k)) : (d.ny >= 32 && k != 1 ? (d.ny >= 4096 ? 8 : 4) : sched_default(sched_len(k))));
  }
  // Times whole steps under every offered schedule -- P in {4, 8, 16} for each of K1, K2, K3,
  // the K1 and K3 splits on or off and the K2 split 0, 1 or 2, up to 324 combinations -- in two
  // passes (the second in reverse order, best of the two), keeps the fastest and records it for
  // this (model, size, dtype) so that later contexts start with it. The state, its stepper
  // history and the clock are restored afterwards. us (nullable) receives 324 step times in
  // microseconds, combination c = ((((i1 3 + i2) 3 + i3) 2 + split_K1) 3 + split_K2) 2 + split_K3
  // with i = 0, 1, 2 for P = 4, 8, 16, NaN where a combination is not offered (split_K3 = 1 where
  // k3_single_column holds for P_K3, split_K2 = 2 where k2_staged_offered holds for P_K2).
  void autotune(int steps, double* us) override {
    if (steps < 1) throw GfError(GF_EINVAL, "gf_autotune: steps must be >= 1");
    constexpr int NC = 324;
    if (us) for (int j = 0; j < NC; ++j) us[j] = std::nan("");
    if (!tunable()) return;   // MultiLayerQG keeps its fixed schedule
    HIPCHECK(hipStreamSynchronize(stream));
    const size_t cb = sizeof(C) * fsz * nl;
    DevBuf save[4];
    DevBuf* keep[4] = {&sol, &racc, &n1, &prevsol};
    for (int j = 0; j < 4; ++j)
      if (keep[j]->p) { save[j].alloc(cb); HIPCHECK(dw::copy(save[j].p, keep[j]->p, cb, hipMemcpyDeviceToDevice)); }
    const double t0 = t;
    const int64_t step0 = stepno;
    const bool pv0 = prevsol_valid;
    const int nf0 = *nf_host;
    const int old[3] = {sched[0], sched[1], sched[2]};
    // plain launches while timing: a schedule change drops the captured graphs, and capturing
    // again inside a timed block would be charged to the schedule
    const bool old_graphs = graphs;
    graphs = false;
    const int old_split[3] = {split[0], split[1], split[2]};
    static constexpr int PV[3] = {4, 8, 16};
    auto decode = [](int c, int* q, int* v) {
      v[2] = c % 2; v[1] = (c / 2) % 3; v[0] = (c / 6) % 2;
      int r = c / 12;
      q[2] = PV[r % 3]; r /= 3;
      q[1] = PV[r % 3]; r /= 3;
      q[0] = PV[r % 3];
    };
    auto apply = [&](int c) {
      int q[3], v[3];
      decode(c, q, v);
      set_schedule(q);
      set_splits(v);
    };
    std::vector<int> combos;
    for (int c = 0; c < NC; ++c) {
      int q[3], v[3];
      decode(c, q, v);
      bool ok = true;
      for (int k = 0; k < 3; ++k) ok &= sched_candidate(sched_len(k), q[k]);
      ok &= !v[2] || k3_single_column(sched_len(2), q[2]);
      ok &= v[1] != 2 || k2_staged_offered(q[1]);
      if (ok) combos.push_back(c);
    }
    double best_t[NC];
    for (double& b : best_t) b = INFINITY;
    hipEvent_t ev[2] = {nullptr, nullptr};
    auto cleanup = [&] {
      for (auto& e : ev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
    };
    try {
      HIPCHECK(hipEventCreate(&ev[0]));
      HIPCHECK(hipEventCreate(&ev[1]));
      for (int pass = 0; pass < 2; ++pass) {
        for (size_t i = 0; i < combos.size(); ++i) {
          const int c = combos[pass ? combos.size() - 1 - i : i];
          apply(c);
          step(2);
          HIPCHECK(hipEventRecord(ev[0], stream));
          step(steps);
          HIPCHECK(hipEventRecord(ev[1], stream));
          HIPCHECK(hipEventSynchronize(ev[1]));
          float ms = 0;
          HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
          best_t[c] = std::min(best_t[c], 1e3 * (double)ms / steps);
        }
      }
    } catch (...) {
      cleanup();
      set_schedule(old);
      set_splits(old_split);
      graphs = old_graphs;
      throw;
    }
    cleanup();
    graphs = old_graphs;
    for (int j = 0; j < 4; ++j)
      if (save[j].p) HIPCHECK(dw::copy(keep[j]->p, save[j].p, cb, hipMemcpyDeviceToDevice));
    t = t0; stepno = step0; prevsol_valid = pv0;
    *(volatile int*)nf_host = nf0;
    int bc = -1;
    for (int c : combos) {
      if (us) us[c] = best_t[c];
      if (bc < 0 || best_t[c] < best_t[bc]) bc = c;
    }
    if (bc < 0) { set_schedule(old); set_splits(old_split); return; }
    apply(bc);
    Sched e;
    for (int k = 0; k < 3; ++k) e.p[k] = sched[k];
    for (int k = 0; k < 3; ++k) e.split[k] = split[k];
    e.step_us = best_t[bc];
    sched_store(sched_key(), e);
  }

  // ---------------- graphs ----------------
  // Every step of a gf_step call but the last has the same launches and arguments (the last
  // also computes the aliased band). That step is captured once into a HIP graph and replayed:
  // one graph launch per step instead of 3 per stage. The ring forcing's step number, the
  // one argument that changes, is read from step_dev, which the graph advances itself.
  static constexpr int GRAPH_BLOCK = 8;   // interior steps in the larger graph
  // Off unless gf_use_graphs(ctx, 1): for one problem the replay gains 1-3% up to 512^2 and
  // loses 0.7% at 2048^2 (the steps are bound by kernel time, not by launches).
  bool graphs = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};   // one interior step, GRAPH_BLOCK interior steps
  DevBuf step_dev;

  bool graph_usable() const { return graphs && !tm.on && !(ab3() && stepno < 3) && !harmonic; }
  // stage times of the configured stepper, as fractions of dt after the step's start
  // ([FF] ETDRK4 and RK4: t, t + dt/2, t + dt/2, t + dt; ForwardEuler and AB3: t)
  static double stage_offset(int k, int ns) { return ns == 1 ? 0.0 : (k == 0 ? 0.0 : (k == 3 ? 1.0 : 0.5)); }
  void interior_launches(const int64_t* fstep_dev) {
    const C* ins[4];
    int epis[4];
    const int ns = stages(ins, epis);
    for (int k = 0; k < ns; ++k) calcN_launch(ins[k], epis[k], false, fstep_dev);
    if (fstep_dev) launch_counter(step_dev.as<int64_t>(), 1, 1, stream);
  }
  void build_graph(int which) {
    const bool ring = d.forcing == GF_FORCING_RING;
    if (ring && !step_dev.p) step_dev.alloc(sizeof(int64_t));
    gexec[which] = dw::capture(stream, [&] {
      for (int s = 0; s < (which ? GRAPH_BLOCK : 1); ++s) interior_launches(ring ? step_dev.as<int64_t>() : nullptr);
    });
  }
  void set_graphs(int enable) override { graphs = enable != 0; }

  void step(int n) override {
    const C* ins[4];
    int epis[4];
    const int ns = stages(ins, epis);
    bool counter_set = false;
    tm.mark(stream);
    for (int s = 0; s < n;) {
      const bool full = s == n - 1;
      // the work diagnostics of stochastic forcing use the state at the start of the step
      // (prevsol, src/twodnavierstokes.jl:311-319); kept for the last step of the call
      if (full && d.forcing == GF_FORCING_RING) {
        if (!prevsol.p) prevsol.alloc(sizeof(C) * fsz * nl);
        HIPCHECK(hipMemcpyAsync(prevsol.p, sol.p, sizeof(C) * fsz * nl, hipMemcpyDeviceToDevice, stream));
        prevsol_valid = true;
        tm.mark(stream);
      }
      int done = 1;
      if (!full && graph_usable()) {
        const int which = (n - 1 - s >= GRAPH_BLOCK) ? 1 : 0;
        if (!gexec[which]) build_graph(which);
        if (d.forcing == GF_FORCING_RING && !counter_set) {
          launch_counter(step_dev.as<int64_t>(), stepno, 0, stream);
          counter_set = true;
        }
        HIPCHECK(hipGraphLaunch(gexec[which], stream));
        done = which ? GRAPH_BLOCK : 1;
      } else {
        for (int k = 0; k < ns; ++k) {
          if (harmonic) {
            cur_fscale = (T)std::cos(f_omega * (t + stage_offset(k, ns) * d.dt) + f_phase);
            last_fscale = (double)cur_fscale;
          }
          calcN_launch(ins[k], epis[k], full);
        }
        cur_fscale = 1;
        counter_set = false;
      }
      s += done;
      for (int j = 0; j < done; ++j) t += d.dt;   // the clock advances one dt at a time, as [FF]'s does
      stepno += done;
    }
  }

  void set_state(const void* src, hipMemcpyKind kind) override {
    const size_t bytes = sizeof(C) * (size_t)g.nkr * d.ny * nl;
    if (kind == hipMemcpyHostToDevice) {
      HIPCHECK(hipMemcpyAsync(jl.p, src, bytes, hipMemcpyHostToDevice, stream));
    } else {
      HIPCHECK(hipMemcpyAsync(jl.p, src, bytes, hipMemcpyDeviceToDevice, stream));
    }
    launch_from_julia<T>(g, jl.as<C>(), sol.as<C>(), nl, stream);
    HIPCHECK(hipStreamSynchronize(stream));
    clear_nonfinite();
  }
  void get_state(void* dst, hipMemcpyKind kind) override {
    const size_t bytes = sizeof(C) * (size_t)g.nkr * d.ny * nl;
    launch_to_julia<T>(g, sol.as<C>(), jl.as<C>(), nl, stream);
    HIPCHECK(hipMemcpyAsync(dst, jl.p, bytes, kind == hipMemcpyDeviceToHost ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  void set_physical(const void* f) override {
    const size_t np = (size_t)d.nx * d.ny * nl;
    HIPCHECK(hipMemcpyAsync(phys.p, f, sizeof(T) * np, hipMemcpyHostToDevice, stream));
    fwd_device(phys.as<T>(), sol.as<C>(), nl, d.model != GF_MODEL_SLQG);
    launch_mask<T>(g, sol.as<C>(), nl, stream);   // updatevars!' dealias!
    HIPCHECK(hipStreamSynchronize(stream));
    clear_nonfinite();
  }
  // set_psi! (src/multilayerqg.jl:873-889): psih = rfft(psi); qh = S psih (pvfromstreamfunction!,
  // 520-552); q = irfft(qh); then set_q!(q): rfft, qh[1, 1, :] = 0, sol = qh, updatevars!'s
  // dealias!.  s1 and s2 (stage scratch between steps) hold psih and qh.
  void set_psi(const void* f) override {
    if (d.model != GF_MODEL_MLQG) throw GfError(GF_EUNSUPPORTED, "set_psi is a MultiLayerQG entry point");
    const size_t np = (size_t)d.nx * d.ny * nl;
    HIPCHECK(hipMemcpyAsync(phys.p, f, sizeof(T) * np, hipMemcpyHostToDevice, stream));
    fwd_device(phys.as<T>(), s1.as<C>(), nl, 0);
    PvArgs<T> a{};
    a.g = g; a.nl = nl; a.psi = s1.as<C>(); a.q = s2.as<C>();
    for (int j = 0; j < nl; ++j)
      for (int k = 0; k < nl; ++k) a.F[j * MAXL + k] = (T)Fcpl[(size_t)j * nl + k];
    launch_pv_from_psi<T>(a, stream);
    inv_device(s2.as<C>(), phys.as<T>(), nl);
    fwd_device(phys.as<T>(), sol.as<C>(), nl, 1);
    launch_mask<T>(g, sol.as<C>(), nl, stream);
    HIPCHECK(hipStreamSynchronize(stream));
    clear_nonfinite();
  }
  int nvars() const override { return d.model == GF_MODEL_TWODNS ? 3 : 4; }
  void updatevars(void* out) override {
    launch_mask<T>(g, sol.as<C>(), nl, stream);
    K1Args<T> a{};
    a.g = g; a.model = d.model; a.fieldset = FS_OUT; a.sol = sol.as<C>(); a.out = mixed.as<C>(); a.fstride = fsz;
    a.ncols = g.kmax; a.eb = eb; a.ell2inv = (T)ell2inv; a.sinv = sinv.as<T>(); a.tw = twy.as<C>();
    launch_k1<T>(d.ny, a, stream);
    const int nf = nvars();
    K2Args<T> b{};
    b.g = g; b.g.nlayers = 1; b.mode = K2_INVERSE_OUT; b.nfields = nf * nl; b.in = mixed.as<C>();
    b.fstride = fsz; b.kin = g.kmax; b.phys = phys.as<T>(); b.pstride = (long)d.nx * d.ny; b.tw = twx.as<C>();
    launch_k2<T>(d.nx, b, stream);
    HIPCHECK(hipGetLastError());
    const size_t np = (size_t)d.nx * d.ny;
    for (int f = 0; f < nf; ++f)
      for (int j = 0; j < nl; ++j)
        HIPCHECK(hipMemcpyAsync((char*)out + sizeof(T) * np * (f * nl + j), phys.as<T>() + np * (j * nf + f),
                                sizeof(T) * np, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  void calcN(void* N, void* solh, double tt) override {
    const size_t bytes = sizeof(C) * (size_t)g.nkr * d.ny * nl;
    HIPCHECK(hipMemcpyAsync(jl.p, solh, bytes, hipMemcpyHostToDevice, stream));
    launch_from_julia<T>(g, jl.as<C>(), cinp.as<C>(), nl, stream);
    if (harmonic) cur_fscale = (T)std::cos(f_omega * tt + f_phase);   // calcF!(Fh, sol, t, ...) at t
    tm.mark(stream);
    calcN_launch(cinp.as<C>(), E_CALCN, true);
    cur_fscale = 1;
    launch_to_julia<T>(g, nout.as<C>(), jl.as<C>(), nl, stream);
    HIPCHECK(hipMemcpyAsync(N, jl.p, bytes, hipMemcpyDeviceToHost, stream));
    if (!(d.model == GF_MODEL_MLQG && d.linear)) launch_mask<T>(g, cinp.as<C>(), nl, stream);   // calcN!'s dealias!
    launch_to_julia<T>(g, cinp.as<C>(), jl.as<C>(), nl, stream);
    HIPCHECK(hipMemcpyAsync(solh, jl.p, bytes, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  bool is_harmonic() const override { return harmonic; }
  void set_forcing_harmonic(const void* F, double omega, double phase) override {
    set_forcing(F);
    harmonic = true;
    f_omega = omega;
    f_phase = phase;
    last_fscale = 0;
  }
  void set_forcing(const void* F) override {
    if (d.forcing != GF_FORCING_FIXED) throw GfError(GF_EINVAL, "gf_set_forcing: problem was not created with GF_FORCING_FIXED");
    harmonic = false;
    const size_t bytes = sizeof(C) * (size_t)g.nkr * d.ny * nl;
    HIPCHECK(hipMemcpyAsync(jl.p, F, bytes, hipMemcpyHostToDevice, stream));
    launch_from_julia<T>(g, jl.as<C>(), Fh.as<C>(), nl, stream);
    HIPCHECK(hipStreamSynchronize(stream));
  }
  void timing(int enable) override { tm.reset(enable, stream); }
  void timing_read(int k, int64_t* n, double* ms) override { tm.read(k, n, ms, stream); }
  // Kernel class k (0 = K1, 1 = K2, 2 = K3 with the calcN epilogue) launched `reps` times back to
  // back on the current state between two events. Consecutive launches of one kernel follow
  // each other without a gap, so the mean is the kernel's execution time (what a kernel trace
  // reports), which an event bound to each launch inside a step overstates by the dispatch gap
  // that its completion signal adds. The state is not changed: K1 writes the mixed fields, K2
  // the fluxes, K3 the calcN output buffer.
  void time_kernel(int k, int reps, double* ms) override {
    if (k < 0 || k > 2) throw GfError(GF_EINVAL, "kernel class must be 0, 1 or 2");
    if (reps < 1) throw GfError(GF_EINVAL, "gf_time_kernel: reps must be >= 1");
    K1Args<T> a; K2Args<T> b; K3Args<T> c;
    make_args(sol.as<C>(), E_CALCN, false, a, b, c);
    const bool was = tm.on;
    tm.on = false;
    launch_k1<T>(d.ny, a, stream);           // the inputs of K2 and K3 from this state
    launch_k2<T>(d.nx, b, stream);
    hipEvent_t e[2] = {nullptr, nullptr};
    auto done = [&] { for (auto& x : e) if (x) { (void)hipEventDestroy(x); x = nullptr; } tm.on = was; };
    try {
      HIPCHECK(hipEventCreate(&e[0]));
      HIPCHECK(hipEventCreate(&e[1]));
      HIPCHECK(hipEventRecord(e[0], stream));
      for (int r = 0; r < reps; ++r) {
        if (k == 0) launch_k1<T>(d.ny, a, stream);
        else if (k == 1) launch_k2<T>(d.nx, b, stream);
        else launch_k3<T>(d.ny, c, stream);
      }
      HIPCHECK(hipEventRecord(e[1], stream));
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipEventSynchronize(e[1]));
      float x = 0;
      HIPCHECK(hipEventElapsedTime(&x, e[0], e[1]));
      *ms = (double)x / reps;
    } catch (...) { done(); throw; }
    done();
  }
  // compulsory HBM bytes per launch of kernel class k
  double kernel_bytes(int k, int full) const override {
    const double cs = sizeof(C), rs = sizeof(T);
    const double ny = d.ny, kmax = g.kmax, nkr = g.nkr;
    const double kept_l = (double)(g.la0 + (d.ny - g.la1));
    const double kept = kmax * kept_l;                // surviving modes
    const double cols = full ? nkr : kmax;
    const double out_modes = full ? nkr * ny : kept;  // modes whose N is used
    const double L = nl;
    if (k == 0) {
      double b = kept * cs * L + 3 * kmax * ny * cs * L;
      if (d.model == GF_MODEL_MLQG) b += kept * L * L * rs;
      return b;
    }
    if (k == 1) {
      double b = 3 * kmax * ny * cs * L + 2 * cols * ny * cs * L;
      if (d.model == GF_MODEL_SLQG && has_eta && !advect_full) b += (double)d.nx * ny * rs;
      if (advect_full && has_eta) b += 2 * (double)d.nx * ny * rs;
      return b;
    }
    // K3: flux read + the average stage epilogue of the configured stepper
    double b = 2 * cols * ny * cs * L;
    const double cf = ccplx ? cs : rs;
    double e;
    if (etd()) {
      const double s1b = kept * cs * 2 + kept * cf * 2 + out_modes * cs;                       // sol,E2,Z -> s1, n1
      const double s2b = s1b;                                                                    // sol,E2,Z -> s2, n23
      const double s3b = kept * cs * 3 + kept * cf * 2 + out_modes * cs * 2;                    // s1,n1,E2,Z -> s2; n23 rmw
      const double s4b = kept * cs + out_modes * cs * 3 + out_modes * cf * 4;                    // sol,n1,n23,E,a,b,G -> sol
      e = (s1b + s2b + s3b + s4b) / 4;
    } else if (d.stepper == GF_RK4 || d.stepper == GF_FILTERED_RK4) {
      const double r1 = kept * cs * 2 + kept * cf + out_modes * cs;                             // sol,L -> s1, racc
      const double r2 = kept * cs * 3 + kept * cf + out_modes * cs * 2;                         // sol,s1,L -> s1; racc rmw
      const double r4 = kept * cs * 2 + out_modes * cs * 2 + out_modes * cf;                    // sol,s1,racc,L -> sol
      e = (r1 + 2 * r2 + r4) / 4;
    } else {
      e = kept * cs + out_modes * cs + out_modes * cf;
    }
    return b + e * L;
  }
  void sync() override { HIPCHECK(hipStreamSynchronize(stream)); }

  // ---------------- diagnostics ----------------
  // fills the diagnostic's kernel arguments and the host-side scale; throws for a kind the
  // problem does not provide
  void diag_setup(int kind, double xi, int nxi, DiagArgs<T>& a, double& scale) {
    a.g = g; a.kind = kind; a.model = d.model; a.nl = nl; a.sol = sol.as<C>();
    a.nout = 1;
    const double np2 = (double)d.nx * d.ny * (double)d.nx * d.ny;
    scale = 1.0 / np2;                                      // 1/(Lx Ly) times parsevalsum's Lx Ly/(nx ny)^2
    const bool slqg = d.model == GF_MODEL_SLQG, twodns = d.model == GF_MODEL_TWODNS;
    auto bad = [&](const char* m) { throw GfError(GF_EUNSUPPORTED, m); };
    if (d.model == GF_MODEL_MLQG) {
      if (kind != DIAG_LAYER_ENERGIES && kind != DIAG_FLUXES) bad("MultiLayerQG provides GF_DIAG_LAYER_ENERGIES and GF_DIAG_FLUXES");
      a.nout = 2 * nl - 1;
      if (kind == DIAG_FLUXES) { scale = 1; return; }   // the factors are applied per point
      double Hs = 0;
      for (double h : Hlay) Hs += h;
      const double V = d.Lx * d.Ly * Hs;
      for (int j = 0; j < nl; ++j) a.lay_ke[j] = Hlay[j] / (2 * V);
      if (nl == 1) a.lay_ke[0] = 1 / (2 * d.Lx * d.Ly);   // src/multilayerqg.jl:953-960
      for (int j = 0; j + 1 < nl; ++j) a.lay_pe[j] = d.f0 * d.f0 / (2 * V * gprime[j]);
      a.sinv = sinv.as<T>();
      scale = d.Lx * d.Ly / np2;
    } else {
      switch (kind) {
        case DIAG_ENERGY: case DIAG_ENSTROPHY: break;
        case DIAG_KINETIC_ENERGY: case DIAG_POTENTIAL_ENERGY: if (!slqg) bad("kinetic/potential energy are SingleLayerQG diagnostics"); break;
        case DIAG_PALINSTROPHY: if (!twodns) bad("palinstrophy is a TwoDNavierStokes diagnostic"); break;
        case DIAG_ENERGY_DISSIPATION: case DIAG_ENSTROPHY_DISSIPATION:
        case DIAG_ENERGY_DRAG: case DIAG_ENSTROPHY_DRAG:
          a.xi = (T)xi; a.nxi = nxi;
          if (slqg && eb) bad("not implemented for finite deformation radius (src/singlelayerqg.jl)");
          break;
        case DIAG_ENERGY_WORK: case DIAG_ENSTROPHY_WORK:
          if (slqg && eb) bad("not implemented for finite deformation radius (src/singlelayerqg.jl)");
          if (d.forcing == GF_FORCING_NONE) bad("the problem has no forcing");
          if (d.forcing == GF_FORCING_FIXED) { a.Fh = Fh.as<C>(); a.fscale = harmonic ? (T)last_fscale : (T)1; }
          if (d.forcing == GF_FORCING_RING && prevsol_valid) {
            // before the first step Fh is still zero, and so is the work (as the reference's
            // zero-initialised vars.Fh gives)
            a.prevsol = prevsol.as<C>();
            a.sqrt_spec = fkx >= 0 ? sqrt_spec.as<T>() : nullptr;
            a.seed = d.seed; a.fstep = stepno - 1; a.sqrt_dt = (T)std::sqrt(d.dt);
            a.fkx = fkx; a.fky = fky; a.instance = d.instance;
          }
          break;
        default: throw GfError(GF_EINVAL, "unknown or inapplicable diagnostic kind");
      }
      a.etah = (slqg && has_eta) ? etah.as<C>() : nullptr;
      a.eb = eb; a.ell2inv = (T)ell2inv;
    }
  }
  // per-block partial sums into diag_part; returns the number of blocks
  int diag_launch(const DiagArgs<T>& a) {
    const int nblocks = (int)std::min<long>(1024, (fsz + 255) / 256);
    if (!diag_part.p) diag_part.alloc(sizeof(double) * 1024 * DIAG_MAXOUT);
    if (a.kind == DIAG_FLUXES && d.model == GF_MODEL_MLQG) {
      flux_launch(diag_part.as<double>(), nblocks);
      return nblocks;
    }
    launch_diag<T>(a, diag_part.as<double>(), nblocks, stream);
    HIPCHECK(hipGetLastError());
    return nblocks;
  }
  // MultiLayerQG fluxes (src/multilayerqg.jl:977-1022) as per-block partials: lateral_1..n, then
  // vertical_{3/2}..{n-1/2}.  fluxes calls updatevars!, which dealiases sol in place; the spectra
  // of psi, v and du/dy go to s1, s2 and n23 (stage scratch between steps), their y transforms to
  // the mixed fields, and the x sums are taken there by Parseval along x.
  void flux_launch(double* part, int nblocks) {
    launch_mask<T>(g, sol.as<C>(), nl, stream);
    FluxArgs<T> f{};
    f.g = g; f.nl = nl; f.sol = sol.as<C>(); f.sinv = nl > 1 ? sinv.as<T>() : nullptr;
    f.f[0] = s1.as<C>(); f.f[1] = s2.as<C>(); f.f[2] = n23.as<C>(); f.fstride = fsz; f.Uy = Uy.as<T>();
    double Hs = 0;
    for (double h : Hlay) Hs += h;
    const double dxdy = (d.Lx / d.nx) * (d.Ly / d.ny);
    const double V = nl > 1 ? d.Lx * d.Ly * Hs : d.Lx * d.Ly;
    for (int j = 0; j < nl; ++j) f.fac[j] = (nl > 1 ? Hlay[j] : 1.0) * dxdy / V * d.nx;
    for (int j = 0; j + 1 < nl; ++j) f.fac[nl + j] = d.f0 * d.f0 / gprime[j] * dxdy / V * d.nx;
    launch_flux_fields<T>(f, stream);
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < nl; ++j) {
        K1Args<T> a{};
        a.g = g; a.g.nlayers = 1; a.model = GF_MODEL_TWODNS; a.fieldset = FS_RAW;
        a.sol = f.f[k] + (size_t)j * fsz; a.out = mixed.as<C>() + (size_t)(k * nl + j) * fsz; a.fstride = fsz;
        a.ncols = g.nkr; a.tw = twy.as<C>();
        launch_k1<T>(d.ny, a, stream);
      }
    for (int k = 0; k < 3; ++k) f.f[k] = mixed.as<C>() + (size_t)k * nl * fsz;
    launch_flux<T>(f, part, nblocks, stream);
    HIPCHECK(hipGetLastError());
  }
  // gf_diagnostic_async without any allocation: the same kernels and block count, the partials in
  // the static scratch of this device instead of diag_part.  Calls that overlap on one GPU race
  // on the scratch.
  int diagnostic_noalloc(int kind, double xi, int nxi, double* out_dev, int nout) override {
    DiagArgs<T> a{};
    double scale = 1;
    diag_setup(kind, xi, nxi, a, scale);
    if (nout < a.nout) throw GfError(GF_EINVAL, "output array too small for this diagnostic");
    const int nblocks = (int)std::min<long>(1024, (fsz + 255) / 256);
    if (nblocks > kScratchBlocks || (size_t)nblocks * a.nout > kScratchDoubles)
      throw GfError(GF_EINVAL, "gf_diagnostic_async_noalloc: " + std::to_string(nblocks) + " x " +
                               std::to_string(a.nout) + " partial sums exceed the static scratch of " +
                               std::to_string(kScratchDoubles) + " doubles");
    if (kind == DIAG_FLUXES && d.model == GF_MODEL_MLQG) {
      flux_launch(scratch_dev, nblocks);
    } else {
      launch_diag<T>(a, scratch_dev, nblocks, stream);
      HIPCHECK(hipGetLastError());
    }
    launch_diag_finish(scratch_dev, nblocks, a.nout, scale, out_dev, stream);
    HIPCHECK(hipGetLastError());
    return a.nout;
  }
  int diagnostic_async(int kind, double xi, int nxi, double* out_dev, int nout) override {
    DiagArgs<T> a{};
    double scale = 1;
    diag_setup(kind, xi, nxi, a, scale);
    if (nout < a.nout) throw GfError(GF_EINVAL, "output array too small for this diagnostic");
    const int nblocks = diag_launch(a);
    launch_diag_finish(diag_part.as<double>(), nblocks, a.nout, scale, out_dev, stream);
    HIPCHECK(hipGetLastError());
    return a.nout;
  }
  int diagnostic(int kind, double xi, int nxi, double* out, int nout) override {
    DiagArgs<T> a{};
    double scale = 1;
    diag_setup(kind, xi, nxi, a, scale);
    if (nout < a.nout) throw GfError(GF_EINVAL, "output array too small for this diagnostic");
    const int nblocks = diag_launch(a);
    std::vector<double> part((size_t)nblocks * a.nout);
    HIPCHECK(hipMemcpyAsync(part.data(), diag_part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    for (int k = 0; k < a.nout; ++k) {
      double s = 0;
      for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * a.nout + k];
      out[k] = s * scale;
    }
    return a.nout;
  }

  // ---------------- adjoint (TwoDNavierStokes, ETDRK4, FP64) ----------------
  DevBuf adj_mixed, adj_flux, lamb, l0b, l1b, lN1b, lN2b, lN3b, nu_part, dtab[6], pool, ckpts;
  // The tape is a pool of segment slots. Segment s of a run is taped into slot s mod P, so after
  // the forward pass the last P segments are still taped, and the backward sweep recomputes only
  // the earlier ones from their checkpoints. P is 2 (the two-stream pipeline) unless a tape
  // budget (gf_adjoint_tape_budget) affords more. With recomputation the sweep is pipelined over
  // two streams: segment s - 1 is recomputed into its slot on rstream while the adjoint of
  // segment s reads its own slot on the main stream.
  hipStream_t rstream = nullptr;
  std::vector<hipEvent_t> adj_done, rec_done;   // per slot
  hipEvent_t fwd_done = nullptr;
  long pool_seg = 0;         // steps per slot of the pool (grown on demand, never shrunk:
  int pool_slots = 0;        // freeing device memory synchronises the whole device)
  long ckpt_slots = 0;
  size_t tape_budget = 0;    // bytes the pool may take; 0: two slots
  int ka3_blocks = 0;
  // A tape field holds only the columns kr < kmax (rounded up to a block): every reader masks,
  // and in the blocked layout those columns are a prefix of the field.
  long kstride() const { return (long)((g.kmax + BW - 1) / BW * BW) * d.ny; }

  void adjoint_setup() {
    if (adj_mixed.p) return;
    adj_mixed.alloc(sizeof(C) * fsz * 5);
    adj_flux.alloc(sizeof(C) * fsz * 3);
    for (DevBuf* b : {&lamb, &l0b, &l1b, &lN1b, &lN2b, &lN3b}) b->alloc(sizeof(C) * fsz);
    ka3_blocks = adj_blocks<double>(d.ny, g.kmax);
    nu_part.alloc(sizeof(double) * ka3_blocks);
    // dJ/dnu tables: d(coefficient)/dL times dL/dnu = -Krsq^nnu (0 at k = 0)
    const int ny = d.ny, nkr = g.nkr;
    const size_t nm = (size_t)ny * nkr;
    std::vector<double> t6[6];
    for (auto& v : t6) v.assign(nm, 0.0);
    const double dt = d.dt;
    parallel_for((long)nm, [&](long i) {
      const int l = (int)(i / nkr), kr = (int)(i % nkr);
      const double L = Lop[i].real();
      const double dLdnu = (kr == 0 && l == 0) ? 0.0 : -std::pow(krsq(kr, l), (double)d.nnu);
      cd dz, da, db, dg;
      etd_coeff_derivs(cd(dt * L, 0), dt, dz, da, db, dg);
      t6[0][i] = dLdnu * dt * std::exp(dt * L);
      t6[1][i] = dLdnu * dt / 2 * std::exp(dt * L / 2);
      t6[2][i] = dLdnu * dz.real(); t6[3][i] = dLdnu * da.real();
      t6[4][i] = dLdnu * db.real(); t6[5][i] = dLdnu * dg.real();
    });
    for (int k = 0; k < 6; ++k) upload_blocked(dtab[k], t6[k]);
  }

  // Step i, field f of pool slot tp: fields s0 (the step's input), N1, N2, N3, N4; step i + 1's
  // s0 follows. The stage inputs s1, s2, s3 are not taped: the adjoint rebuilds them from s0 and
  // the Ns (E2 s0 + Z N1, E2 s0 + Z N2, E2 s1 + Z (2 N3 - N1)), 5 fields per step instead of 8.
  static constexpr int TAPE_FIELDS = 5;
  C* slot(long i, int f, int tp = 0) const {
    return pool.as<C>() + ((size_t)tp * (TAPE_FIELDS * pool_seg + 1) + (size_t)i * TAPE_FIELDS + f) * kstride();
  }
  void set_tape_budget(size_t bytes) override { tape_budget = bytes; }

  // One ETDRK4 step from tape slot i (s0) writing N1..N4 of that slot and the new state to next
  // (kept modes only: every later read masks), on stream st. The stage inputs s1, s2 and s3 (in
  // s2's buffer, as in a plain step) go to the forward path's buffers, like the rest of its
  // scratch (mixed, flux, n1, n23); the adjoint kernels use none of them.
  void tape_step(long i, C* next, int tp, hipStream_t st) {
    C* s0 = slot(i, 0, tp);
    C* b1 = s1.as<C>();
    C* b2 = s2.as<C>();
    K1Args<T> a; K2Args<T> b; K3Args<T> k;
    auto run = [&]() {
      launch_k1<T>(d.ny, a, st); launch_k2<T>(d.nx, b, st); launch_k3<T>(d.ny, k, st);
    };
    make_args(s0, E_ETD1, false, a, b, k);
    k.sol = s0; k.s1 = b1; k.nsave = slot(i, 1, tp); run();
    make_args(b1, E_ETD2, false, a, b, k);
    k.sol = s0; k.s2 = b2; k.nsave = slot(i, 2, tp); run();
    make_args(b2, E_ETD3, false, a, b, k);
    k.s1 = b1; k.s2 = b2; k.nsave = slot(i, 3, tp); run();
    make_args(b2, E_ETD4, false, a, b, k);
    k.sol = s0; k.sol_out = next; k.nsave = slot(i, 4, tp); k.nonfinite = nullptr; run();
    HIPCHECK(hipGetLastError());
  }

  // adjoint of the step taped in slot i of tape tp: lamb (lambda of the new state) -> lambda of s0
  void adjoint_step(long i, bool want_nu, int tp = 0) {
    if constexpr (std::is_same<T, double>::value) {
      for (int stage = 4; stage >= 1; --stage) {
        AdjK1Args<T> a{};
        a.g = g; a.stage = stage;
        a.s0 = slot(i, 0, tp); a.N1 = slot(i, 1, tp); a.N2 = slot(i, 2, tp); a.N3 = slot(i, 3, tp);
        a.E2 = (const T*)coef[0].p; a.Z = (const T*)coef[1].p;
        a.x = stage == 4 ? lamb.as<C>() : stage == 3 ? lN3b.as<C>() : stage == 2 ? lN2b.as<C>() : lN1b.as<C>();
        a.lamc = stage == 4 ? (const T*)coef[5].p : nullptr;
        a.out = adj_mixed.as<C>(); a.fstride = fsz; a.ncols = g.kmax; a.tw = twy.as<C>();
        a.fsplit = split[0];   // the forward problem's tuned workgroup splits
        AdjK2Args<T> b{};
        b.g = g; b.in = adj_mixed.as<C>(); b.fstride = fsz; b.kin = g.kmax;
        b.out = adj_flux.as<C>(); b.ostride = fsz; b.kout = g.kmax; b.tw = twx.as<C>();
        b.hsplit = split[1] != 0;
        AdjK3Args<T> k{};
        k.g = g; k.in = adj_flux.as<C>(); k.fstride = fsz; k.ncols = g.kmax; k.stage = stage;
        k.lam = lamb.as<C>(); k.l0 = l0b.as<C>(); k.l1 = l1b.as<C>();
        k.lN1 = lN1b.as<C>(); k.lN2 = lN2b.as<C>(); k.lN3 = lN3b.as<C>();
        k.E2 = (const T*)coef[0].p; k.Z = (const T*)coef[1].p; k.E = (const T*)coef[2].p;
        k.Al = (const T*)coef[3].p; k.Be = (const T*)coef[4].p;
        k.dE = dtab[0].as<T>(); k.dE2 = dtab[1].as<T>(); k.dZ = dtab[2].as<T>();
        k.dA = dtab[3].as<T>(); k.dB = dtab[4].as<T>(); k.dG = dtab[5].as<T>();
        k.s0 = slot(i, 0, tp);
        k.N1 = slot(i, 1, tp); k.N2 = slot(i, 2, tp); k.N3 = slot(i, 3, tp); k.N4 = slot(i, 4, tp);
        k.nu_part = want_nu ? nu_part.as<T>() : nullptr;
        k.tw = twy.as<C>();
        rec(0, true); launch_adj<T>(1, d.ny, &a, stream); rec(0, false);
        rec(1, true); launch_adj<T>(2, d.nx, &b, stream); rec(1, false);
        rec(2, true); launch_adj<T>(3, d.ny, &k, stream); rec(2, false);
      }
      HIPCHECK(hipGetLastError());
    }
  }

  // host-side Hermitian weight, alias mask and zero-mean handling of a blocked spectral field
  void host_spectral(std::vector<C>& h, bool to_lambda, double scale) const {
    for (int l = 0; l < d.ny; ++l)
      for (int kr = 0; kr < g.nkr; ++kr) {
        C& z = h[boff_h(kr, l, d.ny)];
        const bool keep = kr < g.kmax && (l < g.la0 || l >= g.la1);
        const double w = (kr == 0 || 2 * kr == d.nx) ? 1.0 : 2.0;
        const double f = !keep ? 0.0 : (to_lambda ? w * scale : scale / w);
        z.re = (T)(z.re * f); z.im = (T)(z.im * f);
      }
    if (!to_lambda) h[boff_h(0, 0, d.ny)] = C{};   // set_zeta!'s zero mean
  }

  void adjoint_grad(const double* target, int nsteps, int every, double* grad, double* dnu, double* Jout) override {
    if constexpr (!std::is_same<T, double>::value) {
      throw GfError(GF_EUNSUPPORTED, "the adjoint is provided for FP64 problems");
    } else {
      if (d.model != GF_MODEL_TWODNS || d.stepper != GF_ETDRK4 || d.forcing != GF_FORCING_NONE)
        throw GfError(GF_EUNSUPPORTED, "the adjoint is provided for unforced TwoDNavierStokes ETDRK4 problems");
      if (d.nx != d.ny) throw GfError(GF_EUNSUPPORTED, "the adjoint needs a square grid");
      if (d.nx > 2048 || !is_pow2(d.nx))
        throw GfError(GF_EUNSUPPORTED, "the adjoint kernels are built for power-of-two grids up to 2048^2");
      if (nsteps < 1 || every < 1) throw GfError(GF_EINVAL, "nsteps and checkpoint_every must be >= 1");
      adjoint_setup();
      struct Active {
        int others;
        Active() : others(g_adjoint_active.fetch_add(1)) {}
        ~Active() { g_adjoint_active.fetch_sub(1); }
      } active;
      const long seg = std::min(every, nsteps);
      const long nseg = (nsteps + seg - 1) / seg;
      const size_t kb = sizeof(C) * (size_t)kstride();   // one tape field (a prefix of a state)
      const size_t seg_bytes = kb * (size_t)(TAPE_FIELDS * seg + 1);
      // pool slots: as many whole segments as the budget affords, at least two with the pipeline
      // (which pays only when this is the one gradient in flight), never more than the run has
      long nslot = (long)std::min<size_t>((size_t)nseg, tape_budget / seg_bytes);
      const bool alone = active.others == 0;
      nslot = std::max<long>(nslot, (nseg > 1 && alone) ? 2 : 1);
      const int P = (int)nslot;
      const bool pipe = nseg > P && alone;   // recompute on rstream beside the sweep
      if (pool_seg < seg || pool_slots < P) {
        pool.alloc(0);   // free first: the old and the new pool need not fit together
        pool.alloc(sizeof(C) * (size_t)std::max<long>(P, pool_slots) * (TAPE_FIELDS * seg + 1) * kstride());
        pool_seg = seg;
        pool_slots = std::max(P, pool_slots);
      }
      if (pipe && !rstream) {
        HIPCHECK(dw::stream_create(&rstream));
        HIPCHECK(hipEventCreateWithFlags(&fwd_done, hipEventDisableTiming));
      }
      while (pipe && (int)adj_done.size() < P) {
        hipEvent_t a = nullptr, r = nullptr;
        HIPCHECK(hipEventCreateWithFlags(&a, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&r, hipEventDisableTiming));
        adj_done.push_back(a);
        rec_done.push_back(r);
      }
      if (ckpt_slots < nseg) {
        ckpts.alloc(kb * nseg);
        ckpt_slots = nseg;
      }
      const size_t fb = sizeof(C) * fsz;
      auto len = [&](long s) { return std::min(seg, (long)nsteps - s * seg); };
      // forward: checkpoints at segment starts; segment s is taped into slot s mod P, so the last
      // P segments' tapes stay for the sweep
      auto tape_of = [&](long s) { return (int)(s % P); };
      for (long s = 0; s < nseg; ++s) {
        const int tp = tape_of(s);
        HIPCHECK(hipMemcpyAsync(slot(0, 0, tp), s == 0 ? (const C*)sol.p : slot(len(s - 1), 0, tape_of(s - 1)), kb,
                                hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(ckpts.as<C>() + s * kstride(), slot(0, 0, tp), kb, hipMemcpyDeviceToDevice, stream));
        const long n = len(s);
        for (long i = 0; i < n; ++i) tape_step(i, slot(i + 1, 0, tp), tp, stream);
      }
      // loss and lambda_T = M (w/N) rfft2(zeta_T - target), zeta_T = irfft2(M sol_T); the columns of
      // cinp past the tape's prefix hold stale values, which the mask clears
      const long nlast = len(nseg - 1);
      const int tlast = tape_of(nseg - 1);
      const size_t np = (size_t)d.nx * d.ny;
      HIPCHECK(hipMemcpyAsync(cinp.p, slot(nlast, 0, tlast), kb, hipMemcpyDeviceToDevice, stream));
      launch_mask<T>(g, cinp.as<C>(), 1, stream);
      inv_device(cinp.as<C>(), phys.as<T>(), 1);
      std::vector<T> r(np);
      HIPCHECK(hipMemcpyAsync(r.data(), phys.p, sizeof(T) * np, hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      double J = 0;
      for (size_t i = 0; i < np; ++i) { r[i] -= (T)target[i]; J += 0.5 * (double)r[i] * (double)r[i]; }
      HIPCHECK(hipMemcpyAsync(phys.p, r.data(), sizeof(T) * np, hipMemcpyHostToDevice, stream));
      fwd_device(phys.as<T>(), lamb.as<C>(), 1, 0);
      std::vector<C> h(fsz);
      HIPCHECK(hipMemcpyAsync(h.data(), lamb.p, fb, hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      host_spectral(h, true, 1.0 / (double)np);
      HIPCHECK(hipMemcpyAsync(lamb.p, h.data(), fb, hipMemcpyHostToDevice, stream));
      HIPCHECK(hipMemsetAsync(nu_part.p, 0, sizeof(double) * ka3_blocks, stream));
      // reverse sweep: while the main stream runs the adjoint of segment s from tape s mod 2,
      // rstream recomputes segment s - 1 into the other tape, once the adjoint of segment s + 1
      // (the last reader of that tape) is done
      if (pipe) {
        HIPCHECK(hipEventRecord(fwd_done, stream));
        HIPCHECK(hipStreamWaitEvent(rstream, fwd_done, 0));
      }
      // segments s >= nseg - P are still taped from the forward pass
      const long resident = nseg - P;
      for (long s = nseg - 1; s >= 0 && !pipe; --s) {
        // one stream: recompute segment s into its slot if it is not resident, then its adjoint
        const long n = len(s);
        const int tp = tape_of(s);
        if (s < resident) {
          HIPCHECK(hipMemcpyAsync(slot(0, 0, tp), ckpts.as<C>() + s * kstride(), kb, hipMemcpyDeviceToDevice, stream));
          for (long i = 0; i < n; ++i) tape_step(i, slot(i + 1, 0, tp), tp, stream);
        }
        for (long i = n - 1; i >= 0; --i) adjoint_step(i, dnu != nullptr, tp);
      }
      for (long s = nseg - 1; s >= 0 && pipe; --s) {
        const int tp = tape_of(s);
        if (s >= 1 && s - 1 < resident) {
          // segment s - 1 goes into the slot of segment s - 1 + P, whose adjoint was queued before
          const int tq = tape_of(s - 1);
          HIPCHECK(hipStreamWaitEvent(rstream, adj_done[tq], 0));
          HIPCHECK(hipMemcpyAsync(slot(0, 0, tq), ckpts.as<C>() + (s - 1) * kstride(), kb, hipMemcpyDeviceToDevice,
                                  rstream));
          const long m = len(s - 1);
          for (long i = 0; i < m; ++i) tape_step(i, slot(i + 1, 0, tq), tq, rstream);
          HIPCHECK(hipEventRecord(rec_done[tq], rstream));
        }
        if (s < resident) HIPCHECK(hipStreamWaitEvent(stream, rec_done[tp], 0));
        const long n = len(s);
        for (long i = n - 1; i >= 0; --i) adjoint_step(i, dnu != nullptr, tp);
        HIPCHECK(hipEventRecord(adj_done[tp], stream));
      }
      // dJ/dzeta0 = N irfft2(Z M lambda_0 / w)
      HIPCHECK(hipMemcpyAsync(h.data(), lamb.p, fb, hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      host_spectral(h, false, (double)np);
      HIPCHECK(hipMemcpyAsync(cinp.p, h.data(), fb, hipMemcpyHostToDevice, stream));
      inv_device(cinp.as<C>(), phys.as<T>(), 1);
      if (grad) HIPCHECK(hipMemcpyAsync(grad, phys.p, sizeof(T) * np, hipMemcpyDeviceToHost, stream));
      std::vector<double> parts(ka3_blocks);
      HIPCHECK(hipMemcpyAsync(parts.data(), nu_part.p, sizeof(double) * ka3_blocks, hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      if (dnu) { double sum = 0; for (double v : parts) sum += v; *dnu = sum; }
      if (Jout) *Jout = J;
    }
  }
};

// Ensemble of independent FP64 instances advanced together: per stage, one launch per
// kernel class covers every member (flat (instance, block) work tables, largest
// members first).  Each member is a full Ctx on the shared stream, so the
// single-instance entry points (state, physical fields, updatevars!, calcN!, and a
// member-only step) work on it unchanged.
template <typename T>
struct BatchCtx final : CtxBase {
  using C = cplx<T>;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<Ctx<T>>> inst;
  int nst = 0;             // stages per step
  // BatchItem per member for (stage, variant) and BatchWork table per (kernel class, variant).
  // variant: the last step of a call (1) or an interior step (0); for AB3, whose every step is
  // computed in full, the forward-Euler start-up (1) or the Adams-Bashforth step (0)
  DevBuf items[4][2];
  DevBuf work[3][2];
  int nwork[3][2] = {};
  KernelTimer tm;
  bool ring = false;       // some member has stochastic ring forcing
  bool ab3m = false;       // the members step with AB3 / FilteredAB3
  DevBuf ctr;              // ring forcing: the members' step number, read by K3 (fstep_dev)

  BatchCtx(const gf_problem_desc* ds, int n) {
    HIPCHECK(dw::stream_create(&stream));
    try {
      inst.reserve(n);
      for (int i = 0; i < n; ++i) inst.emplace_back(new Ctx<T>(ds[i], stream));
      build_tables();
    } catch (...) {
      inst.clear();
      (void)dw::stream_destroy(stream);
      stream = nullptr;
      throw;
    }
  }
  ~BatchCtx() override {
    if (stream) (void)hipStreamSynchronize(stream);
    if (gexec) (void)dw::exec_destroy(gexec);
    inst.clear();
    if (stream) (void)dw::stream_destroy(stream);
  }

  bool k2_plain = true;   // every member's x pass is K2_ADVECT (no background gradients)
  void build_tables() {
    const int n = (int)inst.size();
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return inst[a]->d.nx > inst[b]->d.nx; });
    const C* ins[4];
    int epis[4];
    nst = inst[0]->stages(ins, epis);
    ab3m = inst[0]->ab3();
    ring = false;
    for (auto& m : inst) ring = ring || m->d.forcing == GF_FORCING_RING;
    if (ring) ctr.alloc(sizeof(int64_t));
    std::vector<BatchItem<T>> h(n);
    k2_plain = true;
    for (int full = 0; full < 2; ++full) {
      for (int k = 0; k < nst; ++k) {
        for (int i = 0; i < n; ++i) {
          Ctx<T>& m = *inst[i];
          m.stages(ins, epis);
          h[i] = BatchItem<T>{};
          h[i].n = m.d.nx;
          if (ab3m) m.make_args(ins[k], epis[k], true, h[i].k1, h[i].k2, h[i].k3, full);
          else m.make_args(ins[k], epis[k], full != 0, h[i].k1, h[i].k2, h[i].k3);
          h[i].k1.fsplit = 0;   // the batched launches are one-dimensional
          h[i].k2.hsplit = 0;
          h[i].k3.csplit = 0;
          // ring forcing: every step's realisation is drawn from the step number in ctr, which the
          // step sequence sets and advances (so the tables and a captured step stay valid)
          if (ring) h[i].k3.fstep_dev = ctr.as<int64_t>();
          k2_plain = k2_plain && h[i].k2.mode == K2_ADVECT;
        }
        items[k][full].alloc(sizeof(BatchItem<T>) * n);
        HIPCHECK(dw::copy(items[k][full].p, h.data(), sizeof(BatchItem<T>) * n, hipMemcpyHostToDevice));
      }
      for (int kc = 0; kc < 3; ++kc) {
        std::vector<BatchWork> w;
        for (int i : ord) {
          const Ctx<T>& m = *inst[i];
          const int count = kc == 1 ? m.d.ny : (kc == 2 && (full || ab3m) ? m.g.nkr : m.g.kmax);
          const int nb = batch_blocks(kc + 1, m.d.nx, count);
          for (int b = 0; b < nb; ++b) w.push_back(BatchWork{i, b});
        }
        nwork[kc][full] = (int)w.size();
        work[kc][full].alloc(sizeof(BatchWork) * w.size());
        HIPCHECK(dw::copy(work[kc][full].p, w.data(), sizeof(BatchWork) * w.size(), hipMemcpyHostToDevice));
      }
    }
  }

  void step_launches(int full) {
    tm.mark(stream);
    for (int k = 0; k < nst; ++k) {
      const BatchItem<T>* it = items[k][full].as<BatchItem<T>>();
      if (persistent) {
        const BatchStageWork sw{{work[0][full].as<BatchWork>(), work[1][full].as<BatchWork>(), work[2][full].as<BatchWork>()},
                                {nwork[0][full], nwork[1][full], nwork[2][full]}};
        HIPCHECK(launch_batch_stage<T>(it, sw, k2_plain, stream));
        continue;
      }
      tm.rec(0, true, stream);
      launch_batch<T, 1>(it, work[0][full].as<BatchWork>(), nwork[0][full], 0, stream);
      tm.rec(0, false, stream);
      tm.rec(1, true, stream);
      launch_batch<T, 2>(it, work[1][full].as<BatchWork>(), nwork[1][full], k2_plain, stream);
      tm.rec(1, false, stream);
      tm.rec(2, true, stream);
      launch_batch<T, 3>(it, work[2][full].as<BatchWork>(), nwork[2][full], 0, stream);
      tm.rec(2, false, stream);
    }
    if (ring) launch_counter(ctr.as<int64_t>(), 1, 1, stream);   // the next step's realisation
    HIPCHECK(hipGetLastError());
  }
  // the interior step (all stages, every member) captured once and replayed, as for one problem
  bool graphs = true;
  hipGraphExec_t gexec = nullptr;
  void set_graphs(int enable) override { graphs = enable != 0; }
  // each stage as one persistent cooperative launch (launch_batch_stage) instead of three
  // launches; such steps run without graph replay
  bool persistent = false;
  void set_persistent(int enable) override { persistent = enable != 0; }
  void build_graph() {
    gexec = dw::capture(stream, [&] { step_launches(0); });
  }

  void step(int n) override {
    for (auto& m : inst)
      if (m->is_harmonic())
        throw GfError(GF_EUNSUPPORTED, "time-harmonic forcing is not supported in ensembles (step the member alone)");
    if (n <= 0) return;
    if (ring || ab3m) {
      // one step number for every member: the forcing counter and the AB3 start-up are shared
      for (auto& m : inst)
        if (m->stepno != inst[0]->stepno)
          throw GfError(GF_EINVAL, "ensemble members with ring forcing or AB3 must be at the same step "
                                   "(a member was stepped alone); set their clocks with gf_set_clock");
      if (ring) launch_counter(ctr.as<int64_t>(), inst[0]->stepno, 0, stream);
    }
    for (int s = 0; s < n; ++s) {
      const int last = s == n - 1;
      // the tables' variant: the last step of the call, or for AB3 the forward-Euler start-up
      const int v = ab3m ? (inst[0]->stepno < 3) : last;
      if (last && ring) {
        // the work diagnostics of stochastic forcing use the state at the start of the step
        // (prevsol, src/twodnavierstokes.jl:311-319); kept for the last step of the call
        for (auto& m : inst) {
          if (m->d.forcing != GF_FORCING_RING) continue;
          if (!m->prevsol.p) m->prevsol.alloc(sizeof(C) * m->fsz);
          HIPCHECK(hipMemcpyAsync(m->prevsol.p, m->sol.p, sizeof(C) * m->fsz, hipMemcpyDeviceToDevice, stream));
          m->prevsol_valid = true;
        }
      }
      if (!last && v == 0 && graphs && !tm.on && !persistent) {
        if (!gexec) build_graph();
        HIPCHECK(hipGraphLaunch(gexec, stream));
      } else {
        step_launches(v);
      }
      for (auto& m : inst) { m->t += m->d.dt; m->stepno += 1; }
    }
    t = inst[0]->t;
    stepno = inst[0]->stepno;
  }
  [[noreturn]] static void member_only(const char* what) {
    throw GfError(GF_EINVAL, std::string(what) + ": this is an ensemble handle; address one member through gf_ctx_instance");
  }
  void set_state(const void*, hipMemcpyKind) override { member_only("set_state"); }
  void get_state(void*, hipMemcpyKind) override { member_only("get_state"); }
  void set_physical(const void*) override { member_only("set_physical"); }
  void updatevars(void*) override { member_only("updatevars"); }
  void calcN(void*, void*, double) override { member_only("calcN"); }
  void set_forcing(const void*) override { member_only("set_forcing"); }
  int nvars() const override { return inst[0]->nvars(); }
  void timing(int enable) override { tm.reset(enable, stream); }
  void timing_read(int k, int64_t* n, double* ms) override { tm.read(k, n, ms, stream); }
  double kernel_bytes(int k, int full) const override {
    double b = 0;
    for (auto& m : inst) b += m->kernel_bytes(k, full);
    return b;
  }
  void sync() override { HIPCHECK(hipStreamSynchronize(stream)); }
  int ninstances() const override { return (int)inst.size(); }
  CtxBase* instance(int i) override { return (i >= 0 && i < (int)inst.size()) ? inst[i].get() : nullptr; }
  void set_clock(double tt, int64_t s) override {
    t = tt; stepno = s;
    for (auto& m : inst) { m->t = tt; m->stepno = s; }
  }
};

int validate(const gf_problem_desc* d, std::string& msg) {
  if (!d) { msg = "null descriptor"; return GF_EINVAL; }
  if (d->model < 0 || d->model > 2) { msg = "model must be GF_MODEL_TWODNS, GF_MODEL_SLQG or GF_MODEL_MLQG"; return GF_EINVAL; }
  if (d->dtype != GF_F64 && d->dtype != GF_F32) { msg = "dtype must be GF_F64 or GF_F32"; return GF_EINVAL; }
  if (!(size_supported(d->nx) || generic_len_supported(d->nx)) || !(size_supported(d->ny) || generic_len_supported(d->ny))) {
    msg = "nx and ny must each be a power of two between 16 and 4096 or even between 18 and " +
          std::to_string(GEN_MAX) + " (got " + std::to_string(d->nx) + "x" + std::to_string(d->ny) + ")";
    return GF_EUNSUPPORTED;
  }
  if (!(d->Lx > 0) || !(d->Ly > 0)) { msg = "Lx and Ly must be positive"; return GF_EINVAL; }
  if (!(d->dt > 0)) { msg = "dt must be positive"; return GF_EINVAL; }
  if (d->stepper < 0 || d->stepper > 7) { msg = "unknown stepper"; return GF_EINVAL; }
  if (d->aliased_fraction < 0 || d->aliased_fraction >= 1) { msg = "aliased_fraction must be in [0, 1)"; return GF_EINVAL; }
  if (d->model == GF_MODEL_MLQG) {
    if (d->nlayers < 1 || d->nlayers > MAXL) { msg = "MultiLayerQG: nlayers must be 1..8 in this build"; return GF_EUNSUPPORTED; }
    if (d->nlayers > 1 && (!d->H || !d->b)) { msg = "MultiLayerQG: H and b are required"; return GF_EINVAL; }
    for (int j = 0; j + 1 < d->nlayers; ++j)
      if (d->b[j] - d->b[j + 1] == 0) { msg = "MultiLayerQG: equal buoyancy in adjacent layers"; return GF_EINVAL; }
  } else if (d->nlayers != 1 && d->nlayers != 0) {
    msg = "nlayers must be 1 for TwoDNavierStokes and SingleLayerQG"; return GF_EINVAL;
  }
  if (d->forcing < 0 || d->forcing > 2) { msg = "unknown forcing kind"; return GF_EINVAL; }
  if (d->forcing == GF_FORCING_RING && (!(d->dkf > 0) || !(d->eps >= 0))) { msg = "ring forcing needs dkf > 0, eps >= 0"; return GF_EINVAL; }
  if (d->model == GF_MODEL_TWODNS && d->nnu < 0) { msg = "nnu must be >= 0"; return GF_EINVAL; }
  return GF_OK;
}

int validate_batch(const gf_problem_desc* ds, int n, std::string& msg) {
  if (!ds || n < 1) { msg = "an ensemble needs at least one descriptor"; return GF_EINVAL; }
  for (int i = 0; i < n; ++i) {
    const gf_problem_desc& d = ds[i];
    const std::string who = "ensemble member " + std::to_string(i) + ": ";
    if (d.model == GF_MODEL_MLQG && d.nlayers > 1) {
      msg = who + "MultiLayerQG members must have one layer in ensembles"; return GF_EUNSUPPORTED;
    }
    std::string m;
    const int rc = validate(&d, m);
    if (rc != GF_OK) { msg = who + m; return rc; }
    if (d.dtype != ds[0].dtype) { msg = who + "all members must have the same precision"; return GF_EINVAL; }
    if (d.nx != d.ny || !batch_size_supported(d.nx)) {
      msg = who + "members must be square with n in {32, 64, 128, 256, 512}"; return GF_EUNSUPPORTED;
    }
    if (d.stepper != ds[0].stepper) { msg = who + "all members must use the same stepper"; return GF_EINVAL; }
  }
  return GF_OK;
}

}  // namespace

struct gf_ctx {
  std::unique_ptr<CtxBase> owned;
  CtxBase* impl = nullptr;
  gf_ctx* parent = nullptr;                      // set on the member views of an ensemble
  std::vector<std::unique_ptr<gf_ctx>> members;  // views, created on first request
  std::string err;
};

template <typename F>
static int guarded(gf_ctx* c, F f) {
  if (!c) return GF_EINVAL;
  try {
    f();
    c->err.clear();
    return GF_OK;
  } catch (const GfError& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c->err = e.what();
    return GF_EHIP;
  }
}

static thread_local std::string g_create_err;

extern "C" {

const char* gf_version(void) { return "gfhip 0.1 (gfx950)"; }

const char* gf_error_string(int code) {
  switch (code) {
    case GF_OK: return "ok";
    case GF_EINVAL: return "invalid argument";
    case GF_EUNSUPPORTED: return "unsupported configuration";
    case GF_EHIP: return "HIP runtime error";
    case GF_ENOTFINITE: return "non-finite state";
    case GF_ENODEVICE: return "no HIP device";
    default: return "unknown error";
  }
}

// The schedule table shipped beside the library (schedules.json in the directory of
// libgfhip.so), or the file named by GF_SCHEDULES, is read once, before the first context is
// created. A missing default file is not an error; a bad one is reported by gf_schedules_error.
static std::string g_sched_err;
static void autoload_schedules() {
  static std::once_flag once;
  std::call_once(once, [] {
    std::string path;
    if (const char* e = std::getenv("GF_SCHEDULES")) {
      path = e;
    } else {
      Dl_info info;
      if (dladdr((void*)&autoload_schedules, &info) == 0 || !info.dli_fname) return;
      path = info.dli_fname;
      const size_t slash = path.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/schedules.json";
      if (FILE* f = std::fopen(path.c_str(), "rb")) std::fclose(f);
      else return;
    }
    try { gf::sched_load(path); }
    catch (const std::exception& ex) { g_sched_err = ex.what(); }
  });
}

int gf_ctx_create(gf_ctx** out, const gf_problem_desc* desc) {
  if (!out) return GF_EINVAL;
  autoload_schedules();
  *out = nullptr;
  std::string msg;
  int rc = validate(desc, msg);
  if (rc != GF_OK) { g_create_err = msg; return rc; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_err = "no HIP device"; return GF_ENODEVICE; }
  auto* c = new gf_ctx();
  try {
    if (desc->dtype == GF_F64) c->owned.reset(new Ctx<double>(*desc));
    else c->owned.reset(new Ctx<float>(*desc));
  } catch (const GfError& e) {
    g_create_err = e.what(); delete c; return e.code;
  } catch (const std::exception& e) {
    g_create_err = e.what(); delete c; return GF_EHIP;
  }
  c->impl = c->owned.get();
  *out = c;
  return GF_OK;
}

int gf_ctx_create_batch(gf_ctx** out, const gf_problem_desc* descs, int n) {
  if (!out) return GF_EINVAL;
  autoload_schedules();
  *out = nullptr;
  std::string msg;
  int rc = validate_batch(descs, n, msg);
  if (rc != GF_OK) { g_create_err = msg; return rc; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_err = "no HIP device"; return GF_ENODEVICE; }
  auto* c = new gf_ctx();
  try {
    if (descs[0].dtype == GF_F64) c->owned.reset(new BatchCtx<double>(descs, n));
    else c->owned.reset(new BatchCtx<float>(descs, n));
  } catch (const GfError& e) {
    g_create_err = e.what(); delete c; return e.code;
  } catch (const std::exception& e) {
    g_create_err = e.what(); delete c; return GF_EHIP;
  }
  c->impl = c->owned.get();
  c->members.resize(n);
  *out = c;
  return GF_OK;
}

int gf_batch_size(const gf_ctx* c) { return c ? c->impl->ninstances() : GF_EINVAL; }

int gf_ctx_instance(gf_ctx* c, int i, gf_ctx** member) {
  if (!c || !member) return GF_EINVAL;
  *member = nullptr;
  if (c->parent || i < 0 || i >= c->impl->ninstances()) {
    c->err = "gf_ctx_instance: index out of range or handle is not an ensemble";
    return GF_EINVAL;
  }
  if (c->members.empty()) {   // a plain context is its own only member
    *member = c;
    return GF_OK;
  }
  if (!c->members[i]) {
    c->members[i].reset(new gf_ctx());
    c->members[i]->impl = c->impl->instance(i);
    c->members[i]->parent = c;
  }
  *member = c->members[i].get();
  return GF_OK;
}

void gf_ctx_destroy(gf_ctx* ctx) {
  if (ctx && !ctx->parent) delete ctx;   // member views belong to their ensemble
}

const char* gf_last_error(const gf_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int gf_set_state(gf_ctx* c, const void* s) {
  if (!s) return GF_EINVAL;
  return guarded(c, [&] { c->impl->set_state(s, hipMemcpyHostToDevice); });
}
int gf_get_state(gf_ctx* c, void* s) {
  if (!s) return GF_EINVAL;
  return guarded(c, [&] { c->impl->get_state(s, hipMemcpyDeviceToHost); });
}
int gf_set_state_dev(gf_ctx* c, const void* s) {
  if (!s) return GF_EINVAL;
  return guarded(c, [&] { c->impl->set_state(s, hipMemcpyDeviceToDevice); });
}
int gf_get_state_dev(gf_ctx* c, void* s) {
  if (!s) return GF_EINVAL;
  return guarded(c, [&] { c->impl->get_state(s, hipMemcpyDeviceToDevice); });
}
int gf_set_physical(gf_ctx* c, const void* f) {
  if (!f) return GF_EINVAL;
  return guarded(c, [&] { c->impl->set_physical(f); });
}
int gf_set_psi(gf_ctx* c, const void* psi) {
  if (!c || !psi) return GF_EINVAL;
  return guarded(c, [&] { c->impl->set_psi(psi); });
}
int gf_updatevars(gf_ctx* c, void* out) {
  if (!out) return GF_EINVAL;
  return guarded(c, [&] { c->impl->updatevars(out); });
}
int gf_num_vars(const gf_ctx* c) { return c ? c->impl->nvars() : GF_EINVAL; }
int gf_calcN(gf_ctx* c, void* N, void* sol, double t) {
  if (!N || !sol) return GF_EINVAL;
  return guarded(c, [&] { c->impl->calcN(N, sol, t); });
}
int gf_set_forcing_harmonic(gf_ctx* c, const void* F, double omega, double phase) {
  if (!F) return GF_EINVAL;
  return guarded(c, [&] { c->impl->set_forcing_harmonic(F, omega, phase); });
}
int gf_set_forcing(gf_ctx* c, const void* F) {
  if (!F) return GF_EINVAL;
  return guarded(c, [&] { c->impl->set_forcing(F); });
}
static void check_finite(CtxBase* x) {
  for (int i = 0; i < x->ninstances(); ++i)
    if (x->instance(i)->nonfinite())
      throw GfError(GF_ENOTFINITE, x->ninstances() > 1 ? "the state of ensemble member " + std::to_string(i) +
                                                             " has non-finite values"
                                                       : std::string("the state has non-finite values"));
}
int gf_step(gf_ctx* c, int n) {
  if (n < 0) return GF_EINVAL;
  return guarded(c, [&] { c->impl->step(n); check_finite(c->impl); });
}
int gf_clock(const gf_ctx* c, double* t, int64_t* step) {
  if (!c) return GF_EINVAL;
  if (t) *t = c->impl->t;
  if (step) *step = c->impl->stepno;
  return GF_OK;
}
int gf_set_clock(gf_ctx* c, double t, int64_t step) {
  if (!c) return GF_EINVAL;
  c->impl->set_clock(t, step);
  return GF_OK;
}
int gf_synchronize(gf_ctx* c) { return guarded(c, [&] { c->impl->sync(); check_finite(c->impl); }); }
int gf_timing(gf_ctx* c, int enable) { return guarded(c, [&] { c->impl->timing(enable); }); }
int gf_timing_read(gf_ctx* c, int k, int64_t* n, double* ms) {
  if (!n || !ms) return GF_EINVAL;
  return guarded(c, [&] { c->impl->timing_read(k, n, ms); });
}
int gf_time_kernel(gf_ctx* c, int k, int reps, double* ms) {
  if (!ms) return GF_EINVAL;
  return guarded(c, [&] { c->impl->time_kernel(k, reps, ms); });
}
int gf_diagnostic(gf_ctx* c, int kind, double xi, int nxi, double* out, int nout) {
  if (!out || nout < 1) return GF_EINVAL;
  int n = 0;
  const int rc = guarded(c, [&] { n = c->impl->diagnostic(kind, xi, nxi, out, nout); });
  return rc == GF_OK ? n : rc;
}

int64_t gf_snapshot(gf_ctx* c, void* dst) {
  if (!dst) return GF_EINVAL;
  int64_t ticket = 0;
  const int rc = guarded(c, [&] { ticket = c->impl->snapshot(dst); });
  return rc == GF_OK ? ticket : rc;
}
int gf_set_schedule(gf_ctx* c, int p_k1, int p_k2, int p_k3) {
  const int p[3] = {p_k1, p_k2, p_k3};
  return guarded(c, [&] { c->impl->set_schedule(p); });
}
int gf_get_schedule(gf_ctx* c, int* p_k1, int* p_k2, int* p_k3) {
  int p[3] = {0, 0, 0};
  const int rc = guarded(c, [&] { c->impl->get_schedule(p); });
  if (rc == GF_OK) {
    if (p_k1) *p_k1 = p[0];
    if (p_k2) *p_k2 = p[1];
    if (p_k3) *p_k3 = p[2];
  }
  return rc;
}
int gf_set_splits(gf_ctx* c, int k1_fields, int k2_halves, int k3_columns) {
  const int v[3] = {k1_fields, k2_halves, k3_columns};
  return guarded(c, [&] { c->impl->set_splits(v); });
}
int gf_get_splits(gf_ctx* c, int* k1_fields, int* k2_halves, int* k3_columns) {
  int v[3] = {0, 0, 0};
  const int rc = guarded(c, [&] { c->impl->get_splits(v); });
  if (rc == GF_OK) {
    if (k1_fields) *k1_fields = v[0];
    if (k2_halves) *k2_halves = v[1];
    if (k3_columns) *k3_columns = v[2];
  }
  return rc;
}
int gf_autotune(gf_ctx* c, int steps, double* us) {
  return guarded(c, [&] { c->impl->autotune(steps, us); });
}
int gf_schedules_save(const char* path) {
  if (!path) return GF_EINVAL;
  try { gf::sched_save(path); g_sched_err.clear(); return GF_OK; }
  catch (const std::exception& e) { g_sched_err = e.what(); return GF_EINVAL; }
}
int gf_schedules_load(const char* path) {
  if (!path) return GF_EINVAL;
  autoload_schedules();
  try { const int n = gf::sched_load(path); g_sched_err.clear(); return n; }
  catch (const std::exception& e) { g_sched_err = e.what(); return GF_EINVAL; }
}
void gf_schedules_clear(void) { autoload_schedules(); gf::sched_clear(); }
const char* gf_schedules_error(void) { return g_sched_err.c_str(); }
int gf_use_graphs(gf_ctx* c, int enable) {
  return guarded(c, [&] { c->impl->set_graphs(enable); });
}
int gf_batch_persistent(gf_ctx* c, int enable) {
  return guarded(c, [&] { c->impl->set_persistent(enable); });
}
int gf_snapshot_wait(gf_ctx* c, int64_t ticket) {
  return guarded(c, [&] { c->impl->snapshot_wait(ticket); });
}
void* gf_host_alloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0 || dw::host_malloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}
void gf_host_free(void* p) {
  if (p) (void)dw::host_free(p);
}
void* gf_device_alloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0 || dw::malloc(&p, bytes) != hipSuccess) return nullptr;
  if (dw::zero_sync(p, bytes) != hipSuccess) {
    (void)dw::free(p);
    return nullptr;
  }
  return p;
}
void gf_device_free(void* p) {
  if (p) (void)dw::free(p);
}
int gf_device_copy(void* dst, const void* src, size_t bytes) {
  if (!dst || !src) return GF_EINVAL;
  return dw::copy(dst, src, bytes, hipMemcpyDefault) == hipSuccess ? GF_OK : GF_EHIP;
}

int gf_diagnostic_async(gf_ctx* c, int kind, double xi, int nxi, double* out_dev, int nout) {
  if (!out_dev || nout < 1) return GF_EINVAL;
  int n = 0;
  const int rc = guarded(c, [&] { n = c->impl->diagnostic_async(kind, xi, nxi, out_dev, nout); });
  return rc == GF_OK ? n : rc;
}

int gf_diagnostic_async_noalloc(gf_ctx* c, int kind, double xi, int nxi, double* out_dev, int nout) {
  if (!out_dev || nout < 1) return GF_EINVAL;
  int n = 0;
  const int rc = guarded(c, [&] { n = c->impl->diagnostic_noalloc(kind, xi, nxi, out_dev, nout); });
  return rc == GF_OK ? n : rc;
}

int gf_adjoint_tape_budget(gf_ctx* c, size_t bytes) {
  return guarded(c, [&] { c->impl->set_tape_budget(bytes); });
}

int gf_adjoint_grad(gf_ctx* c, const double* target, int nsteps, int checkpoint_every, double* dJ_dzeta0,
                    double* dJ_dnu, double* J) {
  if (!target) return GF_EINVAL;
  return guarded(c, [&] { c->impl->adjoint_grad(target, nsteps, checkpoint_every, dJ_dzeta0, dJ_dnu, J); });
}

int gf_kernel_bytes(const gf_ctx* c, int k, int full, double* bytes) {
  if (!c || !bytes || k < 0 || k > 2) return GF_EINVAL;
  *bytes = c->impl->kernel_bytes(k, full);
  return GF_OK;
}

}  // extern "C"
