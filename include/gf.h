/*
 * gf.h -- C ABI of the MI355X pseudospectral time-stepping library (libgfhip.so).
 *
 * This is the drop-in boundary for GeophysicalFlows.jl's hot path: the
 * per-step nonlinear term calcN! of TwoDNavierStokes, SingleLayerQG and
 * MultiLayerQG and the stepper stages behind FourierFlows' stepforward!.
 * Each entry point names the reference interface it replaces (paths relative
 * to the GeophysicalFlows.jl v0.16.5 source tree; [FF] = FourierFlows.jl,
 * the reference's dependency, called at the cited line).
 *
 * Conventions
 *  - Every call returns GF_OK (0) or a negative GF_E* code; nothing throws or
 *    aborts across the ABI.  gf_last_error(ctx) gives the message.
 *  - Spectral arrays crossing the ABI use the reference's memory layout: the
 *    column-major Julia array sol[kr, l(, layer)] with interleaved complex
 *    (re, im) of the problem's precision (double or float), nkr = nx/2 + 1.
 *    Physical arrays are f[x, y(, layer)], x fastest.
 *  - "_dev" entry points take device pointers (HIP, same process); the others
 *    take host pointers and copy synchronously.
 *  - One context owns one HIP stream and all of its device memory; a context
 *    is not thread-safe, distinct contexts may be used concurrently.
 */
#ifndef GF_H
#define GF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GF_OK = 0,
  GF_EINVAL = -1,      /* invalid argument or descriptor                     */
  GF_EUNSUPPORTED = -2,/* configuration not supported by this build          */
  GF_EHIP = -3,        /* HIP runtime error (allocation, launch, copy)        */
  GF_ENOTFINITE = -4,  /* non-finite values detected in the state            */
  GF_ENODEVICE = -5    /* no HIP device                                        */
};

enum { GF_MODEL_TWODNS = 0, GF_MODEL_SLQG = 1, GF_MODEL_MLQG = 2 };
enum { GF_F64 = 0, GF_F32 = 1 };
enum {
  GF_ETDRK4 = 0, GF_FILTERED_ETDRK4 = 1, GF_RK4 = 2, GF_FILTERED_RK4 = 3,
  GF_FORWARD_EULER = 4, GF_FILTERED_FORWARD_EULER = 5,
  GF_AB3 = 6, GF_FILTERED_AB3 = 7   /* [FF] AB3TimeStepper (test/runtests.jl:58); not in ensembles */
};
enum { GF_FORCING_NONE = 0, GF_FORCING_FIXED = 1, GF_FORCING_RING = 2 };

/* Problem description: the keyword arguments of
 *   TwoDNavierStokes.Problem  (src/twodnavierstokes.jl:71-89),
 *   SingleLayerQG.Problem     (src/singlelayerqg.jl:80-102),
 *   MultiLayerQG.Problem      (src/multilayerqg.jl:87-114),
 * plus the stepper and dt handed to FourierFlows.Problem
 * (src/twodnavierstokes.jl:99). Arrays are borrowed for the call only. */
typedef struct gf_problem_desc {
  int model;                 /* GF_MODEL_*                                         */
  int dtype;                 /* GF_F64 | GF_F32  (keyword T)                        */
  int nx, ny, nlayers;       /* nx, ny each a power of two 16..4096 or even 18..256; nlayers: MLQG 1..8, else 1 */
  double Lx, Ly;
  double aliased_fraction;   /* default 1/3                                         */
  int stepper;               /* GF_ETDRK4 ...                                       */
  double dt;
  double nu; int nnu;        /* (hyper)viscosity nu |k|^(2 nnu)                     */
  double mu; int nmu;        /* TwoDNS: mu |k|^(2 nmu); SLQG: linear drag; MLQG: bottom drag */
  /* SingleLayerQG */
  double beta;               /* SLQG, MLQG                                          */
  double deformation_radius; /* SLQG: <= 0 or inf => barotropic                     */
  double U;                  /* SLQG constant U (used when Uy == NULL)               */
  const double* eta;         /* topographic PV, nx*ny (x fastest), or NULL          */
  const double* Uy;          /* SLQG: U(y), ny values, or NULL.
                                MLQG: U per layer, nlayers*ny (y fastest), or NULL  */
  /* MultiLayerQG */
  double f0;
  const double* H;           /* nlayers (unused, nullable, for one layer)           */
  const double* b;           /* nlayers (Boussinesq buoyancy; as H)                 */
  double topo_grad_x, topo_grad_y;   /* topographic_pv_gradient                     */
  int linear;                /* MLQG linearised equations (src/multilayerqg.jl:358-362, calcNlinear!) */
  /* forcing (calcF / calcFq) */
  int forcing;               /* GF_FORCING_*                                        */
  double kf, dkf, eps;       /* ring spectrum (examples/twodnavierstokes_stochasticforcing.jl:53-66) */
  uint64_t seed;             /* counter-based generator key                         */
  int instance;              /* generator stream id (ensemble member)               */
} gf_problem_desc;

typedef struct gf_ctx gf_ctx;

/* FourierFlows.Problem(Equation(params, grid), stepper, dt, grid, vars, params)
 * with the module's Params/Equation (src/twodnavierstokes.jl:97-99,149-154;
 * src/singlelayerqg.jl:104-120,230-252; src/multilayerqg.jl:120-134,255-377). */
int gf_ctx_create(gf_ctx** out, const gf_problem_desc* desc);
void gf_ctx_destroy(gf_ctx* ctx);
const char* gf_last_error(const gf_ctx* ctx);
const char* gf_error_string(int code);

/* prob.sol read/write (Julia layout, host). */
int gf_set_state(gf_ctx* ctx, const void* sol);
int gf_get_state(gf_ctx* ctx, void* sol);
/* same on device pointers (AMDGPU.jl ROCArray / torch tensors in this process) */
int gf_set_state_dev(gf_ctx* ctx, const void* sol_dev);
int gf_get_state_dev(gf_ctx* ctx, void* sol_dev);

/* set_zeta!/set_q! (src/twodnavierstokes.jl:352-360, src/singlelayerqg.jl:525-534,
 * src/multilayerqg.jl:845-851): forward rfft of the physical field on the
 * device, zero mean for TwoDNS and MLQG, then updatevars!'s in-place dealias. */
int gf_set_physical(gf_ctx* ctx, const void* field);

/* MultiLayerQG set_psi! (src/multilayerqg.jl:873-889): psih = rfft(psi) per layer, qh = S psih
 * per wavenumber (pvfromstreamfunction!, 520-552; one layer: qh = -Krsq psih), then set_q! of
 * irfft(qh): rfft, zero mean, updatevars!'s dealias!.  psi holds nlayers*nx*ny values of the
 * problem's precision, layer slowest, x fastest; host memory.  GF_EUNSUPPORTED for the other
 * models. */
int gf_set_psi(gf_ctx* ctx, const void* psi);

/* updatevars! (src/twodnavierstokes.jl:331-345, src/singlelayerqg.jl:502-518,
 * src/multilayerqg.jl:819-831): dealias sol in place, write the physical fields
 * (TwoDNS: zeta, u, v; SLQG/MLQG: q, psi, u, v -- each nx*ny per layer, layer
 * slowest) into out, which holds nfields*nlayers*nx*ny values. */
int gf_updatevars(gf_ctx* ctx, void* out);
int gf_num_vars(const gf_ctx* ctx);

/* calcN! as a unit hook (src/twodnavierstokes.jl:286-294, src/singlelayerqg.jl:443-451,
 * src/multilayerqg.jl:651-663): N = calcN(sol) with sol dealiased in place (P1).
 * Both arrays in Julia layout on the host; sol is read and written back. */
int gf_calcN(gf_ctx* ctx, void* N, void* sol, double t);

/* fixed forcing table Fh (a calcF! that copies a constant array, e.g.
 * test/test_twodnavierstokes.jl:239-243); Julia layout, host. */
int gf_set_forcing(gf_ctx* ctx, const void* Fh);

/* Time-harmonic forcing F(t) = Fh cos(omega t + phase), evaluated at each stage's time as
 * calcF! is (the calcF! of test/test_twodnavierstokes.jl:141-145 and
 * test/test_singlelayerqg.jl:127-131 is Fh cos(2t)). For problems created with
 * GF_FORCING_FIXED. The work diagnostics use the last evaluation's F, which is zero before the
 * first step. Ensembles do not support it: step such a member alone. */
int gf_set_forcing_harmonic(gf_ctx* ctx, const void* Fh, double omega, double phase);

/* [FF] stepforward!(prob, nsteps): nsteps of the configured stepper, queued on the context's
 * stream (the call returns before they finish; every read of the state waits for them).
 * Failure detection: the last stage of every step flags a non-finite value of the new state.
 * gf_step returns GF_ENOTFINITE (after queuing its steps) when a step that has already finished
 * was flagged; gf_synchronize waits and then returns GF_ENOTFINITE if any step since the state
 * was last set (gf_set_state, gf_set_state_dev, gf_set_physical) was flagged. No counterpart in
 * the reference, which steps non-finite states on silently.
 * With graph replay on, every step but the call's last replays a HIP graph of the whole step
 * (all stages, or all stages of every ensemble member), captured at the first such step. */
int gf_step(gf_ctx* ctx, int nsteps);

/* Turn the graph replay of gf_step on (1) or off (0); results are identical. On by default
 * for ensembles, off by default for single problems. No counterpart in the reference. */
int gf_use_graphs(gf_ctx* ctx, int enable);

/* Ensembles: run the three kernel classes of each stage (K1, K2, K3 over every member) as one
 * persistent cooperative launch of the workgroups the GPU holds at once, with grid-wide
 * barriers between the classes (1), or as three launches (0, the default). Results agree to
 * rounding; the persistent form runs without graph replay and is slower on MI355X (a
 * cooperative launch and its grid barriers cost more than the launches they replace).
 * GF_EINVAL for a single problem. No counterpart in the reference, whose ensemble members are
 * separate problems. */
int gf_batch_persistent(gf_ctx* ctx, int enable);

/* Radix schedules (no counterpart in the reference, whose FFTs are FFTW's or cuFFT's plans).
 * For single-layer problems the FFT of each calcN kernel runs with P points per thread:
 *   K1 (inverse y-pass), K2 (x-pass with the products), K3 (forward y-pass and stage).
 * P is 4 (up to 2048 points), 8 (from 32 points) or 16 (from 1024 points); 0 restores the
 * built-in choice (4 below 32 points, 16 from 4096, else 8). Results differ from another
 * schedule's only by rounding. gf_get_schedule reports the P in use. MultiLayerQG, and grids
 * with a length that is not a power of two, have a fixed schedule. */
int gf_set_schedule(gf_ctx* ctx, int p_k1, int p_k2, int p_k3);
int gf_get_schedule(gf_ctx* ctx, int* p_k1, int* p_k2, int* p_k3);

/* Workgroup splits for single-layer problems (0 off, 1 on), part of the schedule:
 *   k1_fields:  K1 gives each of u, v, q its own row of workgroups;
 *   k2_halves:  K2 forms u q and v q in two rows of workgroups, each transforming q itself (1),
 *               or walks row pairs persistently, each field's rows loaded straight into LDS one
 *               field ahead of the transforms (2; TwoDNavierStokes and SingleLayerQG without a
 *               varying U, nx / P_K2 >= 64; the unsplit pass runs elsewhere);
 *   k3_columns: K3 gives each spectral column its own workgroup instead of a column pair, where
 *               a column's ny / P_K3 threads fill a 64-lane wave (ignored elsewhere).
 * Workgroups are regrouped only: the results are the same bit for bit, except that k2_halves = 2
 * may round FP64 products differently (agreement to about 1e-16 relative). */
int gf_set_splits(gf_ctx* ctx, int k1_fields, int k2_halves, int k3_columns);
int gf_get_splits(gf_ctx* ctx, int* k1_fields, int* k2_halves, int* k3_columns);

/* The autotuner. It times `steps` whole steps under every offered schedule on this problem:
 * P in {4, 8, 16} for each of K1, K2, K3, with the K1 and K3 splits on or off and the K2 split
 * 0, 1 or 2, up to 324 combinations. It runs two passes, the second in reverse order, keeping
 * each combination's best, then applies the fastest. The state, its stepper history and the
 * clock are restored afterwards. The choice is recorded for (model, nlayers, nx, ny, dtype):
 * contexts created later in the process start with it. us (nullable) receives 324 step times in
 * microseconds, indexed ((((i1 * 3 + i2) * 3 + i3) * 2 + split_K1) * 3 + split_K2) * 2 + split_K3,
 * where i = 0, 1, 2 for P = 4, 8, 16; NaN where a combination is not offered. */
int gf_autotune(gf_ctx* ctx, int steps, double* us);

/* The process-wide table of tuned schedules as JSON ("gf-schedules-1": one entry per (model,
 * nlayers, nx, ny, dtype) with its P triple, its splits and the measured step time). gf_schedules_load returns the
 * number of entries read (>= 0), or GF_EINVAL for an unreadable or malformed file, or for a
 * schedule this build does not offer. gf_schedules_error() then names the problem.
 * Loaded entries apply to contexts created afterwards. Before the first context is created,
 * the library reads the file named by the environment variable GF_SCHEDULES, or else
 * schedules.json in the directory of libgfhip.so (the table tuned on MI355X that ships with
 * it) if that file exists. */
int gf_schedules_save(const char* path);
int gf_schedules_load(const char* path);
void gf_schedules_clear(void);
const char* gf_schedules_error(void);

/* [FF] prob.clock */
int gf_clock(const gf_ctx* ctx, double* t, int64_t* step);
int gf_set_clock(gf_ctx* ctx, double t, int64_t step);

/* Wait for all work queued on the context's stream; GF_ENOTFINITE as described at gf_step. */
int gf_synchronize(gf_ctx* ctx);

/* Per-kernel timing of gf_step with HIP events that each launch carries (hipExtLaunchKernel
 * stamps the event with the kernel's end): a launch's time runs from the end of the timed
 * launch before it on the context's stream to its own end, i.e. its execution plus the short
 * dispatch gap, with no event packets between the kernels. gf_timing(ctx, 1) starts recording
 * (and replays no graphs meanwhile), gf_timing_read() returns, for kernel
 * class k (0 = y-inverse, 1 = x-pass, 2 = y-forward + stage), the number of
 * launches and the summed milliseconds since recording started. */
int gf_timing(gf_ctx* ctx, int enable);
int gf_timing_read(gf_ctx* ctx, int k, int64_t* launches, double* ms);

/* Execution time of kernel class k (0 = y-inverse, 1 = x-pass, 2 = y-forward with the calcN
 * epilogue) on the current state: `reps` launches back to back between two events on the
 * context's stream, *ms_per_launch = their mean in milliseconds. Back-to-back launches of one
 * kernel leave no gap between them, so this is the kernel's execution time as a kernel trace
 * reports it; gf_timing's per-launch events add the dispatch gap that a profiled launch's
 * completion brings (about 4.5 us per kernel at 2048^2). The state and clock are unchanged.
 * Single problems only (GF_EUNSUPPORTED for an ensemble). No counterpart in the reference:
 * a measurement hook, like gf_timing. */
int gf_time_kernel(gf_ctx* ctx, int k, int reps, double* ms_per_launch);

/* Bytes of HBM each kernel class must move per launch in the fused
 * algorithm (compulsory traffic of its inputs and outputs), for the given
 * step phase: full = 1 for the last step of a gf_step call. */
int gf_kernel_bytes(const gf_ctx* ctx, int k, int full, double* bytes);

/* Ensembles of independent problems advanced together.
 *
 * No counterpart in the reference: there an ensemble is a Julia loop of
 * stepforward!(prob, n) over separate Problems (one calcN! per member per stage).
 * gf_ctx_create_batch builds n members from n descriptors (one precision, FP64 or
 * FP32, for all; TwoDNS, SingleLayerQG or one-layer MultiLayerQG; square grids with
 * n in 32..512, sizes may differ; one stepper for all, AB3 included; any forcing
 * kind but the time-harmonic one).  gf_step on the ensemble advances every member
 * with one launch per kernel class and stage covering all of them.  With ring
 * forcing or AB3 every member must be at the same step (GF_EINVAL otherwise): the
 * forcing's step number and the AB3 start-up are shared.
 * gf_ctx_instance returns a borrowed handle for member i, valid until the
 * ensemble is destroyed (gf_ctx_destroy on it does nothing); every
 * single-problem entry point works on it, and gf_step on it advances that member
 * alone (the sequential path).  For a plain context gf_batch_size is 1 and
 * gf_ctx_instance(ctx, 0, &m) gives ctx itself. */
int gf_ctx_create_batch(gf_ctx** out, const gf_problem_desc* descs, int n);
int gf_batch_size(const gf_ctx* ctx);
int gf_ctx_instance(gf_ctx* ensemble, int i, gf_ctx** member);

/* Domain-averaged diagnostics, computed on the device from the raw state as the reference
 * computes them. Each is a Parseval sum over the rfft half plane.
 *  - TwoDNavierStokes (src/twodnavierstokes.jl:374-535):
 *      energy, enstrophy, palinstrophy;
 *      energy_dissipation(prob, xi, nxi) and enstrophy_dissipation (GF_DIAG_*_DISSIPATION with
 *        xi, nxi; *_hyperviscosity passes nu, nnu and *_hypoviscosity passes mu, nmu);
 *      energy_work, enstrophy_work.
 *  - SingleLayerQG (src/singlelayerqg.jl:546-715):
 *      kinetic_energy, potential_energy, energy, enstrophy (of q + eta);
 *      energy/enstrophy_dissipation (xi = nu, nxi = nnu);
 *      energy/enstrophy_drag (xi = mu);
 *      energy/enstrophy_work.
 *  - MultiLayerQG energies (src/multilayerqg.jl:918-960): GF_DIAG_LAYER_ENERGIES writes
 *    KE_1..KE_n, then PE_{3/2}..PE_{n-1/2} (one layer: the energy only).
 *  - MultiLayerQG fluxes (src/multilayerqg.jl:977-1022): GF_DIAG_FLUXES writes the lateral
 *    eddy fluxes lateralfluxes_1..n, then the vertical fluxes verticalfluxes_{3/2}..{n-1/2}
 *    (one layer: the lateral flux only). As the reference's fluxes calls updatevars!, it
 *    dealiases the state in place first.
 * For stochastic forcing the work uses (prevsol + sol)/2 and the last step's forcing.
 * Returns the number of values written (>= 1) or a negative GF_E* code; out holds nout
 * doubles. */
enum {
  GF_DIAG_ENERGY = 0, GF_DIAG_KINETIC_ENERGY = 1, GF_DIAG_POTENTIAL_ENERGY = 2, GF_DIAG_ENSTROPHY = 3,
  GF_DIAG_PALINSTROPHY = 4, GF_DIAG_ENERGY_DISSIPATION = 5, GF_DIAG_ENSTROPHY_DISSIPATION = 6,
  GF_DIAG_ENERGY_DRAG = 7, GF_DIAG_ENSTROPHY_DRAG = 8, GF_DIAG_ENERGY_WORK = 9, GF_DIAG_ENSTROPHY_WORK = 10,
  GF_DIAG_LAYER_ENERGIES = 11, GF_DIAG_FLUXES = 12
};
int gf_diagnostic(gf_ctx* ctx, int kind, double xi, int nxi, double* out, int nout);

/* The same diagnostic, written to device memory: out_dev holds nout doubles on the context's
 * device. The reduction is queued on the context's stream and the call returns without waiting,
 * so a time series ([FF] Diagnostic, increment!) stays on the device until it is read. The value
 * equals gf_diagnostic's bit for bit. Returns the number of values written or a GF_E* code. */
int gf_diagnostic_async(gf_ctx* ctx, int kind, double xi, int nxi, double* out_dev, int nout);

/* gf_diagnostic_async without any allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory at the call (code under HIP stream capture, serving loops,
 * harnesses that forbid hipMalloc). gf_diagnostic_async allocates its partial-sum buffer on its
 * first call; this variant keeps the partial sums in a static device array of the library
 * (1024 x 15 doubles, 120 KiB per GPU, held until the process ends), whose address on each device
 * is resolved when the first context on that device is created. Same kernels, same result bit for
 * bit. It calls no hipMalloc, hipFree or allocating routine. Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (on different streams or threads) race, so
 * serialise them. Returns the number of values written or a GF_E* code. */
int gf_diagnostic_async_noalloc(gf_ctx* ctx, int kind, double xi, int nxi, double* out_dev, int nout);

/* Snapshot of the state for output ([FF] Output / saveoutput, which store prob.sol; e.g.
 * examples/twodnavierstokes_decaying.jl:111-115). Queues a copy of sol, in the Julia layout, to
 * the host buffer dst and returns a ticket >= 0 without waiting (or a negative GF_E* code).
 * The copy runs on a second stream: gf_step calls made after gf_snapshot overlap it. dst
 * must stay valid until gf_snapshot_wait(ctx, ticket) returns. The overlap needs pinned memory
 * (gf_host_alloc); with pageable memory the copy is synchronous. */
int64_t gf_snapshot(gf_ctx* ctx, void* dst);
/* Wait until snapshot `ticket` is in host memory; ticket < 0 waits for every snapshot. */
int gf_snapshot_wait(gf_ctx* ctx, int64_t ticket);
/* Pinned host memory for snapshot buffers (hipHostMalloc); NULL on failure. */
void* gf_host_alloc(size_t bytes);
void gf_host_free(void* p);
/* Device memory for callers without a device allocator of their own (the Python mirror's
 * diagnostic series): zeroed hipMalloc, hipFree, and a synchronous copy in any direction. The
 * copy does not wait for a context's stream: call gf_synchronize first. */
void* gf_device_alloc(size_t bytes);
void gf_device_free(void* p);
int gf_device_copy(void* dst, const void* src, size_t bytes);

/* Gradient of J = 1/2 sum_{x,y} (zeta(T) - target)^2 for the problem's current state as
 * zeta0 (set with gf_set_physical), where zeta(T) = irfft2(M sol_T) after nsteps ETDRK4 steps.
 *
 * No counterpart in the reference, which has no adjoint. It is the discrete adjoint of this
 * library's own ETDRK4 step, including calcN!'s in-place dealias!.
 *  - Scope: unforced TwoDNavierStokes, FP64, square grids.
 *  - The forward pass keeps a checkpoint every checkpoint_every steps. The backward sweep
 *    recomputes each segment into a device tape of 5 spectral fields per step (the step's input
 *    and its four calcN outputs; the adjoint rebuilds the stage inputs from them).
 *    checkpoint_every >= nsteps tapes the whole run once.
 *  - Outputs:
 *      dJ_dzeta0: nx*ny doubles, physical layout, x fastest; nullable;
 *      dJ_dnu: scalar; nullable, in which case it is skipped;
 *      J: the loss; nullable.
 *  - The context's state and clock are left unchanged. */
int gf_adjoint_grad(gf_ctx* ctx, const double* target, int nsteps, int checkpoint_every,
                    double* dJ_dzeta0, double* dJ_dnu, double* J);

/* Device memory the tape of gf_adjoint_grad may take (bytes). No counterpart in the reference.
 * The tape is a pool of segment slots: the forward pass tapes segment s into slot s mod P, so the
 * last P segments stay taped and the backward sweep recomputes only the earlier ones from their
 * checkpoints. P is the number of whole segments the budget affords, at least two (one when
 * other gradients are in flight in the process) and at most the number of segments. One
 * segment of k steps takes (5 k + 1) fields of kmax_pair * ny complex values, kmax_pair being
 * the surviving columns rounded up to a pair (170 at 512^2: 349 MB for 50 steps). 0, the
 * default: the minimum. With 288 GB of HBM a budget near the whole run removes recomputation. */
int gf_adjoint_tape_budget(gf_ctx* ctx, size_t bytes);

/* Library version string. */
const char* gf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GF_H */
