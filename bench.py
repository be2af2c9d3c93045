"""Benchmark: TwoDNavierStokes 2048^2 FP64 ETDRK4, stochastic ring forcing.

Workload (the headline configuration): nx = ny = 2048, L = 2 pi, Float64,
ETDRK4, dt = 0.005, hyperviscosity nu = 1e-21 (n_nu = 4), linear drag mu = 0.1
(n_mu = 0), white-in-time forcing on the wavenumber ring k_f = 14 (width 1.5,
injection rate 0.1), zeta0 = 0.  One step is one ETDRK4 step (four calcN!
evaluations and the stage arithmetic).  All state is resident in HBM.

Output: one JSON line with grid-point-steps/s (``value``), the roofline of
the dominant kernel (its execution time from HIP events on the library's
stream around back-to-back launches on the state the timed steps left, the
per-launch events inside the steps beside it; algorithmic bytes from
gf_kernel_bytes), the whole-step rate against
the 64 S-per-step algorithmic traffic model, and the CPU baseline (the NumPy /
SciPy restatement in oracle/, timed on a bounded sample on this host).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NX = 2048
CFG = dict(nx=NX, Lx=2 * np.pi, nu=1e-21, nnu=4, mu=0.1, nmu=0, dt=0.005, kf=14.0, dkf=1.5, eps=0.1, seed=2)
KERNELS = ("y_inverse_K1", "x_pass_K2", "y_forward_stage_K3")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "twodns_2048_fp64_etdrk4_hbm_traffic.md")


def symbols(sched, splits):
    """Kernel symbols of the FP64 2048^2 headline path, as rocprofv3 names them, under the radix
    schedule sched = (P_K1, P_K2, P_K3) (one thread group per workgroup at this size) and the
    workgroup splits (K2 split 2 is the staged x pass; K3's sixth template argument is its
    columns per workgroup, 1 or 2), as a list per kernel class."""
    p1, p2, p3 = sched
    cpw = 1 if splits[2] and 2048 // p3 >= 64 else 2
    if cpw == 1 and p3 == 8:
        # one kernel per ETDRK4 stage (the seventh argument is the stage's epilogue code)
        k3 = [f"k3_kernel<double, 2048, 8, 1, 1, 1, {e}, false, false>" for e in (2, 3, 4, 5)]
    else:
        k3 = [f"k3_kernel<double, 2048, {p3}, 1, 1, {cpw}, -1, false, false>"]
    k2 = f"k2_staged_kernel<double, 2048, {p2}>" if splits[1] == 2 else f"k2_kernel<double, 2048, {p2}, 1, 0, 1>"
    return {"y_inverse_K1": [f"k1_kernel<double, 2048, {p1}, 1, 1, 0, 0, false>"],
            "x_pass_K2": [k2],
            "y_forward_stage_K3": k3}


def measured_traffic(sched, splits):
    """HBM bytes per launch of each headline kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by the copy-kernel calibration); a kernel is left out when the
    summary was taken under another schedule, and {} is returned if the file is absent."""
    try:
        txt = open(TRAFFIC_FILE).read()
        block = txt.split("```json", 1)[1].split("```", 1)[0]
        tb = json.loads(block)["traffic_bytes"]
        # a kernel class launched as several kernels (K3's stages, once each per step): the
        # mean over them is the class's bytes per launch
        return {k: sum(tb[x] for x in syms) / len(syms) for k, syms in symbols(sched, splits).items()
                if all(x in tb for x in syms)}
    except (OSError, IndexError, KeyError, ValueError):
        return {}


# back-to-back launches per kernel class for the execution times behind the roofline, after
# KERNEL_WARM untimed launches of the same kernel (its time falls by several percent while the
# GPU's clocks settle under sustained load)
KERNEL_REPS = 50
KERNEL_WARM = 200

# steps of the sequential comparison of the ensemble (every member stepped alone, one after another)
ENSEMBLE_SEQ_STEPS = 10

# Steps before the untimed warm-up steps, so that the timed steps do not start on a GPU whose
# clocks are still ramping up from idle (about 0.1 s of work at 2048^2)
CLOCK_RAMP_STEPS = 300


def _step(gf, prob, n):
    """gf_step that accepts a non-finite state: the headline configuration (dt = 0.005 at 2048^2)
    exceeds the advective CFL limit within about 200 steps, and the kernels' timing does not
    depend on the values (the library reports it as GF_ENOTFINITE)."""
    rc = gf._lib.lib.gf_step(prob._ctx, int(n))
    if rc not in (gf._lib.GF_OK, gf._lib.GF_ENOTFINITE):
        gf._lib.check(rc, prob._ctx)


def _sync(gf, prob):
    rc = gf._lib.lib.gf_synchronize(prob._ctx)
    if rc not in (gf._lib.GF_OK, gf._lib.GF_ENOTFINITE):
        gf._lib.check(rc, prob._ctx)


def gpu_run(gf, steps, warmup, T, dt=CFG["dt"]):
    lib = gf._lib.lib

    def problem():
        p = gf.TwoDNavierStokes.Problem(
            gf.GPU(), nx=CFG["nx"], Lx=CFG["Lx"], ν=CFG["nu"], nν=CFG["nnu"], μ=CFG["mu"], nμ=CFG["nmu"],
            dt=dt, stepper="ETDRK4", T=T,
            calcF=gf.RingForcing(CFG["kf"], CFG["dkf"], CFG["eps"], seed=CFG["seed"]))
        gf.TwoDNavierStokes.set_ζ(p, np.zeros((NX, NX)))
        return p

    # the measured problem starts at zeta0 = 0; a twin takes the clock-ramp steps, so that the GPU
    # is busy right up to the warm-up steps (setting the state from the host would leave it idle)
    prob, ramp = problem(), problem()
    _step(gf, ramp, CLOCK_RAMP_STEPS)
    _sync(gf, ramp)
    if warmup:
        _step(gf, prob, warmup)
    _sync(gf, prob)
    t0 = time.perf_counter()
    _step(gf, prob, steps)
    _sync(gf, prob)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    finite = bool(np.isfinite(prob.sol).all())
    # second pass with an event bound to each launch on the library's stream (gf_timing): a
    # launch's interval also holds the dispatch gap that its profiled completion adds
    lib.gf_timing(prob._ctx, 1)
    _step(gf, prob, steps)
    _sync(gf, prob)
    kern = {}
    for k, name in enumerate(KERNELS):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        lib.gf_timing_read(prob._ctx, k, ctypes.byref(n), ctypes.byref(ms))
        b0, b1 = ctypes.c_double(), ctypes.c_double()
        lib.gf_kernel_bytes(prob._ctx, k, 0, ctypes.byref(b0))
        lib.gf_kernel_bytes(prob._ctx, k, 1, ctypes.byref(b1))
        in_step_s = ms.value * 1e-3 / max(n.value, 1)
        # the kernel's execution time: KERNEL_REPS launches back to back on the state the timed
        # steps left, between two events on the same stream (gf_time_kernel), on the retained
        # columns of an interior step (K3 with the calcN epilogue)
        iso = ctypes.c_double()
        gf._lib.check(lib.gf_time_kernel(prob._ctx, k, KERNEL_WARM, ctypes.byref(iso)), prob._ctx)
        gf._lib.check(lib.gf_time_kernel(prob._ctx, k, KERNEL_REPS, ctypes.byref(iso)), prob._ctx)
        avg_s = iso.value * 1e-3
        kern[name] = dict(launches=n.value, avg_us=avg_s * 1e6, bytes_per_launch=b0.value,
                          achieved_GBps=b0.value / avg_s / 1e9, timed_launches=KERNEL_REPS,
                          in_step_event_us=in_step_s * 1e6)
    lib.gf_timing(prob._ctx, 0)
    kern["schedule"] = list(gf.schedule(prob))
    kern["splits"] = list(gf.splits(prob))
    return elapsed, kern, finite


# batched SingleLayerQG beta-plane ensemble: 256 x 64^2, 192 x 128^2, 64 x 256^2 (FP64, FilteredRK4);
# initial energy 0.08 as in the reference's beta-plane example (examples/singlelayerqg_betadecay.jl:68):
# unit energy at k_peak = 8 exceeds the RK4 advective limit of the 256^2 members at dt = 0.01
ENSEMBLE = ((64, 256), (128, 192), (256, 64))
ENSEMBLE_E0 = 0.08


def ensemble_members(gf):
    members, q0 = [], []
    i = 0
    for n, count in ENSEMBLE:
        for _ in range(count):
            rng = np.random.default_rng(3_000_000 + i)
            grid = gf.problem.Grid(n, n, 2 * np.pi, 2 * np.pi, 1 / 3, np.float64)
            eta = gf.utils.gaussian_bump(grid, 0.1, 2 * np.pi / 10, rng.uniform(-np.pi, np.pi, 2))
            members.append((gf.SingleLayerQG.Problem, dict(nx=n, Lx=2 * np.pi, β=10.0, μ=0.01, eta=eta, dt=0.01,
                                                           stepper="FilteredRK4")))
            q0.append(gf.peakedisotropicspectrum(grid, 8, ENSEMBLE_E0, rng=rng))
            i += 1
    return members, q0


def ensemble_run(gf, steps, seq_steps):
    """Batched ensemble steps against the same members stepped one after another."""
    lib = gf._lib.lib
    members, q0 = ensemble_members(gf)
    ens = gf.Ensemble(members)
    for p, q in zip(ens.problems, q0):
        gf.SingleLayerQG.set_q(p, q)
    pts = sum(n * n * c for n, c in ENSEMBLE)
    gf.stepforward(ens, 3)
    ens.synchronize()
    t0 = time.perf_counter()
    gf.stepforward(ens, steps)
    ens.synchronize()
    tb = time.perf_counter() - t0
    # the same steps with each stage as one persistent cooperative launch (gf_batch_persistent)
    ens.persistent(True)
    gf.stepforward(ens, 3)
    ens.synchronize()
    t0 = time.perf_counter()
    gf.stepforward(ens, steps)
    ens.synchronize()
    tp = time.perf_counter() - t0
    ens.persistent(False)
    lib.gf_timing(ens._ctx, 1)
    gf.stepforward(ens, steps)
    ens.synchronize()
    kern = {}
    for k, name in enumerate(KERNELS):
        n, ms, b0 = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        lib.gf_timing_read(ens._ctx, k, ctypes.byref(n), ctypes.byref(ms))
        lib.gf_kernel_bytes(ens._ctx, k, 0, ctypes.byref(b0))
        avg_s = ms.value * 1e-3 / max(n.value, 1)
        kern[name] = dict(launches=n.value, avg_us=avg_s * 1e6, bytes_per_launch=b0.value,
                          achieved_GBps=b0.value / avg_s / 1e9)
    lib.gf_timing(ens._ctx, 0)
    for p in ens.problems:                      # one member at a time, single-instance kernels
        gf.stepforward(p, 1)
    ens.synchronize()
    t0 = time.perf_counter()
    for p in ens.problems:
        gf.stepforward(p, seq_steps)
    ens.synchronize()
    ts = time.perf_counter() - t0
    finite = all(bool(np.isfinite(p.sol).all()) for p in ens.problems[::37])
    batched = pts * steps / tb
    sequential = pts * seq_steps / ts
    return {"workload": "SingleLayerQG beta-plane ensemble 256x64^2 + 192x128^2 + 64x256^2, beta=10 mu=0.01 "
                        "FilteredRK4 dt=0.01, Gaussian-bump topography, q0 at k_peak=8 with energy 0.08, FP64",
            "members": sum(c for _, c in ENSEMBLE), "grid_points": pts,
            "batched": {"value": batched, "unit": "grid-pt-steps/s", "steps": steps, "ms_per_step": tb / steps * 1e3},
            "batched_persistent": {"value": pts * steps / tp, "unit": "grid-pt-steps/s", "steps": steps,
                                   "ms_per_step": tp / steps * 1e3,
                                   "form": "each stage as one cooperative launch (gf_batch_persistent)"},
            "sequential": {"value": sequential, "unit": "grid-pt-steps/s", "steps": seq_steps,
                           "ms_per_step": ts / seq_steps * 1e3},
            "batched_over_sequential": batched / sequential, "kernels": kern, "state_finite": finite}


def mlqg_run(gf, steps):
    """MultiLayerQG, 3 layers, 512^2 FP64 ETDRK4 dt = 0.005: H = [0.2, 0.3, 0.5], densities
    [4.0, 5.0, 5.375] with g = 1 mapped to buoyancy through the reduced gravities
    g'_{j+1/2} = g (rho_{j+1} - rho_j) / rho_{j+1} (b = [0, -0.2, -0.2698]; only b_j - b_{j+1}
    enters, src/multilayerqg.jl:297), f0 = 1, beta = 5, U = [1.0, 0.2, 0.0], mu = 0.005,
    q = 1e-2 N(0, 1) per layer (seed 4)."""
    rho = np.array([4.0, 5.0, 5.375])
    gprime = (rho[1:] - rho[:-1]) / rho[1:]
    b = np.concatenate([[0.0], -np.cumsum(gprime)])
    n = 512
    prob = gf.MultiLayerQG.Problem(3, gf.GPU(), nx=n, Lx=2 * np.pi, f0=1.0, β=5.0, U=[1.0, 0.2, 0.0],
                                   H=[0.2, 0.3, 0.5], b=b, μ=0.005, dt=0.005, stepper="ETDRK4")
    q0 = 1e-2 * np.random.default_rng(4).standard_normal((3, n, n))
    gf.MultiLayerQG.set_q(prob, q0)
    gf.stepforward(prob, 10)
    prob.synchronize()
    t0 = time.perf_counter()
    gf.stepforward(prob, steps)
    prob.synchronize()
    el = time.perf_counter() - t0
    return {"workload": "MultiLayerQG 3 layers 512^2 FP64 ETDRK4 dt=0.005, H=[0.2,0.3,0.5], rho=[4,5,5.375] "
                        "(b=[0,-0.2,-0.2698]), f0=1 beta=5 U=[1,0.2,0] mu=0.005",
            "value": n * n * 3 * steps / el, "unit": "grid-pt-steps/s (nx ny nlayers steps / s)", "steps": steps,
            "ms_per_step": el / steps * 1e3, "state_finite": bool(np.isfinite(prob.sol).all())}


def decaying_run(gf, steps=200):
    """TwoDNavierStokes decaying turbulence, 128^2 FP64, nu = 0, FilteredRK4 dt = 0.005, 200 steps,
    zeta0 from peakedisotropicspectrum (k_peak = 6, energy 1, seed 1)."""
    n = 128
    prob = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, Lx=2 * np.pi, dt=0.005, stepper="FilteredRK4")
    grid = gf.problem.Grid(n, n, 2 * np.pi, 2 * np.pi, 1 / 3, np.float64)
    gf.TwoDNavierStokes.set_ζ(prob, gf.peakedisotropicspectrum(grid, 6, 1.0, rng=np.random.default_rng(1)))
    prob.synchronize()
    t0 = time.perf_counter()
    gf.stepforward(prob, steps)
    prob.synchronize()
    el = time.perf_counter() - t0
    return {"workload": "TwoDNavierStokes decaying 128^2 FP64 nu=0 FilteredRK4 dt=0.005, 200 steps",
            "value": n * n * steps / el, "unit": "grid-pt-steps/s", "steps": steps, "ms_per_step": el / steps * 1e3,
            "state_finite": bool(np.isfinite(prob.sol).all())}


def adjoint_run(gf, nsteps=2000, checkpoint=50):
    """Adjoint TwoDNavierStokes 512^2 FP64 ETDRK4 dt = 0.005, nu = 1e-3 (n_nu = 1): gradient of
    J = 1/2 |zeta(T) - target|^2 after 2000 steps w.r.t. zeta0 and nu, against 2000 forward steps.
    zeta0 is a peakedisotropicspectrum field (k_peak = 6, energy 0.5, seed 5); the target is zeta(T)
    of a seeded run from another such field (seed 6, energy 0.3: at 0.5 some seeds exceed the
    advective limit of dt = 0.005 within 2000 steps)."""
    n = 512
    grid = gf.problem.Grid(n, n, 2 * np.pi, 2 * np.pi, 1 / 3, np.float64)
    z0 = gf.peakedisotropicspectrum(grid, 6, 0.5, rng=np.random.default_rng(5))
    prob = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, nν=1, dt=0.005, stepper="ETDRK4")
    gf.TwoDNavierStokes.set_ζ(prob, gf.peakedisotropicspectrum(grid, 6, 0.3, rng=np.random.default_rng(6)))
    gf.stepforward(prob, nsteps)
    gf.TwoDNavierStokes.updatevars(prob)
    target = prob.vars.ζ.copy()
    gf.TwoDNavierStokes.set_ζ(prob, z0)
    gf.stepforward(prob, 20)
    prob.synchronize()
    gf.TwoDNavierStokes.set_ζ(prob, z0)
    t0 = time.perf_counter()
    gf.stepforward(prob, nsteps)
    prob.synchronize()
    tf = time.perf_counter() - t0
    gf.TwoDNavierStokes.set_ζ(prob, z0)
    # untimed: allocates the 2000-step tape (5 retained-column fields per step, 13.9 GB) and the tables
    gf.TwoDNavierStokes.adjoint_gradient(prob, target, nsteps, nsteps, nu=False)
    out = {"workload": "adjoint TwoDNavierStokes 512^2 FP64 ETDRK4 dt=0.005 nu=1e-3 nnu=1, 2000 steps, "
                       "gradient of 1/2|zeta(T)-target|^2 w.r.t. zeta0 and nu",
           "forward_s": tf}
    for name, ck in (("checkpoint_50", checkpoint), ("full_tape", nsteps)):
        t0 = time.perf_counter()
        J, grad, dnu = gf.TwoDNavierStokes.adjoint_gradient(prob, target, nsteps, ck)
        ta = time.perf_counter() - t0
        out[name] = {"gradient_s": ta, "adjoint_over_forward": ta / tf, "J": J, "dJ_dnu": dnu,
                     "grad_finite": bool(np.isfinite(grad).all())}
    return out


def adjoint_batch_run(gf, nb=16, nsteps=2000, checkpoint=50, fd_dirs=32, graphs=False, budget=15e9):
    """Config 5 as a batch: 16 independent 512^2 problems (own contexts and streams), their
    2000-step gradients computed concurrently from 16 host threads, against the same 16 forward
    runs done concurrently.  Each target is zeta(T) of a seeded forward run from another initial
    condition (energy 0.3).  Two tape settings: the minimum (every segment but the last recomputed) and a
    budget of `budget` bytes per problem, which keeps the last segments taped from the forward
    pass (16 x 15 GB of the 288 GB HBM).  Then central differences of one gradient on fd_dirs
    random directions at h = 1e-6, and an h sweep on 4 of them."""
    from concurrent.futures import ThreadPoolExecutor
    n = 512
    grid = gf.problem.Grid(n, n, 2 * np.pi, 2 * np.pi, 1 / 3, np.float64)
    probs, z0s, targets = [], [], []
    for b in range(nb):
        rng = np.random.default_rng(500 + b)
        z0s.append(gf.peakedisotropicspectrum(grid, 6, 0.5, rng=rng))
        p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, nν=1, dt=0.005, stepper="ETDRK4")
        # target: zeta(T) of a seeded run from another initial condition
        gf.TwoDNavierStokes.set_ζ(p, gf.peakedisotropicspectrum(grid, 6, 0.3, rng=np.random.default_rng(900 + b)))
        gf.stepforward(p, nsteps)
        gf.TwoDNavierStokes.updatevars(p)
        targets.append(p.vars.ζ.copy())
        if graphs:
            p.use_graphs(True)
        probs.append(p)
    pool = ThreadPoolExecutor(max_workers=nb)

    def fwd(b):
        gf.TwoDNavierStokes.set_ζ(probs[b], z0s[b])
        gf.stepforward(probs[b], nsteps)
        probs[b].synchronize()

    def grad(b, tape):
        gf.TwoDNavierStokes.set_ζ(probs[b], z0s[b])
        return gf.TwoDNavierStokes.adjoint_gradient(probs[b], targets[b], nsteps, checkpoint, tape_budget=tape)

    list(pool.map(lambda b: gf.TwoDNavierStokes.adjoint_gradient(probs[b], targets[b], 20, 20), range(nb)))
    t0 = time.perf_counter()
    list(pool.map(fwd, range(nb)))
    tf = time.perf_counter() - t0
    out = {"workload": f"{nb} independent adjoint TwoDNavierStokes 512^2 FP64 ETDRK4 problems, {nsteps} steps, "
                       f"checkpoints every {checkpoint}, gradients run concurrently (one context and stream each); "
                       "targets zeta(T) of seeded runs", "forward_batch_s": tf}
    res = None
    for name, tape in (("minimal_tape", 0), ("tape_budget", int(budget))):
        t0 = time.perf_counter()
        r = list(pool.map(lambda b: grad(b, tape), range(nb)))
        tg = time.perf_counter() - t0
        out[name] = {"tape_bytes_per_problem": tape, "gradient_batch_s": tg, "adjoint_over_forward": tg / tf,
                     "grads_finite": all(bool(np.isfinite(x[1]).all()) for x in r)}
        if res is None:
            res = r
        else:
            out[name]["same_bits_as_minimal_tape"] = all(
                np.array_equal(x[1], y[1]) and x[0] == y[0] and x[2] == y[2] for x, y in zip(res, r))
    pool.shutdown()
    best = min(out["minimal_tape"]["adjoint_over_forward"], out["tape_budget"]["adjoint_over_forward"])
    out["adjoint_over_forward"] = best
    # central differences on one member of the batch
    p, z0, tgt = probs[0], z0s[0], targets[0]
    J, g0, _ = res[0]
    rng = np.random.default_rng(7)

    def fd_err(v, h):
        Js = []
        for sgn in (1, -1):
            gf.TwoDNavierStokes.set_ζ(p, z0 + sgn * h * v)
            gf.stepforward(p, nsteps)
            gf.TwoDNavierStokes.updatevars(p)
            Js.append(0.5 * float(np.sum((p.vars.ζ - tgt) ** 2)))
        fd = (Js[0] - Js[1]) / (2 * h)
        return abs(fd - float(np.sum(g0 * v))) / abs(fd)

    dirs = [rng.standard_normal((n, n)) for _ in range(fd_dirs)]
    errs = [fd_err(v, 1e-6) for v in dirs]
    sweep = {f"{h:g}": max(fd_err(v, h) for v in dirs[:4]) for h in (1e-3, 1e-4, 1e-5, 1e-6)}
    out["J"] = J
    out["fd_check"] = {"directions": fd_dirs, "h": 1e-6, "max_rel_error": max(errs),
                       "median_rel_error": float(np.median(errs)), "h_sweep_max_rel_error_4_dirs": sweep}
    out["grads_finite"] = out["minimal_tape"]["grads_finite"] and out["tape_budget"]["grads_finite"]
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(nsteps=20):
    """The oracle (NumPy + SciPy pocketfft with one worker per core) on the same workload: 20
    ETDRK4 steps after one warm-up step.  Workers: the cores this process may run on, at most 16
    (a one-GPU box's share of its host, whatever nproc reports there)."""
    from oracle import fourierflows as off
    from oracle import twodnavierstokes as o2d
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = min(nproc, 16)
    off.set_fft_workers(cores)
    r = o2d.Problem(nx=NX, Lx=CFG["Lx"], nu=CFG["nu"], nnu=CFG["nnu"], mu=CFG["mu"], nmu=CFG["nmu"], dt=CFG["dt"],
                    stepper="ETDRK4", calcF=("ring", CFG["kf"], CFG["dkf"], CFG["eps"], CFG["seed"]))
    off.stepforward(r, 1)
    t0 = time.perf_counter()
    off.stepforward(r, nsteps)
    dt = time.perf_counter() - t0
    return dict(value=NX * NX * nsteps / dt, unit="grid-pt-steps/s", cores=cores, kind="port",
                sample=f"{nsteps} ETDRK4 steps of the 2048^2 FP64 forced workload after 1 warm-up step; "
                       f"scipy.fft rfft2/irfft2 with workers={cores}, NumPy broadcasts single-threaded",
                s_per_step=dt / nsteps, nproc=nproc, os_cpu_count=os.cpu_count(), cpu_model=_cpu_model())


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size (the runtime's copy kernel), run under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to calibrate those counters against bytes moved."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)       # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    dt = (time.perf_counter() - t0) / reps
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps, "read_plus_write_GBps": 2 * nbytes / dt / 1e9}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint-batch", action="store_true",
                    help="with --only-adjoint: the single-problem gradient only (no 16-thread leg)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the Float32 variant")
    ap.add_argument("--graphs", action="store_true",
                    help="with --only-adjoint: graph replay on for the 16 concurrent problems (each thread captures)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--no-stable", action="store_true", help="skip the dt = 0.001 run")
    ap.add_argument("--copy-probe", action="store_true", help="only run the 1 GiB copy used to calibrate counters")
    ap.add_argument("--only-ensemble", action="store_true", help="only run the batched-ensemble measurement")
    ap.add_argument("--only-adjoint", action="store_true", help="only run the adjoint measurement")
    ap.add_argument("--only-mlqg", action="store_true", help="only run the MultiLayerQG 512^2 x 3 measurement")
    ap.add_argument("--only-decaying", action="store_true", help="only run the decaying 128^2 measurement")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.only_adjoint:
        import __graft_entry__ as ge
        gf = ge.load_package()
        out = {"adjoint_512": adjoint_run(gf)}
        if not args.no_adjoint_batch:
            out["adjoint_512_batch16"] = adjoint_batch_run(gf, graphs=args.graphs)
        print(json.dumps(out), flush=True)
        return
    if args.only_mlqg or args.only_decaying:
        import __graft_entry__ as ge
        gf = ge.load_package()
        out = {"mlqg_512x3": mlqg_run(gf, 200)} if args.only_mlqg else {"decaying_128": decaying_run(gf, 200)}
        print(json.dumps(out), flush=True)
        return
    if args.only_ensemble:
        import __graft_entry__ as ge
        print(json.dumps({"ensemble": ensemble_run(ge.load_package(), max(1, args.ensemble_steps), ENSEMBLE_SEQ_STEPS)}),
              flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the decomposition over several GPUs is designed but not built.", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as ge
    gf = ge.load_package()
    steps, warmup = max(1, args.steps), max(0, args.warmup)
    elapsed, kern, finite = gpu_run(gf, steps, warmup, np.float64)
    sched = kern.pop("schedule")
    splits = kern.pop("splits")
    ms_per_step = elapsed / steps * 1e3
    value = NX * NX * steps / elapsed
    dom = max(kern, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
    dk = kern[dom]
    traffic = measured_traffic(tuple(sched), tuple(splits))
    for k in kern:
        kern[k]["hbm_traffic_bytes"] = traffic.get(k)
    S = 16.0 * (NX // 2 + 1) * NX
    step_bytes = 64 * S
    out = {
        "metric": "grid-pt-steps/s 2048^2 FP64 ETDRK4",
        "value": value,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (zeta0 = 0, counter-based ring forcing, seed 2)",
        "config": {"workload": "TwoDNavierStokes 2048x2048 L=2pi ETDRK4 dt=0.005 nu=1e-21 nnu=4 mu=0.1 "
                               "ring forcing kf=14 dkf=1.5 eps=0.1", "nx": NX, "ny": NX, "stepper": "ETDRK4",
                   "parallelism": "single GPU", "fft_points_per_thread_k1_k2_k3": sched,
                   "workgroup_splits_k1_k2_k3": splits,
                   "schedule_source": "geophysicalflows.jl_amd/schedules.json (autotuned on MI355X)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": dk["achieved_GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dk["achieved_GBps"] / HBM_PEAK_GBS,
                     "bytes_per_launch": dk["bytes_per_launch"], "avg_us": dk["avg_us"],
                     "timed": f"{KERNEL_REPS} back-to-back launches after the timed steps and {KERNEL_WARM} "
                              "untimed ones (gf_time_kernel)",
                     "in_step_event_us": dk["in_step_event_us"], "traffic": traffic.get(dom),
                     "traffic_source": "profiles/twodns_2048_fp64_etdrk4_hbm_traffic.md" if traffic else None},
        "kernels": kern,
        "step_model": {"algorithmic_bytes_per_step": step_bytes, "achieved_GBps": step_bytes / (ms_per_step * 1e-3) / 1e9,
                       "frac_of_peak": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "state_finite": finite,
        "clock_ramp_steps": CLOCK_RAMP_STEPS,
    }
    if not args.no_stable:
        # at dt = 0.005 the forced 2048^2 flow exceeds the advective CFL limit within a few hundred
        # steps; the same step count at dt = 0.001 stays finite throughout (its time per step is
        # within a few percent of the main figure)
        es, _, fs = gpu_run(gf, steps, warmup, np.float64, dt=0.001)
        out["stable_dt_0.001"] = {"value": NX * NX * steps / es, "ms_per_step": es / steps * 1e3, "state_finite": fs}
    if not args.no_ensemble:
        out["ensemble"] = ensemble_run(gf, max(1, args.ensemble_steps), ENSEMBLE_SEQ_STEPS)
        out["mlqg_512x3"] = mlqg_run(gf, 200)
        out["decaying_128"] = decaying_run(gf)
        out["adjoint_512"] = adjoint_run(gf)
        out["adjoint_512_batch16"] = adjoint_batch_run(gf)
    if not args.no_fp32:
        e32, k32, f32 = gpu_run(gf, steps, warmup, np.float32, dt=0.001)
        sched32 = k32.pop("schedule")
        splits32 = k32.pop("splits")
        s32 = 8.0 * (NX // 2 + 1) * NX
        out["fp32"] = {"workload": "the same forced 2048^2 problem in Float32 (dt = 0.001, which stays finite)",
                       "value": NX * NX * steps / e32, "ms_per_step": e32 / steps * 1e3, "state_finite": f32,
                       "fft_points_per_thread_k1_k2_k3": sched32, "workgroup_splits_k1_k2_k3": splits32,
                       "step_model_frac_of_peak": 64 * s32 / (e32 / steps) / 1e9 / HBM_PEAK_GBS,
                       "kernels": {k: {"avg_us": v["avg_us"], "achieved_GBps": v["achieved_GBps"]} for k, v in k32.items()}}
    if not args.no_cpu:
        cb = cpu_baseline()
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
