"""Benchmark: TwoDNavierStokes 2048^2 FP64 ETDRK4, stochastic ring forcing.

Workload (the headline configuration): nx = ny = 2048, L = 2 pi, Float64,
ETDRK4, dt = 0.005, hyperviscosity nu = 1e-21 (n_nu = 4), linear drag mu = 0.1
(n_mu = 0), white-in-time forcing on the wavenumber ring k_f = 14 (width 1.5,
injection rate 0.1), zeta0 = 0.  One step is one ETDRK4 step (four calcN!
evaluations and the stage arithmetic).  All state is resident in HBM.

Output: one JSON line with grid-point-steps/s (``value``), the roofline of
the dominant kernel (per-launch durations from HIP events on the library's
stream, algorithmic bytes from gf_kernel_bytes), the whole-step rate against
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


def gpu_run(gf, steps, warmup, T):
    lib = gf._lib.lib
    prob = gf.TwoDNavierStokes.Problem(
        gf.GPU(), nx=CFG["nx"], Lx=CFG["Lx"], ν=CFG["nu"], nν=CFG["nnu"], μ=CFG["mu"], nμ=CFG["nmu"],
        dt=CFG["dt"], stepper="ETDRK4", T=T,
        calcF=gf.RingForcing(CFG["kf"], CFG["dkf"], CFG["eps"], seed=CFG["seed"]))
    gf.TwoDNavierStokes.set_ζ(prob, np.zeros((NX, NX)))
    if warmup:
        gf.stepforward(prob, warmup)
    prob.synchronize()
    t0 = time.perf_counter()
    gf.stepforward(prob, steps)
    prob.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # second pass with per-launch HIP events on the library's stream
    lib.gf_timing(prob._ctx, 1)
    gf.stepforward(prob, steps)
    prob.synchronize()
    kern = {}
    for k, name in enumerate(KERNELS):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        lib.gf_timing_read(prob._ctx, k, ctypes.byref(n), ctypes.byref(ms))
        b0, b1 = ctypes.c_double(), ctypes.c_double()
        lib.gf_kernel_bytes(prob._ctx, k, 0, ctypes.byref(b0))
        lib.gf_kernel_bytes(prob._ctx, k, 1, ctypes.byref(b1))
        per_launch = ((steps - 1) * b0.value + b1.value) / steps
        avg_s = ms.value * 1e-3 / max(n.value, 1)
        kern[name] = dict(launches=n.value, avg_us=avg_s * 1e6, bytes_per_launch=per_launch,
                          achieved_GBps=per_launch / avg_s / 1e9)
    lib.gf_timing(prob._ctx, 0)
    finite = bool(np.isfinite(prob.sol).all())
    return elapsed, kern, finite


def cpu_baseline(nsteps=3):
    """The oracle (NumPy + SciPy pocketfft with one worker per core) on the same workload."""
    from oracle import fourierflows as off
    from oracle import twodnavierstokes as o2d
    cores = min(os.cpu_count() or 1, 16)
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
                s_per_step=dt / nsteps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--fp32", action="store_true", help="also time the Float32 variant")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU decomposition is described in DESIGN.md and not built.", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as ge
    gf = ge.load_package()
    steps, warmup = max(1, args.steps), max(0, args.warmup)
    elapsed, kern, finite = gpu_run(gf, steps, warmup, np.float64)
    ms_per_step = elapsed / steps * 1e3
    value = NX * NX * steps / elapsed
    dom = max(kern, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
    dk = kern[dom]
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
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": dk["achieved_GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dk["achieved_GBps"] / HBM_PEAK_GBS,
                     "bytes_per_launch": dk["bytes_per_launch"], "avg_us": dk["avg_us"], "traffic": None},
        "kernels": kern,
        "step_model": {"algorithmic_bytes_per_step": step_bytes, "achieved_GBps": step_bytes / (ms_per_step * 1e-3) / 1e9,
                       "frac_of_peak": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "state_finite": finite,
    }
    if args.fp32:
        e32, k32, f32 = gpu_run(gf, steps, warmup, np.float32)
        out["fp32"] = {"value": NX * NX * steps / e32, "ms_per_step": e32 / steps * 1e3, "state_finite": f32}
    if not args.no_cpu:
        cb = cpu_baseline()
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
