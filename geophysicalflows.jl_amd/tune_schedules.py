"""Tune the radix schedules on this GPU and write the table the library reads at start-up.

    python geophysicalflows.jl_amd/tune_schedules.py [--out PATH] [--steps K]

from the repository root, on the GPU the table is for.  For every
workload below it builds the problem, lets ``gf_autotune`` time each offered schedule over K
steps, and records the fastest per kernel; the table goes to ``schedules.json`` beside
libgfhip.so unless ``--out`` says otherwise.  Problems of the same model, size and dtype share
an entry; topography, forcing and the stepper do not change which kernels run.
"""
import argparse
import os
import sys

import numpy as np

WORKLOADS = (
    # (module, nx, dtype, problem keywords)
    [("TwoDNavierStokes", n, np.float64, dict(ν=1e-6, nν=2, μ=0.1, dt=0.002, stepper="ETDRK4"))
     for n in (64, 128, 256, 512, 1024, 2048, 4096)]
    + [("TwoDNavierStokes", n, np.float32, dict(ν=1e-6, nν=2, μ=0.1, dt=0.001, stepper="ETDRK4"))
       for n in (256, 512, 1024, 2048, 4096)]
    + [("SingleLayerQG", n, np.float64, dict(β=10.0, μ=0.01, dt=0.002, stepper="FilteredRK4"))
       for n in (64, 128, 256, 512)]
)


def main(argv=None):
    here = os.path.dirname(os.path.abspath(__file__))
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=os.path.join(here, "schedules.json"))
    ap.add_argument("--steps", type=int, default=30)
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(here))
    import __graft_entry__ as ge
    gf = ge.load_package()
    gf._lib.lib.gf_schedules_clear()             # tune from the built-in choices
    for mod, n, T, kw in WORKLOADS:
        m = getattr(gf, mod)
        p = m.Problem(gf.GPU(), nx=n, T=T, **kw)
        rng = np.random.default_rng(n)
        f = 0.1 * rng.standard_normal((n, n))
        (m.set_ζ if mod == "TwoDNavierStokes" else m.set_q)(p, f)
        gf.stepforward(p, 3)
        builtin = gf.schedule(p) + gf.splits(p)
        chosen, us = gf.autotune(p, steps=args.steps)
        base = us.get(builtin)
        best_plain = min(v for k, v in us.items() if k[3:] == (0, 0, 0))
        print(f"{mod:17s} {n:5d}^2 {np.dtype(T).name:8s} P {chosen[:3]} splits {chosen[3:]}  step {us[chosen]:.2f} us"
              f"  (built-in {base:.2f} us, best without splits {best_plain:.2f} us)", flush=True)
    gf.save_schedules(args.out)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
