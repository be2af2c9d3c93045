"""Write tests/golden/oracle_fixtures_32.npz: seeded 32^2 states with the oracle's calcN, one
step of the configured stepper and the ETDRK4 coefficient tables.

These are ORACLE outputs (the CPU restatement under oracle/), not reference outputs: the
reference ships no data files and cannot run here.  They pin the oracle against drift
(tests/test_oracle_fixtures.py recomputes them) and give the device a check that needs no oracle
run (tests/test_gpu_fixtures.py).

    python tests/golden/make_oracle_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fourierflows as off          # noqa: E402
from oracle import multilayerqg as oml          # noqa: E402
from oracle import singlelayerqg as osl         # noqa: E402
from oracle import twodnavierstokes as o2d      # noqa: E402

N = 32


def _state(rng, shape, scale):
    s = scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    s[..., 0, 0] = 0
    return s


def cases():
    """(name, make_problem, calcN module, state shape, scale, steps)."""
    x = -np.pi + 2 * np.pi * np.arange(N) / N
    X, Y = np.meshgrid(x, x)
    eta = 0.2 * np.cos(3 * X) * np.sin(2 * Y)
    return [
        ("twodns_etdrk4", lambda: o2d.Problem(nx=N, nu=1e-3, nnu=2, mu=0.05, nmu=0, dt=0.01, stepper="ETDRK4"),
         o2d, (N, N // 2 + 1), 0.5, 3),
        ("twodns_filteredrk4", lambda: o2d.Problem(nx=N, nu=1e-4, nnu=1, dt=0.01, stepper="FilteredRK4"),
         o2d, (N, N // 2 + 1), 0.5, 3),
        ("slqg_topo_etdrk4", lambda: osl.Problem(nx=N, beta=5.0, U=0.2, eta=eta, nu=1e-3, nnu=1, mu=0.02, dt=0.01,
                                                 stepper="ETDRK4"), osl, (N, N // 2 + 1), 0.3, 3),
        ("mlqg3_etdrk4", lambda: oml.Problem(3, nx=N, f0=1.0, beta=5.0, U=[0.5, 0.1, 0.0], H=[0.2, 0.3, 0.5],
                                             b=[0.0, -0.2, -0.27], mu=0.01, nu=1e-3, dt=0.005, stepper="ETDRK4"),
         oml, (3, N, N // 2 + 1), 0.02, 3),
    ]


def compute():
    out = {}
    for k, (name, mk, mod, shape, scale, steps) in enumerate(cases()):
        rng = np.random.default_rng(1000 + k)
        s = _state(rng, shape, scale)
        r = mk()
        Nr, sr = np.zeros_like(s), s.copy()
        mod.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
        r.sol[...] = s
        off.stepforward(r, steps)
        out[f"{name}/state"] = s
        out[f"{name}/calcN"] = Nr
        out[f"{name}/calcN_dealiased_state"] = sr
        out[f"{name}/after_{steps}_steps"] = r.sol.copy()
    # ETDRK4 coefficient tables for a real L spanning the contour's regimes
    L = -np.logspace(-3, 3, 25).reshape(5, 5)
    for name, table in zip(("zeta", "alpha", "beta", "Gamma"), off.etd_coefficients(0.01, L)):
        out[f"etd/{name}"] = np.asarray(table)
    out["etd/L"] = L
    return out


if __name__ == "__main__":
    path = os.path.join(HERE, "oracle_fixtures_32.npz")
    np.savez_compressed(path, **compute())
    print("wrote", path, os.path.getsize(path), "bytes")
