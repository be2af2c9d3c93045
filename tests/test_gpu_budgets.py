"""Energy and enstrophy budgets on the device: the reference's budget tests restated
(test/test_twodnavierstokes.jl:25-208, test/test_singlelayerqg.jl:53-257), with the diagnostic
series recorded on the device (Diagnostic / gf_diagnostic_async) and the forcing either
time-harmonic (Fh cos(2t), evaluated at each stage's time as calcF! is) or this library's
stochastic ring forcing (the reference draws unit-modulus random phases with Julia's RNG; the
Stratonovich closure holds for both).  Tolerances are the reference's: 1e-4 (deterministic,
centred difference) and 1e-3 (stochastic, forward difference).

Also the time-harmonic forcing against the oracle (which takes the same forcing as a calcF
callable) at 1e-12.
"""
import math

import numpy as np
import pytest

from oracle import fourierflows as off
from oracle import singlelayerqg as osl
from oracle import twodnavierstokes as o2d
from tests.conftest import rel

pytestmark = pytest.mark.gpu
N, L, DT = 256, 2 * math.pi, 0.005
NT = round(0.1 / 0.1 / DT)            # tf = 0.1/mu with mu = 0.1


def _cos_forcing(n, k0=1.0, l0=1.0):
    g = off.TwoDGrid(n, L)
    x, y = g.gridpoints()
    return g.rfft(0.01 * np.cos(4 * k0 * x) * np.cos(5 * l0 * y))


@pytest.mark.parametrize("model,stepper", [("twodns", "RK4"), ("twodns", "ETDRK4"), ("slqg", "RK4"),
                                           ("slqg", "FilteredRK4")])
def test_harmonic_forcing_matches_oracle(gf, model, stepper):
    n = 64
    fh = _cos_forcing(n)
    rng = np.random.default_rng(100)
    s = 0.05 * (rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1)))
    s[0, 0] = 0

    def calcF(Fh, sol, t, *rest):
        Fh[...] = fh * np.cos(2 * t + 0.3)

    if model == "twodns":
        p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-4, nν=2, μ=0.1, dt=0.01, stepper=stepper,
                                        calcF=gf.HarmonicForcing(fh, 2.0, 0.3))
        r = o2d.Problem(nx=n, nu=1e-4, nnu=2, mu=0.1, dt=0.01, stepper=stepper, calcF=calcF)
        work, owork = gf.TwoDNavierStokes.energy_work, o2d.energy_work
    else:
        p = gf.SingleLayerQG.Problem(gf.GPU(), nx=n, β=3.0, ν=1e-4, nν=2, μ=0.1, dt=0.01, stepper=stepper,
                                     calcF=gf.HarmonicForcing(fh, 2.0, 0.3))
        r = osl.Problem(nx=n, beta=3.0, nu=1e-4, nnu=2, mu=0.1, dt=0.01, stepper=stepper, calcF=calcF)
        work = gf.SingleLayerQG.energy_work

        def owork(r):   # src/singlelayerqg.jl:641-655 (ForcedVars): invKrsq sol conj(Fh)
            g = r.grid
            return 1 / (g.Lx * g.Ly) * off.parsevalsum(g.invKrsq * r.sol * np.conj(r.vars.Fh), g)
    assert work(p) == 0                      # no forcing evaluated yet (vars.Fh starts at zero)
    p.sol = s
    r.sol[...] = s
    N1, _ = p.calcN(s, t=0.7)
    Nr, sr = np.zeros_like(s), s.copy()
    (o2d if model == "twodns" else osl).calcN(Nr, sr, 0.7, r.clock, r.vars, r.params, r.grid)
    assert rel(N1, Nr) < 1e-13
    for nsteps in (3, 4):
        gf.stepforward(p, nsteps)
        off.stepforward(r, nsteps)
        assert rel(p.sol, r.sol) < 1e-12
        assert work(p) == pytest.approx(owork(r), rel=1e-11)


def _budget(gf, model, forcing, what):
    if model == "twodns":
        M = gf.TwoDNavierStokes
        p = M.Problem(gf.GPU(), nx=N, Lx=L, ν=1e-7, nν=2, μ=0.1, nμ=0, dt=DT, stepper="RK4", calcF=forcing)
        M.set_ζ(p, np.zeros((N, N)))
        fns = ((M.energy, M.energy_dissipation_hyperviscosity, M.energy_dissipation_hypoviscosity, M.energy_work)
               if what == "energy" else
               (M.enstrophy, M.enstrophy_dissipation_hyperviscosity, M.enstrophy_dissipation_hypoviscosity,
                M.enstrophy_work))
        sign = 1.0                            # dE/dt = W + D + R (D, R signed)
    else:
        M = gf.SingleLayerQG
        p = M.Problem(gf.GPU(), nx=N, Lx=L, ν=1e-7, nν=2, μ=0.1, dt=DT, stepper="RK4", calcF=forcing)
        M.set_q(p, np.zeros((N, N)))
        fns = ((M.kinetic_energy, M.energy_dissipation, M.energy_drag, M.energy_work) if what == "energy" else
               (M.enstrophy, M.enstrophy_dissipation, M.enstrophy_drag, M.enstrophy_work))
        sign = -1.0                           # dE/dt = W - D - R (D, R positive)
    E, D, R, W = (gf.Diagnostic(f, p, nsteps=NT) for f in fns)
    gf.stepforward(p, [E, D, W, R], NT)
    return tuple(np.asarray(x.data) for x in (E, D, R, W)) + (sign,)


@pytest.mark.parametrize("model", ["twodns", "slqg"])
@pytest.mark.parametrize("what", ["energy", "enstrophy"])
def test_deterministic_forcing_budget(gf, model, what):
    # test/test_twodnavierstokes.jl:128-207, test/test_singlelayerqg.jl:112-257: f = 0.01 cos(4x) cos(5y) cos(2t)
    fh = _cos_forcing(N)
    E, D, R, W, sign = _budget(gf, model, gf.HarmonicForcing(fh, 2.0), what)
    assert np.all(np.isfinite(E)) and E[-1] > 0
    dEdt = (E[2:] - E[:-2]) / (2 * DT)
    computed = W[1:-1] + sign * (D[1:-1] + R[1:-1])
    assert np.linalg.norm(dEdt - computed) / np.linalg.norm(dEdt) < 1e-4


@pytest.mark.parametrize("model,what", [("twodns", "enstrophy"), ("slqg", "energy"), ("slqg", "enstrophy")])
def test_stochastic_forcing_budget(gf, model, what):
    # test/test_twodnavierstokes.jl:77-125, test/test_singlelayerqg.jl:53-218 (kf = 12, dkf = 2, eps = 0.1)
    E, D, R, W, sign = _budget(gf, model, gf.RingForcing(12.0, 2.0, 0.1, seed=1234), what)
    assert np.all(np.isfinite(E)) and E[-1] > 0
    dEdt = (E[1:] - E[:-1]) / DT
    computed = W[1:] + sign * (D[:-1] + R[:-1])
    assert np.linalg.norm(dEdt - computed) / np.linalg.norm(dEdt) < 1e-3
