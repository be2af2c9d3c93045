"""Oracle restatement of ``GeophysicalFlows.TwoDNavierStokes`` (``src/twodnavierstokes.jl``).

* ``Equation``: ``L = -nu Krsq^nnu - mu Krsq^nmu``, ``L[0,0] = 0`` (149-154).
* ``calcN_advection``: 253-275; ``calcN``: 286-294 (dealias in place, P1);
  ``addforcing``: 302-319 (stochastic forcing only when ``t == clock.t``).
* ``updatevars``: 331-345; ``set_zeta``: 352-360; diagnostics 374-535.

Forcing kinds supported by the restatement (the reference accepts any Julia
closure ``calcF!``): ``None``, a fixed spectral array ``Fh`` (what the
reference's deterministic tests' closures do, ``test/test_twodnavierstokes.jl:239-243``),
or a ``("ring", kf, dkf, eps, seed[, instance])`` tuple (the stochastic ring of
``examples/twodnavierstokes_stochasticforcing.jl:53-82`` with the counter-based
generator of ``oracle.forcing``).
"""
import numpy as np

from . import forcing as _forcing
from .fourierflows import Equation as _Equation, Problem as _Problem, TwoDGrid, dealias, parsevalsum


class Params:
    def __init__(self, nu, nnu, mu=0.0, nmu=0, calcF=None):
        self.nu, self.nnu, self.mu, self.nmu, self.calcF = nu, nnu, mu, nmu, calcF


class Vars:
    def __init__(self, grid, forced=False, stochastic=False):
        self.zeta = np.zeros((grid.ny, grid.nx))
        self.u = np.zeros((grid.ny, grid.nx))
        self.v = np.zeros((grid.ny, grid.nx))
        self.zetah = np.zeros((grid.nl, grid.nkr), complex)
        self.uh = np.zeros((grid.nl, grid.nkr), complex)
        self.vh = np.zeros((grid.nl, grid.nkr), complex)
        self.Fh = np.zeros((grid.nl, grid.nkr), complex) if forced else None
        self.prevsol = np.zeros((grid.nl, grid.nkr), complex) if stochastic else None


def linear_operator(params, grid):
    """``L = -nu Krsq^nnu - mu Krsq^nmu``; integer powers (Krsq^0 = 1), ``L[1,1] = 0``."""
    L = -params.nu * grid.Krsq ** params.nnu - params.mu * grid.Krsq ** params.nmu
    L[0, 0] = 0
    return L


def calcN_advection(N, sol, t, clock, vars, params, grid):
    """``N = -i kr (u zeta)^ - i l (v zeta)^`` (src/twodnavierstokes.jl:253-275)."""
    vars.uh[...] = 1j * grid.Lw * grid.invKrsq * sol
    vars.vh[...] = -1j * grid.Kr * grid.invKrsq * sol
    vars.zetah[...] = sol
    vars.u[...] = grid.irfft(vars.uh)
    vars.v[...] = grid.irfft(vars.vh)
    vars.zeta[...] = grid.irfft(vars.zetah)
    uz = vars.u * vars.zeta
    vz = vars.v * vars.zeta
    uzh = grid.rfft(uz)
    vzh = grid.rfft(vz)
    N[...] = -1j * grid.Kr * uzh - 1j * grid.Lw * vzh


def addforcing(N, sol, t, clock, vars, params, grid):
    F = params.calcF
    if F is None:
        return None
    if isinstance(F, tuple) and F[0] == "ring":
        if t == clock.t:  # not a substep (src/twodnavierstokes.jl:311)
            vars.prevsol[...] = sol
            vars.Fh[...] = _forcing.ring_forcing(grid, params.sqrt_spec, F[4], clock.step, clock.dt,
                                                 F[5] if len(F) > 5 else 0)
    elif callable(F):
        F(vars.Fh, sol, t, clock, vars, params, grid)
    else:
        vars.Fh[...] = F
    N += vars.Fh
    return None


def calcN(N, sol, t, clock, vars, params, grid):
    dealias(sol, grid)
    calcN_advection(N, sol, t, clock, vars, params, grid)
    addforcing(N, sol, t, clock, vars, params, grid)


def Problem(nx=256, ny=None, Lx=2 * np.pi, Ly=None, nu=0.0, nnu=1, mu=0.0, nmu=0, dt=0.01,
            stepper="RK4", calcF=None, stochastic=False, aliased_fraction=1 / 3):
    grid = TwoDGrid(nx, Lx, ny, Ly, aliased_fraction)
    params = Params(float(nu), int(nnu), float(mu), int(nmu), calcF)
    ring = isinstance(calcF, tuple) and calcF[0] == "ring"
    if ring:
        kf, dkf, eps = calcF[1:4]
        params.sqrt_spec = np.sqrt(_forcing.ring_spectrum(grid, kf, dkf, eps))
        stochastic = True
    vars = Vars(grid, forced=calcF is not None, stochastic=stochastic)
    eq = _Equation(linear_operator(params, grid), calcN, grid)
    return _Problem(eq, stepper, dt, grid, vars, params)


def updatevars(prob):
    vars, grid, sol = prob.vars, prob.grid, prob.sol
    dealias(sol, grid)
    vars.zetah[...] = sol
    vars.uh[...] = 1j * grid.Lw * grid.invKrsq * sol
    vars.vh[...] = -1j * grid.Kr * grid.invKrsq * sol
    vars.zeta[...] = grid.irfft(vars.zetah.copy())
    vars.u[...] = grid.irfft(vars.uh.copy())
    vars.v[...] = grid.irfft(vars.vh.copy())


def set_zeta(prob, zeta):
    prob.sol[...] = prob.grid.rfft(zeta)
    prob.sol[0, 0] = 0
    updatevars(prob)


# diagnostics (src/twodnavierstokes.jl:374-535); they read the raw (undealiased) sol
def energy(prob):
    g = prob.grid
    return 1 / (g.Lx * g.Ly) * parsevalsum(0.5 * g.invKrsq * np.abs(prob.sol) ** 2, g)


def enstrophy(prob):
    g = prob.grid
    return 1 / (2 * g.Lx * g.Ly) * parsevalsum(np.abs(prob.sol) ** 2, g)


def palinstrophy(prob):
    g = prob.grid
    return 1 / (g.Lx * g.Ly) * parsevalsum(0.5 * g.Krsq * np.abs(prob.sol) ** 2, g)


def energy_dissipation(prob, xi, nxi):
    g = prob.grid
    h = -xi * g.Krsq ** (nxi - 1) * np.abs(prob.sol) ** 2 if nxi >= 1 else \
        -xi * np.where(g.Krsq > 0, g.Krsq, 1.0) ** (nxi - 1) * np.abs(prob.sol) ** 2
    h[0, 0] = 0
    return 1 / (g.Lx * g.Ly) * parsevalsum(h, g)


def enstrophy_dissipation(prob, xi, nxi):
    g = prob.grid
    h = -xi * g.Krsq ** nxi * np.abs(prob.sol) ** 2
    h[0, 0] = 0
    return 1 / (g.Lx * g.Ly) * parsevalsum(h, g)


def energy_dissipation_hyperviscosity(prob):
    return energy_dissipation(prob, prob.params.nu, prob.params.nnu)


def energy_dissipation_hypoviscosity(prob):
    return energy_dissipation(prob, prob.params.mu, prob.params.nmu)


def enstrophy_dissipation_hyperviscosity(prob):
    return enstrophy_dissipation(prob, prob.params.nu, prob.params.nnu)


def enstrophy_dissipation_hypoviscosity(prob):
    return enstrophy_dissipation(prob, prob.params.mu, prob.params.nmu)


def energy_work(prob):
    g, v = prob.grid, prob.vars
    s = (v.prevsol + prob.sol) / 2 if v.prevsol is not None else prob.sol
    return 1 / (g.Lx * g.Ly) * parsevalsum(g.invKrsq * s * np.conj(v.Fh), g)


def enstrophy_work(prob):
    g, v = prob.grid, prob.vars
    s = (v.prevsol + prob.sol) / 2 if v.prevsol is not None else prob.sol
    return 1 / (g.Lx * g.Ly) * parsevalsum(s * np.conj(v.Fh), g)
