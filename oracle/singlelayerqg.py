"""Oracle restatement of ``GeophysicalFlows.SingleLayerQG`` (``src/singlelayerqg.jl``).

* ``Params``: 137-207 -- ``etah = rfft(eta)``, ``Qx = irfft(i kr etah)``,
  ``Qy = irfft(i l etah) - U_yy`` (varying-U only).
* ``Equation``: 230-252 -- complex ``L = -mu - nu Krsq^nnu + i beta kr/(Krsq [+ 1/ell^2])
  [- i U kr]``, ``L[0,0] = 0``.
* ``calcN_advection`` constant U: 356-382; varying U(y): 395-432.
* ``calcN``: 443-451; ``addforcing``: 459-477; ``streamfunctionfrompv``: 484-490;
  ``updatevars``: 502-518; ``set_q``: 525-534.
"""
import numpy as np

from . import forcing as _forcing
from .fourierflows import Equation as _Equation, Problem as _Problem, TwoDGrid, dealias, parsevalsum


class Params:
    pass


class Vars:
    def __init__(self, grid, forced=False, stochastic=False):
        z = lambda: np.zeros((grid.ny, grid.nx))
        zh = lambda: np.zeros((grid.nl, grid.nkr), complex)
        self.q, self.psi, self.u, self.v = z(), z(), z(), z()
        self.qh, self.psih, self.uh, self.vh = zh(), zh(), zh(), zh()
        self.Fh = zh() if forced else None
        self.prevsol = zh() if stochastic else None


def make_params(grid, beta, deformation_radius, U, eta, mu, nu, nnu, calcF):
    p = Params()
    p.beta, p.mu, p.nu, p.nnu, p.calcF = float(beta), float(mu), float(nu), int(nnu), calcF
    p.deformation_radius = None if (deformation_radius is None or deformation_radius == np.inf) \
        else float(deformation_radius)
    if eta is None:
        eta = np.zeros((grid.ny, grid.nx))
    elif callable(eta):
        X, Y = grid.gridpoints()
        eta = eta(X, Y)
    p.eta = np.asarray(eta, dtype=float)
    p.etah = grid.rfft(p.eta)
    p.Qx = grid.irfft(1j * grid.Kr * p.etah)
    p.Qy = grid.irfft(1j * grid.Lw * p.etah)
    if np.ndim(U) > 0:  # U(y): broadcast over x (src/singlelayerqg.jl:176-179)
        U = np.asarray(U, dtype=float).reshape(grid.ny)
        U2 = np.broadcast_to(U[:, None], (grid.ny, grid.nx)).copy()
        Uh = grid.rfft(U2)
        Uyy = grid.irfft(-grid.Lw ** 2 * Uh)
        p.Qy = p.Qy - Uyy
        p.U = U2
        p.varyingU = True
    else:
        p.U = float(U)
        p.varyingU = False
    return p


def linear_operator(p, grid):
    if p.deformation_radius is None:
        L = -p.mu - p.nu * grid.Krsq ** p.nnu + 1j * p.beta * grid.Kr * grid.invKrsq
    else:
        L = -p.mu - p.nu * grid.Krsq ** p.nnu + \
            1j * p.beta * grid.Kr / (grid.Krsq + 1 / p.deformation_radius ** 2)
    if not p.varyingU:
        L = L - 1j * p.U * grid.Kr
    L = L.astype(complex)
    L[0, 0] = 0
    return L


def streamfunctionfrompv(psih, qh, p, grid):
    if p.deformation_radius is None:
        psih[...] = -grid.invKrsq * qh
    else:
        psih[...] = -qh / (grid.Krsq + 1 / p.deformation_radius ** 2)


def calcN_advection(N, sol, t, clock, vars, p, grid):
    if not p.varyingU:  # src/singlelayerqg.jl:356-382
        vars.qh[...] = sol
        streamfunctionfrompv(vars.psih, vars.qh, p, grid)
        vars.uh[...] = -1j * grid.Lw * vars.psih
        vars.vh[...] = 1j * grid.Kr * vars.psih
        vars.q[...] = grid.irfft(vars.qh)
        vars.u[...] = grid.irfft(vars.uh)
        vars.v[...] = grid.irfft(vars.vh)
        qe = vars.q + p.eta
        uqh = grid.rfft(vars.u * qe)
        vqh = grid.rfft(vars.v * qe)
        N[...] = -1j * grid.Kr * uqh - 1j * grid.Lw * vqh
        N -= 1j * grid.Kr * p.U * p.etah
    else:  # src/singlelayerqg.jl:395-432
        vars.qh[...] = sol
        streamfunctionfrompv(vars.psih, vars.qh, p, grid)
        vars.uh[...] = -1j * grid.Lw * vars.psih
        vars.vh[...] = 1j * grid.Kr * vars.psih
        vars.u[...] = grid.irfft(vars.uh)
        vars.u += p.U
        N[...] = -grid.rfft(vars.u * p.Qx)
        vars.v[...] = grid.irfft(vars.vh)
        N -= grid.rfft(vars.v * p.Qy)
        vars.q[...] = grid.irfft(vars.qh)
        uqh = grid.rfft(vars.u * vars.q)
        vqh = grid.rfft(vars.v * vars.q)
        N -= 1j * grid.Kr * uqh + 1j * grid.Lw * vqh


def addforcing(N, sol, t, clock, vars, p, grid):
    F = p.calcF
    if F is None:
        return None
    if isinstance(F, tuple) and F[0] == "ring":
        if t == clock.t:
            vars.prevsol[...] = sol
            vars.Fh[...] = _forcing.ring_forcing(grid, p.sqrt_spec, F[4], clock.step, clock.dt,
                                                 getattr(p, "instance", 0))
    elif callable(F):
        F(vars.Fh, sol, t, clock, vars, p, grid)
    else:
        vars.Fh[...] = F
    N += vars.Fh
    return None


def calcN(N, sol, t, clock, vars, p, grid):
    dealias(sol, grid)
    calcN_advection(N, sol, t, clock, vars, p, grid)
    addforcing(N, sol, t, clock, vars, p, grid)


def Problem(nx=256, ny=None, Lx=2 * np.pi, Ly=None, beta=0.0, deformation_radius=np.inf, U=0.0,
            eta=None, nu=0.0, nnu=1, mu=0.0, dt=0.01, stepper="RK4", calcF=None,
            stochastic=False, aliased_fraction=1 / 3):
    grid = TwoDGrid(nx, Lx, ny, Ly, aliased_fraction)
    p = make_params(grid, beta, deformation_radius, U, eta, mu, nu, nnu, calcF)
    if isinstance(calcF, tuple) and calcF[0] == "ring":
        _, kf, dkf, eps, _seed = calcF
        p.sqrt_spec = np.sqrt(_forcing.ring_spectrum(grid, kf, dkf, eps))
        stochastic = True
    vars = Vars(grid, forced=calcF is not None, stochastic=stochastic)
    eq = _Equation(linear_operator(p, grid), calcN, grid)
    return _Problem(eq, stepper, dt, grid, vars, p)


def updatevars(prob):
    sol, vars, p, grid = prob.sol, prob.vars, prob.params, prob.grid
    dealias(sol, grid)
    vars.qh[...] = sol
    streamfunctionfrompv(vars.psih, vars.qh, p, grid)
    vars.uh[...] = -1j * grid.Lw * vars.psih
    vars.vh[...] = 1j * grid.Kr * vars.psih
    vars.q[...] = grid.irfft(vars.qh.copy())
    vars.psi[...] = grid.irfft(vars.psih.copy())
    vars.u[...] = grid.irfft(vars.uh.copy())
    vars.v[...] = grid.irfft(vars.vh.copy())


def set_q(prob, q):
    prob.vars.qh[...] = prob.grid.rfft(q)
    prob.sol[...] = prob.vars.qh
    updatevars(prob)


# diagnostics (src/singlelayerqg.jl:546-715)
def kinetic_energy(prob):
    sol, v, p, g = prob.sol, prob.vars, prob.params, prob.grid
    streamfunctionfrompv(v.psih, sol, p, g)
    return 1 / (2 * g.Lx * g.Ly) * parsevalsum(g.Krsq * np.abs(v.psih) ** 2, g)


def potential_energy(prob):
    sol, v, p, g = prob.sol, prob.vars, prob.params, prob.grid
    if p.deformation_radius is None:
        return 0.0
    streamfunctionfrompv(v.psih, sol, p, g)
    return 1 / (2 * g.Lx * g.Ly) * parsevalsum(np.abs(v.psih) ** 2, g) / p.deformation_radius ** 2


def energy(prob):
    return kinetic_energy(prob) + potential_energy(prob)


def enstrophy(prob):
    sol, v, p, g = prob.sol, prob.vars, prob.params, prob.grid
    return 1 / (2 * g.Lx * g.Ly) * parsevalsum(np.abs(sol + p.etah) ** 2, g)
