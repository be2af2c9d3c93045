"""SingleLayerQG on the device (mirror of ``src/singlelayerqg.jl``).

``Problem`` takes the reference keywords (``src/singlelayerqg.jl:80-102``): complex
``L = -mu - nu |k|^(2 nnu) + i beta kr/(|k|^2 [+ 1/ell^2]) [- i kr U]`` (230-252),
topographic PV ``eta`` (array of shape ``(ny, nx)`` or a function of ``(x, y)``),
constant ``U`` or ``U(y)`` (array of ``ny`` values, the varying-U path 395-432).
"""
import math

import numpy as np

from . import _lib
from .problem import GPU, Problem as _Problem, attach_forcing, dptr, make_desc, stepforward  # noqa: F401
from .spectral import parsevalsum, parsevalsum2


class Params:
    pass


class SLQGProblem(_Problem):
    _names = ("q", "ψ", "u", "v")


def Problem(dev=None, *, nx=256, ny=None, Lx=2 * math.pi, Ly=None, β=0.0, deformation_radius=math.inf,
            U=0.0, eta=None, ν=0.0, nν=1, μ=0.0, dt=0.01, stepper="RK4", calcF=None, stochastic=False,
            aliased_fraction=1 / 3, T=np.float64):
    from .grid import make_grid
    ny = nx if ny is None else ny
    Ly = Lx if Ly is None else Ly
    grid = make_grid(nx, ny, Lx, Ly, aliased_fraction, T)
    d = make_desc(_lib.GF_MODEL_SLQG, nx, ny, Lx, Ly, T, stepper, dt, aliased_fraction)
    d.beta, d.mu, d.nu, d.nnu = float(β), float(μ), float(ν), int(nν)
    ell = deformation_radius
    d.deformation_radius = -1.0 if (ell is None or ell == math.inf) else float(ell)
    keep = []
    if eta is not None:
        if callable(eta):
            X, Y = grid.gridpoints()
            eta = eta(X, Y)
        e = np.ascontiguousarray(np.asarray(eta, dtype=np.float64).reshape(ny, nx))
        keep.append(e)
        d.eta = dptr(e)
    if np.ndim(U) > 0:
        u = np.ascontiguousarray(np.asarray(U, dtype=np.float64).reshape(ny))
        keep.append(u)
        d.Uy = dptr(u)
    else:
        d.U = float(U)
    Fh = attach_forcing(d, calcF)
    p = Params()
    p.β, p.deformation_radius, p.U, p.eta, p.μ, p.ν, p.nν, p.calcF, p.stepper = β, ell, U, eta, μ, ν, nν, calcF, stepper
    prob = SLQGProblem(d, grid, p, keep=keep)
    p.etah = None
    if Fh is not None:
        prob.set_forcing(Fh)
    return prob


def set_q(prob, q):
    """``set_q!``: rfft on the device (no zero-mean), then ``updatevars!``."""
    prob._set_physical(q)
    prob._updatevars()


def updatevars(prob):
    prob._updatevars()


def _psih(prob, sol):
    g, p = prob.grid, prob.params
    if p.deformation_radius is None or p.deformation_radius == math.inf:
        iK = np.where(g.Krsq > 0, 1.0 / np.where(g.Krsq > 0, g.Krsq, 1.0), 0.0)
        return -iK * sol
    return -sol / (g.Krsq + 1 / p.deformation_radius ** 2)


def kinetic_energy(prob):
    g = prob.grid
    return parsevalsum2(np.sqrt(g.Krsq) * _psih(prob, prob.sol), g) / (2 * g.Lx * g.Ly)


def potential_energy(prob):
    g, p = prob.grid, prob.params
    if p.deformation_radius is None or p.deformation_radius == math.inf:
        return 0.0
    return 1 / p.deformation_radius ** 2 * parsevalsum2(_psih(prob, prob.sol), g) / (2 * g.Lx * g.Ly)


def energy(prob):
    return kinetic_energy(prob) + potential_energy(prob)
