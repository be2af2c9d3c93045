"""SingleLayerQG on the device (mirror of ``src/singlelayerqg.jl``).

``Problem`` takes the reference keywords (``src/singlelayerqg.jl:80-102``): complex
``L = -mu - nu |k|^(2 nnu) + i beta kr/(|k|^2 [+ 1/ell^2]) [- i kr U]`` (230-252),
topographic PV ``eta`` (array of shape ``(ny, nx)`` or a function of ``(x, y)``),
constant ``U`` or ``U(y)`` (array of ``ny`` values, the varying-U path 395-432).
"""
import math

import numpy as np

from . import _lib
from .problem import GPU, Grid, Problem as _Problem, attach_forcing, device_diagnostic, dptr, make_desc, stepforward  # noqa: F401


class Params:
    pass


class SLQGProblem(_Problem):
    _names = ("q", "ψ", "u", "v")


def Problem(dev=None, *, nx=256, ny=None, Lx=2 * math.pi, Ly=None, β=0.0, deformation_radius=math.inf,
            U=0.0, eta=None, ν=0.0, nν=1, μ=0.0, dt=0.01, stepper="RK4", calcF=None, stochastic=False,
            aliased_fraction=1 / 3, T=np.float64):
    ny = nx if ny is None else ny
    Ly = Lx if Ly is None else Ly
    grid = Grid(int(nx), int(ny), Lx, Ly, aliased_fraction, T)
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


# Diagnostics (src/singlelayerqg.jl:546-715), computed on the device from the raw state.
@device_diagnostic(lambda p: (_lib.GF_DIAG_KINETIC_ENERGY, 0.0, 0, 1))
def kinetic_energy(prob):
    return prob._diagnostic(_lib.GF_DIAG_KINETIC_ENERGY)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_POTENTIAL_ENERGY, 0.0, 0, 1))
def potential_energy(prob):
    return prob._diagnostic(_lib.GF_DIAG_POTENTIAL_ENERGY)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENERGY, 0.0, 0, 1))
def energy(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENERGY)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENSTROPHY, 0.0, 0, 1))
def enstrophy(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENSTROPHY)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENERGY_DISSIPATION, p.params.ν, p.params.nν, 1))
def energy_dissipation(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENERGY_DISSIPATION, prob.params.ν, prob.params.nν)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENSTROPHY_DISSIPATION, p.params.ν, p.params.nν, 1))
def enstrophy_dissipation(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENSTROPHY_DISSIPATION, prob.params.ν, prob.params.nν)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENERGY_DRAG, p.params.μ, 0, 1))
def energy_drag(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENERGY_DRAG, prob.params.μ, 0)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENSTROPHY_DRAG, p.params.μ, 0, 1))
def enstrophy_drag(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENSTROPHY_DRAG, prob.params.μ, 0)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENERGY_WORK, 0.0, 0, 1))
def energy_work(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENERGY_WORK)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENSTROPHY_WORK, 0.0, 0, 1))
def enstrophy_work(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENSTROPHY_WORK)[0]
