"""TwoDNavierStokes on the device (mirror of ``src/twodnavierstokes.jl``).

``Problem`` takes the reference's keyword arguments (``src/twodnavierstokes.jl:71-89``);
the linear operator ``L = -nu |k|^(2 nnu) - mu |k|^(2 nmu)`` (149-154), ``calcN!``
(286-294) and the time stepper run in the HIP library.  ``set_ζ``/``updatevars``
mirror ``set_ζ!``/``updatevars!`` (331-360).  The diagnostics (374-535) are
Parseval sums evaluated on the host from ``prob.sol``.
"""
import math

import numpy as np

from . import _lib
from .problem import GPU, Grid, Problem as _Problem, attach_forcing, device_diagnostic, make_desc, stepforward  # noqa: F401


class Params:
    def __init__(self, ν, nν, μ, nμ, calcF, stepper):
        self.ν, self.nν, self.μ, self.nμ, self.calcF, self.stepper = ν, nν, μ, nμ, calcF, stepper


class TwoDNSProblem(_Problem):
    _names = ("ζ", "u", "v")


def Problem(dev=None, *, nx=256, ny=None, Lx=2 * math.pi, Ly=None, ν=0.0, nν=1, μ=0.0, nμ=0,
            dt=0.01, stepper="RK4", calcF=None, stochastic=False, aliased_fraction=1 / 3, T=np.float64):
    ny = nx if ny is None else ny
    Ly = Lx if Ly is None else Ly
    d = make_desc(_lib.GF_MODEL_TWODNS, nx, ny, Lx, Ly, T, stepper, dt, aliased_fraction)
    d.nu, d.nnu, d.mu, d.nmu = float(ν), int(nν), float(μ), int(nμ)
    Fh = attach_forcing(d, calcF)
    p = TwoDNSProblem(d, Grid(int(nx), int(ny), Lx, Ly, aliased_fraction, T), Params(float(ν), int(nν), float(μ), int(nμ), calcF, stepper))
    if Fh is not None:
        p.set_forcing(Fh)
    return p


def set_ζ(prob, ζ):
    """``set_ζ!``: rfft on the device, ``sol[1,1] = 0``, then ``updatevars!``."""
    prob._set_physical(ζ)
    prob._updatevars()


def updatevars(prob):
    prob._updatevars()


# Diagnostics (src/twodnavierstokes.jl:374-535), computed on the device from the raw state.
@device_diagnostic(lambda p: (_lib.GF_DIAG_ENERGY, 0.0, 0, 1))
def energy(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENERGY)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENSTROPHY, 0.0, 0, 1))
def enstrophy(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENSTROPHY)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_PALINSTROPHY, 0.0, 0, 1))
def palinstrophy(prob):
    return prob._diagnostic(_lib.GF_DIAG_PALINSTROPHY)[0]


def energy_dissipation(prob, ξ, nξ):
    return prob._diagnostic(_lib.GF_DIAG_ENERGY_DISSIPATION, ξ, nξ)[0]


def enstrophy_dissipation(prob, ξ, nξ):
    return prob._diagnostic(_lib.GF_DIAG_ENSTROPHY_DISSIPATION, ξ, nξ)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENERGY_DISSIPATION, p.params.ν, p.params.nν, 1))
def energy_dissipation_hyperviscosity(prob):
    return energy_dissipation(prob, prob.params.ν, prob.params.nν)


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENERGY_DISSIPATION, p.params.μ, p.params.nμ, 1))
def energy_dissipation_hypoviscosity(prob):
    return energy_dissipation(prob, prob.params.μ, prob.params.nμ)


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENSTROPHY_DISSIPATION, p.params.ν, p.params.nν, 1))
def enstrophy_dissipation_hyperviscosity(prob):
    return enstrophy_dissipation(prob, prob.params.ν, prob.params.nν)


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENSTROPHY_DISSIPATION, p.params.μ, p.params.nμ, 1))
def enstrophy_dissipation_hypoviscosity(prob):
    return enstrophy_dissipation(prob, prob.params.μ, prob.params.nμ)


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENERGY_WORK, 0.0, 0, 1))
def energy_work(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENERGY_WORK)[0]


@device_diagnostic(lambda p: (_lib.GF_DIAG_ENSTROPHY_WORK, 0.0, 0, 1))
def enstrophy_work(prob):
    return prob._diagnostic(_lib.GF_DIAG_ENSTROPHY_WORK)[0]


def adjoint_gradient(prob, target, nsteps, checkpoint=50, nu=True, tape_budget=None):
    """Loss ``J = 1/2 sum (zeta(T) - target)^2`` after ``nsteps`` steps from the current state and
    its gradients with respect to the initial vorticity (physical, ``(ny, nx)``) and, if ``nu``,
    the viscosity: ``(J, dJ/dzeta0, dJ/dnu)``.  The discrete adjoint of the device ETDRK4 step
    (unforced, FP64); checkpoints every ``checkpoint`` steps (no reference counterpart).
    ``tape_budget`` (bytes): device memory the tape may take, so that the last segments stay
    taped and only the earlier ones are recomputed (``gf_adjoint_tape_budget``)."""
    import ctypes
    from ._lib import check, lib
    if tape_budget is not None:
        check(lib.gf_adjoint_tape_budget(prob._ctx, int(tape_budget)), prob._ctx)
    g = prob.grid
    t = np.ascontiguousarray(np.asarray(target, dtype=np.float64).reshape(g.ny, g.nx))
    grad = np.empty_like(t)
    J, dnu = ctypes.c_double(), ctypes.c_double()
    check(lib.gf_adjoint_grad(prob._ctx, t.ctypes.data, int(nsteps), int(checkpoint), grad.ctypes.data,
                              ctypes.byref(dnu) if nu else None, ctypes.byref(J)), prob._ctx)
    return J.value, grad, (dnu.value if nu else None)
