"""MultiLayerQG on the device (mirror of ``src/multilayerqg.jl``).

``Problem(nlayers, dev; ...)`` takes the reference keywords (``src/multilayerqg.jl:87-114``),
for 1 to 8 layers (one layer: ``SingleLayerParams``, 193-219).  The per-wavenumber ``S^-1``
inversion (495-509, 563-583) and ``S`` product (520-552), the layered Jacobian with
background gradients ``Qx``, ``Qy`` (695-731), the bottom drag (658), the energies and the
fluxes (918-1022) run in the HIP library.  ``U`` may be a scalar, one value per layer, or an
``(ny, nlayers)`` array as in the reference.
"""
import math

import numpy as np

from . import _lib
from .problem import GPU, Grid, Problem as _Problem, attach_forcing, device_diagnostic, dptr, make_desc, stepforward  # noqa: F401


class Params:
    pass


class MLQGProblem(_Problem):
    _names = ("q", "ψ", "u", "v")


def _U_layers(U, n, ny):
    U = np.asarray(U, dtype=np.float64)
    if U.ndim == 0:
        return np.full((n, ny), float(U))
    if U.ndim == 1 and U.size == n:
        return np.broadcast_to(U[:, None], (n, ny)).copy()
    if U.ndim == 1:
        return np.broadcast_to(U[None, :], (n, ny)).copy()
    return np.ascontiguousarray(U.T)       # (ny, n) -> (n, ny)


def Problem(nlayers, dev=None, *, nx=128, ny=None, Lx=2 * math.pi, Ly=None, f0=1.0, β=0.0, U=None, H=None,
            b=None, eta=None, topographic_pv_gradient=(0.0, 0.0), μ=0.0, ν=0.0, nν=1, dt=0.01, stepper="RK4",
            calcFq=None, stochastic=False, linear=False, aliased_fraction=1 / 3, T=np.float64):
    n = int(nlayers)
    ny = nx if ny is None else ny
    Ly = Lx if Ly is None else Ly
    grid = Grid(int(nx), int(ny), Lx, Ly, aliased_fraction, T)
    d = make_desc(_lib.GF_MODEL_MLQG, nx, ny, Lx, Ly, T, stepper, dt, aliased_fraction)
    d.nlayers = n
    H = np.ones(n) / n if H is None else H
    b = -(1 + np.arange(n) / n) if b is None else b
    U = np.zeros(n) if U is None else U
    Ha = np.ascontiguousarray(np.asarray(H, dtype=np.float64).reshape(-1))
    ba = np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(-1))
    Ua = np.ascontiguousarray(_U_layers(U, n, ny))
    if n > 1 and (Ha.size != n or ba.size != n):
        raise ValueError("H and b need one value per layer")
    keep = [Ha, ba, Ua]
    d.H, d.b, d.Uy = dptr(Ha), dptr(ba), dptr(Ua)
    d.f0, d.beta, d.mu, d.nu, d.nnu, d.linear = float(f0), float(β), float(μ), float(ν), int(nν), int(bool(linear))
    d.topo_grad_x, d.topo_grad_y = float(topographic_pv_gradient[0]), float(topographic_pv_gradient[1])
    if eta is not None:
        if callable(eta):
            X, Y = grid.gridpoints()
            eta = eta(X, Y)
        e = np.ascontiguousarray(np.asarray(eta, dtype=np.float64).reshape(ny, nx))
        keep.append(e)
        d.eta = dptr(e)
    Fh = attach_forcing(d, calcFq)
    p = Params()
    p.nlayers, p.f0, p.β, p.H, p.b, p.U, p.μ, p.ν, p.nν, p.stepper = n, f0, β, Ha, ba, Ua, μ, ν, nν, stepper
    gp = ba[:-1] - ba[1:]
    p.gprime = gp
    prob = MLQGProblem(d, grid, p, nlayers=n, keep=keep)
    if Fh is not None:
        prob.set_forcing(Fh)
    return prob


def set_q(prob, q):
    """``set_q!``: rfft on the device, ``qh[1,1,:] = 0``, then ``updatevars!``."""
    prob._set_physical(q)
    prob._updatevars()


def set_ψ(prob, ψ):
    """``set_ψ!`` (src/multilayerqg.jl:873-889): ``qh = S ψh`` per wavenumber on the device
    (``pvfromstreamfunction!``, 520-552), then ``set_q!`` of ``irfft(qh)``."""
    a = np.ascontiguousarray(np.asarray(ψ, dtype=prob._rtype).reshape(prob._pshape))
    _lib.check(_lib.lib.gf_set_psi(prob._ctx, a.ctypes.data), prob._ctx)
    prob._updatevars()


set_psi = set_ψ


def updatevars(prob):
    prob._updatevars()


def _split(r):
    if len(r) == 1:
        return r[0]
    return np.array(r[: (len(r) + 1) // 2]), np.array(r[(len(r) + 1) // 2:])


@device_diagnostic(lambda p: (_lib.GF_DIAG_LAYER_ENERGIES, 0.0, 0, 2 * p.nlayers - 1), unpack=_split)
def energies(prob):
    """Layer kinetic energies KE_j and interface potential energies PE_{j+1/2}
    (src/multilayerqg.jl:918-960), computed on the device: returns (KE, PE) arrays, or the
    energy alone for one layer."""
    n = prob.nlayers
    return _split(prob._diagnostic(_lib.GF_DIAG_LAYER_ENERGIES, nout=2 * n - 1))


@device_diagnostic(lambda p: (_lib.GF_DIAG_FLUXES, 0.0, 0, 2 * p.nlayers - 1), unpack=_split)
def fluxes(prob):
    """Lateral eddy fluxes per layer and vertical eddy fluxes per interface
    (src/multilayerqg.jl:977-1022), computed on the device: returns (lateralfluxes,
    verticalfluxes), or the lateral flux alone for one layer.  Dealiases the state in place,
    as the reference's call of ``updatevars!`` does."""
    n = prob.nlayers
    return _split(prob._diagnostic(_lib.GF_DIAG_FLUXES, nout=2 * n - 1))
