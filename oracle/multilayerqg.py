"""Oracle restatement of ``GeophysicalFlows.MultiLayerQG`` (``src/multilayerqg.jl``).

* ``Params``: 255-320 (``g' = b[:-1]-b[1:]``, ``Fm = f0^2/(g' H[1:])``,
  ``Fp = f0^2/(g' H[:-1])``; ``Qx`` bottom layer ``= d eta/dx``; ``Qy = beta - U_yy``
  minus the layer coupling of U, 312-316).
* ``calcS``/``calcS_inv``: 603-634 (``S = -k^2 I + F``, ``S^-1 = I/S`` with k^2=0 -> 1,
  then ``S^-1[0,0] = 0``).
* ``hyperviscosity``/``Equation``: 338-377 (``L = -nu Krsq^nnu`` per layer, ``L[0,0,:] = 0``).
* ``streamfunctionfrompv``/``pv_streamfunction_kernel!``: 495-509, 563-595.
* ``calcN``: 651-663 (dealias P1, bottom drag ``N[n] += mu Krsq psih[n]``);
  ``calcN_advection``: 695-731; ``calcNlinear``/``calcN_linearadvection``: 675-683, 744-779.
* ``updatevars``/``set_q``/``set_psi``: 819-894; ``energies``: 918-960; ``fluxes``: 977-1022.

Array layout: spectral ``(nlayers, nl, nkr)``, physical ``(nlayers, ny, nx)`` -- the
reference's ``(nkr, nl, nlayers)``/``(nx, ny, nlayers)`` column-major arrays.
"""
import numpy as np

from .fourierflows import Equation as _Equation, Problem as _Problem, TwoDGrid, dealias, parsevalsum


class Params:
    pass


class Vars:
    def __init__(self, grid, n, forced=False, stochastic=False):
        z = lambda: np.zeros((n, grid.ny, grid.nx))
        zh = lambda: np.zeros((n, grid.nl, grid.nkr), complex)
        self.q, self.psi, self.u, self.v = z(), z(), z(), z()
        self.qh, self.psih, self.uh, self.vh = zh(), zh(), zh(), zh()
        self.Fqh = zh() if forced else None
        self.prevsol = zh() if stochastic else None


def _layer_coupling(f0, H, b, n):
    b = np.asarray(b, dtype=float)
    H = np.asarray(H, dtype=float)
    gp = b[:-1] - b[1:]
    Fm = f0 ** 2 / (gp * H[1:])
    Fp = f0 ** 2 / (gp * H[:-1])
    F = np.zeros((n, n))
    for j in range(n):
        F[j, j] = -((Fp[j] if j < n - 1 else 0.0) + (Fm[j - 1] if j > 0 else 0.0))
        if j < n - 1:
            F[j, j + 1] = Fp[j]
        if j > 0:
            F[j, j - 1] = Fm[j - 1]
    return gp, Fm, Fp, F


def _U3D(U, n, grid):
    U = np.asarray(U, dtype=float)
    if U.ndim == 0:
        return np.full((n, grid.ny), float(U))
    if U.ndim == 1 and U.size == n:
        return np.broadcast_to(U[:, None], (n, grid.ny)).copy()
    if U.ndim == 1:
        return np.broadcast_to(U[None, :], (n, grid.ny)).copy()
    # (ny, n) as in the reference's U matrix
    return U.T.copy()


def make_params(n, grid, f0, beta, b, H, U, eta, topographic_pv_gradient, mu, nu, nnu, calcFq):
    p = Params()
    p.nlayers, p.f0, p.beta, p.mu, p.nu, p.nnu, p.calcFq = n, float(f0), float(beta), float(mu), \
        float(nu), int(nnu), calcFq
    p.H = np.asarray(H, dtype=float)
    p.U = _U3D(U, n, grid)                                    # (n, ny)
    Uh = np.fft.fft(p.U, axis=1)
    Uyy = np.real(np.fft.ifft(-grid.l[None, :] ** 2 * Uh, axis=1))  # (n, ny)
    if eta is None:
        eta = np.zeros((grid.ny, grid.nx))
    elif callable(eta):
        X, Y = grid.gridpoints()
        eta = eta(X, Y)
    p.eta = np.asarray(eta, dtype=float)
    etah = grid.rfft(p.eta)
    etax = grid.irfft(1j * grid.Kr * etah) + topographic_pv_gradient[0]
    etay = grid.irfft(1j * grid.Lw * etah) + topographic_pv_gradient[1]
    p.Qx = np.zeros((n, grid.ny, grid.nx))
    p.Qx[n - 1] += etax
    p.Qy = np.broadcast_to((p.beta - Uyy)[:, :, None], (n, grid.ny, grid.nx)).copy()
    p.Qy[n - 1] += etay
    if n == 1:
        p.S = p.Sinv = None
        return p
    p.b = np.asarray(b, dtype=float)
    p.gp, p.Fm, p.Fp, F = _layer_coupling(f0, H, b, n)
    Ucol = p.U[:, :, None]
    p.Qy[0] = p.Qy[0] - p.Fp[0] * (Ucol[1] - Ucol[0])
    for j in range(1, n - 1):
        p.Qy[j] = p.Qy[j] - p.Fp[j] * (Ucol[j + 1] - Ucol[j]) - p.Fm[j - 1] * (Ucol[j - 1] - Ucol[j])
    p.Qy[n - 1] = p.Qy[n - 1] - p.Fm[n - 2] * (Ucol[n - 2] - Ucol[n - 1])
    k2 = grid.Krsq
    I = np.eye(n)
    p.S = -k2[:, :, None, None] * I + F                       # (nl, nkr, n, n)
    k2i = np.where(k2 == 0, 1.0, k2)
    p.Sinv = np.linalg.inv(-k2i[:, :, None, None] * I + F)
    p.Sinv[0, 0] = 0
    p.Fmat = F
    return p


def streamfunctionfrompv(psih, qh, p, grid):
    if p.nlayers == 1:
        psih[...] = -grid.invKrsq * qh
        return
    # psih[:, i, j] = Sinv[i, j] @ qh[:, i, j]
    psih[...] = np.einsum("ijab,bij->aij", p.Sinv, qh)


def pvfromstreamfunction(qh, psih, p, grid):
    if p.nlayers == 1:
        qh[...] = -grid.Krsq * psih
        return
    qh[...] = np.einsum("ijab,bij->aij", p.S, psih)


def hyperviscosity(p, grid):
    L = np.empty((p.nlayers, grid.nl, grid.nkr))
    L[...] = -p.nu * grid.Krsq ** p.nnu
    L[:, 0, 0] = 0
    return L


def calcN_advection(N, sol, vars, p, grid):
    vars.qh[...] = sol
    streamfunctionfrompv(vars.psih, vars.qh, p, grid)
    vars.uh[...] = -1j * grid.Lw * vars.psih
    vars.vh[...] = 1j * grid.Kr * vars.psih
    vars.u[...] = grid.irfft(vars.uh)
    vars.u += p.U[:, :, None]
    N[...] = -grid.rfft(vars.u * p.Qx)
    vars.v[...] = grid.irfft(vars.vh)
    N -= grid.rfft(vars.v * p.Qy)
    vars.q[...] = grid.irfft(vars.qh)
    uqh = grid.rfft(vars.u * vars.q)
    vqh = grid.rfft(vars.v * vars.q)
    N -= 1j * grid.Kr * uqh + 1j * grid.Lw * vqh


def calcN_linearadvection(N, sol, vars, p, grid):
    vars.qh[...] = sol
    streamfunctionfrompv(vars.psih, vars.qh, p, grid)
    vars.uh[...] = -1j * grid.Lw * vars.psih
    vars.vh[...] = 1j * grid.Kr * vars.psih
    vars.u[...] = grid.irfft(vars.uh)
    vars.u += p.U[:, :, None]
    N[...] = -grid.rfft(vars.u * p.Qx)
    vars.v[...] = grid.irfft(vars.vh)
    N -= grid.rfft(vars.v * p.Qy)
    vars.q[...] = grid.irfft(vars.qh)
    Uq = np.broadcast_to(p.U[:, :, None], vars.q.shape) * vars.q
    N -= 1j * grid.Kr * grid.rfft(Uq)


def addforcing(N, sol, t, clock, vars, p, grid):
    F = p.calcFq
    if F is None:
        return
    if callable(F):
        F(vars.Fqh, sol, t, clock, vars, p, grid)
    else:
        vars.Fqh[...] = F
    N += vars.Fqh


def calcN(N, sol, t, clock, vars, p, grid):
    n = p.nlayers
    dealias(sol, grid)
    calcN_advection(N, sol, vars, p, grid)
    N[n - 1] += p.mu * grid.Krsq * vars.psih[n - 1]
    addforcing(N, sol, t, clock, vars, p, grid)


def calcNlinear(N, sol, t, clock, vars, p, grid):
    n = p.nlayers
    calcN_linearadvection(N, sol, vars, p, grid)
    N[n - 1] += p.mu * grid.Krsq * vars.psih[n - 1]
    addforcing(N, sol, t, clock, vars, p, grid)


def Problem(nlayers, nx=128, ny=None, Lx=2 * np.pi, Ly=None, f0=1.0, beta=0.0, U=None, H=None,
            b=None, eta=None, topographic_pv_gradient=(0.0, 0.0), mu=0.0, nu=0.0, nnu=1, dt=0.01,
            stepper="RK4", calcFq=None, stochastic=False, linear=False, aliased_fraction=1 / 3):
    n = nlayers
    U = np.zeros(n) if U is None else U
    H = np.ones(n) / n if H is None else H
    b = -(1 + np.arange(n) / n) if b is None else b
    grid = TwoDGrid(nx, Lx, ny, Ly, aliased_fraction)
    p = make_params(n, grid, f0, beta, b, H, U, eta, topographic_pv_gradient, mu, nu, nnu, calcFq)
    vars = Vars(grid, n, forced=calcFq is not None, stochastic=stochastic)
    eq = _Equation(hyperviscosity(p, grid), calcNlinear if linear else calcN, grid)
    return _Problem(eq, stepper, dt, grid, vars, p)


def updatevars(prob):
    vars, p, grid, sol = prob.vars, prob.params, prob.grid, prob.sol
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
    vars, grid = prob.vars, prob.grid
    vars.qh[...] = grid.rfft(np.asarray(q).reshape(vars.q.shape))
    vars.qh[:, 0, 0] = 0
    prob.sol[...] = vars.qh
    updatevars(prob)


def set_psi(prob, psi):
    vars, p, grid = prob.vars, prob.params, prob.grid
    vars.psih[...] = grid.rfft(np.asarray(psi).reshape(vars.psi.shape))
    pvfromstreamfunction(vars.qh, vars.psih, p, grid)
    vars.q[...] = grid.irfft(vars.qh)
    set_q(prob, vars.q.copy())


def energies(prob):
    vars, p, grid, sol = prob.vars, prob.params, prob.grid, prob.sol
    n = p.nlayers
    vars.qh[...] = sol
    streamfunctionfrompv(vars.psih, vars.qh, p, grid)
    a = grid.Krsq * np.abs(vars.psih) ** 2
    if n == 1:
        return 1 / (2 * grid.Lx * grid.Ly) * parsevalsum(a[0], grid)
    V = grid.Lx * grid.Ly * np.sum(p.H)
    KE = np.array([1 / (2 * V) * parsevalsum(a[j], grid) * p.H[j] for j in range(n)])
    PE = np.array([1 / (2 * V) * p.f0 ** 2 / p.gp[j] *
                   parsevalsum(np.abs(vars.psih[j] - vars.psih[j + 1]) ** 2, grid) for j in range(n - 1)])
    return KE, PE


def fluxes(prob):
    """Lateral eddy fluxes per layer and vertical eddy fluxes per interface
    (src/multilayerqg.jl:977-1022): ``updatevars!`` (which dealiases sol), then
    ``du/dy = irfft(i l uh)`` and the physical-space sums times ``dx dy / V``."""
    vars, p, grid = prob.vars, prob.params, prob.grid
    n = p.nlayers
    updatevars(prob)
    dudy = grid.irfft(1j * grid.Lw * vars.uh)
    U = p.U[:, :, None]                                           # (n, ny, 1)
    dxdy = grid.dx * grid.dy
    if n == 1:
        return float(np.sum(U[0] * vars.v[0] * dudy[0])) * dxdy / (grid.Lx * grid.Ly)
    V = grid.Lx * grid.Ly * np.sum(p.H)
    lateral = p.H * np.array([np.sum(U[j] * vars.v[j] * dudy[j]) for j in range(n)]) * dxdy / V
    vertical = np.array([np.sum(p.f0 ** 2 / p.gp[j] * (U[j] - U[j + 1]) * vars.v[j + 1] * vars.psi[j])
                         for j in range(n - 1)]) * dxdy / V
    return lateral, vertical
