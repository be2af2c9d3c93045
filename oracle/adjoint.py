"""Discrete adjoint of TwoDNavierStokes ETDRK4 time stepping (CPU restatement).

Test infrastructure only, like the rest of ``oracle/``.  The reference has no adjoint;
this module differentiates the oracle's own forward step (``fourierflows.ETDRK4Stepper``
with ``twodnavierstokes.calcN``, i.e. src/twodnavierstokes.jl:253-294 and the [FF]
ETDRK4 stages) exactly, operation by operation.  It is pinned by (tests/test_oracle_adjoint.py):

* central finite differences of the forward step and of the loss;
* the dot-product identity <J v, w> = <v, J^T w> between ``calcN_tlm`` and ``calcN_adj``
  and between the tangent and adjoint of a whole step;
* the closed form of the linear case (N = 0: the step is E sol, its adjoint E lambda).

Conventions.  A complex spectral array z (rfft half plane, shape (nl, nkr)) stands for the
real vector (Re z, Im z); its adjoint ("gradient") is lambda = dJ/dRe z + i dJ/dIm z, and the
inner product is <a, b> = sum Re(conj(a) b) over the half plane.  With N = nx ny and the
Hermitian weights w (2 for 0 < kr < nx/2, 1 on the kr = 0 and kr = nx/2 columns):

* adjoint of the normalised c2r transform f = irfft2(z):   lambda_z = (w / N) rfft2(lambda_f);
* adjoint of the unnormalised r2c transform z = rfft2(f):  lambda_f = N irfft2(lambda_z / w).
"""
import numpy as np

from . import fourierflows as off


def hermitian_weights(grid):
    w = np.full((grid.nl, grid.nkr), 2.0)
    w[:, 0] = 1.0
    if grid.nx % 2 == 0:
        w[:, -1] = 1.0
    return w


def _fields(s, grid):
    """u, v, zeta of a (masked) state, as calcN_advection forms them."""
    u = grid.irfft(1j * grid.Lw * grid.invKrsq * s)
    v = grid.irfft(-1j * grid.Kr * grid.invKrsq * s)
    z = grid.irfft(s)
    return u, v, z


def calcN(s, grid):
    """N of a state that is already masked (twodnavierstokes.calcN_advection)."""
    u, v, z = _fields(s, grid)
    return -1j * grid.Kr * grid.rfft(u * z) - 1j * grid.Lw * grid.rfft(v * z)


def calcN_tlm(ds, s, grid):
    """Tangent of ``calcN`` at the masked state s in the (masked) direction ds."""
    u, v, z = _fields(s, grid)
    du, dv, dz = _fields(ds, grid)
    return (-1j * grid.Kr * grid.rfft(du * z + u * dz)
            - 1j * grid.Lw * grid.rfft(dv * z + v * dz))


def calcN_adj(lam, s, grid, w=None):
    """Adjoint of ``calcN`` at the masked state s: maps lambda_N to lambda_s (unmasked)."""
    w = hermitian_weights(grid) if w is None else w
    u, v, z = _fields(s, grid)
    pa = grid.irfft(1j * grid.Kr * lam / w)      # N^{-1} lambda of (u zeta)
    pb = grid.irfft(1j * grid.Lw * lam / w)      # N^{-1} lambda of (v zeta)
    return w * (grid.rfft(u * pa + v * pb)
                - 1j * grid.Lw * grid.invKrsq * grid.rfft(z * pa)
                + 1j * grid.Kr * grid.invKrsq * grid.rfft(z * pb))


def etd_coefficient_derivatives(dt, L, ncirc=32, rcirc=1.0):
    """d/dL of the ETDRK4 coefficients (E, E2, zeta, alpha, beta, Gamma) for real L.

    E' = dt E, E2' = dt/2 E2; the contour means of ``fourierflows.etd_coefficients`` are
    differentiated under the mean: with z = dt L + r exp(i theta), d/dL <f(z)> = dt <f'(z)>,
      zeta:  f = (e^{z/2} - 1)/z,                       f' = (e^{z/2}(z/2 - 1) + 1)/z^2
      alpha: f = (-4 - z + e^z(4 - 3z + z^2))/z^3,      f' = (12 + 2z + e^z(-12 + 10z - 4z^2 + z^3))/z^4
      beta:  f = (2 + z + e^z(z - 2))/z^3,              f' = (-6 - 2z + e^z(6 - 4z + z^2))/z^4
      Gamma: f = (-4 - 3z - z^2 + e^z(4 - z))/z^3,      f' = (12 + 6z + z^2 + e^z(-12 + 6z - z^2))/z^4
    each multiplied by dt (the coefficients carry a factor dt).
    """
    L = np.asarray(L, dtype=complex)
    flat = L.ravel()
    uniq, inv = np.unique(flat, return_inverse=True)
    th = 2 * np.pi / ncirc * (np.arange(ncirc) + 0.5)
    z = dt * uniq[:, None] + rcirc * np.exp(1j * th)[None, :]
    ez, ez2 = np.exp(z), np.exp(z / 2)
    dz = dt * dt * np.mean((ez2 * (z / 2 - 1) + 1) / z ** 2, axis=1)
    da = dt * dt * np.mean((12 + 2 * z + ez * (-12 + 10 * z - 4 * z ** 2 + z ** 3)) / z ** 4, axis=1)
    db = dt * dt * np.mean((-6 - 2 * z + ez * (6 - 4 * z + z ** 2)) / z ** 4, axis=1)
    dg = dt * dt * np.mean((12 + 6 * z + z ** 2 + ez * (-12 + 6 * z - z ** 2)) / z ** 4, axis=1)
    shape = L.shape
    out = [v[inv].reshape(shape) for v in (dz, da, db, dg)]
    if np.all(L.imag == 0):
        out = [np.real(v) for v in out]
    dE = dt * np.exp(dt * L)
    dE2 = dt / 2 * np.exp(dt * L / 2)
    if np.all(L.imag == 0):
        dE, dE2 = dE.real, dE2.real
    return (dE, dE2) + tuple(out)


def twodns_L(grid, nu, nnu, mu=0.0, nmu=0):
    """L = -nu Krsq^nnu - mu Krsq^nmu, L[1,1] = 0 (src/twodnavierstokes.jl:150-151), and dL/dnu."""
    K2 = grid.Krsq
    L = -nu * K2 ** nnu - mu * K2 ** nmu
    dLdnu = -(K2 ** nnu)
    L[0, 0] = 0
    dLdnu[0, 0] = 0
    return L, dLdnu


class ETDStep:
    """One ETDRK4 step of the oracle with the stage states kept for the adjoint.

    Follows ``fourierflows.ETDRK4Stepper.step`` with calcN's in-place ``dealias!`` (P1):
    s0 = M sol, N1 = N(s0), s1 = E2 s0 + Z N1, N2 = N(M s1), s2 = E2 s0 + Z N2,
    N3 = N(M s2), s2' = E2 (M s1) + Z (2 N3 - N1), N4 = N(M s2'),
    sol' = E s0 + a N1 + 2 b (N2 + N3) + G N4.
    """

    def __init__(self, grid, L, dt, dLdnu=None):
        self.grid, self.dt, self.L = grid, dt, L
        self.E = np.exp(dt * L)
        self.E2 = np.exp(dt * L / 2)
        self.Z, self.A, self.B, self.G = off.etd_coefficients(dt, L)
        self.M = grid.dealias_mask()
        self.w = hermitian_weights(grid)
        self.dLdnu = dLdnu
        if dLdnu is not None:
            d = etd_coefficient_derivatives(dt, L)
            self.dE, self.dE2, self.dZ, self.dA, self.dB, self.dG = (dLdnu * v for v in d)

    def forward(self, sol):
        M, E2, Z = self.M, self.E2, self.Z
        s0 = M * sol
        N1 = calcN(s0, self.grid)
        s1 = M * (E2 * s0 + Z * N1)
        N2 = calcN(s1, self.grid)
        s2 = M * (E2 * s0 + Z * N2)
        N3 = calcN(s2, self.grid)
        s3 = M * (E2 * s1 + Z * (2 * N3 - N1))
        N4 = calcN(s3, self.grid)
        new = self.E * s0 + self.A * N1 + 2 * self.B * (N2 + N3) + self.G * N4
        return new, (s0, s1, s2, s3, N1, N2, N3, N4)

    def tlm(self, dsol, stages):
        M, E2, Z, g = self.M, self.E2, self.Z, self.grid
        s0, s1, s2, s3 = stages[:4]
        d0 = M * dsol
        dN1 = calcN_tlm(d0, s0, g)
        d1 = M * (E2 * d0 + Z * dN1)
        dN2 = calcN_tlm(d1, s1, g)
        d2 = M * (E2 * d0 + Z * dN2)
        dN3 = calcN_tlm(d2, s2, g)
        d3 = M * (E2 * d1 + Z * (2 * dN3 - dN1))
        dN4 = calcN_tlm(d3, s3, g)
        return self.E * d0 + self.A * dN1 + 2 * self.B * (dN2 + dN3) + self.G * dN4

    def adjoint(self, lam, stages, dnu=None):
        """lambda of sol' -> lambda of sol (the coefficients are real for TwoDNS).  With
        ``dnu`` (a one-element list) the step's contribution to dJ/dnu is added to dnu[0]."""
        M, E2, Z, g, w = self.M, self.E2, self.Z, self.grid, self.w
        s0, s1, s2, s3 = stages[:4]
        l0 = self.E * lam
        lN1 = self.A * lam
        lN2 = 2 * self.B * lam
        lN3 = 2 * self.B * lam
        lN4 = self.G * lam
        l3 = M * calcN_adj(lN4, s3, g, w)
        l1 = E2 * l3
        lN3 = lN3 + 2 * Z * l3
        lN1 = lN1 - Z * l3
        l2 = M * calcN_adj(lN3, s2, g, w)
        l0 = l0 + E2 * l2
        lN2 = lN2 + Z * l2
        l1 = M * (l1 + calcN_adj(lN2, s1, g, w))
        l0 = l0 + E2 * l1
        lN1 = lN1 + Z * l1
        if dnu is not None:
            N1, N2, N3, N4 = stages[4:]
            dnu[0] += (inner(lam, self.dE * s0 + self.dA * N1 + 2 * self.dB * (N2 + N3) + self.dG * N4)
                       + inner(l3, self.dE2 * s1 + self.dZ * (2 * N3 - N1))
                       + inner(l2, self.dE2 * s0 + self.dZ * N2)
                       + inner(l1, self.dE2 * s0 + self.dZ * N1))
        l0 = l0 + calcN_adj(lN1, s0, g, w)
        return M * l0


def inner(a, b):
    """<a, b> = sum Re(conj(a) b) over the half plane (or over the grid for real arrays)."""
    return float(np.sum(np.real(np.conj(a) * b)))


def set_zeta(zeta0, grid):
    """set_zeta! + updatevars!' dealias: sol0 = M Z rfft2(zeta0) (src/twodnavierstokes.jl:352-360)."""
    s = grid.rfft(zeta0)
    s[0, 0] = 0
    return grid.dealias_mask() * s


def set_zeta_adj(lam, grid, w=None):
    w = hermitian_weights(grid) if w is None else w
    m = grid.dealias_mask() * lam
    m[0, 0] = 0
    return grid.nx * grid.ny * grid.irfft(m / w)


def loss_and_gradient(zeta0, target, grid, nu, nnu, dt, nsteps, checkpoint=50, mu=0.0, nmu=0):
    """J = 1/2 sum (zeta(T) - target)^2 with zeta(T) = irfft2(M sol_T), its gradient with respect
    to zeta0 and dJ/dnu (checkpoints every ``checkpoint`` steps, segments recomputed)."""
    L, dLdnu = twodns_L(grid, nu, nnu, mu, nmu)
    st = ETDStep(grid, L, dt, dLdnu)
    sol = set_zeta(zeta0, grid)
    ckpts = {}
    for n in range(nsteps):
        if n % checkpoint == 0:
            ckpts[n] = sol.copy()
        sol, _ = st.forward(sol)
    zT = grid.irfft(st.M * sol)
    r = zT - target
    J = 0.5 * float(np.sum(r * r))
    lam = st.M * (st.w / (grid.nx * grid.ny)) * grid.rfft(r)
    dnu = [0.0]
    for c in sorted(ckpts, reverse=True):
        s = ckpts[c].copy()
        seg = []
        for n in range(c, min(c + checkpoint, nsteps)):
            s_new, stages = st.forward(s)
            seg.append(stages)
            s = s_new
        for stages in reversed(seg):
            lam = st.adjoint(lam, stages, dnu)
    return J, set_zeta_adj(lam, grid, st.w), dnu[0]
