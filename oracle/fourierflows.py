"""Oracle restatement of the FourierFlows.jl 0.10 pieces the reference calls.

FourierFlows.jl is a dependency of the reference (``Project.toml:25``, compat
``0.10.5``) and is not vendored; the rules below restate its published
algorithm and are anchored on the reference call sites:

* ``TwoDGrid``        -- call sites ``src/twodnavierstokes.jl:91``,
  ``src/singlelayerqg.jl:105``, ``src/multilayerqg.jl:126``.
* ``dealias!``        -- ``src/twodnavierstokes.jl:287``, ``src/singlelayerqg.jl:444``,
  ``src/multilayerqg.jl:654``.
* rfft plan ``mul!``/``ldiv!`` -- ``src/twodnavierstokes.jl:258-270`` (forward
  unnormalised, inverse normalised by ``1/(nx*ny)``; ``docs/src/index.md:25-35``).
* ``parsevalsum``     -- imported at ``src/twodnavierstokes.jl:27``.
* time steppers ``ETDRK4``, ``FilteredETDRK4``, ``RK4``, ``FilteredRK4``,
  ``ForwardEuler`` -- stepper names at ``test/runtests.jl:58``; ``prob.timestepper.N``
  at ``test/test_twodnavierstokes.jl:316``; ``filter`` at
  ``examples/twodnavierstokes_decaying.jl:65``.

Layout: physical ``(ny, nx)``, spectral ``(nl, nkr)`` (the reference's
column-major ``(nx, ny)`` / ``(nkr, nl)`` arrays, same bytes).
"""
import math
import os

import numpy as np

try:  # pocketfft with threads, the stand-in for FFTW threads (README.md:78-86)
    import scipy.fft as _sfft
except Exception:  # pragma: no cover
    _sfft = None

_WORKERS = int(os.environ.get("GF_ORACLE_FFT_WORKERS", "1"))


def set_fft_workers(n):
    """Threads used by the oracle's 2D FFTs (1 by default; bench sets nproc)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def _rfft2(f):
    if _sfft is not None:
        return _sfft.rfft2(f, workers=_WORKERS)
    return np.fft.rfft2(f)


def _irfft2(fh, nx):
    if _sfft is not None:
        return _sfft.irfft2(fh, s=(fh.shape[-2], nx), workers=_WORKERS)
    return np.fft.irfft2(fh, s=(fh.shape[-2], nx))


# ---------------------------------------------------------------------------
# Grid
# ---------------------------------------------------------------------------

def aliased_wavenumbers(nk, nkr, aliased_fraction):
    """0-based index ranges zeroed by ``dealias!`` ([FF] getaliasedwavenumbers).

    ``iL = floor((1-f)/2 * n) + 1``, ``iR = ceil((1+f)/2 * n)`` (1-based); the
    full-axis alias range is ``iL:iR`` and the rfft-axis range ``iL:nkr``.
    Returns ``(kalias, kralias)`` as Python ranges (0-based), empty when f == 0.
    """
    if aliased_fraction <= 0:
        return range(0), range(0)
    L = (1 - aliased_fraction) / 2
    R = (1 + aliased_fraction) / 2
    iL = math.floor(L * nk) + 1
    iR = math.ceil(R * nk)
    return range(iL - 1, iR), range(iL - 1, nkr)


class TwoDGrid:
    """Restates [FF] ``TwoDGrid(dev; nx, Lx, ny, Ly, aliased_fraction, T)``."""

    def __init__(self, nx, Lx, ny=None, Ly=None, aliased_fraction=1 / 3, T=np.float64):
        ny = nx if ny is None else ny
        Ly = Lx if Ly is None else Ly
        self.nx, self.ny, self.Lx, self.Ly = int(nx), int(ny), float(Lx), float(Ly)
        self.T = T
        self.nkr, self.nl = self.nx // 2 + 1, self.ny
        self.dx, self.dy = self.Lx / self.nx, self.Ly / self.ny
        self.x = -self.Lx / 2 + np.arange(self.nx) * self.dx
        self.y = -self.Ly / 2 + np.arange(self.ny) * self.dy
        # rfftfreq / fftfreq scaled by 2pi/L: kr = 0..nx/2, l in FFT order
        self.kr = np.arange(self.nkr) * (2 * np.pi / self.Lx)
        lidx = np.arange(self.ny)
        lidx = np.where(lidx < self.ny // 2, lidx, lidx - self.ny) if self.ny % 2 == 0 else \
            np.where(lidx <= self.ny // 2, lidx, lidx - self.ny)
        self.l = lidx * (2 * np.pi / self.Ly)
        self.Kr = np.broadcast_to(self.kr[None, :], (self.nl, self.nkr)).copy()
        self.Lw = np.broadcast_to(self.l[:, None], (self.nl, self.nkr)).copy()
        self.Krsq = self.Kr ** 2 + self.Lw ** 2
        with np.errstate(divide="ignore"):
            self.invKrsq = 1.0 / self.Krsq
        self.invKrsq[0, 0] = 0.0
        self.aliased_fraction = aliased_fraction
        _, self.kralias = aliased_wavenumbers(self.nx, self.nkr, aliased_fraction)
        self.lalias, _ = aliased_wavenumbers(self.ny, self.ny // 2 + 1, aliased_fraction)

    def gridpoints(self):
        X = np.broadcast_to(self.x[None, :], (self.ny, self.nx)).copy()
        Y = np.broadcast_to(self.y[:, None], (self.ny, self.nx)).copy()
        return X, Y

    def dealias_mask(self):
        """1 where a state value survives ``dealias!``, 0 in the aliased band."""
        m = np.ones((self.nl, self.nkr))
        if len(self.kralias):
            m[:, self.kralias.start:self.kralias.stop] = 0
        if len(self.lalias):
            m[self.lalias.start:self.lalias.stop, :] = 0
        return m

    # FFT conventions of the reference's rfftplan (src/twodnavierstokes.jl:258-270)
    def rfft(self, f):
        return _rfft2(f)

    def irfft(self, fh):
        return _irfft2(fh, self.nx)


def dealias(fh, grid):
    """In-place ``dealias!`` (P1): zero ``kralias`` columns and ``lalias`` rows."""
    if len(grid.kralias):
        fh[..., :, grid.kralias.start:grid.kralias.stop] = 0
    if len(grid.lalias):
        fh[..., grid.lalias.start:grid.lalias.stop, :] = 0
    return None


def parsevalsum(uh, grid):
    """[FF] ``parsevalsum``: kr=0 column once, kr>0 columns twice (rfft layout)."""
    U = np.sum(uh[..., :, 0]) + 2 * np.sum(uh[..., :, 1:])
    norm = grid.Lx * grid.Ly / (grid.nx ** 2 * grid.ny ** 2)
    return norm * np.real(U)


def parsevalsum2(uh, grid):
    U = np.sum(np.abs(uh[..., :, 0]) ** 2) + 2 * np.sum(np.abs(uh[..., :, 1:]) ** 2)
    norm = grid.Lx * grid.Ly / (grid.nx ** 2 * grid.ny ** 2)
    return norm * U


def makefilter(grid, order=4, innerK=0.65, outerK=1.0, tol=1e-15):
    """[FF] ``makefilter(g::TwoDGrid)``: ``exp(-decay (K-innerK)^order)`` for K >= innerK.

    ``K = sqrt((kr dx/pi)^2 + (l dy/pi)^2)``, ``decay = -log(tol)/(outerK-innerK)^order``.
    """
    K = np.sqrt((grid.Kr * grid.dx / np.pi) ** 2 + (grid.Lw * grid.dy / np.pi) ** 2)
    decay = -np.log(tol) / (outerK - innerK) ** order
    filt = np.exp(-decay * (K - innerK) ** order)
    filt[K < innerK] = 1.0
    return filt


def etd_coefficients(dt, L, ncirc=32, rcirc=1.0):
    """[FF] ``getetdcoeffs``: Kassam--Trefethen contour means of the phi functions.

    ``z = dt L + r exp(2 pi i (m + 1/2)/M)``, ``m = 0..M-1``;
    ``zeta = dt <(e^{z/2}-1)/z>``, ``alpha = dt <(-4-z+e^z(4-3z+z^2))/z^3>``,
    ``beta = dt <(2+z+e^z(z-2))/z^3>``, ``Gamma = dt <(-4-3z-z^2+e^z(4-z))/z^3>``;
    real parts are taken when ``L`` is real.
    """
    L = np.asarray(L)
    real = not np.iscomplexobj(L)
    # the coefficients depend on L pointwise: evaluate once per distinct value
    uniq, inv = np.unique(L.ravel(), return_inverse=True)
    if uniq.size < L.size:
        out = etd_coefficients(dt, uniq, ncirc, rcirc)
        return tuple(c[inv].reshape(L.shape) for c in out)
    circ = rcirc * np.exp(2j * np.pi / ncirc * (np.arange(ncirc) + 0.5))
    zeta = np.zeros(L.shape, complex)
    alpha = np.zeros(L.shape, complex)
    beta = np.zeros(L.shape, complex)
    gamma = np.zeros(L.shape, complex)
    dtL = dt * L.astype(complex)
    for c in circ:  # sequential accumulation over contour points
        z = dtL + c
        ez = np.exp(z)
        z3 = z ** 3
        zeta += (np.exp(z / 2) - 1) / z
        alpha += (-4 - z + ez * (4 - 3 * z + z ** 2)) / z3
        beta += (2 + z + ez * (-2 + z)) / z3
        gamma += (-4 - 3 * z - z ** 2 + ez * (4 - z)) / z3
    out = []
    for c in (zeta, alpha, beta, gamma):
        c = dt * (c / ncirc)
        out.append(np.real(c).copy() if real else c)
    return tuple(out)


# ---------------------------------------------------------------------------
# Problem / Clock / steppers
# ---------------------------------------------------------------------------

class Clock:
    def __init__(self, dt, t=0.0, step=0):
        self.dt, self.t, self.step = dt, t, step


class Equation:
    """[FF] ``Equation(L, calcN!, grid)``."""

    def __init__(self, L, calcN, grid):
        self.L, self.calcN, self.grid = L, calcN, grid


class ETDRK4Stepper:
    def __init__(self, eq, dt, filtered=False):
        L = eq.L
        self.expLdt = np.exp(dt * L)
        self.expLdt2 = np.exp(dt * L / 2)
        self.zeta, self.alpha, self.beta, self.gamma = etd_coefficients(dt, L)
        shape = L.shape
        self.sol1 = np.zeros(shape, complex)
        self.sol2 = np.zeros(shape, complex)
        self.N1, self.N2, self.N3, self.N4 = (np.zeros(shape, complex) for _ in range(4))
        self.filter = makefilter(eq.grid) if filtered else None

    def step(self, sol, clock, eq, vars, params, grid):
        t, dt = clock.t, clock.dt
        eq.calcN(self.N1, sol, t, clock, vars, params, grid)
        self.sol1[...] = self.expLdt2 * sol + self.zeta * self.N1
        eq.calcN(self.N2, self.sol1, t + dt / 2, clock, vars, params, grid)
        self.sol2[...] = self.expLdt2 * sol + self.zeta * self.N2
        eq.calcN(self.N3, self.sol2, t + dt / 2, clock, vars, params, grid)
        self.sol2[...] = self.expLdt2 * self.sol1 + self.zeta * (2 * self.N3 - self.N1)
        eq.calcN(self.N4, self.sol2, t + dt, clock, vars, params, grid)
        sol[...] = (self.expLdt * sol + self.alpha * self.N1
                    + 2 * self.beta * (self.N2 + self.N3) + self.gamma * self.N4)
        if self.filter is not None:
            sol *= self.filter


class RK4Stepper:
    def __init__(self, eq, dt, filtered=False):
        shape = eq.L.shape
        self.sol1 = np.zeros(shape, complex)
        self.RHS1, self.RHS2, self.RHS3, self.RHS4 = (np.zeros(shape, complex) for _ in range(4))
        self.filter = makefilter(eq.grid) if filtered else None

    def step(self, sol, clock, eq, vars, params, grid):
        t, dt, L = clock.t, clock.dt, eq.L
        eq.calcN(self.RHS1, sol, t, clock, vars, params, grid)
        self.RHS1 += L * sol
        self.sol1[...] = sol + dt / 2 * self.RHS1
        eq.calcN(self.RHS2, self.sol1, t + dt / 2, clock, vars, params, grid)
        self.RHS2 += L * self.sol1
        self.sol1[...] = sol + dt / 2 * self.RHS2
        eq.calcN(self.RHS3, self.sol1, t + dt / 2, clock, vars, params, grid)
        self.RHS3 += L * self.sol1
        self.sol1[...] = sol + dt * self.RHS3
        eq.calcN(self.RHS4, self.sol1, t + dt, clock, vars, params, grid)
        self.RHS4 += L * self.sol1
        upd = sol + dt * (self.RHS1 / 6 + self.RHS2 / 3 + self.RHS3 / 3 + self.RHS4 / 6)
        if self.filter is not None:
            upd = self.filter * upd
        sol[...] = upd


class ForwardEulerStepper:
    def __init__(self, eq, dt, filtered=False):
        self.N = np.zeros(eq.L.shape, complex)
        self.filter = makefilter(eq.grid) if filtered else None

    def step(self, sol, clock, eq, vars, params, grid):
        eq.calcN(self.N, sol, clock.t, clock, vars, params, grid)
        sol += clock.dt * (self.N + eq.L * sol)
        if self.filter is not None:
            sol *= self.filter


class AB3Stepper:
    """[FF] ``AB3TimeStepper`` ("AB3", "FilteredAB3" at ``test/runtests.jl:58``): with
    ``RHS = N + L sol``, forward Euler ``sol += dt RHS`` while ``clock.step < 3``, then
    ``sol += dt (23/12 RHS - 16/12 RHS_1 + 5/12 RHS_2)``; the RHS history shifts every step.
    Parity: the linear part is pinned by the reference's Rossby-wave tests; the number of
    forward Euler start-up steps (3, recalled from FourierFlows 0.10) is parity unpinned."""

    def __init__(self, eq, dt, filtered=False):
        shape = eq.L.shape
        self.RHS, self.RHS1, self.RHS2 = (np.zeros(shape, complex) for _ in range(3))
        self.filter = makefilter(eq.grid) if filtered else None

    def step(self, sol, clock, eq, vars, params, grid):
        eq.calcN(self.RHS, sol, clock.t, clock, vars, params, grid)
        self.RHS += eq.L * sol
        if clock.step < 3:
            sol += clock.dt * self.RHS
        else:
            sol += clock.dt * (23 / 12 * self.RHS - 16 / 12 * self.RHS1 + 5 / 12 * self.RHS2)
        if self.filter is not None:
            sol *= self.filter
        self.RHS2[...] = self.RHS1
        self.RHS1[...] = self.RHS


STEPPERS = {
    "ETDRK4": (ETDRK4Stepper, False), "FilteredETDRK4": (ETDRK4Stepper, True),
    "RK4": (RK4Stepper, False), "FilteredRK4": (RK4Stepper, True),
    "ForwardEuler": (ForwardEulerStepper, False), "FilteredForwardEuler": (ForwardEulerStepper, True),
    "AB3": (AB3Stepper, False), "FilteredAB3": (AB3Stepper, True),
}


class Problem:
    """[FF] ``Problem(eqn, stepper, dt, grid, vars, params)``."""

    def __init__(self, eq, stepper, dt, grid, vars, params):
        if stepper not in STEPPERS:
            raise ValueError(f"unknown stepper {stepper!r}")
        cls, filtered = STEPPERS[stepper]
        self.eqn, self.grid, self.vars, self.params = eq, grid, vars, params
        self.sol = np.zeros(eq.L.shape, complex)
        self.clock = Clock(dt)
        self.timestepper = cls(eq, dt, filtered)
        self.stepper = stepper


def stepforward(prob, nsteps=1, diags=None):
    """[FF] ``stepforward!(prob[, diags], n)``; ``diags`` is a list of callables
    ``f(prob)`` whose values are appended after every step (``increment!``)."""
    for _ in range(nsteps):
        prob.timestepper.step(prob.sol, prob.clock, prob.eqn, prob.vars, prob.params, prob.grid)
        prob.clock.t += prob.clock.dt
        prob.clock.step += 1
        if diags is not None:
            for d in diags:
                d.increment(prob)


class Diagnostic:
    """Minimal [FF] ``Diagnostic``: value at construction, then one per step."""

    def __init__(self, fn, prob):
        self.fn = fn
        self.data = [fn(prob)]
        self.t = [prob.clock.t]

    def increment(self, prob):
        self.data.append(self.fn(prob))
        self.t.append(prob.clock.t)
