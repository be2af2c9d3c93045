"""Oracle forcing: the counter-based random numbers and the ring spectrum.

The reference's stochastic forcing draws from Julia's global RNG inside a user
closure (``test/test_twodnavierstokes.jl:45-53``,
``examples/twodnavierstokes_stochasticforcing.jl:70-82``); those streams cannot
be reproduced outside Julia.  This restatement therefore fixes one generator that
the oracle and the HIP kernels both implement bit-for-bit:

* Philox-4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
  1, 2, 3"), key ``(seed_lo, seed_hi)``, counter ``(index, instance, step_lo, step_hi)``
  where ``index`` is the spectral-array offset ``l*nkr + kr`` (Julia linear index
  minus one);
* two uniform doubles from the four 32-bit words, ``u = (w >> 11 + 1) 2^-53``;
* a complex normal with variance 1/2 per part (what ``randn!`` gives a complex
  array in ``examples/twodnavierstokes_stochasticforcing.jl:79``) by Box--Muller:
  ``sqrt(-log u1) * (cos 2 pi u2 + i sin 2 pi u2)``.

The forcing itself follows ``examples/twodnavierstokes_stochasticforcing.jl:53-82``:
``Fh = xi * sqrt(spectrum) / sqrt(dt)``, regenerated once per time step (when
``t == clock.t``, ``src/twodnavierstokes.jl:311-319``).
"""
import numpy as np

from .fourierflows import parsevalsum

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox-4x32 on uint64 arrays holding 32-bit words."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) & _MASK for c in (c0, c1, c2, c3))
    k0 = np.uint64(k0) & _MASK
    k1 = np.uint64(k1) & _MASK
    for _ in range(rounds):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & _MASK, lo1, (hi0 ^ c3 ^ k1) & _MASK, lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def complex_normal(seed, step, index, instance=0):
    """Complex N(0, 1/2 + 1/2) sample for each spectral offset ``index``."""
    index = np.asarray(index, dtype=np.uint64)
    seed = int(seed)
    step = int(step)
    w0, w1, w2, w3 = philox4x32(index, np.full_like(index, instance),
                                np.full_like(index, step & 0xFFFFFFFF),
                                np.full_like(index, (step >> 32) & 0xFFFFFFFF),
                                seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    a = (w0 << np.uint64(32)) | w1
    b = (w2 << np.uint64(32)) | w3
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = ((b >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    r = np.sqrt(-np.log(u1))
    th = 2.0 * np.pi * u2
    return r * np.cos(th) + 1j * (r * np.sin(th))


def ring_spectrum(grid, kf, dkf, eps):
    """Forcing covariance spectrum of ``examples/twodnavierstokes_stochasticforcing.jl:53-66``:
    ``exp(-(K-kf)^2/(2 dkf^2))``, zero at K=0, normalised so that
    ``parsevalsum(spec * invKrsq / 2) / (Lx Ly) = eps``."""
    K = np.sqrt(grid.Krsq)
    spec = np.exp(-(K - kf) ** 2 / (2 * dkf ** 2))
    spec[grid.Krsq == 0] = 0.0
    eps0 = parsevalsum(spec * grid.invKrsq / 2, grid) / (grid.Lx * grid.Ly)
    return spec * (eps / eps0)


def ring_forcing(grid, sqrt_spec, seed, step, dt, instance=0):
    """One realisation ``Fh = xi * sqrt(spec) / sqrt(dt)`` on the rfft grid."""
    idx = np.arange(grid.nl * grid.nkr, dtype=np.uint64).reshape(grid.nl, grid.nkr)
    xi = complex_normal(seed, step, idx, instance)
    return xi * sqrt_spec / np.sqrt(dt)
