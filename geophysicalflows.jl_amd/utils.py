"""Initial-condition utilities of the reference (``src/utils.jl``).

These are host-side setup helpers, not part of the device hot path.
"""
import numpy as np


def peakedisotropicspectrum(grid, kpeak, E0, mask=None, allones=False, rng=None):
    """Random field with an isotropic spectrum peaked at ``kpeak`` (``src/utils.jl:29-58``).

    ``|psih| = (K^2 (1 + (K/k0)^4))^(-1/2)`` with ``k0 = kpeak 2 pi/Lx`` and ``psih[0, 0] = 0``,
    times complex normal phases (unit variance; each part variance 1/2 as Julia's
    ``randn(Complex{T})``) unless ``allones``, times ``mask``, scaled so that
    ``sum(Krsq |psih|^2)/(nx ny)^2 = E0``.  Returns ``irfft(-Krsq psih)`` as an array of
    shape ``(ny, nx)``.  The phases come from NumPy's generator ``rng`` (Julia's random
    stream cannot be reproduced), so fields agree with the reference in distribution only.
    """
    if grid.Lx != grid.Ly:
        raise ValueError("the domain is not square")
    k0 = kpeak * 2 * np.pi / grid.Lx
    K2 = grid.Krsq
    with np.errstate(divide="ignore"):
        modpsi = (K2 * (1 + (np.sqrt(K2) / k0) ** 4)) ** -0.5
    modpsi[0, 0] = 0.0
    if allones:
        psih = modpsi.astype(np.complex128)
    else:
        rng = np.random.default_rng() if rng is None else rng
        phases = (rng.standard_normal(K2.shape) + 1j * rng.standard_normal(K2.shape)) / np.sqrt(2)
        psih = phases * modpsi
    if mask is not None:
        psih = psih * mask
    e = np.sum(K2 * np.abs(psih) ** 2) / (grid.nx * grid.ny) ** 2
    psih = psih * np.sqrt(E0 / e)
    return np.fft.irfft2(-K2 * psih, s=(grid.ny, grid.nx))


def gaussian_bump(grid, amplitude, width, centre):
    """Periodic Gaussian ``amplitude exp(-r^2 / (2 width^2))`` about ``centre``, summed over the
    nearest periodic images, on the grid points (shape ``(ny, nx)``)."""
    X, Y = grid.gridpoints()
    cx, cy = centre
    out = np.zeros_like(X)
    for ix in (-1, 0, 1):
        for iy in (-1, 0, 1):
            out += np.exp(-((X - cx - ix * grid.Lx) ** 2 + (Y - cy - iy * grid.Ly) ** 2) / (2 * width ** 2))
    return amplitude * out
