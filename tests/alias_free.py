"""Exact (alias-free) TwoDNavierStokes advection term by direct convolution.

Test helper, independent of the oracle's transforms.  For a state sol on the rfft half
plane (shape (ny, nx//2 + 1), L = 2 pi so wavenumbers are integers), the physical
fields that irfft2 forms are band-limited trigonometric polynomials.  The product of two
of them has the exact spectrum

    (f g)^(k) = (1/N) sum_{p + q = k} f^(p) g^(q),      N = nx ny,

with p, q, k integer wave vectors (no wrap-around).  The pseudospectral product computes the
same sum with p + q = k (mod n); the two agree on a mode k exactly when no pair of
retained modes aliases onto k.  calcN (src/twodnavierstokes.jl:253-275) is

    N = -i kr (u zeta)^ - i l (v zeta)^,   u^ = i l zeta^ / |k|^2,  v^ = -i kr zeta^ / |k|^2.

``exact_calcN`` evaluates it by the direct pair sum, for comparison on the retained modes.
"""
import numpy as np


def full_plane(sol, nx):
    """Full-plane unnormalised spectrum of the real field irfft2(sol): fft2(irfft2(sol)).
    This is the Hermitian part of sol (irfft2 ignores the anti-Hermitian part of the
    kr = 0 column), i.e. the spectrum of the field the pseudospectral product sees."""
    return np.fft.fft2(np.fft.irfft2(sol, s=(sol.shape[0], nx)))


def exact_calcN(sol, nx, ny):
    F = full_plane(sol, nx)
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    KX, KY = np.meshgrid(kx, ky)
    K2 = KX ** 2 + KY ** 2
    inv = np.where(K2 == 0, 0.0, 1.0 / np.where(K2 == 0, 1.0, K2))
    U = 1j * KY * inv * F
    V = -1j * KX * inv * F
    nz = np.abs(F) > np.abs(F).max() * 1e-14   # the zeroed band comes back as rounding noise
    px, py = KX[nz].astype(int), KY[nz].astype(int)
    fz, fu, fv = F[nz], U[nz], V[nz]
    # every pair (p, q) of nonzero modes; exact integer sums p + q
    sx = (px[:, None] + px[None, :]).ravel()
    sy = (py[:, None] + py[None, :]).ravel()
    W = 2 * max(nx, ny)
    idx = (sx + W) * (2 * W + 1) + (sy + W)
    A = np.zeros((2 * W + 1) * (2 * W + 1), complex)
    B = np.zeros_like(A)
    np.add.at(A, idx, (fu[:, None] * fz[None, :]).ravel())
    np.add.at(B, idx, (fv[:, None] * fz[None, :]).ravel())
    A /= nx * ny
    B /= nx * ny
    # N on the rfft half plane: kr = 0..nx/2, l in FFT order
    nkr = nx // 2 + 1
    kr = np.arange(nkr)
    KR, L = np.meshgrid(kr, ky.astype(int))
    j = (KR + W) * (2 * W + 1) + (L + W)
    return -1j * KR * A[j] - 1j * L * B[j]
