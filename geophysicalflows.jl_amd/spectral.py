"""Host-side Parseval sums over rfft-layout arrays ([FF] ``parsevalsum``/``parsevalsum2``)."""
import numpy as np


def parsevalsum(uh, grid):
    U = np.sum(uh[..., :, 0]) + 2 * np.sum(uh[..., :, 1:])
    return grid.Lx * grid.Ly / (grid.nx ** 2 * grid.ny ** 2) * np.real(U)


def parsevalsum2(uh, grid):
    U = np.sum(np.abs(uh[..., :, 0]) ** 2) + 2 * np.sum(np.abs(uh[..., :, 1:]) ** 2)
    return grid.Lx * grid.Ly / (grid.nx ** 2 * grid.ny ** 2) * U
