"""Grid description shared by the model modules ([FF] ``TwoDGrid``)."""
from .problem import Grid


def make_grid(nx, ny, Lx, Ly, aliased_fraction, T):
    return Grid(int(nx), int(ny), Lx, Ly, aliased_fraction, T)
