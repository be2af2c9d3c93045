"""What the mathematics fixes of the dealias band ([FF] dealias!, called at
src/twodnavierstokes.jl:287 and in every calcN!).

The reference documents the rule only as "zeroing out the highest-aliased_fraction
wavenumber components" (docs/src/aliasing.md:8-10).  An aliasing argument pins part of it:
on every mode that survives dealias!, the pseudospectral N of a quadratic nonlinearity must
equal the exact convolution of the retained modes (tests/alias_free.py).  These tests check
that for the oracle, and that a band one index too narrow at the top of the l axis (which
lets a pair of retained modes alias onto a retained one at n = 16 and 64 with f = 1/3) is
caught.  A band one index too narrow at the bottom of either axis keeps the product
alias-free at every power-of-two n (3 floor(n/3) < n), so the lower band edge (682 at 2048)
stays parity unpinned, recorded from the recalled FourierFlows rule.
"""
import numpy as np
import pytest

from oracle import twodnavierstokes as o2d
from tests.alias_free import exact_calcN


def _kept_error(n, f, mutate=None, seed=0):
    r = o2d.Problem(nx=n, stepper="ETDRK4", aliased_fraction=f)
    g = r.grid
    if mutate:
        mutate(g)
    s = np.fft.rfft2(np.random.default_rng(seed + n).standard_normal((n, n)))
    N = np.zeros_like(s)
    o2d.calcN(N, s, 0.0, r.clock, r.vars, r.params, g)
    E = exact_calcN(s, n, n)
    m = g.dealias_mask()
    return float(np.linalg.norm((N - E) * m) / np.linalg.norm(E * m))


@pytest.mark.parametrize("n", [16, 32, 64])
@pytest.mark.parametrize("f", [1 / 3, 1 / 2])
def test_retained_modes_are_alias_free(n, f):
    assert _kept_error(n, f) < 1e-13


@pytest.mark.parametrize("n", [16, 64])
def test_band_one_row_too_narrow_is_caught(n):
    # planted mistake: the l band ends one index early (l = -floor(n/3) - 1 retained)
    def narrow(g):
        g.lalias = range(g.lalias.start, g.lalias.stop - 1)
    assert _kept_error(n, 1 / 3, narrow) > 1e-3


def test_no_dealiasing_aliases():
    # aliased_fraction = 0 keeps every mode: the products alias and the check sees it
    assert _kept_error(16, 0.0) > 1e-3
