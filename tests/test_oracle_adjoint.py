"""Pins of the oracle's discrete adjoint (oracle/adjoint.py) on the CPU.

The adjoint has no reference counterpart.  What fixes it: central finite differences of
the oracle's forward step (itself pinned against the reference's tests), the dot-product
identity between tangent and adjoint, and the linear closed form.  A planted error (a
dropped Hermitian weight) must fail the dot-product test.
"""
import numpy as np
import pytest

from oracle import adjoint as oad
from oracle import fourierflows as off
from oracle import twodnavierstokes as o2d


def _grid(n=32):
    return off.TwoDGrid(n, 2 * np.pi)


def _rand_state(g, rng, scale=1.0):
    s = g.rfft(rng.standard_normal((g.ny, g.nx)))
    return scale * g.dealias_mask() * s


def _rand_cplx(g, rng):
    return rng.standard_normal((g.nl, g.nkr)) + 1j * rng.standard_normal((g.nl, g.nkr))


def test_calcN_matches_oracle_calcN():
    g = _grid()
    rng = np.random.default_rng(0)
    s = _rand_state(g, rng)
    p = o2d.Problem(nx=32, dt=0.01, stepper="ETDRK4")
    N = np.zeros_like(s)
    s2 = s.copy()
    o2d.calcN(N, s2, 0.0, p.clock, p.vars, p.params, p.grid)
    assert np.allclose(oad.calcN(s, g), N, rtol=0, atol=1e-12 * np.abs(N).max())


def test_calcN_tangent_is_the_derivative():
    g = _grid()
    rng = np.random.default_rng(1)
    s, ds = _rand_state(g, rng), _rand_state(g, rng)
    h = 1e-6
    fd = (oad.calcN(s + h * ds, g) - oad.calcN(s - h * ds, g)) / (2 * h)
    tl = oad.calcN_tlm(ds, s, g)
    assert np.linalg.norm(fd - tl) / np.linalg.norm(tl) < 1e-8


@pytest.mark.parametrize("n", [16, 32, 64])
def test_calcN_adjoint_dot_product(n):
    g = _grid(n)
    rng = np.random.default_rng(2)
    s, v = _rand_state(g, rng), _rand_state(g, rng)
    w = _rand_cplx(g, rng)
    lhs = oad.inner(w, oad.calcN_tlm(v, s, g))
    rhs = oad.inner(oad.calcN_adj(w, s, g), v)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_dropped_hermitian_weight_fails_dot_product():
    # the planted mistake: unit weights instead of w (kr > 0 columns counted once)
    g = _grid(32)
    rng = np.random.default_rng(3)
    s, v = _rand_state(g, rng), _rand_state(g, rng)
    w = _rand_cplx(g, rng)
    lhs = oad.inner(w, oad.calcN_tlm(v, s, g))
    rhs = oad.inner(oad.calcN_adj(w, s, g, w=np.ones((g.nl, g.nkr))), v)
    assert abs(lhs - rhs) > 1e-3 * abs(lhs)


def _step(n=32, nu=1e-3):
    g = _grid(n)
    L, dLdnu = oad.twodns_L(g, nu, 1)
    return g, oad.ETDStep(g, L, 0.005, dLdnu)


def test_step_forward_equals_oracle_stepper():
    g, st = _step()
    rng = np.random.default_rng(4)
    sol = g.rfft(rng.standard_normal((g.ny, g.nx)))
    p = o2d.Problem(nx=32, nu=1e-3, nnu=1, dt=0.005, stepper="ETDRK4")
    p.sol[...] = sol
    off.stepforward(p, 1)
    new, _ = st.forward(sol)
    assert np.linalg.norm(new - p.sol) <= 1e-13 * np.linalg.norm(p.sol)


def test_step_tangent_and_adjoint():
    g, st = _step()
    rng = np.random.default_rng(5)
    sol = _rand_state(g, rng, 0.05)
    d = _rand_cplx(g, rng)
    w = _rand_cplx(g, rng)
    _, stages = st.forward(sol)
    h = 1e-6
    fd = (st.forward(sol + h * d)[0] - st.forward(sol - h * d)[0]) / (2 * h)
    tl = st.tlm(d, stages)
    assert np.linalg.norm(fd - tl) / np.linalg.norm(tl) < 1e-7
    lhs = oad.inner(w, tl)
    rhs = oad.inner(st.adjoint(w, stages), d)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_linear_step_adjoint_closed_form():
    # at the zero state N and its derivative vanish: step = E M sol, adjoint = M E lambda
    g, st = _step()
    rng = np.random.default_rng(6)
    lam = _rand_cplx(g, rng)
    _, stages = st.forward(np.zeros((g.nl, g.nkr), complex))
    got = st.adjoint(lam, stages)
    assert np.allclose(got, st.M * st.E * lam, rtol=0, atol=1e-15)


def test_loss_gradient_matches_central_differences():
    g = _grid(16)
    rng = np.random.default_rng(7)
    z0 = 0.3 * rng.standard_normal((16, 16))
    target = 0.3 * rng.standard_normal((16, 16))
    args = (g, 1e-2, 1, 0.01, 12)
    J, grad, dnu = oad.loss_and_gradient(z0, target, *args, checkpoint=5)
    h = 1e-6
    for k in range(4):
        v = rng.standard_normal((16, 16))
        Jp = oad.loss_and_gradient(z0 + h * v, target, *args, checkpoint=5)[0]
        Jm = oad.loss_and_gradient(z0 - h * v, target, *args, checkpoint=5)[0]
        assert (Jp - Jm) / (2 * h) == pytest.approx(float(np.sum(grad * v)), rel=1e-7)
    # dJ/dnu against central differences in nu (relative step 1e-4)
    hn = 1e-6
    Jp = oad.loss_and_gradient(z0, target, g, 1e-2 + hn, 1, 0.01, 12, checkpoint=5)[0]
    Jm = oad.loss_and_gradient(z0, target, g, 1e-2 - hn, 1, 0.01, 12, checkpoint=5)[0]
    assert (Jp - Jm) / (2 * hn) == pytest.approx(dnu, rel=1e-6)


@pytest.mark.parametrize("dtL", [-0.5, -1.0, -3.0, -20.0, -400.0])
def test_etd_coefficient_derivatives(dtL):
    # d/dL of the contour-mean coefficients against central differences of the coefficients
    dt = 0.01
    L = np.array([[dtL / dt]])
    d = oad.etd_coefficient_derivatives(dt, L)
    h = 1e-5 * abs(L[0, 0])
    cp = off.etd_coefficients(dt, L + h)
    cm = off.etd_coefficients(dt, L - h)
    for k in range(4):
        fd = (cp[k] - cm[k]) / (2 * h)
        assert d[2 + k][0, 0] == pytest.approx(fd[0, 0], rel=1e-6, abs=1e-12 * dt * dt)
    assert d[0][0, 0] == pytest.approx(dt * np.exp(dt * L[0, 0]))
