"""MultiLayerQG on the device beyond 2-4 layers: one layer (SingleLayerParams), 5 to 8 layers,
set_psi! (S psih per wavenumber) and the eddy fluxes, against the oracle and the reference's
literal values (test/test_multilayerqg.jl:502-655, 809-838).

Tolerances (FP64): 1e-13 relative for one calcN and for set_psi!; 1e-12 relative L2 on the
state after 3 steps; the reference's own rtol 1e-6 / atol 1e-7 on its printed flux values, and
1e-12 relative (1e-15 absolute where the value is zero) against the oracle's fluxes.
"""
import math

import numpy as np
import pytest

from oracle import fourierflows as off
from oracle import multilayerqg as oml
from tests.conftest import rel

pytestmark = pytest.mark.gpu
TWO_PI = 2 * math.pi


def rand_spec(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _layers(n):
    H = list(np.linspace(1.0, 2.0, n) / np.sum(np.linspace(1.0, 2.0, n)))
    b = list(-1.0 - 0.1 * np.arange(n) - 0.01 * np.arange(n) ** 2)
    return H, b


@pytest.mark.parametrize("nlayers,n", [(1, 64), (5, 64), (8, 64), (5, 512)])
def test_mlqg_layer_counts_calcN_steps_updatevars(gf, nlayers, n):
    H, b = _layers(nlayers)
    g = off.TwoDGrid(n, TWO_PI)
    X, Y = g.gridpoints()
    eta = 0.3 * np.exp(-((X + 1) ** 2 + (Y - 0.5) ** 2) / 0.3)
    Uy = np.stack([0.4 * np.cos(g.y) * (nlayers - j) / nlayers for j in range(nlayers)], axis=1)   # (ny, n)
    s = rand_spec(np.random.default_rng(100 + nlayers), (nlayers, n, n // 2 + 1)) * 0.05
    kw = dict(nx=n, Lx=TWO_PI, H=H, b=b, U=Uy, eta=eta, dt=0.005, stepper="ETDRK4",
              topographic_pv_gradient=(0.01, 0.1))
    p = gf.MultiLayerQG.Problem(nlayers, gf.GPU(), f0=1.0, β=5.0, μ=0.01, ν=1e-3, **kw)
    r = oml.Problem(nlayers, f0=1.0, beta=5.0, mu=0.01, nu=1e-3, **kw)
    shape = r.sol.shape
    N, _ = p.calcN(s.reshape(p._sshape))
    Nr, sr = np.zeros_like(s), s.copy()
    oml.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N.reshape(shape), Nr) < 1e-13
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 3)
    off.stepforward(r, 3)
    assert rel(p.sol.reshape(shape), r.sol) < 1e-12
    gf.MultiLayerQG.updatevars(p)
    oml.updatevars(r)
    for a, bb in ((p.vars.q, r.vars.q), (p.vars.ψ, r.vars.psi), (p.vars.u, r.vars.u), (p.vars.v, r.vars.v)):
        assert rel(np.reshape(a, bb.shape), bb) < 1e-13


@pytest.mark.parametrize("stepper", ["RK4", "ForwardEuler", "FilteredETDRK4"])
def test_mlqg_single_layer_other_steppers_and_linear(gf, stepper):
    n = 64
    g = off.TwoDGrid(n, TWO_PI)
    Uy = 0.3 * np.sin(2 * g.y)
    s = rand_spec(np.random.default_rng(7), (1, n, n // 2 + 1)) * 0.05
    for linear in (False, True):
        kw = dict(nx=n, Lx=TWO_PI, U=Uy, dt=0.004, stepper=stepper, linear=linear)
        p = gf.MultiLayerQG.Problem(1, gf.GPU(), β=3.0, μ=0.05, ν=1e-3, **kw)
        r = oml.Problem(1, beta=3.0, mu=0.05, nu=1e-3, **kw)
        p.sol = s
        r.sol[...] = s
        gf.stepforward(p, 3)
        off.stepforward(r, 3)
        assert rel(p.sol.reshape(r.sol.shape), r.sol) < 1e-12, linear


def test_mlqg_single_layer_rossby_wave(gf):
    # test/test_multilayerqg.jl:809-838: one layer, RK4, dt = 0.01, 20 steps
    n, beta, U = 64, 2.0, 0.5
    p = gf.MultiLayerQG.Problem(1, gf.GPU(), nx=n, Lx=TWO_PI, U=U, β=beta, stepper="RK4", dt=1e-2)
    x, y = p.grid.gridpoints()
    ampl, kw, lw = 1e-2, 3.0, 2.0
    omega = kw * (U - beta / (kw ** 2 + lw ** 2))
    gf.MultiLayerQG.set_q(p, ampl * np.cos(kw * x) * np.cos(lw * y))
    gf.stepforward(p, 20)
    gf.MultiLayerQG.updatevars(p)
    qth = ampl * np.cos(kw * (x - omega / kw * p.clock.t)) * np.cos(lw * y)
    assert rel(p.vars.q, qth) < n * n * 20 * 1e-12


@pytest.mark.parametrize("nlayers", [1, 2, 3, 5])
def test_mlqg_set_psi_matches_oracle(gf, nlayers):
    n = 64
    H, b = _layers(nlayers)
    rng = np.random.default_rng(200 + nlayers)
    psi = rng.standard_normal((nlayers, n, n))
    p = gf.MultiLayerQG.Problem(nlayers, gf.GPU(), nx=n, Lx=TWO_PI, f0=1.0, H=H, b=b, dt=0.01, stepper="RK4")
    r = oml.Problem(nlayers, nx=n, Lx=TWO_PI, f0=1.0, H=H, b=b)
    gf.MultiLayerQG.set_ψ(p, psi)
    oml.set_psi(r, psi)
    assert rel(p.sol.reshape(r.sol.shape), r.sol) < 1e-13
    assert rel(np.reshape(p.vars.ψ, r.vars.psi.shape), r.vars.psi) < 1e-13
    assert rel(np.reshape(p.vars.q, r.vars.q.shape), r.vars.q) < 1e-13


def test_mlqg_energies_through_device_set_psi(gf):
    # test/test_multilayerqg.jl:502-536 entering through set_psi! (restated on 64 x 64)
    n = 64
    p = gf.MultiLayerQG.Problem(2, gf.GPU(), nx=n, Lx=TWO_PI, f0=1, H=[2.5, 7.5], b=[-1.0, -1.5], dt=0.01,
                                stepper="RK4")
    x, y = p.grid.gridpoints()
    gf.MultiLayerQG.set_ψ(p, np.array([np.cos(2 * x) * np.cos(2 * y), 0.5 * np.cos(2 * x) * np.cos(2 * y)]))
    KE, PE = gf.MultiLayerQG.energies(p)
    assert abs(KE[0] - 1 / 4) < 1e-13 and abs(KE[1] - 3 / 16) < 1e-13 and abs(PE[0] - 1 / 160) < 1e-13


def test_mlqg_single_layer_energy(gf):
    # test/test_multilayerqg.jl:538-560: one layer, Ly = 3 pi, energy 29/9
    nx = ny = 64
    p = gf.MultiLayerQG.Problem(1, gf.GPU(), nx=nx, Lx=TWO_PI, ny=ny, Ly=3 * math.pi, U=np.zeros(ny),
                                stepper="ForwardEuler", dt=0.01)
    x, y = p.grid.gridpoints()
    k0, l0 = 1.0, 2 / 3
    q0 = (-((2 * k0) ** 2 + (2 * l0) ** 2) * np.sin(2 * k0 * x) * np.cos(2 * l0 * y)
          - (k0 ** 2 + (3 * l0) ** 2) * 2 * np.sin(k0 * x) * np.cos(3 * l0 * y))
    gf.MultiLayerQG.set_q(p, q0)
    e = gf.MultiLayerQG.energies(p)
    assert abs(e - 29 / 9) < 1e-13 * 29 / 9


@pytest.mark.parametrize("nlayers", [2, 3])
def test_mlqg_fluxes_reference_values_and_oracle(gf, nlayers):
    # test/test_multilayerqg.jl:567-625 restated on 128 x 128 (the reference's 128 x 126 gives
    # the same values: the sums are trapezoid rules of an analytic periodic integrand)
    nx = ny = 128
    if nlayers == 2:
        H, b = [0.2, 0.8], [-1.0, -1.2]
    else:
        H, b = [0.2, 0.35, 0.45], [-1.0, -1.2, -1.2 + 1e-9]
    g = off.TwoDGrid(nx, TWO_PI)
    U = np.zeros((ny, nlayers))
    U[:, 0] = 1 / np.cosh(g.y / 0.2) ** 2
    kw = dict(nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, f0=1, H=H, b=b, U=U)
    p = gf.MultiLayerQG.Problem(nlayers, gf.GPU(), dt=0.01, stepper="RK4", **kw)
    r = oml.Problem(nlayers, **kw)
    x, y = r.grid.gridpoints()
    psi2 = np.cos(x + np.pi / 10) * np.cos(y)
    psi = np.array([np.cos(x) * np.cos(y) + np.sin(x), psi2] + ([psi2] if nlayers == 3 else []))
    gf.MultiLayerQG.set_ψ(p, psi)
    oml.set_psi(r, psi)
    lat, ver = gf.MultiLayerQG.fluxes(p)
    latr, verr = oml.fluxes(r)
    assert abs(lat[0] - 0.00626267) <= 1e-6 * 0.00626267
    assert abs(lat[1]) <= 1e-12
    assert abs(ver[0] + 0.196539) <= 1e-6 * 0.196539
    if nlayers == 3:
        assert abs(lat[2]) <= 1e-12 and abs(ver[1]) <= 1e-12
    # three layers: g'_{5/2} = -1e-9 couples layers 2 and 3 with F ~ 1e9, so S^-1 has a condition
    # number ~1e9 and two correct inversions (the host's Gauss-Jordan, NumPy's LU) agree to ~1e-7
    tol = 1e-12 if nlayers == 2 else 1e-6
    assert abs(lat[0] - latr[0]) <= tol * abs(latr[0]) and abs(ver[0] - verr[0]) <= tol * abs(verr[0])
    assert np.all(np.abs(lat[1:] - latr[1:]) <= 1e-15) and np.all(np.abs(ver[1:] - verr[1:]) <= 1e-15)


def test_mlqg_fluxes_single_layer(gf):
    # test/test_multilayerqg.jl:632-655 restated on 128 x 128: 0.0313134 (atol 1e-7)
    n = 128
    g = off.TwoDGrid(n, TWO_PI)
    U = 1 / np.cosh(g.y / 0.2) ** 2
    p = gf.MultiLayerQG.Problem(1, gf.GPU(), nx=n, Lx=TWO_PI, U=U, dt=0.01, stepper="RK4")
    r = oml.Problem(1, nx=n, Lx=TWO_PI, U=U)
    x, y = r.grid.gridpoints()
    psi = np.cos(x) * np.cos(y) + np.sin(x)
    gf.MultiLayerQG.set_ψ(p, psi)
    oml.set_psi(r, psi[None])
    f = gf.MultiLayerQG.fluxes(p)
    assert abs(f - 0.0313134) <= 1e-7
    assert abs(f - oml.fluxes(r)) <= 1e-12 * abs(f)


@pytest.mark.parametrize("nlayers,n", [(4, 64), (2, 256)])
def test_mlqg_fluxes_random_state_match_oracle(gf, nlayers, n):
    # a raw random state (the fluxes dealias it in place first) and U(y) in every layer
    H, b = _layers(nlayers)
    g = off.TwoDGrid(n, TWO_PI)
    U = np.stack([np.cos(g.y + j) * 0.5 for j in range(nlayers)], axis=1)
    kw = dict(nx=n, Lx=TWO_PI, f0=1.3, H=H, b=b, U=U)
    p = gf.MultiLayerQG.Problem(nlayers, gf.GPU(), dt=0.01, stepper="RK4", **kw)
    r = oml.Problem(nlayers, **kw)
    s = rand_spec(np.random.default_rng(300 + n), (nlayers, n, n // 2 + 1))
    p.sol = s
    r.sol[...] = s
    lat, ver = gf.MultiLayerQG.fluxes(p)
    latr, verr = oml.fluxes(r)
    assert rel(lat, latr) < 1e-12 and rel(ver, verr) < 1e-12
    assert rel(p.sol, r.sol) == 0.0     # both dealiased sol in place
