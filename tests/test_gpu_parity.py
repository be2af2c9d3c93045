"""GPU parity tests: the HIP library through the C ABI against the CPU oracle,
the reference's closed forms, and size-independent invariants.

Tolerances (FP64): 1e-13 relative for one calcN and for the reference's
manufactured steady solutions (the reference's own rtol, test/runtests.jl:22-24);
1e-12 relative L2 on the state after a few steps against the oracle; FP32
states within 2e-5 of the FP64 oracle after 10 steps.
"""
import math

import numpy as np
import pytest

from oracle import fourierflows as off
from oracle import multilayerqg as oml
from oracle import singlelayerqg as osl
from oracle import twodnavierstokes as o2d
from tests.conftest import rel

pytestmark = pytest.mark.gpu
TWO_PI = 2 * math.pi


def rand_spec(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# ----------------------------------------------------------------------------- TwoDNS
@pytest.mark.parametrize("nx,ny,Ly", [(64, 64, TWO_PI), (128, 64, 3 * math.pi), (32, 128, 1.7)])
def test_twodns_set_zeta_updatevars(gf, nx, ny, Ly):
    rng = np.random.default_rng(nx + ny)
    z0 = rng.standard_normal((ny, nx))
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=nx, ny=ny, Lx=TWO_PI, Ly=Ly, stepper="ETDRK4")
    gf.TwoDNavierStokes.set_ζ(p, z0)
    r = o2d.Problem(nx=nx, ny=ny, Lx=TWO_PI, Ly=Ly, stepper="ETDRK4")
    o2d.set_zeta(r, z0)
    assert rel(p.sol, r.sol) < 1e-14
    assert rel(p.vars.ζ, r.vars.zeta) < 1e-13
    assert rel(p.vars.u, r.vars.u) < 1e-13 and rel(p.vars.v, r.vars.v) < 1e-13


def test_twodns_calcN_and_in_place_dealias(gf):
    n = 128
    rng = np.random.default_rng(2)
    s = rand_spec(rng, (n, n // 2 + 1))
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, stepper="ETDRK4")
    N, sm = p.calcN(s)
    r = o2d.Problem(nx=n, stepper="ETDRK4")
    Nr, sr = np.zeros_like(s), s.copy()
    o2d.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < 1e-13
    assert np.array_equal(sm, sr)  # P1: the input is dealiased in place, bit-for-bit


@pytest.mark.parametrize("stepper", ["ETDRK4", "FilteredETDRK4", "RK4", "FilteredRK4", "ForwardEuler",
                                     "FilteredForwardEuler"])
def test_twodns_steppers_match_oracle(gf, stepper):
    n = 64
    rng = np.random.default_rng(3)
    s0 = rand_spec(rng, (n, n // 2 + 1)) * 0.1
    s0[0, 0] = 0
    kw = dict(nx=n, Lx=TWO_PI, dt=0.005, stepper=stepper)
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), ν=1e-3, nν=2, μ=0.05, nμ=0, **kw)
    r = o2d.Problem(nu=1e-3, nnu=2, mu=0.05, nmu=0, **kw)
    p.sol = s0
    r.sol[...] = s0
    gf.stepforward(p, 5)
    off.stepforward(r, 5)
    assert rel(p.sol, r.sol) < 1e-12
    assert p.clock.step == 5 and abs(p.clock.t - 0.025) < 1e-15


def test_twodns_ring_forcing_matches_oracle(gf):
    n = 128
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-7, nν=2, μ=0.1, nμ=0, dt=0.005, stepper="ETDRK4",
                                    calcF=gf.RingForcing(12.0, 2.0, 0.1, seed=1234))
    r = o2d.Problem(nx=n, nu=1e-7, nnu=2, mu=0.1, nmu=0, dt=0.005, stepper="ETDRK4",
                    calcF=("ring", 12.0, 2.0, 0.1, 1234))
    gf.stepforward(p, 6)
    off.stepforward(r, 6)
    assert rel(p.sol, r.sol) < 1e-12


def _twodns_manufactured(gf, stepper, dt, nt, n=128):
    g = off.TwoDGrid(n, TWO_PI)
    x, y = g.gridpoints()
    nu = 1e-2
    zf = -8 * np.sin(2 * x) * np.cos(2 * y) - 20 * np.sin(x) * np.cos(3 * y)
    Ff = -(nu * (64 * np.sin(2 * x) * np.cos(2 * y) + 200 * np.sin(x) * np.cos(3 * y))
           + 8 * (np.cos(x) * np.cos(3 * y) * np.sin(2 * x) * np.sin(2 * y)
                  - 3 * np.cos(2 * x) * np.cos(2 * y) * np.sin(x) * np.sin(3 * y)))
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, Lx=TWO_PI, ν=nu, nν=1, dt=dt, stepper=stepper, calcF=g.rfft(Ff))
    gf.TwoDNavierStokes.set_ζ(p, zf)
    gf.stepforward(p, nt)
    gf.TwoDNavierStokes.updatevars(p)
    return rel(p.vars.ζ, zf)


@pytest.mark.parametrize("stepper,dt,nt", [("ForwardEuler", 5e-4, 2000), ("ETDRK4", 5e-3, 200),
                                           ("FilteredRK4", 5e-3, 200)])
def test_twodns_reference_advection_test(gf, stepper, dt, nt):
    """test/test_twodnavierstokes.jl:221-257 at the reference's size and tolerance (1e-13)."""
    assert _twodns_manufactured(gf, stepper, dt, nt) < 1e-13


def test_twodns_closed_form_diagnostics(gf):
    """test/test_twodnavierstokes.jl:259-317 (128x128 grid, Lx = 2pi, Ly = 3pi)."""
    nx, ny, Lx, Ly = 128, 128, TWO_PI, 3 * math.pi
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=nx, ny=ny, Lx=Lx, Ly=Ly, stepper="ForwardEuler")
    x, y = p.grid.gridpoints()
    k0, l0 = TWO_PI / Lx, TWO_PI / Ly
    z0 = (-((2 * k0) ** 2 + (2 * l0) ** 2) * np.sin(2 * k0 * x) * np.cos(2 * l0 * y)
          - (k0 ** 2 + (3 * l0) ** 2) * 2 * np.sin(k0 * x) * np.cos(3 * l0 * y))
    gf.TwoDNavierStokes.set_ζ(p, z0)
    assert abs(gf.TwoDNavierStokes.energy(p) / (29 / 9) - 1) < 1e-13
    assert abs(gf.TwoDNavierStokes.enstrophy(p) / (2701 / 162) - 1) < 1e-13
    assert abs(gf.TwoDNavierStokes.palinstrophy(p) / (126277 / 1458) - 1) < 1e-13


def test_twodns_fp32_tolerance(gf):
    n = 128
    rng = np.random.default_rng(4)
    s0 = rand_spec(rng, (n, n // 2 + 1)) * 0.05
    s0[0, 0] = 0
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, nν=1, dt=0.005, stepper="ETDRK4", T=np.float32)
    r = o2d.Problem(nx=n, nu=1e-3, nnu=1, dt=0.005, stepper="ETDRK4")
    p.sol = s0
    r.sol[...] = p.sol.astype(np.complex128)
    gf.stepforward(p, 10)
    off.stepforward(r, 10)
    assert p.sol.dtype == np.complex64
    assert rel(p.sol, r.sol) < 2e-5


def test_state_roundtrip_and_lazy_band(gf):
    """set/get is bit-exact; stepping n steps in one call equals n calls of one step
    (the aliased band is produced by the last step of a call, P1)."""
    n = 64
    rng = np.random.default_rng(5)
    s0 = rand_spec(rng, (n, n // 2 + 1))
    a = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, nν=1, dt=0.005, stepper="ETDRK4")
    b = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, nν=1, dt=0.005, stepper="ETDRK4")
    a.sol = s0
    assert np.array_equal(a.sol, s0)
    b.sol = s0
    gf.stepforward(a, 4)
    for _ in range(4):
        gf.stepforward(b, 1)
    assert np.array_equal(a.sol, b.sol)
    gf.TwoDNavierStokes.updatevars(a)
    m = off.TwoDGrid(n, TWO_PI).dealias_mask()
    assert np.all(a.sol[m == 0] == 0)  # updatevars! dealiases sol in place


def test_set_forcing_error_path(gf):
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=32, stepper="RK4")
    with pytest.raises(gf._lib.GFError) as e:
        p.set_forcing(np.zeros((32, 17), complex))
    assert e.value.code == -1


# ----------------------------------------------------------------------------- SingleLayerQG
@pytest.mark.parametrize("ell,U,topo", [(math.inf, 0.0, False), (math.inf, 0.3, True), (1.23, 0.0, True),
                                        (1.23, 0.3, True)])
def test_slqg_calcN_and_steps(gf, ell, U, topo):
    n = 64
    rng = np.random.default_rng(6)
    g = off.TwoDGrid(n, TWO_PI)
    X, Y = g.gridpoints()
    eta = 0.4 * np.cos(3 * X) * np.cos(2 * Y) if topo else None
    kw = dict(nx=n, Lx=TWO_PI, U=U, eta=eta, dt=0.005, stepper="ETDRK4")
    p = gf.SingleLayerQG.Problem(gf.GPU(), β=2.0, deformation_radius=ell, ν=1e-3, μ=0.01, **kw)
    r = osl.Problem(beta=2.0, deformation_radius=ell, nu=1e-3, mu=0.01, **kw)
    s = rand_spec(rng, (n, n // 2 + 1)) * 0.1
    N, sm = p.calcN(s)
    Nr, sr = np.zeros_like(s), s.copy()
    osl.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < 1e-13
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 4)
    off.stepforward(r, 4)
    assert rel(p.sol, r.sol) < 1e-12


def test_slqg_varying_U_calcN(gf):
    n = 64
    g = off.TwoDGrid(n, TWO_PI)
    X, Y = g.gridpoints()
    eta = 0.4 * np.cos(3 * X) * np.cos(2 * Y)
    Uy = 0.2 / np.cosh(2 * g.y) ** 2
    rng = np.random.default_rng(7)
    s = rand_spec(rng, (n, n // 2 + 1)) * 0.1
    p = gf.SingleLayerQG.Problem(gf.GPU(), nx=n, β=2.0, U=Uy, eta=eta, ν=1e-3, dt=0.005, stepper="RK4")
    r = osl.Problem(nx=n, beta=2.0, U=Uy, eta=eta, nu=1e-3, dt=0.005, stepper="RK4")
    N, _ = p.calcN(s)
    Nr, sr = np.zeros_like(s), s.copy()
    osl.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < 1e-13


def _slqg_forcing(g, eta0, U, Uyy, ell=None):
    x, y = g.gridpoints()
    nu = 1e-2
    s2, c2 = np.sin, np.cos
    base = (-8 * (c2(x) * c2(3 * y) * s2(2 * x) * s2(2 * y) - 3 * c2(2 * x) * c2(2 * y) * s2(x) * s2(3 * y))
            - 20 * eta0 * (s2(10 * x) * c2(10 * y) * (s2(2 * x) * s2(2 * y) + 3 * s2(x) * s2(3 * y))
                           + c2(10 * x) * s2(10 * y) * (c2(2 * x) * c2(2 * y) + c2(x) * c2(3 * y))))
    visc = 64 * s2(2 * x) * c2(2 * y) + 200 * s2(x) * c2(3 * y)
    if ell is not None:
        visc = visc + (8 * s2(2 * x) * c2(2 * y) + 20 * s2(x) * c2(3 * y)) / ell ** 2
    F = -nu * visc + base
    F = F + (-U * (16 * c2(2 * x) * c2(2 * y) + 20 * c2(x) * c2(3 * y))
             - Uyy * (2 * c2(2 * x) * c2(2 * y) + 2 * c2(x) * c2(3 * y))
             - eta0 * U * 10 * s2(10 * x) * c2(10 * y))
    return g.rfft(F)


@pytest.mark.parametrize("topo,bg,vary,ell", [(False, False, False, None), (True, False, False, None),
                                              (False, True, False, None), (True, True, False, None),
                                              (False, True, True, None), (True, False, False, 1.23)])
def test_slqg_reference_nonlinear_advection(gf, topo, bg, vary, ell):
    """test/test_singlelayerqg.jl:276-399 at the reference's size (128^2, ForwardEuler,
    dt = 5e-4, 2000 steps) and tolerance 1e-13."""
    n = 128
    g = off.TwoDGrid(n, TWO_PI)
    x, y = g.gridpoints()
    eta0 = 0.4 if topo else 0.0
    eta = eta0 * np.cos(10 * x) * np.cos(10 * y)
    psif = np.sin(2 * x) * np.cos(2 * y) + 2 * np.sin(x) * np.cos(3 * y)
    qf = -8 * np.sin(2 * x) * np.cos(2 * y) - 20 * np.sin(x) * np.cos(3 * y)
    if ell is not None:
        qf = qf - psif / ell ** 2
    U, Uyy = 0.0, 0.0
    if bg and not vary:
        U = 0.2
    elif bg:
        U = 0.2 / np.cosh(2 * y) ** 2
        Uyy = g.irfft(-g.Lw ** 2 * g.rfft(U))
    Fh = _slqg_forcing(g, eta0, U, Uyy, ell)
    U0 = U[:, 0] if (bg and vary) else U
    p = gf.SingleLayerQG.Problem(gf.GPU(), nx=n, Lx=TWO_PI, U=U0, eta=eta, ν=1e-2, nν=1, dt=5e-4,
                                 stepper="ForwardEuler", calcF=Fh,
                                 deformation_radius=math.inf if ell is None else ell)
    gf.SingleLayerQG.set_q(p, qf)
    gf.stepforward(p, 2000)
    gf.SingleLayerQG.updatevars(p)
    assert rel(p.vars.q, qf) < 1e-13


# ----------------------------------------------------------------------------- MultiLayerQG
@pytest.mark.parametrize("nlayers,n", [(2, 64), (3, 64), (4, 64), (2, 512), (3, 512), (4, 512)])
def test_mlqg_calcN_steps_updatevars(gf, nlayers, n):
    # from 512^2 the multi-layer K3 runs one spectral column per workgroup
    H = {2: [0.2, 0.8], 3: [0.2, 0.3, 0.5], 4: [0.1, 0.2, 0.3, 0.4]}[nlayers]
    b = {2: [-1.0, -1.2], 3: [-1.0, -1.2, -1.3], 4: [-1.0, -1.1, -1.25, -1.4]}[nlayers]
    U = list(np.linspace(1.0, 0.0, nlayers))
    g = off.TwoDGrid(n, TWO_PI)
    X, Y = g.gridpoints()
    eta = 0.3 * np.exp(-((X + 1) ** 2 + (Y - 0.5) ** 2) / 0.3)
    rng = np.random.default_rng(nlayers)
    s = rand_spec(rng, (nlayers, n, n // 2 + 1)) * 0.05
    kw = dict(nx=n, Lx=TWO_PI, H=H, b=b, U=U, eta=eta, dt=0.005, stepper="ETDRK4")
    p = gf.MultiLayerQG.Problem(nlayers, gf.GPU(), f0=1.0, β=5.0, μ=0.005, ν=1e-3, **kw)
    r = oml.Problem(nlayers, f0=1.0, beta=5.0, mu=0.005, nu=1e-3, **kw)
    N, _ = p.calcN(s)
    Nr, sr = np.zeros_like(s), s.copy()
    oml.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < 1e-13
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 3)
    off.stepforward(r, 3)
    assert rel(p.sol, r.sol) < 1e-12
    gf.MultiLayerQG.updatevars(p)
    oml.updatevars(r)
    for a, bb in ((p.vars.q, r.vars.q), (p.vars.ψ, r.vars.psi), (p.vars.u, r.vars.u), (p.vars.v, r.vars.v)):
        assert rel(a, bb) < 1e-13


@pytest.mark.parametrize("stepper", ["AB3", "FilteredAB3"])
@pytest.mark.parametrize("model", ["twodns", "slqg"])
def test_ab3_matches_oracle(gf, stepper, model):
    # [FF] AB3 (test/runtests.jl:58): forward Euler start-up, then 3rd-order Adams-Bashforth;
    # 7 steps cross the start-up, split over calls to exercise the carried RHS history
    n = 64
    rng = np.random.default_rng(80)
    s = rand_spec(rng, (n, n // 2 + 1)) * 0.05
    if model == "twodns":
        p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, nν=1, μ=0.05, dt=0.002, stepper=stepper)
        r = o2d.Problem(nx=n, nu=1e-3, nnu=1, mu=0.05, dt=0.002, stepper=stepper)
    else:
        g = off.TwoDGrid(n, TWO_PI)
        X, Y = g.gridpoints()
        eta = 0.3 * np.cos(2 * X) * np.sin(Y)
        p = gf.SingleLayerQG.Problem(gf.GPU(), nx=n, β=2.0, U=0.2, eta=eta, ν=1e-3, dt=0.002, stepper=stepper)
        r = osl.Problem(nx=n, beta=2.0, U=0.2, eta=eta, nu=1e-3, dt=0.002, stepper=stepper)
    p.sol = s
    r.sol[...] = s
    for k in (2, 1, 4):
        gf.stepforward(p, k)
        off.stepforward(r, k)
        assert rel(p.sol, r.sol) < 1e-12, k
    assert p.clock.step == 7


@pytest.mark.parametrize("nlayers", [2, 3])
def test_mlqg_linear_calcN_and_steps(gf, nlayers):
    # calcNlinear! (src/multilayerqg.jl:675-683, 744-779): no dealiasing of the input
    n = 64
    H = {2: [0.2, 0.8], 3: [0.2, 0.3, 0.5]}[nlayers]
    b = {2: [-1.0, -1.2], 3: [-1.0, -1.2, -1.3]}[nlayers]
    g = off.TwoDGrid(n, TWO_PI)
    X, Y = g.gridpoints()
    eta = 0.3 * np.exp(-((X + 1) ** 2 + (Y - 0.5) ** 2) / 0.3)
    Uy = np.stack([0.3 * np.cos(g.y) * (j + 1) / nlayers for j in range(nlayers)], axis=1)   # (ny, n)
    s = rand_spec(np.random.default_rng(70 + nlayers), (nlayers, n, n // 2 + 1)) * 0.05
    kw = dict(nx=n, Lx=TWO_PI, H=H, b=b, U=Uy, eta=eta, dt=0.005, stepper="ETDRK4", linear=True)
    p = gf.MultiLayerQG.Problem(nlayers, gf.GPU(), f0=1.0, β=5.0, μ=0.01, ν=1e-3, **kw)
    r = oml.Problem(nlayers, f0=1.0, beta=5.0, mu=0.01, nu=1e-3, **kw)
    N, s_out = p.calcN(s)
    Nr, sr = np.zeros_like(s), s.copy()
    oml.calcNlinear(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < 1e-13
    assert np.array_equal(s_out, s)           # the input is not dealiased
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 3)
    off.stepforward(r, 3)
    assert rel(p.sol, r.sol) < 1e-12


def test_mlqg_linear_manufactured(gf):
    # test/test_multilayerqg.jl:413-494 restated on 64 x 64: the steady solution stays put
    from tests.test_oracle_qg import mlqg_manufactured_setup
    kw, Ff, qf = mlqg_manufactured_setup(2, 64, linear=True)
    p = gf.MultiLayerQG.Problem(2, gf.GPU(), nx=64, ny=64, Lx=kw["Lx"], Ly=kw["Ly"], f0=1.0, H=kw["H"], b=kw["b"],
                                U=kw["U"], eta=kw["eta"], β=kw["beta"], μ=kw["mu"], ν=kw["nu"], nν=1,
                                linear=True, dt=0.005, stepper="ForwardEuler", calcFq=Ff)
    gf.MultiLayerQG.set_q(p, qf)
    gf.stepforward(p, 100)
    gf.MultiLayerQG.updatevars(p)
    assert rel(p.vars.q, qf) < 1e-12


@pytest.mark.parametrize("stepper", ["RK4", "FilteredRK4", "ForwardEuler"])
def test_mlqg_other_steppers(gf, stepper):
    n = 64
    rng = np.random.default_rng(11)
    s = rand_spec(rng, (2, n, n // 2 + 1)) * 0.05
    kw = dict(nx=n, Lx=TWO_PI, H=[0.4, 0.6], b=[-1.0, -1.3], U=[0.5, 0.0], dt=0.002, stepper=stepper)
    p = gf.MultiLayerQG.Problem(2, gf.GPU(), β=1.0, μ=0.05, ν=1e-3, **kw)
    r = oml.Problem(2, beta=1.0, mu=0.05, nu=1e-3, **kw)
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 3)
    off.stepforward(r, 3)
    assert rel(p.sol, r.sol) < 1e-12


# ----------------------------------------------------------------------------- full size
def test_full_size_2048_forced_steps_match_oracle(gf):
    """The benchmark workload (2048^2 FP64 ETDRK4, ring forcing) under the shipped tuned
    schedule (P = (16, 16, 8), K3 column split): 5 steps from a random state in one call (4
    interior steps on the retained columns, the step counter on the device, then the full last
    step) against the oracle at 1e-12, then bit-exact determinism of the device path."""
    n = 2048
    kw = dict(nx=n, Lx=TWO_PI, dt=0.005, stepper="ETDRK4")
    rng = np.random.default_rng(20)
    z = rng.standard_normal((n, n))
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), ν=1e-21, nν=4, μ=0.1, nμ=0, calcF=gf.RingForcing(14.0, 1.5, 0.1, seed=2), **kw)
    assert gf.schedule(p) == (16, 16, 8) and gf.splits(p) == (0, 0, 1)
    gf.TwoDNavierStokes.set_ζ(p, z)
    s0 = p.sol
    r = o2d.Problem(nu=1e-21, nnu=4, mu=0.1, nmu=0, calcF=("ring", 14.0, 1.5, 0.1, 2), **kw)
    r.sol[...] = s0
    gf.stepforward(p, 5)
    off.stepforward(r, 5)
    assert rel(p.sol, r.sol) < 1e-12
    s5 = p.sol
    p.sol = s0
    p.set_clock(0.0, 0)
    gf.stepforward(p, 5)
    assert np.array_equal(p.sol, s5)


def test_full_size_2048_fp32_shipped_schedule_matches_fp64_oracle(gf):
    """The Float32 variant of the benchmark workload under the shipped tuned schedule
    (P = (8, 16, 4), K1 field split with the per-transform address refresh that FP32 K1 uses
    from 2048 points, K2 staged loads, K3 column split): 3 steps in one call against the FP64 oracle started
    from the same (FP32-rounded) state.  Tolerance 2e-5 relative L2, as for the other FP32
    checks: a few hundred FP32 roundings deep (eps 6e-8) per step."""
    n = 2048
    kw = dict(nx=n, Lx=TWO_PI, dt=0.005, stepper="ETDRK4")
    z = np.random.default_rng(22).standard_normal((n, n))
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), ν=1e-21, nν=4, μ=0.1, nμ=0, T=np.float32,
                                    calcF=gf.RingForcing(14.0, 1.5, 0.1, seed=2), **kw)
    assert gf.schedule(p) == (8, 16, 4) and gf.splits(p) == (1, 2, 1)
    gf.TwoDNavierStokes.set_ζ(p, z)
    s0 = p.sol.astype(np.complex128)
    r = o2d.Problem(nu=1e-21, nnu=4, mu=0.1, nmu=0, calcF=("ring", 14.0, 1.5, 0.1, 2), **kw)
    r.sol[...] = s0
    gf.stepforward(p, 3)
    off.stepforward(r, 3)
    assert np.isfinite(p.sol).all()
    assert rel(p.sol, r.sol) < 2e-5


def test_4096_calcN_and_step_match_oracle(gf):
    """Beyond the benchmark size: 4096^2, under the shipped tuned schedules (FP64 (8, 8, 8), FP32
    (8, 16, 8) with the K3 column split) and the built-in one (16 points per thread in every pass;
    the y passes' column-pair workgroups then take 147 KB of LDS in FP64).  FP64 calcN at 1e-13
    and one RK4 step at 1e-12 against the oracle, FP32 calcN at 2e-6."""
    n = 4096
    kw = dict(nx=n, Lx=TWO_PI, dt=0.001, stepper="RK4")
    rng = np.random.default_rng(40)
    s = 0.01 * (rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1)))
    s[0, 0] = 0
    r = o2d.Problem(nu=1e-25, nnu=4, mu=0.1, nmu=0, **kw)
    Nr, sr = np.zeros_like(s), s.copy()
    o2d.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    for dtype, tol, shipped in ((np.float32, 2e-6, (8, 16, 8)), (np.float64, 1e-13, (8, 8, 8))):
        p = gf.TwoDNavierStokes.Problem(gf.GPU(), ν=1e-25, nν=4, μ=0.1, nμ=0, T=dtype, **kw)
        assert gf.schedule(p) == shipped
        N, _ = p.calcN(s)
        assert rel(N, Nr) < tol, dtype
        gf.set_schedule(p, 16, 16, 16)
        gf.set_splits(p, 0, 0, 0)
        N, _ = p.calcN(s)
        assert rel(N, Nr) < tol, dtype
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 1)
    off.stepforward(r, 1)
    assert rel(p.sol, r.sol) < 1e-12


@pytest.mark.parametrize("n", [16, 64])
@pytest.mark.parametrize("f", [1 / 3, 1 / 2])
def test_device_calcN_retained_modes_are_alias_free(gf, n, f):
    """On every retained mode the device N equals the exact convolution of the retained
    modes (tests/alias_free.py); at n = 16 and 64 with f = 1/3 this fails if the l band is
    one index too narrow at its top."""
    from tests.alias_free import exact_calcN
    s = np.fft.rfft2(np.random.default_rng(n).standard_normal((n, n)))
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, stepper="ETDRK4", aliased_fraction=f)
    N, sm = p.calcN(s)
    m = off.TwoDGrid(n, TWO_PI, aliased_fraction=f).dealias_mask()
    E = exact_calcN(sm, n, n)
    assert rel(N * m, E * m) < 1e-13


def test_translation_equivariance_512(gf):
    """Shifting the initial vorticity by one grid cell in x shifts the solution: the
    spectral state picks up exp(-i kr dx) exactly (size-independent property)."""
    n = 512
    rng = np.random.default_rng(21)
    z = rng.standard_normal((n, n))
    kw = dict(nx=n, Lx=TWO_PI, ν=1e-6, nν=2, dt=0.002, stepper="ETDRK4")
    a = gf.TwoDNavierStokes.Problem(gf.GPU(), **kw)
    b = gf.TwoDNavierStokes.Problem(gf.GPU(), **kw)
    gf.TwoDNavierStokes.set_ζ(a, z)
    gf.TwoDNavierStokes.set_ζ(b, np.roll(z, 1, axis=1))
    gf.stepforward(a, 5)
    gf.stepforward(b, 5)
    g = off.TwoDGrid(n, TWO_PI)
    shift = np.exp(-1j * g.Kr * g.dx)
    m = g.dealias_mask()
    assert rel(b.sol * m, a.sol * shift * m) < 1e-12


@pytest.mark.parametrize("n", [64, 512])
def test_mlqg_ring_forcing_matches_oracle(gf, n):
    # the ring forcing's realisation is added to every layer; with one layer per workgroup row each
    # row draws it into a box slice of its own (64^2: column pairs, 512^2: one column per workgroup)
    from oracle import forcing as ofo
    kf, dkf, eps, seed = 6.0, 1.5, 0.1, 21
    kw = dict(nx=n, Lx=TWO_PI, H=[0.2, 0.8], b=[-1.0, -1.2], U=[1.0, 0.0], dt=0.002, stepper="ETDRK4")
    p = gf.MultiLayerQG.Problem(2, gf.GPU(), f0=1.0, β=5.0, μ=0.005, ν=1e-4,
                                calcFq=gf.RingForcing(kf, dkf, eps, seed=seed), **kw)
    cache = {}

    def ring(Fqh, sol, t, clock, vars, prm, grid):
        if "s" not in cache:
            cache["s"] = np.sqrt(ofo.ring_spectrum(grid, kf, dkf, eps))
        Fqh[...] = ofo.ring_forcing(grid, cache["s"], seed, clock.step, clock.dt)[None]

    r = oml.Problem(2, f0=1.0, beta=5.0, mu=0.005, nu=1e-4, calcFq=ring, stochastic=True, **kw)
    s = rand_spec(np.random.default_rng(n), (2, n, n // 2 + 1)) * 0.02
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 4)
    off.stepforward(r, 4)
    assert rel(p.sol, r.sol) < 1e-12


# ----------------------------------------------------------------------------- failure detection
def test_nonfinite_state_is_reported(gf):
    """An over-large dt on a small grid drives the state to non-finite values: gf_synchronize
    (and gf_step, once a finished step has flagged it) returns GF_ENOTFINITE; setting the state
    clears the flag, and a stable run reports nothing."""
    n = 32
    z = 50.0 * np.random.default_rng(30).standard_normal((n, n))
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, dt=5.0, stepper="RK4")
    gf.TwoDNavierStokes.set_ζ(p, z)
    with pytest.raises(gf._lib.GFError) as e:
        for _ in range(200):
            gf.stepforward(p, 1)
        p.synchronize()
    assert e.value.code == gf._lib.GF_ENOTFINITE
    assert not np.isfinite(p.sol).all()
    with pytest.raises(gf._lib.GFError):
        p.synchronize()                     # the flag stays until the state is set
    gf.TwoDNavierStokes.set_ζ(p, 1e-3 * z)
    p.synchronize()
    q = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, dt=1e-3, stepper="ETDRK4")
    gf.TwoDNavierStokes.set_ζ(q, z / 50)
    gf.stepforward(q, 20)
    q.synchronize()
    assert np.isfinite(q.sol).all()
