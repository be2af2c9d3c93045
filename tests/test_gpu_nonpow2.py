"""Grid lengths that are not powers of two (even, 18..256). A column length ny (the reference's
own grids 64 x 66, 128 x 126 and 32 x 34, test/test_multilayerqg.jl:232-701) runs in the y-pass
kernels of the next power of two, with a direct DFT in place of the Stockham transform; a row
length nx runs in the x pass of direct sums (one workgroup per row).

Tolerances: FP64 1e-13 relative for one calcN and for updatevars/set_psi!, 1e-12 relative L2
after 3-4 steps; FP32 2e-6 for calcN and 2e-5 after 3 steps, as the other FP32 checks; the
reference's own rtol 1e-6 / atol 1e-7 on its printed flux values and 1e-13 on its energies.
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


@pytest.mark.parametrize("nx,ny,dtype", [(64, 66, np.float64), (128, 126, np.float64), (32, 34, np.float64),
                                         (256, 250, np.float64), (64, 18, np.float64), (64, 66, np.float32),
                                         (66, 64, np.float64), (126, 126, np.float64), (254, 128, np.float64),
                                         (18, 32, np.float64), (68, 66, np.float64), (66, 64, np.float32),
                                         (18, 18, np.float64), (250, 66, np.float64), (126, 126, np.float32)])
def test_twodns_ring_forced_calcN_steps_updatevars(gf, nx, ny, dtype):
    Ly = TWO_PI * ny / nx
    kw = dict(nx=nx, ny=ny, Lx=TWO_PI, Ly=Ly, dt=0.002, stepper="ETDRK4")
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), ν=1e-4, nν=1, μ=0.05, T=dtype,
                                    calcF=gf.RingForcing(6.0, 1.5, 0.1, seed=3), **kw)
    r = o2d.Problem(nu=1e-4, nnu=1, mu=0.05, calcF=("ring", 6.0, 1.5, 0.1, 3), **kw)
    s = 0.05 * rand_spec(np.random.default_rng(ny), (ny, nx // 2 + 1))
    s[0, 0] = 0
    s = s.astype(np.complex64 if dtype == np.float32 else np.complex128).astype(np.complex128)
    t1, t3 = (1e-13, 1e-12) if dtype == np.float64 else (2e-6, 2e-5)
    N, _ = p.calcN(s)
    Nr, sr = np.zeros_like(s), s.copy()
    o2d.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < t1
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 3)
    off.stepforward(r, 3)
    assert np.isfinite(p.sol).all() and rel(p.sol, r.sol) < t3
    gf.TwoDNavierStokes.updatevars(p)
    o2d.updatevars(r)
    assert rel(p.vars.ζ, r.vars.zeta) < t1 and rel(p.vars.u, r.vars.u) < t1


@pytest.mark.parametrize("nx,ny", [(64, 66), (66, 64), (90, 90)])
def test_twodns_set_zeta_round_trip_and_energy(gf, nx, ny):
    # physical -> spectral (K2 forward, K3 store) and back, and the device energy of a state
    # with a closed form: E = <zeta^2>/(2 |k|^2) per mode, (1/4)/13/2 + (1/8)/5/2
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, dt=0.01, stepper="RK4")
    r = o2d.Problem(nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, dt=0.01, stepper="RK4")
    x, y = r.grid.gridpoints()
    z = np.cos(2 * x) * np.cos(3 * y) + 0.5 * np.sin(x - 2 * y)
    gf.TwoDNavierStokes.set_ζ(p, z)
    o2d.set_zeta(r, z)
    assert rel(p.sol, r.sol) < 1e-13
    gf.TwoDNavierStokes.updatevars(p)
    assert rel(p.vars.ζ, z) < 1e-13
    e = gf.TwoDNavierStokes.energy(p)
    assert abs(e - (0.25 / 13 / 2 + 0.125 / 5 / 2)) < 1e-13


@pytest.mark.parametrize("nx,ny,ell,U,topo", [(64, 66, None, 0.0, False), (64, 66, 1.5, 0.3, True),
                                              (66, 64, None, 0.0, False), (66, 64, 1.5, 0.3, True),
                                              (66, 64, None, "y", True)])
def test_slqg_calcN_and_steps(gf, nx, ny, ell, U, topo):
    g = off.TwoDGrid(nx, TWO_PI, ny, TWO_PI)
    X, Y = g.gridpoints()
    eta = 0.4 * np.cos(3 * X) * np.cos(2 * Y) if topo else None
    if U == "y":   # U(y): the x pass with background gradients
        U = 0.3 / np.cosh(2 * g.y) ** 2
    kw = dict(nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, U=U, eta=eta, dt=0.005, stepper="FilteredRK4")
    p = gf.SingleLayerQG.Problem(gf.GPU(), β=2.0, deformation_radius=ell, ν=1e-3, μ=0.01, **kw)
    r = osl.Problem(beta=2.0, deformation_radius=ell, nu=1e-3, mu=0.01, **kw)
    s = rand_spec(np.random.default_rng(6), (ny, nx // 2 + 1)) * 0.1
    N, _ = p.calcN(s)
    Nr, sr = np.zeros_like(s), s.copy()
    osl.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < 1e-13
    p.sol = s
    r.sol[...] = s
    gf.stepforward(p, 4)
    off.stepforward(r, 4)
    assert rel(p.sol, r.sol) < 1e-12


@pytest.mark.parametrize("nlayers,nx,ny", [(1, 64, 66), (2, 64, 66), (3, 64, 66), (1, 66, 66), (2, 66, 66),
                                           (3, 66, 66), (2, 126, 64)])
def test_mlqg_calcN_steps_updatevars(gf, nlayers, nx, ny):
    # the reference's 64 x 66 grid (test/test_multilayerqg.jl:232-300), and non-power-of-two rows,
    # with U(y), topography, beta and bottom drag
    H = [0.2, 0.3, 0.5][:nlayers] if nlayers > 1 else None
    b = [-1.0, -1.2, -1.3][:nlayers] if nlayers > 1 else None
    g = off.TwoDGrid(nx, TWO_PI, ny, TWO_PI)
    X, Y = g.gridpoints()
    eta = 0.3 * np.exp(-((X + 1) ** 2 + (Y - 0.5) ** 2) / 0.3)
    Uy = np.stack([0.4 * np.cos(g.y) * (nlayers - j) / nlayers for j in range(nlayers)], axis=1)
    kw = dict(nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, U=Uy, eta=eta, dt=0.005, stepper="ETDRK4")
    if nlayers > 1:
        kw.update(H=H, b=b)
    p = gf.MultiLayerQG.Problem(nlayers, gf.GPU(), f0=1.0, β=5.0, μ=0.01, ν=1e-3, **kw)
    r = oml.Problem(nlayers, f0=1.0, beta=5.0, mu=0.01, nu=1e-3, **kw)
    shape = r.sol.shape
    s = rand_spec(np.random.default_rng(40 + nlayers), shape) * 0.05
    N, _ = p.calcN(s.reshape(p._sshape))
    Nr, sr = np.zeros_like(s), s.copy()
    oml.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N.reshape(shape), Nr) < 1e-13
    p.sol = s.reshape(p._sshape)
    r.sol[...] = s
    gf.stepforward(p, 3)
    off.stepforward(r, 3)
    assert rel(p.sol.reshape(shape), r.sol) < 1e-12
    gf.MultiLayerQG.updatevars(p)
    oml.updatevars(r)
    for a, bb in ((p.vars.q, r.vars.q), (p.vars.ψ, r.vars.psi), (p.vars.u, r.vars.u), (p.vars.v, r.vars.v)):
        assert rel(np.reshape(a, bb.shape), bb) < 1e-13


@pytest.mark.parametrize("nlayers", [1, 2])
def test_mlqg_linear_on_nonpow2_rows(gf, nlayers):
    # the linearised equations (U q in place of u q) on 66 x 64
    nx, ny = 66, 64
    g = off.TwoDGrid(nx, TWO_PI, ny, TWO_PI)
    Uy = np.stack([0.3 * np.sin(2 * g.y) * (nlayers - j) / nlayers for j in range(nlayers)], axis=1)
    kw = dict(nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, U=Uy if nlayers > 1 else Uy[:, 0], dt=0.004,
              stepper="RK4", linear=True)
    if nlayers > 1:
        kw.update(H=[0.4, 0.6], b=[-1.0, -1.2])
    p = gf.MultiLayerQG.Problem(nlayers, gf.GPU(), β=3.0, μ=0.05, ν=1e-3, **kw)
    r = oml.Problem(nlayers, beta=3.0, mu=0.05, nu=1e-3, **kw)
    s = rand_spec(np.random.default_rng(70 + nlayers), r.sol.shape) * 0.05
    p.sol = s.reshape(p._sshape)
    r.sol[...] = s
    gf.stepforward(p, 3)
    off.stepforward(r, 3)
    assert rel(p.sol.reshape(r.sol.shape), r.sol) < 1e-12


@pytest.mark.parametrize("topo,vary", [(True, False), (False, True)])
def test_slqg_reference_nonlinear_advection_126x128(gf, topo, vary):
    """test/test_singlelayerqg.jl:276-399 (ForwardEuler, dt = 5e-4, 2000 steps, tolerance 1e-13)
    on a 126 x 128 grid: the steady state's closed form holds whatever the row length."""
    from tests.test_gpu_parity import _slqg_forcing
    nx, ny = 126, 128
    g = off.TwoDGrid(nx, TWO_PI, ny, TWO_PI)
    x, y = g.gridpoints()
    eta0 = 0.4 if topo else 0.0
    eta = eta0 * np.cos(10 * x) * np.cos(10 * y)
    qf = -8 * np.sin(2 * x) * np.cos(2 * y) - 20 * np.sin(x) * np.cos(3 * y)
    U, Uyy = 0.2, 0.0
    if vary:
        U = 0.2 / np.cosh(2 * y) ** 2
        Uyy = g.irfft(-g.Lw ** 2 * g.rfft(U))
    Fh = _slqg_forcing(g, eta0, U, Uyy)
    p = gf.SingleLayerQG.Problem(gf.GPU(), nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, U=U[:, 0] if vary else U,
                                 eta=eta, ν=1e-2, nν=1, dt=5e-4, stepper="ForwardEuler", calcF=Fh)
    gf.SingleLayerQG.set_q(p, qf)
    gf.stepforward(p, 2000)
    gf.SingleLayerQG.updatevars(p)
    assert rel(p.vars.q, qf) < 1e-13


def test_mlqg_energies_on_the_reference_grid(gf):
    # test/test_multilayerqg.jl:502-536 on its own 64 x 66 grid, entering through set_psi!:
    # KE = (1/4, 3/16), PE = 1/160
    p = gf.MultiLayerQG.Problem(2, gf.GPU(), nx=64, ny=66, Lx=TWO_PI, Ly=TWO_PI, f0=1, H=[2.5, 7.5],
                                b=[-1.0, -1.5], dt=0.01, stepper="RK4")
    x, y = p.grid.gridpoints()
    gf.MultiLayerQG.set_ψ(p, np.array([np.cos(2 * x) * np.cos(2 * y), 0.5 * np.cos(2 * x) * np.cos(2 * y)]))
    KE, PE = gf.MultiLayerQG.energies(p)
    assert abs(KE[0] - 1 / 4) < 1e-13 and abs(KE[1] - 3 / 16) < 1e-13 and abs(PE[0] - 1 / 160) < 1e-13


@pytest.mark.parametrize("nlayers", [2, 3])
def test_mlqg_fluxes_on_the_reference_grid(gf, nlayers):
    # test/test_multilayerqg.jl:567-625 on its own 128 x 126 grid: lateral 0.00626267, vertical
    # -0.196539 (rtol 1e-6), the third layer's fluxes zero
    nx, ny = 128, 126
    if nlayers == 2:
        H, b = [0.2, 0.8], [-1.0, -1.2]
    else:
        H, b = [0.2, 0.35, 0.45], [-1.0, -1.2, -1.2 + 1e-9]
    g = off.TwoDGrid(nx, TWO_PI, ny, TWO_PI)
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
    tol = 1e-12 if nlayers == 2 else 1e-6    # 3 layers: g' = -1e-9, S^-1 conditioned ~1e9
    assert abs(lat[0] - latr[0]) <= tol * abs(latr[0]) and abs(ver[0] - verr[0]) <= tol * abs(verr[0])


def test_mlqg_fluxes_on_nonpow2_rows(gf):
    # the flux diagnostic (x sums by Parseval over a row length that is not a power of two)
    # against the oracle on 126 x 128, two layers, U(y) in the upper layer; the fields are
    # resolved on both grids, so the reference's values for its 128 x 126 grid hold too
    # (test/test_multilayerqg.jl:567-625: 0.00626267 and -0.196539, rtol 1e-6)
    nx, ny = 126, 128
    g = off.TwoDGrid(nx, TWO_PI, ny, TWO_PI)
    U = np.zeros((ny, 2))
    U[:, 0] = 1 / np.cosh(g.y / 0.2) ** 2
    kw = dict(nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, f0=1, H=[0.2, 0.8], b=[-1.0, -1.2], U=U)
    p = gf.MultiLayerQG.Problem(2, gf.GPU(), dt=0.01, stepper="RK4", **kw)
    r = oml.Problem(2, **kw)
    x, y = r.grid.gridpoints()
    psi = np.array([np.cos(x) * np.cos(y) + np.sin(x), np.cos(x + np.pi / 10) * np.cos(y)])
    gf.MultiLayerQG.set_ψ(p, psi)
    oml.set_psi(r, psi)
    lat, ver = gf.MultiLayerQG.fluxes(p)
    latr, verr = oml.fluxes(r)
    assert abs(lat[0] - latr[0]) <= 1e-12 * abs(latr[0]) and abs(ver[0] - verr[0]) <= 1e-12 * abs(verr[0])
    assert abs(lat[1] - latr[1]) <= 1e-12 * abs(latr[0])
    assert abs(lat[0] - 0.00626267) <= 1e-6 * 0.00626267 and abs(ver[0] + 0.196539) <= 1e-6 * 0.196539


def test_mlqg_single_layer_fluxes_on_the_reference_grid(gf):
    # test/test_multilayerqg.jl:632-655 on 128 x 126: 0.0313134 (atol 1e-7)
    nx, ny = 128, 126
    g = off.TwoDGrid(nx, TWO_PI, ny, TWO_PI)
    U = 1 / np.cosh(g.y / 0.2) ** 2
    p = gf.MultiLayerQG.Problem(1, gf.GPU(), nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, U=U, dt=0.01, stepper="RK4")
    x, y = p.grid.gridpoints()
    gf.MultiLayerQG.set_ψ(p, np.cos(x) * np.cos(y) + np.sin(x))
    assert abs(gf.MultiLayerQG.fluxes(p) - 0.0313134) <= 1e-7


@pytest.mark.parametrize("nx,ny", [(32, 34), (34, 32)])
def test_mlqg_set_psi_set_q_round_trip(gf, nx, ny):
    # test/test_multilayerqg.jl:663-701 on its 32 x 34 grid (and 34 x 32): set_psi! gives psi back
    # through updatevars!, set_q! gives q back
    p = gf.MultiLayerQG.Problem(2, gf.GPU(), nx=nx, ny=ny, Lx=TWO_PI, Ly=TWO_PI, f0=1, H=[0.2, 0.8],
                                b=[-1.0, -1.2], dt=0.01, stepper="RK4")
    x, y = p.grid.gridpoints()
    f = np.array([2 * np.cos(x) * np.cos(y), np.cos(x + np.pi / 10) * np.cos(2 * y)])
    gf.MultiLayerQG.set_ψ(p, f)
    assert rel(np.reshape(p.vars.ψ, f.shape), f) < 1e-13
    gf.MultiLayerQG.set_q(p, f)
    gf.MultiLayerQG.updatevars(p)
    assert rel(np.reshape(p.vars.q, f.shape), f) < 1e-13


def test_unsupported_grid_lengths_and_schedules(gf):
    for nx, ny in ((64, 65), (64, 258), (65, 64), (258, 64), (8192, 64)):
        with pytest.raises(gf._lib.GFError, match="even between 18 and 256"):
            gf.TwoDNavierStokes.Problem(gf.GPU(), nx=nx, ny=ny, stepper="RK4")
    for nx, ny in ((64, 66), (66, 64)):
        p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=nx, ny=ny, stepper="RK4")
        assert gf.schedule(p) == (8, 8, 8)        # the 128-point kernels' built-in P
        with pytest.raises(gf._lib.GFError, match="not offered"):
            gf.set_schedule(p, 4, 8, 4)
        with pytest.raises(gf._lib.GFError, match="nx and ny powers of two"):
            gf.set_splits(p, 1, 0, 0)
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=66, ny=66, dt=0.01, stepper="ETDRK4")
    with pytest.raises(gf._lib.GFError, match="power-of-two grids"):
        gf.TwoDNavierStokes.adjoint_gradient(p, np.zeros((66, 66)), 2)
