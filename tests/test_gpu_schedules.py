"""Radix schedules and the autotuner on the device.  Every offered schedule computes calcN and
steps within the oracle tolerance (schedules differ from each other only by rounding), and the
autotuner leaves the state, the stepper history and the clock as it found them."""
import math

import numpy as np
import pytest

from oracle import singlelayerqg as osl
from oracle import twodnavierstokes as o2d
from tests.conftest import rel

pytestmark = pytest.mark.gpu
SCHEDULES = [(4, 4, 4), (8, 8, 8), (4, 8, 4), (8, 4, 8)]


def _spec(rng, n):
    return rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))


@pytest.mark.parametrize("n", [64, 256])
@pytest.mark.parametrize("sched", SCHEDULES)
@pytest.mark.parametrize("split", [(0, 0, 0), (1, 1, 1), (1, 0, 0), (0, 0, 1)])
def test_twodns_every_schedule_matches_oracle(gf, n, sched, split):
    rng = np.random.default_rng(80 + n)
    s = _spec(rng, n)
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, nν=1, μ=0.05, dt=0.001, stepper="ETDRK4",
                                    calcF=gf.RingForcing(6.0, 1.5, 0.1, seed=4))
    gf.set_schedule(p, *sched)
    gf.set_splits(p, *split)
    assert gf.schedule(p) == sched and gf.splits(p) == split
    r = o2d.Problem(nx=n, nu=1e-3, nnu=1, mu=0.05, dt=0.001, stepper="ETDRK4", calcF=("ring", 6.0, 1.5, 0.1, 4))
    N, _ = p.calcN(s)
    Nr, sr = np.zeros_like(s), s.copy()
    o2d.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < 1e-13
    s0 = 0.05 * s
    s0[0, 0] = 0
    p.sol = s0
    r.sol[...] = s0
    gf.stepforward(p, 5)
    from oracle import fourierflows as off
    off.stepforward(r, 5)
    assert rel(p.sol, r.sol) < 1e-12


@pytest.mark.parametrize("sched", SCHEDULES)
def test_slqg_schedules_topography_and_varying_U(gf, sched):
    n = 128
    x = -math.pi + 2 * math.pi * np.arange(n) / n
    X, Y = np.meshgrid(x, x)
    eta = 0.2 * np.cos(3 * X) * np.sin(2 * Y)
    Uy = 0.1 * np.cos(Y[:, 0])
    for kw in (dict(U=0.3, eta=eta), dict(U=Uy, eta=eta)):
        p = gf.SingleLayerQG.Problem(gf.GPU(), nx=n, β=2.0, μ=0.01, dt=0.002, stepper="FilteredRK4", **kw)
        gf.set_schedule(p, *sched)
        gf.set_splits(p, 1, 1, 1)
        rk = dict(kw)
        r = osl.Problem(nx=n, beta=2.0, mu=0.01, dt=0.002, stepper="FilteredRK4", **rk)
        s = _spec(np.random.default_rng(90), n)
        N, _ = p.calcN(s)
        Nr, sr = np.zeros_like(s), s.copy()
        osl.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
        assert rel(N, Nr) < 1e-13, kw.keys()


@pytest.mark.parametrize("n,dtype", [(512, np.float64), (1024, np.float64), (2048, np.float64), (2048, np.float32)])
def test_k3_column_split_is_bitwise_the_pair_layout(gf, n, dtype):
    # one column per workgroup regroups the same arithmetic: calcN and steps agree bit for bit
    rng = np.random.default_rng(95)
    T = gf.TwoDNavierStokes
    kw = dict(nx=n, ν=1e-8, nν=2, μ=0.1, dt=0.001, stepper="ETDRK4", T=dtype,
              calcF=gf.RingForcing(10.0, 1.5, 0.1, seed=7))
    z = 0.1 * rng.standard_normal((n, n))
    out = {}
    for p3 in (4, 8):
        for cs in (0, 1):
            p = T.Problem(gf.GPU(), **kw)
            gf.set_schedule(p, 8, 8, p3)
            gf.set_splits(p, 1, 0, cs)
            assert gf.splits(p) == (1, 0, cs)
            T.set_ζ(p, z)
            N, _ = p.calcN(p.sol)
            gf.stepforward(p, 3)
            out[p3, cs] = (N, p.sol)
        assert np.array_equal(out[p3, 0][0], out[p3, 1][0]) and np.array_equal(out[p3, 0][1], out[p3, 1][1])
        assert np.isfinite(out[p3, 1][1]).all()


def test_schedule_not_offered_is_rejected(gf):
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=16, stepper="RK4")
    with pytest.raises(gf._lib.GFError, match="not offered"):
        gf.set_schedule(p, 4, 8, 4)            # P = 8 needs 32 points
    with pytest.raises(gf._lib.GFError, match="not offered"):
        gf.set_schedule(p, 16, 4, 4)
    assert gf.schedule(p) == (4, 4, 4)         # the built-in choice below 32 points
    assert gf.schedule(gf.TwoDNavierStokes.Problem(gf.GPU(), nx=32, stepper="RK4")) == (8, 8, 8)
    m = gf.MultiLayerQG.Problem(2, gf.GPU(), nx=64, f0=1, H=[1.0, 1.0], b=[0, -1.0], dt=0.01, stepper="RK4")
    with pytest.raises(gf._lib.GFError, match="not offered"):
        gf.set_schedule(m, 4, 4, 4)
    with pytest.raises(gf._lib.GFError, match="TwoDNavierStokes and SingleLayerQG"):
        gf.set_splits(m, 1, 0)
    # a one-layer MultiLayerQG problem keeps the fixed schedule too
    m1 = gf.MultiLayerQG.Problem(1, gf.GPU(), nx=64, dt=0.01, stepper="RK4")
    with pytest.raises(gf._lib.GFError, match="not offered"):
        gf.set_schedule(m1, 4, 4, 4)
    with pytest.raises(gf._lib.GFError, match="TwoDNavierStokes and SingleLayerQG"):
        gf.set_splits(m1, 1, 0)


def test_autotune_restores_state_and_records_the_choice(gf, tmp_path):
    n = 256
    T = gf.TwoDNavierStokes
    # AB3 treats L explicitly: nu Krsq^2 dt stays below 0.01 at the largest kept wavenumber
    kw = dict(nx=n, ν=1e-9, nν=2, μ=0.1, dt=0.002, stepper="AB3", calcF=gf.RingForcing(8.0, 1.5, 0.1, seed=6))
    p, q = T.Problem(gf.GPU(), **kw), T.Problem(gf.GPU(), **kw)
    z = 0.2 * np.random.default_rng(91).standard_normal((n, n))
    T.set_ζ(p, z)
    T.set_ζ(q, z)
    gf.stepforward(p, 5)                       # AB3 history and prevsol in use
    gf.stepforward(q, 5)
    before, t0, k0 = p.sol, p.clock.t, p.clock.step
    w0 = T.energy_work(p)
    chosen, us = gf.autotune(p, steps=6)
    assert np.array_equal(p.sol, before) and p.clock.t == t0 and p.clock.step == k0
    assert T.energy_work(p) == w0
    assert all(c in (4, 8) for c in chosen[:3]) and gf.schedule(p) + gf.splits(p) == chosen   # no 16 at 256
    # 8 P triples x 4 (K1, K2) splits, and the K3 column split where 256 / P_K3 >= 64 (P_K3 = 4)
    assert len(us) == 48 and min(us.values()) > 0 and us[chosen] == min(us.values())
    assert all(k[5] == 0 or k[2] == 4 for k in us)
    gf.set_schedule(q, *chosen[:3])
    gf.set_splits(q, *chosen[3:])
    gf.stepforward(p, 7)                       # the history survived: same steps as the twin
    gf.stepforward(q, 7)
    assert np.isfinite(p.sol).all()
    assert np.array_equal(p.sol, q.sol)
    # later problems of the same kind start with the tuned schedule; the table round-trips
    r = T.Problem(gf.GPU(), **kw)
    assert gf.schedule(r) + gf.splits(r) == chosen
    path = tmp_path / "sched.json"
    gf.save_schedules(path)
    gf._lib.lib.gf_schedules_clear()
    r = T.Problem(gf.GPU(), **kw)
    assert gf.schedule(r) == (8, 8, 8) and gf.splits(r) == (0, 0, 0)
    assert gf.load_schedules(path) >= 1
    r = T.Problem(gf.GPU(), **kw)
    assert gf.schedule(r) + gf.splits(r) == chosen
    gf._lib.lib.gf_schedules_clear()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
@pytest.mark.parametrize("sched", [(16, 16, 16), (16, 8, 4), (8, 16, 16)])
def test_sixteen_point_schedules_at_1024(gf, dtype, tol, sched):
    n = 1024
    rng = np.random.default_rng(95)
    s = _spec(rng, n)
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, nν=1, dt=0.001, stepper="ETDRK4", T=dtype)
    gf.set_schedule(p, *sched)
    assert gf.schedule(p) == sched
    r = o2d.Problem(nx=n, nu=1e-3, nnu=1, dt=0.001, stepper="ETDRK4")
    N, _ = p.calcN(s)
    Nr, sr = np.zeros_like(s), s.copy()
    o2d.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(N, Nr) < tol


@pytest.mark.parametrize("nx,ny,Ly", [(128, 512, 2.5), (512, 256, 4.0)])
def test_k3_column_split_on_rectangular_grids(gf, nx, ny, Ly):
    # columns number nx/2 + 1 (kept: about nx/3) while a column's length is ny: the split is
    # offered from ny / P_K3 >= 64 and must agree with the pair layout bit for bit and with the oracle
    rng = np.random.default_rng(97)
    T = gf.TwoDNavierStokes
    kw = dict(nx=nx, ny=ny, Lx=2 * math.pi, Ly=Ly, ν=1e-4, nν=1, μ=0.05, dt=0.002, stepper="ETDRK4")
    s = rng.standard_normal((ny, nx // 2 + 1)) + 1j * rng.standard_normal((ny, nx // 2 + 1))
    s *= 0.05
    s[0, 0] = 0
    out = []
    for cs in (0, 1):
        p = T.Problem(gf.GPU(), **kw)
        gf.set_schedule(p, 4, 4, 4)
        gf.set_splits(p, 0, 0, cs)
        N, _ = p.calcN(s)
        p.sol = s
        gf.stepforward(p, 3)
        out.append((N, p.sol))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    r = o2d.Problem(nx=nx, ny=ny, Lx=2 * math.pi, Ly=Ly, nu=1e-4, nnu=1, mu=0.05, dt=0.002, stepper="ETDRK4")
    Nr, sr = np.zeros_like(s), s.copy()
    o2d.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(out[1][0], Nr) < 1e-13
    r.sol[...] = s
    from oracle import fourierflows as off
    off.stepforward(r, 3)
    assert rel(out[1][1], r.sol) < 1e-12
