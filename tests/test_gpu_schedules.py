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
    # 8 P triples x K1 split x K2 split 0, 1 (and 2, the staged pass, where 256 / P_K2 >= 64) x
    # K3 split 0 (and 1 where 256 / P_K3 >= 64): 2 x 2 (P_K1, split) x 5 (P_K2, split) x 3 (P_K3, split)
    assert len(us) == 60 and min(us.values()) > 0 and us[chosen] == min(us.values())
    assert all(k[5] == 0 or k[2] == 4 for k in us) and all(k[4] != 2 or k[1] == 4 for k in us)
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


@pytest.mark.parametrize("n,ny,dtype,p2", [(256, 256, np.float64, 4), (512, 512, np.float64, 8),
                                           (1024, 1024, np.float32, 16), (1024, 1024, np.float64, 8),
                                           (2048, 2048, np.float64, 8), (2048, 2048, np.float32, 8),
                                           (2048, 2048, np.float64, 16), (512, 256, np.float64, 4)])
def test_k2_staged_loads_match_the_plain_pass_and_oracle(gf, n, ny, dtype, p2):
    # K2 split 2 stages each field's rows through LDS one field ahead.  The arithmetic is the
    # unsplit pass's; the compiler may contract the products into the first transform pass
    # differently in the two kernels, so FP64 results agree to rounding (FP32 bit for bit here)
    rng = np.random.default_rng(99)
    T = gf.TwoDNavierStokes
    Ly = 2 * math.pi * ny / n
    kw = dict(nx=n, ny=ny, Ly=Ly, ν=1e-8, nν=2, μ=0.1, dt=0.001, stepper="ETDRK4", T=dtype)
    z = 0.1 * rng.standard_normal((ny, n))
    out = []
    for k2 in (0, 2):
        p = T.Problem(gf.GPU(), **kw)
        gf.set_schedule(p, 8 if n >= 512 else 4, p2, 8 if n >= 512 else 4)
        gf.set_splits(p, 1, k2, 0)
        assert gf.splits(p) == (1, k2, 0)
        T.set_ζ(p, z)
        s0 = p.sol
        N, _ = p.calcN(s0)
        gf.stepforward(p, 4)
        out.append((N, p.sol))
    tol = 1e-13 if dtype == np.float64 else 2e-6
    assert rel(out[1][0], out[0][0]) < tol / 10 and rel(out[1][1], out[0][1]) < tol / 10
    if dtype == np.float32:
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.isfinite(out[1][1]).all()
    r = o2d.Problem(nx=n, ny=ny, Ly=Ly, nu=1e-8, nnu=2, mu=0.1, dt=0.001, stepper="ETDRK4")
    Nr, sr = np.zeros_like(s0, dtype=np.complex128), s0.astype(np.complex128)
    o2d.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(out[1][0], Nr) < tol


def test_k2_staged_loads_with_topography(gf):
    # SingleLayerQG with topography adds eta to q inside the staged pass
    n = 512
    x = -math.pi + 2 * math.pi * np.arange(n) / n
    X, Y = np.meshgrid(x, x)
    eta = 0.2 * np.cos(3 * X) * np.sin(2 * Y)
    s = 0.05 * _spec(np.random.default_rng(98), n)
    s[0, 0] = 0
    out = []
    for k2 in (0, 2):
        p = gf.SingleLayerQG.Problem(gf.GPU(), nx=n, β=2.0, μ=0.01, U=0.3, eta=eta, dt=0.002, stepper="FilteredRK4")
        gf.set_schedule(p, 8, 8, 8)
        gf.set_splits(p, 0, k2, 0)
        N, _ = p.calcN(s)
        p.sol = s
        gf.stepforward(p, 3)
        out.append((N, p.sol))
    assert rel(out[1][0], out[0][0]) < 1e-14 and rel(out[1][1], out[0][1]) < 1e-14
    r = osl.Problem(nx=n, beta=2.0, mu=0.01, U=0.3, eta=eta, dt=0.002, stepper="FilteredRK4")
    Nr, sr = np.zeros_like(s), s.copy()
    osl.calcN(Nr, sr, 0.0, r.clock, r.vars, r.params, r.grid)
    assert rel(out[1][0], Nr) < 1e-13


def test_kernel_timer_counts_each_launch_and_stays_within_the_step(gf):
    # gf_timing: every K1/K2/K3 launch of the ETDRK4 stages is counted once; the summed kernel
    # times (each kernel's own start to end) fit inside the wall time of the steps
    import ctypes
    import time
    lib = gf._lib.lib
    n, steps = 512, 20
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-8, nν=2, μ=0.1, dt=0.001, stepper="ETDRK4",
                                    calcF=gf.RingForcing(10.0, 1.5, 0.1, seed=7))
    gf.TwoDNavierStokes.set_ζ(p, 0.1 * np.random.default_rng(3).standard_normal((n, n)))
    gf.stepforward(p, 2)
    ref = p.sol
    q = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-8, nν=2, μ=0.1, dt=0.001, stepper="ETDRK4",
                                    calcF=gf.RingForcing(10.0, 1.5, 0.1, seed=7))
    gf.TwoDNavierStokes.set_ζ(q, 0.1 * np.random.default_rng(3).standard_normal((n, n)))
    assert lib.gf_timing(q._ctx, 1) == 0
    t0 = time.perf_counter()
    gf.stepforward(q, 2)
    lib.gf_synchronize(q._ctx)
    wall_ms = (time.perf_counter() - t0) * 1e3
    assert np.array_equal(q.sol, ref)            # timing changes no result
    gf.stepforward(q, steps - 2)
    lib.gf_synchronize(q._ctx)
    tot = 0.0
    for k in range(3):
        cnt, ms = ctypes.c_int64(), ctypes.c_double()
        assert lib.gf_timing_read(q._ctx, k, ctypes.byref(cnt), ctypes.byref(ms)) == 0
        assert cnt.value == 4 * steps and ms.value > 0
        tot += ms.value
    assert lib.gf_timing(q._ctx, 0) == 0
    # the first two steps alone: their 24 kernels ran within the measured wall time
    assert tot / steps * 2 < wall_ms


def test_time_kernel_leaves_the_state_and_rejects_bad_classes(gf):
    import ctypes
    lib = gf._lib.lib
    n = 256
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-8, nν=2, μ=0.1, dt=0.001, stepper="ETDRK4",
                                    calcF=gf.RingForcing(10.0, 1.5, 0.1, seed=7))
    gf.TwoDNavierStokes.set_ζ(p, 0.1 * np.random.default_rng(4).standard_normal((n, n)))
    gf.stepforward(p, 3)
    s0, t0, k0 = p.sol, p.clock.t, p.clock.step
    ms = ctypes.c_double()
    for k in range(3):
        assert lib.gf_time_kernel(p._ctx, k, 5, ctypes.byref(ms)) == 0 and 0 < ms.value < 10
    assert np.array_equal(p.sol, s0) and p.clock.t == t0 and p.clock.step == k0
    gf.stepforward(p, 2)                       # stepping goes on as from the saved state
    q = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-8, nν=2, μ=0.1, dt=0.001, stepper="ETDRK4",
                                    calcF=gf.RingForcing(10.0, 1.5, 0.1, seed=7))
    q.sol = s0
    q.set_clock(t0, k0)
    gf.stepforward(q, 2)
    assert np.array_equal(p.sol, q.sol)
    assert lib.gf_time_kernel(p._ctx, 3, 5, ctypes.byref(ms)) == gf._lib.GF_EINVAL
    assert lib.gf_time_kernel(p._ctx, 1, 0, ctypes.byref(ms)) == gf._lib.GF_EINVAL
    # a three-layer MultiLayerQG problem (layered kernels, bottom drag) and an ensemble
    m = gf.MultiLayerQG.Problem(3, gf.GPU(), nx=128, f0=1.0, H=[0.2, 0.3, 0.5], b=[-1.0, -1.2, -1.3], μ=0.01,
                                U=[0.3, 0.1, 0.0], dt=0.005, stepper="ETDRK4")
    m.sol = 1e-3 * (np.random.default_rng(5).standard_normal(m._sshape) + 0j)
    gf.stepforward(m, 2)
    m0 = m.sol
    for k in range(3):
        assert lib.gf_time_kernel(m._ctx, k, 3, ctypes.byref(ms)) == 0 and ms.value > 0
    assert np.array_equal(m.sol, m0)
    e = gf.Ensemble([(gf.TwoDNavierStokes.Problem, dict(nx=64, dt=0.01, stepper="RK4")) for _ in range(2)])
    assert lib.gf_time_kernel(e._ctx, 1, 3, ctypes.byref(ms)) == gf._lib.GF_EUNSUPPORTED
