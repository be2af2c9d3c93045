"""Batched ensembles on the GPU: every member of an ensemble stepped together must
equal the same problem stepped on its own (the single-instance path, itself
checked against the oracle in test_gpu_parity.py), and one member is checked
against the oracle directly.

Tolerances (FP64): member vs single run 1e-14 relative L2 (same kernels bodies,
same arithmetic per column and row; only the launch grouping differs); member vs
oracle 1e-12 relative L2 after a few steps, as for single problems.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import fourierflows as off
from oracle import singlelayerqg as osl
from tests.conftest import rel

pytestmark = pytest.mark.gpu
TWO_PI = 2 * math.pi


def _slqg_kw(gf, n, rng, stepper="FilteredRK4", dt=0.01, **extra):
    grid = gf.problem.Grid(n, n, TWO_PI, TWO_PI, 1 / 3, np.float64)
    eta = gf.utils.gaussian_bump(grid, 0.1, TWO_PI / 10, rng.uniform(-math.pi, math.pi, 2))
    kw = dict(nx=n, Lx=TWO_PI, β=10.0, μ=0.01, eta=eta, dt=dt, stepper=stepper)
    kw.update(extra)
    return kw


def _q0(gf, n, rng):
    # energy 0.02 keeps dt = 0.01 inside the advective stability limit up to 512^2
    grid = gf.problem.Grid(n, n, TWO_PI, TWO_PI, 1 / 3, np.float64)
    return gf.peakedisotropicspectrum(grid, min(8, n // 8), 0.02, rng=rng)


def test_ensemble_members_match_single_runs(gf):
    rng = np.random.default_rng(30)
    sizes = [64, 32, 512, 64, 128, 256, 64, 128]
    kws = [_slqg_kw(gf, n, rng) for n in sizes]
    q0 = [_q0(gf, n, rng) for n in sizes]
    ens = gf.Ensemble([(gf.SingleLayerQG.Problem, kw) for kw in kws])
    assert len(ens) == len(sizes)
    singles = []
    for i, kw in enumerate(kws):
        gf.SingleLayerQG.set_q(ens[i], q0[i])
        p = gf.SingleLayerQG.Problem(gf.GPU(), **kw)
        gf.SingleLayerQG.set_q(p, q0[i])
        singles.append(p)
    # two calls: the lazily produced aliased band must agree after each
    for nsteps in (7, 3):
        gf.stepforward(ens, nsteps)
        for p in singles:
            gf.stepforward(p, nsteps)
        for i, p in enumerate(singles):
            a, b = ens[i].sol, p.sol
            assert np.isfinite(b).all(), (i, sizes[i])
            e = rel(a, b)
            print(f"member {i} n={sizes[i]} rel diff {e:.2e}")
            assert e < 1e-14, (i, sizes[i], e)
    assert ens.t == pytest.approx(10 * 0.01)
    assert ens[3].clock.step == 10


def test_ensemble_member_matches_oracle(gf):
    rng = np.random.default_rng(31)
    sizes = [128, 64, 32]
    kws = [_slqg_kw(gf, n, rng, ν=1e-3, nν=2) for n in sizes]
    ens = gf.Ensemble([(gf.SingleLayerQG.Problem, kw) for kw in kws])
    n = 64
    kw = kws[1]
    r = osl.Problem(nx=n, Lx=TWO_PI, beta=10.0, mu=0.01, nu=1e-3, nnu=2, eta=kw["eta"], dt=0.01,
                    stepper="FilteredRK4")
    s = (rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))) * 0.1
    ens[1].sol = s
    r.sol[...] = s
    gf.stepforward(ens, 5)
    off.stepforward(r, 5)
    assert rel(ens[1].sol, r.sol) < 1e-12


def test_mixed_models_forcing_and_alias_fraction(gf):
    rng = np.random.default_rng(32)
    members = []
    # TwoDNS with a fixed forcing table, TwoDNS without dealiasing, SLQG equivalent barotropic
    n0 = 64
    Fh = (rng.standard_normal((n0, n0 // 2 + 1)) + 1j * rng.standard_normal((n0, n0 // 2 + 1))) * 0.01
    Fh[0, 0] = 0
    members.append((gf.TwoDNavierStokes.Problem, dict(nx=n0, ν=1e-3, nν=1, μ=0.05, dt=0.005, stepper="RK4",
                                                       calcF=Fh)))
    members.append((gf.TwoDNavierStokes.Problem, dict(nx=128, ν=1e-7, nν=2, dt=0.005, stepper="RK4",
                                                       aliased_fraction=0.0)))
    members.append((gf.SingleLayerQG.Problem, _slqg_kw(gf, 256, rng, stepper="RK4", dt=0.005,
                                                       deformation_radius=0.7, U=0.2)))
    members.append((gf.SingleLayerQG.Problem, dict(nx=32, β=1.0, U=0.3 * np.tanh(np.linspace(-3, 3, 32)),
                                                   dt=0.005, stepper="RK4")))
    ens = gf.Ensemble(members)
    singles = [ctor(gf.GPU(), **kw) for ctor, kw in members]
    for i, p in enumerate(singles):
        n = p.grid.nx
        s = (rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))) * 0.05
        ens[i].sol = s
        p.sol = s
    gf.stepforward(ens, 6)
    for p in singles:
        gf.stepforward(p, 6)
    for i, p in enumerate(singles):
        a, b = ens[i].sol, p.sol
        assert np.isfinite(b).all(), i
        e = rel(a, b)
        print(f"member {i} rel diff {e:.2e}")
        assert e < 1e-14, (i, e)
    # updatevars on a member view gives the single run's physical fields
    gf.TwoDNavierStokes.updatevars(ens[0])
    gf.TwoDNavierStokes.updatevars(singles[0])
    assert rel(ens[0].vars.ζ, singles[0].vars.ζ) < 1e-14


def test_member_steps_alone_and_handle_rules(gf):
    rng = np.random.default_rng(33)
    kws = [_slqg_kw(gf, n, rng) for n in (64, 128)]
    ens = gf.Ensemble([(gf.SingleLayerQG.Problem, kw) for kw in kws])
    single = gf.SingleLayerQG.Problem(gf.GPU(), **kws[0])
    q0 = _q0(gf, 64, rng)
    gf.SingleLayerQG.set_q(ens[0], q0)
    gf.SingleLayerQG.set_q(single, q0)
    untouched = ens[1].sol
    gf.stepforward(ens[0], 4)            # member 0 alone, through the single-instance kernels
    gf.stepforward(single, 4)
    assert np.array_equal(ens[0].sol, single.sol)
    assert np.array_equal(ens[1].sol, untouched)
    assert ens[0].clock.step == 4 and ens[1].clock.step == 0
    lib = gf._lib.lib
    assert lib.gf_batch_size(ens._ctx) == 2
    buf = np.zeros((64, 33), np.complex128)
    assert lib.gf_get_state(ens._ctx, buf.ctypes.data) == gf._lib.GF_EINVAL
    assert b"gf_ctx_instance" in lib.gf_last_error(ens._ctx)
    m = ctypes.c_void_p()
    assert lib.gf_ctx_instance(ens._ctx, 2, ctypes.byref(m)) == gf._lib.GF_EINVAL
    assert lib.gf_ctx_instance(ens._ctx, 1, ctypes.byref(m)) == 0
    lib.gf_ctx_destroy(m)                # a member view: no effect
    assert np.array_equal(ens[1].sol, untouched)
    # a plain context is an ensemble of one whose only member is itself
    assert lib.gf_batch_size(single._ctx) == 1
    assert lib.gf_ctx_instance(single._ctx, 0, ctypes.byref(m)) == 0 and m.value == single._ctx.value
    assert lib.gf_set_clock(ens._ctx, 1.5, 9) == 0
    assert ens[1].clock.step == 9 and ens[0].clock.t == 1.5


def test_ensemble_ring_forcing_members(gf):
    # the reference's stochastic recipe (examples/twodnavierstokes_stochasticforcing.jl:53-89) for
    # every member: counter-based ring forcing keyed on (seed, step, instance); the batched step
    # reads the step number from a device counter, so interior steps replay from a graph
    rng = np.random.default_rng(34)
    sizes = [64, 128, 32, 256]
    kws = [dict(nx=n, ν=1e-4, nν=2, μ=0.1, dt=0.005, stepper="ETDRK4",
                calcF=gf.RingForcing(6.0, 1.5, 0.1, seed=40 + i, instance=i)) for i, n in enumerate(sizes)]
    ens = gf.Ensemble([(gf.TwoDNavierStokes.Problem, kw) for kw in kws])
    singles = [gf.TwoDNavierStokes.Problem(gf.GPU(), **kw) for kw in kws]
    for i, n in enumerate(sizes):
        s = (rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))) * 0.01
        ens[i].sol = s
        singles[i].sol = s
    for nsteps in (7, 3):
        gf.stepforward(ens, nsteps)
        for p in singles:
            gf.stepforward(p, nsteps)
        for i, p in enumerate(singles):
            e = rel(ens[i].sol, p.sol)
            assert np.isfinite(p.sol).all() and e < 1e-14, (i, sizes[i], e)
    # the work of the last step uses prevsol, kept by the batched step as by the single one
    for i, p in enumerate(singles):
        a = gf.TwoDNavierStokes.energy_work(ens[i])
        b = gf.TwoDNavierStokes.energy_work(p)
        assert a == pytest.approx(b, rel=1e-13, abs=1e-18), i
    # member 1 against the oracle from the same state
    from oracle import twodnavierstokes as o2d
    r = o2d.Problem(nx=128, nu=1e-4, nnu=2, mu=0.1, dt=0.005, stepper="ETDRK4",
                    calcF=("ring", 6.0, 1.5, 0.1, 41, 1))
    r.sol[...] = ens[1].sol
    r.clock.step = ens[1].clock.step
    r.clock.t = ens[1].clock.t
    gf.stepforward(ens, 4)
    off.stepforward(r, 4)
    assert rel(ens[1].sol, r.sol) < 1e-12
    # a member stepped alone leaves the ensemble's shared step number behind
    gf.stepforward(ens[2], 1)
    with pytest.raises(gf._lib.GFError, match="same step"):
        gf.stepforward(ens, 1)


@pytest.mark.parametrize("stepper", ["AB3", "FilteredAB3"])
def test_ensemble_ab3_members(gf, stepper):
    # AB3: every step in full, the forward-Euler start-up for the first three steps, then the
    # Adams-Bashforth step with the carried right-hand sides; calls split across the start-up
    rng = np.random.default_rng(35)
    sizes = [64, 128, 32]
    kws = [_slqg_kw(gf, n, rng, stepper=stepper, dt=0.002, ν=1e-3) for n in sizes]
    ens = gf.Ensemble([(gf.SingleLayerQG.Problem, kw) for kw in kws])
    singles = [gf.SingleLayerQG.Problem(gf.GPU(), **kw) for kw in kws]
    for i, n in enumerate(sizes):
        q0 = _q0(gf, n, rng)
        gf.SingleLayerQG.set_q(ens[i], q0)
        gf.SingleLayerQG.set_q(singles[i], q0)
    for nsteps in (2, 1, 4, 6):
        gf.stepforward(ens, nsteps)
        for p in singles:
            gf.stepforward(p, nsteps)
        for i, p in enumerate(singles):
            e = rel(ens[i].sol, p.sol)
            assert np.isfinite(p.sol).all() and e < 1e-14, (nsteps, i, e)
    assert ens[0].clock.step == 13
    # one member against the oracle over the start-up
    r = osl.Problem(nx=32, Lx=TWO_PI, beta=10.0, mu=0.01, nu=1e-3, eta=kws[2]["eta"], dt=0.002, stepper=stepper)
    ens2 = gf.Ensemble([(gf.SingleLayerQG.Problem, kws[2])])
    s = (rng.standard_normal((32, 17)) + 1j * rng.standard_normal((32, 17))) * 0.05
    ens2[0].sol = s
    r.sol[...] = s
    gf.stepforward(ens2, 5)
    off.stepforward(r, 5)
    assert rel(ens2[0].sol, r.sol) < 1e-12


def test_ensemble_fp32_members(gf):
    # one precision per ensemble; FP32 members against the same problems stepped alone under the
    # built-in schedule (the batched kernels' 8 points per thread)
    rng = np.random.default_rng(36)
    sizes = [64, 256, 32, 128]
    kws = [dict(nx=n, ν=1e-4, nν=2, μ=0.05, dt=0.004, stepper="ETDRK4", T=np.float32) for n in sizes]
    ens = gf.Ensemble([(gf.TwoDNavierStokes.Problem, kw) for kw in kws])
    singles = []
    for i, n in enumerate(sizes):
        p = gf.TwoDNavierStokes.Problem(gf.GPU(), **kws[i])
        gf.set_schedule(p, 8 if n >= 32 else 4, 8 if n >= 32 else 4, 8 if n >= 32 else 4)
        gf.set_splits(p, 0, 0, 0)
        z0 = gf.peakedisotropicspectrum(gf.problem.Grid(n, n, TWO_PI, TWO_PI, 1 / 3, np.float64), min(8, n // 8),
                                        0.02, rng=rng)
        gf.TwoDNavierStokes.set_ζ(ens[i], z0)
        gf.TwoDNavierStokes.set_ζ(p, z0)
        singles.append(p)
    gf.stepforward(ens, 6)
    for p in singles:
        gf.stepforward(p, 6)
    for i, p in enumerate(singles):
        e = rel(ens[i].sol, p.sol)
        assert np.isfinite(p.sol).all() and e < 1e-6, (i, sizes[i], e)
    # and against the FP64 oracle at the FP32 tolerance of the single-problem tests (2e-5)
    from oracle import twodnavierstokes as o2d
    n = 128
    r = o2d.Problem(nx=n, nu=1e-4, nnu=2, mu=0.05, dt=0.004, stepper="ETDRK4")
    r.sol[...] = ens[3].sol.astype(np.complex128)
    gf.stepforward(ens, 5)
    off.stepforward(r, 5)
    assert rel(ens[3].sol.astype(np.complex128), r.sol) < 2e-5
    with pytest.raises(gf._lib.GFError, match="same precision"):
        gf.Ensemble([(gf.TwoDNavierStokes.Problem, kws[0]),
                     (gf.TwoDNavierStokes.Problem, dict(kws[1], T=np.float64))])


def test_ensemble_single_layer_mlqg_members(gf):
    # one-layer MultiLayerQG members (SingleLayerParams: beta and U(y) in the advection, the bottom
    # drag in N) beside a SingleLayerQG member
    from oracle import multilayerqg as oml
    rng = np.random.default_rng(37)

    def mlqg1(dev, **kw):
        return gf.MultiLayerQG.Problem(1, dev, **kw)

    members = []
    for n in (64, 128):
        g = off.TwoDGrid(n, TWO_PI)
        members.append((mlqg1, dict(nx=n, U=0.3 * np.sin(2 * g.y), β=3.0, μ=0.05, ν=1e-3, dt=0.004,
                                    stepper="RK4")))
    members.append((gf.SingleLayerQG.Problem, _slqg_kw(gf, 64, rng, stepper="RK4", dt=0.004)))
    ens = gf.Ensemble(members)
    singles = [ctor(gf.GPU(), **kw) for ctor, kw in members]
    for i, p in enumerate(singles):
        n = p.grid.nx
        s = (rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))) * 0.05
        ens[i].sol = s
        p.sol = s
    gf.stepforward(ens, 5)
    for p in singles:
        gf.stepforward(p, 5)
    for i, p in enumerate(singles):
        e = rel(ens[i].sol, p.sol)
        assert np.isfinite(p.sol).all() and e < 1e-14, (i, e)
    n = 64
    g = off.TwoDGrid(n, TWO_PI)
    r = oml.Problem(1, nx=n, U=0.3 * np.sin(2 * g.y), beta=3.0, mu=0.05, nu=1e-3, dt=0.004, stepper="RK4")
    r.sol[...] = ens[0].sol.reshape(r.sol.shape)
    gf.stepforward(ens, 3)
    off.stepforward(r, 3)
    assert rel(ens[0].sol.reshape(r.sol.shape), r.sol) < 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_persistent_stage_launches_match_three_launches(gf, dtype):
    # gf_batch_persistent: each stage's K1, K2 and K3 over every member in one cooperative
    # launch with grid-wide barriers; the same bodies as the three launches, so the members agree
    # to rounding (1e-14 FP64, 1e-6 FP32 relative L2) after each call
    rng = np.random.default_rng(36)
    sizes = [64, 512, 32, 128, 256, 64]
    members = [(gf.SingleLayerQG.Problem, _slqg_kw(gf, n, rng, T=dtype)) for n in sizes[:3]]
    members.append((gf.TwoDNavierStokes.Problem, dict(nx=128, ν=1e-4, nν=2, μ=0.1, dt=0.005, stepper="FilteredRK4",
                                                       T=dtype, calcF=gf.RingForcing(6.0, 1.5, 0.1, seed=5, instance=3))))
    members.append((gf.SingleLayerQG.Problem, dict(nx=256, β=1.0, U=0.3 * np.tanh(np.linspace(-3, 3, 256)), dt=0.005,
                                                   stepper="FilteredRK4", T=dtype)))
    members.append((gf.TwoDNavierStokes.Problem, dict(nx=64, ν=1e-4, nν=2, μ=0.1, dt=0.005, stepper="FilteredRK4",
                                                       T=dtype, calcF=gf.RingForcing(5.0, 1.5, 0.1, seed=6, instance=5))))
    a = gf.Ensemble(members)
    b = gf.Ensemble(members)
    b.persistent(True)
    for i, n in enumerate(sizes):
        f = _q0(gf, n, rng)
        mod = gf.TwoDNavierStokes if members[i][0] is gf.TwoDNavierStokes.Problem else gf.SingleLayerQG
        for e in (a, b):
            (mod.set_ζ if mod is gf.TwoDNavierStokes else mod.set_q)(e[i], f)
    tol = 1e-14 if dtype == np.float64 else 1e-6
    for nsteps in (7, 3):
        gf.stepforward(a, nsteps)
        gf.stepforward(b, nsteps)
        for i in range(len(sizes)):
            x, y = a[i].sol, b[i].sol
            assert np.isfinite(x).all() and np.isfinite(y).all(), i
            e = rel(y, x)
            print(f"member {i} n={sizes[i]} persistent vs three launches {e:.2e} bitwise {np.array_equal(x, y)}")
            assert e < tol, (i, e)
    b.persistent(False)           # and back
    gf.stepforward(a, 2)
    gf.stepforward(b, 2)
    assert all(rel(b[i].sol, a[i].sol) < tol for i in range(len(sizes)))
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=64, dt=0.01, stepper="RK4")
    with pytest.raises(gf._lib.GFError, match="for ensembles"):
        gf._lib.check(gf._lib.lib.gf_batch_persistent(p._ctx, 1), p._ctx)
