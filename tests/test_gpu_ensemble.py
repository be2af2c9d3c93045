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
