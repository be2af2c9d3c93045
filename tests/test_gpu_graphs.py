"""Graph replay of gf_step against plain launches: every step but a call's last replays one
captured HIP graph.  The two paths run the same kernels with the same arguments (the ring
forcing's step number is read from device memory inside the graph), so the states must agree
bit for bit after every call, with eager work and host synchronisation between the replays."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(ctor, **kw):
    a, b = ctor(**kw), ctor(**kw)
    a.use_graphs(True)
    b.use_graphs(False)
    return a, b


def _check_calls(gf, a, b, calls, diag=None):
    for n in calls:
        gf.stepforward(a, n)
        gf.stepforward(b, n)
        sa, sb = a.sol, b.sol                    # eager reads (and waits) between replays
        assert np.isfinite(sa).all()
        assert np.array_equal(sa, sb), (n, np.abs(sa - sb).max())
        if diag is not None:
            assert diag(a) == diag(b)
        assert a.clock.step == b.clock.step


@pytest.mark.parametrize("n,stepper", [(64, "ETDRK4"), (256, "ETDRK4"), (128, "RK4")])
def test_ring_forced_twodns_graph_equals_launches(gf, n, stepper):
    T = gf.TwoDNavierStokes
    mk = lambda: T.Problem(gf.GPU(), nx=n, ν=1e-6, nν=2, μ=0.1, dt=0.005, stepper=stepper,
                           calcF=gf.RingForcing(6.0, 1.5, 0.1, seed=11))
    a, b = _pair(mk)
    rng = np.random.default_rng(70)
    z = rng.standard_normal((n, n))
    T.set_ζ(a, z)
    T.set_ζ(b, z)
    _check_calls(gf, a, b, [5, 1, 7, 2, 12], diag=T.energy_work)
    a.set_clock(0.5, 100)                        # the forcing counter follows the clock
    b.set_clock(0.5, 100)
    _check_calls(gf, a, b, [4, 3], diag=T.energy_work)


@pytest.mark.parametrize("stepper", ["AB3", "FilteredAB3"])
def test_ab3_graph_after_start_steps(gf, stepper):
    S = gf.SingleLayerQG
    n = 64
    mk = lambda: S.Problem(gf.GPU(), nx=n, β=5.0, μ=0.01, dt=0.002, stepper=stepper)
    a, b = _pair(mk)
    rng = np.random.default_rng(71)
    q = 0.1 * rng.standard_normal((n, n))
    S.set_q(a, q)
    S.set_q(b, q)
    _check_calls(gf, a, b, [1, 4, 6, 1, 9])


def test_slqg_topography_and_mlqg_graphs(gf):
    S, M = gf.SingleLayerQG, gf.MultiLayerQG
    n = 128
    g = gf.problem.Grid(n, n, 2 * math.pi, 2 * math.pi, 1 / 3, np.float64)
    eta = gf.utils.gaussian_bump(g, 0.1, 2 * math.pi / 10, (0.3, -0.2))
    mk = lambda: S.Problem(gf.GPU(), nx=n, β=10.0, μ=0.01, eta=eta, U=0.2, dt=0.005, stepper="FilteredRK4")
    a, b = _pair(mk)
    q = 0.5 * np.random.default_rng(72).standard_normal((n, n))
    S.set_q(a, q)
    S.set_q(b, q)
    _check_calls(gf, a, b, [3, 8, 1])
    mk = lambda: M.Problem(3, gf.GPU(), nx=64, f0=1, β=5.0, H=[0.2, 0.3, 0.5], b=[0, -0.2, -0.27],
                           U=[1.0, 0.2, 0.0], μ=0.005, dt=0.005, stepper="ETDRK4")
    a, b = _pair(mk)
    s = 1e-2 * np.random.default_rng(73).standard_normal((3, 64, 64))
    M.set_q(a, s)
    M.set_q(b, s)
    _check_calls(gf, a, b, [6, 2, 5])


def test_ensemble_graph_equals_launches(gf):
    S = gf.SingleLayerQG
    members = [(S.Problem, dict(nx=n, β=10.0, μ=0.01, dt=0.01, stepper="FilteredRK4")) for n in (64, 128, 64, 256)]
    ea, eb = gf.Ensemble(members), gf.Ensemble(members)     # graphs are on by default for ensembles
    eb.use_graphs(False)
    rng = np.random.default_rng(74)
    for pa, pb in zip(ea.problems, eb.problems):
        q = 0.3 * rng.standard_normal((pa.grid.ny, pa.grid.nx))
        S.set_q(pa, q)
        S.set_q(pb, q)
    for n in (4, 1, 9):
        gf.stepforward(ea, n)
        gf.stepforward(eb, n)
        for pa, pb in zip(ea.problems, eb.problems):
            assert np.array_equal(pa.sol, pb.sol)
    # a member stepped alone replays its own graph
    ea.problems[1].use_graphs(True)
    gf.stepforward(ea.problems[1], 5)
    gf.stepforward(eb.problems[1], 5)
    assert np.array_equal(ea.problems[1].sol, eb.problems[1].sol)
