"""Snapshots (gf_snapshot) and the Output mirror on the device: every snapshot equals the state
at the moment it was queued, while later steps overlap its copy."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(gf, n=128, seed=0):
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-4, nν=2, dt=0.005, stepper="ETDRK4")
    rng = np.random.default_rng(seed)
    gf.TwoDNavierStokes.set_ζ(p, rng.standard_normal((n, n)))
    return p


def test_snapshot_equals_state_when_queued(gf):
    lib = gf._lib.lib
    p = _problem(gf)
    nbytes = p._sshape_bytes()
    bufs = [lib.gf_host_alloc(nbytes) for _ in range(3)]
    assert all(bufs)
    try:
        views = [np.ctypeslib.as_array((ctypes.c_char * nbytes).from_address(b)).view(np.complex128).reshape(p._sshape)
                 for b in bufs]
        want, tickets = [], []
        for k in range(3):                      # the third snapshot reuses the first staging buffer
            gf.stepforward(p, 4)
            want.append(p.sol)
            t = lib.gf_snapshot(p._ctx, ctypes.c_void_p(bufs[k]))
            assert t == k
            tickets.append(t)
            gf.stepforward(p, 7)                # queued behind the snapshot, overlapping its copy
        for k in range(3):
            assert lib.gf_snapshot_wait(p._ctx, tickets[k]) == 0
            assert np.array_equal(views[k], want[k])
        assert lib.gf_snapshot_wait(p._ctx, -1) == 0
    finally:
        for b in bufs:
            lib.gf_host_free(ctypes.c_void_p(b))


def test_snapshot_of_pageable_memory_and_layers(gf):
    lib = gf._lib.lib
    n = 64
    p = gf.MultiLayerQG.Problem(3, gf.GPU(), nx=n, f0=1, H=[0.2, 0.3, 0.5], b=[0, -0.2, -0.27], U=[0.1, 0.0, 0.0],
                                dt=0.005, stepper="RK4")
    rng = np.random.default_rng(1)
    s = 1e-3 * (rng.standard_normal((3, n, n // 2 + 1)) + 1j * rng.standard_normal((3, n, n // 2 + 1)))
    p.sol = s
    gf.stepforward(p, 3)
    want = p.sol
    out = np.zeros_like(want)
    t = lib.gf_snapshot(p._ctx, out.ctypes.data_as(ctypes.c_void_p))
    assert t >= 0
    assert lib.gf_snapshot_wait(p._ctx, t) == 0
    assert np.array_equal(out, want)


def test_output_file_holds_the_reference_keys(gf, tmp_path):
    T = gf.TwoDNavierStokes
    p, q = _problem(gf, seed=3), _problem(gf, seed=3)
    path = tmp_path / "decaying.npz"
    get_zeta = lambda prob: (T.updatevars(prob), prob.vars.ζ.copy())[1]
    out = gf.Output(p, path, ("sol", gf.get_sol), ("zeta", get_zeta))
    gf.saveproblem(out)
    E = gf.Diagnostic(T.energy, p, nsteps=30)
    gf.saveoutput(out)
    for _ in range(3):
        gf.stepforward(p, E, 10)
        gf.saveoutput(out)
    out.close()
    gf.savediagnostic(E, "energy", path)
    f = np.load(path)
    assert int(f["grid/nx"]) == 128 and float(f["grid/Lx"]) == pytest.approx(2 * np.pi)
    assert float(f["clock/dt"]) == 0.005 and float(f["params/ν"]) == 1e-4
    for step in (0, 10, 20, 30):
        assert float(f[f"snapshots/t/{step}"]) == pytest.approx(0.005 * step, abs=1e-15)
        s = q.sol
        assert np.array_equal(f[f"snapshots/sol/{step}"], s)
        T.updatevars(q)
        assert np.array_equal(f[f"snapshots/zeta/{step}"], q.vars.ζ)
        gf.stepforward(q, 10)
    assert np.array_equal(f["diagnostics/energy/steps"], np.arange(31))
    assert f["diagnostics/energy/data"].shape == (31,)
    assert float(f["diagnostics/energy/data"][0]) == T.energy(_problem(gf, seed=3))
