"""gf_diagnostic_async_noalloc (the allocation-free variant of gf_diagnostic_async) through the C
ABI: its values must equal gf_diagnostic_async's bit for bit (same kernels, same block count, same
summation order), over two calls in a row with the second on a smaller problem, whose fewer
partial sums share the static scratch with the first call's leftovers.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _on_device(gf, entry, prob, kind, nout, xi=0.0, nxi=0):
    lib = gf._lib.lib
    buf = lib.gf_device_alloc(8 * nout)
    assert buf
    try:
        n = entry(prob._ctx, int(kind), float(xi), int(nxi), ctypes.c_void_p(buf), int(nout))
        gf._lib.check(min(n, 0), prob._ctx)
        assert n == nout
        prob.synchronize()
        host = np.zeros(nout)
        assert lib.gf_device_copy(host.ctypes.data, ctypes.c_void_p(buf), 8 * nout) == 0
        return host
    finally:
        lib.gf_device_free(ctypes.c_void_p(buf))


def _both(gf, prob, kind, nout, xi=0.0, nxi=0):
    lib = gf._lib.lib
    a = _on_device(gf, lib.gf_diagnostic_async, prob, kind, nout, xi, nxi)
    b = _on_device(gf, lib.gf_diagnostic_async_noalloc, prob, kind, nout, xi, nxi)
    return a, b


def test_noalloc_diagnostic_equals_the_allocating_one(gf):
    L = gf._lib
    rng = np.random.default_rng(17)
    # call 1: 512^2 ring-forced TwoDNavierStokes (1024 blocks of partials), two kinds
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=512, ν=1e-3, nν=1, dt=0.002, stepper="ETDRK4",
                                    calcF=gf.RingForcing(14.0, 1.5, 0.1, seed=3))
    p.sol = _rand(rng, (512, 257)) * 1e-2
    gf.stepforward(p, 2)
    for kind, xi, nxi in ((L.GF_DIAG_ENERGY, 0.0, 0), (L.GF_DIAG_ENSTROPHY_DISSIPATION, 1e-3, 1),
                          (L.GF_DIAG_ENERGY_WORK, 0.0, 0)):
        a, b = _both(gf, p, kind, 1, xi, nxi)
        assert np.array_equal(a, b), (kind, a, b)
        assert np.array_equal(b, np.array(p._diagnostic(kind, xi, nxi, 1)))
    # a 3-layer 256^2 MultiLayerQG problem: 5 energies, 5 fluxes per call
    H, b3 = [0.2, 0.3, 0.5], [-1.0, -1.2, -1.3]
    m = gf.MultiLayerQG.Problem(3, gf.GPU(), nx=256, f0=1.0, H=H, b=b3, U=[1.0, 0.3, 0.0], dt=0.01,
                                stepper="RK4")
    m.sol = _rand(rng, (3, 256, 129)) * 1e-2
    for kind in (L.GF_DIAG_LAYER_ENERGIES, L.GF_DIAG_FLUXES):
        a, b = _both(gf, m, kind, 5)
        assert np.array_equal(a, b), (kind, a, b)
    # call 2: smaller problems (fewer blocks than the leftovers in the scratch)
    q = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=64, dt=0.01, stepper="RK4")
    q.sol = _rand(rng, (64, 33))
    a, b = _both(gf, q, L.GF_DIAG_ENERGY, 1)
    assert np.array_equal(a, b)
    assert np.array_equal(b, np.array(q._diagnostic(L.GF_DIAG_ENERGY, nout=1)))
    m2 = gf.MultiLayerQG.Problem(2, gf.GPU(), nx=32, f0=1.0, H=[0.4, 0.6], b=[-1.0, -1.2], U=[0.5, 0.0],
                                 dt=0.01, stepper="RK4")
    m2.sol = _rand(rng, (2, 32, 17))
    for kind in (L.GF_DIAG_LAYER_ENERGIES, L.GF_DIAG_FLUXES):
        a, b = _both(gf, m2, kind, 3)
        assert np.array_equal(a, b), (kind, a, b)


def test_noalloc_diagnostic_rejects_bad_arguments(gf):
    lib = gf._lib.lib
    q = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=32, dt=0.01, stepper="RK4")
    buf = lib.gf_device_alloc(8 * 4)
    try:
        assert lib.gf_diagnostic_async_noalloc(q._ctx, gf._lib.GF_DIAG_ENERGY, 0.0, 0, None, 1) == gf._lib.GF_EINVAL
        m = gf.MultiLayerQG.Problem(3, gf.GPU(), nx=32, f0=1.0, H=[0.2, 0.3, 0.5], b=[-1.0, -1.2, -1.3], dt=0.01,
                                    stepper="RK4")
        # 5 values do not fit in 4
        rc = lib.gf_diagnostic_async_noalloc(m._ctx, gf._lib.GF_DIAG_FLUXES, 0.0, 0, ctypes.c_void_p(buf), 4)
        assert rc == gf._lib.GF_EINVAL and b"too small" in lib.gf_last_error(m._ctx)
        # a kind the model does not provide
        rc = lib.gf_diagnostic_async_noalloc(q._ctx, gf._lib.GF_DIAG_FLUXES, 0.0, 0, ctypes.c_void_p(buf), 4)
        assert rc == gf._lib.GF_EINVAL
    finally:
        lib.gf_device_free(ctypes.c_void_p(buf))
