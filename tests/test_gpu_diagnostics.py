"""Device diagnostics (gf_diagnostic) against the oracle's restatement of the reference's
diagnostics and against the reference tests' closed forms.

Tolerance: 1e-12 relative (one Parseval reduction in FP64, summed in a different order).
States are raw (their aliased band filled), as the reference's diagnostics read sol as is.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import fourierflows as off
from oracle import multilayerqg as oml
from oracle import singlelayerqg as osl
from oracle import twodnavierstokes as o2d

pytestmark = pytest.mark.gpu


def close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(abs(b), 1e-300)


def _rand(rng, shape, amp=1.0):
    return amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def test_twodns_diagnostics_match_oracle(gf):
    n = 64
    rng = np.random.default_rng(60)
    kw = dict(nx=n, Lx=2 * math.pi, ny=n, Ly=3 * math.pi, nu=2e-3, nnu=2, mu=0.1, nmu=-1, dt=0.01)
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, Lx=kw["Lx"], ny=n, Ly=kw["Ly"], ν=kw["nu"], nν=kw["nnu"],
                                    μ=kw["mu"], nμ=kw["nmu"], dt=0.01, stepper="RK4")
    r = o2d.Problem(stepper="RK4", **kw)
    s = _rand(rng, (n, n // 2 + 1))
    p.sol = s
    r.sol[...] = s
    T, O = gf.TwoDNavierStokes, o2d
    for a, b in [(T.energy(p), O.energy(r)), (T.enstrophy(p), O.enstrophy(r)), (T.palinstrophy(p), O.palinstrophy(r)),
                 (T.energy_dissipation_hyperviscosity(p), O.energy_dissipation_hyperviscosity(r)),
                 (T.energy_dissipation_hypoviscosity(p), O.energy_dissipation_hypoviscosity(r)),
                 (T.enstrophy_dissipation_hyperviscosity(p), O.enstrophy_dissipation_hyperviscosity(r)),
                 (T.enstrophy_dissipation_hypoviscosity(p), O.enstrophy_dissipation_hypoviscosity(r)),
                 (T.energy_dissipation(p, 0.3, 3), O.energy_dissipation(r, 0.3, 3))]:
        assert close(a, b), (a, b)


def test_twodns_work_fixed_and_stochastic(gf):
    n = 64
    rng = np.random.default_rng(61)
    F = _rand(rng, (n, n // 2 + 1), 0.01)
    F[0, 0] = 0
    p = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, dt=0.01, stepper="RK4", calcF=F)
    r = o2d.Problem(nx=n, nu=1e-3, dt=0.01, stepper="RK4", calcF=F)
    s = _rand(rng, (n, n // 2 + 1), 0.1)
    p.sol = s
    r.sol[...] = s
    r.vars.Fh[...] = F
    assert close(gf.TwoDNavierStokes.energy_work(p), o2d.energy_work(r))
    assert close(gf.TwoDNavierStokes.enstrophy_work(p), o2d.enstrophy_work(r))
    # stochastic ring forcing: (prevsol + sol)/2 and the last step's realisation
    ps = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=1e-3, μ=0.1, dt=0.01, stepper="ETDRK4",
                                     calcF=gf.RingForcing(8.0, 1.5, 0.1, seed=5))
    rs = o2d.Problem(nx=n, nu=1e-3, mu=0.1, dt=0.01, stepper="ETDRK4", calcF=("ring", 8.0, 1.5, 0.1, 5))
    ps.sol = s
    rs.sol[...] = s
    gf.stepforward(ps, 3)
    off.stepforward(rs, 3)
    assert close(gf.TwoDNavierStokes.energy_work(ps), o2d.energy_work(rs), 1e-10)
    assert close(gf.TwoDNavierStokes.enstrophy_work(ps), o2d.enstrophy_work(rs), 1e-10)


@pytest.mark.parametrize("ell", [math.inf, 1.23])
def test_slqg_diagnostics_match_oracle_and_closed_forms(gf, ell):
    nx, Lx, ny, Ly = 64, 2 * np.pi, 64, 3 * np.pi
    k0, l0 = 2 * np.pi / Lx, 2 * np.pi / Ly
    grid = off.TwoDGrid(nx, Lx, ny, Ly)
    x, y = grid.gridpoints()
    eta = np.cos(10 * k0 * x) * np.cos(10 * l0 * y)
    p = gf.SingleLayerQG.Problem(gf.GPU(), nx=nx, Lx=Lx, ny=ny, Ly=Ly, deformation_radius=ell, eta=eta,
                                 ν=1e-3, nν=2, μ=0.05, dt=0.01, stepper="ForwardEuler")
    psi0 = np.sin(2 * k0 * x) * np.cos(2 * l0 * y) + 2 * np.sin(k0 * x) * np.cos(3 * l0 * y)
    q0 = (-((2 * k0) ** 2 + (2 * l0) ** 2) * np.sin(2 * k0 * x) * np.cos(2 * l0 * y)
          - (k0 ** 2 + (3 * l0) ** 2) * 2 * np.sin(k0 * x) * np.cos(3 * l0 * y))
    if ell != math.inf:
        q0 = q0 - psi0 / ell ** 2
    gf.SingleLayerQG.set_q(p, q0)
    S = gf.SingleLayerQG
    # test/test_singlelayerqg.jl:406-474
    assert close(S.kinetic_energy(p), float(Fraction("29/9")), 1e-13)
    if ell == math.inf:
        assert S.potential_energy(p) == 0
        assert close(S.enstrophy(p), float(Fraction("10885/648")), 1e-13)
    else:
        assert close(S.potential_energy(p), 5 / (8 * ell ** 2), 1e-13)
    # a raw random state against the oracle
    r = osl.Problem(nx=nx, Lx=Lx, ny=ny, Ly=Ly, deformation_radius=ell, eta=eta, nu=1e-3, nnu=2, mu=0.05,
                    dt=0.01, stepper="ForwardEuler")
    s = _rand(np.random.default_rng(62), (ny, nx // 2 + 1))
    p.sol = s
    r.sol[...] = s
    pairs = [(S.kinetic_energy(p), osl.kinetic_energy(r)), (S.potential_energy(p), osl.potential_energy(r)),
             (S.energy(p), osl.energy(r)), (S.enstrophy(p), osl.enstrophy(r))]
    for a, b in pairs:
        assert close(a, b), (a, b)
    if ell == math.inf:
        # src/singlelayerqg.jl:617-715 restated on the host from the same state
        g = r.grid
        nrm = 1 / (g.Lx * g.Ly)
        iK = g.invKrsq
        K2 = g.Krsq
        h = 1e-3 * K2 ** 1 * np.abs(s) ** 2; h[0, 0] = 0
        assert close(S.energy_dissipation(p), nrm * off.parsevalsum(h, g))
        h = 1e-3 * K2 ** 2 * np.abs(s) ** 2; h[0, 0] = 0
        assert close(S.enstrophy_dissipation(p), nrm * off.parsevalsum(h, g))
        h = 0.05 * iK * np.abs(s) ** 2; h[0, 0] = 0
        assert close(S.energy_drag(p), nrm * off.parsevalsum(h, g))
        h = 0.05 * np.abs(s) ** 2; h[0, 0] = 0
        assert close(S.enstrophy_drag(p), nrm * off.parsevalsum(h, g))
    else:
        with pytest.raises(gf._lib.GFError):
            S.energy_dissipation(p)


def test_mlqg_energies_closed_form_and_oracle(gf):
    # test/test_multilayerqg.jl:502-536: KE1 = 1/4, KE2 = 3/16, PE = 1/160
    n = 64
    p = gf.MultiLayerQG.Problem(2, gf.GPU(), nx=n, Lx=2 * np.pi, f0=1, H=[2.5, 7.5], b=[-1.0, -1.5], dt=0.01,
                                stepper="RK4")
    r = oml.Problem(2, nx=n, Lx=2 * np.pi, f0=1, H=[2.5, 7.5], b=[-1.0, -1.5])
    x, y = r.grid.gridpoints()
    psi = np.array([np.cos(2 * x) * np.cos(2 * y), 0.5 * np.cos(2 * x) * np.cos(2 * y)])
    oml.set_psi(r, psi)
    p.sol = r.sol
    KE, PE = gf.MultiLayerQG.energies(p)
    assert abs(KE[0] - 1 / 4) < 1e-13 and abs(KE[1] - 3 / 16) < 1e-13 and abs(PE[0] - 1 / 160) < 1e-13
    # three layers, random raw state
    p3 = gf.MultiLayerQG.Problem(3, gf.GPU(), nx=n, Lx=2 * np.pi, f0=1, H=[0.2, 0.3, 0.5], b=[0, -0.2, -0.27],
                                 dt=0.01, stepper="RK4")
    r3 = oml.Problem(3, nx=n, Lx=2 * np.pi, f0=1, H=[0.2, 0.3, 0.5], b=[0, -0.2, -0.27])
    s = _rand(np.random.default_rng(63), (3, n, n // 2 + 1))
    p3.sol = s
    r3.sol[...] = s
    KE, PE = gf.MultiLayerQG.energies(p3)
    KEr, PEr = oml.energies(r3)
    for a, b in zip(list(KE) + list(PE), list(KEr) + list(PEr)):
        assert close(a, b), (a, b)


def test_diagnostic_series_on_device_equals_synchronous_values(gf):
    # [FF] Diagnostic / increment!: a series recorded on the device (gf_diagnostic_async) holds
    # the values gf_diagnostic returns after each step, bit for bit
    n, nsteps = 64, 12
    T = gf.TwoDNavierStokes
    kw = dict(nx=n, ν=1e-4, nν=2, μ=0.1, dt=0.005, stepper="RK4")
    s = _rand(np.random.default_rng(64), (n, n // 2 + 1), 0.5)
    s[0, 0] = 0
    a = T.Problem(gf.GPU(), calcF=gf.RingForcing(6.0, 1.5, 0.1, seed=9), **kw)
    b = T.Problem(gf.GPU(), calcF=gf.RingForcing(6.0, 1.5, 0.1, seed=9), **kw)
    a.sol = s
    b.sol = s
    fns = [T.energy, T.enstrophy, T.energy_dissipation_hyperviscosity, T.energy_dissipation_hypoviscosity]
    diags = [gf.Diagnostic(f, a, nsteps=nsteps) for f in fns]
    work = gf.Diagnostic(T.energy_work, a, freq=3, nsteps=nsteps)
    assert work.data[0] == 0          # no forcing realised before the first step (vars.Fh = 0)
    gf.stepforward(a, diags + [work], nsteps)
    want = [[f(b)] for f in fns]
    wwork = []
    for k in range(1, nsteps + 1):
        gf.stepforward(b, 1)
        for w, f in zip(want, fns):
            w.append(f(b))
        if k % 3 == 0:
            wwork.append(T.energy_work(b))
    for d, w in zip(diags, want):
        assert d.i == nsteps + 1
        assert np.array_equal(d.data, np.array(w))
        assert np.allclose(d.t, 0.005 * np.arange(nsteps + 1), rtol=0, atol=1e-15)
    assert np.array_equal(work.data[1:], np.array(wwork))
    assert list(work.steps[:work.i]) == [0, 3, 6, 9, 12]


def test_diagnostic_series_grows_and_unpacks_layer_energies(gf):
    n = 64
    p = gf.MultiLayerQG.Problem(3, gf.GPU(), nx=n, Lx=2 * np.pi, f0=1, H=[0.2, 0.3, 0.5], b=[0, -0.2, -0.27],
                                dt=0.005, stepper="ETDRK4", U=[0.1, 0.02, 0.0], μ=0.01)
    s = _rand(np.random.default_rng(65), (3, n, n // 2 + 1), 1e-3)
    p.sol = s
    d = gf.Diagnostic(gf.MultiLayerQG.energies, p, nsteps=2)        # ndata 3, grown twice below
    gf.stepforward(p, d, 9)
    assert d.i == 10 and d.ndata >= 10
    KE, PE = d.data[-1]
    KEn, PEn = gf.MultiLayerQG.energies(p)
    assert np.array_equal(KE, KEn) and np.array_equal(PE, PEn)
    assert len(KE) == 3 and len(PE) == 2


def test_stochastic_energy_budget_on_device(gf):
    """test/test_twodnavierstokes.jl:25-72 on the device: n = 256, dt = 0.005, nu = 1e-7 (nnu = 2),
    mu = 0.1, kf = 12, dkf = 2, eps = 0.1, 200 steps of RK4 from rest; dE/dt = W + D + R to 1e-3
    (Stratonovich work with prevsol).  The forcing is this library's ring forcing with complex
    Gaussian amplitudes (the reference draws unit-modulus phases); the closure holds for both."""
    n, dt, nt = 256, 0.005, 200
    T = gf.TwoDNavierStokes
    p = T.Problem(gf.GPU(), nx=n, Lx=2 * math.pi, ν=1e-7, nν=2, μ=0.1, nμ=0, dt=dt, stepper="RK4",
                  calcF=gf.RingForcing(12.0, 2.0, 0.1, seed=1234))
    T.set_ζ(p, np.zeros((n, n)))
    E = gf.Diagnostic(T.energy, p, nsteps=nt)
    D = gf.Diagnostic(T.energy_dissipation_hyperviscosity, p, nsteps=nt)
    R = gf.Diagnostic(T.energy_dissipation_hypoviscosity, p, nsteps=nt)
    W = gf.Diagnostic(T.energy_work, p, nsteps=nt)
    gf.stepforward(p, [E, D, W, R], nt)
    Ev, Dv, Rv, Wv = (np.asarray(x.data) for x in (E, D, R, W))
    dEdt = (Ev[1:] - Ev[:-1]) / dt
    computed = Wv[1:] + Dv[:-1] + Rv[:-1]
    assert np.all(np.isfinite(Ev)) and Ev[-1] > 0
    assert np.linalg.norm(dEdt - computed) / np.linalg.norm(dEdt) < 1e-3


def test_energy_budget_series_matches_oracle(gf):
    # the same budget series at 64^2 against the oracle's (60 steps, 1e-10 relative)
    n, dt, nt = 64, 0.005, 60
    T = gf.TwoDNavierStokes
    p = T.Problem(gf.GPU(), nx=n, ν=1e-5, nν=2, μ=0.1, nμ=0, dt=dt, stepper="RK4",
                  calcF=gf.RingForcing(6.0, 1.5, 0.1, seed=1234))
    r = o2d.Problem(nx=n, nu=1e-5, nnu=2, mu=0.1, nmu=0, dt=dt, stepper="RK4", calcF=("ring", 6.0, 1.5, 0.1, 1234))
    s = _rand(np.random.default_rng(66), (n, n // 2 + 1), 0.05)
    s[0, 0] = 0
    p.sol = s
    r.sol[...] = s
    fns = [(T.energy, o2d.energy), (T.energy_dissipation_hyperviscosity, o2d.energy_dissipation_hyperviscosity),
           (T.energy_dissipation_hypoviscosity, o2d.energy_dissipation_hypoviscosity)]
    dg = [gf.Diagnostic(f, p, nsteps=nt) for f, _ in fns]
    do = [off.Diagnostic(g, r) for _, g in fns]
    gf.stepforward(p, dg, nt)
    off.stepforward(r, nt, diags=do)
    for a, b in zip(dg, do):
        assert np.linalg.norm(np.asarray(a.data) - np.asarray(b.data)) <= 1e-10 * np.linalg.norm(b.data)
