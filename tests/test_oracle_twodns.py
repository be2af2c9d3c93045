"""Pins the oracle's TwoDNavierStokes and FourierFlows restatement (CPU only).

Each check is one the reference's own tests make, or a closed form /
invariant that a plausible mistake (dropped term, wrong sign, transposed axis,
wrong normalisation) would fail.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import fourierflows as ff
from oracle import twodnavierstokes as T2

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_values.json")))


def _manufactured(stepper, dt, nt, n=64):
    """test/test_twodnavierstokes.jl:221-257 restated: psi, zeta and the forcing that
    makes zeta a steady solution (nu = 1e-2); stepping must keep zeta to 1e-13."""
    g = ff.TwoDGrid(n, 2 * np.pi)
    x, y = g.gridpoints()
    nu = 1e-2
    zf = -8 * np.sin(2 * x) * np.cos(2 * y) - 20 * np.sin(x) * np.cos(3 * y)
    Ff = -(nu * (64 * np.sin(2 * x) * np.cos(2 * y) + 200 * np.sin(x) * np.cos(3 * y))
           + 8 * (np.cos(x) * np.cos(3 * y) * np.sin(2 * x) * np.sin(2 * y)
                  - 3 * np.cos(2 * x) * np.cos(2 * y) * np.sin(x) * np.sin(3 * y)))
    prob = T2.Problem(nx=n, Lx=2 * np.pi, nu=nu, nnu=1, dt=dt, stepper=stepper, calcF=g.rfft(Ff))
    T2.set_zeta(prob, zf)
    ff.stepforward(prob, nt)
    T2.updatevars(prob)
    return np.linalg.norm(prob.vars.zeta - zf) / np.linalg.norm(zf)


@pytest.mark.parametrize("stepper,dt,nt", [("ForwardEuler", 5e-4, 400), ("ETDRK4", 5e-3, 100),
                                           ("RK4", 5e-3, 100), ("FilteredRK4", 5e-3, 100),
                                           ("FilteredETDRK4", 5e-3, 100)])
def test_manufactured_steady_solution(stepper, dt, nt):
    assert _manufactured(stepper, dt, nt) < GOLD["rtol_manufactured"]["value"]


def test_manufactured_catches_sign_error(monkeypatch):
    # a planted mistake in the oracle (wrong sign of the kr term) must fail the test above
    orig = T2.calcN_advection

    def bad(N, sol, t, clock, vars, params, grid):
        orig(N, sol, t, clock, vars, params, grid)
        N += 2j * grid.Kr * grid.rfft(vars.u * vars.zeta)

    monkeypatch.setattr(T2, "calcN_advection", bad)
    assert _manufactured("ForwardEuler", 5e-4, 50) > 1e-6


def test_energy_enstrophy_palinstrophy_closed_forms():
    """test/test_twodnavierstokes.jl:259-317 (restated on 128x128, Lx=2pi, Ly=3pi)."""
    nx, Lx, ny, Ly = 128, 2 * np.pi, 128, 3 * np.pi
    prob = T2.Problem(nx=nx, Lx=Lx, ny=ny, Ly=Ly, stepper="ForwardEuler")
    x, y = prob.grid.gridpoints()
    k0, l0 = 2 * np.pi / Lx, 2 * np.pi / Ly
    z0 = (-((2 * k0) ** 2 + (2 * l0) ** 2) * np.sin(2 * k0 * x) * np.cos(2 * l0 * y)
          - (k0 ** 2 + (3 * l0) ** 2) * 2 * np.sin(k0 * x) * np.cos(3 * l0 * y))
    T2.set_zeta(prob, z0)
    T2.updatevars(prob)
    for fn, key in ((T2.energy, "twodns_energy_128x126_2pi_3pi"), (T2.enstrophy, "twodns_enstrophy_128x126_2pi_3pi"),
                    (T2.palinstrophy, "twodns_palinstrophy_128x126_2pi_3pi")):
        exact = float(Fraction(GOLD[key]["value"]))
        assert abs(fn(prob) / exact - 1) < 1e-13


def test_fft_convention_matches_library_dft():
    # forward unnormalised, inverse normalised (docs/src/index.md:25-35): check against
    # an explicit DFT sum on a tiny grid, layout (ny, nx) with kr along the last axis
    g = ff.TwoDGrid(8, 2 * np.pi, ny=4, Ly=3.0)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((4, 8))
    F = g.rfft(f)
    yy, xx = np.meshgrid(np.arange(4), np.arange(8), indexing="ij")
    for l in range(4):
        for k in range(5):
            ref = np.sum(f * np.exp(-2j * np.pi * (k * xx / 8 + l * yy / 4)))
            assert abs(F[l, k] - ref) < 1e-12
    assert np.allclose(g.irfft(F), f, atol=1e-14)


def test_grid_wavenumbers_and_dealias_rule():
    g = ff.TwoDGrid(2048, 2 * np.pi)
    assert g.kr[1] == 1.0 and g.l[1] == 1.0 and g.l[1024] == -1024.0 and g.l[-1] == -1.0
    # 2/3 rule as FourierFlows implements it: kr >= 682 and l in 682..1365 are zeroed
    assert (g.kralias.start, g.kralias.stop) == (682, 1025)
    assert (g.lalias.start, g.lalias.stop) == (682, 1366)
    m = g.dealias_mask()
    assert m[0, 681] == 1 and m[0, 682] == 0 and m[681, 0] == 1 and m[682, 0] == 0 and m[1366, 0] == 1
    g0 = ff.TwoDGrid(64, 2 * np.pi, aliased_fraction=0)
    assert g0.dealias_mask().all()


def test_parsevalsum_matches_physical_mean():
    g = ff.TwoDGrid(32, 2 * np.pi, ny=16, Ly=5.0)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((16, 32))
    fh = g.rfft(f)
    # Nyquist column counted twice by parsevalsum: compare on a field with no Nyquist content
    fh[:, -1] = 0
    fh[8, :] = 0
    f = g.irfft(fh)
    fh = g.rfft(f)
    lhs = ff.parsevalsum(np.abs(fh) ** 2, g)
    rhs = np.sum(f ** 2) * g.dx * g.dy
    assert abs(lhs / rhs - 1) < 1e-12


def test_etdrk4_exact_for_linear_problem():
    # N = 0: ETDRK4 gives sol(t) = exp(L t) sol0 (closed form); decaying problem with no advection
    g = ff.TwoDGrid(32, 2 * np.pi)
    prob = T2.Problem(nx=32, Lx=2 * np.pi, nu=0.05, nnu=1, mu=0.2, nmu=0, dt=0.01, stepper="ETDRK4")
    prob.eqn.calcN = lambda N, sol, *a: N.__setitem__(Ellipsis, 0)
    rng = np.random.default_rng(1)
    s0 = (rng.standard_normal((32, 17)) + 1j * rng.standard_normal((32, 17))) * g.dealias_mask()
    prob.sol[...] = s0
    ff.stepforward(prob, 10)
    assert np.allclose(prob.sol, np.exp(prob.eqn.L * 0.1) * s0, rtol=1e-13, atol=1e-15)


def test_etd_coefficients_against_phi_functions():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 50
    dt = 0.005
    for Lv in (0.0, -1.0, -37.5, -2000.0, 3.0, -5 + 4j):
        L = np.array([Lv])
        z, a, b, G = ff.etd_coefficients(dt, L if isinstance(Lv, complex) else L.real)
        h = mpmath.mpf(dt)
        zz = h * mpmath.mpc(Lv)
        if zz == 0:
            ex = [h / 2, h / 6, h / 6, h / 6]
        else:
            e = mpmath.exp(zz)
            ex = [h * (mpmath.exp(zz / 2) - 1) / zz,
                  h * (-4 - zz + e * (4 - 3 * zz + zz ** 2)) / zz ** 3,
                  h * (2 + zz + e * (zz - 2)) / zz ** 3,
                  h * (-4 - 3 * zz - zz ** 2 + e * (4 - zz)) / zz ** 3]
        for got, want in zip((z, a, b, G), ex):
            want = complex(want)
            assert abs(complex(got[0]) - want) <= 1e-13 * abs(want)


@pytest.mark.parametrize("stepper,order", [("ETDRK4", 4), ("RK4", 4)])
def test_time_order_of_accuracy(stepper, order):
    # dt-halving on a smooth decaying flow: error ratio ~ 2^order
    n = 32
    g = ff.TwoDGrid(n, 2 * np.pi)
    x, y = g.gridpoints()
    z0 = np.sin(2 * x) * np.cos(y) + 0.5 * np.cos(3 * x + y) + 0.3 * np.sin(x - 2 * y)

    def run(dt, T=0.4):
        p = T2.Problem(nx=n, Lx=2 * np.pi, nu=0.02, nnu=1, dt=dt, stepper=stepper)
        T2.set_zeta(p, z0)
        ff.stepforward(p, int(round(T / dt)))
        # the aliased band holds the last step's N-combination (P1), not a converging
        # quantity: compare the dealiased state
        return p.sol * g.dealias_mask()

    ref = run(0.4 / 256)
    e1 = np.linalg.norm(run(0.4 / 8) - ref)
    e2 = np.linalg.norm(run(0.4 / 16) - ref)
    assert abs(np.log2(e1 / e2) - order) < 0.4


def test_stochastic_energy_budget_closes():
    """test/test_twodnavierstokes.jl:25-72 restated with the counter-based generator:
    dE/dt = W + D + R (Stratonovich work with prevsol) within 1e-3 (RK4, 256^2 -> 64^2)."""
    n, dt, nsteps = 64, 0.005, 60
    nu, nnu, mu, nmu = 1e-5, 2, 0.1, 0
    prob = T2.Problem(nx=n, Lx=2 * np.pi, nu=nu, nnu=nnu, mu=mu, nmu=nmu, dt=dt, stepper="RK4",
                      calcF=("ring", 6.0, 1.5, 0.1, 1234))
    T2.set_zeta(prob, np.zeros((n, n)))
    E = ff.Diagnostic(T2.energy, prob)
    D = ff.Diagnostic(T2.energy_dissipation_hyperviscosity, prob)
    R = ff.Diagnostic(T2.energy_dissipation_hypoviscosity, prob)
    W = ff.Diagnostic(T2.energy_work, prob)
    ff.stepforward(prob, nsteps, diags=[E, D, R, W])
    Ev, Dv, Rv, Wv = (np.array(d.data) for d in (E, D, R, W))
    dEdt = (Ev[1:] - Ev[:-1]) / dt
    computed = Wv[1:] + Dv[:-1] + Rv[:-1]
    assert np.linalg.norm(dEdt - computed) / np.linalg.norm(dEdt) < 1e-3


def test_ring_forcing_injection_rate():
    # the normalisation makes <energy injection> = eps: parsevalsum(spec invK / 2)/(LxLy) = eps
    from oracle import forcing
    g = ff.TwoDGrid(128, 2 * np.pi)
    spec = forcing.ring_spectrum(g, 14.0, 1.5, 0.1)
    assert abs(ff.parsevalsum(spec * g.invKrsq / 2, g) / (g.Lx * g.Ly) - 0.1) < 1e-14
    assert spec[0, 0] == 0


def test_philox_known_answer_vectors():
    from oracle.forcing import philox4x32
    for c in GOLD["philox4x32_10_kat"]["cases"]:
        ctr = [int(x, 16) for x in c["ctr"]]
        key = [int(x, 16) for x in c["key"]]
        out = [int(v) for v in philox4x32(*ctr, *key)]
        assert out == [int(x, 16) for x in c["out"]]


def test_complex_normal_statistics():
    from oracle.forcing import complex_normal
    z = complex_normal(5, 3, np.arange(200000, dtype=np.uint64))
    assert abs(np.mean(z.real)) < 0.01 and abs(np.mean(z.imag)) < 0.01
    assert abs(np.var(z.real) - 0.5) < 0.01 and abs(np.var(z.imag) - 0.5) < 0.01
    assert abs(np.mean(z.real * z.imag)) < 0.01
