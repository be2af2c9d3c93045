"""GPU adjoint of the TwoDNavierStokes ETDRK4 step against the oracle's adjoint
(oracle/adjoint.py, pinned on the CPU by finite differences and dot-product identities),
against central differences of the GPU's own loss, and for checkpoint invariance.

Tolerances (FP64): gradient vs oracle 1e-10 relative L2, dJ/dnu 1e-9 relative, J 1e-12;
central differences (h = 1e-4 on unit-variance directions) 1e-6 relative.
"""
import numpy as np
import pytest

from oracle import adjoint as oad
from oracle import fourierflows as off

pytestmark = pytest.mark.gpu


def _setup(gf, n, nu, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    grid = gf.problem.Grid(n, n, 2 * np.pi, 2 * np.pi, 1 / 3, np.float64)
    z0 = gf.peakedisotropicspectrum(grid, min(6, n // 8), scale, rng=rng)
    target = gf.peakedisotropicspectrum(grid, min(6, n // 8), scale, rng=rng)
    prob = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=n, ν=nu, nν=1, dt=0.005, stepper="ETDRK4")
    gf.TwoDNavierStokes.set_ζ(prob, z0)
    return prob, z0, target


@pytest.mark.parametrize("n,nsteps,ckpt", [(32, 3, 3), (64, 4, 2)])
def test_adjoint_loss_equals_forward_steps(gf, n, nsteps, ckpt):
    # the taped forward of the adjoint and stepforward give the same zeta(T)
    prob, z0, target = _setup(gf, n, 1e-3, 30 + n)
    J, _, _ = gf.TwoDNavierStokes.adjoint_gradient(prob, target, nsteps, ckpt, nu=False)
    gf.stepforward(prob, nsteps)
    gf.TwoDNavierStokes.updatevars(prob)
    Jf = 0.5 * float(np.sum((prob.vars.ζ - target) ** 2))
    assert J == pytest.approx(Jf, rel=1e-13)


@pytest.mark.parametrize("n,nsteps,ckpt", [(32, 12, 5), (64, 10, 10), (128, 6, 4)])
def test_gradient_matches_oracle(gf, n, nsteps, ckpt):
    prob, z0, target = _setup(gf, n, 1e-3, 40 + n)
    J, grad, dnu = gf.TwoDNavierStokes.adjoint_gradient(prob, target, nsteps, ckpt)
    g = off.TwoDGrid(n, 2 * np.pi)
    Jr, gr, dnur = oad.loss_and_gradient(z0, target, g, 1e-3, 1, 0.005, nsteps, checkpoint=ckpt)
    assert J == pytest.approx(Jr, rel=1e-12)
    assert np.linalg.norm(grad - gr) <= 1e-10 * np.linalg.norm(gr)
    assert dnu == pytest.approx(dnur, rel=1e-9)


def test_gradient_512_two_checkpoint_segments_matches_oracle(gf):
    """The adjoint workload's grid (512^2, nu = 1e-3, nnu = 1, dt = 0.005) over 100 steps
    with a checkpoint every 50: the first segment is recomputed from its checkpoint on the
    second stream while the sweep runs the second.  Against the oracle's adjoint (gradient
    1e-10, dJ/dnu 1e-9, J 1e-12), then central differences of the device loss on 8 random
    directions (h = 1e-4, 1e-6 relative)."""
    n, nsteps, ckpt = 512, 100, 50
    prob, z0, target = _setup(gf, n, 1e-3, 512)
    J, grad, dnu = gf.TwoDNavierStokes.adjoint_gradient(prob, target, nsteps, ckpt)
    g = off.TwoDGrid(n, 2 * np.pi)
    Jr, gr, dnur = oad.loss_and_gradient(z0, target, g, 1e-3, 1, 0.005, nsteps, checkpoint=ckpt)
    assert J == pytest.approx(Jr, rel=1e-12)
    assert np.linalg.norm(grad - gr) <= 1e-10 * np.linalg.norm(gr)
    assert dnu == pytest.approx(dnur, rel=1e-9)
    rng = np.random.default_rng(8)
    h = 1e-4
    for _ in range(8):
        v = rng.standard_normal((n, n))
        Js = []
        for sgn in (1, -1):
            gf.TwoDNavierStokes.set_ζ(prob, z0 + sgn * h * v)
            Js.append(gf.TwoDNavierStokes.adjoint_gradient(prob, target, nsteps, nsteps, nu=False)[0])
        fd = (Js[0] - Js[1]) / (2 * h)
        assert fd == pytest.approx(float(np.sum(grad * v)), rel=1e-6)


def test_checkpoint_interval_does_not_change_the_gradient(gf):
    prob, z0, target = _setup(gf, 64, 1e-3, 5)
    base = gf.TwoDNavierStokes.adjoint_gradient(prob, target, 13, 13)
    for ckpt in (1, 4, 6, 3, 5, 12):          # 13, 4, 3, 5, 3 and 2 segments, uneven last ones
        other = gf.TwoDNavierStokes.adjoint_gradient(prob, target, 13, ckpt)
        assert other[0] == base[0]
        assert np.array_equal(other[1], base[1])
        assert other[2] == base[2]


def test_gradient_matches_central_differences(gf):
    n, nsteps = 128, 20
    prob, z0, target = _setup(gf, n, 2e-3, 6)
    J, grad, dnu = gf.TwoDNavierStokes.adjoint_gradient(prob, target, nsteps, 8)
    rng = np.random.default_rng(7)
    # J is O(1e6) here: rounding noise in J of ~1e-8 needs h well above 1e-6; the O(h^2)
    # truncation of central differences stays far below the tolerance at h = 1e-4
    h = 1e-4
    for _ in range(3):
        v = rng.standard_normal((n, n))
        Js = []
        for sgn in (1, -1):
            gf.TwoDNavierStokes.set_ζ(prob, z0 + sgn * h * v)
            Js.append(gf.TwoDNavierStokes.adjoint_gradient(prob, target, nsteps, nsteps, nu=False)[0])
        fd = (Js[0] - Js[1]) / (2 * h)
        assert fd == pytest.approx(float(np.sum(grad * v)), rel=1e-6)


def test_state_unchanged_and_unsupported_problems(gf):
    prob, z0, target = _setup(gf, 64, 1e-3, 8)
    before = prob.sol
    gf.TwoDNavierStokes.adjoint_gradient(prob, target, 5, 2)
    assert np.array_equal(prob.sol, before)
    assert prob.clock.step == 0
    rk = gf.TwoDNavierStokes.Problem(gf.GPU(), nx=64, ν=1e-3, dt=0.005, stepper="RK4")
    with pytest.raises(gf._lib.GFError) as e:
        gf.TwoDNavierStokes.adjoint_gradient(rk, target, 5)
    assert e.value.code == gf._lib.GF_EUNSUPPORTED
    q = gf.SingleLayerQG.Problem(gf.GPU(), nx=64, dt=0.005, stepper="ETDRK4")
    with pytest.raises(gf._lib.GFError) as e:
        gf.TwoDNavierStokes.adjoint_gradient(q, target, 5)
    assert e.value.code == gf._lib.GF_EUNSUPPORTED


def test_concurrent_gradients_equal_sequential_ones(gf):
    # a gradient that starts while another runs keeps its sweep on one stream; alone, the
    # recomputation runs on a second stream beside the sweep: both give the same bits
    from concurrent.futures import ThreadPoolExecutor
    setups = [_setup(gf, 64, 1e-3, 70 + k) for k in range(4)]
    seq = [gf.TwoDNavierStokes.adjoint_gradient(p, tgt, 40, 6) for p, _, tgt in setups]
    with ThreadPoolExecutor(max_workers=4) as pool:
        par = list(pool.map(lambda s: gf.TwoDNavierStokes.adjoint_gradient(s[0], s[2], 40, 6), setups))
    for a, b in zip(seq, par):
        assert a[0] == b[0] and a[2] == b[2]
        assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n", [128, 512])
def test_workgroup_splits_give_the_same_gradient(gf, n):
    # the splits only regroup the same arithmetic: forward tape and adjoint agree bit for bit
    # (at 512 K3's column split applies: 512 / 8 threads per column)
    prob, z0, target = _setup(gf, n, 1e-3, 80)
    gf.set_splits(prob, 0, 0, 0)
    a = gf.TwoDNavierStokes.adjoint_gradient(prob, target, 9, 4)
    gf.set_splits(prob, 1, 1, 1)
    b = gf.TwoDNavierStokes.adjoint_gradient(prob, target, 9, 4)
    assert a[0] == b[0] and a[2] == b[2]
    assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n", [64, 512])
def test_tape_budget_keeps_segments_and_gives_the_same_gradient(gf, n):
    # the recomputed tape holds the forward pass's bits, so the number of segments kept resident
    # (pool slots from the budget) changes no bit of the gradient: 7 segments of 3 steps with 2
    # (default), 4 and all 7 slots, alone (two streams) and beside another gradient (one stream)
    from concurrent.futures import ThreadPoolExecutor
    prob, z0, target = _setup(gf, n, 1e-3, 90)
    base = gf.TwoDNavierStokes.adjoint_gradient(prob, target, 20, 3)
    kmax2 = {64: 22, 512: 170}[n]   # surviving columns kr < kmax rounded up to a pair
    seg_bytes = 16 * kmax2 * n * (8 * 3 + 1)
    for slots in (4, 7):
        got = gf.TwoDNavierStokes.adjoint_gradient(prob, target, 20, 3, tape_budget=slots * seg_bytes)
        assert got[0] == base[0] and got[2] == base[2], slots
        assert np.array_equal(got[1], base[1]), slots
    other, _, tgt2 = _setup(gf, n, 1e-3, 91)
    with ThreadPoolExecutor(max_workers=2) as pool:
        fa = pool.submit(gf.TwoDNavierStokes.adjoint_gradient, prob, target, 20, 3, True, 4 * seg_bytes)
        fb = pool.submit(gf.TwoDNavierStokes.adjoint_gradient, other, tgt2, 20, 3)
        got, _ = fa.result(), fb.result()
    assert got[0] == base[0] and got[2] == base[2] and np.array_equal(got[1], base[1])
    # back to the minimum: the pool keeps its size, the gradient its bits
    got = gf.TwoDNavierStokes.adjoint_gradient(prob, target, 20, 3, tape_budget=0)
    assert np.array_equal(got[1], base[1])
