"""The device against the committed oracle fixtures (tests/golden/oracle_fixtures_32.npz): calcN,
its in-place dealiasing and three steps for TwoDNavierStokes (ETDRK4, FilteredRK4),
SingleLayerQG with topography and MultiLayerQG (3 layers) at 32^2, without running the oracle.
Tolerances as in the parity tests: calcN 1e-13, the state after three steps 1e-12 (relative L2)."""
import math
import os

import numpy as np
import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu
PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_fixtures_32.npz")
N = 32


def _problems(gf):
    x = -math.pi + 2 * math.pi * np.arange(N) / N
    X, Y = np.meshgrid(x, x)
    eta = 0.2 * np.cos(3 * X) * np.sin(2 * Y)
    return {
        "twodns_etdrk4": lambda: gf.TwoDNavierStokes.Problem(gf.GPU(), nx=N, ν=1e-3, nν=2, μ=0.05, nμ=0, dt=0.01,
                                                             stepper="ETDRK4"),
        "twodns_filteredrk4": lambda: gf.TwoDNavierStokes.Problem(gf.GPU(), nx=N, ν=1e-4, nν=1, dt=0.01,
                                                                  stepper="FilteredRK4"),
        "slqg_topo_etdrk4": lambda: gf.SingleLayerQG.Problem(gf.GPU(), nx=N, β=5.0, U=0.2, eta=eta, ν=1e-3, nν=1,
                                                             μ=0.02, dt=0.01, stepper="ETDRK4"),
        "mlqg3_etdrk4": lambda: gf.MultiLayerQG.Problem(3, gf.GPU(), nx=N, f0=1.0, β=5.0, U=[0.5, 0.1, 0.0],
                                                        H=[0.2, 0.3, 0.5], b=[0.0, -0.2, -0.27], μ=0.01, ν=1e-3,
                                                        dt=0.005, stepper="ETDRK4"),
    }


@pytest.mark.parametrize("name", ["twodns_etdrk4", "twodns_filteredrk4", "slqg_topo_etdrk4", "mlqg3_etdrk4"])
def test_device_matches_oracle_fixtures(gf, name):
    f = np.load(PATH)
    p = _problems(gf)[name]()
    s = f[f"{name}/state"]
    Nd, sd = p.calcN(s)
    assert rel(Nd, f[f"{name}/calcN"]) < 1e-13
    assert np.array_equal(sd, f[f"{name}/calcN_dealiased_state"])
    p.sol = s
    gf.stepforward(p, 3)
    assert rel(p.sol, f[f"{name}/after_3_steps"]) < 1e-12
