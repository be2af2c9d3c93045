"""The oracle against its committed fixtures (tests/golden/oracle_fixtures_32.npz, written by
tests/golden/make_oracle_fixtures.py): a change to the oracle that moves any of these values
shows up here.  The fixtures are oracle outputs, not reference outputs."""
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "oracle_fixtures_32.npz")


def _generator():
    spec = importlib.util.spec_from_file_location("make_oracle_fixtures", os.path.join(HERE, "golden", "make_oracle_fixtures.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_oracle_reproduces_its_fixtures():
    want = np.load(PATH)
    got = _generator().compute()
    assert sorted(got) == sorted(want.files)
    for k in want.files:
        a, b = got[k], want[k]
        assert a.shape == b.shape, k
        assert np.allclose(a, b, rtol=1e-13, atol=1e-15 * max(np.abs(b).max(), 1e-300)), k


def test_fixture_coefficients_against_closed_forms():
    # phi-function closed forms at the fixture's dt L (away from the cancellation region)
    f = np.load(PATH)
    dt = 0.01
    z = dt * f["etd/L"]
    sel = np.abs(z) > 1e-2
    ez = np.exp(z)
    closed = {"zeta": dt * (np.exp(z / 2) - 1) / z,
              "alpha": dt * (-4 - z + ez * (4 - 3 * z + z ** 2)) / z ** 3,
              "beta": dt * (2 + z + ez * (z - 2)) / z ** 3,
              "Gamma": dt * (-4 - 3 * z - z ** 2 + ez * (4 - z)) / z ** 3}
    for k, v in closed.items():
        assert np.allclose(f[f"etd/{k}"][sel], v[sel], rtol=1e-9), k
