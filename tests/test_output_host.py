"""Host side of the Output mirror (no device): the key layout of the reference's saved files
(examples/singlelayerqg_betaforced.jl:216-231 reads snapshots/t/<step>, grid/x, ...), the
one-write-per-key rule, and the diagnostic series."""
import types

import numpy as np
import pytest


class _Clock:
    def __init__(self):
        self.step, self.t, self.dt = 0, 0.0, 0.25


def _stub():
    g = types.SimpleNamespace(nx=4, ny=2, Lx=2.0, Ly=1.0, x=np.linspace(-1, 0.5, 4), y=np.array([-0.5, 0.0]))
    params = types.SimpleNamespace(ν=1e-3, nν=2, stepper="RK4", calcF=None)
    return types.SimpleNamespace(grid=g, params=params, clock=_Clock(), _ctx=None, counter=0)


def test_output_keys_and_values(gf, tmp_path):
    out_mod = gf.output
    p = _stub()
    path = tmp_path / "run.npz"
    field = lambda prob: np.full((prob.grid.ny, prob.grid.nx), float(prob.clock.step))
    out = out_mod.Output(p, path, ("f", field))
    out_mod.saveproblem(out)
    for step in (0, 5):
        p.clock.step, p.clock.t = step, 0.25 * step
        out_mod.saveoutput(out)
    out.close()
    f = np.load(path)
    assert sorted(f.files) == sorted(["grid/nx", "grid/ny", "grid/Lx", "grid/Ly", "grid/x", "grid/y", "clock/dt",
                                      "params/ν", "params/nν", "snapshots/t/0", "snapshots/t/5",
                                      "snapshots/f/0", "snapshots/f/5"])
    assert float(f["snapshots/t/5"]) == 1.25 and np.all(f["snapshots/f/5"] == 5.0)
    assert np.array_equal(f["grid/x"], p.grid.x) and int(f["params/nν"]) == 2
    with pytest.raises(ValueError, match="already exists"):
        out_mod.saveoutput(out)


def test_savediagnostic_layout(gf, tmp_path):
    out_mod = gf.output
    d = types.SimpleNamespace(i=3, t=np.array([0.0, 0.1, 0.2, 0.0]), steps=np.array([0, 1, 2, 0]),
                              data=[1.0, 0.5, 0.25])
    out_mod.savediagnostic(d, "energy", tmp_path / "d.npz")
    f = np.load(tmp_path / "d.npz")
    assert np.array_equal(f["diagnostics/energy/data"], [1.0, 0.5, 0.25])
    assert np.array_equal(f["diagnostics/energy/steps"], [0, 1, 2])
