"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/gf.h, the descriptor layout matches, and invalid
descriptors are rejected with the documented codes before any device work."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "gf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gf_[A-Za-z_0-9]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("gf_ctx_create", "gf_calcN", "gf_step", "gf_set_state", "gf_get_state", "gf_updatevars",
                 "gf_set_physical", "gf_clock", "gf_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(gf):
    lib = ctypes.CDLL(gf._lib.LIBPATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(gf._lib.EXPORTS) == set(_declared())


def test_descriptor_layout_matches_header(gf, tmp_path):
    # compile a probe against include/gf.h with the host C compiler and compare every
    # field offset and the struct size with the ctypes mirror
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    D = gf._lib.ProblemDesc
    fields = [f[0] for f in D._fields_]
    body = "\n".join(f'  printf("%s %zu\\n", "{f}", offsetof(gf_problem_desc, {f}));' for f in fields)
    src = tmp_path / "probe.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "gf.h"\nint main(void){\n' + body +
                   '\n  printf("sizeof %zu\\n", sizeof(gf_problem_desc));\n  return 0;\n}\n')
    exe = tmp_path / "probe"
    subprocess.check_call([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = dict(line.split() for line in subprocess.check_output([str(exe)]).decode().splitlines())
    for f in fields:
        assert int(out[f]) == getattr(D, f).offset, f
    assert int(out["sizeof"]) == ctypes.sizeof(D)


def _desc(gf, **kw):
    d = gf._lib.ProblemDesc()
    d.model, d.dtype, d.nx, d.ny, d.nlayers = 0, 0, 64, 64, 1
    d.Lx = d.Ly = 6.283185307179586
    d.aliased_fraction, d.stepper, d.dt = 1 / 3, 0, 0.01
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("bad,code", [
    (dict(nx=8192), -2), (dict(ny=8192), -2), (dict(ny=65), -2), (dict(ny=258), -2), (dict(nx=67), -2),
    (dict(nx=260), -2), (dict(nx=16, ny=14), -2),
    (dict(dt=0.0), -1), (dict(model=7), -1),
    (dict(stepper=9), -1), (dict(dtype=3), -1), (dict(model=2, nlayers=9), -2),
    (dict(aliased_fraction=1.5), -1), (dict(Lx=-1.0), -1), (dict(nlayers=2), -1),
    (dict(forcing=2, dkf=0.0), -1),
])
def test_invalid_descriptors_are_rejected(gf, bad, code):
    lib = gf._lib.lib
    ctx = ctypes.c_void_p()
    rc = lib.gf_ctx_create(ctypes.byref(ctx), ctypes.byref(_desc(gf, **bad)))
    assert rc == code
    assert not ctx.value
    assert lib.gf_last_error(None)


def test_null_arguments(gf):
    lib = gf._lib.lib
    assert lib.gf_ctx_create(ctypes.byref(ctypes.c_void_p()), None) == -1
    assert lib.gf_step(None, 1) == -1
    assert lib.gf_error_string(-2) == b"unsupported configuration"


def test_python_mirror_rejects_unknown_stepper(gf):
    with pytest.raises(ValueError):
        gf.TwoDNavierStokes.Problem(gf.GPU(), nx=64, stepper="LSRK54")
    with pytest.raises(NotImplementedError):
        gf.TwoDNavierStokes.Problem(gf.GPU(), nx=64, calcF=lambda *a: None)


def _slqg(gf, n=64, **kw):
    return _desc(gf, model=1, nx=n, ny=n, stepper=3, **kw)


@pytest.mark.parametrize("bad,code", [
    (dict(dtype=1), -1), (dict(model=2, nlayers=2), -2), (dict(ny=64), -2), (dict(nx=1024, ny=1024), -2),
    (dict(nx=16, ny=16), -2), (dict(stepper=6), -1), (dict(stepper=2), -1), (dict(dt=0.0), -1),
])
def test_invalid_ensembles_are_rejected(gf, bad, code):
    lib = gf._lib.lib
    descs = (gf._lib.ProblemDesc * 3)(_slqg(gf), _slqg(gf, 128), _slqg(gf))
    for k, v in bad.items():
        setattr(descs[1], k, v)
    ctx = ctypes.c_void_p()
    assert lib.gf_ctx_create_batch(ctypes.byref(ctx), descs, 3) == code
    assert not ctx.value
    assert b"ensemble member 1" in lib.gf_last_error(None)


def test_empty_ensemble_and_null_handles(gf):
    lib = gf._lib.lib
    ctx = ctypes.c_void_p()
    assert lib.gf_ctx_create_batch(ctypes.byref(ctx), None, 0) == -1
    assert lib.gf_batch_size(None) == -1
    assert lib.gf_ctx_instance(None, 0, ctypes.byref(ctx)) == -1
    lib.gf_ctx_destroy(None)


def test_peakedisotropicspectrum_shape_and_energy(gf):
    # src/utils.jl:29-58 with allones = true: |psih| = (K^2 (1 + (K/k0)^4))^(-1/2) scaled to E0
    import numpy as np
    g = gf.problem.Grid(128, 128, 2 * np.pi, 2 * np.pi, 1 / 3, np.float64)
    q = gf.peakedisotropicspectrum(g, 6, 0.5, allones=True)
    qh = np.fft.rfft2(q)
    with np.errstate(divide="ignore", invalid="ignore"):
        psih = np.where(g.Krsq > 0, -qh / g.Krsq, 0)
    assert np.sum(g.Krsq * np.abs(psih) ** 2) / (128 * 128) ** 2 == pytest.approx(0.5, rel=1e-12)
    K = np.sqrt(g.Krsq)
    k0 = 6.0
    sel = (K > 3) & (K < 40)
    expect = (g.Krsq[sel] * (1 + (K[sel] / k0) ** 4)) ** -0.5
    ratio = np.abs(psih[sel]) / expect
    assert np.ptp(ratio) / ratio.mean() < 1e-10
    # the reference's radial-spectrum check: |qh|/Krsq ~ K^-2 well above k0 (test/test_utils.jl:4-15)
    hi = (K > 15) & (K <= 17.5)
    r = (np.abs(qh[hi]) / g.Krsq[hi]) * K[hi] ** 2
    assert np.ptp(r) / r.mean() < 0.2
    with pytest.raises(ValueError):
        gf.peakedisotropicspectrum(gf.problem.Grid(32, 34, 2 * np.pi, 3 * np.pi, 1 / 3, np.float64), 6, 0.5)
