"""ctypes binding of libgfhip.so (the C ABI declared in include/gf.h).

The library is loaded from this package's directory.  There is no fallback:
if the shared library is missing or cannot be loaded, importing the package
raises ``ImportError`` naming the file, so no CPU path can stand in for it.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libgfhip.so")

GF_OK, GF_EINVAL, GF_EUNSUPPORTED, GF_EHIP, GF_ENOTFINITE, GF_ENODEVICE = 0, -1, -2, -3, -4, -5
GF_MODEL_TWODNS, GF_MODEL_SLQG, GF_MODEL_MLQG = 0, 1, 2
GF_F64, GF_F32 = 0, 1
STEPPERS = {"ETDRK4": 0, "FilteredETDRK4": 1, "RK4": 2, "FilteredRK4": 3,
            "ForwardEuler": 4, "FilteredForwardEuler": 5, "AB3": 6, "FilteredAB3": 7}
GF_FORCING_NONE, GF_FORCING_FIXED, GF_FORCING_RING = 0, 1, 2
(GF_DIAG_ENERGY, GF_DIAG_KINETIC_ENERGY, GF_DIAG_POTENTIAL_ENERGY, GF_DIAG_ENSTROPHY, GF_DIAG_PALINSTROPHY,
 GF_DIAG_ENERGY_DISSIPATION, GF_DIAG_ENSTROPHY_DISSIPATION, GF_DIAG_ENERGY_DRAG, GF_DIAG_ENSTROPHY_DRAG,
 GF_DIAG_ENERGY_WORK, GF_DIAG_ENSTROPHY_WORK, GF_DIAG_LAYER_ENERGIES, GF_DIAG_FLUXES) = range(13)

EXPORTS = ("gf_ctx_create", "gf_ctx_destroy", "gf_last_error", "gf_error_string",
           "gf_set_state", "gf_get_state", "gf_set_state_dev", "gf_get_state_dev",
           "gf_set_physical", "gf_updatevars", "gf_num_vars", "gf_calcN", "gf_set_forcing",
           "gf_step", "gf_clock", "gf_set_clock", "gf_synchronize", "gf_timing",
           "gf_timing_read", "gf_time_kernel", "gf_kernel_bytes", "gf_version", "gf_ctx_create_batch",
           "gf_batch_size", "gf_ctx_instance", "gf_adjoint_grad", "gf_diagnostic",
           "gf_diagnostic_async", "gf_snapshot", "gf_snapshot_wait", "gf_host_alloc", "gf_host_free",
           "gf_device_alloc", "gf_device_free", "gf_device_copy", "gf_use_graphs",
           "gf_set_schedule", "gf_get_schedule", "gf_autotune", "gf_schedules_save", "gf_schedules_load",
           "gf_schedules_clear", "gf_schedules_error", "gf_set_splits", "gf_get_splits",
           "gf_set_forcing_harmonic", "gf_set_psi", "gf_diagnostic_async_noalloc", "gf_adjoint_tape_budget",
           "gf_batch_persistent")


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int), ("dtype", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nlayers", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
        ("aliased_fraction", ctypes.c_double),
        ("stepper", ctypes.c_int), ("dt", ctypes.c_double),
        ("nu", ctypes.c_double), ("nnu", ctypes.c_int), ("mu", ctypes.c_double), ("nmu", ctypes.c_int),
        ("beta", ctypes.c_double), ("deformation_radius", ctypes.c_double), ("U", ctypes.c_double),
        ("eta", ctypes.POINTER(ctypes.c_double)), ("Uy", ctypes.POINTER(ctypes.c_double)),
        ("f0", ctypes.c_double), ("H", ctypes.POINTER(ctypes.c_double)), ("b", ctypes.POINTER(ctypes.c_double)),
        ("topo_grad_x", ctypes.c_double), ("topo_grad_y", ctypes.c_double),
        ("linear", ctypes.c_int),
        ("forcing", ctypes.c_int), ("kf", ctypes.c_double), ("dkf", ctypes.c_double), ("eps", ctypes.c_double),
        ("seed", ctypes.c_uint64), ("instance", ctypes.c_int),
    ]


class GFError(RuntimeError):
    """Error returned across the C ABI (mirrors the reference's ErrorException)."""

    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def _load():
    if not os.path.exists(LIBPATH):
        raise ImportError(f"libgfhip.so not found at {LIBPATH}: build it with `make -C {_HERE}` "
                          "or __graft_entry__.build(); there is no CPU fallback")
    try:
        lib = ctypes.CDLL(LIBPATH)
    except OSError as e:
        raise ImportError(f"cannot load {LIBPATH}: {e}") from e
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    i64p = ctypes.POINTER(ctypes.c_int64)
    lib.gf_ctx_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ProblemDesc)]
    lib.gf_ctx_destroy.argtypes = [vp]
    lib.gf_ctx_destroy.restype = None
    lib.gf_last_error.argtypes = [vp]
    lib.gf_last_error.restype = ctypes.c_char_p
    lib.gf_error_string.argtypes = [ip]
    lib.gf_error_string.restype = ctypes.c_char_p
    for name in ("gf_set_state", "gf_get_state", "gf_set_state_dev", "gf_get_state_dev",
                 "gf_set_physical", "gf_updatevars", "gf_set_forcing"):
        getattr(lib, name).argtypes = [vp, vp]
    lib.gf_set_psi.argtypes = [vp, vp]
    lib.gf_num_vars.argtypes = [vp]
    lib.gf_calcN.argtypes = [vp, vp, vp, ctypes.c_double]
    lib.gf_step.argtypes = [vp, ip]
    lib.gf_clock.argtypes = [vp, dp, i64p]
    lib.gf_set_clock.argtypes = [vp, ctypes.c_double, ctypes.c_int64]
    lib.gf_synchronize.argtypes = [vp]
    lib.gf_timing.argtypes = [vp, ip]
    lib.gf_timing_read.argtypes = [vp, ip, i64p, dp]
    lib.gf_time_kernel.argtypes = [vp, ip, ip, dp]
    lib.gf_kernel_bytes.argtypes = [vp, ip, ip, dp]
    lib.gf_version.restype = ctypes.c_char_p
    lib.gf_ctx_create_batch.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ProblemDesc), ip]
    lib.gf_batch_size.argtypes = [vp]
    lib.gf_ctx_instance.argtypes = [vp, ip, ctypes.POINTER(vp)]
    lib.gf_adjoint_grad.argtypes = [vp, vp, ip, ip, vp, dp, dp]
    lib.gf_adjoint_tape_budget.argtypes = [vp, ctypes.c_size_t]
    lib.gf_diagnostic.argtypes = [vp, ip, ctypes.c_double, ip, dp, ip]
    lib.gf_diagnostic_async.argtypes = [vp, ip, ctypes.c_double, ip, vp, ip]
    lib.gf_diagnostic_async_noalloc.argtypes = [vp, ip, ctypes.c_double, ip, vp, ip]
    lib.gf_snapshot.argtypes = [vp, vp]
    lib.gf_snapshot.restype = ctypes.c_int64
    lib.gf_snapshot_wait.argtypes = [vp, ctypes.c_int64]
    lib.gf_host_alloc.argtypes = [ctypes.c_size_t]
    lib.gf_host_alloc.restype = ctypes.c_void_p
    lib.gf_host_free.argtypes = [vp]
    lib.gf_host_free.restype = None
    lib.gf_device_alloc.argtypes = [ctypes.c_size_t]
    lib.gf_device_alloc.restype = ctypes.c_void_p
    lib.gf_device_free.argtypes = [vp]
    lib.gf_device_free.restype = None
    lib.gf_device_copy.argtypes = [vp, vp, ctypes.c_size_t]
    lib.gf_use_graphs.argtypes = [vp, ip]
    lib.gf_batch_persistent.argtypes = [vp, ip]
    lib.gf_set_schedule.argtypes = [vp, ip, ip, ip]
    lib.gf_get_schedule.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int)] * 3
    lib.gf_autotune.argtypes = [vp, ip, dp]
    lib.gf_set_splits.argtypes = [vp, ip, ip, ip]
    lib.gf_set_forcing_harmonic.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double]
    lib.gf_get_splits.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int)]
    lib.gf_schedules_save.argtypes = [ctypes.c_char_p]
    lib.gf_schedules_load.argtypes = [ctypes.c_char_p]
    lib.gf_schedules_clear.argtypes = []
    lib.gf_schedules_clear.restype = None
    lib.gf_schedules_error.argtypes = []
    lib.gf_schedules_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        getattr(lib, name)
    return lib


lib = _load()


def check(rc, ctx=None):
    if rc != GF_OK:
        msg = lib.gf_last_error(ctx).decode() if ctx is not None else lib.gf_last_error(None).decode()
        raise GFError(rc, msg or lib.gf_error_string(rc).decode())
