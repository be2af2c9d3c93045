"""MI355X-native pseudospectral time stepping for GeophysicalFlows.jl's hot path.

The per-step nonlinear term ``calcN!`` of TwoDNavierStokes, SingleLayerQG and
MultiLayerQG and the ETDRK4 / RK4 / ForwardEuler (and Filtered) stepper stages
behind ``stepforward!`` run as fused HIP kernels for gfx950 in ``libgfhip.so``
(C ABI: ``include/gf.h``).  This Python package mirrors the reference's
user-facing API on top of that ABI; it has no CPU compute path.
"""
from . import _lib  # noqa: F401  (loads libgfhip.so or raises ImportError)
from . import multilayerqg as MultiLayerQG  # noqa: F401
from . import singlelayerqg as SingleLayerQG  # noqa: F401
from . import twodnavierstokes as TwoDNavierStokes  # noqa: F401
from . import output  # noqa: F401
from . import utils  # noqa: F401
from .output import Output, get_sol, savediagnostic, saveoutput, saveproblem  # noqa: F401
from .problem import (CPU, GPU, Diagnostic, Ensemble, HarmonicForcing, RingForcing, autotune,  # noqa: F401
                      load_schedules, save_schedules, schedule, set_schedule, set_splits, splits, stepforward)
from .utils import peakedisotropicspectrum  # noqa: F401

__all__ = ["TwoDNavierStokes", "SingleLayerQG", "MultiLayerQG", "GPU", "CPU", "Ensemble", "RingForcing", "HarmonicForcing", "Diagnostic", "stepforward",
           "peakedisotropicspectrum", "Output", "get_sol", "saveproblem", "saveoutput", "savediagnostic",
           "autotune", "schedule", "set_schedule", "splits", "set_splits", "save_schedules", "load_schedules"]
