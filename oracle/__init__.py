"""CPU oracle for the pseudospectral time-stepping path of GeophysicalFlows.jl.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy/SciPy restatement of the
reference algorithm (GeophysicalFlows.jl v0.16.5 plus the FourierFlows.jl 0.10
grid/FFT/dealias/time-stepper rules it calls).  It exists to check the HIP path:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and then only as the checker or as the timed CPU
baseline.  The product path (the HIP library and its Python host mirror) never
imports it and fails loudly when its native library is missing.

Conventions (identical memory layout to the reference's column-major arrays):

* physical fields are NumPy arrays of shape ``(ny, nx)`` (``(nlayers, ny, nx)``
  for MultiLayerQG) -- Julia ``f[x, y]`` is ``f[y, x]`` here, same bytes;
* spectral fields are ``(nl, nkr)`` (``(nlayers, nl, nkr)``), ``nkr = nx//2 + 1``.

Parity status of each function is stated in the module docstrings; the summary:

* ``calcN`` of all three models, ``L``, FFT normalisation, ``parsevalsum``,
  PV inversion: pinned by the reference's closed-form and manufactured-solution
  tests, restated in ``tests/test_oracle_*.py``.
* ETDRK4 coefficients: pinned against closed-form phi-functions (mpmath).
* Dealias index rule, ``makefilter`` constants, ETD contour parameters: recalled
  from FourierFlows.jl 0.10 (not vendored in the reference); the tests pin what
  the reference's tests pin at low wavenumbers; the exact aliased-band index and
  filter constants are "parity unpinned" against the real package.
"""
