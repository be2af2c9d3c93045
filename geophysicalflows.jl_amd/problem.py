"""Host-side mirror of the FourierFlows.jl objects the reference builds around its hot path.

``Problem`` holds what ``FourierFlows.Problem(eqn, stepper, dt, grid, vars, params)``
holds for the reference (``src/twodnavierstokes.jl:99``): the grid description,
the parameters, a ``Clock``, ``sol`` and ``vars``.  The state itself lives in
device memory owned by the C ABI context; ``prob.sol`` copies it out (Julia
memory layout, NumPy shape ``(nl, nkr)`` or ``(nlayers, nl, nkr)``).

Array conventions: a reference array ``f[x, y]`` (column-major, ``(nx, ny)``)
is a NumPy array of shape ``(ny, nx)`` here -- the same bytes.
"""
import ctypes
import math
import threading

import numpy as np

from . import _lib
from ._lib import check, lib


class CPU:
    """The reference's ``CPU()`` device: not provided by this library (HIP only)."""

    def __init__(self):
        raise NotImplementedError("this library runs on the GPU only; use GPU()")


class GPU:
    """The device handle (one HIP device, the current one)."""


class RingForcing:
    """Stochastic forcing on a wavenumber ring, regenerated once per step.

    Restates ``examples/twodnavierstokes_stochasticforcing.jl:53-82``: spectrum
    ``exp(-(K - kf)^2 / (2 dkf^2))`` (zero at K = 0) normalised to inject energy at
    rate ``eps``; ``Fh = xi sqrt(spectrum) / sqrt(dt)`` with ``xi`` complex normal
    (variance 1/2 per part) from a counter-based generator keyed on ``seed``.
    """

    def __init__(self, kf, dkf, eps, seed=0, instance=0):
        self.kf, self.dkf, self.eps, self.seed, self.instance = float(kf), float(dkf), float(eps), int(seed), int(instance)


class HarmonicForcing:
    """Time-harmonic forcing ``Fh cos(omega t + phase)``, evaluated at each stage's time as the
    reference's ``calcF!`` is (``test/test_twodnavierstokes.jl:141-145``:
    ``Fh = fh * cos(2t)``).  ``Fh`` is a spectral array in the layout of ``sol``."""

    def __init__(self, Fh, omega, phase=0.0):
        self.Fh, self.omega, self.phase = np.asarray(Fh), float(omega), float(phase)


class Clock:
    """[FF] ``Clock``: ``dt``, ``t``, ``step`` (read from the device context)."""

    def __init__(self, prob):
        self._p = prob

    @property
    def dt(self):
        return self._p._dt

    @property
    def t(self):
        t = ctypes.c_double()
        s = ctypes.c_int64()
        check(lib.gf_clock(self._p._ctx, ctypes.byref(t), ctypes.byref(s)), self._p._ctx)
        return t.value

    @property
    def step(self):
        t = ctypes.c_double()
        s = ctypes.c_int64()
        check(lib.gf_clock(self._p._ctx, ctypes.byref(t), ctypes.byref(s)), self._p._ctx)
        return s.value


class Grid:
    """Description of [FF] ``TwoDGrid``: sizes, spacings, wavenumbers, alias ranges."""

    def __init__(self, nx, ny, Lx, Ly, aliased_fraction, T):
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, float(Lx), float(Ly)
        self.nkr, self.nl = nx // 2 + 1, ny
        self.dx, self.dy = self.Lx / nx, self.Ly / ny
        self.T = T
        self.aliased_fraction = aliased_fraction
        self.x = -self.Lx / 2 + np.arange(nx) * self.dx
        self.y = -self.Ly / 2 + np.arange(ny) * self.dy
        self.kr = np.arange(self.nkr) * ((2 * math.pi / self.Lx * nx) / nx)
        li = np.arange(ny)
        self.l = np.where(li < (ny + 1) // 2, li, li - ny) * ((2 * math.pi / self.Ly * ny) / ny)
        self.Krsq = self.kr[None, :] ** 2 + self.l[:, None] ** 2

    def gridpoints(self):
        X = np.broadcast_to(self.x[None, :], (self.ny, self.nx)).copy()
        Y = np.broadcast_to(self.y[:, None], (self.ny, self.nx)).copy()
        return X, Y


class Vars:
    """Physical fields written by ``updatevars`` (names as in the reference's Vars)."""


_DEFER = threading.local()


class Problem:
    """Device-resident problem; see ``TwoDNavierStokes.Problem`` etc. for the constructors."""

    _names = ()

    def __init__(self, desc, grid, params, nlayers=1, keep=()):
        self._desc = desc
        self._dt = desc.dt
        self._keep = list(keep)          # arrays the descriptor points into
        self._owner = None
        self._pending_Fh = None
        self.grid = grid
        self.params = params
        self.nlayers = nlayers
        self.vars = Vars()
        self.stepper = params.stepper
        self._ctype = np.complex128 if desc.dtype == _lib.GF_F64 else np.complex64
        self._rtype = np.float64 if desc.dtype == _lib.GF_F64 else np.float32
        self.clock = Clock(self)
        self._ctx = None
        if getattr(_DEFER, "on", False):
            return                       # an Ensemble binds this problem to one of its members
        ctx = ctypes.c_void_p()
        check(lib.gf_ctx_create(ctypes.byref(ctx), ctypes.byref(desc)))
        self._ctx = ctx

    def _bind(self, member, owner):
        self._ctx = member
        self._owner = owner              # keeps the ensemble alive while this view is used
        if self._pending_Fh is not None:
            Fh, self._pending_Fh = self._pending_Fh, None
            self.set_forcing(Fh)

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value and getattr(self, "_owner", None) is None:
            lib.gf_ctx_destroy(ctx)
        self._ctx = None

    # ---- state ----
    @property
    def _sshape(self):
        g = self.grid
        return (g.nl, g.nkr) if self.nlayers == 1 else (self.nlayers, g.nl, g.nkr)

    def _sshape_bytes(self):
        return int(np.prod(self._sshape)) * np.dtype(self._ctype).itemsize

    @property
    def _pshape(self):
        g = self.grid
        return (g.ny, g.nx) if self.nlayers == 1 else (self.nlayers, g.ny, g.nx)

    @property
    def sol(self):
        out = np.empty(self._sshape, dtype=self._ctype)
        check(lib.gf_get_state(self._ctx, out.ctypes.data), self._ctx)
        return out

    @sol.setter
    def sol(self, value):
        a = np.ascontiguousarray(np.asarray(value, dtype=self._ctype).reshape(self._sshape))
        check(lib.gf_set_state(self._ctx, a.ctypes.data), self._ctx)

    def set_state_device(self, ptr):
        check(lib.gf_set_state_dev(self._ctx, ctypes.c_void_p(ptr)), self._ctx)

    def get_state_device(self, ptr):
        check(lib.gf_get_state_dev(self._ctx, ctypes.c_void_p(ptr)), self._ctx)

    def _set_physical(self, f):
        a = np.ascontiguousarray(np.asarray(f, dtype=self._rtype).reshape(self._pshape))
        check(lib.gf_set_physical(self._ctx, a.ctypes.data), self._ctx)

    def _updatevars(self):
        nf = lib.gf_num_vars(self._ctx)
        out = np.empty((nf,) + self._pshape, dtype=self._rtype)
        check(lib.gf_updatevars(self._ctx, out.ctypes.data), self._ctx)
        for name, arr in zip(self._names, out):
            setattr(self.vars, name, arr)

    def calcN(self, sol, t=None):
        """``calcN!(N, sol, t, ...)``: returns (N, sol dealiased in place as the reference does)."""
        s = np.ascontiguousarray(np.asarray(sol, dtype=self._ctype).reshape(self._sshape)).copy()
        N = np.empty(self._sshape, dtype=self._ctype)
        check(lib.gf_calcN(self._ctx, N.ctypes.data, s.ctypes.data, float(self.clock.t if t is None else t)), self._ctx)
        return N, s

    def set_forcing(self, Fh):
        """A fixed forcing table, or a ``HarmonicForcing``."""
        if self._ctx is None:
            self._pending_Fh = Fh
            return
        if isinstance(Fh, HarmonicForcing):
            a = np.ascontiguousarray(np.asarray(Fh.Fh, dtype=self._ctype).reshape(self._sshape))
            check(lib.gf_set_forcing_harmonic(self._ctx, a.ctypes.data, Fh.omega, Fh.phase), self._ctx)
            return
        a = np.ascontiguousarray(np.asarray(Fh, dtype=self._ctype).reshape(self._sshape))
        check(lib.gf_set_forcing(self._ctx, a.ctypes.data), self._ctx)

    def synchronize(self):
        check(lib.gf_synchronize(self._ctx), self._ctx)

    def use_graphs(self, enable=True):
        """Replay a captured graph of the step in ``stepforward`` (the default for ensembles) or
        launch every kernel (the default for single problems)."""
        check(lib.gf_use_graphs(self._ctx, int(bool(enable))), self._ctx)

    def _diagnostic(self, kind, xi=0.0, nxi=0, nout=1):
        out = (ctypes.c_double * nout)()
        n = lib.gf_diagnostic(self._ctx, int(kind), float(xi), int(nxi), out, nout)
        check(min(n, 0), self._ctx)
        return list(out)[:n]

    def set_clock(self, t, step):
        check(lib.gf_set_clock(self._ctx, float(t), int(step)), self._ctx)


def stepforward(prob, nsteps=1, diags=None):
    """[FF] ``stepforward!(prob, nsteps)`` and ``stepforward!(prob, diags, nsteps)``.

    Advances ``nsteps`` on the device.  ``prob`` may be an ``Ensemble`` (every member
    advances, batched) or one of its members (that member alone advances).  With ``diags``
    (a ``Diagnostic`` or a list of them; the reference's argument order
    ``stepforward(prob, diags, nsteps)`` is accepted too) the problem advances one step at
    a time and every diagnostic is incremented after each step, as [FF] does."""
    if not isinstance(nsteps, (int, np.integer)):
        diags, nsteps = nsteps, (1 if diags is None else diags)
    if diags is None:
        check(lib.gf_step(prob._ctx, int(nsteps)), prob._ctx)
        return
    if isinstance(diags, Diagnostic):
        diags = [diags]
    for _ in range(int(nsteps)):
        check(lib.gf_step(prob._ctx, 1), prob._ctx)
        for d in diags:
            d.increment()


def set_schedule(prob, p_k1=0, p_k2=0, p_k3=0):
    """FFT points per thread of the three calcN kernels (K1 inverse y-pass, K2 x-pass, K3 forward
    y-pass and stage): 4, 8 or 16, 0 for the built-in choice.  No counterpart in the reference."""
    check(lib.gf_set_schedule(prob._ctx, int(p_k1), int(p_k2), int(p_k3)), prob._ctx)


def schedule(prob):
    """The radix schedule in use, ``(P_K1, P_K2, P_K3)``."""
    p = [ctypes.c_int() for _ in range(3)]
    check(lib.gf_get_schedule(prob._ctx, *[ctypes.byref(x) for x in p]), prob._ctx)
    return tuple(x.value for x in p)


def set_splits(prob, k1_fields=0, k2_halves=0, k3_columns=0):
    """Workgroup splits of a single-layer problem: K1 one field per workgroup row; K2 u q and v q
    in two rows (1: more, shorter workgroups for small grids) or row pairs walked with each
    field's rows loaded into LDS one field ahead (2: nx / P_K2 >= 64, no varying U); K3 one
    column per workgroup instead of a column pair (where a column fills a wave).  No counterpart
    in the reference."""
    check(lib.gf_set_splits(prob._ctx, int(k1_fields), int(k2_halves), int(k3_columns)), prob._ctx)


def splits(prob):
    """The workgroup splits in use, ``(k1_fields, k2_halves, k3_columns)``."""
    v = [ctypes.c_int() for _ in range(3)]
    check(lib.gf_get_splits(prob._ctx, *[ctypes.byref(x) for x in v]), prob._ctx)
    return tuple(x.value for x in v)


def autotune(prob, steps=20):
    """Time whole steps of ``prob`` under every offered schedule (radix and splits; state and
    clock restored), apply the fastest and record it for later problems of the same model, size
    and dtype.  Returns the chosen ``schedule + splits`` 6-tuple and the step time in
    microseconds of each combination tried, keyed the same way."""
    us = (ctypes.c_double * 324)()
    check(lib.gf_autotune(prob._ctx, int(steps), us), prob._ctx)
    times = {}
    pv = (4, 8, 16)
    for c in range(324):
        if math.isfinite(us[c]):
            r = c // 12
            key = (pv[r // 9], pv[(r // 3) % 3], pv[r % 3], (c // 6) % 2, (c // 2) % 3, c % 2)
            times[key] = us[c]
    return schedule(prob) + splits(prob), times


def save_schedules(path):
    """Write the process's tuned schedules as a JSON table."""
    if lib.gf_schedules_save(str(path).encode()) != 0:
        raise _lib.GFError(_lib.GF_EINVAL, lib.gf_schedules_error().decode())


def load_schedules(path):
    """Read a table written by ``save_schedules``; problems created afterwards use its entries."""
    n = lib.gf_schedules_load(str(path).encode())
    if n < 0:
        raise _lib.GFError(n, lib.gf_schedules_error().decode())
    return n


def device_diagnostic(spec, unpack=None):
    """Marks a diagnostic function as computable on the device: ``spec(prob)`` returns
    ``(kind, xi, nxi, nout)`` for ``gf_diagnostic_async``; ``unpack`` turns one recorded row of
    ``nout`` values into what the function itself returns."""
    def wrap(fn):
        fn._gf_device = spec
        fn._gf_unpack = unpack
        return fn
    return wrap


class Diagnostic:
    """[FF] ``Diagnostic(calc, prob; freq=1, nsteps=100, ndata=ceil((nsteps+1)/freq))``.

    ``calc(prob)`` is recorded at construction and then by ``increment()`` whenever
    ``prob.clock.step`` is a multiple of ``freq``; ``d.i`` counts the records and ``d.t``,
    ``d.steps`` hold their times and step numbers.  Indexing is 0-based (``d[0]`` is the
    reference's ``d[1]``).

    For this package's own diagnostics (``energy``, ``enstrophy``, ``energy_work``, ...)
    each value is reduced on the device straight into a device-resident series
    (``gf_diagnostic_async`` into ``gf_device_alloc`` memory): stepping with diagnostics copies
    nothing to the host until ``data`` is read.  Any other callable is evaluated as ``calc(prob)`` after each step.
    The series grows past ``ndata`` when needed, as the reference's does.
    """

    def __init__(self, calc, prob, freq=1, nsteps=100, ndata=None):
        if freq < 1:
            raise ValueError("freq must be >= 1")
        self.calc, self.prob, self.freq = calc, prob, int(freq)
        n = int(math.ceil((nsteps + 1) / freq)) if ndata is None else int(ndata)
        n = max(n, 1)
        self.i = 0
        self.t = np.zeros(n)
        self.steps = np.zeros(n, dtype=np.int64)
        spec = getattr(calc, "_gf_device", None)
        self._spec = spec(prob) if spec is not None and prob._ctx is not None else None
        self._dev = None
        if self._spec is not None:
            self._dev = self._alloc(n)
        else:
            self._host = [None] * n
        self.increment(force=True)

    def _alloc(self, n):
        p = lib.gf_device_alloc(n * self._spec[3] * 8)
        if not p:
            raise MemoryError("gf_device_alloc: device memory for a diagnostic series")
        return p

    def __del__(self):
        dev = getattr(self, "_dev", None)
        if not dev:
            return
        self._dev = None
        try:
            if getattr(self.prob, "_ctx", None) is not None:
                self.prob.synchronize()        # no record may still be in flight
        except Exception:
            pass                               # (hipFree below waits for the device anyway)
        finally:
            try:
                lib.gf_device_free(ctypes.c_void_p(dev))
            except Exception:
                pass

    @property
    def ndata(self):
        return len(self.t)

    def _grow(self):
        n = 2 * self.ndata
        self.t = np.concatenate([self.t, np.zeros(n - len(self.t))])
        self.steps = np.concatenate([self.steps, np.zeros(n - len(self.steps), dtype=np.int64)])
        if self._spec is not None:
            self.prob.synchronize()
            dev = self._alloc(n)
            check(lib.gf_device_copy(ctypes.c_void_p(dev), ctypes.c_void_p(self._dev), self.i * self._spec[3] * 8))
            lib.gf_device_free(ctypes.c_void_p(self._dev))
            self._dev = dev
        else:
            self._host += [None] * (n - len(self._host))

    def increment(self, force=False):
        """[FF] ``increment!(diag)``: record ``calc(prob)`` if the step is a multiple of ``freq``."""
        step = self.prob.clock.step
        if not force and step % self.freq != 0:
            return
        if self.i == self.ndata:
            self._grow()
        if self._spec is not None:
            kind, xi, nxi, nout = self._spec
            ptr = self._dev + self.i * nout * 8
            n = lib.gf_diagnostic_async(self.prob._ctx, int(kind), float(xi), int(nxi), ctypes.c_void_p(ptr), nout)
            check(min(n, 0), self.prob._ctx)
        else:
            self._host[self.i] = self.calc(self.prob)
        self.t[self.i] = self.prob.clock.t
        self.steps[self.i] = step
        self.i += 1

    @property
    def data(self):
        """The recorded values, ``data[:i]`` (device series copied out once per read)."""
        if self._spec is None:
            return list(self._host[: self.i])
        self.prob.synchronize()
        rows = np.empty((self.i, self._spec[3]))
        check(lib.gf_device_copy(rows.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self._dev), rows.nbytes))
        unpack = getattr(self.calc, "_gf_unpack", None)
        if unpack is not None:
            return [unpack(r) for r in rows]
        return rows[:, 0].copy() if self._spec[3] == 1 else rows.copy()

    def __getitem__(self, idx):
        return self.data[idx]

    @property
    def value(self):
        """The latest recorded value (the reference's ``diag.value``)."""
        return self.data[self.i - 1]


class Ensemble:
    """Independent problems advanced together on the device.

    ``members`` is a list of ``(constructor, kwargs)`` pairs, for example
    ``(SingleLayerQG.Problem, dict(nx=64, β=10.0, stepper="FilteredRK4", dt=0.01))``;
    sizes may differ.  ``ens.problems[i]`` is an ordinary ``Problem`` bound to member
    ``i`` (``set_q``, ``updatevars``, ``sol`` and diagnostics work on it);
    ``stepforward(ens, n)`` advances every member with one launch per kernel class
    and stage, ``stepforward(ens.problems[i], n)`` advances member ``i`` alone.
    The reference has no such object: an ensemble there is a loop over Problems.
    """

    def __init__(self, members):
        probs = []
        _DEFER.on = True
        try:
            for ctor, kw in members:
                probs.append(ctor(GPU(), **kw))
        finally:
            _DEFER.on = False
        n = len(probs)
        descs = (_lib.ProblemDesc * n)(*[p._desc for p in probs])
        ctx = ctypes.c_void_p()
        check(lib.gf_ctx_create_batch(ctypes.byref(ctx), descs, n))
        self._ctx = ctx
        self.problems = probs
        for i, p in enumerate(probs):
            m = ctypes.c_void_p()
            check(lib.gf_ctx_instance(ctx, i, ctypes.byref(m)), ctx)
            p._bind(m, self)

    def __len__(self):
        return len(self.problems)

    def __getitem__(self, i):
        return self.problems[i]

    @property
    def t(self):
        t = ctypes.c_double()
        check(lib.gf_clock(self._ctx, ctypes.byref(t), None), self._ctx)
        return t.value

    def synchronize(self):
        check(lib.gf_synchronize(self._ctx), self._ctx)

    def use_graphs(self, enable=True):
        """Replay a captured graph of the step in ``stepforward`` (the default for ensembles) or
        launch every kernel (the default for single problems)."""
        check(lib.gf_use_graphs(self._ctx, int(bool(enable))), self._ctx)

    def persistent(self, enable=True):
        """Run each stage's three kernel classes over every member as one persistent cooperative
        launch (``gf_batch_persistent``) instead of three launches; results are identical."""
        check(lib.gf_batch_persistent(self._ctx, int(bool(enable))), self._ctx)

    def _diagnostic(self, kind, xi=0.0, nxi=0, nout=1):
        out = (ctypes.c_double * nout)()
        n = lib.gf_diagnostic(self._ctx, int(kind), float(xi), int(nxi), out, nout)
        check(min(n, 0), self._ctx)
        return list(out)[:n]

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            for p in getattr(self, "problems", ()):
                p._ctx = None
            lib.gf_ctx_destroy(ctx)
        self._ctx = None


def make_desc(model, nx, ny, Lx, Ly, T, stepper, dt, aliased_fraction):
    if stepper not in _lib.STEPPERS:
        raise ValueError(f"unknown stepper {stepper!r}; supported: {sorted(_lib.STEPPERS)}")
    d = _lib.ProblemDesc()
    d.model = model
    if T in (np.float64, float, "Float64"):
        d.dtype = _lib.GF_F64
    elif T in (np.float32, "Float32"):
        d.dtype = _lib.GF_F32
    else:
        raise ValueError("T must be np.float64 or np.float32")
    d.nx, d.ny, d.nlayers = int(nx), int(ny), 1
    d.Lx, d.Ly = float(Lx), float(Ly)
    d.aliased_fraction = float(aliased_fraction)
    d.stepper = _lib.STEPPERS[stepper]
    d.dt = float(dt)
    return d


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def attach_forcing(d, calcF):
    """Translate a calcF argument into the descriptor; returns what ``set_forcing`` takes
    (a fixed table or a ``HarmonicForcing``) or None."""
    if calcF is None:
        d.forcing = _lib.GF_FORCING_NONE
        return None
    if isinstance(calcF, RingForcing):
        d.forcing = _lib.GF_FORCING_RING
        d.kf, d.dkf, d.eps, d.seed, d.instance = calcF.kf, calcF.dkf, calcF.eps, calcF.seed, calcF.instance
        return None
    if isinstance(calcF, HarmonicForcing):
        d.forcing = _lib.GF_FORCING_FIXED
        return calcF
    if callable(calcF):
        raise NotImplementedError(
            "arbitrary calcF closures cannot run on the device path; pass a fixed spectral "
            "array, a HarmonicForcing or a RingForcing")
    d.forcing = _lib.GF_FORCING_FIXED
    return np.asarray(calcF)
