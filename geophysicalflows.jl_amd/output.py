"""[FF] ``Output``, ``saveproblem``, ``saveoutput`` and ``savediagnostic``: snapshots of a
problem's fields in the key layout the reference's examples read back
(``examples/twodnavierstokes_decaying.jl:111-115``, ``examples/singlelayerqg_betaforced.jl:174-231``):

    grid/nx, grid/ny, grid/Lx, grid/Ly, grid/x, grid/y, clock/dt, params/<name>
    snapshots/t/<step>, snapshots/<field>/<step>
    diagnostics/<name>/t, diagnostics/<name>/steps, diagnostics/<name>/data

The container is a NumPy ``.npz`` archive (a zip of ``.npy`` arrays): ``np.load(path)["snapshots/sol/100"]``
reads what ``file["snapshots/sol/100"]`` reads from the reference's JLD2 file.  JLD2 itself
(an HDF5 dialect) is not written: no HDF5 library is available to this package.

The state field ``get_sol`` does not stop the stepper: ``saveoutput`` queues a device snapshot
(``gf_snapshot``) into pinned host memory and returns; the copy runs on a second stream while
later steps compute, and the array is written to the file at the next ``saveoutput`` or at
``close()``.  Other fields are evaluated when ``saveoutput`` is called.
"""
import ctypes
import zipfile

import numpy as np

from ._lib import check, lib


def get_sol(prob):
    """The reference's ``get_sol(prob) = prob.sol``; saved through the asynchronous snapshot path."""
    return prob.sol


def _write(path, items, written):
    with zipfile.ZipFile(path, mode="a", compression=zipfile.ZIP_STORED, allowZip64=True) as zf:
        for key, val in items:
            if key in written:
                raise ValueError(f"{path}: {key} already exists")
            with zf.open(key + ".npy", "w", force_zip64=True) as f:
                np.lib.format.write_array(f, np.asanyarray(val), allow_pickle=False)
            written.add(key)


class Output:
    """[FF] ``Output(prob, filename, (:sol, get_sol), (:u, get_u), ...)``: ``fields`` are
    ``(name, function)`` pairs; ``out[name]`` evaluates one."""

    def __init__(self, prob, path, *fields):
        self.prob = prob
        self.path = str(path)
        self.fields = dict(fields)
        self._written = set()
        self._pending = None        # (key, buffer index, ticket) of the snapshot in flight
        self._nsnap = 0
        self._bufs = []
        if any(f is get_sol for f in self.fields.values()):
            nbytes = prob._sshape_bytes()
            for _ in range(2):
                p = lib.gf_host_alloc(nbytes)
                if not p:
                    self._free()
                    raise MemoryError("gf_host_alloc: pinned host memory for snapshots")
                arr = np.ctypeslib.as_array((ctypes.c_char * nbytes).from_address(p))
                self._bufs.append((p, arr.view(prob._ctype).reshape(prob._sshape)))

    def __getitem__(self, name):
        return self.fields[name](self.prob)

    def _flush_pending(self):
        if self._pending is None:
            return
        key, k, ticket = self._pending
        self._pending = None
        if self.prob._ctx is not None:
            check(lib.gf_snapshot_wait(self.prob._ctx, ticket), self.prob._ctx)
        # (a destroyed context has already drained its copy stream)
        _write(self.path, [(key, self._bufs[k][1])], self._written)

    def flush(self):
        """Write the snapshot still in flight, if any."""
        self._flush_pending()

    def _free(self):
        for p, _ in self._bufs:
            lib.gf_host_free(ctypes.c_void_p(p))
        self._bufs = []

    def close(self):
        self._flush_pending()
        self._free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def saveproblem(out):
    """[FF] ``saveproblem(out)``: the grid, ``clock/dt`` and the scalar parameters."""
    p, g = out.prob, out.prob.grid
    items = [("grid/nx", np.int64(g.nx)), ("grid/ny", np.int64(g.ny)), ("grid/Lx", np.float64(g.Lx)),
             ("grid/Ly", np.float64(g.Ly)), ("grid/x", np.asarray(g.x)), ("grid/y", np.asarray(g.y)),
             ("clock/dt", np.float64(p.clock.dt))]
    for name, v in vars(p.params).items():
        if isinstance(v, (int, float, np.integer, np.floating)) and not isinstance(v, bool):
            items.append((f"params/{name}", np.asarray(v)))
        elif isinstance(v, np.ndarray) and v.dtype != object:
            items.append((f"params/{name}", v))
    _write(out.path, items, out._written)


def saveoutput(out):
    """[FF] ``saveoutput(out)``: ``snapshots/t/<step>`` and ``snapshots/<field>/<step>``."""
    p = out.prob
    step, t = p.clock.step, p.clock.t
    items = [(f"snapshots/t/{step}", np.float64(t))]
    queued = None
    for name, fn in out.fields.items():
        key = f"snapshots/{name}/{step}"
        if fn is get_sol and p._ctx is not None:
            k = out._nsnap & 1
            ticket = lib.gf_snapshot(p._ctx, ctypes.c_void_p(out._bufs[k][0]))
            check(int(min(ticket, 0)), p._ctx)
            out._nsnap += 1
            queued = (key, k, ticket)
        else:
            items.append((key, fn(p)))
    # the previous snapshot used the other buffer; write it before that buffer is reused
    out._flush_pending()
    _write(out.path, items, out._written)
    out._pending = queued


def savediagnostic(diag, diagname, path):
    """[FF] ``savediagnostic(diag, diagname, filename)``."""
    i = diag.i
    data = diag.data
    arr = np.asarray(data) if not isinstance(data[0], tuple) else np.array([np.concatenate(r) for r in data])
    _write(path, [(f"diagnostics/{diagname}/t", diag.t[:i]), (f"diagnostics/{diagname}/steps", diag.steps[:i]),
                  (f"diagnostics/{diagname}/data", arr)], set())
