"""Python handle over the poe2d C ABI (include/poe2d.h).

`Problem` describes one simulation: grid and C-PML parameters, the eight coefficient arrays, the
source and the receivers.  `Solver` creates a libpoe2d handle for it, runs steps and returns the
traces and snapshots as numpy arrays.  The same call sequence the reference's main() performs
(set-up, upload, stepping, gather, D2H) maps onto create / set_source / set_receivers / run /
seismograms.
"""
import ctypes
import dataclasses
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _abi
from ._abi import COEFF_NAMES


@dataclass
class SourceSpec:
    kind: int = _abi.SRC_FORCE_Y_RICKER
    ix: int = 0
    iz: int = 0
    amp: float = 1.0e10
    t0: float = 0.0      # <= 0: 1/f0
    t_cut: float = 0.0   # <= 0: no cut


@dataclass
class ReceiverSpec:
    ix: np.ndarray
    iz: np.ndarray
    comp_mask: int = _abi.VX | _abi.VY | _abi.P
    decimate: int = 1
    nsamp: int = 0

    @property
    def n(self):
        return int(len(self.ix))

    @property
    def ncomp(self):
        return bin(self.comp_mask & 7).count("1")


@dataclass
class Problem:
    nx: int
    nz: int
    dx: float
    dz: float
    dt: float
    nt: int
    coeffs: dict                 # name -> (nz+1, nx+1) array, names in COEFF_NAMES
    source: Optional[SourceSpec] = None
    receivers: Optional[ReceiverSpec] = None
    npml: int = 20
    npower: float = 2.0
    rcoef: float = 1.0e-3
    kmax: float = 1.0
    alpha_max: float = 0.0       # <= 0: pi*f0
    f0: float = 80.0
    cp_max: float = 0.0          # <= 0: from the coefficients
    check_cfl: int = 1
    stress_law: int = _abi.STRESS_AS_WRITTEN
    staggered_material: int = _abi.MATERIAL_AS_WRITTEN
    wavelet: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def npoints(self):
        return (self.nx + 1) * (self.nz + 1)

    def grid_struct(self, precision):
        return _abi.Grid(self.nx, self.nz, self.dx, self.dz, self.dt, self.npml, self.npower,
                         self.rcoef, self.kmax, self.alpha_max, self.f0, self.cp_max,
                         precision, self.check_cfl, self.stress_law, self.staggered_material)


class CoeffBuffers:
    """Keeps the coefficient arrays alive while a Coeffs struct points at them."""

    def __init__(self, coeffs, nx, nz, dtype=np.float32):
        self.arrays = {}
        for name in COEFF_NAMES:
            a = np.ascontiguousarray(coeffs[name], dtype=dtype)
            if a.shape != (nz + 1, nx + 1):
                raise ValueError(f"coefficient {name} has shape {a.shape}, need {(nz + 1, nx + 1)}")
            self.arrays[name] = a
        code = _abi.F32 if dtype == np.float32 else _abi.F64
        self.struct = _abi.Coeffs(code, *[self.arrays[n].ctypes.data for n in COEFF_NAMES], nx + 1)


class HandleSolver:
    """A solver handle behind a C ABI with the poe2d call set; `prefix` selects the library's
    symbol family."""

    prefix = "poe2d_"

    def __init__(self, lib, problem: Problem, precision=_abi.F32, coeff_dtype=None):
        self._lib = lib
        self.problem = problem
        self.precision = precision
        self.real = np.float32 if precision == _abi.F32 else np.float64
        cdt = coeff_dtype or (np.float32 if precision == _abi.F32 else np.float64)
        bufs = CoeffBuffers(problem.coeffs, problem.nx, problem.nz, cdt)
        grid = problem.grid_struct(precision)
        h = ctypes.c_void_p()
        self._check(self._fn("create")(ctypes.byref(grid), ctypes.byref(bufs.struct), ctypes.byref(h)),
                    "create")
        self._h = h
        if problem.source is not None:
            self.set_source(problem.source)
        if problem.wavelet is not None:
            self.set_wavelet(problem.wavelet)
        if problem.receivers is not None:
            self.set_receivers(problem.receivers, problem.nt)

    # -- plumbing ----------------------------------------------------------------------------
    def _fn(self, name):
        return getattr(self._lib, self.prefix + name)

    def _check(self, status, what):
        if status != 0:
            msg = self._fn("last_error")()
            raise _abi.Poe2dError(status, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- API ---------------------------------------------------------------------------------
    def set_source(self, src: SourceSpec):
        s = _abi.Source(src.kind, src.ix, src.iz, src.amp, src.t0, src.t_cut)
        self._check(self._fn("set_source")(self._h, ctypes.byref(s)), "set_source")

    def set_wavelet(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        self._check(self._fn("set_wavelet")(self._h, w.ctypes.data_as(_abi.c_double_p), len(w)),
                    "set_wavelet")

    def set_receivers(self, rec: ReceiverSpec, nt=None):
        self._rix = np.ascontiguousarray(rec.ix, dtype=np.int32)
        self._riz = np.ascontiguousarray(rec.iz, dtype=np.int32)
        nsamp = rec.nsamp or -(-(nt or self.problem.nt) // rec.decimate)
        self._rec = rec
        self._nsamp = nsamp
        r = _abi.Receivers(len(self._rix), self._rix.ctypes.data_as(_abi.c_int_p),
                           self._riz.ctypes.data_as(_abi.c_int_p), rec.comp_mask, rec.decimate, nsamp)
        self._check(self._fn("set_receivers")(self._h, ctypes.byref(r)), "set_receivers")

    def run(self, nsteps=None):
        n = self.problem.nt if nsteps is None else nsteps
        self._check(self._fn("run")(self._h, int(n)), "run")

    def step_count(self):
        n = ctypes.c_int()
        self._check(self._fn("step_count")(self._h, ctypes.byref(n)), "step_count")
        return n.value

    def seismograms(self):
        """Traces as an array [ncomp, nrec, nsamp] (components VX, VY, P within comp_mask)."""
        rec = self._rec
        out = np.zeros((rec.ncomp, rec.n, self._nsamp), dtype=self.real)
        self._check(self._fn("seismograms")(self._h, out.ctypes.data, out.nbytes), "seismograms")
        return out

    def set_field(self, fieldname, values):
        """Overwrite one state field (vx .. sxy) of the current state with an (nz+1, nx+1) array."""
        f = _abi.FIELD_NAMES[fieldname] if isinstance(fieldname, str) else fieldname
        a = np.ascontiguousarray(values, dtype=self.real)
        if a.shape != (self.problem.nz + 1, self.problem.nx + 1):
            raise ValueError(f"set_field: shape {a.shape}, need {(self.problem.nz + 1, self.problem.nx + 1)}")
        self._check(self._fn("set_field")(self._h, f, a.ctypes.data, a.nbytes), "set_field")

    def snapshot(self, fieldname):
        f = _abi.FIELD_NAMES[fieldname] if isinstance(fieldname, str) else fieldname
        out = np.zeros((self.problem.nz + 1, self.problem.nx + 1), dtype=self.real)
        self._check(self._fn("snapshot")(self._h, f, out.ctypes.data, out.nbytes), "snapshot")
        return out


class Solver(HandleSolver):
    """A libpoe2d (GPU) handle."""

    prefix = "poe2d_"

    def __init__(self, problem: Problem, precision=_abi.F32, coeff_dtype=None):
        super().__init__(_abi.load_library(), problem, precision, coeff_dtype)

    def last_run_ms(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        self._check(self._lib.poe2d_last_run_ms(self._h, ctypes.byref(ms), ctypes.byref(n)),
                    "last_run_ms")
        return ms.value, n.value

    def reset(self):
        self._check(self._lib.poe2d_reset(self._h), "reset")

    def set_schedule(self, schedule=0, block_w=256, tile_rows=14):
        """0: single fused launch per step (tiles of block_w-8 columns x tile_rows rows), reading
        the coefficients through the material palette when the grid has one;
        1: separate stress and velocity launches;
        2: as 0, always reading the per-point coefficient arrays."""
        self._check(self._lib.poe2d_set_schedule(self._h, schedule, block_w, tile_rows), "set_schedule")

    def set_graphs(self, on=True):
        """Replay captured 32-step graphs (default) or launch every step directly."""
        self._check(self._lib.poe2d_set_graphs(self._h, int(bool(on))), "set_graphs")

    def clone(self, source=None, receivers=None, nt=None):
        """A handle on the same grid and coefficients (shared on the device) with its own
        wavefield, and `source` / `receivers` set on it."""
        h = ctypes.c_void_p()
        self._check(self._lib.poe2d_clone(self._h, ctypes.byref(h)), "clone")
        c = Solver.__new__(Solver)
        c._lib = self._lib
        c.problem = dataclasses.replace(self.problem, source=source, receivers=receivers, wavelet=None,
                                        nt=nt or self.problem.nt)
        c.precision = self.precision
        c.real = self.real
        c._h = h
        if source is not None:
            c.set_source(source)
        if receivers is not None:
            c.set_receivers(receivers, c.problem.nt)
        return c

    def set_coeffs(self, coeffs, coeff_dtype=None):
        """Replace the grid's coefficients (dict of (nz+1, nx+1) arrays) and reset the handle: the
        same handle then solves another model of the same grid without device allocations."""
        cdt = coeff_dtype or self.real
        bufs = CoeffBuffers(coeffs, self.problem.nx, self.problem.nz, cdt)
        self._check(self._lib.poe2d_set_coeffs(self._h, ctypes.byref(bufs.struct)), "set_coeffs")
        self.problem = dataclasses.replace(self.problem, coeffs=coeffs)

    def autotune(self, nsteps=10):
        """Time the single-pass tile shapes on this grid, keep the fastest, reset the handle;
        returns (block_w, tile_rows)."""
        w, h = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.poe2d_autotune(self._h, nsteps, ctypes.byref(w), ctypes.byref(h)), "autotune")
        return w.value, h.value

    def material_palette(self):
        """Distinct coefficient tuples of the grid when at most 256, else 0."""
        n = ctypes.c_int()
        self._check(self._lib.poe2d_material_palette(self._h, ctypes.byref(n)), "material_palette")
        return n.value

    def checkpoint(self):
        n = ctypes.c_size_t()
        self._check(self._lib.poe2d_checkpoint_size(self._h, ctypes.byref(n)), "checkpoint_size")
        blob = np.zeros(n.value, dtype=np.uint8)
        self._check(self._lib.poe2d_checkpoint(self._h, blob.ctypes.data, blob.nbytes), "checkpoint")
        return blob

    def restore(self, blob):
        blob = np.ascontiguousarray(blob, dtype=np.uint8)
        self._check(self._lib.poe2d_restore(self._h, blob.ctypes.data, blob.nbytes), "restore")


class Batch:
    """Several handles advanced in lock-step, one launch per time step over all of them (ragged
    grids allowed); each keeps its own results."""

    def __init__(self, solvers):
        self._lib = _abi.load_library()
        self.members = list(solvers)
        arr = (ctypes.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = ctypes.c_void_p()
        _abi.check(self._lib, self._lib.poe2d_batch_create(len(self.members), arr, ctypes.byref(h)), "batch_create")
        self._h = h

    def set_schedule(self, block_w=256, tile_rows=14):
        _abi.check(self._lib, self._lib.poe2d_batch_set_schedule(self._h, block_w, tile_rows), "batch_set_schedule")

    def set_graphs(self, on=True):
        _abi.check(self._lib, self._lib.poe2d_batch_set_graphs(self._h, int(bool(on))), "batch_set_graphs")

    def run(self, nsteps):
        _abi.check(self._lib, self._lib.poe2d_batch_run(self._h, int(nsteps)), "batch_run")

    def last_run_ms(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        _abi.check(self._lib, self._lib.poe2d_batch_last_run_ms(self._h, ctypes.byref(ms), ctypes.byref(n)),
                   "batch_last_run_ms")
        return ms.value, n.value

    def autotune(self, nsteps=10):
        w, t = ctypes.c_int(), ctypes.c_int()
        _abi.check(self._lib, self._lib.poe2d_batch_autotune(self._h, nsteps, ctypes.byref(w), ctypes.byref(t)),
                   "batch_autotune")
        return w.value, t.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.poe2d_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
