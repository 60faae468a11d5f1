"""Patchy-saturation media: seeded random disks (libpoe2d) and the lattice layouts of
patchymodel.m."""
import ctypes

import numpy as np

from . import _abi


def patchy_mask(nx, nz, dx, dz, gas_fraction, seed, rmin, rmax):
    """Bool mask on (nz+1, nx+1) points: disks with uniform random centres and log-uniform radii
    in [rmin, rmax] metres, painted until at least `gas_fraction` of the points are covered."""
    lib = _abi.load_library()
    mask = np.zeros((nz + 1, nx + 1), dtype=np.uint8)
    realised = ctypes.c_double()
    _abi.check(lib, lib.poe2d_patchy_mask(nx, nz, dx, dz, gas_fraction, ctypes.c_uint64(seed), rmin, rmax,
                                          mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                          ctypes.byref(realised)), "patchy_mask")
    return mask.astype(bool)


def patchy_lattice(kind, rows=240, cols=5000, r=8):
    """One of the six lattice layers of patchymodel.m (1 inside a disk, 0 outside), rows x cols."""
    lib = _abi.load_library()
    layer = np.zeros((rows, cols), dtype=np.uint8)
    _abi.check(lib, lib.poe2d_patchy_lattice(kind, rows, cols, r,
                                             layer.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
               "patchy_lattice")
    return layer


def patchymodel_medium(kind, rows=240, cols=5000, model_rows=2800, first_row=2000, r=8):
    """Medium-index map of patchymodel.m for one lattice type: 1 everywhere, and in rows
    [first_row, first_row+rows) value 2 inside disks and 3 between them."""
    layer = patchy_lattice(kind, rows, cols, r)
    model = np.ones((model_rows, cols), dtype=np.float32)
    model[first_row:first_row + rows] = np.where(layer == 1, 2.0, 3.0)
    return model
