"""Reference-format input and output through the C ABI, and the `poe2d-run` program.

The reference program reads input1.txt (NSTEP ISOURCE JSOURCE a0_pos a1_pos), input2.txt (f0)
and four float32 model grids, and writes vz%05i.vzo / vx%05i.vxo snapshots and ASCII seismograms
(seismic_CPML_2D_poroelastic_withCUDA.cu:472-1905).  `poe2d-run` (built beside libpoe2d.so from
csrc/driver.cpp) does the same run on the GPU; the functions here expose its building blocks.
"""
import ctypes
import os

import numpy as np

from . import _abi

RUNNER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "poe2d-run")
X_FASTEST, Z_FASTEST = 0, 1


def _b(path):
    return os.fsencode(path)


def read_input1(path):
    """(nstep, isource, jsource, a0_pos, a1_pos) from input1.txt."""
    lib = _abi.load_library()
    v = [ctypes.c_int() for _ in range(5)]
    _abi.check(lib, lib.poe2d_read_input1(_b(path), *[ctypes.byref(x) for x in v]), "read_input1")
    return tuple(x.value for x in v)


def read_input2(path):
    """f0 from input2.txt (read as a float, as the reference does)."""
    lib = _abi.load_library()
    f = ctypes.c_double()
    _abi.check(lib, lib.poe2d_read_input2(_b(path), ctypes.byref(f)), "read_input2")
    return f.value


def read_grid(path, nx1, nz1, layout=X_FASTEST):
    """A float32 grid file as an (nz1, nx1) array (x fastest in memory)."""
    lib = _abi.load_library()
    out = np.zeros((nz1, nx1), dtype=np.float32)
    _abi.check(lib, lib.poe2d_read_grid_f32(_b(path), nx1, nz1, layout, out.ctypes.data_as(_abi.c_float_p)),
               "read_grid_f32")
    return out


def write_grid(path, arr):
    """An (nz1, nx1) array as a float32 file, x fastest (the .vzo / .vxo format)."""
    lib = _abi.load_library()
    a = np.ascontiguousarray(arr, dtype=np.float32)
    _abi.check(lib, lib.poe2d_write_grid_f32(_b(path), a.shape[1], a.shape[0], a.ctypes.data_as(_abi.c_float_p)),
               "write_grid_f32")


def write_traces(path, arr, one_line=False):
    """A (rows, cols) table (or, with one_line, a 1-D trace) in the reference's ASCII format."""
    lib = _abi.load_library()
    a = np.ascontiguousarray(arr, dtype=np.float32)
    rows, cols = (1, a.size) if one_line else a.shape
    _abi.check(lib, lib.poe2d_write_traces(_b(path), rows, cols, a.ctypes.data_as(_abi.c_float_p), int(one_line)),
               "write_traces")


def read_traces(path):
    """Values of an ASCII trace file: a (rows, cols) array for a table, 1-D for a single line."""
    text = open(path).read()
    lines = [ln for ln in text.split("\n") if ln.strip()]
    rows = [np.array(ln.split(), dtype=np.float64) for ln in lines]
    if not text.endswith("\n") and len(rows) == 1:
        return rows[0]
    return np.array(rows)


def reference_receivers(nx, nz, isource, jsource, a0_pos, a1_pos, panj=122, panj2=680):
    """The reference layout: (ix, iz, n_vsp, n_hor, n_pts); positions are VSP line, horizontal
    line, then the a0_pos, a1_pos and JSOURCE point traces."""
    lib = _abi.load_library()
    nv, nh, npt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    args = (nx, nz, isource, jsource, a0_pos, a1_pos, panj, panj2)
    _abi.check(lib, lib.poe2d_reference_receivers(*args, 0, None, None, ctypes.byref(nv), ctypes.byref(nh),
                                                  ctypes.byref(npt)), "reference_receivers")
    n = nv.value + nh.value + npt.value
    ix = np.zeros(n, dtype=np.int32)
    iz = np.zeros(n, dtype=np.int32)
    _abi.check(lib, lib.poe2d_reference_receivers(*args, n, ix.ctypes.data_as(_abi.c_int_p),
                                                  iz.ctypes.data_as(_abi.c_int_p), ctypes.byref(nv),
                                                  ctypes.byref(nh), ctypes.byref(npt)), "reference_receivers")
    return ix, iz, nv.value, nh.value, npt.value


def output_names(f0):
    """File names of the seismograms for frequency f0 (the reference's naming by int(f0))."""
    fi = int(f0)
    names = {"vsp_vz": f"recvspvz{fi:08d}.vsp", "vsp_vx": f"recvspvx{fi:08d}.vsp",
             "hor_vz": f"rechorvz{fi:08d}.hor", "hor_vx": f"rechorvx{fi:08d}.hor"}
    if f0 < 1.0:
        names.update(src="recsrc%08f.src" % f0, ats="recats%08f.ats" % f0, bwh="recbwh%08f.bwh" % f0)
    else:
        names.update(src=f"recsrc{fi:08d}.src", ats=f"recats{fi:08d}.ats", bwh=f"recbwh{fi:08d}.bwh")
    return names
