"""ctypes mirror of include/poe2d.h (structs, enums) and the loader of libpoe2d.so.

The library is the product: a missing or unloadable libpoe2d.so raises instead of falling back
to anything else.
"""
import ctypes
import os

POE2D_OK = 0
POE2D_E_INVALID = 1
POE2D_E_CFL = 2
POE2D_E_NOMEM = 3
POE2D_E_HIP = 4
POE2D_E_IO = 5
POE2D_E_UNSTABLE = 6
POE2D_E_STATE = 7
STATUS_NAMES = {
    0: "POE2D_OK", 1: "POE2D_E_INVALID", 2: "POE2D_E_CFL", 3: "POE2D_E_NOMEM",
    4: "POE2D_E_HIP", 5: "POE2D_E_IO", 6: "POE2D_E_UNSTABLE", 7: "POE2D_E_STATE",
}

F32 = 32
F64 = 64

# palette entries up to which schedule 3 (two steps per launch) runs; larger palettes run schedule 0
MAX_PALETTE_TWO_STEP = 28

STRESS_AS_WRITTEN = 0
STRESS_DRAINED = 1
MATERIAL_AS_WRITTEN = 0
MATERIAL_AVERAGED = 1

SRC_PRESSURE_DGAUSS = 0
SRC_FORCE_Y_RICKER = 1

VX = 1
VY = 2
P = 4

FIELD_VX, FIELD_VY, FIELD_WX, FIELD_WY = 0, 1, 2, 3
FIELD_EXX, FIELD_EYY, FIELD_ZETA, FIELD_SXY = 4, 5, 6, 7
FIELD_P, FIELD_SXX, FIELD_SYY = 8, 9, 10
FIELD_NAMES = {
    "vx": FIELD_VX, "vy": FIELD_VY, "wx": FIELD_WX, "wy": FIELD_WY, "exx": FIELD_EXX,
    "eyy": FIELD_EYY, "zeta": FIELD_ZETA, "sxy": FIELD_SXY, "p": FIELD_P, "sxx": FIELD_SXX,
    "syy": FIELD_SYY,
}

c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)


class Grid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("npml", ctypes.c_int),
        ("npower", ctypes.c_double), ("rcoef", ctypes.c_double), ("kmax", ctypes.c_double),
        ("alpha_max", ctypes.c_double), ("f0", ctypes.c_double), ("cp_max", ctypes.c_double),
        ("precision", ctypes.c_int), ("check_cfl", ctypes.c_int),
        ("stress_law", ctypes.c_int), ("staggered_material", ctypes.c_int),
    ]


class Coeffs(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int),
        ("rho", ctypes.c_void_p), ("rhof", ctypes.c_void_p), ("m", ctypes.c_void_p),
        ("mu", ctypes.c_void_p), ("lambdac", ctypes.c_void_p), ("M", ctypes.c_void_p),
        ("alpha", ctypes.c_void_p), ("b", ctypes.c_void_p),
        ("pitch", ctypes.c_size_t),
    ]


class Source(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("ix", ctypes.c_int), ("iz", ctypes.c_int),
        ("amp", ctypes.c_double), ("t0", ctypes.c_double), ("t_cut", ctypes.c_double),
    ]


class Receivers(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("ix", c_int_p), ("iz", c_int_p),
        ("comp_mask", ctypes.c_uint), ("decimate", ctypes.c_int), ("nsamp", ctypes.c_int),
    ]


COEFF_NAMES = ("rho", "rhof", "m", "mu", "lambdac", "M", "alpha", "b")

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpoe2d.so")
_lib = None


def load_library():
    """Load the in-tree libpoe2d.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    _declare(lib)
    _lib = lib
    return lib


def _declare(lib):
    vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    sigs = {
        "poe2d_abi_version": ([], i),
        "poe2d_device_info": ([c_int_p, ctypes.c_char_p, sz], i),
        "poe2d_create": ([ctypes.POINTER(Grid), ctypes.POINTER(Coeffs), ctypes.POINTER(vp)], i),
        "poe2d_set_source": ([vp, ctypes.POINTER(Source)], i),
        "poe2d_set_wavelet": ([vp, c_double_p, i], i),
        "poe2d_set_receivers": ([vp, ctypes.POINTER(Receivers)], i),
        "poe2d_run": ([vp, i], i),
        "poe2d_set_schedule": ([vp, i, i, i], i),
        "poe2d_material_palette": ([vp, c_int_p], i),
        "poe2d_autotune": ([vp, i, c_int_p, c_int_p], i),
        "poe2d_set_graphs": ([vp, i], i),
        "poe2d_clone": ([vp, ctypes.POINTER(vp)], i),
        "poe2d_set_coeffs": ([vp, ctypes.POINTER(Coeffs)], i),
        "poe2d_batch_create": ([i, ctypes.POINTER(vp), ctypes.POINTER(vp)], i),
        "poe2d_batch_set_schedule": ([vp, i, i], i),
        "poe2d_batch_set_graphs": ([vp, i], i),
        "poe2d_batch_run": ([vp, i], i),
        "poe2d_batch_last_run_ms": ([vp, c_double_p, c_int_p], i),
        "poe2d_batch_autotune": ([vp, i, c_int_p, c_int_p], i),
        "poe2d_batch_destroy": ([vp], None),
        "poe2d_mlp_create": ([i, c_int_p, i, i, ctypes.c_uint64, ctypes.POINTER(vp)], i),
        "poe2d_mlp_get_params": ([vp, i, i, c_float_p, c_float_p], i),
        "poe2d_mlp_set_params": ([vp, i, i, c_float_p, c_float_p], i),
        "poe2d_mlp_get_bf16": ([vp, i, i, vp, vp], i),
        "poe2d_mlp_forward": ([vp, i, i, c_float_p, c_float_p], i),
        "poe2d_mlp_gradients": ([vp, i, i, c_float_p, c_float_p, c_float_p, ctypes.POINTER(c_float_p),
                                 ctypes.POINTER(c_float_p)], i),
        "poe2d_mlp_set_dataset": ([vp, i, c_float_p, c_float_p], i),
        "poe2d_mlp_train": ([vp, i, i, i, c_int_p, ctypes.c_float, c_float_p], i),
        "poe2d_mlp_predict": ([vp, i, c_float_p, c_float_p, c_float_p], i),
        "poe2d_mlp_set_output_scale": ([vp, c_float_p], i),
        "poe2d_mlp_train_many": ([vp, i, c_int_p, i, i, c_int_p, ctypes.c_float, c_float_p], i),
        "poe2d_mlp_time": ([vp, i, i, c_double_p, c_double_p], i),
        "poe2d_gemm_bf16": ([i, i, i, c_float_p, c_float_p, c_float_p, i], i),
        "poe2d_gemm_bf16_noalloc": ([i, i, i, c_float_p, c_float_p, c_float_p, i], i),
        "poe2d_mlp_destroy": ([vp], None),
        "poe2d_host_alloc": ([ctypes.c_size_t, ctypes.POINTER(vp)], i),
        "poe2d_host_free": ([vp], None),
        "poe2d_device_alloc": ([ctypes.c_size_t, ctypes.POINTER(vp)], i),
        "poe2d_device_free": ([vp], None),
        "poe2d_copy": ([vp, vp, ctypes.c_size_t], i),
        "poe2d_read_input1": ([ctypes.c_char_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p], i),
        "poe2d_read_input2": ([ctypes.c_char_p, c_double_p], i),
        "poe2d_read_grid_f32": ([ctypes.c_char_p, i, i, i, c_float_p], i),
        "poe2d_write_grid_f32": ([ctypes.c_char_p, i, i, c_float_p], i),
        "poe2d_write_traces": ([ctypes.c_char_p, i, i, c_float_p, i], i),
        "poe2d_reference_receivers": ([i, i, i, i, i, i, i, i, i, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p], i),
        "poe2d_last_run_ms": ([vp, c_double_p, c_int_p], i),
        "poe2d_step_count": ([vp, c_int_p], i),
        "poe2d_seismograms": ([vp, vp, sz], i),
        "poe2d_snapshot": ([vp, i, vp, sz], i),
        "poe2d_set_field": ([vp, i, vp, sz], i),
        "poe2d_checkpoint_size": ([vp, ctypes.POINTER(sz)], i),
        "poe2d_checkpoint": ([vp, vp, sz], i),
        "poe2d_restore": ([vp, vp, sz], i),
        "poe2d_reset": ([vp], i),
        "poe2d_destroy": ([vp], None),
        "poe2d_last_error": ([], ctypes.c_char_p),
        "poe2d_cpml_profile": ([i, i, d, d, i, d, d, d, d, d, i] + [c_double_p] * 6, i),
        "poe2d_material_biot": ([sz] + [c_double_p] * 18, i),
        "poe2d_material_reference": ([sz, c_float_p, c_float_p, c_float_p, c_float_p, i]
                                     + [c_float_p] * 14, i),
        "poe2d_wavelet": ([i, d, d, d, d, d, i, c_double_p], i),
        "poe2d_patchy_mask": ([i, i, d, d, d, ctypes.c_uint64, d, d,
                               ctypes.POINTER(ctypes.c_uint8), c_double_p], i),
        "poe2d_patchy_lattice": ([i, i, i, i, ctypes.POINTER(ctypes.c_uint8)], i),
        "poe2d_patchy_samples": ([i, i, d, d, ctypes.c_uint64, i, c_double_p, c_double_p, i,
                                  ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint8), i], i),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


class Poe2dError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def check(lib, status, what=""):
    if status != POE2D_OK:
        msg = lib.poe2d_last_error()
        raise Poe2dError(status, (what + ": " if what else "") + (msg.decode() if msg else ""))
