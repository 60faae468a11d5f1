"""Host set-up helpers of libpoe2d: material front ends, C-PML profiles, wavelets."""
import numpy as np

from . import _abi

BIOT_INPUTS = ("rho_s", "rho_f", "phi", "kappa", "eta", "tort", "K_s", "K_f", "K_fr", "mu_fr")


def material_biot(**p):
    """Per-point Biot parameters (arrays or scalars, names in BIOT_INPUTS) -> dict of the eight
    coefficient arrays (FP64) the time loop reads."""
    lib = _abi.load_library()
    arrs = np.broadcast_arrays(*[np.asarray(p[k], dtype=np.float64) for k in BIOT_INPUTS])
    shape = arrs[0].shape
    ins = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in arrs]
    outs = [np.zeros(ins[0].size) for _ in range(8)]
    _abi.check(lib, lib.poe2d_material_biot(ins[0].size, *[a.ctypes.data_as(_abi.c_double_p) for a in ins],
                                            *[o.ctypes.data_as(_abi.c_double_p) for o in outs]),
               "material_biot")
    return {k: o.reshape(shape) for k, o in zip(_abi.COEFF_NAMES, outs)}


# Medium tables of the reference (seismic_CPML_2d_poroelastic.cpp:106-116, "elastic condition"):
# K_f, eta, kappa, a per medium index 0..3, with phi = 1e-6 and rho_f = 10.
REFERENCE_TABLES = dict(kf=[2.4e9, 2.4e9, 0.7e9, 0.02e9], eta=[0.003, 0.003, 0.03, 1.0e-5],
                        kappa=[1.5e-17, 500.0e-15, 500.0e-15, 500.0e-15], a=[2.0, 2.0, 2.0, 2.0],
                        phi=[1.0e-6] * 4, rhof=[10.0] * 4)


def material_reference(cp, cs, rho_sat, medium, tables=REFERENCE_TABLES):
    """The reference's own front end: per-point c_p, c_s, saturated density and medium index plus
    per-medium fluid tables -> the eight float32 coefficient arrays (reference float arithmetic)."""
    lib = _abi.load_library()
    ins = [np.ascontiguousarray(a, dtype=np.float32).ravel() for a in (cp, cs, rho_sat, medium)]
    tabs = [np.ascontiguousarray(tables[k], dtype=np.float32)
            for k in ("kf", "eta", "kappa", "phi", "a", "rhof")]
    n = ins[0].size
    outs = [np.zeros(n, dtype=np.float32) for _ in range(8)]
    _abi.check(lib, lib.poe2d_material_reference(
        n, *[a.ctypes.data_as(_abi.c_float_p) for a in ins], len(tabs[0]),
        *[t.ctypes.data_as(_abi.c_float_p) for t in tabs],
        *[o.ctypes.data_as(_abi.c_float_p) for o in outs]), "material_reference")
    return {k: o.reshape(np.shape(cp)) for k, o in zip(_abi.COEFF_NAMES, outs)}


def cpml_profile(axis, n, delta, dt, npml, npower, rcoef, kmax, alpha_max, cp_max, precision=_abi.F32):
    lib = _abi.load_library()
    outs = [np.zeros(n) for _ in range(6)]
    _abi.check(lib, lib.poe2d_cpml_profile(axis, n, delta, dt, npml, npower, rcoef, kmax, alpha_max,
                                           cp_max, precision,
                                           *[o.ctypes.data_as(_abi.c_double_p) for o in outs]),
               "cpml_profile")
    return dict(zip(("a", "b", "K", "a_half", "b_half", "K_half"), outs))


def wavelet(kind, f0, t0, amp, t_cut, dt, n):
    lib = _abi.load_library()
    out = np.zeros(n)
    _abi.check(lib, lib.poe2d_wavelet(kind, f0, t0, amp, t_cut, dt, n, out.ctypes.data_as(_abi.c_double_p)),
               "wavelet")
    return out
