"""Workload definitions: Biot media and geometry of the benchmark configurations.

Configuration 0: 401 x 401 points, dx = dz = 0.5 m, homogeneous brine sandstone, Ricker vertical
force (f0 = 80 Hz) at the centre, dt = 5e-5 s, nt = 3000, 20-point C-PML (NPOWER 2, R 1e-3), 101
receivers 40 points below the source at x = 50 + 3k recording vx, vy and p.
Configuration 1: 4001 x 4001 points, patchy gas (fraction 0.15) over the brine background,
nt = 6000, 801 receivers at x = 400 + 4k, 40 points below the central source.
"""
import numpy as np

from . import _abi
from .solver import Problem, ReceiverSpec, SourceSpec

# Water-saturated sandstone (the configs' brine rock) and the gas that fills the patches.
BRINE = dict(rho_s=2650.0, rho_f=1000.0, phi=0.25, kappa=1.0e-12, eta=1.0e-3, tort=2.0,
             K_s=37.0e9, K_f=2.25e9, K_fr=9.6e9, mu_fr=8.0e9)
GAS = dict(BRINE, K_f=0.05e9, rho_f=100.0, eta=1.5e-5)
BIOT_NAMES = ("rho_s", "rho_f", "phi", "kappa", "eta", "tort", "K_s", "K_f", "K_fr", "mu_fr")


def biot_fields(nx, nz, gas_mask=None, background=BRINE, patch=GAS):
    """Per-point Biot parameters on (nz+1, nx+1); gas_mask (bool/uint8) selects the patch medium."""
    shape = (nz + 1, nx + 1)
    out = {}
    for k in BIOT_NAMES:
        a = np.full(shape, background[k], dtype=np.float64)
        if gas_mask is not None:
            a[np.asarray(gas_mask, dtype=bool)] = patch[k]
        out[k] = a
    return out


def default_material():
    from .material import material_biot
    return material_biot


def coefficients(params, material_fn=None):
    fn = material_fn or default_material()
    return fn(**params)


def line_receivers(x0, dxr, n, z, comp=_abi.VX | _abi.VY | _abi.P, decimate=1):
    ix = x0 + dxr * np.arange(n, dtype=np.int32)
    iz = np.full(n, z, dtype=np.int32)
    return ReceiverSpec(ix=ix, iz=iz, comp_mask=comp, decimate=decimate)


def config0(nt=3000, material_fn=None, n=400, f0=80.0):
    """Configuration 0 (homogeneous brine rock, 401 x 401 points).  `n` scales the grid for
    quick variants: the source stays at the centre, receivers 40 points below it."""
    coeffs = coefficients(biot_fields(n, n), material_fn)
    c = n // 2
    nrec = 101 if n >= 400 else max(5, (n - 100) // 3)
    x0 = c - 3 * (nrec // 2)
    return Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5.0e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=c, iz=c, amp=1.0e10),
                   receivers=line_receivers(x0, 3, nrec, c + 40), npml=20, f0=f0,
                   meta={"config": 0, "workload": f"biot-cpml {n + 1}x{n + 1} homogeneous brine"})


def config1(nt=6000, material_fn=None, n=4000, gas_fraction=0.15, seed=1, mask=None):
    """Configuration 1 (patchy gas over brine, 4001 x 4001 points)."""
    if mask is None:
        from .patchy import patchy_mask
        mask = patchy_mask(n, n, 0.5, 0.5, gas_fraction, seed, 2.0, 20.0)
    coeffs = coefficients(biot_fields(n, n, mask), material_fn)
    c = n // 2
    nrec = 801 if n >= 4000 else max(5, (n - 200) // 4)
    x0 = c - 4 * (nrec // 2)
    return Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5.0e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=c, iz=c, amp=1.0e10),
                   receivers=line_receivers(x0, 4, nrec, c + 40), npml=20, f0=80.0,
                   meta={"config": 1, "workload": f"biot-cpml {n + 1}x{n + 1} patchy gas {gas_fraction}",
                         "gas_fraction_realised": float(np.mean(mask))})
