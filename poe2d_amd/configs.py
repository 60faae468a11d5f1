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


def config2a(nt=3000, material_fn=None, n=1200, nshots=64, gas_fraction=0.15, seed=1):
    """Configuration 2, batch A: `nshots` Ricker force shots over one (n+1)^2 patchy model.

    Returns (problem of the model with shot 0, list of SourceSpec, ReceiverSpec per shot).  The
    configuration's 10 m spacing (20 points) does not fit 64 shots into the 1201-point model's
    interior between the C-PML strips, so the shots are 9 m (18 points) apart, centred, on the first
    row below the top strip; each shot records 101 receivers two rows below, spread over the
    interior."""
    base = config1(nt=nt, material_fn=material_fn, n=n, gas_fraction=gas_fraction, seed=seed)
    npml = base.npml
    lo, hi = npml + 3, n - npml - 2  # interior columns
    span = 18 * (nshots - 1)
    x0 = lo + max(0, (hi - lo - span) // 2)
    zs = npml + 4
    shots = [SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=int(x0 + 18 * k), iz=zs, amp=1.0e10) for k in range(nshots)]
    rec = line_receivers(lo + 10, (hi - lo - 20) // 100, 101, zs + 2)
    base.source = shots[0]
    base.receivers = rec
    base.meta.update(config="2A", workload=f"biot-cpml {nshots} shots over {n + 1}x{n + 1} patchy gas {gas_fraction}")
    return base, shots, rec


def config2b(nt=3000, material_fn=None, nmodels=32, nmin=300, nmax=1500, seed=1):
    """Configuration 2, batch B: `nmodels` independent patchy models, nx and nz uniform in
    [nmin, nmax] points, gas fraction uniform in [0, 0.3]; Ricker force at the centre, 101
    receivers 40 points below it.  Model k uses seed `seed + k` for its shape and patches."""
    from .patchy import patchy_mask
    probs = []
    for k in range(nmodels):
        rng = np.random.default_rng(seed + k)
        nx, nz = (int(v) for v in rng.integers(nmin, nmax + 1, size=2))
        frac = float(rng.uniform(0.0, 0.3))
        mask = patchy_mask(nx, nz, 0.5, 0.5, frac, seed + k, 2.0, 20.0) if frac > 0 else None
        coeffs = coefficients(biot_fields(nx, nz, mask), material_fn)
        cx, cz = nx // 2, nz // 2
        nrec = 101
        step = max(1, (nx - 60) // nrec)
        x0 = cx - step * (nrec // 2)
        probs.append(Problem(nx=nx, nz=nz, dx=0.5, dz=0.5, dt=5.0e-5, nt=nt, coeffs=coeffs,
                             source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=cx, iz=cz, amp=1.0e10),
                             receivers=line_receivers(x0, step, nrec, min(cz + 40, nz - 25)), npml=20, f0=80.0,
                             meta={"config": "2B", "gas_fraction": frac, "seed": seed + k}))
    return probs


def continuous_model(nx, nz):
    """Biot parameters of a medium that varies continuously in depth and offset (every point its own
    coefficients, so no material palette): the brine rock with porosity falling from 0.30 to 0.18
    with depth, frame moduli rising with it (K_fr 9.6 -> 19.2 GPa, mu_fr 8 -> 14.4 GPa), and a
    lateral undulation of +-0.03 in porosity and +-10 % in the moduli."""
    z = np.linspace(0.0, 1.0, nz + 1)[:, None]
    x = np.linspace(0.0, 1.0, nx + 1)[None, :]
    wave = np.sin(2 * np.pi * (3.0 * x + 1.5 * z))
    out = biot_fields(nx, nz)
    out["phi"] = 0.30 - 0.12 * z + 0.03 * wave
    out["K_fr"] = BRINE["K_fr"] * (1.0 + z) * (1.0 + 0.1 * wave)
    out["mu_fr"] = BRINE["mu_fr"] * (1.0 + 0.8 * z) * (1.0 + 0.1 * wave)
    return {k: np.broadcast_to(v, (nz + 1, nx + 1)).copy() for k, v in out.items()}


def config1_continuous(nt=6000, n=4000, material_fn=None):
    """Configuration 1's geometry (4001 x 4001 points, Ricker force at the centre, 801 receivers 40
    points below it) over continuous_model's medium."""
    coeffs = coefficients(continuous_model(n, n), material_fn)
    c = n // 2
    nrec = 801 if n >= 4000 else max(5, (n - 200) // 4)
    x0 = c - 4 * (nrec // 2)
    return Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5.0e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=c, iz=c, amp=1.0e10),
                   receivers=line_receivers(x0, 4, nrec, c + 40), npml=20, f0=80.0,
                   meta={"config": "1c", "workload": f"biot-cpml {n + 1}x{n + 1} continuous porosity and frame moduli"})


# Per-medium fluid tables for runs of the reference program's grid over synthetic models (K_f, eta,
# kappa, phi, tortuosity, rho_f of media 0..3).  The reference's own tables (porosity 1e-6) put the
# fluid-wave speed sqrt(M / m) of media 0 and 1 near 11 km/s, above that grid's Courant limit.
SYNTHETIC_MEDIA = dict(kf=[2.4e9, 2.25e9, 0.05e9, 0.02e9], eta=[0.003, 0.001, 1.5e-5, 1.0e-5],
                       kappa=[1.5e-13, 1.0e-12, 1.0e-12, 5.0e-13], phi=[0.2, 0.25, 0.25, 0.01],
                       a=[2.0, 2.0, 2.0, 2.0], rhof=[1000.0, 1000.0, 100.0, 160.0])

# The reference program's grid (seismic_CPML_2D_poroelastic_withCUDA.cu:486-510): NX = 13599,
# NZ = 2439, dx = dz = 1.25 m, dt = 95 us, 10 C-PML points.
REFERENCE_GRID = dict(nx=13599, nz=2439, dx=1.25, dt=95e-6, npml=10)


def marmousi_like(nx=REFERENCE_GRID["nx"], nz=REFERENCE_GRID["nz"]):
    """Synthetic float32 model grids in the form of the reference program's inputs, each (nz+1,
    nx+1): medium index 0..3 in undulating layers, c_p rising from 1800 to 4500 m/s with depth plus
    a lateral undulation of +-200 m/s, c_s = c_p / 1.9, saturated density 1900 + 0.12 c_p kg/m^3.
    Every point has its own c_p, so the grid has no material palette."""
    z = np.linspace(0.0, 1.0, nz + 1, dtype=np.float64)[:, None]
    x = np.linspace(0.0, 1.0, nx + 1, dtype=np.float64)[None, :]
    wave = np.sin(2.0 * np.pi * (4.0 * x + 2.0 * z))
    depth = z + 0.03 * wave
    medium = np.select([depth < 0.25, depth < 0.5, depth < 0.75], [1.0, 2.0, 0.0], 3.0)
    cp = 1800.0 + 2700.0 * z + 200.0 * wave
    cs = cp / 1.9
    rho = 1900.0 + 0.12 * cp
    return tuple(np.ascontiguousarray(a, dtype=np.float32) for a in (medium, cp, cs, rho))


def reference_grid(nt=300, f0=30.0, isource=6800, jsource=40, a0_pos=50, a1_pos=70, material_reference_fn=None):
    """The reference program's run on its own grid over marmousi_like's model with SYNTHETIC_MEDIA:
    the pressure source of the reference (derivative of a Gaussian, gated at 0.4 s) at (isource,
    jsource), its 122 VSP, 680 horizontal and 3 point receivers recording v_x and v_z.
    material_reference_fn(cp, cs, rho_sat, medium, tables) replaces the library's front end."""
    from . import refio
    from .material import material_reference
    g = REFERENCE_GRID
    medium, cp, cs, rho = marmousi_like(g["nx"], g["nz"])
    coeffs = (material_reference_fn or material_reference)(cp, cs, rho, medium, SYNTHETIC_MEDIA)
    ix, iz, nv, nh, npt = refio.reference_receivers(g["nx"], g["nz"], isource, jsource, a0_pos, a1_pos)
    return Problem(nx=g["nx"], nz=g["nz"], dx=g["dx"], dz=g["dx"], dt=g["dt"], nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_PRESSURE_DGAUSS, ix=isource, iz=jsource, amp=1e10, t_cut=0.4),
                   receivers=ReceiverSpec(ix=ix, iz=iz, comp_mask=_abi.VX | _abi.VY), npml=g["npml"], f0=f0,
                   cp_max=float(cp.max()), check_cfl=0,
                   meta={"config": "reference grid", "workload": f"biot-cpml {g['nx'] + 1}x{g['nz'] + 1} "
                         "reference grid, synthetic layered model (no palette)"})
