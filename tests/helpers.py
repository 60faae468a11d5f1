"""Shared test helpers: small problems, Biot closed forms, trace error measure."""
import numpy as np

from poe2d_amd.configs import BRINE, GAS, biot_fields
from poe2d_amd.solver import Problem


def oracle_coeffs(nx, nz, mask=None, background=BRINE, patch=GAS):
    from oracle import pyoracle
    return pyoracle.material_biot(**biot_fields(nx, nz, mask, background, patch))


def small_problem(nx=120, nz=120, dx=0.5, dt=5e-5, nt=100, coeffs=None, src=None, rec=None,
                  npml=20, f0=80.0, **kw):
    if coeffs is None:
        coeffs = oracle_coeffs(nx, nz)
    return Problem(nx=nx, nz=nz, dx=dx, dz=dx, dt=dt, nt=nt, coeffs=coeffs, source=src,
                   receivers=rec, npml=npml, f0=f0, **kw)


def random_patchy_mask(nx, nz, frac, seed, rmin=4, rmax=12):
    rng = np.random.default_rng(seed)
    mask = np.zeros((nz + 1, nx + 1), dtype=bool)
    zz, xx = np.mgrid[0:nz + 1, 0:nx + 1]
    while mask.mean() < frac:
        cx, cz = rng.uniform(0, nx), rng.uniform(0, nz)
        r = np.exp(rng.uniform(np.log(rmin), np.log(rmax)))
        mask |= (xx - cx) ** 2 + (zz - cz) ** 2 <= r * r
    return mask


def biot_constants(p):
    """Closed-form Biot moduli for a parameter dict (textbook definitions)."""
    alpha = 1.0 - p["K_fr"] / p["K_s"]
    M = 1.0 / (p["phi"] / p["K_f"] + (alpha - p["phi"]) / p["K_s"])
    rho = (1 - p["phi"]) * p["rho_s"] + p["phi"] * p["rho_f"]
    m = p["tort"] * p["rho_f"] / p["phi"]
    mu = p["mu_fr"]
    lam_c = p["K_fr"] - 2.0 * mu / 3.0 + alpha ** 2 * M
    return dict(alpha=alpha, M=M, rho=rho, m=m, mu=mu, lam_c=lam_c, rhof=p["rho_f"],
                b=p["eta"] / p["kappa"])


def biot_speeds_hf(p):
    """Fast-P, slow-P and S speeds of Biot's equations without viscous coupling (b = 0), for the
    stress law sigma_xx = (lambda_c + 2 mu) e_xx + lambda_c e_yy - alpha p,
    p = -alpha M (e_xx + e_yy) + M zeta: roots of det(K - v^2 Mass) = 0 with
    K = [[H, C], [C, M]], H = lambda_c + 2 mu + alpha^2 M, C = alpha M, Mass = [[rho, rho_f], [rho_f, m]]."""
    c = biot_constants(p)
    H = c["lam_c"] + 2 * c["mu"] + c["alpha"] ** 2 * c["M"]
    C = c["alpha"] * c["M"]
    K = np.array([[H, C], [C, c["M"]]])
    Mass = np.array([[c["rho"], c["rhof"]], [c["rhof"], c["m"]]])
    ev = np.sort(np.real(np.linalg.eigvals(np.linalg.solve(Mass, K))))
    vs = np.sqrt(c["mu"] / (c["rho"] - c["rhof"] ** 2 / c["m"]))
    return np.sqrt(ev[1]), np.sqrt(ev[0]), vs


def biot_speeds_lf(p):
    """Low-frequency (fluid locked, w = 0) P and S speeds: sqrt(H/rho), sqrt(mu/rho)."""
    c = biot_constants(p)
    H = c["lam_c"] + 2 * c["mu"] + c["alpha"] ** 2 * c["M"]
    return np.sqrt(H / c["rho"]), np.sqrt(c["mu"] / c["rho"])


def trace_rel_l2(g, o, floor=1e-3):
    """Per-trace relative L2 error ||g - o|| / max(||o||, floor * max trace norm) over the last
    axis; g, o shaped [..., ntrace, nsamp]."""
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    num = np.linalg.norm(g - o, axis=-1)
    den = np.linalg.norm(o, axis=-1)
    fl = floor * den.max(axis=-1, keepdims=True)
    return num / np.maximum(np.maximum(den, fl), 1e-300)

