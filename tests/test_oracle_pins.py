"""Pins of the oracle (oracle/poe2d_oracle.cpp) against what mathematics and physics fix.

The reference ships no tests, fixtures or golden vectors, and its CPU file cannot be built here
without writing a stand-in for its missing conio.h, so every check below comes from closed forms,
invariants or Biot's theory, chosen so that a plausible slip (a wrong stencil neighbour, a
dropped coupling term, a flipped sign, a transposed material index) fails one of them.
"""
import numpy as np
import pytest

from oracle import pyoracle
from oracle.pyoracle import OracleSolver
from poe2d_amd import _abi
from poe2d_amd.configs import BRINE, GAS, biot_fields
from poe2d_amd.solver import ReceiverSpec, SourceSpec

from .helpers import (biot_constants, biot_speeds_hf, biot_speeds_lf, oracle_coeffs,
                      random_patchy_mask, small_problem, trace_rel_l2)

F64 = _abi.F64
F32 = _abi.F32


# ----------------------------------------------------------------------------- C-PML profiles
def _profile(axis, n=2001, npml=400, delta=0.5, dt=5e-5, cp=3605.0, f0=80.0, prec=F64):
    return pyoracle.cpml_profile(axis, n, delta, dt, npml, 2.0, 1e-3, 1.0, np.pi * f0, cp, prec)


@pytest.mark.parametrize("axis", [0, 1])
def test_cpml_reflection_integral_and_alpha(axis):
    """Recover d and alpha from (a, b): b = exp(-(d + alpha) dt), a = d (b - 1)/(d + alpha) when
    K = 1.  The theoretical reflection exp(-2/c int_0^L d ds) must equal Rcoef, and alpha must
    fall linearly from alpha_max at the inner edge (plus 0.1 alpha_max on the top layer)."""
    n, npml, delta, dt, cp, f0 = 2001, 400, 0.5, 5e-5, 3605.0, 80.0
    t = _profile(axis, n, npml, delta, dt, cp, f0)
    a, b, K = t["a"], t["b"], t["K"]
    assert np.all(K == 1.0)
    L = npml * delta
    nz = a != 0
    d = np.zeros(n)
    d[nz] = -a[nz] * np.log(b[nz]) / (dt * (b[nz] - 1.0))
    alpha = -np.log(b) / dt - d
    i = np.arange(n)
    s_low = (L + (3 * delta if axis == 0 else 0.0)) - i * delta  # distance into the low layer
    inside = (s_low >= 0) & (s_low <= L)
    s = s_low[inside]
    order = np.argsort(s)
    integral = np.trapezoid(d[inside][order], s[order])
    R = np.exp(-2.0 * integral / cp)
    assert abs(np.log(R) - np.log(1e-3)) < 1e-4 * abs(np.log(1e-3))
    extra = 0.1 * np.pi * f0 if axis == 1 else 0.0
    np.testing.assert_allclose(alpha[inside], np.pi * f0 * (1 - s / L) + extra, rtol=1e-6, atol=1e-6)
    # high-index layer (right / bottom): origin (n-1-2) delta - L, no extra alpha
    s_high = i * delta - ((n - 3) * delta - L)
    ins2 = (s_high >= 0) & (s_high <= L)
    s2 = s_high[ins2]
    o2 = np.argsort(s2)
    R2 = np.exp(-2.0 * np.trapezoid(d[ins2][o2], s2[o2]) / cp)
    assert abs(np.log(R2) - np.log(1e-3)) < 1e-4 * abs(np.log(1e-3))
    np.testing.assert_allclose(alpha[ins2], np.pi * f0 * (1 - s2 / L), rtol=1e-6, atol=1e-6)
    # interior: no absorption at all
    mid = (s_low < 0) & (s_high < 0)
    assert np.all(a[mid] == 0) and np.all(b[mid] == 1.0)


def test_cpml_layer_extents_config0():
    """Layer extents at configuration 0 (401 points, npml 20) from the origins of the reference
    (left L+3dx, right (NX-2)dx-L, top L, bottom (NZ-2)dz-L): counts of nonzero a."""
    counts = {}
    for axis in (0, 1):
        t = _profile(axis, 401, 20, 0.5, 5e-5, 3605.0, 80.0, prec=F32)
        for key in ("a", "a_half"):
            nzmask = t[key] != 0
            counts[(axis, key, "low")] = int(nzmask[:200].sum())
            counts[(axis, key, "high")] = int(nzmask[200:].sum())
    assert counts[(0, "a", "low")] == 23 and counts[(0, "a_half", "low")] == 23
    assert counts[(0, "a", "high")] == 22 and counts[(0, "a_half", "high")] == 23
    assert counts[(1, "a", "low")] == 20 and counts[(1, "a_half", "low")] == 20
    assert counts[(1, "a", "high")] == 22 and counts[(1, "a_half", "high")] == 23


def test_cpml_float_mirrors_double():
    for axis in (0, 1):
        t32 = _profile(axis, 401, 20, prec=F32)
        t64 = _profile(axis, 401, 20, prec=F64)
        for k in t32:
            np.testing.assert_allclose(t32[k], t64[k], rtol=3e-5, atol=1e-9)


# ----------------------------------------------------------------------------- materials
def test_material_biot_config_values_and_gassmann():
    for p, exp in ((BRINE, dict(rho=2237.5, alpha=0.74054, M=8.0406e9, lambdac=8.6761e9, m=8000.0, b=1e9)),
                   (GAS, dict(rho=2012.5, alpha=0.74054, M=1.9947e8, lambdac=4.3761e9, m=800.0, b=1.5e7))):
        c = pyoracle.material_biot(**p)
        for k, v in exp.items():
            assert abs(float(c[k]) - v) <= 1e-4 * abs(v), (k, float(c[k]), v)
        # Gassmann: K_sat = K_dry + (1 - K_dry/K_0)^2 / (phi/K_fl + (1 - phi)/K_0 - K_dry/K_0^2)
        kd, k0, kf, ph = p["K_fr"], p["K_s"], p["K_f"], p["phi"]
        ksat = kd + (1 - kd / k0) ** 2 / (ph / kf + (1 - ph) / k0 - kd / k0 ** 2)
        assert abs(float(c["lambdac"]) + 2.0 / 3.0 * float(c["mu"]) - ksat) < 1e-9 * ksat
        assert float(c["mu"]) == p["mu_fr"]


def test_material_biot_zero_viscosity_or_permeability():
    c = pyoracle.material_biot(**dict(BRINE, eta=0.0))
    assert float(c["b"]) == 0.0
    c = pyoracle.material_biot(**dict(BRINE, kappa=0.0))
    assert float(c["b"]) == 0.0


def test_material_reference_route_identities():
    """Reference route (K_s = K_b/(1-phi)): alpha = phi, M = K_f/phi, mu = rho c_s^2,
    lambda_c = rho (c_p^2 - 2 c_s^2) + phi K_f, m = a rho_f/phi, b = eta/kappa."""
    rng = np.random.default_rng(3)
    n = 64
    cp = rng.uniform(2500, 4500, n).astype(np.float32)
    cs = (cp / rng.uniform(1.6, 2.0, n)).astype(np.float32)
    rho = rng.uniform(2000, 2600, n).astype(np.float32)
    medium = rng.integers(0, 4, n).astype(np.float32)
    tables = dict(kf=[2.4e9, 2.4e9, 0.7e9, 0.02e9], eta=[0.003, 0.003, 0.03, 1e-5],
                  kappa=[1.5e-17, 5e-13, 5e-13, 5e-13], phi=[0.01, 0.25, 0.25, 0.1],
                  a=[2.0, 2.0, 2.0, 2.0], rhof=[1000.0, 1000.0, 700.0, 160.0])
    c = pyoracle.material_reference(cp, cs, rho, medium, tables)
    mi = medium.astype(int)
    phi = np.array(tables["phi"], dtype=np.float64)[mi]
    kf = np.array(tables["kf"], dtype=np.float64)[mi]
    cp64, cs64, r64 = cp.astype(np.float64), cs.astype(np.float64), rho.astype(np.float64)
    np.testing.assert_allclose(c["alpha"], phi, rtol=2e-5)
    np.testing.assert_allclose(c["M"], kf / phi, rtol=2e-5)
    np.testing.assert_allclose(c["mu"], r64 * cs64 ** 2, rtol=1e-6)
    np.testing.assert_allclose(c["lambdac"], r64 * (cp64 ** 2 - 2 * cs64 ** 2) + phi * kf, rtol=2e-5)
    np.testing.assert_allclose(c["m"], 2.0 * np.array(tables["rhof"])[mi] / phi, rtol=1e-6)
    np.testing.assert_allclose(c["b"], np.array(tables["eta"])[mi] / np.array(tables["kappa"])[mi], rtol=1e-6)
    with pytest.raises(_abi.Poe2dError):
        pyoracle.material_reference(cp, cs, rho, np.full(n, 7, np.float32), tables)


# ----------------------------------------------------------------------------- wavelets
def test_wavelet_closed_forms():
    f0, dt = 80.0, 1e-6
    t0 = 1.0 / f0
    n = int(0.03 / dt)
    t = np.arange(n) * dt
    r = pyoracle.wavelet(_abi.SRC_FORCE_Y_RICKER, f0, 0.0, 1.0, 0.0, dt, n)
    k0 = int(round(t0 / dt))
    assert abs(r[k0] - 1.0) < 1e-12 and r.argmax() == k0
    tz = t0 + 1.0 / (np.pi * f0 * np.sqrt(2.0))  # Ricker zero crossing
    kz = int(tz / dt)
    assert r[kz] > 0 > r[kz + 1]
    assert abs(r.min() + 2.0 * np.exp(-1.5)) < 1e-6  # Ricker minimum -2 e^(-3/2)
    g = pyoracle.wavelet(_abi.SRC_PRESSURE_DGAUSS, f0, 0.0, 1.0, 0.0, dt, n)
    a = (np.pi * f0) ** 2
    assert abs(g[k0]) < 1e-9 * g.max()  # zero at t0 (k0*dt equals t0 up to rounding)
    kp = int(round((t0 + 1.0 / np.sqrt(2 * a)) / dt))  # first-derivative-of-Gaussian peak
    assert g.argmax() == kp and abs(g.max() - np.sqrt(2 * a) * np.exp(-0.5)) < 1e-3 * g.max()
    gc = pyoracle.wavelet(_abi.SRC_PRESSURE_DGAUSS, f0, 0.0, 1e10, 0.012, dt, n)
    assert np.all(gc[t >= 0.012] == 0) and np.allclose(gc[t < 0.012], 1e10 * g[t < 0.012], rtol=1e-14)


# ----------------------------------------------------------------------------- one-step exactness
CUBIC = (1.0, 0.3, -0.02, 0.001)


def cubic(s):
    c0, c1, c2, c3 = CUBIC
    return c0 + c1 * s + c2 * s ** 2 + c3 * s ** 3


def dcubic(s):
    _, c1, c2, c3 = CUBIC
    return c1 + 2 * c2 * s + 3 * c3 * s ** 2


NXS = 100
DX = 0.5
DT = 5e-5
INNER = (slice(30, 71), slice(30, 71))


def _one_step(init, coeffs=None, **opts):
    prob = small_problem(NXS, NXS, DX, DT, nt=1, coeffs=coeffs, npml=10, **opts)
    s = OracleSolver(prob, precision=F64, threads=1)
    for name, arr in init.items():
        s.set_field(name, arr)
    s.run(1)
    return s


def _grids():
    j, i = np.mgrid[0:NXS + 1, 0:NXS + 1].astype(np.float64)
    return i * DX, j * DX  # x, z positions of integer points


@pytest.mark.parametrize("case", ["exx_from_vx", "eyy_from_vy", "zeta_from_wx", "zeta_from_wy",
                                  "sxy_from_vy", "sxy_from_vx"])
def test_stress_stencils_exact_on_cubics(case):
    """Each velocity -> stress difference is the derivative of a cubic, exactly, at the staggered
    point it belongs to: vx at (i, j), vy at (i+1/2, j+1/2), normal stress at (i+1/2, j),
    shear stress at (i, j+1/2)."""
    x, z = _grids()
    mu = BRINE["mu_fr"]
    spec = {
        "exx_from_vx": ("vx", cubic(x), "exx", DT * dcubic(x + DX / 2)),
        "eyy_from_vy": ("vy", cubic(z + DX / 2), "eyy", DT * dcubic(z)),
        "zeta_from_wx": ("wx", cubic(x), "zeta", -DT * dcubic(x + DX / 2)),
        "zeta_from_wy": ("wy", cubic(z + DX / 2), "zeta", -DT * dcubic(z)),
        "sxy_from_vy": ("vy", cubic(x + DX / 2), "sxy", DT * mu * dcubic(x)),
        "sxy_from_vx": ("vx", cubic(z), "sxy", DT * mu * dcubic(z + DX / 2)),
    }[case]
    s = _one_step({spec[0]: spec[1]})
    got = s.snapshot(spec[2])[INNER]
    np.testing.assert_allclose(got, spec[3][INNER], rtol=1e-9, atol=1e-12 * np.abs(spec[3]).max())


def _cn_solve(rho, rhof, m, b, T, P):
    """One Crank-Nicolson step of [rho rho_f; rho_f m] d/dt [v; w] = [T; -P - b w] from rest."""
    A = np.array([[rho, rhof], [rhof, m + DT * b / 2.0]])
    return np.linalg.solve(A, DT * np.array([T, -P]))


@pytest.mark.parametrize("case", ["vx_from_exx", "vy_from_eyy", "vx_from_sxy", "vy_from_sxy",
                                  "vx_from_zeta"])
def test_velocity_stencils_and_biot_coupling(case):
    """Velocities after one step from a cubic stress state equal the solution of Biot's 2x2
    momentum system (solid + relative fluid, viscous term Crank-Nicolson) with the exact
    derivative of the stress at the velocity point."""
    x, z = _grids()
    c = biot_constants(BRINE)
    H = c["lam_c"] + 2 * c["mu"] + c["alpha"] ** 2 * c["M"]
    aM = c["alpha"] * c["M"]
    # (initial field, its value, velocity read, fluid read, T(derivative), P(derivative))
    spec = {
        "vx_from_exx": ("exx", cubic(x + DX / 2), "vx", "wx", H * dcubic(x), -aM * dcubic(x)),
        "vy_from_eyy": ("eyy", cubic(z), "vy", "wy", H * dcubic(z + DX / 2), -aM * dcubic(z + DX / 2)),
        "vx_from_sxy": ("sxy", cubic(z + DX / 2), "vx", "wx", dcubic(z), 0.0 * z),
        "vy_from_sxy": ("sxy", cubic(x), "vy", "wy", dcubic(x + DX / 2), 0.0 * x),
        "vx_from_zeta": ("zeta", cubic(x + DX / 2), "vx", "wx", -aM * dcubic(x), c["M"] * dcubic(x)),
    }[case]
    s = _one_step({spec[0]: spec[1]})
    v = s.snapshot(spec[2])[INNER]
    w = s.snapshot(spec[3])[INNER]
    T = spec[4][INNER]
    P = spec[5][INNER]
    ev = np.zeros_like(T)
    ew = np.zeros_like(T)
    for idx in np.ndindex(T.shape):
        ev[idx], ew[idx] = _cn_solve(c["rho"], c["rhof"], c["m"], c["b"], T[idx], P[idx])
    np.testing.assert_allclose(v, ev, rtol=1e-8, atol=1e-12 * np.abs(ev).max())
    np.testing.assert_allclose(w, ew, rtol=1e-8, atol=1e-12 * np.abs(ew).max())
    other = {"vx": "vy", "vy": "vx"}[spec[2]]
    assert np.abs(s.snapshot(other)[INNER]).max() == 0.0


def test_vy_uses_material_of_the_point_below():
    """The v_y / w_y update takes rho, rho_f, m, b from point (x, z+1) (reference vz_kernel reads
    rho[top]); in a medium layered at row 50 the row-49 update uses the row-50 material."""
    mask = np.zeros((NXS + 1, NXS + 1), dtype=bool)
    mask[50:, :] = True
    coeffs = oracle_coeffs(NXS, NXS, mask)
    x, _ = _grids()
    s = _one_step({"sxy": cubic(x)}, coeffs)
    vy = s.snapshot("vy")
    cg = biot_constants(GAS)
    cb = biot_constants(BRINE)
    T = dcubic(x[49, 40] + DX / 2)
    v_gas = _cn_solve(cg["rho"], cg["rhof"], cg["m"], cg["b"], T, 0.0)[0]
    v_brine = _cn_solve(cb["rho"], cb["rhof"], cb["m"], cb["b"], T, 0.0)[0]
    assert abs(vy[49, 40] - v_gas) < 1e-9 * abs(v_gas)
    assert abs(vy[49, 40] - v_brine) > 1e-3 * abs(v_brine)
    T2 = dcubic(x[48, 40] + DX / 2)
    assert abs(vy[48, 40] - _cn_solve(cb["rho"], cb["rhof"], cb["m"], cb["b"], T2, 0.0)[0]) < 1e-9 * abs(vy[48, 40])


def test_drained_law_one_step():
    """Under the drained law the normal stress uses lambda_o = lambda_c - alpha^2 M, so a cubic
    e_xx drives v_x with the Biot/Gassmann modulus lambda_c + 2 mu (not lambda_c + 2 mu + alpha^2
    M as written), while the pressure gradient (-alpha M e_xx) is the same in both laws."""
    x, _ = _grids()
    c = biot_constants(BRINE)
    aM = c["alpha"] * c["M"]
    for law, H in ((_abi.STRESS_DRAINED, c["lam_c"] + 2 * c["mu"]),
                   (_abi.STRESS_AS_WRITTEN, c["lam_c"] + 2 * c["mu"] + c["alpha"] ** 2 * c["M"])):
        s = _one_step({"exx": cubic(x + DX / 2)}, stress_law=law)
        v = s.snapshot("vx")[INNER]
        T = H * dcubic(x)[INNER]
        P = -aM * dcubic(x)[INNER]
        ev = np.array([_cn_solve(c["rho"], c["rhof"], c["m"], c["b"], t, p)[0] for t, p in zip(T.ravel(), P.ravel())])
        np.testing.assert_allclose(v.ravel(), ev, rtol=1e-8, atol=1e-12 * np.abs(ev).max())


def test_averaged_material_one_step():
    """Averaged staggered material, in a medium layered at row 50 (rows >= 50: gas with a softer
    frame, mu 4 GPa against 8 GPa): the row-49 v_y update uses the means of rho, rho_f, m, b of
    rows 49 and 50, and the row-49 shear stress uses the harmonic mean of their mu; rows away
    from the interface are unchanged."""
    mask = np.zeros((NXS + 1, NXS + 1), dtype=bool)
    mask[50:, :] = True
    soft = dict(GAS, mu_fr=4.0e9)
    coeffs = oracle_coeffs(NXS, NXS, mask, patch=soft)
    x, z = _grids()
    avg = _abi.MATERIAL_AVERAGED
    cb, cg = biot_constants(BRINE), biot_constants(soft)
    mean = {k: 0.5 * (cb[k] + cg[k]) for k in ("rho", "rhof", "m", "b")}
    s = _one_step({"sxy": cubic(x)}, coeffs, staggered_material=avg)
    vy = s.snapshot("vy")
    T = dcubic(x[49, 40] + DX / 2)
    want = _cn_solve(mean["rho"], mean["rhof"], mean["m"], mean["b"], T, 0.0)[0]
    assert abs(vy[49, 40] - want) < 1e-9 * abs(want)
    T2 = dcubic(x[48, 40] + DX / 2)
    assert abs(vy[48, 40] - _cn_solve(cb["rho"], cb["rhof"], cb["m"], cb["b"], T2, 0.0)[0]) < 1e-9 * abs(vy[48, 40])
    # shear stress at (x, z + 1/2) from v_x = cubic(z)
    s = _one_step({"vx": cubic(z)}, coeffs, staggered_material=avg)
    sxy = s.snapshot("sxy")
    h = 2.0 / (1.0 / cb["mu"] + 1.0 / cg["mu"])
    for row, mu in ((49, h), (48, cb["mu"]), (51, cg["mu"])):
        want = DT * mu * dcubic(z[row, 40] + DX / 2)
        assert abs(sxy[row, 40] - want) < 1e-9 * abs(want), (row, sxy[row, 40], want)
    # as written: mu of the point below
    s = _one_step({"vx": cubic(z)}, coeffs)
    want = DT * cg["mu"] * dcubic(z[49, 40] + DX / 2)
    assert abs(s.snapshot("sxy")[49, 40] - want) < 1e-9 * abs(want)


def test_options_rejected_when_unknown():
    for opts in (dict(stress_law=2), dict(staggered_material=-1)):
        with pytest.raises(_abi.Poe2dError):
            OracleSolver(small_problem(40, 40, DX, DT, nt=1, npml=5, **opts), precision=F64)


def test_sources_first_step():
    """Pressure source: p at the source after step 0 equals the wavelet sample; vertical force:
    v_y there equals Biot's momentum solve with T = force."""
    f0 = 80.0
    for kind in (_abi.SRC_PRESSURE_DGAUSS, _abi.SRC_FORCE_Y_RICKER):
        rec = ReceiverSpec(ix=np.array([50]), iz=np.array([50]), comp_mask=_abi.P | _abi.VY, nsamp=2)
        prob = small_problem(NXS, NXS, DX, DT, nt=2, npml=10,
                             src=SourceSpec(kind=kind, ix=50, iz=50, amp=1e10), rec=rec)
        s = OracleSolver(prob, precision=F64, threads=1)
        s.run(1)
        w0 = pyoracle.wavelet(kind, f0, 0.0, 1e10, 0.0, DT, 1)[0]
        if kind == _abi.SRC_PRESSURE_DGAUSS:
            assert s.seismograms()[0, 0, 0] == 0.0  # v_y at slot 0 is the initial state
            assert abs(s.seismograms()[1, 0, 0] - w0) <= 1e-12 * abs(w0)
        else:
            c = biot_constants(BRINE)
            ev = _cn_solve(c["rho"], c["rhof"], c["m"], c["b"], w0, 0.0)[0]
            assert abs(s.snapshot("vy")[50, 50] - ev) <= 1e-9 * abs(ev)


# ----------------------------------------------------------------------------- invariants
def _impulse(s, coeffs, comp, ix, iz):
    """Unit force impulse at a velocity point: [v; w] = M+^{-1} [1; 0] with the point's
    material (v_y points use the material of (x, z+1))."""
    zz = iz + 1 if comp == "vy" else iz
    rho, rhof, m, b = (float(coeffs[k][zz, ix]) for k in ("rho", "rhof", "m", "b"))
    A = np.array([[rho, rhof], [rhof, m + DT * b / 2.0]])
    v, w = np.linalg.solve(A, [1.0, 0.0])
    shape = (s.problem.nz + 1, s.problem.nx + 1)
    fv = np.zeros(shape); fv[iz, ix] = v
    fw = np.zeros(shape); fw[iz, ix] = w
    s.set_field(comp, fv)
    s.set_field("wx" if comp == "vx" else "wy", fw)


@pytest.mark.parametrize("ca,cb", [("vx", "vx"), ("vy", "vy"), ("vx", "vy")])
def test_reciprocity_heterogeneous(ca, cb):
    """In a random poroelastic medium (gas patches in brine rock: rho, rho_f, m, b, moduli all
    vary), the velocity response at B to a unit force impulse at A equals the response at A to
    the impulse at B, sample by sample.  Holds when the velocity stencils are the negative
    transposes of the stress stencils and the stress law and mass matrices are symmetric."""
    n = 120
    mask = random_patchy_mask(n, n, 0.3, seed=11)
    coeffs = oracle_coeffs(n, n, mask)
    A, B = (52, 55), (67, 63)
    nt = 110
    traces = {}
    for (src, scomp), (rcv, rcomp) in (((A, ca), (B, cb)), ((B, cb), (A, ca))):
        mask_bit = _abi.VX if rcomp == "vx" else _abi.VY
        rec = ReceiverSpec(ix=np.array([rcv[0]]), iz=np.array([rcv[1]]), comp_mask=mask_bit, nsamp=nt)
        prob = small_problem(n, n, DX, DT, nt=nt, coeffs=coeffs, rec=rec, npml=20)
        s = OracleSolver(prob, precision=F64, threads=4)
        _impulse(s, coeffs, scomp, src[0], src[1])
        s.run(nt)
        traces[(src, rcv)] = s.seismograms()[0, 0]
    g_ab = traces[(A, B)]
    g_ba = traces[(B, A)]
    scale = np.abs(g_ab).max()
    assert scale > 0
    assert np.abs(g_ab - g_ba).max() <= 1e-9 * scale


def test_mirror_symmetry_vertical_force():
    n, c, nt = 120, 60, 60
    prob = small_problem(n, n, DX, DT, nt=nt, npml=20,
                         src=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=c, iz=c, amp=1e10))
    s = OracleSolver(prob, precision=F64, threads=4)
    s.run(nt)
    vy, vx, p = s.snapshot("vy"), s.snapshot("vx"), s.snapshot("p")
    k = np.arange(1, 35)
    for f, left, right, sign in ((vy, c - k, c + k, 1), (p, c - k, c + k, 1), (vx, c - k, c + 1 + k, -1)):
        scale = np.abs(f).max()
        assert scale > 0
        np.testing.assert_allclose(f[:, right], sign * f[:, left], rtol=0, atol=1e-10 * scale)


def _plane_wave_run(params, inits, nt=400, nx=640, nz=480, x0=150, dt=7.5e-5, **opts):
    coeffs = pyoracle.material_biot(**biot_fields(nx, nz, None, params))
    row = nz // 2
    rec = ReceiverSpec(ix=np.arange(nx + 1, dtype=np.int32), iz=np.full(nx + 1, row, np.int32),
                       comp_mask=_abi.VX | _abi.VY | _abi.P, nsamp=nt)
    prob = small_problem(nx, nz, DX, dt, nt=nt, coeffs=coeffs, rec=rec, npml=5, **opts)
    s = OracleSolver(prob, precision=F64)
    x = np.arange(nx + 1) * DX
    w = 4.0
    for name, shift in inits:
        prof = np.exp(-(((x + shift) - x0 * DX) / w) ** 2)
        s.set_field(name, np.tile(prof, (nz + 1, 1)))
    s.run(nt)
    return s.seismograms(), x0


def _centroid(trace_row, lo, hi):
    seg = trace_row[lo:hi] ** 2
    return (np.arange(lo, hi) * seg).sum() / seg.sum()


def _speed(seis, comp, k1, k2, win1, win2, dt=7.5e-5):
    c1 = _centroid(seis[comp, :, k1], *win1)
    c2 = _centroid(seis[comp, :, k2], *win2)
    return (c2 - c1) * DX / ((k2 - k1) * dt)


@pytest.mark.slow
def test_plane_wave_speeds_low_frequency():
    """Brine rock, b = eta/kappa = 1e9 (Biot critical frequency ~20 kHz, pulse ~300 Hz): the
    fluid is locked to the frame, so P travels at sqrt(H/rho) and S at sqrt(mu/rho)."""
    seis, x0 = _plane_wave_run(BRINE, [("exx", DX / 2), ("vy", DX / 2)])
    vp, vs = biot_speeds_lf(BRINE)
    vp_hf, _, vs_hf = biot_speeds_hf(BRINE)
    got_p = _speed(seis, 0, 150, 399, (x0 + 60, 640), (x0 + 150, 640))
    got_s = _speed(seis, 1, 150, 399, (x0 + 20, 640), (x0 + 60, 640))
    assert abs(got_p / vp - 1) < 2e-3, (got_p, vp)
    assert abs(got_s / vs - 1) < 2e-3, (got_s, vs)
    assert abs(got_p / vp_hf - 1) > 5e-3 and abs(got_s / vs_hf - 1) > 1e-2


@pytest.mark.slow
def test_plane_wave_speed_drained_law_is_gassmann():
    """Drained law, same brine rock in the fluid-locked regime: P travels at Gassmann's
    sqrt((lambda_c + 2 mu) / rho) = sqrt((K_fr + 4/3 mu_fr + alpha^2 M) / rho) (plane strain), about
    8 % slower than with the stress law as written; S is unchanged."""
    seis, x0 = _plane_wave_run(BRINE, [("exx", DX / 2), ("vy", DX / 2)], stress_law=_abi.STRESS_DRAINED)
    c = biot_constants(BRINE)
    K_sat = BRINE["K_fr"] + c["alpha"] ** 2 * c["M"]
    vp = np.sqrt((K_sat + 4.0 / 3.0 * c["mu"]) / c["rho"])
    _, vs = biot_speeds_lf(BRINE)
    vp_written, _ = biot_speeds_lf(BRINE)
    got_p = _speed(seis, 0, 150, 399, (x0 + 60, 640), (x0 + 150, 640))
    got_s = _speed(seis, 1, 150, 399, (x0 + 20, 640), (x0 + 60, 640))
    assert abs(got_p / vp - 1) < 2e-3, (got_p, vp)
    assert abs(got_s / vs - 1) < 2e-3, (got_s, vs)
    assert vp_written / vp - 1 > 0.05


@pytest.mark.slow
def test_plane_wave_speeds_inviscid_biot():
    """b = 0: Biot's fast P, slow P and S waves all propagate, at the roots of
    det(K - v^2 Mass) = 0 and sqrt(mu / (rho - rho_f^2/m))."""
    params = dict(BRINE, eta=0.0)
    seis, x0 = _plane_wave_run(params, [("exx", DX / 2), ("zeta", DX / 2), ("vy", DX / 2)])
    vfast, vslow, vs = biot_speeds_hf(params)
    got_fast = _speed(seis, 0, 150, 399, (x0 + 60, 640), (x0 + 150, 640))
    got_slow = _speed(seis, 2, 250, 399, (x0 + 12, x0 + 70), (x0 + 25, x0 + 90))
    got_s = _speed(seis, 1, 150, 399, (x0 + 20, 640), (x0 + 60, 640))
    assert abs(got_fast / vfast - 1) < 2e-3, (got_fast, vfast)
    assert abs(got_slow / vslow - 1) < 5e-3, (got_slow, vslow)
    assert abs(got_s / vs - 1) < 2e-3, (got_s, vs)


def _small_run(amp=1e10, prec=F32, nt=150, wavelet=None, kind=_abi.SRC_FORCE_Y_RICKER, n=100):
    rec = ReceiverSpec(ix=np.array([30, 50, 70, 50]), iz=np.array([60, 70, 60, 40]),
                       comp_mask=_abi.VX | _abi.VY | _abi.P, nsamp=nt)
    prob = small_problem(n, n, DX, DT, nt=nt, npml=10, wavelet=wavelet,
                         src=SourceSpec(kind=kind, ix=50, iz=50, amp=amp), rec=rec)
    s = OracleSolver(prob, precision=prec, threads=2)
    s.run(nt)
    return s


def test_linearity_exact_scaling_and_zero():
    a = _small_run(1e10).seismograms()
    b = _small_run(2e10).seismograms()
    assert np.array_equal(b, 2 * a)  # scaling by 2 is exact in binary floating point
    assert np.abs(a).max() > 0
    z = _small_run(0.0).seismograms()
    assert np.all(z == 0)


def test_time_shift_invariance_bitwise():
    nt, k = 150, 7
    w = pyoracle.wavelet(_abi.SRC_FORCE_Y_RICKER, 80.0, 0.0, 1.0, 0.0, DT, nt)
    a = _small_run(1e10, wavelet=w, nt=nt).seismograms()
    b = _small_run(1e10, wavelet=np.concatenate([np.zeros(k), w]), nt=nt).seismograms()
    assert np.array_equal(b[..., k:], a[..., :nt - k])
    assert np.all(b[..., :k] == 0)


def test_gather_equals_snapshot():
    s = _small_run(nt=40, kind=_abi.SRC_PRESSURE_DGAUSS)
    seis = s.seismograms()
    rec = s._rec
    s2 = _small_run(nt=23, kind=_abi.SRC_PRESSURE_DGAUSS)
    vx, vy = s2.snapshot("vx"), s2.snapshot("vy")
    for r in range(rec.n):
        assert seis[0, r, 23] == vx[rec.iz[r], rec.ix[r]]
        assert seis[1, r, 23] == vy[rec.iz[r], rec.ix[r]]
    s3 = _small_run(nt=24, kind=_abi.SRC_PRESSURE_DGAUSS)
    p = s3.snapshot("p")
    for r in range(rec.n):
        assert seis[2, r, 23] == p[rec.iz[r], rec.ix[r]]


def test_thread_count_does_not_change_results():
    a = _small_run().seismograms()
    rec = ReceiverSpec(ix=np.array([30, 50, 70, 50]), iz=np.array([60, 70, 60, 40]),
                       comp_mask=_abi.VX | _abi.VY | _abi.P, nsamp=150)
    prob = small_problem(100, 100, DX, DT, nt=150, npml=10,
                         src=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=50, iz=50, amp=1e10), rec=rec)
    s = OracleSolver(prob, precision=F32, threads=1)
    s.run(150)
    assert np.array_equal(a, s.seismograms())


def test_float_oracle_tracks_double_oracle():
    """The FP32 restatement (reference arithmetic) stays within 1e-4 per-trace relative L2 of the
    FP64 restatement over 800 steps of a configuration-0-like run."""
    from poe2d_amd.configs import config0
    from oracle.pyoracle import material_biot
    prob = config0(nt=800, material_fn=material_biot, n=200)
    r = {}
    for prec in (F32, F64):
        s = OracleSolver(prob, precision=prec)
        s.run(800)
        r[prec] = s.seismograms()
    err = trace_rel_l2(r[F32], r[F64])
    assert err.max() < 1e-4, err.max()


def test_reference_scans_record_what_the_gather_records():
    """The reference-structured scans (three full-grid passes per step, CPP:657-739; the CPU
    baseline runs them for their cost) record the same samples as the receiver gather at the
    reference's receiver positions: the VSP line at column is, the horizontal line at row 10 and
    the point traces at column (NX+1)/2 - 1."""
    from oracle.pyoracle import OracleSolver, material_biot
    from poe2d_amd.configs import biot_fields
    from poe2d_amd.solver import Problem, ReceiverSpec, SourceSpec
    nx, nz, nt, panj, panj2, a0, a1 = 130, 119, 60, 10, 13, 70, 90
    src = SourceSpec(kind=_abi.SRC_PRESSURE_DGAUSS, ix=60, iz=40, amp=1e10)
    len2, length = (nz + 1) // panj, (nx + 1) // panj2
    xc = (nx + 1) // 2 - 1
    vsp = [(src.ix, (m + 1) * len2 - 1) for m in range(panj)]
    hor = [((m + 1) * length - 1, 10) for m in range(panj2)]
    pts = [(xc, 50), (xc, a0), (xc, a1), (xc, src.iz)]
    pos = vsp + hor + pts
    rec = ReceiverSpec(ix=np.array([p[0] for p in pos], np.int32), iz=np.array([p[1] for p in pos], np.int32),
                       comp_mask=_abi.VX | _abi.VY)
    coeffs = material_biot(**biot_fields(nx, nz))
    prob = Problem(nx=nx, nz=nz, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs, source=src, receivers=rec,
                   npml=10, f0=200.0)
    o = OracleSolver(prob, _abi.F32, threads=2)
    o.set_reference_scans(panj, panj2, a0, a1)
    o.run(nt)
    g = o.seismograms()  # [vx, vy][receiver][slot]
    assert np.abs(g[1]).max() > 0
    for m in range(panj):
        assert np.array_equal(o.reference_scan(0, m, nt), g[1, m].astype(np.float64)), m
        assert np.array_equal(o.reference_scan(1, m, nt), g[0, m].astype(np.float64)), m
    for m in range(panj2):
        assert np.array_equal(o.reference_scan(2, m, nt), g[1, panj + m].astype(np.float64)), m
        assert np.array_equal(o.reference_scan(3, m, nt), g[0, panj + m].astype(np.float64)), m
    for q in range(4):
        assert np.array_equal(o.reference_scan(4 + q, 0, nt), g[1, panj + panj2 + q].astype(np.float64)), q
    # the compared samples carry signal on every line
    assert np.abs(o.reference_scan(0, 3, nt)).max() > 0 and np.abs(o.reference_scan(3, 4, nt)).max() > 0
    assert np.abs(o.reference_scan(7, 0, nt)).max() > 0
    o.close()
