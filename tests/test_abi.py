"""CPU checks of the product library boundary: it loads, exports what include/poe2d.h declares,
and its host-side set-up routines agree with the oracle's independent restatements."""
import os
import re

import numpy as np
import pytest

from oracle import pyoracle
from poe2d_amd import _abi
from poe2d_amd import material as pm
from poe2d_amd.configs import biot_fields
from poe2d_amd.patchy import patchy_lattice, patchy_mask, patchymodel_medium

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(poe2d_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    lib = _abi.load_library()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.poe2d_abi_version() == 1
    assert len(declared_functions()) >= 20


def test_cpml_profile_matches_oracle():
    for prec in (_abi.F32, _abi.F64):
        for axis in (0, 1):
            for n, npml in ((401, 20), (4001, 20), (97, 10)):
                args = (axis, n, 0.5, 5e-5, npml, 2.0, 1e-3, 1.0, np.pi * 80, 3605.4, prec)
                a = pm.cpml_profile(*args)
                b = pyoracle.cpml_profile(*args)
                for k in a:
                    if prec == _abi.F32:
                        assert np.array_equal(a[k], b[k]), (prec, axis, n, k)
                    else:
                        np.testing.assert_allclose(a[k], b[k], rtol=1e-14, atol=1e-300)


def test_material_biot_matches_oracle():
    rng = np.random.default_rng(5)
    mask = rng.random((30, 40)) < 0.3
    p = biot_fields(39, 29, mask)
    p["phi"] = p["phi"] * rng.uniform(0.6, 1.4, p["phi"].shape)
    a = pm.material_biot(**p)
    b = pyoracle.material_biot(**p)
    for k in a:
        np.testing.assert_allclose(a[k], b[k], rtol=1e-15)


def test_material_reference_matches_oracle_bitwise():
    rng = np.random.default_rng(7)
    n = 500
    cp = rng.uniform(1500, 4500, n).astype(np.float32)
    cs = (cp / rng.uniform(1.5, 2.2, n)).astype(np.float32)
    rho = rng.uniform(1000, 2700, n).astype(np.float32)
    med = rng.integers(0, 4, n).astype(np.float32)
    a = pm.material_reference(cp, cs, rho, med)
    b = pyoracle.material_reference(cp, cs, rho, med, pm.REFERENCE_TABLES)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    with pytest.raises(_abi.Poe2dError):
        pm.material_reference(cp, cs, rho, med + 9)


def test_wavelet_matches_oracle():
    for kind in (_abi.SRC_PRESSURE_DGAUSS, _abi.SRC_FORCE_Y_RICKER):
        a = pm.wavelet(kind, 80.0, 0.0, 1e10, 0.1, 5e-5, 3000)
        b = pyoracle.wavelet(kind, 80.0, 0.0, 1e10, 0.1, 5e-5, 3000)
        np.testing.assert_allclose(a, b, rtol=1e-15, atol=0)


def test_patchy_mask_fraction_and_seed():
    m1 = patchy_mask(400, 300, 0.5, 0.5, 0.15, 1, 2.0, 20.0)
    m2 = patchy_mask(400, 300, 0.5, 0.5, 0.15, 1, 2.0, 20.0)
    m3 = patchy_mask(400, 300, 0.5, 0.5, 0.15, 2, 2.0, 20.0)
    assert m1.shape == (301, 401)
    assert np.array_equal(m1, m2) and not np.array_equal(m1, m3)
    # painting stops at the first disk that reaches the target: overshoot below one disk area
    assert 0.15 <= m1.mean() < 0.15 + np.pi * 40 ** 2 / m1.size
    assert not patchy_mask(100, 100, 0.5, 0.5, 0.0, 1, 2.0, 20.0).any()
    # a tiny target paints a single disk: its point count is pi r^2 within one perimeter
    m = patchy_mask(200, 200, 1.0, 1.0, 1e-9, 3, 8.0, 8.0)
    assert abs(m.sum() - np.pi * 64) < 2 * np.pi * 8


def _lattice_restated(kind, rows=240, cols=5000, r=8):
    """patchymodel.m's layer of type `kind`, restated from its loops: disk centres i = start:step:cols
    and j = start:step:rows (1-based), every point within distance r of a centre set to 1 (the
    MATLAB sums the disks and clamps values above 1)."""
    dia = 2 * r
    lattices = {1: [(r + 1, dia, r + 1, dia)],
                2: [(r + 1, 2 * dia, r + 1, 2 * dia)],
                3: [(r + 1, 4 * dia, r + 1, 2 * dia)],
                4: [(r + 1, 2 * dia, r + 1, 2 * dia), (3 * r + 1, 2 * dia, 3 * r + 1, 2 * dia)],
                5: [(r + 1, 2 * dia, 3 * r + 1, 2 * dia), (r + 1, dia, r + 1, 2 * dia)],
                6: [(r + 1, 4 * dia, r + 1, 2 * dia), (5 * r + 1, 4 * dia, 3 * r + 1, 2 * dia)]}[kind]
    out = np.zeros((rows, cols), np.int32)
    dj, di = np.mgrid[-r:r + 1, -r:r + 1]
    disk = (np.sqrt(di.astype(float) ** 2 + dj ** 2) <= r).astype(np.int32)
    pad = np.zeros((rows + 2 * r, cols + 2 * r), np.int32)
    for x0, sx, z0, sz in lattices:
        for i in range(x0, cols + 1, sx):
            for j in range(z0, rows + 1, sz):
                # 1-based centre (i, j) -> padded 0-based window rows j-1 .. j-1+2r
                pad[j - 1:j + 2 * r, i - 1:i + 2 * r] += disk
    out = pad[r:r + rows, r:r + cols]
    return (out > 0).astype(np.uint8)


def test_patchymodel_lattice_counts():
    # type 1: radius-8 disks centred every 16 points; a lone disk has 197 points
    one = patchy_lattice(1, rows=17, cols=17, r=8)
    assert one.sum() == 197
    t1 = patchy_lattice(1)
    assert t1.shape == (240, 5000)
    # centres: (5000-9)//16+1 = 312 along x, (240-9)//16+1 = 15 along z
    assert t1[8, 8] == 1 and t1[8, 8 + 16 * 311] == 1 and t1[8 + 16 * 14, 8] == 1
    for k in range(2, 7):
        lay = patchy_lattice(k)
        assert 0 < lay.sum() < t1.sum()
        assert np.array_equal(lay, _lattice_restated(k)), k
    assert np.array_equal(t1, _lattice_restated(1))
    med = patchymodel_medium(1)
    assert med.shape == (2800, 5000)
    assert set(np.unique(med)) == {1.0, 2.0, 3.0}
    assert np.all(med[:2000] == 1) and np.all(med[2240:] == 1)


def _patchy_restated(nx, nz, dx, dz, gas, seed, rmin, rmax):
    """The random patchy mask restated from its definition: SplitMix64 (Steele, Lea and Flood 2014)
    draws each disk's centre uniform over [0, nx dx] x [0, nz dz] and its radius log-uniform in
    [rmin, rmax]; every grid point within the radius of the centre is painted; disks are drawn
    until ceil(gas * points) points are painted."""
    import math
    mask64 = (1 << 64) - 1
    state = [seed & mask64]

    def uniform():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & mask64
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask64
        return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0)

    mask = np.zeros((nz + 1, nx + 1), np.uint8)
    target = math.ceil(gas * (nx + 1) * (nz + 1))
    lr0, lr1 = math.log(rmin), math.log(rmax)
    xs = np.arange(nx + 1) * dx
    zs = np.arange(nz + 1) * dz
    while mask.sum() < target:
        cx, cz = uniform() * nx * dx, uniform() * nz * dz
        r = math.exp(lr0 + (lr1 - lr0) * uniform())
        inside = (xs[None, :] - cx) ** 2 + (zs[:, None] - cz) ** 2 <= r * r
        mask[inside] = 1
    return mask


@pytest.mark.parametrize("nx,nz,dx,dz,gas,seed,rmin,rmax", [
    (120, 90, 0.5, 0.5, 0.2, 1, 2.0, 20.0),
    (97, 131, 1.25, 0.75, 0.12, 7, 1.0, 6.0),
    (80, 80, 0.5, 0.5, 0.3, 11, 3.3, 3.3),
    (150, 40, 2.0, 1.0, 0.05, 5, 0.6, 30.0),
])
def test_patchy_mask_equals_disk_restatement(nx, nz, dx, dz, gas, seed, rmin, rmax):
    """The library's painter against the restatement point for point: every disk's points, its
    radius (exp of a log-uniform draw, also when rmin = rmax) and the stopping count."""
    got = patchy_mask(nx, nz, dx, dz, gas, seed, rmin, rmax)
    want = _patchy_restated(nx, nz, dx, dz, gas, seed, rmin, rmax)
    assert got.shape == want.shape and np.array_equal(got, want)
