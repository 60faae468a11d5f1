"""Reference-format I/O (host only): text inputs, float32 grids in both layouts, the ASCII trace
format, the receiver layout against a brute-force restatement of the reference's gather
conditions, and the error exits of the poe2d-run program that precede any device work."""
import os
import subprocess

import numpy as np
import pytest

from poe2d_amd import _abi, refio


def test_input_files(tmp_path):
    p1 = tmp_path / "input1.txt"
    p1.write_text("3000\n200 180\n  240 260 \n")
    assert refio.read_input1(p1) == (3000, 200, 180, 240, 260)
    p2 = tmp_path / "input2.txt"
    p2.write_text("80.5\n")
    assert refio.read_input2(p2) == pytest.approx(80.5)
    p2.write_text("0.1")  # read as float, as the reference does
    assert refio.read_input2(p2) == float(np.float32(0.1))
    p1.write_text("3000 200 180 240")
    with pytest.raises(_abi.Poe2dError) as ei:
        refio.read_input1(p1)
    assert ei.value.status == _abi.POE2D_E_IO
    with pytest.raises(_abi.Poe2dError) as ei:
        refio.read_input2(tmp_path / "missing.txt")
    assert ei.value.status == _abi.POE2D_E_IO


def test_grid_layouts(tmp_path):
    rng = np.random.default_rng(3)
    g = rng.standard_normal((7, 11)).astype(np.float32)  # (nz+1, nx+1)
    f = tmp_path / "g.bin"
    refio.write_grid(f, g)
    assert f.stat().st_size == g.size * 4
    assert np.array_equal(np.fromfile(f, dtype="<f4").reshape(7, 11), g)  # x fastest, little-endian
    assert np.array_equal(refio.read_grid(f, 11, 7), g)
    # patchymodel.m writes the z-by-x matrix column-major: z fastest
    fz = tmp_path / "gz.bin"
    g.T.astype("<f4").tofile(fz)
    assert np.array_equal(refio.read_grid(fz, 11, 7, refio.Z_FASTEST), g)
    with pytest.raises(_abi.Poe2dError) as ei:
        refio.read_grid(f, 11, 8)  # too short
    assert ei.value.status == _abi.POE2D_E_IO
    with pytest.raises(_abi.Poe2dError):
        refio.read_grid(f, 10, 7)  # too long


def test_trace_text_format(tmp_path):
    vals = np.array([0.0, -0.0, 1.5, -3.14159265, 1e-12, 2.5e-5, 123456789.0, -1e20, 7.0, 0.1,
                     3.0517578e-05, 1.0e-45], dtype=np.float32)
    t = vals.reshape(3, 4)
    f = tmp_path / "a.vsp"
    refio.write_traces(f, t)
    expect = "".join("".join(f"{float(v):g} " for v in row) + "\n" for row in t)
    assert f.read_text() == expect
    one = tmp_path / "a.ats"
    refio.write_traces(one, vals, one_line=True)
    assert one.read_text() == "".join(f"{float(v):g} " for v in vals)
    back = refio.read_traces(f)
    assert back.shape == (3, 4) and np.allclose(back, t.astype(np.float64), rtol=1e-5, atol=0)
    assert refio.read_traces(one).shape == (vals.size,)


def brute_force_layout(nx, nz, is_, js, a0, a1, panj, panj2):
    """The reference's gather conditions (save1_kernel) evaluated at every grid point: column
    ISOURCE with (y+1) % ((NZ+1)/panj) == 0 feeds VSP slot y/((NZ+1)/panj); row 10 with
    (x+1) % ((NX+1)/panj2) == 0 feeds horizontal slot x/((NX+1)/panj2); column (NX+1)/2-1 at rows
    a0_pos, a1_pos, JSOURCE feeds the point traces.  Slots at or beyond panj / panj2 (which the
    reference writes into the next sample's row) are dropped."""
    len2, length = (nz + 1) // panj, (nx + 1) // panj2
    vsp, hor = {}, {}
    pts = [None, None, None]
    for y in range(nz + 1):
        for x in range(nx + 1):
            if x == is_ and (y + 1) % len2 == 0 and y // len2 < panj:
                vsp[y // len2] = (x, y)
            if y == 10 and (x + 1) % length == 0 and x // length < panj2:
                hor[x // length] = (x, y)
            if x == (nx + 1) // 2 - 1:
                for k, row in enumerate((a0, a1, js)):
                    if y == row:
                        pts[k] = (x, y)
    v = [vsp[k] for k in sorted(vsp)]
    h = [hor[k] for k in sorted(hor)]
    assert sorted(vsp) == list(range(len(v))) and sorted(hor) == list(range(len(h)))
    return v + h + pts, len(v), len(h)


@pytest.mark.parametrize("nx,nz,is_,js,a0,a1,panj,panj2", [
    (399, 399, 200, 180, 240, 260, 122, 100),  # 400/122 = 3, not a divisor: clipped at 122
    (299, 244, 150, 100, 60, 70, 61, 300),     # 245/61 = 4: 61 slots, the last at row 243
    (99, 99, 10, 50, 0, 99, 100, 100),         # spacing 1: every row / column
    (1279, 600, 640, 300, 310, 320, 122, 680),
])
def test_reference_receiver_layout(nx, nz, is_, js, a0, a1, panj, panj2):
    ix, iz, nv, nh, npt = refio.reference_receivers(nx, nz, is_, js, a0, a1, panj, panj2)
    want, wv, wh = brute_force_layout(nx, nz, is_, js, a0, a1, panj, panj2)
    assert (nv, nh, npt) == (wv, wh, 3)
    assert list(zip(ix.tolist(), iz.tolist())) == want


def test_reference_receiver_errors():
    with pytest.raises(_abi.Poe2dError):
        refio.reference_receivers(99, 99, 10, 50, 0, 99, 200, 100)  # spacing (NZ+1)/panj = 0
    with pytest.raises(_abi.Poe2dError):
        refio.reference_receivers(99, 99, 10, 50, 0, 150, 100, 100)  # a1_pos outside the grid


def test_output_names():
    n = refio.output_names(80.0)
    assert n["vsp_vz"] == "recvspvz00000080.vsp" and n["hor_vx"] == "rechorvx00000080.hor"
    assert n["src"] == "recsrc00000080.src" and n["ats"] == "recats00000080.ats" and n["bwh"] == "recbwh00000080.bwh"
    assert refio.output_names(0.5)["src"] == "recsrc0.500000.src"


def _run_program(cwd, *args):
    return subprocess.run([refio.RUNNER, *args], cwd=cwd, capture_output=True, text=True, timeout=60)


def test_program_missing_inputs(tmp_path):
    assert os.path.exists(refio.RUNNER), "poe2d-run is not built"
    r = _run_program(tmp_path)
    assert r.returncode == 1 and r.stdout.startswith("File not opened...")


def test_program_courant_exit(tmp_path):
    """A time step above the reference's Courant limit (largest input c_p) ends the run with the
    reference's message and status 0, before any device work, and writes no output."""
    n = 40
    (tmp_path / "input1.txt").write_text("10 20 20 25 30\n")
    (tmp_path / "input2.txt").write_text("80\n")
    for name, v in (("m.bin", 1.0), ("cp.bin", 3000.0), ("cs.bin", 1700.0), ("rho.bin", 2300.0)):
        np.full((n + 1, n + 1), v, dtype="<f4").tofile(tmp_path / name)
    r = _run_program(tmp_path, "--nx", str(n), "--nz", str(n), "--dx", "1.25", "--dt", "5e-4",
                     "--medium", "m.bin", "--cp", "cp.bin", "--cs", "cs.bin", "--rho", "rho.bin",
                     "--panj", "10", "--panj2", "10")
    assert r.returncode == 0, r.stderr
    assert "time step is too large, simulation will be unstable" in r.stdout
    assert not list(tmp_path.glob("*.vsp")) and not list(tmp_path.glob("*.vzo"))


def test_reference_default_media_exceed_the_reference_grid_courant_limit():
    """The reference's default fluid tables (porosity 1e-6) give media 0 and 1 a fluid modulus
    M = K_f / phi of 2.4e15 Pa against m = a rho_f / phi = 2e7 kg/m^3: a fluid-wave speed
    sqrt(M / m) of about 11 km/s, above the Courant limit 6/7/sqrt(2) dx/dt = 7.98 km/s of the
    reference grid (1.25 m, 95 us), which the reference's check against the input c_p does not
    see.  Runs of its default grid over synthetic models therefore take a media table."""
    from poe2d_amd.material import REFERENCE_TABLES, material_reference
    n = 4
    one = lambda v: np.full(n, v, dtype=np.float32)
    c = material_reference(one(3000.0), one(1600.0), one(2260.0), np.arange(n, dtype=np.float32),
                           REFERENCE_TABLES)
    speed = np.sqrt(c["M"].astype(np.float64) / c["m"])
    limit = 6.0 / 7.0 / np.sqrt(2.0) * 1.25 / 95e-6
    assert 7974.0 < limit < 7975.0
    assert (speed[:2] > 1.35 * limit).all() and (speed[2:] < limit).all()
