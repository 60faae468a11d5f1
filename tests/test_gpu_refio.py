"""The reference-format program end to end on the GPU: poe2d-run reads input1.txt, input2.txt and
four float32 model grids (z-fastest files, as patchymodel.m writes them), and its seismogram and
snapshot files must match the oracle run of the same problem (material from the oracle's own
restatement of the reference front end, receivers from a brute-force restatement of the
reference's gather conditions).  Bound: per-trace relative L2 <= 1e-4 (the ASCII files carry 6
significant digits, ~1e-6 relative)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle
from oracle.pyoracle import OracleSolver
from poe2d_amd import _abi, refio
from poe2d_amd.configs import SYNTHETIC_MEDIA, marmousi_like
from poe2d_amd.solver import Problem, ReceiverSpec, SourceSpec

from .helpers import random_patchy_mask, trace_rel_l2
from .test_refio import brute_force_layout

pytestmark = pytest.mark.gpu

MEDIA = SYNTHETIC_MEDIA


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)


def make_case(d, n=199, nstep=400, src=(100, 90), a0=130, a1=150, f0=80):
    mask = random_patchy_mask(n, n, 0.2, seed=31)
    medium = np.where(mask, 2.0, 1.0).astype(np.float32)
    cp = np.where(mask, 3100.0, 3300.0).astype(np.float32)
    cs = np.where(mask, 1900.0, 1850.0).astype(np.float32)
    rho = np.where(mask, 2200.0, 2350.0).astype(np.float32)
    for name, a in (("medium.bin", medium), ("cp.bin", cp), ("cs.bin", cs), ("rho.bin", rho)):
        a.T.astype("<f4").tofile(d / name)  # z fastest
    with open(d / "media.txt", "w") as f:
        for k in range(4):
            f.write(" ".join(repr(float(MEDIA[c][k])) for c in ("kf", "eta", "kappa", "phi", "a", "rhof")) + "\n")
    (d / "input1.txt").write_text(f"{nstep}\n{src[0]} {src[1]}\n{a0} {a1}\n")
    (d / "input2.txt").write_text(f"{f0}\n")
    return dict(n=n, nstep=nstep, src=src, a0=a0, a1=a1, f0=f0, medium=medium, cp=cp, cs=cs, rho=rho)


@pytest.mark.parametrize("law,material", [("written", "written"), ("drained", "averaged")])
def test_reference_program_matches_oracle(tmp_path, law, material):
    c = make_case(tmp_path)
    n, nstep, (is_, js) = c["n"], c["nstep"], c["src"]
    panj, panj2, samp = 50, 40, 150
    out = tmp_path / "out"
    out.mkdir()
    say(f"poe2d-run {n + 1}^2 NSTEP={nstep} stress law {law}, material {material}")
    r = subprocess.run([refio.RUNNER, "--nx", str(n), "--nz", str(n), "--medium", "medium.bin", "--cp", "cp.bin",
                        "--cs", "cs.bin", "--rho", "rho.bin", "--layout", "z", "--media", "media.txt",
                        "--panj", str(panj), "--panj2", str(panj2), "--samp", str(samp), "--outdir", "out",
                        "--tile", "256x14", "--stress-law", law, "--material", material],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    say(r.stdout.strip().replace("\n", " | "))
    assert r.returncode == 0, r.stderr

    # the same problem through the oracle
    pos, nv, nh = brute_force_layout(n, n, is_, js, c["a0"], c["a1"], panj, panj2)
    ix = np.array([p[0] for p in pos], dtype=np.int32)
    iz = np.array([p[1] for p in pos], dtype=np.int32)
    coeffs = pyoracle.material_reference(c["cp"], c["cs"], c["rho"], c["medium"], MEDIA)
    prob = Problem(nx=n, nz=n, dx=1.25, dz=1.25, dt=95e-6, nt=nstep, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_PRESSURE_DGAUSS, ix=is_, iz=js, amp=1e10, t_cut=0.4),
                   receivers=ReceiverSpec(ix=ix, iz=iz, comp_mask=_abi.VX | _abi.VY), npml=10, f0=80.0,
                   cp_max=float(c["cp"].max()), check_cfl=0,
                   stress_law=_abi.STRESS_DRAINED if law == "drained" else _abi.STRESS_AS_WRITTEN,
                   staggered_material=_abi.MATERIAL_AVERAGED if material == "averaged" else _abi.MATERIAL_AS_WRITTEN)
    o = OracleSolver(prob, precision=_abi.F32)
    o.run()
    seis = o.seismograms()  # [vx, vy][receiver][sample]
    vz_last = o.snapshot("vy")

    names = refio.output_names(c["f0"])
    files = {k: refio.read_traces(out / v) for k, v in names.items()}
    assert files["vsp_vz"].shape == (nstep, panj) and files["hor_vx"].shape == (nstep, panj2)
    assert files["ats"].shape == (nstep,)
    got = {
        "vsp_vz": files["vsp_vz"][:, :nv].T, "vsp_vx": files["vsp_vx"][:, :nv].T,
        "hor_vz": files["hor_vz"][:, :nh].T, "hor_vx": files["hor_vx"][:, :nh].T,
        "points": np.stack([files["ats"], files["bwh"], files["src"]]),
    }
    want = {
        "vsp_vz": seis[1, :nv], "vsp_vx": seis[0, :nv], "hor_vz": seis[1, nv:nv + nh],
        "hor_vx": seis[0, nv:nv + nh], "points": seis[1, nv + nh:],
    }
    assert not files["vsp_vz"][:, nv:].any() and not files["hor_vz"][:, nh:].any()
    for k in got:
        err = trace_rel_l2(got[k][None], want[k][None])
        say(f"{k}: {got[k].shape[0]} traces, max rel L2 vs oracle {err.max():.2e}")
        assert np.abs(want[k]).max() > 0, k
        assert err.max() <= 1e-4, k
    # snapshots every samp steps and after the last
    for step in (150, 300, 400):
        assert (out / f"vz{step:05d}.vzo").stat().st_size == 4 * (n + 1) ** 2
        assert (out / f"vx{step:05d}.vxo").stat().st_size == 4 * (n + 1) ** 2
    snap = refio.read_grid(out / "vz00400.vzo", n + 1, n + 1)
    rel = np.linalg.norm(snap - vz_last) / np.linalg.norm(vz_last)
    say(f"vz00400.vzo vs oracle: rel L2 {rel:.2e}")
    assert rel <= 1e-4


def test_reference_grid_shape(tmp_path):
    """poe2d-run on the reference program's own grid (its defaults: 13600 x 2440 points, 1.25 m,
    95 us, 10 C-PML points, 122 VSP and 680 horizontal receivers) over synthetic float32 model
    files in its format (x fastest, mar_*_2440x13600.bin).  The model varies at every point, so the
    run takes the two-step kernel on the coefficient groups.  The media come from a table file:
    the reference's default tables (porosity 1e-6) give media 0 and 1 a fluid-wave speed of about
    11 km/s, above this grid's Courant limit of 7.98 km/s (test_refio pins this), and the oracle
    diverges on them as well.  300 steps: the default schedule (two steps per launch, autotuned tile) equals
    the single-step kernel bitwise in every output file, and matches the FP32 oracle within 1e-4
    per trace and in the final snapshots."""
    nx, nz, nstep, is_, js, a0, a1, f0 = 13599, 2439, 300, 6800, 40, 50, 70, 30
    medium, cp, cs, rho = marmousi_like(nx, nz)
    for name, a in (("poro", medium), ("cp", cp), ("cs", cs), ("rho", rho)):
        a.astype("<f4").tofile(tmp_path / f"mar_{name}_2440x13600.bin")  # x fastest
    with open(tmp_path / "media.txt", "w") as f:
        for k in range(4):
            f.write(" ".join(repr(float(MEDIA[c][k])) for c in ("kf", "eta", "kappa", "phi", "a", "rhof")) + "\n")
    (tmp_path / "input1.txt").write_text(f"{nstep}\n{is_} {js}\n{a0} {a1}\n")
    (tmp_path / "input2.txt").write_text(f"{f0}\n")
    runs = {"default": [], "single-step": ["--schedule", "0", "--tile", "256x8"]}
    for tag, extra in runs.items():
        (tmp_path / tag).mkdir()
        say(f"poe2d-run 13600 x 2440, NSTEP={nstep}, {tag} schedule")
        r = subprocess.run([refio.RUNNER, "--media", "media.txt", "--outdir", tag, *extra], cwd=tmp_path,
                           capture_output=True, text=True, timeout=300)
        say(r.stdout.strip().replace("\n", " | "))
        assert r.returncode == 0, r.stderr
    names = sorted(p.name for p in (tmp_path / "default").iterdir())
    assert names == sorted(p.name for p in (tmp_path / "single-step").iterdir())
    assert f"vz{nstep:05d}.vzo" in names and f"vx{nstep:05d}.vxo" in names
    for name in names:
        assert (tmp_path / "default" / name).read_bytes() == (tmp_path / "single-step" / name).read_bytes(), name
    say(f"{len(names)} output files bitwise equal between the schedules")

    say("oracle FP32, same problem, all host threads")
    ix, iz, nv, nh, npt = refio.reference_receivers(nx, nz, is_, js, a0, a1)
    coeffs = pyoracle.material_reference(cp, cs, rho, medium, MEDIA)
    prob = Problem(nx=nx, nz=nz, dx=1.25, dz=1.25, dt=95e-6, nt=nstep, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_PRESSURE_DGAUSS, ix=is_, iz=js, amp=1e10, t_cut=0.4),
                   receivers=ReceiverSpec(ix=ix, iz=iz, comp_mask=_abi.VX | _abi.VY), npml=10, f0=float(f0),
                   cp_max=float(cp.max()), check_cfl=0)
    o = OracleSolver(prob, precision=_abi.F32, threads=min(os.cpu_count() or 1, 16))
    o.run()
    seis = o.seismograms()
    out = tmp_path / "default"
    files = {k: refio.read_traces(out / v) for k, v in refio.output_names(f0).items()}
    got = np.concatenate([files["vsp_vz"][:, :nv].T, files["hor_vz"][:, :nh].T,
                          np.stack([files["ats"], files["bwh"], files["src"]])])
    err = trace_rel_l2(got[None], seis[1][None])
    say(f"v_z traces: max rel L2 vs oracle {err.max():.2e} over {got.shape[0]} traces")
    assert err.max() <= 1e-4
    for fname, field in ((f"vz{nstep:05d}.vzo", "vy"), (f"vx{nstep:05d}.vxo", "vx")):
        snap = refio.read_grid(out / fname, nx + 1, nz + 1)
        want = o.snapshot(field)
        rel = np.linalg.norm(snap - want) / np.linalg.norm(want)
        say(f"{fname} vs oracle: rel L2 {rel:.2e}")
        assert rel <= 1e-4
