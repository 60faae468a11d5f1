"""The reference-format program end to end on the GPU: poe2d-run reads input1.txt, input2.txt and
four float32 model grids (z-fastest files, as patchymodel.m writes them), and its seismogram and
snapshot files must match the oracle run of the same problem (material from the oracle's own
restatement of the reference front end, receivers from a brute-force restatement of the
reference's gather conditions).  Bound: per-trace relative L2 <= 1e-4 (the ASCII files carry 6
significant digits, ~1e-6 relative)."""
import subprocess

import numpy as np
import pytest

from oracle import pyoracle
from oracle.pyoracle import OracleSolver
from poe2d_amd import _abi, refio
from poe2d_amd.solver import Problem, ReceiverSpec, SourceSpec

from .helpers import random_patchy_mask, trace_rel_l2
from .test_refio import brute_force_layout

pytestmark = pytest.mark.gpu

MEDIA = dict(kf=[2.4e9, 2.25e9, 0.05e9, 0.02e9], eta=[0.003, 0.001, 1.5e-5, 1.0e-5],
             kappa=[1.5e-13, 1.0e-12, 1.0e-12, 5.0e-13], phi=[0.2, 0.25, 0.25, 0.01],
             a=[2.0, 2.0, 2.0, 2.0], rhof=[1000.0, 1000.0, 100.0, 160.0])


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
