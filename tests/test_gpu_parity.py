"""GPU parity: libpoe2d (HIP, gfx950) against the oracle on identical inputs, through the C ABI.

Bound (FP32): per-trace relative L2 <= 1e-4, ||g - o|| / max(||o||, 1e-3 max_trace ||o||), over
every recorded sample.  FP64 kernels are held to 1e-9 against the FP64 oracle.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleSolver
from poe2d_amd import _abi
from poe2d_amd.configs import biot_fields, config0
from poe2d_amd.solver import Problem, ReceiverSpec, Solver, SourceSpec

from .helpers import random_patchy_mask, trace_rel_l2

pytestmark = pytest.mark.gpu

F32, F64 = _abi.F32, _abi.F64
TOL32 = 1e-4


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)


def assert_parity(g, prob, label, o32=None, o64=None, exempt=0):
    """Per-trace bound 1e-4 against the FP32 oracle.  A trace on which FP32 arithmetic itself is
    noisy (a near-nodal trace, e.g. v_x on the axis of a pressure source, whose value is a small
    difference of large symmetric fields) is recognised by the FP32 oracle's own distance to the
    FP64 oracle: where that exceeds 5e-5, no FP32 evaluation order can promise 1e-4 against
    another FP32 evaluation.  Such a trace may miss the 1e-4 bound only if the GPU result is no
    further from the FP64 oracle than 1.25 x the FP32 oracle is; the number of traces that use
    this exemption must equal `exempt`.  o64 may be an FP64 GPU result of the same problem (the
    FP64 kernels match the FP64 oracle to ~1e-14).  Returns (max error, exempt count)."""
    o32 = run_oracle(prob, F32) if o32 is None else o32
    o64 = run_oracle(prob, F64) if o64 is None else o64
    err = trace_rel_l2(g, o32)
    noise = trace_rel_l2(o32, o64)
    g64 = trace_rel_l2(g, o64)
    noisy = noise > 5e-5
    used = noisy & (err > TOL32)
    names = [f"{'vx vy p'.split()[c]}[{r}]: {err[c, r]:.2e} vs FP32, {g64[c, r]:.2e} vs FP64, oracle FP32 "
             f"{noise[c, r]:.2e} vs FP64 (ratio {g64[c, r] / noise[c, r]:.2f})" for c, r in np.argwhere(used)]
    say(f"{label}: max rel L2 vs FP32 oracle = {err.max():.3e}, regular traces "
        f"{err[~noisy].max() if (~noisy).any() else 0:.3e}; {int(noisy.sum())} FP32-noisy traces, "
        f"{int(used.sum())} exempt" + "".join(f"; {n}" for n in names))
    assert np.isfinite(g).all()
    assert (err[~noisy] <= TOL32).all(), (label, float(err[~noisy].max()))
    assert (g64[used] <= 1.25 * noise[used]).all(), label
    assert int(used.sum()) == exempt, (label, int(used.sum()), exempt)
    return float(err.max()), int(used.sum())


def run_gpu(prob, precision=F32, nsteps=None):
    with Solver(prob, precision=precision) as s:
        s.run(nsteps)
        return s.seismograms()


def run_oracle(prob, precision=F32, nsteps=None):
    s = OracleSolver(prob, precision=precision)
    s.run(nsteps)
    return s.seismograms()


def test_device_is_gfx950():
    import ctypes
    lib = _abi.load_library()
    n = ctypes.c_int()
    arch = ctypes.create_string_buffer(64)
    assert lib.poe2d_device_info(ctypes.byref(n), arch, 64) == 0
    assert n.value >= 1 and arch.value.decode().startswith("gfx950")


def test_parity_config0_full_fp32():
    """Configuration 0 as specified: 401 x 401 points, 3000 steps, 101 receivers x (vx, vy, p)."""
    say("config0 401x401 nt=3000 FP32 vs oracle")
    prob = config0()
    g = run_gpu(prob)
    assert np.abs(g).max() > 0
    assert_parity(g, prob, "config0")


# traces outside the 1e-4 bound whose FP32 noise floor exceeds it: v_x on the source column of
# the homogeneous pressure-source case (receiver 15, x = 120)
EXEMPT = {"force_homog": 0, "pressure_homog": 1, "force_patchy": 0, "pressure_patchy_tcut": 0}


@pytest.mark.parametrize("case", ["force_homog", "pressure_homog", "force_patchy", "pressure_patchy_tcut"])
def test_parity_variants_fp32(case):
    n, nt = 240, 1200
    mask = random_patchy_mask(n, n, 0.2, seed=4) if "patchy" in case else None
    from poe2d_amd.material import material_biot
    coeffs = material_biot(**biot_fields(n, n, mask))
    kind = _abi.SRC_PRESSURE_DGAUSS if "pressure" in case else _abi.SRC_FORCE_Y_RICKER
    rec = ReceiverSpec(ix=np.arange(30, 211, 6, dtype=np.int32), iz=np.full(31, 160, np.int32))
    prob = Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=kind, ix=120, iz=100, amp=1e10,
                                     t_cut=0.03 if "tcut" in case else 0.0),
                   receivers=rec, npml=20, f0=80.0)
    say(f"variant {case} {n + 1}^2 nt={nt}")
    g = run_gpu(prob)
    assert_parity(g, prob, f"variant {case}", exempt=EXEMPT[case])


def test_parity_config0_full_fp64():
    """Configuration 0 in FP64 as specified (401 x 401 points, 3000 steps): the FP64 kernels
    against the FP64 oracle."""
    say("config0 401x401 nt=3000 FP64 kernels vs FP64 oracle")
    prob = config0()
    g = run_gpu(prob, F64)
    o = run_oracle(prob, F64)
    err = trace_rel_l2(g, o)
    say(f"config0 FP64 max rel L2 = {err.max():.3e}")
    assert np.abs(g).max() > 0
    assert err.max() <= 1e-9


def test_fp64_vs_fp32_gpu():
    say("FP64 vs FP32 GPU 201^2 nt=1500")
    prob = config0(nt=1500, n=200)
    e = trace_rel_l2(run_gpu(prob, F32), run_gpu(prob, F64))
    say(f"FP32 vs FP64 GPU max rel L2 = {e.max():.3e}")
    assert e.max() <= TOL32


def test_snapshots_match_oracle():
    n, nt = 160, 300
    mask = random_patchy_mask(n, n, 0.25, seed=9)
    from poe2d_amd.material import material_biot
    coeffs = material_biot(**biot_fields(n, n, mask))
    prob = Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_PRESSURE_DGAUSS, ix=80, iz=70, amp=1e10),
                   npml=20, f0=80.0)
    say("snapshots of all 11 fields after 300 steps")
    gs = Solver(prob, F32)
    gs.run(nt)
    os_ = OracleSolver(prob, F32)
    os_.run(nt)
    for name in _abi.FIELD_NAMES:
        a, b = gs.snapshot(name), os_.snapshot(name)
        scale = np.abs(b).max()
        assert scale > 0, name
        assert np.abs(a - b).max() <= 2e-4 * scale, (name, np.abs(a - b).max() / scale)
    gs.close()


def _small(nt=200, amp=1e10, kind=_abi.SRC_FORCE_Y_RICKER, decimate=1):
    prob = config0(nt=nt, n=128)
    prob.source.amp = amp
    prob.source.kind = kind
    prob.receivers.decimate = decimate
    return prob


def test_checkpoint_resume_bitwise():
    say("checkpoint/resume")
    prob = _small(nt=300)
    with Solver(prob) as a:
        a.run(300)
        full = a.seismograms()
    with Solver(prob) as b:
        b.run(120)
        blob = b.checkpoint()
    with Solver(prob) as c:
        c.restore(blob)
        assert c.step_count() == 120
        c.run(180)
        assert np.array_equal(c.seismograms(), full)


def test_reset_rerun_bitwise():
    say("reset/rerun")
    prob = _small(nt=150)
    with Solver(prob) as a:
        a.run()
        first = a.seismograms()
        a.reset()
        assert a.step_count() == 0 and not a.seismograms().any()
        a.run()
        assert np.array_equal(a.seismograms(), first)


def test_linearity_gpu():
    say("linearity x2")
    a = run_gpu(_small(amp=1e10))
    b = run_gpu(_small(amp=2e10))
    assert np.abs(a).max() > 0
    # Scaling by 2 is exact in binary floating point only while no result rounds in the
    # subnormal range.  Ahead of the wavefront the precursor starts at ~1e-45 and grows by ~35
    # orders of magnitude as the wave arrives; subnormals sit on a fixed absolute grid, so the
    # two runs round differently there and the difference rides up with the growth to the last
    # bit of normal samples.  Linearity therefore holds to FP32 rounding, per trace.
    assert trace_rel_l2(b, 2 * a).max() <= 1e-5
    assert not run_gpu(_small(amp=0.0)).any()


def test_decimation_and_gather_equals_snapshot():
    say("decimation / gather")
    full = run_gpu(_small(nt=120))
    dec = run_gpu(_small(nt=120, decimate=6))
    assert np.array_equal(dec, full[..., ::6])
    prob = _small(nt=120)
    with Solver(prob) as s:
        s.run(37)
        vx, vy = s.snapshot("vx"), s.snapshot("vy")
    rec = prob.receivers
    assert np.array_equal(full[0, :, 37], vx[rec.iz, rec.ix])
    assert np.array_equal(full[1, :, 37], vy[rec.iz, rec.ix])


def test_instability_is_reported():
    """Courant 1.44 with the check switched off: the field grows without bound and run() reports
    POE2D_E_UNSTABLE instead of returning garbage."""
    say("instability detection")
    prob = config0(nt=1500, n=128)
    prob.dt = 2e-4
    prob.check_cfl = 0
    with Solver(prob) as s:
        with pytest.raises(_abi.Poe2dError) as ei:
            s.run(1500)
        assert ei.value.status == _abi.POE2D_E_UNSTABLE


def test_errors_without_compute():
    prob = config0(nt=10, n=128)
    prob.dt = 2e-4
    with pytest.raises(_abi.Poe2dError) as ei:
        Solver(prob)
    assert ei.value.status == _abi.POE2D_E_CFL
    prob = config0(nt=10, n=128)
    s = Solver(prob)
    for bad in (SourceSpec(ix=0, iz=50), SourceSpec(ix=50, iz=128),
                SourceSpec(kind=_abi.SRC_PRESSURE_DGAUSS, ix=50, iz=1)):
        with pytest.raises(_abi.Poe2dError) as ei:
            s.set_source(bad)
        assert ei.value.status == _abi.POE2D_E_INVALID
    with pytest.raises(_abi.Poe2dError):
        s.set_receivers(ReceiverSpec(ix=np.array([200]), iz=np.array([5])), 10)
    s.close()


def _run_sched(prob, schedule, block_w=256, tile_rows=64, precision=F32):
    with Solver(prob, precision=precision) as s:
        s.set_schedule(schedule, block_w, tile_rows)
        s.run()
        return s.seismograms(), s.snapshot("vx"), s.snapshot("sxy")


def test_schedules_agree_and_tiles_are_exact():
    """The single-pass kernel and the two-pass kernels agree to rounding; within the single-pass
    kernel, the tile height does not change a single bit (each point's arithmetic is the same
    wherever the tile boundaries fall), and other block widths agree to rounding."""
    from poe2d_amd.material import material_biot
    n, nt = 300, 500
    mask = random_patchy_mask(n, n, 0.25, seed=21)
    coeffs = material_biot(**biot_fields(n, n, mask))
    rec = ReceiverSpec(ix=np.arange(25, 276, 10, dtype=np.int32), iz=np.full(26, 200, np.int32))
    prob = Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=150, iz=140, amp=1e10),
                   receivers=rec, npml=20, f0=80.0)
    say("schedules: two-pass reference run")
    two = _run_sched(prob, 1)
    base = _run_sched(prob, 0, 256, 64)
    for a, b in zip(two, base):
        scale = np.abs(a).max()
        assert scale > 0 and np.abs(a - b).max() <= 1e-5 * scale
    for th in (7, 16, 33, 301):
        say(f"schedules: fused bw=256 th={th}")
        other = _run_sched(prob, 0, 256, th)
        for a, b in zip(base, other):
            assert np.array_equal(a, b), th
    for bw in (128, 512):
        say(f"schedules: fused bw={bw}")
        other = _run_sched(prob, 0, bw, 48)
        for a, b in zip(base, other):
            scale = np.abs(a).max()
            assert np.abs(a - b).max() <= 1e-5 * scale, bw


def test_palette_and_groups_agree():
    """A two-material patchy grid is read through the one-byte material palette by default;
    forcing the per-point coefficient arrays (schedule 2) gives the same seismograms and fields."""
    from poe2d_amd.material import material_biot
    n, nt = 300, 500
    coeffs = material_biot(**biot_fields(n, n, random_patchy_mask(n, n, 0.25, seed=22)))
    rec = ReceiverSpec(ix=np.arange(25, 276, 10, dtype=np.int32), iz=np.full(26, 200, np.int32))
    prob = Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_PRESSURE_DGAUSS, ix=150, iz=140, amp=1e10),
                   receivers=rec, npml=20, f0=80.0)
    with Solver(prob) as s:
        assert s.material_palette() == 2
    for precision in (F32, F64):
        say(f"palette vs coefficient groups, precision {precision}")
        pal = _run_sched(prob, 0, 256, 64, precision)
        grp = _run_sched(prob, 2, 256, 64, precision)
        for a, b in zip(pal, grp):
            scale = np.abs(a).max()
            say(f"  bitwise equal: {np.array_equal(a, b)}, max diff / scale {np.abs(a - b).max() / scale:.2e}")
            assert scale > 0 and np.abs(a - b).max() <= 1e-6 * scale


def test_parity_graded_medium_without_palette():
    """A medium whose porosity grades with depth and offset has more than 256 distinct
    coefficient tuples: the fused step reads the per-point coefficient arrays, and matches the
    oracle."""
    from poe2d_amd.material import material_biot
    n, nt = 240, 1000
    fields = biot_fields(n, n)
    ramp = np.linspace(-1, 1, n + 1)
    fields["phi"] = fields["phi"] * (1.0 + 0.15 * ramp[:, None]) * (1.0 + 0.05 * ramp[None, :])
    coeffs = material_biot(**fields)
    rec = ReceiverSpec(ix=np.arange(30, 211, 6, dtype=np.int32), iz=np.full(31, 160, np.int32))
    prob = Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=120, iz=100, amp=1e10),
                   receivers=rec, npml=20, f0=80.0)
    with Solver(prob) as s:
        assert s.material_palette() == 0
    say(f"graded medium {n + 1}^2 nt={nt}, no palette")
    assert_parity(run_gpu(prob), prob, "graded medium")


def test_autotune_picks_a_shape_and_resets():
    """autotune leaves the handle at step 0 with a valid shape, whose results equal a run with the
    same shape set by hand; on a handle past step 0 it refuses."""
    from poe2d_amd.material import material_biot
    n, nt = 200, 300
    coeffs = material_biot(**biot_fields(n, n, random_patchy_mask(n, n, 0.2, seed=9)))
    rec = ReceiverSpec(ix=np.arange(30, 171, 7, dtype=np.int32), iz=np.full(21, 140, np.int32))
    prob = Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=100, iz=90, amp=1e10),
                   receivers=rec, npml=20, f0=80.0)
    with Solver(prob) as s:
        s.set_schedule(0, 256, 14)  # the single-step kernel's shapes (FP32 handles start on 3)
        bw, th = s.autotune(3)
        say(f"autotune {n + 1}^2: block {bw}, tile rows {th}")
        assert bw in (128, 256, 512) and th in (8, 14, 20, 26, 32, 44, 62, 98)
        assert s.step_count() == 0
        s.run()
        tuned = s.seismograms()
        with pytest.raises(_abi.Poe2dError) as ei:
            s.autotune(3)
        assert ei.value.status == _abi.POE2D_E_STATE
    manual = _run_sched(prob, 0, bw, th)[0]
    assert np.array_equal(tuned, manual)
    # the shape is cached for this grid class: a second handle gets it without timing, and a
    # negative step count tunes again
    import time
    with Solver(prob) as s2:
        s2.set_schedule(0, 256, 14)
        t0 = time.perf_counter()
        assert s2.autotune(3) == (bw, th)
        cached = time.perf_counter() - t0
        t0 = time.perf_counter()
        bw2, th2 = s2.autotune(-3)
        timed = time.perf_counter() - t0
        say(f"autotune cached {cached * 1e3:.1f} ms, re-timed {timed * 1e3:.1f} ms -> {bw2}x{th2}")
        assert cached < timed and s2.step_count() == 0


@pytest.mark.parametrize("case", ["pressure_patchy", "force_patchy"])
def test_two_pass_schedule_parity(case):
    from poe2d_amd.material import material_biot
    n, nt = 200, 800
    coeffs = material_biot(**biot_fields(n, n, random_patchy_mask(n, n, 0.2, seed=5)))
    kind = _abi.SRC_PRESSURE_DGAUSS if "pressure" in case else _abi.SRC_FORCE_Y_RICKER
    rec = ReceiverSpec(ix=np.arange(30, 171, 5, dtype=np.int32), iz=np.full(29, 140, np.int32))
    prob = Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=kind, ix=97, iz=90, amp=1e10),
                   receivers=rec, npml=20, f0=80.0)
    say(f"two-pass schedule {case}")
    g = _run_sched(prob, 1)[0]
    assert_parity(g, prob, f"two-pass {case}")
