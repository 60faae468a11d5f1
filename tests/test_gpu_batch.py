"""Batched launches, graph replays and clones on the GPU.

- Graph replays equal direct launches bitwise, run after run, with snapshots (eager work) and
  synchronisations between the replays.
- A batch of ragged models and clones advanced in lock-step equals each member run alone with the
  same kernels, bitwise (seismograms and wavefields).
- A clone shares the model and equals its parent for the same source.
"""
import numpy as np
import pytest

from poe2d_amd import _abi
from poe2d_amd.configs import biot_fields
from poe2d_amd.material import material_biot
from poe2d_amd.solver import Batch, Problem, ReceiverSpec, Solver, SourceSpec

from .helpers import random_patchy_mask

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)


def make_problem(nx, nz, seed, src_x=None, kind=_abi.SRC_FORCE_Y_RICKER, nt=250, dec=1):
    mask = random_patchy_mask(nx, nz, 0.2, seed=seed)
    coeffs = material_biot(**biot_fields(nx, nz, mask))
    nrec = 20
    rec = ReceiverSpec(ix=np.linspace(25, nx - 25, nrec).astype(np.int32), iz=np.full(nrec, nz // 2 + 30, np.int32),
                       decimate=dec)
    sx = nx // 2 if src_x is None else src_x
    return Problem(nx=nx, nz=nz, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=kind, ix=sx, iz=nz // 2, amp=1e10),
                   receivers=rec, npml=20, f0=80.0)


def state(s):
    return s.seismograms(), s.snapshot("vy"), s.snapshot("sxy"), s.snapshot("p")


def test_graph_replays_equal_direct_launches():
    prob = make_problem(200, 180, seed=41, kind=_abi.SRC_PRESSURE_DGAUSS, nt=320, dec=3)
    with Solver(prob) as g, Solver(prob) as d:
        d.set_graphs(False)
        for k in (40, 7, 70, 100, 64, 39):
            g.run(k)
            d.run(k)
            a, b = state(g), state(d)
            say(f"after {g.step_count()} steps: bitwise {all(np.array_equal(x, y) for x, y in zip(a, b))}")
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        assert np.abs(a[0]).max() > 0


def test_batch_equals_single_runs():
    probs = [make_problem(160, 200, seed=1), make_problem(230, 170, seed=2, kind=_abi.SRC_PRESSURE_DGAUSS),
             make_problem(199, 199, seed=3, dec=2)]
    base = [Solver(p) for p in probs]
    # two shots over the first model, sharing its coefficients on the device
    shots = [SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=x, iz=100, amp=1e10) for x in (50, 110)]
    clones = [base[0].clone(source=src, receivers=probs[0].receivers) for src in shots]
    members = base + clones
    with Batch(members) as b:
        b.set_schedule(256, 14)
        b.run(100)
        b.run(150)
        ms, n = b.last_run_ms()
        say(f"batch of {len(members)}: {n} steps in {ms:.2f} ms")
    singles = probs + [Problem(**{**probs[0].__dict__, "source": src}) for src in shots]
    for i, (m, p) in enumerate(zip(members, singles)):
        assert m.step_count() == 250
        with Solver(p) as s:
            s.set_schedule(0, 256, 26)
            s.run(250)
            a, b_ = state(m), state(s)
            same = all(np.array_equal(x, y) for x, y in zip(a, b_))
            say(f"member {i} ({p.nx + 1}x{p.nz + 1}): equal to the single run bitwise: {same}")
            assert np.abs(a[0]).max() > 0
            for x, y in zip(a, b_):
                assert np.array_equal(x, y)
    for m in members:
        m.close()


def test_clone_shares_the_model():
    prob = make_problem(200, 200, seed=5, nt=200)
    with Solver(prob) as s:
        c = s.clone(source=prob.source, receivers=prob.receivers)
        assert c.material_palette() == s.material_palette() == 2
        s.run()
        c.run()
        for x, y in zip(state(s), state(c)):
            assert np.array_equal(x, y)
        c.close()


def test_batch_errors_and_autotune():
    p1, p2 = make_problem(150, 150, seed=7, nt=60), make_problem(170, 140, seed=8, nt=60)
    a, b = Solver(p1), Solver(p2)
    with pytest.raises(_abi.Poe2dError):
        Batch([a, a])
    b.set_schedule(1)
    with pytest.raises(_abi.Poe2dError):
        Batch([a, b])
    b.set_schedule(0)
    with Batch([a, b]) as bt:
        bw, th = bt.autotune(2)
        say(f"batch autotune: {bw}x{th}")
        assert a.step_count() == 0 and b.step_count() == 0
        a.run(3)
        with pytest.raises(_abi.Poe2dError) as ei:
            bt.run(5)
        assert ei.value.status == _abi.POE2D_E_STATE
        b.run(3)
        bt.run(10)
        assert a.step_count() == b.step_count() == 13
    a.close()
    b.close()


def _parity_vs_oracle(prob, g, label):
    """The member's traces against the FP32 oracle, with an FP64 GPU run of the same problem as the
    FP64 reference for FP32-noisy traces."""
    from .test_gpu_parity import assert_parity
    with Solver(prob, precision=_abi.F64) as s:
        s.run()
        g64 = s.seismograms()
    assert_parity(g, prob, label, o64=g64)


def test_config2a_batch_at_shape():
    """Configuration 2, batch A at its shape: 64 Ricker shots over one 1201 x 1201 patchy model
    (clones sharing the model), advanced in lock-step on the default schedule for 301 steps
    (150 step pairs and one single step).  Shots 0 and 63 equal the same shots run alone bitwise;
    shot 0 matches the oracle."""
    import dataclasses
    from poe2d_amd.configs import config2a
    nt = 301
    base, shots, rec = config2a(nt=nt)
    first = Solver(base)
    members = [first] + [first.clone(source=sh, receivers=rec, nt=nt) for sh in shots[1:]]
    assert len(members) == 64 and base.nx == base.nz == 1200
    with Batch(members) as b:
        say("config2a: 64 shots x 1201^2, 301 steps in one batch")
        b.run(nt)
        ms, n = b.last_run_ms()
        say(f"config2a: {n} steps in {ms:.1f} ms = {64 * base.npoints * n / ms / 1e6:.1f} Gcell-updates/s")
    results = {i: state(members[i]) for i in (0, 63)}
    for m in members:
        m.close()
    for i in (0, 63):
        p = dataclasses.replace(base, source=shots[i], receivers=rec, nt=nt)
        with Solver(p) as s:
            s.run()
            alone = state(s)
        same = all(np.array_equal(x, y) for x, y in zip(results[i], alone))
        say(f"config2a shot {i}: batch equals the shot alone bitwise: {same}")
        assert np.abs(alone[0]).max() > 0
        for x, y in zip(results[i], alone):
            assert np.array_equal(x, y)
    _parity_vs_oracle(dataclasses.replace(base, source=shots[0], receivers=rec, nt=nt), results[0][0],
                      "config2a shot 0")


def test_config2b_batch_at_shape():
    """Configuration 2, batch B at its shape: 32 patchy models with sides uniform in [300, 1500]
    points and gas fraction uniform in [0, 0.3], in one lock-step batch for 301 steps.  Models 0 and
    31 equal the same models run alone bitwise; model 5 matches the oracle."""
    from poe2d_amd.configs import config2b
    nt = 301
    probs = config2b(nt=nt)
    assert len(probs) == 32
    sides = [(p.nx + 1, p.nz + 1) for p in probs]
    say(f"config2b: sides {min(min(s) for s in sides)} .. {max(max(s) for s in sides)}, "
        f"{sum(p.npoints for p in probs)} points")
    members = [Solver(p) for p in probs]
    with Batch(members) as b:
        b.run(nt)
        ms, n = b.last_run_ms()
        say(f"config2b: {n} steps in {ms:.1f} ms = {sum(p.npoints for p in probs) * n / ms / 1e6:.1f} Gcell-updates/s")
    results = {i: state(members[i]) for i in (0, 5, 31)}
    for m in members:
        m.close()
    for i in (0, 31):
        with Solver(probs[i]) as s:
            s.run()
            alone = state(s)
        same = all(np.array_equal(x, y) for x, y in zip(results[i], alone))
        say(f"config2b model {i} ({sides[i][0]}x{sides[i][1]}): batch equals the model alone bitwise: {same}")
        assert np.abs(alone[0]).max() > 0
        for x, y in zip(results[i], alone):
            assert np.array_equal(x, y)
    _parity_vs_oracle(probs[5], results[5][0], f"config2b model 5 ({sides[5][0]}x{sides[5][1]})")


def test_set_coeffs_reuses_a_handle():
    """A handle given another model's coefficients (poe2d_set_coeffs) equals a handle created for
    that model, bitwise: palette to palette, palette to none and back, after a run and a
    checkpoint; a refused set (Courant check, bad coefficients, a model shared with a clone) leaves
    the handle as it was."""
    import dataclasses
    from poe2d_amd.configs import config1_continuous
    a = make_problem(200, 180, seed=41, kind=_abi.SRC_PRESSURE_DGAUSS, nt=121)
    b = make_problem(200, 180, seed=42, nt=121)
    b = dataclasses.replace(b, source=a.source, receivers=a.receivers)
    c = dataclasses.replace(a, coeffs=config1_continuous(n=200, nt=121).coeffs)  # no palette
    c = dataclasses.replace(c, coeffs={k: v[:181, :] for k, v in c.coeffs.items()})
    refs = {}
    for name, p in (("a", a), ("b", b), ("c", c)):
        with Solver(p) as s:
            s.run()
            refs[name] = state(s)
    with Solver(a) as s:
        s.run(50)
        for name, p in (("b", b), ("c", c), ("a", a), ("c", c)):
            s.set_coeffs(p.coeffs)
            assert s.step_count() == 0 and not s.seismograms().any()
            s.run()
            got = state(s)
            say(f"set_coeffs -> model {name}: bitwise equal {all(np.array_equal(x, y) for x, y in zip(got, refs[name]))}")
            for x, y in zip(got, refs[name]):
                assert np.array_equal(x, y), name
        # refusals leave the handle as it was: a stiffer medium past the Courant bound, a bad mass
        stiff = {k: v.copy() for k, v in a.coeffs.items()}
        stiff["mu"] = stiff["mu"] * 4.0
        stiff["lambdac"] = stiff["lambdac"] * 4.0
        with pytest.raises(_abi.Poe2dError) as ei:
            s.set_coeffs(stiff)
        assert ei.value.status == _abi.POE2D_E_CFL
        bad = {k: v.copy() for k, v in a.coeffs.items()}
        bad["m"][50, 50] = 0.0
        with pytest.raises(_abi.Poe2dError) as ei:
            s.set_coeffs(bad)
        assert ei.value.status == _abi.POE2D_E_INVALID
        s.reset()
        s.run()
        for x, y in zip(state(s), refs["c"]):
            assert np.array_equal(x, y)
        clone = s.clone(source=a.source, receivers=a.receivers)
        with pytest.raises(_abi.Poe2dError) as ei:
            s.set_coeffs(a.coeffs)
        assert ei.value.status == _abi.POE2D_E_STATE
        clone.close()
