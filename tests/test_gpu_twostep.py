"""The two-step kernel (schedule 3: steps n and n+1 in one launch, the intermediate state on
chip) against the single-step kernel (schedule 0) it must reproduce: seismograms, every state
field and the C-PML memory variables (through checkpoints), over odd and even step counts, run
continuations, tile shapes, sources and receivers on tile edges, decimated receivers and the
physics variants; and parity with the oracle."""
import numpy as np
import pytest

from poe2d_amd import _abi
from poe2d_amd.configs import BRINE, GAS, biot_fields, config0
from poe2d_amd.material import material_biot
from poe2d_amd.solver import Problem, ReceiverSpec, Solver, SourceSpec

from .helpers import random_patchy_mask
from .test_gpu_parity import assert_parity, say

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "wx", "wy", "exx", "eyy", "zeta", "sxy", "p")


def _problem(n=300, nt=301, kind=_abi.SRC_PRESSURE_DGAUSS, src=(150, 140), decimate=1, seed=21, npml=20, **opts):
    mask = random_patchy_mask(n, n, 0.25, seed=seed)
    coeffs = material_biot(**biot_fields(n, n, mask, BRINE, dict(GAS, mu_fr=6.0e9)))
    # receivers on a line through the source row and on the first tile-column edges (bw 256:
    # 244 columns per tile; bw 128: 116)
    ix = np.unique(np.r_[np.arange(25, n - 24, 7), [115, 116, 117, 243, 244, 245]]).astype(np.int32)
    ix = ix[ix < n - 20]
    rec = ReceiverSpec(ix=ix, iz=np.full(len(ix), 2 * n // 3, np.int32), decimate=decimate)
    return Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=kind, ix=src[0], iz=src[1], amp=1e10), receivers=rec, npml=npml,
                   f0=80.0, **opts)


def _state(s):
    return s.seismograms(), {f: s.snapshot(f) for f in FIELDS}, s.checkpoint()


def _run(prob, schedule, bw=256, th=12, steps=None, graphs=True):
    with Solver(prob) as s:
        s.set_schedule(schedule, bw, th)
        s.set_graphs(graphs)
        for k in (steps or [prob.nt]):
            s.run(k)
        return _state(s)


def _same(a, b, label):
    (sa, fa, ca), (sb, fb, cb) = a, b
    assert np.array_equal(sa, sb), f"{label}: seismograms differ, max {np.abs(sa - sb).max():.3e} of {np.abs(sa).max():.3e}"
    for f in FIELDS:
        assert np.array_equal(fa[f], fb[f]), f"{label}: field {f} differs"
    assert np.array_equal(ca, cb), f"{label}: checkpoint (state, memory variables, traces) differs"


def test_device_palette_grid():
    with Solver(_problem(nt=2)) as s:
        assert s.material_palette() == 2


@pytest.mark.parametrize("kind", [_abi.SRC_PRESSURE_DGAUSS, _abi.SRC_FORCE_Y_RICKER])
def test_two_step_equals_single_step_bitwise(kind):
    prob = _problem(kind=kind)
    ref = _run(prob, 0, 256, 14)
    for bw, th in ((256, 12), (256, 6), (128, 30), (64, 18), (256, 301)):
        say(f"two-step bw={bw} th={th} kind={kind}")
        _same(_run(prob, 3, bw, th), ref, f"bw {bw} th {th}")


def test_sources_and_receivers_on_tile_edges():
    # the source on the column shared by the first two tiles' halos, odd step count, decimation 3
    for src, dec in (((244, 150), 3), ((116, 40), 2), ((30, 260), 1)):
        prob = _problem(nt=151, src=src, decimate=dec, kind=_abi.SRC_FORCE_Y_RICKER)
        ref = _run(prob, 0, 256, 14)
        for bw, th in ((256, 12), (128, 18), (64, 24)):
            say(f"source {src} decimate {dec}: bw {bw} th {th}")
            _same(_run(prob, 3, bw, th), ref, f"source {src} bw {bw}")


def test_continuations_and_graphs():
    prob = _problem(nt=200)
    ref = _run(prob, 0, 256, 14, steps=[200])
    for steps in ([1, 199], [67, 66, 67], [33, 1, 100, 66], [200]):
        for graphs in (True, False):
            say(f"two-step runs {steps} graphs {graphs}")
            _same(_run(prob, 3, 256, 12, steps=steps, graphs=graphs), ref, f"runs {steps}")


@pytest.mark.parametrize("law,material", [(_abi.STRESS_DRAINED, _abi.MATERIAL_AS_WRITTEN),
                                          (_abi.STRESS_AS_WRITTEN, _abi.MATERIAL_AVERAGED)])
def test_two_step_physics_variants(law, material):
    prob = _problem(nt=121, stress_law=law, staggered_material=material)
    _same(_run(prob, 3, 256, 18), _run(prob, 0, 256, 14), f"law {law} material {material}")


def test_checkpoint_across_schedules():
    prob = _problem(nt=160)
    ref = _run(prob, 0, 256, 14)
    with Solver(prob) as a:
        a.set_schedule(3, 256, 12)
        a.run(81)
        blob = a.checkpoint()
    with Solver(prob) as b:
        b.restore(blob)
        b.set_schedule(0, 256, 14)
        b.run(79)
        _same(_state(b), ref, "restore after two-step, finish single-step")


def test_two_step_parity_config0():
    prob = config0(nt=1000)
    with Solver(prob) as s:
        s.set_schedule(3, 256, 12)
        s.run()
        g = s.seismograms()
    assert_parity(g, prob, "two-step config0 1000 steps")


def test_fp64_falls_back_to_the_single_step_kernel():
    prob = _problem(nt=41)
    with Solver(prob, precision=_abi.F64) as s:
        s.set_schedule(3, 256, 12)
        s.run()
        g64 = s.seismograms()
    with Solver(prob, precision=_abi.F64) as s:
        s.run()
        assert np.array_equal(s.seismograms(), g64)
    with pytest.raises(_abi.Poe2dError):
        with Solver(prob) as s:
            s.set_schedule(3, 512, 12)


def _layered(nlayers, n=160, nt=61):
    # horizontal bands of distinct porosity: nlayers palette entries
    fields = biot_fields(n, n, None, BRINE, GAS)
    rows = np.linspace(0, n + 1, nlayers + 1).round().astype(int)
    for k in range(nlayers):
        fields["phi"][rows[k]:rows[k + 1], :] = 0.18 + 0.004 * k
    coeffs = material_biot(**fields)
    ix = np.arange(25, n - 24, 9, dtype=np.int32)
    rec = ReceiverSpec(ix=ix, iz=np.full(len(ix), 100, np.int32))
    return Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_PRESSURE_DGAUSS, ix=n // 2, iz=n // 2, amp=1e10), receivers=rec,
                   npml=20, f0=80.0)


@pytest.mark.parametrize("extra", [0, 1])
def test_palette_size_limit(extra):
    # at the two-step kernel's palette limit it reads the palette; one entry more and it reads the
    # coefficient groups: either way the results equal schedule 0's bit for bit
    prob = _layered(_abi.MAX_PALETTE_TWO_STEP + extra)
    with Solver(prob) as s:
        assert s.material_palette() == _abi.MAX_PALETTE_TWO_STEP + extra
        s.set_schedule(3, 64, 12)
        s.run()
        g3 = s.seismograms()
    with Solver(prob) as s:
        s.set_schedule(0, 256, 14)
        s.run()
        assert np.array_equal(s.seismograms(), g3)
    assert np.abs(g3).max() > 0


def test_autotune_two_step():
    prob = _problem(nt=20)
    with Solver(prob) as s:
        s.set_schedule(3, 256, 12)
        w, h = s.autotune(4)
        say(f"two-step autotune picked {w}x{h}")
        assert w in (64, 128, 256) and h % 6 == 0
        s.run()
        g = s.seismograms()
    assert np.array_equal(g, _run(prob, 0, 256, 14)[0])


def test_batch_two_step_equals_members_alone():
    """A batch of ragged grids and a clone on the two-step kernel equals each member run alone on the
    single-step kernel, bit for bit (odd step count: pairs, then one single step)."""
    from poe2d_amd.solver import Batch
    probs = [_problem(n=200, nt=91, seed=3, src=(100, 90)), _problem(n=260, nt=91, seed=4, src=(130, 120))]
    alone = []
    for p in probs:
        alone.append(_run(p, 0, 256, 14))
    with Solver(probs[0]) as a, Solver(probs[1]) as b:
        c = a.clone(source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=60, iz=150, amp=1e10),
                    receivers=probs[0].receivers)
        try:
            for s in (a, b, c):
                s.set_schedule(3, 256, 12)
            with Batch([a, b, c]) as batch:
                batch.set_schedule(128, 18)
                batch.run(91)
            _same(_state(a), alone[0], "batch member 0")
            _same(_state(b), alone[1], "batch member 1")
            with Solver(probs[0]) as ref:
                ref.set_source(SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=60, iz=150, amp=1e10))
                ref.run(91)
                _same(_state(c), _state(ref), "batch clone")
        finally:
            c.close()


def _graded(n=300, nt=151, kind=_abi.SRC_PRESSURE_DGAUSS, npml=20, **opts):
    # porosity grading with depth and offset: every row and column its own coefficients (no palette)
    fields = biot_fields(n, n, random_patchy_mask(n, n, 0.2, seed=31), BRINE, dict(GAS, mu_fr=6.0e9))
    ramp = np.linspace(-1, 1, n + 1)
    fields["phi"] = fields["phi"] * (1.0 + 0.15 * ramp[:, None]) * (1.0 + 0.05 * ramp[None, :])
    coeffs = material_biot(**fields)
    ix = np.unique(np.r_[np.arange(25, n - 24, 7), [115, 116, 117, 243, 244, 245]]).astype(np.int32)
    ix = ix[ix < n - 20]
    rec = ReceiverSpec(ix=ix, iz=np.full(len(ix), 2 * n // 3, np.int32))
    return Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=kind, ix=150, iz=140, amp=1e10), receivers=rec, npml=npml, f0=80.0, **opts)


@pytest.mark.parametrize("material", [_abi.MATERIAL_AS_WRITTEN, _abi.MATERIAL_AVERAGED])
def test_two_step_without_palette_equals_single_step(material):
    """A grid with more coefficient tuples than the palette holds runs the two-step kernel through
    the coefficient groups (material read where used, velocity constants per point): bitwise equal
    to the single-step kernel, over tile shapes, an odd step count and both staggered-material
    settings."""
    prob = _graded(staggered_material=material)
    with Solver(prob) as s:
        assert s.material_palette() == 0
    ref = _run(prob, 0, 256, 14)
    for bw, th in ((256, 12), (128, 18), (64, 24)):
        say(f"two-step on the coefficient groups, material {material}: bw {bw} th {th}")
        _same(_run(prob, 3, bw, th), ref, f"groups bw {bw} th {th}")


@pytest.mark.parametrize("npml", [1, 2])
@pytest.mark.parametrize("palette", [True, False])
def test_thin_cpml_strips(npml, palette):
    """C-PML strips of one or two points leave tiles next to the grid's outer two rows and columns
    outside the strips; those must still run as edge tiles, whose updates keep the grid's border
    points fixed (interior tiles without a palette compile those guards out): bitwise equal to the
    single-step kernel over tile shapes whose interior tiles come close to every side."""
    prob = _problem(nt=151, npml=npml) if palette else _graded(npml=npml)
    with Solver(prob) as s:
        assert (s.material_palette() > 0) == palette
    ref = _run(prob, 0, 256, 14)
    assert np.abs(ref[0]).max() > 0
    for bw, th in ((64, 12), (128, 18), (256, 24)):
        say(f"two-step, npml {npml}, palette {palette}: bw {bw} th {th}")
        _same(_run(prob, 3, bw, th), ref, f"npml {npml} bw {bw} th {th}")


def test_two_step_continuous_model():
    """A medium continuous in depth and offset (porosity and frame moduli graded, every point its
    own coefficients): the two-step kernel equals the single-step kernel."""
    from poe2d_amd.configs import config1_continuous
    prob = config1_continuous(nt=121, n=300)
    ref = _run(prob, 0, 256, 14)
    assert np.abs(ref[0]).max() > 0
    _same(_run(prob, 3, 256, 24), ref, "continuous model")


def test_batch_mixing_palette_and_groups():
    """A batch of a palette grid and a grid without one on the two-step kernel (the batch reads
    every member through the groups) equals each member alone on the single-step kernel."""
    from poe2d_amd.solver import Batch
    probs = [_problem(n=200, nt=91, seed=3, src=(100, 90)), _graded(n=240, nt=91)]
    alone = [_run(p, 0, 256, 14) for p in probs]
    with Solver(probs[0]) as a, Solver(probs[1]) as b:
        for s in (a, b):
            s.set_schedule(3, 256, 12)
        with Batch([a, b]) as batch:
            batch.set_schedule(128, 18)
            batch.run(91)
        _same(_state(a), alone[0], "palette member")
        _same(_state(b), alone[1], "groups member")
