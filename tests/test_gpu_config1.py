"""Configuration 1 at its full shape: 4001 x 4001 points, patchy gas (fraction 0.15) over the
brine rock, Ricker force at the centre, 801 receivers x (vx, vy, p), on the default FP32 schedule
(two time steps per launch).

- Whole grid from a random state: every point of the grid carries signal, so every tile of the
  launch takes part (edge and interior tiles, each XCD's share, the last row group, byte offsets
  up to the far corner of each field group).  21 steps (ten step pairs and a single step) on the
  two-step kernel equal the single-step kernel bitwise (seismograms, all fields, memory variables
  through the checkpoint) and the oracle within FP32 rounding.
- The source-driven run against the FP32 oracle over its first 1000 steps (the wave has passed
  about 160 of the 801 receivers by then), per trace (1e-4), with the FP64 build of the same
  kernels as the FP64 reference that recognises FP32-noisy traces.
- The full 6000-step run against the FP64 build, per trace (1e-4).
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleSolver
from poe2d_amd import _abi
from poe2d_amd.configs import config1
from poe2d_amd.solver import Solver

from .helpers import trace_rel_l2
from .test_gpu_parity import assert_parity, say

pytestmark = pytest.mark.gpu

F32, F64 = _abi.F32, _abi.F64
STATE = ("vx", "vy", "wx", "wy", "exx", "eyy", "zeta", "sxy")
FIELDS = STATE + ("p", "sxx", "syy")
# magnitudes of a plane wave of velocity amplitude 1e-3 m/s in the brine rock: strain ~ v / c_p,
# shear stress ~ mu x strain
SCALE = {"vx": 1e-3, "vy": 1e-3, "wx": 1e-3, "wy": 1e-3, "exx": 3e-7, "eyy": 3e-7, "zeta": 3e-7, "sxy": 2.4e3}


@pytest.fixture(scope="module")
def c1():
    say("config1: building the 4001 x 4001 patchy model")
    prob = config1()
    assert prob.nx == prob.nz == 4000 and prob.receivers.n == 801 and prob.nt == 6000
    return prob


def random_state(n, seed=7):
    """Uniform noise on rows and columns 2 .. n-2 of each state field, zero elsewhere: the points the
    reference's updates never reach (and the strains of row n-1, which its normal-stress pass reads
    but its strain pass never writes) keep zero, as in every state the reference can reach."""
    rng = np.random.default_rng(seed)
    out = {}
    for f in STATE:
        a = np.zeros((n + 1, n + 1), np.float32)
        a[2:n - 1, 2:n - 1] = rng.uniform(-SCALE[f], SCALE[f], (n - 3, n - 3)).astype(np.float32)
        out[f] = a
    return out


def test_config1_whole_grid_from_random_state(c1):
    nsteps = 21
    init = random_state(c1.nx)

    def gpu(schedule, bw, th):
        with Solver(c1) as s:
            s.set_schedule(schedule, bw, th)
            for f, a in init.items():
                s.set_field(f, a)
            s.run(nsteps)
            return (s.seismograms()[..., :nsteps], {f: s.snapshot(f) for f in FIELDS}, s.checkpoint())

    say("config1 random state: two-step kernel 256x24, 21 steps")
    two = gpu(3, 256, 24)
    say("config1 random state: single-step kernel 256x8, 21 steps")
    one = gpu(0, 256, 8)
    assert np.array_equal(two[0], one[0]), "seismograms differ between the kernels"
    for f in FIELDS:
        assert np.array_equal(two[1][f], one[1][f]), f"field {f} differs between the kernels"
        assert np.abs(two[1][f]).max() > 0, f
    assert np.array_equal(two[2], one[2]), "checkpoints (state, memory variables, traces) differ"
    # the noise reaches every region of the grid, the absorbing strips included
    for f in ("vx", "sxy"):
        a = two[1][f]
        for region in (a[:20, :20], a[-20:, -20:], a[2000:2020, :20], a[:20, 3990:], a[3990:, 2000:2020]):
            assert np.abs(region).max() > 0, f

    say("config1 random state: oracle, 21 steps")
    o = OracleSolver(c1, F32)
    for f, a in init.items():
        o.set_field(f, a)
    o.run(nsteps)
    for f in FIELDS:
        a, b = two[1][f].astype(np.float64), o.snapshot(f).astype(np.float64)
        err = np.abs(a - b).max() / np.abs(b).max()
        say(f"  field {f}: max |gpu - oracle| / max |oracle| = {err:.2e}")
        assert err <= 2e-5, f
    e = trace_rel_l2(two[0], o.seismograms()[..., :nsteps])
    say(f"  traces: max rel L2 = {e.max():.2e}")
    assert e.max() <= 1e-4
    o.close()


def test_config1_oracle_parity_first_1000_steps(c1):
    nsteps = 1000
    say(f"config1 source run: GPU FP32 (default schedule) and FP64, {nsteps} steps")
    with Solver(c1) as s:
        s.run(nsteps)
        g = s.seismograms()[..., :nsteps]
    with Solver(c1, precision=F64) as s:
        s.run(nsteps)
        g64 = s.seismograms()[..., :nsteps]
    say(f"config1 source run: oracle FP32, {nsteps} steps on all host threads")
    o = OracleSolver(c1, F32)
    o.run(nsteps)
    o32 = o.seismograms()[..., :nsteps]
    o.close()
    reached = int((np.linalg.norm(o32[1], axis=-1) > 1e-3 * np.linalg.norm(o32[1], axis=-1).max()).sum())
    say(f"  {reached} of 801 v_y traces above 1e-3 of the largest")
    assert reached >= 120
    assert_parity(g, c1, "config1 first 1000 steps", o32=o32, o64=g64, exempt=0)


def test_config1_full_run_fp32_vs_fp64(c1):
    say("config1 full run: 6000 steps, FP32 (default schedule) and FP64")
    with Solver(c1) as s:
        s.run()
        g32 = s.seismograms()
        ms, n = s.last_run_ms()
        say(f"  FP32: {n} steps in {ms:.1f} ms ({c1.npoints * n / ms / 1e6:.1f} Gcell-updates/s)")
    with Solver(c1, precision=F64) as s:
        s.run()
        g64 = s.seismograms()
        ms, n = s.last_run_ms()
        say(f"  FP64: {n} steps in {ms:.1f} ms")
    assert np.isfinite(g32).all() and np.isfinite(g64).all()
    assert np.abs(g64).max() > 0
    e = trace_rel_l2(g32, g64)
    c, r = np.unravel_index(e.argmax(), e.shape)
    say(f"  FP32 vs FP64 max rel L2 = {e.max():.3e} (component {c}, receiver {r}); median {np.median(e):.2e}")
    assert e.max() <= 1e-4


def test_continuous_model_whole_grid_two_step_vs_single_step():
    """The continuously varying medium at configuration 1's size (no palette: the two-step kernel
    reads the coefficient groups), from a random state: 21 steps on the two-step kernel equal the
    single-step kernel bitwise over the whole grid."""
    from poe2d_amd.configs import config1_continuous
    prob = config1_continuous(nt=21)
    init = random_state(prob.nx, seed=11)

    def gpu(schedule, bw, th):
        with Solver(prob) as s:
            assert s.material_palette() == 0
            s.set_schedule(schedule, bw, th)
            for f, a in init.items():
                s.set_field(f, a)
            s.run(21)
            return s.seismograms(), {f: s.snapshot(f) for f in ("vx", "vy", "wy", "sxy", "p")}, s.checkpoint()

    say("continuous 4001^2 random state: two-step kernel on the coefficient groups, 21 steps")
    two = gpu(3, 256, 24)
    say("continuous 4001^2 random state: single-step kernel, 21 steps")
    one = gpu(0, 256, 8)
    assert np.array_equal(two[0], one[0])
    for f in two[1]:
        assert np.abs(two[1][f]).max() > 0, f
        assert np.array_equal(two[1][f], one[1][f]), f
    assert np.array_equal(two[2], one[2])
