"""GPU parity of the opt-in physics variants (drained stress law, averaged staggered material)
against the oracle in the same mode, on every schedule, and their handling in batches."""
import numpy as np
import pytest

from poe2d_amd import _abi
from poe2d_amd.configs import BRINE, GAS, biot_fields
from poe2d_amd.material import material_biot
from poe2d_amd.solver import Batch, Problem, ReceiverSpec, Solver, SourceSpec

from .helpers import random_patchy_mask
from .test_gpu_parity import assert_parity, run_oracle, say

pytestmark = pytest.mark.gpu

F32, F64 = _abi.F32, _abi.F64


def _problem(law, material, kind=_abi.SRC_FORCE_Y_RICKER, n=240, nt=1000, seed=5):
    # gas patches with a softer frame, so that mu changes across patch edges
    mask = random_patchy_mask(n, n, 0.25, seed=seed)
    coeffs = material_biot(**biot_fields(n, n, mask, BRINE, dict(GAS, mu_fr=5.0e9)))
    ix = np.arange(30, n - 29, 6, dtype=np.int32)
    rec = ReceiverSpec(ix=ix, iz=np.full(len(ix), 2 * n // 3, np.int32))
    return Problem(nx=n, nz=n, dx=0.5, dz=0.5, dt=5e-5, nt=nt, coeffs=coeffs,
                   source=SourceSpec(kind=kind, ix=n // 2, iz=5 * n // 12, amp=1e10), receivers=rec,
                   npml=20, f0=80.0, stress_law=law, staggered_material=material)


def _run(prob, schedule=0, precision=F32, block_w=256, tile_rows=14):
    with Solver(prob, precision=precision) as s:
        s.set_schedule(schedule, block_w, tile_rows)
        s.run()
        return s.seismograms()


@pytest.mark.parametrize("law,material", [(_abi.STRESS_DRAINED, _abi.MATERIAL_AS_WRITTEN),
                                          (_abi.STRESS_AS_WRITTEN, _abi.MATERIAL_AVERAGED),
                                          (_abi.STRESS_DRAINED, _abi.MATERIAL_AVERAGED)])
def test_options_parity_every_schedule(law, material):
    prob = _problem(law, material)
    o32 = run_oracle(prob, F32)
    for schedule, label in ((0, "fused, palette"), (2, "fused, coefficient groups"), (1, "two-pass")):
        say(f"law {law} material {material}: {label}")
        assert_parity(_run(prob, schedule), prob, f"law {law} material {material} {label}", o32=o32)
    # the variant changes the result
    base = _problem(_abi.STRESS_AS_WRITTEN, _abi.MATERIAL_AS_WRITTEN)
    a, b = run_oracle(base, F32), o32
    change = np.abs(a - b).max() / np.abs(a).max()
    say(f"law {law} material {material}: change against the reference's law and material {change:.2e}")
    assert change > 1e-4


def test_options_fp64():
    prob = _problem(_abi.STRESS_DRAINED, _abi.MATERIAL_AVERAGED, kind=_abi.SRC_PRESSURE_DGAUSS, nt=600)
    g = _run(prob, 0, F64)
    o = run_oracle(prob, F64)
    err = np.abs(g - o).max() / np.abs(o).max()
    say(f"FP64 drained + averaged: max |g - o| / max |o| = {err:.2e}")
    assert err < 1e-9


def test_batch_members_must_share_the_material_setting():
    a = Solver(_problem(_abi.STRESS_AS_WRITTEN, _abi.MATERIAL_AVERAGED, n=120, nt=50))
    b = Solver(_problem(_abi.STRESS_DRAINED, _abi.MATERIAL_AVERAGED, n=140, nt=50, seed=6))
    c = Solver(_problem(_abi.STRESS_AS_WRITTEN, _abi.MATERIAL_AS_WRITTEN, n=120, nt=50))
    try:
        # stress laws may differ (they only change coefficients); results equal the members alone
        with Batch([a, b]) as batch:
            batch.run(50)
            together = [a.seismograms(), b.seismograms()]
        for s, t in zip((a, b), together):
            s.reset()
            s.run(50)
            assert np.array_equal(s.seismograms(), t)
        with Batch([a, c]) as batch:
            with pytest.raises(_abi.Poe2dError):
                batch.run(10)
    finally:
        for s in (a, b, c):
            s.close()


def test_unknown_option_values_are_rejected():
    for law, material in ((3, 0), (0, 2)):
        with pytest.raises(_abi.Poe2dError):
            Solver(_problem(law, material, n=60, nt=5))
