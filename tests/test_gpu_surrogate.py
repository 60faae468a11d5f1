"""Surrogate screening on the GPU at the reduced geometry (121 x 121 points, 600 steps decimated by
6, 21 receivers): data from the batched solver, ensemble training, the spread threshold, and what
screening gains over solving every candidate."""
import numpy as np
import pytest

from poe2d_amd import surrogate as S
from poe2d_amd.mlp import DeviceArray
from poe2d_amd.solver import Solver

pytestmark = pytest.mark.gpu

GEOM = S.SMALL


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.fixture(scope="module")
def trained():
    say("surrogate: simulate 1280 models")
    params = S.draw_params(1280, seed=21)
    gathers, timing = S.simulate(params, geom=GEOM)
    say(f"solver alone: {timing.models_per_second:.0f} models/s wall, {timing.gpu_models_per_second:.0f} on the GPU")
    sur = S.Surrogate(GEOM, members=5, max_batch=1024, receivers=False)
    curves = sur.fit(params[:1024], gathers[:1024], steps=3000, lr=1e-3)
    say("losses " + ", ".join(f"{c[:20].mean():.3f} -> {c[-20:].mean():.4f}" for c in curves))
    yield sur, params, gathers, timing, curves
    sur.close()


def test_simulate_matches_single_solves(trained):
    """simulate's two alternating pools of reused handles give every model the gather a handle
    created for it gives: models of the first use of each pool, of reloads, and of a short last
    batch."""
    _, params, gathers, _, _ = trained
    for k in (0, 64, 130, 200, 777, 1279):
        with Solver(S.model_problem(params[k], geom=GEOM)) as s:
            s.run()
            assert np.array_equal(s.seismograms(), gathers[k]), k
    part, _ = S.simulate(params[:150], geom=GEOM)  # batches of 64, 64 and 22
    assert np.array_equal(part, gathers[:150])


def test_training_converges_and_generalises(trained):
    sur, params, gathers, _, curves = trained
    for c in curves:
        assert np.all(np.isfinite(c)) and c[-20:].mean() < 0.05 * c[:20].mean()
    pred, spread = sur.predict(params[1024:])
    err = sur.errors(pred, gathers[1024:])
    say(f"validation rel L2 p50 {np.median(err):.4f} p90 {np.quantile(err, 0.9):.4f}; spread p50 {np.median(spread):.4f}")
    assert np.median(err) < 0.12
    # the spread ranks the errors: the low-spread half is the more accurate one
    low = spread <= np.median(spread)
    assert np.median(err[low]) < np.median(err[~low])


def test_screening_gains_over_the_solver_alone(trained):
    sur, params, gathers, timing, _ = trained
    thr, err, spread = sur.calibrate(params[1024:], gathers[1024:], bound=0.05, quantile=0.5)
    say(f"threshold {thr:.4f} accepts {np.mean(spread <= thr):.2f} of the validation models")
    assert np.isfinite(thr)
    cand = S.draw_params(512, seed=22)
    res = sur.screen(cand, thr, judge_frac=0.25)
    n_acc = res.n - len(res.fallback)
    ffb = len(res.fallback) / res.n
    p50 = S.mixture_quantile(res.judge_errors, ffb, 0.5)
    say(f"screened {res.n}: {len(res.fallback)} re-solved, {n_acc} accepted, {len(res.judged)} judged; "
        f"{res.models_per_second:.0f} models/s vs {timing.models_per_second:.0f} solving all "
        f"(predict {res.predict_seconds:.3f} s); judged p50 {np.median(res.judge_errors):.4f}, "
        f"p50 after fallback {p50:.4f}")
    assert len(res.judged) == max(1, round(0.25 * n_acc))
    assert not set(res.judged) & set(res.fallback)
    assert res.models_per_second > 1.5 * timing.models_per_second
    # the threshold holds the median error after fallback within the 5 % acceptance bound ...
    assert p50 <= 0.05
    # the screen returns every candidate's gather: the solver's own for the re-solved ones
    assert res.gathers.shape == (res.n,) + S.gather_shape(GEOM)
    for k in res.fallback[:3]:
        with Solver(S.model_problem(cand[k], geom=GEOM)) as s:
            s.run()
            assert np.array_equal(s.seismograms(), res.gathers[k])
    acc = np.setdiff1d(np.arange(res.n), res.fallback)[:2]
    pred, _ = sur.predict(cand[acc])
    assert np.array_equal(pred, res.gathers[acc])
    # ... and the spread sends the worse models to the solver
    assert np.median(res.judge_errors) < np.median(err)
    # the surrogate's own rate, fallback aside
    assert res.n / res.predict_seconds > 20 * timing.gpu_models_per_second
    # the same screen with its gathers kept in device memory: the same gathers, the same judged
    # errors
    dev = DeviceArray((res.n,) + S.gather_shape(GEOM))
    rd = sur.screen(cand, thr, judge_frac=0.25, out=dev)
    say(f"gathers in device memory: {rd.models_per_second:.0f} models/s (predict {rd.predict_seconds:.3f} s)")
    assert rd.gathers is dev
    assert np.array_equal(rd.fallback, res.fallback) and np.array_equal(rd.judged, res.judged)
    assert np.array_equal(rd.judge_errors, res.judge_errors)
    assert np.array_equal(dev.to_host(), res.gathers)
    # judge_frac 0 judges nothing
    assert len(sur.screen(cand[:64], thr, judge_frac=0.0).judged) == 0


def test_receiver_inputs_lower_the_error(trained):
    # the per-receiver gas inputs tell the network where pore pressure drops ~40-fold
    sur, params, gathers, _, _ = trained
    rec = S.Surrogate(GEOM, members=2, max_batch=1024, receivers=True)
    try:
        rec.fit(params[:1024], gathers[:1024], steps=3000, lr=1e-3)
        e_base = sur.errors(sur.predict(params[1024:])[0], gathers[1024:])
        e_rec = rec.errors(rec.predict(params[1024:])[0], gathers[1024:])
        say(f"validation rel L2 p50 / p90: four inputs {np.median(e_base):.4f} / {np.quantile(e_base, 0.9):.4f}, "
            f"with receiver gas {np.median(e_rec):.4f} / {np.quantile(e_rec, 0.9):.4f}")
        assert np.median(e_rec) < np.median(e_base)
        assert np.quantile(e_rec, 0.9) < np.quantile(e_base, 0.9)
    finally:
        rec.close()
