"""Configurations 3 and 4 at their shapes on the GPU: 4096 training models and 256 held-out ones at
the configuration-3 geometry (401 x 401 points, 3000 steps decimated by 6, 101 receivers x (v_x,
v_y, p) = 151 500 outputs) from the batched solver, the 5-member ensemble of the 207 -> 512 -> 512
-> 512 -> 151 500 network trained on them, and the 16 384 candidates of configuration 4 screened.
The training schedule is short (the test checks the pipeline at full size; the accuracy it reaches
with 60 000 steps is the bench's): losses fall, the screen returns every candidate's gather, the
re-solved ones equal to single solves and the accepted ones to the ensemble's prediction."""
import numpy as np
import pytest

from poe2d_amd import surrogate as S
from poe2d_amd.solver import Solver

pytestmark = pytest.mark.gpu

GEOM = S.CONFIG3
NTRAIN, NVAL, NCAND = 4096, 256, 16384


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)


def test_configurations_3_and_4_at_full_shape():
    say(f"configuration 3: simulate {NTRAIN} + {NVAL} models at 401^2, 3000 steps")
    params = S.draw_params(NTRAIN + NVAL, seed=31)
    gathers, timing = S.simulate(params, geom=GEOM)
    assert gathers.shape == (NTRAIN + NVAL,) + S.gather_shape(GEOM)
    assert np.isfinite(gathers).all()
    say(f"solver alone: {timing.models_per_second:.1f} models/s wall, {timing.gpu_models_per_second:.1f} on the GPU")
    sur = S.Surrogate(GEOM, members=5, max_batch=1024, receivers=True)
    try:
        assert sur.net_sizes == [4 + 2 * GEOM["nrec"] + 1, 512, 512, 512, 151500]
        curves = sur.fit(params[:NTRAIN], gathers[:NTRAIN], steps=600)
        say("losses " + ", ".join(f"{c[:20].mean():.4f} -> {c[-20:].mean():.4f}" for c in curves))
        for c in curves:
            assert np.all(np.isfinite(c)) and c[-20:].mean() < 0.5 * c[:20].mean()
        pred, spread = sur.predict(params[NTRAIN:])
        err = sur.errors(pred, gathers[NTRAIN:])
        say(f"held-out rel L2 p50 {np.median(err):.4f}")
        assert np.isfinite(err).all() and np.median(err) < 0.5
        # a threshold that re-solves about 1 % of the candidates
        thr = float(np.quantile(spread, 0.99))
        del gathers, pred

        say(f"configuration 4: screen {NCAND} candidates")
        cand = S.draw_params(NCAND, seed=32)
        res = sur.screen(cand, thr, judge_frac=0.01)
        nfb = len(res.fallback)
        say(f"{nfb} re-solved, {len(res.judged)} judged; {res.models_per_second:.0f} models/s "
            f"(predict {res.predict_seconds:.2f} s); judged p50 {np.median(res.judge_errors):.4f}")
        assert res.n == NCAND and res.gathers.shape == (NCAND,) + S.gather_shape(GEOM)
        assert np.isfinite(res.gathers).all()
        assert 0 < nfb < 0.05 * NCAND
        assert res.models_per_second > timing.models_per_second
        # re-solved candidates carry the solver's own gather
        for k in res.fallback[:2]:
            with Solver(S.model_problem(cand[k], geom=GEOM)) as s:
                s.run()
                assert np.array_equal(s.seismograms(), res.gathers[k])
        # accepted candidates carry the ensemble's prediction (rows do not depend on the batch)
        acc = np.setdiff1d(np.arange(NCAND), res.fallback)[[0, -1]]
        g, _ = sur.predict(cand[acc])
        assert np.array_equal(g, res.gathers[acc])
    finally:
        sur.close()
