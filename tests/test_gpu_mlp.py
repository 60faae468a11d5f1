"""The surrogate network on the GPU against the FP64 reference with bf16 emulation
(tests/mlp_reference.py): the matrix-core GEMM, forward pass, gradients, training and the
ensemble spread."""
import numpy as np
import pytest

from poe2d_amd.mlp import Network, gemm_bf16

from .mlp_reference import bf16, forward, loss_and_grads

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("m,n,k,ksplit", [(256, 512, 512, 1), (200, 300, 100, 1), (256, 512, 4096, 8),
                                          (130, 129, 64, 2), (1, 5, 7, 1)])
def test_gemm_matches_bf16_product(m, n, k, ksplit):
    rng = np.random.default_rng(m + n + k)
    a = rng.normal(size=(m, k)).astype(np.float32)
    b = rng.normal(size=(n, k)).astype(np.float32)
    c = gemm_bf16(a, b, ksplit)
    ref = bf16(a) @ bf16(b).T
    err = rel(c, ref)
    say(f"gemm {m}x{n}x{k} split {ksplit}: rel {err:.2e}")
    assert err < 1e-6


def random_params(sizes, rng):
    return [(rng.normal(0, 1 / np.sqrt(a), (b, a)).astype(np.float32), rng.normal(0, 0.1, b).astype(np.float32))
            for a, b in zip(sizes[:-1], sizes[1:])]


def test_forward_and_gradients_match_reference():
    rng = np.random.default_rng(11)
    sizes = [4, 96, 64, 160, 300]
    params = random_params(sizes, rng)
    x = rng.uniform(-1, 1, (64, 4)).astype(np.float32)
    t = rng.normal(size=(64, 300)).astype(np.float32)
    with Network(sizes, members=1, max_batch=128) as net:
        net.set_params(params)
        got = net.get_params()
        for (w, b), (w2, b2) in zip(params, got):
            assert np.array_equal(w, w2) and np.array_equal(b, b2)
        y = net.forward(x)
        y_ref, _ = forward(params, x, emulate=True)
        ey = rel(y, y_ref)
        loss, grads = net.gradients(x, t)
        loss_ref, grads_ref = loss_and_grads(params, x, t, emulate=True)
        say(f"forward rel {ey:.2e}; loss {loss:.6f} vs {loss_ref:.6f}")
        assert ey < 1e-4
        assert abs(loss - loss_ref) <= 1e-4 * loss_ref
        for l, ((gw, gb), (rw, rb)) in enumerate(zip(grads, grads_ref)):
            ew, eb = rel(gw, rw), rel(gb, rb)
            say(f"layer {l}: dW rel {ew:.2e}, db rel {eb:.2e}")
            assert ew < 2e-3 and eb < 2e-3


def test_training_fits_a_smooth_map_and_spread_tracks_disagreement():
    rng = np.random.default_rng(12)
    n, nin, nout = 1024, 4, 384
    x = rng.uniform(-1, 1, (n, nin)).astype(np.float32)
    freq = rng.normal(size=(nin, nout))
    t = np.sin(x @ freq).astype(np.float32)
    with Network([nin, 128, 128, 128, nout], members=3, max_batch=128, seed=3) as net:
        net.set_dataset(x, t)
        first = last = None
        for member in range(3):
            idx = np.concatenate([rng.permutation(n) for _ in range(40)])[: 300 * 64]
            losses = net.train(idx, 64, 2e-3, member)
            say(f"member {member}: loss {losses[:5].mean():.4f} -> {losses[-5:].mean():.4f}")
            assert np.all(np.isfinite(losses))
            assert losses[-5:].mean() < 0.1 * losses[:5].mean()
            first, last = losses[:5].mean(), losses[-5:].mean()
        mean, spread = net.predict(x[:200])
        fit = rel(mean, t[:200])
        say(f"ensemble fit rel L2 {fit:.3f}, spread median {np.median(spread):.3f}")
        assert fit < 0.3
        # far outside the training inputs the members disagree more than inside
        _, far = net.predict(3.0 * np.sign(x[:200]), want_mean=False)
        assert np.median(far) > 1.5 * np.median(spread)
        # identical members: zero spread
        p = net.get_params(0)
        for member in (1, 2):
            net.set_params(p, member)
        _, same = net.predict(x[:64], want_mean=False)
        assert np.all(same <= 1e-6)
