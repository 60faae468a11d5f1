"""The FP64 network reference is pinned by central finite differences of its own loss, by the
closed form of GELU and its derivative, and by bf16 rounding's defining properties."""
import numpy as np

from .mlp_reference import bf16, forward, gelu, gelu_grad, loss_and_grads


def random_params(sizes, rng):
    return [(rng.normal(0, 1 / np.sqrt(a), (b, a)), rng.normal(0, 0.1, b)) for a, b in zip(sizes[:-1], sizes[1:])]


def test_gelu_closed_form():
    z = np.linspace(-6, 6, 1201)
    # gelu(z) - gelu(-z) = z and the derivative by finite differences
    assert np.allclose(gelu(z) - gelu(-z), z, atol=1e-12)
    h = 1e-6
    assert np.allclose((gelu(z + h) - gelu(z - h)) / (2 * h), gelu_grad(z), atol=1e-8)
    assert abs(gelu(1.0) - 0.8413447460685429) < 1e-12  # 1 * Phi(1)


def test_bf16_rounding():
    assert bf16(1.0) == 1.0 and bf16(-2.5) == -2.5
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3 * 2^-8 ties up to 1 + 2^-6
    assert bf16(1 + 2 ** -8) == 1.0 and bf16(1 + 3 * 2 ** -8) == 1 + 2 ** -6
    x = np.random.default_rng(0).normal(size=10000)
    r = bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8 + 1e-300)
    assert np.array_equal(bf16(r), r)


def test_reference_gradients_by_finite_differences():
    rng = np.random.default_rng(5)
    sizes = [4, 16, 12, 8, 20]
    params = random_params(sizes, rng)
    x = rng.uniform(-1, 1, (32, 4))
    t = rng.normal(size=(32, 20))
    loss, grads = loss_and_grads(params, x, t)
    h = 1e-6
    for l, (w, b) in enumerate(params):
        for (i, j) in [(0, 0), (w.shape[0] - 1, w.shape[1] - 1), (w.shape[0] // 2, w.shape[1] // 3)]:
            wp = [(ww.copy(), bb.copy()) for ww, bb in params]
            wm = [(ww.copy(), bb.copy()) for ww, bb in params]
            wp[l][0][i, j] += h
            wm[l][0][i, j] -= h
            fd = (loss_and_grads(wp, x, t)[0] - loss_and_grads(wm, x, t)[0]) / (2 * h)
            assert abs(fd - grads[l][0][i, j]) <= 1e-6 * max(1.0, abs(fd)), (l, i, j, fd, grads[l][0][i, j])
        k = b.shape[0] // 2
        bp = [(ww.copy(), bb.copy()) for ww, bb in params]
        bm = [(ww.copy(), bb.copy()) for ww, bb in params]
        bp[l][1][k] += h
        bm[l][1][k] -= h
        fd = (loss_and_grads(bp, x, t)[0] - loss_and_grads(bm, x, t)[0]) / (2 * h)
        assert abs(fd - grads[l][1][k]) <= 1e-6 * max(1.0, abs(fd))


def test_emulation_is_close_to_exact():
    rng = np.random.default_rng(6)
    sizes = [4, 64, 64, 10]
    params = random_params(sizes, rng)
    x = rng.uniform(-1, 1, (32, 4))
    y0, _ = forward(params, x)
    y1, _ = forward(params, x, emulate=True)
    assert np.linalg.norm(y1 - y0) / np.linalg.norm(y0) < 2e-2
