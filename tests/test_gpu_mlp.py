"""The surrogate network on the GPU against the FP64 reference with bf16 emulation
(tests/mlp_reference.py): the matrix-core GEMM, forward pass, gradients, training and the
ensemble spread."""
import numpy as np
import pytest

from poe2d_amd.mlp import DeviceArray, Network, gemm_bf16, pinned_empty

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


def test_screening_network_shape_forward_and_gradients():
    """The screening network at its full shape, 207 -> 512 -> 512 -> 512 -> 151 500 (four model
    parameters and 203 layout inputs in; 101 receivers x 3 components x 500 samples out), on a batch
    of 32 samples: forward pass, loss and every layer's gradients against the FP64 reference."""
    rng = np.random.default_rng(13)
    sizes = [207, 512, 512, 512, 151500]
    params = random_params(sizes, rng)
    x = rng.uniform(-1, 1, (32, sizes[0])).astype(np.float32)
    t = rng.normal(size=(32, sizes[-1])).astype(np.float32)
    with Network(sizes, members=1, max_batch=32) as net:
        net.set_params(params)
        y = net.forward(x)
        y_ref, _ = forward(params, x, emulate=True)
        ey = rel(y, y_ref)
        loss, grads = net.gradients(x, t)
    loss_ref, grads_ref = loss_and_grads(params, x, t, emulate=True)
    say(f"207-512-512-512-151500: forward rel {ey:.2e}; loss {loss:.6f} vs {loss_ref:.6f}")
    # The hidden products are split over k (64 blocks instead of 8) and their FP32 partial sums
    # added in another order than the FP64 reference's sums: a pre-activation within an FP32
    # rounding of a bf16 rounding boundary can round the other way, one bf16 ulp (2^-8 relative),
    # which moves this output by about 1e-4 relative per such activation.  The bound allows a few.
    assert ey < 1e-3
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    for l, ((gw, gb), (rw, rb)) in enumerate(zip(grads, grads_ref)):
        ew, eb = rel(gw, rw), rel(gb, rb)
        say(f"  layer {l} ({sizes[l]} -> {sizes[l + 1]}): dW rel {ew:.2e}, db rel {eb:.2e}")
        assert ew < 2e-3 and eb < 2e-3


def test_adam_matches_the_bias_corrected_update():
    """Five training steps of one member against the bias-corrected Adam update (beta 0.9 / 0.999,
    eps 1e-8) restated in FP64 from the GPU's own gradients of each step's batch (the gradients
    themselves are checked against the reference above), and the step's loss against theirs; after
    every step the bf16 copy of each layer equals the round-to-nearest-even of its master weights
    and the transposed copy equals its transpose, bit for bit."""
    rng = np.random.default_rng(21)
    sizes = [4, 96, 64, 160, 300]
    params = random_params(sizes, rng)
    x = rng.uniform(-1, 1, (96, sizes[0])).astype(np.float32)
    t = rng.normal(size=(96, sizes[-1])).astype(np.float32)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    with Network(sizes, members=2, max_batch=64) as net:
        net.set_params(params, member=1)
        net.set_dataset(x, t)
        untouched = net.get_params(0)
        w = [(w.astype(np.float64), b.astype(np.float64)) for w, b in params]
        m = [(np.zeros_like(a), np.zeros_like(b)) for a, b in w]
        v = [(np.zeros_like(a), np.zeros_like(b)) for a, b in w]
        for step in range(1, 6):
            idx = rng.permutation(96)[:64].astype(np.int32)
            loss_ref, grads = net.gradients(x[idx], t[idx], member=1)
            # the training step forms the loss gradient in the output product's epilogue and the
            # output layer's Adam step in its weight-gradient product's epilogue
            loss = net.train(idx, 64, lr, member=1)[0]
            assert abs(loss - loss_ref) <= 1e-5 * loss_ref, (loss, loss_ref)
            got = net.get_params(1)
            copies = net.get_bf16(1)
            c1, c2 = 1.0 - b1 ** step, 1.0 - b2 ** step
            for l in range(len(sizes) - 1):
                for k in range(2):
                    g = grads[l][k].astype(np.float64)
                    mk = b1 * m[l][k] + (1 - b1) * g
                    vk = b2 * v[l][k] + (1 - b2) * g * g
                    upd = lr * (mk / c1) / (np.sqrt(vk / c2) + eps)
                    ref = got[l][k].astype(np.float64)
                    # restated from this step's own start point, so errors do not accumulate
                    start = w[l][k]
                    err = np.abs(ref - (start - upd)).max() / max(np.abs(upd).max(), 1e-30)
                    assert err < 1e-3, (step, l, k, err)
                    m[l] = (mk, m[l][1]) if k == 0 else (m[l][0], mk)
                    v[l] = (vk, v[l][1]) if k == 0 else (v[l][0], vk)
                    w[l] = (ref, w[l][1]) if k == 0 else (w[l][0], ref)
                w16, w16t = copies[l]
                expect = (bf16(got[l][0]).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
                assert np.array_equal(w16, expect), (step, l)
                assert np.array_equal(w16t, w16.T), (step, l)
            say(f"Adam step {step}: master weights within 1e-3 of the update, bf16 copies exact")
        # the other member's weights are untouched
        for (a, c), (b, d) in zip(untouched, net.get_params(0)):
            assert np.array_equal(a, b) and np.array_equal(c, d)


def test_predict_into_page_locked_memory_with_output_scale():
    """predict writes the ensemble mean into a caller's page-locked array, the same bits as into
    its own; the device-side output scale multiplies the mean as a host FP32 product would and
    leaves the spread (of the unscaled outputs) alone."""
    rng = np.random.default_rng(17)
    with Network([8, 64, 200], members=3, max_batch=64, seed=5) as net:
        x = rng.standard_normal((100, 8)).astype(np.float32)  # two batches of at most 64 rows
        m0, s0 = net.predict(x)
        out = pinned_empty((100, 200))
        m1, s1 = net.predict(x, out=out)
        assert m1 is out and np.array_equal(m1, m0) and np.array_equal(s1, s0)
        scale = rng.uniform(0.5, 2.0, 200).astype(np.float32)
        net.set_output_scale(scale)
        m2, s2 = net.predict(x)
        assert np.array_equal(m2, m0 * scale[None, :]) and np.array_equal(s2, s0)
        net.set_output_scale(None)
        m3, _ = net.predict(x)
        assert np.array_equal(m3, m0)


def test_predict_into_device_memory():
    """predict writes the ensemble mean into device memory, whole or into a view of some rows, the
    same bits as into host memory; rows move between host and device unchanged."""
    rng = np.random.default_rng(19)
    with Network([8, 64, 200], members=3, max_batch=64, seed=6) as net:
        x = rng.standard_normal((100, 8)).astype(np.float32)
        m0, s0 = net.predict(x)
        dev = DeviceArray((130, 200))
        m1, s1 = net.predict(x, out=dev.rows(10, 110))
        assert np.array_equal(s1, s0)
        assert isinstance(m1, DeviceArray) and m1.shape == (100, 200)
        assert np.array_equal(m1.to_host(), m0)
        assert np.array_equal(dev.to_host([10, 109, 50]), m0[[0, 99, 40]])
        vals = rng.standard_normal((2, 200)).astype(np.float32)
        dev.set_rows([0, 129], vals)
        assert np.array_equal(dev.to_host([0, 129]), vals)
        assert np.array_equal(dev.to_host()[10:110], m0)  # rows outside the set ones untouched
        with pytest.raises(ValueError):
            net.predict(x, out=dev)  # 130 rows for 100 inputs
        with pytest.raises(IndexError):
            dev.rows(0, 131)


def test_train_many_equals_training_one_member_at_a_time():
    """train_many (members side by side on two streams, an odd count so that each member's steps
    alternate between them) leaves every member's weights and losses as train() does, bit for bit."""
    rng = np.random.default_rng(23)
    sizes = [16, 128, 128, 300]
    x = rng.standard_normal((512, 16)).astype(np.float32)
    t = np.tanh(x @ rng.standard_normal((16, 300))).astype(np.float32)
    idx = [rng.integers(0, 512, size=40 * 64) for _ in range(3)]
    with Network(sizes, members=3, max_batch=64, seed=9) as a, Network(sizes, members=3, max_batch=64, seed=9) as b:
        a.set_dataset(x, t)
        b.set_dataset(x, t)
        la = [a.train(idx[k], 64, 1e-3, member=k) for k in range(3)]
        lb = b.train_many(np.stack(idx), 64, 1e-3, members=[0, 1, 2])
        for k in range(3):
            assert np.array_equal(la[k], lb[k]), k
            for layer, ((wa, ba), (wb, bb)) in enumerate(zip(a.get_params(k), b.get_params(k))):
                assert np.array_equal(wa, wb) and np.array_equal(ba, bb), (k, layer)
        # a second call continues from there (the Adam step counts and moments carry on)
        a.train(idx[1][:640], 64, 1e-3, member=1)
        b.train_many(idx[1][None, :640], 64, 1e-3, members=[1])
        for (wa, ba), (wb, bb) in zip(a.get_params(1), b.get_params(1)):
            assert np.array_equal(wa, wb) and np.array_equal(ba, bb)
