"""Host side of the surrogate screening (no GPU): parameter draws, network inputs, the models built
from parameters, and the choice of the spread threshold."""
import numpy as np
import pytest

from poe2d_amd import surrogate as S


def test_draws_cover_the_ranges():
    p = S.draw_params(20000, seed=5)
    gas, radius, phi, kappa = p.T
    assert gas.min() >= 0 and gas.max() <= 0.3
    assert radius.min() >= 2 and radius.max() <= 20
    assert phi.min() >= 0.15 and phi.max() <= 0.35
    assert kappa.min() >= 1e-14 and kappa.max() <= 1e-11
    # log-uniform: the median sits at the geometric mean of the ends
    assert abs(np.median(np.log(radius)) - np.log(np.sqrt(40.0))) < 0.05
    assert abs(np.median(np.log10(kappa)) + 12.5) < 0.05
    assert abs(np.median(gas) - 0.15) < 0.01
    assert np.array_equal(p, S.draw_params(20000, seed=5))


def test_features_map_range_ends_to_the_input_scale():
    # most patches: largest gas fraction, smallest radius; none: no gas
    lo = [[0.0, 2.0, 0.15, 1e-14]]
    hi = [[0.3, 20.0, 0.35, 1e-11]]
    f_lo, f_hi = S.features(lo)[0], S.features(hi)[0]
    np.testing.assert_allclose(f_lo[1:], -S.INPUT_SCALE, atol=1e-6)
    np.testing.assert_allclose(f_hi[1:], S.INPUT_SCALE, atol=1e-6)
    assert f_lo[0] == pytest.approx(-S.INPUT_SCALE)
    assert S.features([[0.3, 2.0, 0.25, 1e-12]])[0][0] == pytest.approx(S.INPUT_SCALE)
    np.testing.assert_allclose(S.features([[0.15, np.sqrt(40.0), 0.25, 10 ** -12.5]])[0][1:], 0.0, atol=1e-6)
    # the nominal patch count: area x gas / (pi r^2) over the 200 m x 200 m model
    assert S.patch_count([[0.1, 5.0, 0.2, 1e-12]])[0] == pytest.approx(0.1 * 200.0 ** 2 / (np.pi * 25.0))
    # monotone in the count, which grows with gas and falls with radius
    a = S.features([[0.1, 5.0, 0.2, 1e-12], [0.2, 5.0, 0.2, 1e-12], [0.2, 10.0, 0.2, 1e-12]])[:, 0]
    assert a[0] < a[1] and a[2] < a[1]


def test_model_is_a_function_of_its_parameters():
    geom = S.SMALL
    p = np.array([0.2, 5.0, 0.3, 1e-13])
    a, b = S.model_problem(p, geom=geom), S.model_problem(p, geom=geom)
    for k in a.coeffs:
        assert np.array_equal(a.coeffs[k], b.coeffs[k])
    c = S.model_problem(np.array([0.2, 6.0, 0.3, 1e-13]), geom=geom)
    assert not np.array_equal(a.coeffs["M"], c.coeffs["M"])
    # the gas fraction is reached
    gas = np.isclose(a.coeffs["M"], a.coeffs["M"].min())
    assert 0.2 <= gas.mean() < 0.25
    # a different layout seed moves the patches
    d = S.model_problem(p, seed=S.LAYOUT_SEED + 1, geom=geom)
    assert not np.array_equal(a.coeffs["M"], d.coeffs["M"])
    # no gas: the plain rock with the drawn porosity and permeability
    e = S.model_problem(np.array([0.0, 5.0, 0.3, 1e-13]), geom=geom)
    assert np.ptp(e.coeffs["M"]) == 0 and np.ptp(e.coeffs["b"]) == 0


@pytest.mark.parametrize("geom", [S.SMALL, S.CONFIG3])
def test_geometry(geom):
    prob = S.model_problem(np.array([0.1, 4.0, 0.25, 1e-12]), geom=geom)
    n, npml = geom["n"], geom["npml"]
    rx, rz = prob.receivers.ix, prob.receivers.iz
    assert rx.min() > npml and rx.max() < n - npml and np.all(rz < n - npml)
    assert prob.source.ix == n // 2 and prob.source.iz == n // 2
    assert S.gather_shape(geom) == (3, geom["nrec"], geom["nt"] // geom["dec"])
    if geom is S.CONFIG3:
        assert S.gather_shape(geom) == (3, 101, 500) and np.prod(S.gather_shape(geom)) == 151500
        assert rx[0] == 50 and rx[-1] == 350 and rz[0] == 240


def test_after_fallback_zeroes_the_re_solved_models():
    err = np.array([0.01, 0.2, 0.03])
    spread = np.array([0.1, 0.5, 0.2])
    np.testing.assert_array_equal(S.after_fallback(err, spread, 0.2), [0.01, 0.0, 0.03])
    np.testing.assert_array_equal(S.after_fallback(err, spread, -np.inf), [0, 0, 0])


def test_spread_threshold_known_cases():
    err = np.array([0.01, 0.02, 0.2, 0.03])
    spread = np.array([0.1, 0.2, 0.3, 0.4])
    assert S.spread_threshold(err, spread, 0.05, 0.99) == 0.2   # accepting model 2 breaks p99
    assert S.spread_threshold(err, spread, 0.05, 0.5) == 0.4    # the median holds with all accepted
    assert S.spread_threshold(np.full(4, 0.5), spread) == -np.inf


def test_spread_threshold_is_the_largest_that_keeps_the_bound():
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = 50
        spread = rng.uniform(0, 1, n)
        err = np.clip(0.08 * spread + rng.normal(0, 0.02, n), 0, None)
        for q in (0.5, 0.9, 0.99):
            t = S.spread_threshold(err, spread, 0.05, q)
            ok = lambda th: np.quantile(np.where(spread <= th, err, 0.0), q) <= 0.05
            if np.isfinite(t):
                assert ok(t)
            assert not any(ok(th) for th in spread[spread > t])


def test_quantile_upper_rank_is_the_binomial_bound():
    from scipy import stats
    for n, q, c in ((2048, 0.99, 0.95), (2048, 0.5, 0.95), (100, 0.9, 0.9), (50, 0.5, 0.99), (10, 0.5, 0.5)):
        k = S.quantile_upper_rank(n, q, c)
        # the smallest k with P(Binomial(n, q) <= k - 1) >= c
        assert stats.binom.cdf(k - 1, n, q) >= c and stats.binom.cdf(k - 2, n, q) < c
    # coverage: X_(k) of n uniform samples lies above the true q-quantile (= q) in >= c of trials
    rng = np.random.default_rng(8)
    n, q, c = 400, 0.9, 0.95
    k = S.quantile_upper_rank(n, q, c)
    hits = np.mean([np.sort(rng.uniform(0, 1, n))[k - 1] >= q for _ in range(4000)])
    assert hits >= c - 0.015


def test_spread_threshold_with_confidence_uses_the_upper_order_statistic():
    rng = np.random.default_rng(4)
    n = 300
    spread = rng.uniform(0, 1, n)
    err = np.clip(0.08 * spread + rng.normal(0, 0.02, n), 0, None)
    k = S.quantile_upper_rank(n, 0.9, 0.95)
    t = S.spread_threshold(err, spread, 0.05, 0.9, confidence=0.95)
    ok = lambda th: np.sort(np.where(spread <= th, err, 0.0))[k - 1] <= 0.05
    assert np.isfinite(t) and ok(t)
    assert not any(ok(th) for th in spread[spread > t])
    # stricter than the plain estimate
    assert t <= S.spread_threshold(err, spread, 0.05, 0.9)


def test_mixture_quantile():
    je = np.linspace(0.0, 0.1, 101)
    assert S.mixture_quantile(je, 0.5, 0.4) == 0.0
    assert S.mixture_quantile(je, 0.0, 0.5) == pytest.approx(0.05)
    # half re-solved: the overall p75 is the judged median
    assert S.mixture_quantile(je, 0.5, 0.75) == pytest.approx(0.05)
    assert S.mixture_quantile(np.zeros(0), 0.1, 0.99) == 0.0
    # agrees with the quantile of an explicit mixture
    rng = np.random.default_rng(0)
    e = rng.uniform(0, 0.2, 4000)
    full = np.concatenate([e, np.zeros(1000)])
    for q in (0.3, 0.5, 0.9, 0.99):
        assert abs(S.mixture_quantile(e, 0.2, q) - np.quantile(full, q)) < 2e-3


def test_receiver_inputs_follow_the_model_material():
    # the per-receiver gas input is +INPUT_SCALE exactly where the solver's model holds gas at the
    # receiver (the gas fluid modulus is ~40 times the brine's smaller), -INPUT_SCALE elsewhere
    params = S.draw_params(12, seed=11)
    params[0, 0] = 0.0  # one model without gas
    ix, iz = S.receiver_points(S.SMALL)
    f = S.features(params, S.SMALL, receivers=True)
    nrec = S.SMALL["nrec"]
    assert f.shape == (12, 4 + 2 * nrec + 1)
    np.testing.assert_array_equal(f[:, :4], S.features(params, S.SMALL))
    seen_gas = 0
    for k, p in enumerate(params):
        M = S.model_problem(p, geom=S.SMALL).coeffs["M"][iz, ix]
        M_brine = S.model_problem([0.0, *p[1:]], geom=S.SMALL).coeffs["M"][iz, ix]
        gas_here = M < 0.5 * M_brine
        np.testing.assert_array_equal(f[k, 4:4 + nrec], np.where(gas_here, S.INPUT_SCALE, -S.INPUT_SCALE))
        # a receiver under gas has gas at the end of its path; inputs stay within the scale
        assert np.all(f[k, 4 + nrec:4 + 2 * nrec][gas_here] > -S.INPUT_SCALE)
        assert np.all(np.abs(f[k]) <= S.INPUT_SCALE + 1e-6)
        seen_gas += int(gas_here.sum())
    assert seen_gas > 0
    assert np.all(f[0, 4:] == -S.INPUT_SCALE)


def test_path_inputs_count_the_gas_on_the_segment():
    # brute force on one model: walk the source-receiver segment through the mask
    p = S.draw_params(40, seed=12)
    p = p[p[:, 0] > 0.2][:1]
    g = S.SMALL
    lay = S.layout_inputs(p, g)[0]
    mask = S.patchy_mask(g["n"], g["n"], g["dx"], g["dx"], float(p[0, 0]), S.LAYOUT_SEED, float(p[0, 1]), float(p[0, 1]))
    ix, iz = S.receiver_points(g)
    c = g["n"] // 2
    for k in (0, len(ix) // 2, len(ix) - 1):
        share = np.mean([mask[int(np.rint(c + (iz[k] - c) * (i + 0.5) / S.PATH_SAMPLES)),
                              int(np.rint(c + (ix[k] - c) * (i + 0.5) / S.PATH_SAMPLES))]
                         for i in range(S.PATH_SAMPLES)])
        assert lay[g["nrec"] + k] == pytest.approx(share)
    assert lay[-1] == float(mask[c, c])


def test_model_coefficients_equal_the_whole_grid_evaluation():
    # a model's two materials, evaluated once and placed by the mask, give the bits the material
    # front end gives over the whole grid
    from poe2d_amd.configs import BRINE, GAS, biot_fields
    from poe2d_amd.material import material_biot
    g = S.SMALL
    n = g["n"]
    for p in list(S.draw_params(3, seed=13)) + [[0.0, 5.0, 0.2, 1e-12]]:
        gas, radius, phi, kappa = (float(v) for v in p)
        mask = S.patchy_mask(n, n, g["dx"], g["dx"], gas, S.LAYOUT_SEED, radius, radius) if gas > 0 else None
        whole = material_biot(**biot_fields(n, n, mask, dict(BRINE, phi=phi, kappa=kappa),
                                            dict(GAS, phi=phi, kappa=kappa)))
        built = S.model_problem(p, geom=g).coeffs
        assert set(built) == set(whole)
        for k in whole:
            assert built[k].shape == whole[k].shape and np.array_equal(built[k], whole[k]), k


def test_patchy_samples_equal_one_mask_at_a_time():
    """The library's many-model sampler (threads sharing the models) reads each model's mask at the
    given points exactly as painting that model's mask alone and indexing it does."""
    import ctypes

    from poe2d_amd import _abi
    from poe2d_amd.patchy import patchy_mask

    rng = np.random.default_rng(9)
    n, dx = 120, 0.5
    gas = np.concatenate([[0.0], rng.uniform(0.0, 0.3, 12)])
    radius = np.exp(rng.uniform(np.log(2.0), np.log(20.0), len(gas)))
    idx = np.ascontiguousarray(rng.integers(0, (n + 1) ** 2, 300), np.int64)
    out = np.full((len(gas), len(idx)), 7, np.uint8)
    lib = _abi.load_library()
    for threads in (1, 3, 0):
        _abi.check(lib, lib.poe2d_patchy_samples(
            n, n, dx, dx, ctypes.c_uint64(S.LAYOUT_SEED), len(gas), gas.ctypes.data_as(_abi.c_double_p),
            radius.ctypes.data_as(_abi.c_double_p), len(idx), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), threads), "patchy_samples")
        for k in range(len(gas)):
            want = np.zeros(len(idx), np.uint8) if gas[k] <= 0 else \
                patchy_mask(n, n, dx, dx, gas[k], S.LAYOUT_SEED, radius[k], radius[k]).ravel()[idx]
            assert np.array_equal(out[k], want), (threads, k)
    bad = idx.copy()
    bad[5] = (n + 1) ** 2
    with pytest.raises(_abi.Poe2dError):
        _abi.check(lib, lib.poe2d_patchy_samples(
            n, n, dx, dx, ctypes.c_uint64(1), 1, gas.ctypes.data_as(_abi.c_double_p),
            radius.ctypes.data_as(_abi.c_double_p), len(bad), bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), 1), "patchy_samples")
