"""The oracle against the reference's own known-answer tests and closed forms (no GPU).

Known answers are those of the reference's tests: Philox and splitmix64 (tests/test_util.cpp:159-176,
tests/test_fx.cpp:153-154), the uniform map's end points (test_util.cpp:171-175), FIR radii, the
Young-van Vliet normalisation and the reflect folds (test_fx.cpp:140-149, test_util.cpp:152-155).
"""
import math

import numpy as np
import pytest


def test_philox_known_answers(oracle):
    out = oracle.philox([[0, 0, 0, 0], [0, 200, 0, 0]], [[0, 0], [0x5EED, 0]])
    assert [hex(v) for v in out[0]] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(v) for v in out[1]] == ["0x15cecb3f", "0xf63dfa39", "0x3ec2dc0e", "0xf78f6c18"]


def splitmix_draws(n):
    state, mask = 0x243F6A8885A308D3, (1 << 64) - 1
    out = []
    for _ in range(n):
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        out.append((z ^ (z >> 31)) & 0xFFFFFFFF)
    return out


def test_splitmix_known_answers():
    assert splitmix_draws(4) == [0x4ABEA221, 0x23148989, 0x609DFE03, 0xDEB12800]


def test_uniform_map_end_points(oracle):
    lib = oracle.lib()
    assert lib.oracle_uniform(0) == 2.0 ** -25
    assert lib.oracle_uniform(0xFFFFFEFF) < 1.0
    assert lib.oracle_uniform(0xFFFFFFFF) == 1.0


def test_fir_radius_and_taps(oracle):
    lib = oracle.lib()
    assert [lib.oracle_fir_radius(s) for s in (2.9, 0.65, 0.5, 1e-6, 0.0)] == [9, 2, 2, 0, 0]
    for s in (0.5, 0.65, 2.0, 8.0, 16.0):
        t = oracle.fir_taps(s)
        r = len(t) // 2
        assert abs(t.sum() - 1.0) < 1e-15
        k = np.arange(-r, r + 1)
        ref = np.exp(-0.5 * (k / s) ** 2)
        np.testing.assert_allclose(t, ref / ref.sum(), rtol=1e-14)


def test_yvv_coefficients_normalised(oracle):
    for s in (3.0, 4.6, 13.9, 127.6, 261.0):
        B, B1, B2, B3 = oracle.yvv(s)
        assert abs(B + B1 + B2 + B3 - 1.0) < 1e-14


def test_reflect_folds(oracle):
    lib = oracle.lib()
    assert lib.oracle_reflect(-1, 4) == 0 and lib.oracle_reflect(4, 4) == 3
    assert lib.oracle_reflect(-5, 4) == 3 and lib.oracle_reflect(9, 4) == 1
    assert lib.oracle_reflect(5, 1) == 0


def test_sampler_log_is_faithful(oracle):
    """The grain sampler's table-driven log: within 1 ulp everywhere, and equal to the correctly
    rounded result except on the rare inputs within 2^-50 of a rounding midpoint."""
    fn = oracle.lib().oracle_logf
    rng = np.random.default_rng(5)
    xs = np.concatenate([
        (rng.uniform(1.0, 2.0, 200000) * 2.0 ** rng.integers(-149, 128, 200000)),
        1.0 + rng.uniform(-1e-3, 1e-3, 50000),  # near 1, where e ln2 + logc cancels
        rng.uniform(0.0, 1.0, 50000), rng.uniform(10.0, 700.0, 50000)]).astype(np.float32)
    xs = xs[(xs > 0) & np.isfinite(xs)]
    got = np.array([fn(float(x)) for x in xs], np.float32)
    want = np.array([math.log(float(x)) for x in xs]).astype(np.float32)
    ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32))
    assert ulp.max() <= 1 and np.mean(ulp != 0) < 1e-4, (ulp.max(), np.mean(ulp != 0))
    assert fn(1.0) == 0.0 and fn(0.0) == -math.inf and math.isnan(fn(-1.0)) and math.isnan(fn(math.nan))
    assert fn(math.inf) == math.inf and fn(1e-45) == np.float32(math.log(1.401298464324817e-45))


@pytest.mark.parametrize("name,ref", [("oracle_expf", math.exp), ("oracle_log10f", math.log10)])
def test_shared_math_correctly_rounded(oracle, name, ref):
    """The reproducible float transcendentals round like the correctly rounded result."""
    fn = getattr(oracle.lib(), name)
    rng = np.random.default_rng(3)
    if name == "oracle_expf":
        xs = rng.uniform(-87.0, 88.0, 4000).astype(np.float32)
    else:
        xs = (rng.uniform(1.0, 2.0, 4000) * 2.0 ** rng.integers(-100, 100, 4000)).astype(np.float32)
    bad = 0
    for x in xs:
        want = np.float32(ref(float(x)))
        bad += fn(float(x)) != want
    assert bad == 0


def test_poisson_statistics(oracle):
    """Mean and variance of the sampler match lambda (reference test_fx.cpp:546-566 bounds)."""
    n = 20000
    for lam in (0.5, 3.0, 9.5, 10.0, 37.0, 400.0):
        draws = oracle.poisson(np.full(n, lam, np.float32), stream=3, seed=1234)
        assert np.all(draws == np.floor(draws)) and draws.min() >= 0
        m, v = draws.mean(), draws.var()
        assert abs(m - lam) < 5 * math.sqrt(lam / n), (lam, m)
        assert abs(v / lam - 1.0) < 0.06, (lam, v)


def test_poisson_nan_draws_zero(oracle):
    assert oracle.poisson(np.array([np.nan, 0.0], np.float32), 0, 1).tolist() == [0.0, 0.0]


def test_transposed_recursion_is_the_adjoint(oracle):
    """<G x, y> = <x, G^T y> for the oracle's transposed Gaussian (recursive and FIR channels, lines
    of 1 to 300 samples): the identity fixes every edge term of the transposed recursion; a wrong
    weight or index at either end moves the two sides apart by far more than binary32 rounding."""
    rng = np.random.default_rng(0)
    for w, h, sig in [(37, 29, (5.0, 8.0, 32.0)), (64, 300, (3.0, 3.0, 3.0)), (7, 5, (4.0, 40.0, 3.5)),
                      (2, 2, (10.0, 10.0, 10.0)), (1, 1, (5.0, 5.0, 5.0)), (19, 3, (2.0, 1.0, 6.0))]:
        x = rng.normal(size=(h, w, 3)).astype(np.float32)
        y = rng.normal(size=(h, w, 3)).astype(np.float32)
        a = np.sum(oracle.gaussian(x, sig, 3.0).astype(np.float64) * y)
        b = np.sum(x.astype(np.float64) * oracle.gaussian_adjoint(y, sig, 3.0))
        scale = np.sum(np.abs(x)) * np.abs(y).max()
        assert abs(a - b) <= 1e-6 * scale, (w, h, sig, a, b)
