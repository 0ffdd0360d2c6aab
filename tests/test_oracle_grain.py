"""The oracle's grain models beyond the sub-layer sum, pinned on the CPU by what the mathematics
fixes: the film side's cosine against libm, the micro-structure lognormal against its closed form
and its moments, the PLANES draw with unblurred dye planes and the micro-structure with a unit
field reducing exactly to the SUM draw, and the single-layer draw's mean equal to the density it
is drawn at (src/fx/grain.cpp:23-152 and :155-267, kernels/fx_grain.metal:39-116,
kernels/fx_random.h:91-115)."""
import math

import numpy as np
import pytest

import spektralab_engine_amd as spk
from conftest import delta
from spektralab_engine_amd import synth

SEED = 0x4ABEA221


def test_cosine_matches_libm(oracle):
    xs = np.concatenate([np.linspace(-7.0, 7.0, 20001), np.arange(-4, 5) * (math.pi / 2)])
    xs = np.concatenate([xs, np.nextafter(xs, np.inf), np.nextafter(xs, -np.inf)])
    err = max(abs(oracle.cos(float(x)) - math.cos(float(x))) for x in xs)
    assert err <= 2.3e-16, err
    assert oracle.cos(0.0) == 1.0


def test_lognormal_closed_form(oracle):
    """exp(ln m - sig^2/2 + sig sqrt(-2 ln u1) cos(2 pi u2)), sig^2 = ln(1 + s^2/m^2), evaluated in
    binary64 on the same binary32 inputs: within a few binary32 roundings."""
    rng = np.random.default_rng(5)
    for m, s in [(1.0, 0.06144), (1.0, 0.3), (2.5, 0.7)]:
        u = ((rng.integers(0, 1 << 24, (500, 2)) + 0.5) * 2.0 ** -24).astype(np.float32)
        for u1, u2 in u:
            got = oracle.lognormal(m, s, float(u1), float(u2))
            sig = math.sqrt(math.log(1.0 + (s * s) / (m * m)))
            two_pi_u2 = float(np.float32(np.float32(2 * math.pi) * u2))  # the binary32 argument
            z = math.sqrt(-2.0 * math.log(float(u1))) * math.cos(two_pi_u2)
            want = math.exp(math.log(m) - 0.5 * sig * sig + sig * z)
            assert abs(got - want) <= 2e-6 * want, (m, s, u1, u2, got, want)
    # u1 = 1: z = 0, the median exp(mu)
    assert oracle.lognormal(1.0, 0.3, 1.0, 0.3) == pytest.approx(math.exp(-0.5 * math.log(1.09)), rel=2e-7)


def test_micro_field_moments(oracle):
    """Mean 1, standard deviation sigma, log normal with variance ln(1 + sigma^2), and the three
    channels independent (streams 100, 101, 102)."""
    n, sigma = 300_000, 0.2
    f = oracle.micro_field(n, sigma, SEED).astype(np.float64)
    assert np.all(f > 0)
    assert np.abs(f.mean(axis=1) - 1.0).max() <= 5 * sigma / math.sqrt(n)
    assert np.abs(f.std(axis=1) / sigma - 1.0).max() <= 0.02
    lg = np.log(f)
    var = math.log(1.0 + sigma * sigma)
    assert np.abs(lg.mean(axis=1) + 0.5 * var).max() <= 5 * math.sqrt(var / n)
    assert np.abs(lg.var(axis=1) / var - 1.0).max() <= 0.02
    c = np.corrcoef(f)
    assert np.abs(c[np.triu_indices(3, 1)]).max() <= 5 / math.sqrt(n)


@pytest.fixture(scope="module")
def sum_case(resources):
    """A sub-layer SUM plan (8 mm film, 512 x 384: 15.6 um pitch, no dye clouds or micro-structure)."""
    d = delta(grain_active=True, halation_active=False, film_format_mm=8.0)
    plan = spk.host_plan(resources, d, 512, 384, SEED)
    img = synth.make_frame(512, 384, 31)
    return plan, img


def edited(plan, **fields):
    import ctypes
    cp = ctypes.create_string_buffer(bytes(plan), len(plan))
    from oracle import oracle as o
    for k, v in fields.items():
        o.set_field(cp, k, v)
    return cp


def test_planes_with_unblurred_dye_equal_the_sum(oracle, sum_case):
    """The PLANES draw with every dye sigma 0 sums the same three values in the same order and shifts
    once afterwards: bit for bit the SUM draw."""
    plan, img = sum_case
    assert int(oracle.field(plan, "grain_mode")[0]) == 0 and int(oracle.field(plan, "grain_shift")[0]) == 1
    ref = oracle.develop(plan, img)
    p = edited(plan, grain_mode=1, grain_dye=1, grain_shift=0, grain_dye_sigma=np.zeros(9))
    assert np.array_equal(oracle.develop(p, img), ref)


def test_unit_micro_field_equals_the_sum(oracle, sum_case):
    """sigma 0 makes every lognormal value exactly 1: G 1 + shift is the SUM draw's G + shift."""
    plan, img = sum_case
    assert oracle.lognormal(1.0, 0.0, 0.37, 0.81) == 1.0
    ref = oracle.develop(plan, img)
    p = edited(plan, grain_micro=1, grain_micro_sigma=0.0, grain_micro_blur=0.0, grain_shift=0)
    assert np.array_equal(oracle.develop(p, img), ref)


def test_dye_blur_moves_the_picture_not_its_mean(oracle, sum_case):
    """With dye sigmas the picture changes, its mean stays (the blurs are normalised)."""
    plan, img = sum_case
    ref = oracle.develop(plan, img)
    p = edited(plan, grain_mode=1, grain_dye=1, grain_shift=0, grain_dye_sigma=np.full(9, 1.5))
    got = oracle.develop(p, img)
    assert not np.array_equal(got, ref)
    assert np.abs(got.mean(axis=(1, 2)) - ref.mean(axis=(1, 2))).max() <= 2e-3
    assert got.std() < ref.std()  # the clouds smooth the grain


@pytest.mark.parametrize("fmt,n_min", [(35.0, 5000.0), (4.0, 50.0)])
def test_single_layer_draw_is_unbiased(oracle, resources, fmt, n_min):
    """E[k (D/n) sat] = lam (D/n) sat = q D = d: the single-layer draw's expectation is the density
    it is drawn at, so the grain negative's mean is the grain-free negative's."""
    d = dict(grain_sublayers_active=False, film_format_mm=fmt)
    w, h = 512, 384
    img = synth.make_frame(w, h, 32)
    plan = spk.host_plan(resources, delta(grain_active=True, halation_active=False, **d), w, h, SEED)
    assert int(oracle.field(plan, "grain_mode")[0]) == 2
    assert oracle.field(plan, "count", 3).min() >= n_min
    assert not oracle.field(plan, "count", 9)[3:].any()
    neg, lam, cnt = oracle.develop_draws(plan, img)
    assert not cnt.reshape(3, 3, h, w)[:, 1:].any() and not lam.reshape(3, 3, h, w)[:, 1:].any()
    plain = oracle.develop(spk.host_plan(resources, delta(grain_active=False, halation_active=False, **d), w, h, SEED),
                           img)
    n = oracle.field(plan, "count", 3)
    dmax = oracle.field(plan, "sat_d", 3)
    tol = 5 * dmax / np.sqrt(n * w * h)  # 5 standard errors of the mean of the draws
    assert np.all(np.abs(neg.mean(axis=(1, 2)) - plain.mean(axis=(1, 2))) <= tol)
