"""Known answers for the oracle's single stages (CPU).

Each stage function of oracle/chain.cpp is called on its own through a probe and held to what the
reference's own tests or the mathematics fix:

- bicubic upsample sample: a constant table is reproduced, a linear ramp stays linear at a node,
  and the mirrored end samples give the closed-form edge values (reference tests/test_print.cpp:217-233);
- output encode: the five transfer functions at the reference's test values within 1 LSB of the
  closed forms, NaN -> 0, +inf -> 65535, -inf -> 0 without a transfer (test_print.cpp:270-319);
- film curve lookup: exact at the nodes, linear between them, clamped to the end plateaus, and the
  DIR couplers' pre-inverted curves reproduce the measured curve on a uniform field
  (test_film.cpp:83-102);
- highlight boost: 3 EV takes the frame maximum 10 to 80, values below the knee unchanged
  (test_film.cpp:319-344);
- paper curve lookup: exact at the nodes and clamped;
- CAM16-UCS: the oracle's forward and inverse against an independent numpy statement of the
  published equations (Li et al. 2017), in float64 and float32; the knee leaves values at or
  below its threshold unchanged, is continuous there and stays below its limit;
- the recursive Gaussian: its impulse response has variance sigma^2 (to the accuracy of the
  Young-van Vliet approximation) and unit sum, and is close to the exact Gaussian.
"""
import math

import numpy as np
import pytest

import spektralab_engine_amd as spk
from conftest import PLAIN, delta
from test_setup import cam16_consts, cam16_forward, cam16_inverse


# ---- bicubic upsample ---------------------------------------------------------------------------
def ramp_table(L=8):
    t = np.zeros((L, L, 3))
    for i in range(L):
        for j in range(L):
            t[i, j] = [0.25, i, 2.0 * j]
    return t


@pytest.mark.parametrize("f32", [False, True])
def test_mitchell_constant_and_ramp(oracle, f32):
    v = oracle.mitchell_sample(ramp_table(), 3.0 / 7.0, 4.0 / 7.0, f32)
    tol = 1e-6 if f32 else 1e-14
    np.testing.assert_allclose(v, [0.25, 3.0, 8.0], atol=tol)


@pytest.mark.parametrize("f32", [False, True])
def test_mitchell_edges_mirror_the_end_samples(oracle, f32):
    """At t = 0 the taps are rows 1, 0, 1 (weights 1/18, 16/18, 1/18); at t = 1 columns 6, 7, 6."""
    e = oracle.mitchell_sample(ramp_table(), 0.0, 1.0, f32)
    tol = 1e-6 if f32 else 1e-14
    np.testing.assert_allclose(e, [0.25, 2.0 / 18.0, (12.0 + 16.0 * 14.0 + 12.0) / 18.0], atol=tol)


def test_mitchell_weights_sum_to_one_between_nodes(oracle):
    """A constant table is reproduced at any (t0, t1), between the nodes too."""
    t = np.full((16, 16, 3), 0.7)
    rng = np.random.default_rng(3)
    for t0, t1 in rng.uniform(0, 1, (20, 2)):
        np.testing.assert_allclose(oracle.mitchell_sample(t, t0, t1, False), 0.7, atol=1e-14)


# ---- output encode ------------------------------------------------------------------------------
ENC_IN = [0.5, -0.25, 2.0, 0.25, 0.001, 1.0, 0.18, 0.0031308, 0.9]


def encode_closed_form(v, mode):
    if mode == 0:
        e = 12.92 * v if v <= 0.0031308 else 1.055 * math.copysign(abs(v) ** (1 / 2.4), v) - 0.055
    elif mode == 1:
        e = 16 * v if v < 1.0 / 512 else (v ** (1 / 1.8) if v >= 0 else float("nan"))
    elif mode == 2:
        e = v ** (256.0 / 563.0) if v >= 0 else float("nan")
    elif mode == 3:
        e = 4.5 * v if v < 0.018 else 1.099 * v ** 0.45 - 0.099
    else:
        e = v
    return 0.0 if math.isnan(e) else math.floor(min(max(e, 0.0), 1.0) * 65535 + 0.5)


@pytest.mark.parametrize("mode", range(5))
def test_encode_known_values(oracle, mode):
    for v in ENC_IN:
        got = oracle.encode16(np.float32(v), mode)
        want = encode_closed_form(float(np.float32(v)), mode)
        assert abs(got - want) <= 1, (mode, v, got, want)


def test_encode_non_finite_without_transfer(oracle):
    assert oracle.encode16(float("nan"), 4, active=False) == 0
    assert oracle.encode16(float("inf"), 4, active=False) == 65535
    assert oracle.encode16(float("-inf"), 4, active=False) == 0


# ---- film curves, DIR couplers ------------------------------------------------------------------
def film_curves(oracle, plan):
    K = int(oracle.field(plan, "K")[0])
    c = oracle.field(plan, "curves", 8 * K).astype(np.float32)
    return K, c[:K], c[K:2 * K], c[2 * K:5 * K].reshape(3, K), c[5 * K:].reshape(3, K)


def test_curve_lookup_nodes_midpoints_and_plateaus(oracle):
    x = np.array([-1.0, 0.0, 0.5, 2.0], np.float32)
    y = np.array([0.1, 0.4, 1.4, 2.0], np.float32)
    inv = np.append(np.float32(1.0) / np.diff(x), np.float32(0)).astype(np.float32)
    for xi, yi in zip(x, y):
        assert oracle.curve_lookup(x, inv, y, xi) == yi
    assert abs(oracle.curve_lookup(x, inv, y, 0.25) - 0.9) < 1e-6
    assert abs(oracle.curve_lookup(x, inv, y, 1.25) - 1.7) < 1e-6
    assert oracle.curve_lookup(x, inv, y, -50.0) == y[0]
    assert oracle.curve_lookup(x, inv, y, 50.0) == y[-1]


def test_coupler_curves_reproduce_the_measured_curve(resources, oracle):
    """Gamma 1, amount 1: on a uniform field D1 = F(L), chi = M^T D1 and F0(L - chi) = D1 to 5e-3."""
    plan = spk.host_plan(resources, delta(grain_active=False, halation_active=False, dir_couplers_active=True,
                                          dir_couplers_amount=1.0), 64, 64)
    assert oracle.field(plan, "couplers_active")[0] == 1
    K, x, inv, y, y0 = film_curves(oracle, plan)
    m = oracle.field(plan, "coupler_m", 9).astype(np.float32)
    worst = 0.0
    for L in np.linspace(-2.5, 3.5, 601).astype(np.float32):
        d1 = [oracle.curve_lookup(x, inv, y[c], L) for c in range(3)]
        for c in range(3):
            chi = (np.float32(d1[0]) * m[c] + np.float32(d1[1]) * m[3 + c]) + np.float32(d1[2]) * m[6 + c]
            d = oracle.curve_lookup(x, inv, y0[c], np.float32(L - chi))
            worst = max(worst, abs(d - d1[c]))
    assert worst < 5e-3, worst
    # without couplers the pre-inverted curves are the curves themselves
    plan0 = spk.host_plan(resources, delta(grain_active=False, halation_active=False, dir_couplers_active=True,
                                           dir_couplers_amount=0.0), 64, 64)
    _, _, _, ya, y0a = film_curves(oracle, plan0)
    assert np.array_equal(ya, y0a)


# ---- highlight boost ----------------------------------------------------------------------------
def test_boost_three_stops_takes_ten_to_eighty(oracle):
    v = [0.01 * i for i in range(300)] + [10.0]
    while len(v) % 3:
        v.append(1.0)
    v = np.array(v, np.float32)
    b = oracle.boost(v, 3.0)
    assert abs(b.max() - 80.0) < 80.0 * 1e-5
    low = v <= 2.944
    assert np.array_equal(b[low], v[low])
    assert np.all(np.diff(b[:300]) >= 0)  # monotone


# ---- paper curves -------------------------------------------------------------------------------
def test_paper_lookup_nodes_and_plateaus(resources, oracle):
    plan = spk.host_plan(resources, PLAIN, 64, 64)
    K = int(oracle.field(plan, "paper_k")[0])
    p = oracle.field(plan, "paper", 5 * K).astype(np.float32)
    x, y = p[:K], p[2 * K:].reshape(K, 3)
    for i in (0, 7, K // 2, K - 2):
        for c in range(3):
            assert oracle.paper_lookup(plan, c, x[i]) == y[i, c]
    for c in range(3):
        assert oracle.paper_lookup(plan, c, x[0] - 5.0) == y[0, c]
        assert oracle.paper_lookup(plan, c, x[-1] + 5.0) == y[-1, c]
        mid = np.float32((x[10] + x[11]) / 2)
        assert abs(oracle.paper_lookup(plan, c, mid) - (y[10, c] + y[11, c]) / 2) < 1e-5


# ---- CAM16-UCS ----------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def plan1(resources):
    return spk.host_plan(resources, PLAIN, 64, 64)


def test_cam16_forward_matches_the_published_equations(oracle, plan1):
    q = cam16_consts(oracle.field(plan1, "cam64", 22))
    rng = np.random.default_rng(11)
    for xyz in rng.uniform(0.02, 0.95, (40, 3)):
        want = np.array(cam16_forward(q, xyz))
        np.testing.assert_allclose(oracle.cam16_forward(plan1, xyz), want, rtol=1e-12, atol=1e-12)
        got32 = oracle.cam16_forward(plan1, xyz, f32=True)
        assert abs(got32[0] - want[0]) < 2e-4 * max(1.0, want[0]) and abs(got32[1] - want[1]) < 2e-3
        assert abs(math.remainder(got32[2] - want[2], 2 * math.pi)) < 1e-4


def test_cam16_inverse_matches_the_published_equations(oracle, plan1):
    """The oracle solves for a, b by a different route and applies the published 8-digit M16
    inverse (relative accuracy ~1e-9); the numpy statement uses the exact inverse."""
    q = cam16_consts(oracle.field(plan1, "cam64", 22))
    for Jp, Mp, h in [(40.0, 10.0, 0.3), (70.0, 25.0, -2.0), (20.0, 5.0, 2.9), (90.0, 2.0, -0.7)]:
        np.testing.assert_allclose(oracle.cam16_inverse(plan1, Jp, Mp, h), cam16_inverse(q, Jp, Mp, h), rtol=1e-8,
                                   atol=1e-12)


def test_knee_identity_below_threshold_and_bounded(oracle):
    for d in (0.0, 0.3, 0.7):
        assert oracle.knee(d, 0.7, 1.0, 2.2) == d
    assert abs(oracle.knee(0.7 + 1e-9, 0.7, 1.0, 2.2) - 0.7) < 1e-8  # continuous at the threshold
    vals = [oracle.knee(d, 0.7, 1.0, 2.2) for d in np.linspace(0.7, 50.0, 200)]
    assert all(v < 1.0 for v in vals) and np.all(np.diff(vals) > 0)
    # power 1: the knee is x / (1 + x) of the excess
    assert abs(oracle.knee(1.7, 0.7, 1.0, 1.0) - (0.7 + 0.3 * ((1.0 / 0.3) / (1.0 + 1.0 / 0.3)))) < 1e-14


# ---- recursive Gaussian ---------------------------------------------------------------------------
def yvv_published(sigma):
    """Young and van Vliet (Signal Processing 44, 1995): q(sigma), b0..b3, normalised to
    (B, b1/b0, b2/b0, b3/b0) with B = 1 - (b1 + b2 + b3) / b0."""
    q = 0.98711 * sigma - 0.96330 if sigma >= 2.5 else 3.97156 - 4.14554 * math.sqrt(1.0 - 0.26891 * sigma)
    b0 = 1.57825 + 2.44413 * q + 1.4281 * q * q + 0.422205 * q ** 3
    b1 = 2.44413 * q + 2.85619 * q * q + 1.26661 * q ** 3
    b2 = -(1.4281 * q * q + 1.26661 * q ** 3)
    b3 = 0.422205 * q ** 3
    return 1.0 - (b1 + b2 + b3) / b0, b1 / b0, b2 / b0, b3 / b0


@pytest.mark.parametrize("sigma", [0.8, 2.0, 2.6, 4.0, 8.0, 32.0])
def test_recursive_coefficients_are_the_published_ones(oracle, sigma):
    np.testing.assert_allclose(oracle.yvv(sigma), yvv_published(sigma), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("sigma", [4.0, 8.0, 20.0])
def test_recursive_gaussian_impulse_moments(oracle, sigma):
    """An impulse through the recursive pass (sigma above the crossover): unit gain; variance twice
    that of the causal recursion's own impulse response (computed here from the published
    coefficients), i.e. forward and backward passes of one filter; and within 4 % of the exact
    Gaussian's peak everywhere (the 1995 recursion is broader than G_sigma by up to ~23 % in
    variance over this range, a property of the published coefficients)."""
    n = 20 * int(sigma) + 1
    img = np.zeros((1, n, 3), np.float32)
    img[0, n // 2, :] = 1.0
    out = oracle.gaussian(img, (sigma, sigma, sigma), crossover=1.0)[0, :, 0].astype(np.float64)
    x = np.arange(n) - n // 2
    assert abs(out.sum() - 1.0) < 1e-5
    B, b1, b2, b3 = yvv_published(sigma)
    h, w = np.zeros(8000), [0.0, 0.0, 0.0]
    for i in range(h.size):
        v = B * (1.0 if i == 0 else 0.0) + b1 * w[0] + b2 * w[1] + b3 * w[2]
        w = [v, w[0], w[1]]
        h[i] = v
    k = np.arange(h.size)
    var_causal = (h * k * k).sum() - (h * k).sum() ** 2
    assert abs((out * x * x).sum() / (2.0 * var_causal) - 1.0) < 2e-3
    exact = np.exp(-0.5 * (x / sigma) ** 2) / (sigma * math.sqrt(2 * math.pi))
    assert np.abs(out - exact).max() < 0.04 * exact.max()


# ---- print-preview table --------------------------------------------------------------------------
def test_print_lut_trilinear_reproduces_an_affine_table(oracle):
    """A table affine in the three densities is reproduced by trilinear lookup at any density inside
    the axes (to the 16-bit encode's rounding), and densities outside clamp to the table's faces;
    alpha is opaque (reference docs/RESOURCE_FORMAT.md: per-channel axes, [i][j][k][ch] order)."""
    S = 9
    axes = np.array([np.linspace(0.0, 3.0, S), np.linspace(-0.5, 2.5, S), np.linspace(0.2, 4.2, S)], np.float32)
    g = np.stack(np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij"), -1) / (S - 1.0)
    coef = np.array([[0.5, 0.2, 0.1], [0.1, 0.6, 0.05], [0.05, 0.1, 0.7]])
    table = (0.05 + g @ coef.T).astype(np.float32)  # table[i, j, k] = 0.05 + coef . (i, j, k)/(S-1)
    rng = np.random.default_rng(2)
    neg = np.stack([rng.uniform(axes[c, 0], axes[c, -1], (7, 11)) for c in range(3)]).astype(np.float32)
    out = oracle.print_lut(neg, table, axes)
    u = np.stack([(neg[c] - axes[c, 0]) / (axes[c, -1] - axes[c, 0]) for c in range(3)], -1)
    want = np.floor(np.clip(0.05 + u @ coef.T, 0, 1) * 65535 + 0.5)
    assert np.abs(out[..., :3].astype(np.int64) - want).max() <= 1
    assert np.all(out[..., 3] == 65535)
    far = np.stack([np.full((1, 2), v, np.float32) for v in (-9.0, 9.0, 9.0)])
    o = oracle.print_lut(far, table, axes)
    np.testing.assert_array_equal(o[0, 0, :3], np.floor(np.clip(table[0, -1, -1], 0, 1) * 65535 + 0.5))


# ---- halation mixture: list-order accumulation of the pinned Gaussians --------------------------
@pytest.mark.parametrize("comps,crossover", [
    ([(2.0, 0.6), (8.0, 0.3), (32.0, 0.1)], 16.5),
    ([((2.0, 3.0, 4.0), (0.5, 0.6, 0.7)), (8.0, 0.3)], 16.5),
    ([(21.0, 0.5), (3.0, 0.5)], 30.0),
    ([(0.3, 0.4), (1.4, 0.6)], 3.0),
])
def test_mixture_is_the_list_order_sum_of_gaussians(oracle, comps, crossover):
    """oracle.mixture = ((0 + w0 G0) + w1 G1) + ... in binary32, each G the oracle's per-channel
    Gaussian (gaussian.cpp:254-273: the first component ZERO-mode, the rest accumulated)."""
    rng = np.random.default_rng(5)
    img = rng.uniform(0.0, 2.0, (37, 53, 3)).astype(np.float32)
    got = oracle.mixture(img, comps, crossover)
    acc = None
    for s, w in comps:
        g = oracle.gaussian(img, np.broadcast_to(s, 3), crossover)
        term = np.broadcast_to(np.float32(w), (3,)).astype(np.float32) * g
        acc = np.float32(0.0) + term if acc is None else acc + term
    assert acc.dtype == np.float32
    np.testing.assert_array_equal(got, acc)
