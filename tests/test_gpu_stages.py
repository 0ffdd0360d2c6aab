"""GPU parity of the stages around the benchmark path: tiers (downscale), auto exposure, DIR
couplers, glare, and a render with the reference's own default parameters.

Bars: the downscale is bit-exact (reference test_session.cpp:58-88 holds its GPU and CPU copies to
equality); the metered EV equals an independent float64 statement of meter.cpp; negatives are
identical except where a recursive blur rounds 1 ulp apart; RGBA16 within +-1 LSB of the binary32
oracle on the engine's own negative.
"""
import math

import numpy as np
import pytest

from conftest import CONFIG1, delta
from spektralab_engine_amd import synth

pytestmark = pytest.mark.gpu


def downscale_size(w, h, L):
    m = max(w, h)
    if L == 0 or m <= L:
        return w, h
    s = L / m
    return max(1, math.floor(w * s)), max(1, math.floor(h * s))


@pytest.mark.parametrize("shape,L", [((2000, 3000), 1600), ((300, 5000), 1600), ((1437, 2156), 1600),
                                     ((4000, 6000), 3400), ((768, 1024), 3400), ((7, 4001), 800)])
def test_downscale_bit_exact(engine, oracle, shape, L):
    rng = np.random.default_rng(L + shape[0])
    img = (rng.uniform(0.0, 4.0, shape + (3,)) ** 2).astype(np.float32)
    gpu = engine.downscale(img, L)
    cpu = oracle.downscale(img, L)
    w, h = downscale_size(shape[1], shape[0], L)
    assert gpu.shape == cpu.shape == (h, w, 3)
    assert np.array_equal(gpu, cpu)


def test_live_and_preview_tiers(engine, oracle):
    img = synth.make_frame(2400, 1600, 7)
    with engine.open(img, CONFIG1) as s:
        out, flags = s.render("live")
        assert (flags["width"], flags["height"]) == downscale_size(2400, 1600, 1600)
        plan = s.last_plan()
        neg = s.negative("live")
        pv, fp = s.render("preview")  # 2400 <= 3400: the frame itself
        assert (fp["width"], fp["height"]) == (2400, 1600)
    src = oracle.downscale(img, 1600)
    assert abs(oracle.field(plan, "rho_tier")[0] - max(src.shape[:2]) / 2400) < 1e-15
    assert abs(oracle.field(plan, "pitch_um")[0] - 35000.0 / max(src.shape[:2])) < 1e-12
    neg_o = oracle.develop(plan, src)
    assert np.array_equal(neg, neg_o)
    d = np.abs(out.astype(np.int32) - oracle.print_(plan, neg).astype(np.int32))
    assert d.max() <= 1


def prophoto_rgb_to_xyz():
    """ProPhoto RGB (ROMM) primaries and D50 white, by the standard primaries construction."""
    prim = np.array([[0.7347, 0.2653], [0.1596, 0.8404], [0.0366, 0.0001]])
    wxy = np.array([0.3457, 0.3585])
    P = np.array([[x / y, 1.0, (1 - x - y) / y] for x, y in prim]).T
    W = np.array([wxy[0] / wxy[1], 1.0, (1 - wxy[0] - wxy[1]) / wxy[1]])
    return P * np.linalg.solve(P, W)


def center_weighted_ev(rgb):
    """meter.cpp center_weighted, stated independently: -log2(weighted mean Y / 0.184)."""
    h, w, _ = rgb.shape
    m = prophoto_rgb_to_xyz()[1]
    Y = (m[0] * rgb[..., 0].astype(np.float64) + m[1] * rgb[..., 1]) + m[2] * rgb[..., 2]
    M = max(w, h)
    x = (np.arange(w) / w - 0.5) * (w / M)
    y = (np.arange(h) / h - 0.5) * (h / M)
    wgt = np.exp(-(x[None, :] ** 2 + y[:, None] ** 2) / (2 * 0.2 ** 2))
    return -math.log2((wgt * Y).sum() / wgt.sum() / 0.184)


def test_auto_exposure(engine, oracle):
    img = synth.make_frame(2000, 1200, 8)
    with engine.open(img, delta(grain_active=False, halation_active=False, auto_exposure=True)) as s:
        out, _ = s.render("full")
        plan = s.last_plan()
        neg = s.negative("full")
    assert oracle.field(plan, "ae_active")[0] == 1
    ev = oracle.field(plan, "ae_ev")[0]
    want = center_weighted_ev(oracle.downscale(img, 1600))
    assert abs(ev - want) < 1e-9, (ev, want)
    assert oracle.field(plan, "ae_gain")[0] == np.float32(2.0 ** ev)
    assert np.array_equal(neg, oracle.develop(plan, img))
    d = np.abs(out.astype(np.int32) - oracle.print_(plan, neg).astype(np.int32))
    assert d.max() <= 1


def compare(engine, oracle, dl, w, h, seed, mism_frac):
    img = synth.make_frame(w, h, seed)
    with engine.open(img, dl) as s:
        out, _ = s.render("full")
        plan = s.last_plan()
        neg = s.negative("full")
    neg_o = oracle.develop(plan, img)
    diff = np.abs(neg.astype(np.float64) - neg_o)
    mism = int(np.count_nonzero(diff))
    print(f"{mism}/{neg.size} negative values differ, max {diff.max():.3g}")
    assert mism <= neg.size * mism_frac
    d = np.abs(out.astype(np.int32) - oracle.print_(plan, neg).astype(np.int32))
    assert d.max() <= 1, f"max {d.max()} LSB"
    return plan


def test_dir_couplers(engine, oracle):
    plan = compare(engine, oracle, delta(grain_active=False, halation_active=False, dir_couplers_active=True),
                   600, 400, 9, 1e-3)
    assert oracle.field(plan, "couplers_active")[0] == 1


def test_glare(engine, oracle):
    plan = compare(engine, oracle, delta(grain_active=False, halation_active=False, glare_active=True), 500, 330, 10, 0)
    assert oracle.field(plan, "glare_active")[0] == 1
    # grain off: no film seed is drawn, the print seed is the sequence's first value
    assert oracle.field(plan, "glare_seed")[0] == 0x4ABEA221


def test_reference_defaults(engine, oracle):
    """Every parameter at the reference default (auto exposure, scatter + bounce halation, grain,
    DIR couplers, glare), on the synthetic stocks."""
    plan = compare(engine, oracle, dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK), 640, 480, 11, 2e-3)
    for f in ("ae_active", "couplers_active", "glare_active", "grain_active", "scatter_active", "bounce_active"):
        assert oracle.field(plan, f)[0] == 1, f
