"""The scanned-film path (scan_film): the negative is scanned directly, with no enlarger and no paper
(reference src/print/run.cpp:100-188, src/print/build.cpp:133-150, src/print/setup.cpp:267-270).

Bars: the negative is the printed path's (the film side does not change); RGBA16 within +-1 LSB of
the binary32 oracle's scan of that negative; with FP16 spectral products, within the precision
mode's bounds of the float64 oracle.  Glare draws its print seed but does not act; EDR is off.
"""
import numpy as np
import pytest

from conftest import PLAIN, GRAIN_HALATION, MIX3_OVERRIDES, delta
from spektralab_engine_amd import colour, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,w,h,seed", [("plain", 512, 384, 61), ("grain_halation", 640, 427, 62)])
def test_scan_film_matches_oracle(engine, oracle, cfg, w, h, seed):
    img = synth.make_frame(w, h, seed)
    base = PLAIN if cfg == "plain" else GRAIN_HALATION
    with engine.open(img, dict(base, scan_film=True)) as s:
        if cfg == "grain_halation":
            s.set_overrides(MIX3_OVERRIDES)
        out, _ = s.render("full")
        plan = s.last_plan()
        neg = s.negative("full")
    assert int(oracle.field(plan, "printed")[0]) == 0
    neg_o = oracle.develop(plan, img)
    if cfg == "plain":
        assert np.array_equal(neg, neg_o)
    out_o = oracle.print_(plan, neg)
    d = np.abs(out.astype(np.int32) - out_o.astype(np.int32))
    assert d.max() <= 1, f"{d.max()} LSB"
    assert np.all(out[..., 3] == 65535)


def test_scan_film_negative_is_the_printed_one(engine):
    """scan_film is a print-layer choice: the cached negative stays, the picture changes."""
    img = synth.make_frame(300, 200, 63)
    with engine.open(img, PLAIN) as s:
        printed, _ = s.render("full")
        neg = s.negative("full")
        r = s.set_params({"scan_film": True})
        scanned, f = s.render("full", reprint=True)
        assert np.array_equal(s.negative("full"), neg)
    assert r["invalidated"] in ("print", "shoot")
    assert not np.array_equal(printed, scanned)


def test_scan_film_glare_draws_but_does_not_act(engine, oracle):
    """Under scan_film the print seed is still drawn when glare could draw (API_SPEC 2.8), and the
    picture is the glare-free scan."""
    img = synth.make_frame(256, 160, 64)
    with engine.open(img, dict(PLAIN, scan_film=True, glare_active=True)) as s:
        a, _ = s.render("full")
        plan = s.last_plan()
    with engine.open(img, dict(PLAIN, scan_film=True, glare_active=False)) as s:
        b, _ = s.render("full")
    assert int(oracle.field(plan, "glare_active")[0]) == 0
    assert np.array_equal(a, b)


def test_scan_film_fp16_against_float64(engine, oracle):
    img = synth.make_frame(512, 384, 65)
    with engine.open(img, dict(PLAIN, scan_film=True)) as s:
        o32, _ = s.render("full")
        plan = s.last_plan()
        neg = s.negative("full")
        s.set_overrides({"spectral_precision": "fp16"})
        o16, _ = s.render("full", reprint=True)
    ref64 = oracle.print_(plan, neg, f64=True)
    e32, e16 = colour.lsb_stats(o32, ref64), colour.lsb_stats(o16, ref64)
    de = colour.delta_e_ucs(o16, ref64)
    print(f"scan_film fp32 {e32} fp16 {e16} dE p99 {np.percentile(de, 99):.3g}")
    assert e32["max"] <= 2
    assert e16["max"] <= 64 and e16["p99"] <= 8 and np.percentile(de, 99) <= 0.05
