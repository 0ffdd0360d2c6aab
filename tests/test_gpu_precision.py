"""Precision mode: the print's two 81-bin spectral integrals as FP16 matrix-core products with FP32
accumulation (override spectral_precision = "fp16").

The FP32 path is held to +-1 LSB of the binary32 oracle elsewhere; here both paths are compared
with the float64 oracle on the same negative, and the FP16 error is bounded in 16-bit LSB and in
CAM16-UCS colour difference.  The negative must not change when only the precision changes (it is
a print-side choice).
"""
import numpy as np
import pytest

from conftest import PLAIN, GRAIN_HALATION, MIX3_OVERRIDES
from spektralab_engine_amd import colour, synth

pytestmark = pytest.mark.gpu


def errors(out, ref):
    s = colour.lsb_stats(out, ref)
    de = colour.delta_e_ucs(out, ref)
    s.update(de_max=float(de.max()), de_p99=float(np.percentile(de, 99)), de_mean=float(de.mean()))
    return s


@pytest.mark.parametrize("cfg,w,h,seed", [("plain", 512, 512, 0), ("grain_halation", 640, 427, 5)])
def test_fp16_spectral_against_float64(engine, oracle, cfg, w, h, seed):
    img = synth.make_frame(w, h, seed)
    dl = PLAIN if cfg == "plain" else GRAIN_HALATION
    with engine.open(img, dl) as s:
        if cfg == "grain_halation":
            s.set_overrides(MIX3_OVERRIDES)
        o32, _ = s.render("full")
        plan32 = s.last_plan()
        neg = s.negative("full")
        s.set_overrides({"spectral_precision": "fp16"})
        o16, _ = s.render("full", reprint=True)
        plan16 = s.last_plan()
        assert np.array_equal(s.negative("full"), neg)
        s.set_overrides({"spectral_precision": "fp32"})
        back, _ = s.render("full", reprint=True)
    assert oracle.field(plan16, "spectral_fp16")[0] == 1 and oracle.field(plan32, "spectral_fp16")[0] == 0
    assert np.array_equal(back, o32)
    ref64 = oracle.print_(plan32, neg, f64=True)
    e32, e16 = errors(o32, ref64), errors(o16, ref64)
    print(f"{cfg}: fp32 vs f64 {e32}\n{cfg}: fp16 vs f64 {e16}")
    assert e32["max"] <= 2
    assert e16["max"] <= 64 and e16["p99"] <= 8 and e16["de_p99"] <= 0.05, e16


def test_fp16_matches_fp32_on_a_flat_field(engine):
    """A uniform frame prints to a uniform image in both precisions (the matrix steps mix 64 pixels
    per wavefront: a lane-layout slip would show as spatial structure), and both agree."""
    img = np.full((70, 130, 3), 0.18, np.float32)
    with engine.open(img, PLAIN) as s:
        s.set_overrides({"spectral_precision": "fp16"})
        o16, _ = s.render("full")
        s.set_overrides({"spectral_precision": "fp32"})
        o32, _ = s.render("full", reprint=True)
    for o in (o16, o32):
        assert np.all(o == o[0, 0])
    assert np.abs(o16.astype(np.int64) - o32.astype(np.int64)).max() <= 16


def test_fp16_pixel_order(engine, oracle):
    """Every pixel of a ragged frame (not a multiple of the 256-pixel chunk) lands where the FP32 path
    puts it: a ramp with a distinct value per pixel."""
    w, h = 97, 61
    ramp = np.exp2(np.linspace(-6.0, 3.0, w * h)).reshape(h, w)
    img = np.stack([ramp, ramp[::-1, ::-1], np.roll(ramp, 7, axis=1)], axis=-1).astype(np.float32)
    with engine.open(img, PLAIN) as s:
        o32, _ = s.render("full")
        s.set_overrides({"spectral_precision": "fp16"})
        o16, _ = s.render("full", reprint=True)
    d = np.abs(o16.astype(np.int64) - o32.astype(np.int64))[..., :3]
    assert d.max() <= 64, d.max()
