"""GPU parity of the stages around the benchmark path: tiers (downscale), input decode, geometry,
DIR couplers, glare, and a render with the reference's own default parameters.

Bars: the downscale is bit-exact (reference test_session.cpp:58-88 holds its GPU and CPU copies to
equality); negatives are identical except where a recursive blur rounds 1 ulp apart; RGBA16 within
+-1 LSB of the binary32 oracle on the engine's own negative.  The auto-exposure meter is not built:
a develop under auto_exposure refuses.
"""
import math

import numpy as np
import pytest

from conftest import PLAIN, delta
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
    with engine.open(img, PLAIN) as s:
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


def test_auto_exposure_is_refused_at_develop(engine):
    """The auto-exposure meter is not built: a session opens with auto_exposure on (the schema
    default), and a render that has to develop refuses with a message naming the stage; turning
    auto_exposure off renders."""
    import spektralab_engine_amd as spk
    img = synth.make_frame(160, 120, 8)
    with engine.open(img, delta(grain_active=False, halation_active=False, auto_exposure=True)) as s:
        with pytest.raises(spk.SpkError) as e:
            s.render("full")
        assert e.value.status == -5 and "auto exposure meter" in e.value.message
        s.set_params({"auto_exposure": False})
        out, _ = s.render("full")
        assert out.shape == (120, 160, 4)


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
    """Every parameter at the reference default except auto exposure (scatter + bounce halation,
    grain, DIR couplers, glare), on the synthetic stocks."""
    dl = dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK, auto_exposure=False)
    plan = compare(engine, oracle, dl, 640, 480, 11, 2e-3)
    assert oracle.field(plan, "ae_active")[0] == 0
    for f in ("couplers_active", "glare_active", "grain_active", "scatter_active", "bounce_active"):
        assert oracle.field(plan, f)[0] == 1, f


# ---- the stages every configured workload leaves off -------------------------------------------

def srgb_decode(v):
    return np.where(v <= 0.040449936, v / 12.92, np.sign((v + 0.055) / 1.055) * np.abs((v + 0.055) / 1.055) ** 2.4)


@pytest.mark.parametrize("space", ["sRGB", "ProPhoto RGB", "Adobe RGB (1998)", "ITU-R BT.709", "ACEScg"])
def test_input_decode(engine, oracle, space):
    """The tier image decoded by its colour space's transfer function before the upsample (a space
    without one, ACEScg, decodes as the identity)."""
    dl = delta(grain_active=False, halation_active=False, input_cctf_decoding=True, input_color_space=space)
    plan = compare(engine, oracle, dl, 320, 200, 12, 0)
    assert oracle.field(plan, "decode_active")[0] == (0 if space == "ACEScg" else 1)


GEOMETRIES = [
    dict(geometry_crop_x=0.1, geometry_crop_y=0.05, geometry_crop_w=0.7, geometry_crop_h=0.8),
    dict(geometry_rotation_deg=7.5),
    dict(geometry_quarter_turns=1),
    dict(geometry_quarter_turns=2, geometry_flip_h=True),
    dict(geometry_quarter_turns=3, geometry_flip_v=True, geometry_crop_w=0.5, geometry_rotation_deg=-30.0),
]


@pytest.mark.parametrize("geo", GEOMETRIES)
def test_geometry(engine, oracle, geo):
    """Crop, rotation, quarter turns and flips: bilinear at float64 source coordinates; the output
    takes the geometry's size.  With a decode, the taps are decoded first."""
    dl = delta(grain_active=False, halation_active=False, input_cctf_decoding=True, input_color_space="sRGB", **geo)
    plan = compare(engine, oracle, dl, 300, 180, 14, 0)
    assert oracle.field(plan, "geometry_active")[0] == 1
    w, h = oracle.plan_size(plan)
    if geo.get("geometry_quarter_turns", 0) % 2:
        assert (w, h) == (round(geo.get("geometry_crop_h", 1.0) * 180), round(geo.get("geometry_crop_w", 1.0) * 300))


def test_geometry_with_grain_and_halation(engine, oracle):
    dl = delta(grain_active=True, halation_active=True, geometry_quarter_turns=1, geometry_crop_w=0.9)
    compare(engine, oracle, dl, 400, 260, 15, 2e-3)


@pytest.mark.parametrize("ev", [0.5, 2.0])
def test_highlight_boost(engine, oracle, ev):
    plan = compare(engine, oracle, delta(grain_active=False, halation_active=False, halation_boost_ev=ev),
                   320, 200, 16, 0)
    assert oracle.field(plan, "boost_active")[0] == 1


def test_back_reflection_headroom(engine, oracle):
    """A film without antihalation: the back-reflection halation with its highlight headroom."""
    plan = compare(engine, oracle, delta(grain_active=False, halation_active=True, antihalation_removed=True),
                   320, 200, 17, 1e-3)
    assert oracle.field(plan, "headroom_active")[0] == 1


def test_camera_lens_blur(engine, oracle):
    plan = compare(engine, oracle, delta(grain_active=False, halation_active=False, lens_blur_um=25.0),
                   320, 200, 18, 1e-3)
    assert oracle.field(plan, "lens_blur_px")[0] > 0


@pytest.mark.parametrize("black,white", [(True, False), (False, True), (True, True)])
def test_scanner_levels(engine, oracle, black, white):
    dl = delta(grain_active=False, halation_active=False, scanner_black_correction=black,
               scanner_white_correction=white)
    plan = compare(engine, oracle, dl, 320, 200, 19, 0)
    assert oracle.field(plan, "levels_active")[0] == 1
    m, q = oracle.field(plan, "levels_m")[0], oracle.field(plan, "levels_q")[0]
    assert np.isfinite(m) and np.isfinite(q) and m > 0


def compare_edr(engine, oracle, dl, w, h, seed, mism_frac):
    """As compare(), with the EDR map's conditioning taken into account: it divides by the
    luminance y = r_y . rgb of the unsharp-masked RGB, which the mask can bring near 0 by
    cancellation between channels of opposite sign.  Where y is below 1 % of the largest channel
    magnitude (condition number above 100), a 1-ulp difference of the print's RGB is amplified
    past 1 LSB; those pixels are counted (at most 0.1 %), every other pixel is within 1 LSB."""
    img = synth.make_frame(w, h, seed)
    with engine.open(img, dl) as s:
        out, _ = s.render("full")
        plan = s.last_plan()
        neg = s.negative("full")
    neg_o = oracle.develop(plan, img)
    assert np.count_nonzero(neg != neg_o) <= neg.size * mism_frac
    import ctypes
    off = bytearray(bytes(plan))
    o = oracle.lib().oracle_plan_offset(b"edr_active")
    off[o:o + 4] = (0).to_bytes(4, "little")
    _, rgb = oracle.print_(ctypes.create_string_buffer(bytes(off), len(off)), neg, want_rgb=True)  # the map's input
    ry = oracle.field(plan, "edr_r_y", 3).astype(np.float64)
    y = np.maximum(np.tensordot(ry, rgb.astype(np.float64), axes=1), 0.0)
    ok = y >= 0.01 * np.abs(rgb).max(axis=0)
    d = np.abs(out.astype(np.int32) - oracle.print_(plan, neg).astype(np.int32))[..., :3].max(axis=2)
    print(f"EDR: {np.count_nonzero(~ok)} ill-conditioned pixels, max {d[~ok].max() if (~ok).any() else 0} LSB there; "
          f"max {d[ok].max()} LSB elsewhere")
    assert d[ok].max() <= 1
    assert np.count_nonzero(d[~ok] > 1) <= d.size * 1e-3
    return plan


def test_edr_tone_map(engine, oracle):
    dl = delta(grain_active=False, halation_active=False, extended_dynamic_range=True,
               print_stock=synth.PAPER_STOCK_EDR)
    plan = compare_edr(engine, oracle, dl, 320, 200, 20, 0)
    assert oracle.field(plan, "edr_active")[0] == 1 and oracle.field(plan, "edr_n")[0] == 57


def test_every_stage_together(engine, oracle):
    dl = dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK_EDR, auto_exposure=False,
              input_cctf_decoding=True, input_color_space="ProPhoto RGB", geometry_rotation_deg=3.0, geometry_crop_w=0.95, halation_boost_ev=1.0,
              antihalation_removed=True, lens_blur_um=10.0, scanner_black_correction=True,
              scanner_white_correction=True, extended_dynamic_range=True)
    compare_edr(engine, oracle, dl, 480, 320, 21, 3e-3)


def test_print_preview_table(engine, oracle):
    """spk_preview_stock_lut: the tier's cached negative through the stock's shipped table
    (trilinear, 16-bit), bit-identical to the oracle's lookup; spk_print_lut_table serves the
    blob's table itself."""
    import spektralab_engine_amd as spk
    img = synth.make_frame(640, 400, 22)
    lut, axes = synth.print_lut()
    assert np.array_equal(engine.print_lut_table(synth.PAPER_STOCK), lut)
    assert engine.print_lut_catalog()[synth.PAPER_STOCK]["lut_size"] == synth.LUT_SIZE
    with engine.open(img, delta(grain_active=True, halation_active=True)) as s:
        px, flags, reply = s.preview_stock_lut(None, "live")  # develops the live tier first
        neg = s.negative("live")
        px2, _, _ = s.preview_stock_lut(synth.PAPER_STOCK, "live")  # reuses it
        with pytest.raises(spk.SpkError) as e:
            s.preview_stock_lut("no_such_paper", "live")
        assert e.value.status == -2 and synth.PAPER_STOCK in e.value.message
    h, w = neg.shape[1:]
    assert px.shape == (h, w, 4) and flags["reprint"] and flags["negative_was_cached"]
    assert np.array_equal(px, oracle.print_lut(neg, lut, axes))
    assert np.array_equal(px, px2)
    assert reply["paired_film"] == synth.FILM_STOCK and reply["print_stock"] == synth.PAPER_STOCK
    assert reply["tier"] == "live" and "warning" not in reply
    with pytest.raises(spk.SpkError):
        engine.print_lut_table(synth.PAPER_STOCK_EDR)
