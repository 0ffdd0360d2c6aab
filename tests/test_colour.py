"""The precision study's error measures (spektralab_engine_amd.colour), pinned on the CPU."""
import numpy as np

from spektralab_engine_amd import colour


def test_cam16_published_example():
    # Li et al. 2017 worked example as published with the colour-science CAM16 model:
    # XYZ (19.01, 20.00, 21.78), white (95.05, 100, 108.88), L_A 318.31, Y_b 20, average surround.
    m = colour.Cam16Ucs([95.05, 100.0, 108.88], L_A=318.31, Y_b=20.0)
    J, C, h, M = m.jchm(np.array([19.01, 20.00, 21.78]))
    assert abs(J - 41.7312079) < 1e-6
    assert abs(C - 0.1033557) < 1e-6
    assert abs(np.degrees(h) % 360.0 - 217.0679597) < 1e-6
    assert abs(M - 0.1074367) < 1e-6


def test_white_is_lightness_100_and_achromatic_when_fully_adapted():
    code = np.full((1, 1, 4), 65535, np.uint16)
    jab = colour.prophoto16_to_ucs(code)[0, 0]
    assert abs(jab[0] - 100.0) < 1e-9  # A = A_w -> J = 100 -> J' = 170 / 1.7
    white = 100.0 * (colour.PROPHOTO_TO_XYZ @ np.ones(3))
    J, C, h, M = colour.Cam16Ucs(white, D=1.0).jchm(white)
    assert abs(J - 100.0) < 1e-9 and C < 1e-9


def test_prophoto_transfer_and_primaries():
    # ROMM RGB: linear segment below 16/512 (v / 16), power 1.8 above; D50 white maps to Y = 1
    v = np.array([[[0, 512, 65535, 65535]]], np.uint16)
    lin = colour.decode_prophoto16(v)[0, 0]
    assert lin[0] == 0.0 and abs(lin[1] - (512 / 65535) / 16.0) < 1e-15 and lin[2] == 1.0
    xyz_w = colour.PROPHOTO_TO_XYZ @ np.ones(3)
    assert abs(xyz_w[1] - 1.0) < 1e-12 and abs(xyz_w[0] / xyz_w.sum() - 0.3457) < 1e-12


def test_delta_e_and_lsb_stats():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 65536, (8, 9, 4), dtype=np.uint16)
    assert np.all(colour.delta_e_ucs(a, a) == 0.0)
    b = a.copy()
    b[..., 0] = np.minimum(b[..., 0].astype(np.int64) + 3, 65535).astype(np.uint16)
    s = colour.lsb_stats(a, b)
    assert s["max"] == 3
    assert np.all(colour.delta_e_ucs(a, b) >= 0.0)
