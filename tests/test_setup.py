"""The engine's float64 set-up (host logic, no GPU) through spkx_plan.

Pinned by the reference's known answers (CAM16 constants, tests/test_print.cpp:79-104), by closed
forms (scan normalisation, fog split, renormalised halation mix, FIR/IIR crossover) and by the
CAM16 round trip and C_max in-gamut property (test_print.cpp:106-146) with an independent numpy
statement of the published CAM16 equations (Li et al., Color Research and Application, 2017).
"""
import math

import numpy as np
import pytest

import spektralab_engine_amd as spk
from spektralab_engine_amd import synth
from conftest import PLAIN, GRAIN_HALATION, MIX3_OVERRIDES, delta


@pytest.fixture(scope="module")
def plan1(resources):
    return spk.host_plan(resources, PLAIN, 512, 512)


def test_cam16_prophoto_constants(oracle, plan1):
    k = oracle.field(plan1, "cam64", 22)
    assert abs(k[0] - 0.6839903845696502) < 1e-15   # F_L
    assert abs(k[3] - 0.2) == 0.0                  # n
    assert abs(k[1] - 1.0003040045593807) < 1e-15   # N_bb
    assert abs(k[4] - 1.9272135954999579) < 1e-15   # z
    assert abs(k[5] - 37.17349142011997) < 1e-13    # A_w
    np.testing.assert_allclose(k[19:22], [1.0048858995308805, 0.9991651597854243, 1.18239029867048], atol=1e-15)
    assert abs(k[18] - 100.0) < 1e-12               # J' of the adopted white
    assert k[8] == 1.0 and k[9] == 110.0 and abs(k[10] + math.pi) < 1e-15 and abs(k[11] - 2 * math.pi / 720) < 1e-15
    np.testing.assert_allclose(k[12:18], [0, 1, 6, 0.7, 1.0, 2.2])


# ---- an independent numpy statement of CAM16-UCS (Li et al. 2017) for the ProPhoto white ----------
M16 = np.array([[0.401288, 0.650173, -0.051461], [-0.250268, 1.204414, 0.045854], [-0.002079, 0.048952, 0.953127]])


def cam16_consts(k):
    return dict(F_L=k[0], N_bb=k[1], N_cb=k[2], n=k[3], z=k[4], A_w=k[5], c=k[6], N_c=k[7], D=np.array(k[19:22]))


def cam16_forward(q, xyz):
    rgb = M16 @ (100.0 * np.asarray(xyz))
    v = q["D"] * rgb
    f = (q["F_L"] * np.abs(v) / 100.0) ** 0.42
    ra, ga, ba = 400.0 * np.sign(v) * f / (27.13 + f) + 0.1
    a = ra - 12.0 * ga / 11.0 + ba / 11.0
    b = (ra + ga - 2.0 * ba) / 9.0
    h = math.atan2(b, a)
    et = 0.25 * (math.cos(h + 2.0) + 3.8)
    A = (2.0 * ra + ga + ba / 20.0 - 0.305) * q["N_bb"]
    J = 100.0 * (A / q["A_w"]) ** (q["c"] * q["z"])
    t = (50000.0 / 13.0 * q["N_c"] * q["N_cb"] * et * math.hypot(a, b)) / (ra + ga + 21.0 * ba / 20.0)
    C = t ** 0.9 * math.sqrt(J / 100.0) * (1.64 - 0.29 ** q["n"]) ** 0.73
    M = C * q["F_L"] ** 0.25
    return 1.7 * J / (1.0 + 0.007 * J), math.log(1.0 + 0.0228 * M) / 0.0228, h


def cam16_inverse(q, Jp, Mp, h):
    M = (math.exp(0.0228 * Mp) - 1.0) / 0.0228
    J = Jp / (1.7 - 0.007 * Jp)
    C = M / q["F_L"] ** 0.25
    t = (C / (math.sqrt(J / 100.0) * (1.64 - 0.29 ** q["n"]) ** 0.73)) ** (1.0 / 0.9)
    et = 0.25 * (math.cos(h + 2.0) + 3.8)
    A = q["A_w"] * (J / 100.0) ** (1.0 / (q["c"] * q["z"]))
    p1 = 50000.0 / 13.0 * q["N_c"] * q["N_cb"] * et
    p2 = A / q["N_bb"] + 0.305
    gamma = 23.0 * p2 * t / (23.0 * p1 + 11.0 * t * math.cos(h) + 108.0 * t * math.sin(h))
    a, b = gamma * math.cos(h), gamma * math.sin(h)
    ra = (460.0 * p2 + 451.0 * a + 288.0 * b) / 1403.0
    ga = (460.0 * p2 - 891.0 * a - 261.0 * b) / 1403.0
    ba = (460.0 * p2 - 220.0 * a - 6300.0 * b) / 1403.0
    out = []
    for x, d in zip((ra, ga, ba), q["D"]):
        v = x - 0.1
        out.append(np.sign(v) * 100.0 / q["F_L"] * (27.13 * abs(v) / (400.0 - abs(v))) ** (1.0 / 0.42) / d)
    return np.linalg.solve(M16, np.array(out)) / 100.0


def test_cam16_round_trip(oracle, plan1):
    q = cam16_consts(oracle.field(plan1, "cam64", 22))
    for xyz in ([0.2, 0.18, 0.1], [0.05, 0.1, 0.4], [0.6, 0.5, 0.3], [0.9, 1.0, 0.8]):
        Jp, Mp, h = cam16_forward(q, xyz)
        np.testing.assert_allclose(cam16_inverse(q, Jp, Mp, h), xyz, atol=1e-10)


def test_cmax_cells_in_gamut(oracle, plan1):
    """A C_max cell is in gamut; one bisection step (x4) beyond it is not, for most cells."""
    q = cam16_consts(oracle.field(plan1, "cam64", 22))
    cmax = oracle.field(plan1, "cmax64", 64 * 720).reshape(64, 720)
    to_rgb = oracle.field(plan1, "m_to_rgb64", 9).reshape(3, 3)
    assert cmax.min() >= 0.0 and 20.0 < cmax.max() < 150.0
    step = 150.0 / 2 ** 18

    def in_gamut(Jp, M, h):
        rgb = to_rgb @ cam16_inverse(q, Jp, M, h)
        return bool(np.all((rgb >= -1e-6) & (rgb <= 1 + 1e-6)))

    bad, outside = 0, 0
    cells = [(i, j) for i in range(3, 64, 7) for j in range(0, 720, 37)]
    for i, j in cells:
        Jp, h = 1.0 + 109.0 * i / 63.0, -math.pi + 2.0 * math.pi * j / 720.0
        m = cmax[i, j]
        if m > 0 and not in_gamut(Jp, m, h):
            bad += 1
        if m < 149.0 and not in_gamut(Jp, m + 4 * step, h):
            outside += 1
    assert bad == 0
    assert outside >= 0.9 * len(cells)


def test_scan_normalised_to_white(oracle, plan1):
    """wt_Y sums to 1: a zero-density, base-free print scans to Y = 1 (setup.cpp:236-256)."""
    wt = oracle.field(plan1, "scan_wt64", 243).reshape(81, 3)
    assert abs(wt[:, 1].sum() - 1.0) < 1e-12


def test_grain_constants(resources, oracle):
    plan = spk.host_plan(resources, GRAIN_HALATION, 600, 400, film_seed=7)
    assert oracle.field(plan, "grain_active")[0] == 1 and oracle.field(plan, "grain_seed")[0] == 7
    dmin_l = oracle.field(plan, "dmin_l", 9).reshape(3, 3)
    np.testing.assert_allclose(dmin_l.sum(axis=0), 0.03, rtol=1e-6)   # the fog split by layer fractions
    assert oracle.field(plan, "grain_final_blur")[0] == 0.65
    count = oracle.field(plan, "count", 9).reshape(3, 3)
    # count ~ p^2 f / (0.2 s_c s_l): the slow layer (s_l 0.5) holds 4x the fast one (s_l 2) per unit fraction
    frac_ratio = (dmin_l[2] / dmin_l[0])
    np.testing.assert_allclose(count[2] / count[0], 4.0 * frac_ratio, rtol=1e-5)


def test_halation_mixture_override(resources, oracle):
    plan = spk.host_plan(resources, GRAIN_HALATION, 600, 400, overrides=MIX3_OVERRIDES)
    assert oracle.field(plan, "scatter_active")[0] == 0
    assert oracle.field(plan, "bounce_active")[0] == 1 and oracle.field(plan, "n_bounce")[0] == 3
    comps = oracle.field(plan, "bounce", 3 * 6).reshape(3, 2, 3)
    np.testing.assert_array_equal(comps[:, 1, 0], [2.0, 8.0, 32.0])
    np.testing.assert_array_equal(comps[:, 0, 0], [0.6, 0.3, 0.1])
    A = np.float32(0.08) * np.float32(1.0)  # 'weak' antihalation, red
    assert abs(oracle.field(plan, "mix_a", 3)[0] - 1.0 / (1.0 + float(A))) < 1e-15


def test_scan_film_plan(oracle, resources):
    """scan_film: no enlarger or paper in the plan; the scan tables are the film's own under its
    viewing light (not the paper's); glare draws no field and the levels of a scanned negative are
    a pass-through (src/print/setup.cpp:267-270, run.cpp:153-161)."""
    d = dict(glare_active=True, scanner_black_correction=True, scanner_white_correction=True)
    scanned = spk.host_plan(resources, delta(scan_film=True, **d), 64, 64)
    printed = spk.host_plan(resources, delta(**d), 64, 64)
    assert oracle.field(printed, "printed")[0] == 1 and oracle.field(scanned, "printed")[0] == 0
    assert oracle.field(printed, "glare_active")[0] == 1 and oracle.field(scanned, "glare_active")[0] == 0
    assert oracle.field(printed, "levels_active")[0] == 1 and oracle.field(scanned, "levels_active")[0] == 0
    assert oracle.field(scanned, "edr_active")[0] == 0
    assert not np.array_equal(oracle.field(scanned, "scan_eps64", 243), oracle.field(printed, "scan_eps64", 243))


def test_refuses_edr_without_a_tone_map(resources):
    with pytest.raises(spk.SpkError, match="extended_dynamic_range"):
        spk.host_plan(resources, delta(extended_dynamic_range=True), 64, 64)


def test_geometry_plan_sizes(oracle, resources):
    """Output sizes (geometry.cpp:45-62): round(crop * source), swapped by an odd quarter turn."""
    plan = spk.host_plan(resources, delta(geometry_crop_w=0.5, geometry_crop_h=0.25, geometry_quarter_turns=1),
                         640, 480)
    assert oracle.plan_size(plan) == (120, 320)
    assert oracle.source_size(plan) == (640, 480)
    plan = spk.host_plan(resources, delta(geometry_rotation_deg=10.0), 640, 480)
    assert oracle.plan_size(plan) == (640, 480) and oracle.field(plan, "geometry_active")[0] == 1
    plan = spk.host_plan(resources, delta(), 640, 480)
    assert oracle.field(plan, "geometry_active")[0] == 0


def test_levels_and_edr_plan_fields(oracle, resources):
    """Levels: the black and white targets B = 0.01 / 12.92 and W = ((0.98 + 0.055) / 1.055)^2.4 map the
    film's Dmin and white density (print/setup.cpp:373-381): m y_k + q = B and m y_w + q = W, so
    with both on, m = (W - B) / (y_w - y_k) > 0.  EDR: the paper's tone-map nodes, narrowed."""
    plan = spk.host_plan(resources, delta(scanner_black_correction=True, scanner_white_correction=True), 64, 64)
    assert oracle.field(plan, "levels_active")[0] == 1 and oracle.field(plan, "levels_m")[0] > 0
    plan = spk.host_plan(resources, delta(extended_dynamic_range=True, print_stock=synth.PAPER_STOCK_EDR), 64, 64)
    assert oracle.field(plan, "edr_active")[0] == 1 and oracle.field(plan, "edr_n")[0] == 57
    want = np.float32(synth.edr_tone_map()["target_log2_y"])
    assert np.array_equal(oracle.field(plan, "edr_table", 57), want)


def test_coupler_and_glare_plan_fields(oracle, resources):
    """DIR diffusion = (1 - 0.06) G(20 um / p) + exponential surrogate(0.06, 200 um / p)
    (dir_blur.hpp:15-19); glare mean 0.03 amount, std 0.7 mean, sigma 0.5 rho sqrt(amount)."""
    w, h = 640, 480
    plan = spk.host_plan(resources, delta(dir_couplers_active=True, glare_active=True, glare_amount=2.0), w, h)
    assert oracle.field(plan, "couplers_active")[0] == 1 and oracle.field(plan, "n_dir")[0] == 4
    p = 1000.0 * 35.0 / w
    comps = oracle.field(plan, "dir", 4 * 6).reshape(4, 2, 3)
    assert np.allclose(comps[0, 1], 20.0 / p, rtol=1e-15) and np.allclose(comps[0, 0], 0.94, rtol=1e-15)
    lam = 200.0 / p
    np.testing.assert_allclose(comps[1:, 1, 0], np.array([0.5360, 1.5236, 2.7684]) * lam, rtol=1e-15)
    np.testing.assert_allclose(comps[1:, 0, 0], 0.06 * np.array([0.1633, 0.6496, 0.1870]), rtol=1e-15)
    assert abs(comps[:, 0, 0].sum() - (0.94 + 0.06 * 0.9999)) < 1e-15
    assert oracle.field(plan, "glare_active")[0] == 1
    assert np.isclose(oracle.field(plan, "glare_sigma")[0], 0.5 * np.sqrt(2.0), rtol=1e-15)


# ---- the plan bytes of the benchmark configurations ------------------------------------------------
def test_plan_bytes_unchanged(resources):
    """spkx_plan of configurations 1, 2 (600 x 400 and 6000 x 4000) and 5 hashes to the recorded
    digests (tests/golden/make_plan_digests.py): the set-up computes exactly what it computed when
    the digests were recorded."""
    import hashlib
    import json
    from pathlib import Path
    import sys
    sys.path.insert(0, str(Path(__file__).parent / "golden"))
    import make_plan_digests as mk
    want = json.loads((Path(__file__).parent / "golden" / "plan_digests.json").read_text())
    assert mk.digests(resources) == want


def test_plan_is_a_function_of_its_inputs(resources):
    """Two plans of the same inputs are byte-identical (the unused table tails are zero)."""
    a = bytes(spk.host_plan(resources, PLAIN, 96, 64))
    b = bytes(spk.host_plan(resources, PLAIN, 96, 64))
    assert a == b


# ---- spk_output_transform (host form) -------------------------------------------------------------
def test_output_transform_identity(resources):
    """sRGB -> sRGB: the identity matrix (no adaptation between equal whites), sRGB curve modes,
    the default gamut options."""
    r, cmax = spk.output_transform_host(resources, "sRGB", "sRGB")
    np.testing.assert_allclose(np.array(r["matrix"]).reshape(3, 3), np.eye(3), atol=1e-12)
    assert r["source_cctf_mode"] == 0 and r["target_cctf_mode"] == 0
    assert r["source_color_space"] == "sRGB" and r["target_color_space"] == "sRGB"
    g = r["gamut_compress"]
    assert g["algorithm"] == "cam16ucs" and g["cmax_rows"] == 64 and g["cmax_cols"] == 720
    assert g["lightness_compression_active"] is True
    np.testing.assert_allclose(g["k"][12:18], [0, 1, 6, 0.7, 1.0, 2.2])
    m_xyz, m_rgb = np.array(g["m_to_xyz"]).reshape(3, 3), np.array(g["m_to_rgb"]).reshape(3, 3)
    np.testing.assert_allclose(m_rgb @ m_xyz, np.eye(3), atol=1e-6)
    assert cmax.shape == (64, 720) and np.all(cmax >= 0) and cmax.max() > 0


@pytest.mark.parametrize("src,dst", [("ProPhoto RGB", "sRGB"), ("sRGB", "ProPhoto RGB"), ("ACEScg", "Display P3")])
def test_output_transform_maps_white_to_white(resources, src, dst):
    """Between spaces of different whites, CAT02 makes RGB white (1, 1, 1) of the source land on
    (1, 1, 1) of the target (von Kries adaptation maps white to white); the mode is the space's
    transfer (0 sRGB, 1 ProPhoto, 4 none)."""
    r, _ = spk.output_transform_host(resources, src, dst)
    m = np.array(r["matrix"]).reshape(3, 3)
    np.testing.assert_allclose(m @ np.ones(3), np.ones(3), atol=1e-9)
    modes = {"sRGB": 0, "ProPhoto RGB": 1, "ACEScg": 4, "Display P3": 0}
    assert r["source_cctf_mode"] == modes[src] and r["target_cctf_mode"] == modes[dst]


def test_output_transform_options_and_refusals(resources):
    r, _ = spk.output_transform_host(resources, "sRGB", "sRGB",
                                     {"algorithm": "off", "knee": [0.1, 0.9, 4], "lightness_compression_active": False})
    g = r["gamut_compress"]
    assert g["algorithm"] == "off" and g["lightness_compression_active"] is False
    np.testing.assert_allclose(g["k"][12:18], [0.1, 0.9, 4, 0, 1, 1])
    r, _ = spk.output_transform_host(resources, "sRGB", "sRGB", {"knee": [1, 2], "lightness_compression": [0.5, 1, 3]})
    np.testing.assert_allclose(r["gamut_compress"]["k"][12:18], [0, 1, 6, 0.5, 1, 3])  # a 2-array is ignored
    for gamut, status in [({"algorithm": "reinhard"}, -2), ("[1, 2]", -1), ("{", -1)]:
        with pytest.raises(spk.SpkError) as e:
            spk.output_transform_host(resources, "sRGB", "sRGB", gamut)
        assert e.value.status == status
    with pytest.raises(spk.SpkError) as e:
        spk.output_transform_host(resources, "sRGB", "Rec. 9999")
    assert e.value.status == -2


def test_plan_tables_filled_and_tails_zero(resources, oracle):
    """Every table of a grain + halation plan holds its data (not all zero) and nothing past it: the
    zeroing of unused tails must not reach a table another step filled."""
    plan = spk.host_plan(resources, GRAIN_HALATION, 120, 80, 0x4ABEA221, MIX3_OVERRIDES)
    L, K = int(oracle.field(plan, "L")[0]), int(oracle.field(plan, "K")[0])
    gk, pk = int(oracle.field(plan, "grain_k")[0]), int(oracle.field(plan, "paper_k")[0])
    assert gk == K > 0 and pk > 0
    for name, n, cap in [("table", L * L * 3, 0), ("grain_axis", 3 * gk, 3 * 1024), ("grain_inv", 3 * gk, 3 * 1024),
                         ("grain_layers", 9 * gk, 9 * 1024), ("paper", 5 * pk, 5 * 1024), ("cmax", 46080, 0)]:
        v = oracle.field(plan, name, max(n, cap)).view(np.float32)
        assert np.count_nonzero(v[:n]) > n // 2, name
        assert not np.any(v[n:]), name


FILTERS = ["c_filter_neutral", "m_filter_neutral", "y_filter_neutral", "m_filter_shift", "y_filter_shift"]


def test_filter_weights_and_their_jacobian(oracle, resources):
    """The enlarger weights of a CMY filtration (the differentiable render's filter quantities): at
    the parameters' own filtration they are the plan's weights bit for bit; their Jacobian equals
    central differences of plans built at moved filter parameters (the set-up path itself), the
    weights scaled by the moved exposure gain over the held one."""
    d = delta(grain_active=False, halation_active=False, c_filter_neutral=5.0, m_filter_shift=0.2,
              y_filter_shift=-0.3, filter_shift_scale=1.5)
    wt, jac, f0 = spk.filter_weights_host(resources, d)
    plan = spk.host_plan(resources, d, 64, 64)
    assert f0[0] == 5.0 and list(f0[3:]) == [0.2, -0.3]
    assert np.array_equal(wt, oracle.field(plan, "enl_wt64", 243))
    g0 = oracle.field(plan, "enl_g64", 3)
    h = 0.01
    for j, name in enumerate(FILTERS):
        def moved(v):
            dd = dict(d)
            dd[name] = float(f0[j] + v)
            p = spk.host_plan(resources, dd, 64, 64)
            return (oracle.field(p, "enl_wt64", 243).reshape(81, 3) * oracle.field(p, "enl_g64", 3) / g0).ravel()
        fd = (moved(h) - moved(-h)) / (2 * h)
        assert np.abs(jac[:, j] - fd).max() <= 1e-6 * np.abs(fd).max(), name
    # a shift moves the filter it shifts by filter_shift_scale times a neutral step
    assert np.allclose(jac[:, 3], 1.5 * jac[:, 1], rtol=1e-12, atol=0)
    assert np.allclose(jac[:, 4], 1.5 * jac[:, 2], rtol=1e-12, atol=0)
    # the weights at another filtration are the plan's at it
    f1 = f0 + np.array([3.0, -2.0, 4.0, 0.1, -0.1])
    wt1, _, _ = spk.filter_weights_host(resources, d, f1)
    p1 = spk.host_plan(resources, dict(d, **{n: float(v) for n, v in zip(FILTERS, f1)}), 64, 64)
    want = (oracle.field(p1, "enl_wt64", 243).reshape(81, 3) * oracle.field(p1, "enl_g64", 3) / g0).ravel()
    assert np.allclose(wt1, want, rtol=1e-14, atol=0)
    with pytest.raises(spk.SpkError, match="preflash"):
        spk.filter_weights_host(resources, dict(d, preflash_exposure=0.1))
