"""Synthetic resources folder and synthetic input frames.

The reference engine ships no film or paper data (reference README.md:5), so every run needs a
resources folder.  This module writes one in the reference's on-disk format
(docs/RESOURCE_FORMAT.md:204-308 for the constants blob, :31-150 for the profile JSON):

  <dir>/spektralab_constants.bin   SPKR v1 blob: wavelengths, CIE 1931 CMFs (the analytic
                                   multi-lobe fit of Wyman, Sloan and Shirley, JCGT 2013), CAT
                                   matrices, the eight baked colour spaces, a synthetic
                                   hanatos/spectra_lut (L x L x 81) and placeholder entries
  <dir>/profiles/<stock>.json      a 3-layer negative film and a colour paper

Stock model (the synthetic stock of the benchmark configurations):
  film   Gaussian sensitivities at 650/550/450 nm, sigma 30 nm; logistic D-logE sub-layer curves
         (fast/mid/slow) summing to the density curves (mid slope ~0.6, Dmin 0.2, Dmax 2.5, K=256);
         Gaussian dye densities; reference illuminant BB5500, viewing illuminant BB5000
  paper  the same sensitivities; a cdfs density_curves_model; Gaussian dyes; viewing BB5000
Every value is a closed-form function of the wavelength grid, so the folder is reproducible
byte for byte.
"""
from __future__ import annotations

import json
import os
import struct
from pathlib import Path

import numpy as np

WL = 380.0 + 5.0 * np.arange(81)
FILM_STOCK = "spk_synth_negative"
PAPER_STOCK = "spk_synth_paper"
PAPER_STOCK_EDR = "spk_synth_paper_edr"  # the same paper with an EDR tone map

DTYPE_CODE = {np.dtype("float64"): 0, np.dtype("float32"): 1, np.dtype("float16"): 2,
              np.dtype("int32"): 3, np.dtype("uint8"): 4}


# ---- the SPKR v1 blob (docs/RESOURCE_FORMAT.md:204-240) -------------------------------------
def write_blob(path: Path, entries: dict[str, np.ndarray]) -> None:
    """Header (magic, version 1, count, reserved), 104-byte entries, 8-aligned C-order data."""
    names = list(entries)
    header_size = 16 + 104 * len(names)
    offset = (header_size + 7) & ~7
    table = bytearray()
    blobs = []
    for name in names:
        a = np.ascontiguousarray(entries[name])
        if a.dtype not in DTYPE_CODE:
            raise ValueError(f"{name}: unsupported dtype {a.dtype}")
        raw = a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
        enc = name.encode()
        if len(enc) > 63 or not 1 <= a.ndim <= 4:
            raise ValueError(f"{name}: bad name or rank")
        dims = list(a.shape) + [0] * (4 - a.ndim)
        table += enc.ljust(64, b"\0")
        table += struct.pack("<II4IQQ", DTYPE_CODE[a.dtype], a.ndim, *dims, offset, len(raw))
        blobs.append((offset, raw))
        offset = (offset + len(raw) + 7) & ~7
    out = bytearray(offset)
    out[0:16] = struct.pack("<4sIII", b"SPKR", 1, len(names), 0)
    out[16:16 + len(table)] = table
    for off, raw in blobs:
        out[off:off + len(raw)] = raw
    Path(path).write_bytes(bytes(out))


# ---- colour science --------------------------------------------------------------------------
def _lobe(x, mu, s1, s2):
    s = np.where(x < mu, s1, s2)
    return np.exp(-0.5 * ((x - mu) / s) ** 2)


def cmfs_1931() -> np.ndarray:
    """81 x 3 CIE 1931 2-degree CMFs, multi-lobe Gaussian fit (Wyman, Sloan, Shirley 2013)."""
    x = 1.056 * _lobe(WL, 599.8, 37.9, 31.0) + 0.362 * _lobe(WL, 442.0, 16.0, 26.7) \
        - 0.065 * _lobe(WL, 501.1, 20.4, 26.2)
    y = 0.821 * _lobe(WL, 568.8, 46.9, 40.5) + 0.286 * _lobe(WL, 530.9, 16.3, 31.1)
    z = 1.217 * _lobe(WL, 437.0, 11.8, 36.0) + 0.681 * _lobe(WL, 459.0, 26.0, 13.8)
    return np.stack([x, y, z], axis=1)


CAT = {
    "CAT02": [0.7328, 0.4296, -0.1624, -0.7036, 1.6975, 0.0061, 0.0030, 0.0136, 0.9834],
    "CAT16": [0.401288, 0.650173, -0.051461, -0.250268, 1.204414, 0.045854, -0.002079, 0.048952, 0.953127],
    "Bradford": [0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296],
    "Von Kries": [0.40024, 0.70760, -0.08081, -0.22630, 1.16532, 0.04570, 0.0, 0.0, 0.91822],
    "XYZ Scaling": [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0],
}

D65 = (0.3127, 0.3290)
D50 = (0.3457, 0.3585)
ACES_W = (0.32168, 0.33767)
SPACES = {
    "sRGB": ((0.64, 0.33, 0.30, 0.60, 0.15, 0.06), D65),
    "Display P3": ((0.680, 0.320, 0.265, 0.690, 0.150, 0.060), D65),
    "ProPhoto RGB": ((0.7347, 0.2653, 0.1596, 0.8404, 0.0366, 0.0001), D50),
    "Adobe RGB (1998)": ((0.64, 0.33, 0.21, 0.71, 0.15, 0.06), D65),
    "ITU-R BT.709": ((0.64, 0.33, 0.30, 0.60, 0.15, 0.06), D65),
    "ITU-R BT.2020": ((0.708, 0.292, 0.170, 0.797, 0.131, 0.046), D65),
    "ACEScg": ((0.713, 0.293, 0.165, 0.830, 0.128, 0.044), ACES_W),
    "ACES2065-1": ((0.7347, 0.2653, 0.0, 1.0, 0.0001, -0.0770), ACES_W),
}


def rgb_to_xyz_matrix(prim, white) -> np.ndarray:
    """Normalised primary matrix: columns are the primaries' XYZ scaled so RGB(1,1,1) = white, Y=1."""
    xr, yr, xg, yg, xb, yb = prim
    P = np.array([[xr / yr, xg / yg, xb / yb], [1.0, 1.0, 1.0],
                  [(1 - xr - yr) / yr, (1 - xg - yg) / yg, (1 - xb - yb) / yb]])
    W = np.array([white[0] / white[1], 1.0, (1 - white[0] - white[1]) / white[1]])
    return P * np.linalg.solve(P, W)[None, :]


def blackbody(kelvin: float) -> np.ndarray:
    lam = WL * 1e-9
    c1, c2 = 3.741771e-16, 1.4388e-2
    return (c1 / (np.pi * lam ** 5)) / np.expm1(c2 / (lam * kelvin))


def chromaticity(spd: np.ndarray, cmf: np.ndarray) -> tuple[float, float]:
    X, Y, Z = spd @ cmf
    return X / (X + Y + Z), Y / (X + Y + Z)


# ---- the synthetic spectral-upsampling table -------------------------------------------------
def spectra_lut(L: int, cmf: np.ndarray) -> np.ndarray:
    """L x L x 81 spectra; node (i, j) sits at chromaticity x = 1 - sqrt(t0), y = t1 sqrt(t0)
    (the inverse of the table coordinates of docs/RESOURCE_FORMAT.md:262-266).  Each spectrum is a
    non-negative mix of three broad Gaussian bases whose XYZ matches (x, y, 1-x-y) where that is
    reachable, normalised to sum(s (xbar + ybar + zbar)) = 1."""
    basis = np.stack([np.exp(-0.5 * ((WL - mu) / s) ** 2) for mu, s in ((620.0, 50.0), (540.0, 45.0), (450.0, 40.0))],
                     axis=1)  # 81 x 3
    bxyz = cmf.T @ basis  # 3 x 3: XYZ of each basis
    t = np.linspace(0.0, 1.0, L)
    t0, t1 = np.meshgrid(t, t, indexing="ij")
    sq = np.sqrt(t0)
    x = 1.0 - sq
    y = t1 * sq
    z = 1.0 - x - y
    target = np.stack([x, y, z], axis=-1)  # L x L x 3
    w = np.einsum("ij,abj->abi", np.linalg.inv(bxyz), target)
    w = np.maximum(w, 0.0) + 1e-3
    s = np.einsum("lk,abk->abl", basis, w)
    norm = s @ cmf.sum(axis=1)
    return s / norm[..., None]


# ---- stocks ----------------------------------------------------------------------------------
def _gauss(mu, sigma):
    return np.exp(-0.5 * ((WL - mu) / sigma) ** 2)


def _logistic(x, x0, amp, slope):
    k = 4.0 * slope / amp
    return amp / (1.0 + np.exp(-k * (x - x0)))


def film_profile(cmf: np.ndarray, table: np.ndarray) -> dict:
    """3-layer negative: density curves = Dmin + sum of three logistic sub-layers."""
    K = 256
    le = np.linspace(-4.0, 3.0, K)
    x0 = np.log10(0.184)
    dmin, dmax = 0.2, 2.5
    amp = (dmax - dmin) / 3.0
    offsets = (-0.55, 0.0, 0.55)  # fast, mid, slow
    chan_shift = (0.03, 0.0, -0.04)
    layers = np.zeros((K, 3, 3))
    for l, off in enumerate(offsets):
        for c in range(3):
            layers[:, l, c] = _logistic(le, x0 + off + chan_shift[c], amp, 0.23)
    curves = dmin + layers.sum(axis=1)
    # sensitivities: Gaussians at 650 / 550 / 450 nm (sigma 30 nm), scaled so a grey of value v
    # exposes each layer to about v (the table at the reference white's node, times 1/3)
    ref = blackbody(5500.0)
    xw, yw = chromaticity(ref / ref.mean(), cmf)
    L = table.shape[0]
    o = 1.0 - xw
    i = int(round(min(max(o * o, 0.0), 1.0) * (L - 1)))
    j = int(round(min(max(yw / o, 0.0), 1.0) * (L - 1)))
    spec = table[i, j]
    sens = np.stack([_gauss(650.0, 30.0), _gauss(550.0, 30.0), _gauss(450.0, 30.0)], axis=1)
    sens = sens * ((1.0 / 3.0) / (spec @ sens))[None, :]
    dyes = np.stack([_gauss(640.0, 45.0), _gauss(545.0, 40.0), _gauss(445.0, 35.0)], axis=1)
    base = 0.08 + 0.12 * _gauss(430.0, 60.0)
    return {
        "info": {"stock": FILM_STOCK, "name": "Synthetic 3-layer negative", "type": "negative",
                 "support": "film", "use": "still", "antihalation": "weak",
                 "reference_illuminant": "BB5500", "viewing_illuminant": "BB5000"},
        "data": {
            "wavelengths": WL.tolist(),
            "log_sensitivity": np.log10(sens).tolist(),
            "channel_density": dyes.tolist(),
            "base_density": base.tolist(),
            "log_exposure": le.tolist(),
            "density_curves": curves.tolist(),
            "density_curves_layers": layers.tolist(),
        },
    }


def paper_profile() -> dict:
    """Colour paper: cdfs density model centred so a mid-grey print lands near density 0.75."""
    K = 256
    le = np.linspace(-3.0, 2.0, K)
    centers = np.array([[-0.25, 0.10, 0.45]] * 3) + np.array([[0.02], [0.0], [-0.02]])
    amplitudes = np.array([[0.75, 0.75, 0.75]] * 3)
    sigmas = np.array([[0.22, 0.22, 0.22]] * 3)
    from math import erfc, sqrt
    curves = np.zeros((K, 3))
    for c in range(3):
        for l in range(3):
            curves[:, c] += amplitudes[c, l] * np.array([0.5 * erfc(-((v - centers[c, l]) / sigmas[c, l]) / sqrt(2.0))
                                                         for v in le])
    curves += 0.06
    sens = np.stack([_gauss(650.0, 30.0), _gauss(550.0, 30.0), _gauss(450.0, 30.0)], axis=1)
    dyes = np.stack([_gauss(650.0, 40.0), _gauss(550.0, 38.0), _gauss(445.0, 35.0)], axis=1)
    return {
        "info": {"stock": PAPER_STOCK, "name": "Synthetic colour paper", "type": "negative",
                 "support": "paper", "use": "still", "viewing_illuminant": "BB5000",
                 "reference_illuminant": "BB5000"},
        "data": {
            "wavelengths": WL.tolist(),
            "log_sensitivity": np.log10(sens).tolist(),
            "channel_density": dyes.tolist(),
            "base_density": (0.04 + 0.0 * WL).tolist(),
            "log_exposure": le.tolist(),
            "density_curves": curves.tolist(),
            "density_curves_model": {"model_type": "cdfs", "centers": centers.tolist(),
                                     "amplitudes": amplitudes.tolist(), "sigmas": sigmas.tolist()},
        },
    }


def edr_tone_map() -> dict:
    """A tone map over log2 Y in [-10, 4]: identity up to the shoulder at Y = 0.5, then half the slope;
    the toe (Y < 2^-8) lifted by a quarter stop."""
    x = np.linspace(-10.0, 4.0, 57)
    t = np.where(x > -1.0, -1.0 + 0.5 * (x + 1.0), x)
    t = np.where(x < -8.0, x + 0.25 * (-8.0 - x) / 2.0, t)
    return {"version": 1, "domain_log2_y": [-10.0, 4.0], "toe_y": 2.0 ** -8, "shoulder_y": 0.5,
            "target_log2_y": t.tolist()}


SYNTH_VERSION = "3"  # bumped when the folder's contents change
LUT_SIZE = 17


def print_lut(S: int = LUT_SIZE) -> tuple[np.ndarray, np.ndarray]:
    """A print-preview table for the synthetic paper: (S, S, S, 3) float32 over negative densities
    -0.2 .. 3.0 per channel (axes (3, S)); each output channel a smooth transmittance of the three
    densities (not the engine's print: a stand-in with the table's format)."""
    ax = np.linspace(-0.2, 3.0, S)
    d0, d1, d2 = np.meshgrid(ax, ax, ax, indexing="ij")
    mix = np.array([[0.9, 0.08, 0.02], [0.07, 0.85, 0.08], [0.03, 0.1, 0.87]])
    t = np.stack([10.0 ** (-(mix[c, 0] * d0 + mix[c, 1] * d1 + mix[c, 2] * d2) * 0.6) for c in range(3)], axis=-1)
    return t.astype(np.float32), np.stack([ax, ax, ax]).astype(np.float32)


def make_resources(directory, L: int = 192) -> Path:
    """Writes the folder (idempotent: an existing complete folder of the same L and version is kept)."""
    d = Path(directory)
    stamp = d / ".synth_L"
    tag = f"{L}-{SYNTH_VERSION}"
    if stamp.exists() and stamp.read_text().strip() == tag and (d / "spektralab_constants.bin").exists():
        return d
    (d / "profiles").mkdir(parents=True, exist_ok=True)
    cmf = cmfs_1931()
    table = spectra_lut(L, cmf)
    entries: dict[str, np.ndarray] = {
        "spectral/wavelengths": WL.astype(np.float64),
        "colour/cmfs_1931_2deg": cmf.astype(np.float64),
        "colour/cmfs_lms": (cmf @ np.array(CAT["CAT16"]).reshape(3, 3).T).astype(np.float64),
        "colour/matrix_16": np.array(CAT["CAT16"], dtype=np.float64).reshape(3, 3),
        "colour/mallett_srgb_basis": np.stack([_gauss(610, 40), _gauss(545, 35), _gauss(460, 30)], 1),
        "hanatos/spectra_lut": table.astype(np.float32),
        "filters/kg3": (0.2 + 0.8 * _gauss(500.0, 120.0)).astype(np.float64),
        "filters/lens_canon": (1.0 - 0.5 * np.exp(-(WL - 380.0) / 20.0)).astype(np.float64),
    }
    for name, m in CAT.items():
        entries[f"colour/cat/{name}"] = np.array(m, dtype=np.float64).reshape(3, 3)
    for name, (prim, white) in SPACES.items():
        M = rgb_to_xyz_matrix(prim, white)
        p = f"colour/cs/{name}/"
        entries[p + "primaries"] = np.array(prim, dtype=np.float64)
        entries[p + "whitepoint"] = np.array(white, dtype=np.float64)
        entries[p + "matrix_RGB_to_XYZ"] = M
        entries[p + "matrix_XYZ_to_RGB"] = np.linalg.inv(M)
    lut, axes = print_lut()
    entries[f"print_lut/{PAPER_STOCK}"] = lut
    entries[f"print_lut_axes/{PAPER_STOCK}"] = axes
    write_blob(d / "spektralab_constants.bin", entries)
    (d / "print_luts.json").write_text(json.dumps({PAPER_STOCK: {
        "paired_film": FILM_STOCK, "lut_size": LUT_SIZE, "declared_pairing": True,
        "output_color_space": "ProPhoto RGB", "output_cctf_encoding": True}}))
    (d / "profiles" / f"{FILM_STOCK}.json").write_text(json.dumps(film_profile(cmf, table)))
    (d / "profiles" / f"{PAPER_STOCK}.json").write_text(json.dumps(paper_profile()))
    edr = paper_profile()
    edr["info"]["stock"] = PAPER_STOCK_EDR
    edr["info"]["name"] = "Synthetic colour paper (EDR tone map)"
    edr["data"]["edr_tone_map"] = edr_tone_map()
    (d / "profiles" / f"{PAPER_STOCK_EDR}.json").write_text(json.dumps(edr))
    stamp.write_text(tag)
    return d


def default_resources_dir(L: int = 192) -> Path:
    """A per-user cache location for the generated folder."""
    base = Path(os.environ.get("SPK_SYNTH_DIR", Path(os.environ.get("TMPDIR", "/tmp")) / f"spk_synth_L{L}"))
    return make_resources(base, L)


# ---- synthetic input frames ------------------------------------------------------------------
ASPECTS = (1.0, 4.0 / 3.0, 3.0 / 2.0, 16.0 / 9.0)


def batch_workload(n: int = 64, seed: int = 0, min_area: int = 256 * 256, max_area: int = 4096 * 2730):
    """The batched workload: n frames with log-uniform areas in [min_area, max_area] (mean about
    2.2 MP for the defaults), aspect ratios 1, 4:3, 3:2 and 16:9 in either orientation, and
    per-frame parameters: print_exposure U(0.5, 2), magenta and yellow filter shifts U(-0.3, 0.3)
    (the reference's live filter controls; it has no cyan shift), and the halation mixture
    sigma 2/8/32 px each scaled by U(0.5, 2), weights 0.6/0.3/0.1.  Returns a list of dicts
    {width, height, frame_seed, params, mixture}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        area = float(np.exp(rng.uniform(np.log(min_area), np.log(max_area))))
        r = float(ASPECTS[rng.integers(0, len(ASPECTS))])
        if rng.integers(0, 2):
            r = 1.0 / r
        w = max(16, int(round(np.sqrt(area * r))))
        h = max(16, int(round(area / w)))
        params = {"print_exposure": float(rng.uniform(0.5, 2.0)), "m_filter_shift": float(rng.uniform(-0.3, 0.3)),
                  "y_filter_shift": float(rng.uniform(-0.3, 0.3))}
        mixture = [[s * float(rng.uniform(0.5, 2.0)), wt] for s, wt in ((2.0, 0.6), (8.0, 0.3), (32.0, 0.1))]
        out.append({"width": w, "height": h, "frame_seed": 1000 + i, "params": params, "mixture": mixture})
    return out
def make_frame(width: int, height: int, seed: int = 0, rects: int = 64) -> np.ndarray:
    """Linear ProPhoto float32 H x W x 3: per-channel smooth gradients in log2 space spanning
    -8..+4 (values 2^-8..2^4), then `rects` axis-aligned rectangles with log-uniform colours."""
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.linspace(0.0, 1.0, height), np.linspace(0.0, 1.0, width), indexing="ij")
    img = np.empty((height, width, 3), dtype=np.float64)
    for c in range(3):
        a, b, ph = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0), rng.uniform(0, 2 * np.pi)
        t = 0.5 + 0.25 * np.sin(2 * np.pi * a * xx + ph) + 0.25 * np.cos(2 * np.pi * b * yy - ph)
        img[..., c] = -8.0 + 12.0 * t
    for _ in range(rects):
        x0, x1 = np.sort(rng.integers(0, width, 2))
        y0, y1 = np.sort(rng.integers(0, height, 2))
        img[y0:y1 + 1, x0:x1 + 1, :] = rng.uniform(-8.0, 4.0, 3)
    return np.exp2(img).astype(np.float32)
