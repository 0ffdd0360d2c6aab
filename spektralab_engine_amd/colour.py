"""Error measures of the precision study: 16-bit codes -> linear ProPhoto RGB -> CIE XYZ ->
CAM16-UCS (J', a', b'), and the CAM16-UCS colour difference of two RGBA16 images.

CAM16 follows Li et al., "Comprehensive color solutions: CAM16, CAT16, and CAM16-UCS", Color
Res. Appl. 42 (2017): the CAT16 matrix M16, average surround (F 1, c 0.69, N_c 1), adapting
luminance L_A 64 cd/m^2 and background Y_b 20 (the viewing conditions the reference's gamut
stage uses, src/print/internal.hpp:87-94), white = the ProPhoto (D50) white.  The ProPhoto RGB
primaries and transfer function are ROMM RGB's (ISO 22028-2).
"""
from __future__ import annotations

import numpy as np

M16 = np.array([[0.401288, 0.650173, -0.051461],
                [-0.250268, 1.204414, 0.045854],
                [-0.002079, 0.048952, 0.953127]])


def _rgb_to_xyz_matrix(prim, white):
    P = np.array([[x / y, 1.0, (1.0 - x - y) / y] for x, y in prim]).T
    W = np.array([white[0] / white[1], 1.0, (1.0 - white[0] - white[1]) / white[1]])
    return P * np.linalg.solve(P, W)


PROPHOTO_WHITE = (0.3457, 0.3585)
PROPHOTO_TO_XYZ = _rgb_to_xyz_matrix([(0.7347, 0.2653), (0.1596, 0.8404), (0.0366, 0.0001)], PROPHOTO_WHITE)


def decode_prophoto16(rgba16):
    """RGBA16 codes (H x W x 4, ProPhoto transfer) -> linear ProPhoto RGB, float64 (H x W x 3)."""
    v = np.asarray(rgba16)[..., :3].astype(np.float64) / 65535.0
    return np.where(v < 16.0 / 512.0, v / 16.0, v ** 1.8)


class Cam16Ucs:
    def __init__(self, white_xyz100, L_A=64.0, Y_b=20.0, F=1.0, c=0.69, N_c=1.0, D=None):
        X_w = np.asarray(white_xyz100, dtype=np.float64)
        Y_w = X_w[1]
        k = 1.0 / (5.0 * L_A + 1.0)
        self.F_L = 0.2 * k ** 4 * (5.0 * L_A) + 0.1 * (1.0 - k ** 4) ** 2 * (5.0 * L_A) ** (1.0 / 3.0)
        n = Y_b / Y_w
        self.n = n
        self.z = 1.48 + np.sqrt(n)
        self.N_bb = self.N_cb = 0.725 * n ** -0.2
        self.c, self.N_c = c, N_c
        if D is None:  # degree of adaptation
            D = np.clip(F * (1.0 - (1.0 / 3.6) * np.exp((-L_A - 42.0) / 92.0)), 0.0, 1.0)
        rgb_w = M16 @ X_w
        self.D_rgb = D * Y_w / rgb_w + 1.0 - D
        self.A_w = self._achromatic(self._adapt(rgb_w * self.D_rgb))

    def _adapt(self, rgb_c):
        f = (self.F_L * np.abs(rgb_c) / 100.0) ** 0.42
        return 400.0 * np.sign(rgb_c) * f / (27.13 + f) + 0.1

    def _achromatic(self, ra):
        return (2.0 * ra[..., 0] + ra[..., 1] + 0.05 * ra[..., 2] - 0.305) * self.N_bb

    def ucs(self, xyz100):
        """CIE XYZ (Y of the white = 100), ... x 3 -> CAM16-UCS J', a', b' (... x 3)."""
        J, _, h, M = self.jchm(xyz100)
        Jp = 1.7 * J / (1.0 + 0.007 * J)
        Mp = np.log1p(0.0228 * M) / 0.0228
        return np.stack([Jp, Mp * np.cos(h), Mp * np.sin(h)], axis=-1)

    def jchm(self, xyz100):
        """CAM16 lightness J, chroma C, hue angle h (radians) and colourfulness M."""
        rgb = np.asarray(xyz100, dtype=np.float64) @ M16.T
        ra = self._adapt(rgb * self.D_rgb)
        a = ra[..., 0] - 12.0 * ra[..., 1] / 11.0 + ra[..., 2] / 11.0
        b = (ra[..., 0] + ra[..., 1] - 2.0 * ra[..., 2]) / 9.0
        h = np.arctan2(b, a)
        e_t = 0.25 * (np.cos(h + 2.0) + 3.8)
        A = self._achromatic(ra)
        J = 100.0 * np.maximum(A / self.A_w, 0.0) ** (self.c * self.z)
        den = ra[..., 0] + ra[..., 1] + 21.0 / 20.0 * ra[..., 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            t = np.where(den != 0.0, (50000.0 / 13.0 * self.N_c * self.N_cb) * e_t * np.hypot(a, b) / den, 0.0)
        C = np.maximum(t, 0.0) ** 0.9 * np.sqrt(J / 100.0) * (1.64 - 0.29 ** self.n) ** 0.73
        M = C * self.F_L ** 0.25
        return J, C, h, M


def prophoto16_to_ucs(rgba16):
    white = 100.0 * (PROPHOTO_TO_XYZ @ np.ones(3))
    xyz = 100.0 * decode_prophoto16(rgba16) @ PROPHOTO_TO_XYZ.T
    return Cam16Ucs(white).ucs(xyz)


def delta_e_ucs(rgba16_a, rgba16_b):
    """Per-pixel CAM16-UCS colour difference of two ProPhoto RGBA16 images (float64, H x W)."""
    return np.linalg.norm(prophoto16_to_ucs(rgba16_a) - prophoto16_to_ucs(rgba16_b), axis=-1)


def lsb_stats(a, b):
    """Max, p99 and mean |code difference| over the three colour channels of two RGBA16 images."""
    d = np.abs(np.asarray(a)[..., :3].astype(np.int64) - np.asarray(b)[..., :3].astype(np.int64)).ravel()
    return {"max": int(d.max()), "p99": float(np.percentile(d, 99)), "mean": float(d.mean())}
