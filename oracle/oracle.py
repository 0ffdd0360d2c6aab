"""ORACLE -- test infrastructure only: ctypes loader of oracle/liboracle.so (oracle/chain.cpp).

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg to check or time the CPU
restatement of the render chain.  The product library never imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build():
    r = subprocess.run(["make", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        F, D, U, U16 = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.POINTER(C.c_uint16)
        V = C.c_void_p
        l.oracle_develop_f32.argtypes = [V, F, F]
        l.oracle_develop_f64.argtypes = [V, F, D]
        l.oracle_print_f32.argtypes = [V, F, U16, F]
        l.oracle_print_f64.argtypes = [V, D, U16, D]
        l.oracle_gaussian_f32.argtypes = [F, F, C.c_int, C.c_int, D, C.c_double]
        l.oracle_gaussian_adj_f32.argtypes = [F, F, C.c_int, C.c_int, D, C.c_double]
        l.oracle_mixture_f32.argtypes = [F, F, C.c_int, C.c_int, D, D, C.c_int, C.c_double]
        l.oracle_mixture_f32.restype = None
        l.oracle_gaussian_adj_f32.restype = None
        l.oracle_philox.argtypes = [U, U, U, C.c_uint32]
        l.oracle_poisson.argtypes = [F, C.c_uint32, C.c_uint32, C.c_uint32, F]
        l.oracle_uniform.argtypes = [C.c_uint32]
        l.oracle_uniform.restype = C.c_float
        l.oracle_fir_tap.argtypes = [C.c_double, C.c_int]
        l.oracle_fir_tap.restype = C.c_double
        l.oracle_fir_radius.argtypes = [C.c_double]
        l.oracle_yvv.argtypes = [C.c_double, D]
        l.oracle_reflect.argtypes = [C.c_int, C.c_int]
        l.oracle_print_lut.argtypes = [F, C.c_uint32, C.c_uint32, F, C.c_uint32, F, F, U16]
        l.oracle_print_lut.restype = None
        l.oracle_downscale.argtypes = [F, C.c_uint32, C.c_uint32, C.c_uint32, F, U, U]
        l.oracle_downscale.restype = None
        l.oracle_plan_offset.argtypes = [C.c_char_p]
        l.oracle_plan_offset.restype = C.c_long
        l.oracle_grain_counts.argtypes = [V, F, F]
        l.oracle_grain_counts.restype = None
        l.oracle_develop_draws.argtypes = [V, F, F, F, F]
        l.oracle_develop_draws.restype = None
        l.oracle_poisson_at.argtypes = [F, U, U, C.c_size_t, C.c_uint32, F]
        l.oracle_poisson_at.restype = None
        l.oracle_mitchell_sample.argtypes = [D, C.c_uint32, C.c_double, C.c_double, C.c_int, D]
        l.oracle_mitchell_sample.restype = None
        l.oracle_curve_lookup.argtypes = [F, F, F, C.c_uint32, C.c_float]
        l.oracle_curve_lookup.restype = C.c_float
        l.oracle_paper_lookup.argtypes = [V, C.c_uint32, C.c_float]
        l.oracle_paper_lookup.restype = C.c_float
        l.oracle_boost.argtypes = [F, C.c_size_t, C.c_double]
        l.oracle_boost.restype = None
        l.oracle_encode16.argtypes = [C.c_float, C.c_uint32, C.c_int]
        l.oracle_encode16.restype = C.c_uint16
        l.oracle_cam16_forward.argtypes = [V, D, C.c_int, D]
        l.oracle_cam16_forward.restype = None
        l.oracle_cam16_inverse.argtypes = [V, C.c_double, C.c_double, C.c_double, D]
        l.oracle_cam16_inverse.restype = None
        l.oracle_knee.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        l.oracle_knee.restype = C.c_double
        l.oracle_cos.argtypes = [C.c_double]
        l.oracle_cos.restype = C.c_double
        l.oracle_lognormal.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float]
        l.oracle_lognormal.restype = C.c_float
        l.oracle_micro_field.argtypes = [C.c_uint32, C.c_float, C.c_uint32, F]
        l.oracle_micro_field.restype = None
        l.oracle_set_threads.argtypes = [C.c_int]
        l.oracle_set_threads.restype = None
        l.oracle_frozen_f64.argtypes = [V, F, F, F, C.c_int, C.c_double, D, D, D]
        l.oracle_frozen_f64.restype = C.c_double
        l.oracle_frozen_grad.argtypes = [V, F, F, F, D, D, C.c_int, D]
        l.oracle_frozen_grad.restype = C.c_double
        for n in ("oracle_log10f", "oracle_logf", "oracle_expf"):
            getattr(l, n).argtypes = [C.c_float]
            getattr(l, n).restype = C.c_float
        for fn in (l.oracle_develop_f32, l.oracle_develop_f64, l.oracle_print_f32, l.oracle_print_f64,
                   l.oracle_gaussian_f32, l.oracle_philox, l.oracle_poisson):
            fn.restype = None
        _lib = l
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


_DT = {"width": "u4", "height": "u4", "L": "u4", "K": "u4", "scatter_active": "u4", "n_scatter": "u4",
       "bounce_active": "u4", "n_bounce": "u4", "grain_active": "u4", "grain_mode": "u4", "grain_seed": "u4",
       "grain_k": "u4", "printed": "u4", "paper_k": "u4", "cctf_active": "u4", "cctf_mode": "u4",
       "a": "f4", "ec_gain": "f4", "dmin_l": "f4", "sat_d": "f4", "count": "f4", "uniformity": "f4",
       "dmin_c": "f4", "k_gain": "f4", "table": "f4", "curves": "f4", "paper": "f4", "cmax": "f4",
       "grain_axis": "f4", "grain_inv": "f4", "grain_layers": "f4", "ae_active": "u4", "ae_gain": "f4",
       "couplers_active": "u4", "glare_active": "u4", "glare_seed": "u4", "n_dir": "u4", "spectral_fp16": "u4",
       "decode_active": "u4", "decode_mode": "u4", "geometry_active": "u4", "src_w": "u4", "src_h": "u4",
       "boost_active": "u4", "headroom_active": "u4", "levels_active": "u4", "levels_m": "f4", "levels_q": "f4",
       "edr_active": "u4", "edr_n": "u4", "edr_table": "f4", "edr_r_y": "f4", "coupler_m": "f4",
       "coupler_dmax": "f4", "coupler_positive": "u4", "grain_shift": "u4", "grain_dye": "u4", "grain_micro": "u4",
       "grain_micro_sigma": "f4", "grain_n_sub": "u4", "grain_inv_n_sub": "f4"}


def field(plan, name, count=1):
    """A member of an spk_plan buffer as a numpy array (count elements)."""
    off = lib().oracle_plan_offset(name.encode())
    if off < 0:
        raise KeyError(name)
    dt = np.dtype(_DT.get(name, "f8"))
    return np.frombuffer(bytes(plan[off:off + dt.itemsize * count]), dtype=dt)


def set_field(plan, name, values):
    """Overwrites a member of a writable spk_plan buffer (a test's edited copy of a plan)."""
    off = lib().oracle_plan_offset(name.encode())
    if off < 0:
        raise KeyError(name)
    v = np.atleast_1d(np.asarray(values, dtype=np.dtype(_DT.get(name, "f8"))))
    np.frombuffer(plan, dtype=np.uint8)[off:off + v.nbytes] = np.frombuffer(v.tobytes(), dtype=np.uint8)


def plan_size(plan) -> tuple[int, int]:
    """(width, height) of a plan buffer (spk_plan: magic, version, width, height, ...)."""
    head = np.frombuffer(bytes(plan[:16]), dtype=np.uint32)
    assert head[0] == 0x504B5053, "not an spk_plan"
    return int(head[2]), int(head[3])


def source_size(plan) -> tuple[int, int]:
    """(width, height) of the tier image a plan develops (the geometry's source when it is on)."""
    if int(field(plan, "geometry_active")[0]):
        return int(field(plan, "src_w")[0]), int(field(plan, "src_h")[0])
    return plan_size(plan)


def develop(plan, rgb, f64=False):
    """rgb: the tier image, float32 (source_size(plan)[1] x [0] x 3).  Returns the negative, 3 x H x W."""
    w, h = plan_size(plan)
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    sw, sh = source_size(plan)
    assert rgb.shape == (sh, sw, 3)
    out = np.zeros(3 * w * h, dtype=np.float64 if f64 else np.float32)
    if f64:
        lib().oracle_develop_f64(C.cast(plan, C.c_void_p), _p(rgb, C.c_float), _p(out, C.c_double))
    else:
        lib().oracle_develop_f32(C.cast(plan, C.c_void_p), _p(rgb, C.c_float), _p(out, C.c_float))
    return out.reshape(3, h, w)


def print_(plan, negative, f64=False, want_rgb=False):
    """negative: 3 x H x W.  Returns the RGBA16 image H x W x 4 (and the linear RGB)."""
    w, h = plan_size(plan)
    out = np.zeros(4 * w * h, dtype=np.uint16)
    if f64:
        neg = np.ascontiguousarray(negative, dtype=np.float64).ravel()
        rgb = np.zeros(3 * w * h, dtype=np.float64) if want_rgb else None
        lib().oracle_print_f64(C.cast(plan, C.c_void_p), _p(neg, C.c_double), _p(out, C.c_uint16),
                               _p(rgb, C.c_double) if want_rgb else None)
    else:
        neg = np.ascontiguousarray(negative, dtype=np.float32).ravel()
        rgb = np.zeros(3 * w * h, dtype=np.float32) if want_rgb else None
        lib().oracle_print_f32(C.cast(plan, C.c_void_p), _p(neg, C.c_float), _p(out, C.c_uint16),
                               _p(rgb, C.c_float) if want_rgb else None)
    img = out.reshape(h, w, 4)
    return (img, rgb.reshape(3, h, w)) if want_rgb else img


def render(plan, rgb, f64=False):
    return print_(plan, develop(plan, rgb, f64), f64)


def gaussian(image, sigma3, crossover=3.0):
    img = np.ascontiguousarray(image, dtype=np.float32)
    h, w, _ = img.shape
    out = np.zeros_like(img)
    s = (C.c_double * 3)(*[float(v) for v in sigma3])
    lib().oracle_gaussian_f32(_p(img, C.c_float), _p(out, C.c_float), w, h, s, crossover)
    return out


def mixture(image, components, crossover=3.0):
    """The halation mixture of an H x W x 3 image: components = [(sigma, weight), ...] (scalars or
    3 per-channel values), accumulated in list order."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    h, w, _ = img.shape
    out = np.zeros_like(img)
    n = len(components)
    sg = (C.c_double * (3 * n))(*[float(v) for s, _ in components for v in np.broadcast_to(s, 3)])
    wt = (C.c_double * (3 * n))(*[float(v) for _, x in components for v in np.broadcast_to(x, 3)])
    lib().oracle_mixture_f32(_p(img, C.c_float), _p(out, C.c_float), w, h, sg, wt, n, crossover)
    return out


def gaussian_adjoint(image, sigma3, crossover=3.0):
    """The transpose of gaussian() (the FIR passes are symmetric; the recursion is transposed)."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    h, w, _ = img.shape
    out = np.zeros_like(img)
    s = (C.c_double * 3)(*[float(v) for v in sigma3])
    lib().oracle_gaussian_adj_f32(_p(img, C.c_float), _p(out, C.c_float), w, h, s, crossover)
    return out


def print_lut(negative, table, axes):
    """negative: 3 x H x W float32; table: (S, S, S, 3); axes: (3, S).  Returns H x W x 4 uint16."""
    neg = np.ascontiguousarray(negative, dtype=np.float32)
    _, h, w = neg.shape
    t = np.ascontiguousarray(table, dtype=np.float32)
    S = t.shape[0]
    lo = np.ascontiguousarray(axes[:, 0], dtype=np.float32)
    s = np.array([np.float32(1.0 / (float(axes[c, -1]) - float(axes[c, 0]))) for c in range(3)], np.float32)
    out = np.zeros((h, w, 4), np.uint16)
    lib().oracle_print_lut(_p(neg, C.c_float), w, h, _p(t, C.c_float), S, _p(lo, C.c_float), _p(s, C.c_float),
                           _p(out, C.c_uint16))
    return out


def downscale(image, long_edge):
    """Tier downscale of an H x W x 3 float32 image to the long edge (the oracle's restatement)."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    h, w, _ = img.shape
    ow, oh = C.c_uint32(), C.c_uint32()
    lib().oracle_downscale(_p(img, C.c_float), w, h, long_edge, None, C.byref(ow), C.byref(oh))
    out = np.zeros((oh.value, ow.value, 3), np.float32)
    lib().oracle_downscale(_p(img, C.c_float), w, h, long_edge, _p(out, C.c_float), C.byref(ow), C.byref(oh))
    return out


# ---- the differentiable (frozen-count) render ---------------------------------------------------
AD_QUANTITIES = 40   # derivative directions of the parameters (layout in oracle/chain.cpp)
AD_INPUT_DIRS = 8    # further directions along the input pixels


def grain_counts(plan, rgb):
    """The Poisson counts the binary32 develop draws: float32 H x W x 9 ([3 c + l])."""
    w, h = plan_size(plan)
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    out = np.zeros((h, w, 9), np.float32)
    lib().oracle_grain_counts(C.cast(plan, C.c_void_p), _p(rgb, C.c_float), _p(out, C.c_float))
    return out


def develop_draws(plan, rgb):
    """The binary32 develop with its grain draws: (negative 3 x H x W, lambdas 9 x H x W,
    counts 9 x H x W), draw 3 c + l (channel c, sub-layer l) -- the layout spkx_grain_record uses."""
    w, h = plan_size(plan)
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    neg = np.zeros(3 * w * h, np.float32)
    cnt = np.zeros((h * w, 9), np.float32)
    lam = np.zeros((h * w, 9), np.float32)
    lib().oracle_develop_draws(C.cast(plan, C.c_void_p), _p(rgb, C.c_float), _p(neg, C.c_float), _p(cnt, C.c_float),
                               _p(lam, C.c_float))
    return (neg.reshape(3, h, w), np.ascontiguousarray(lam.T).reshape(9, h, w),
            np.ascontiguousarray(cnt.T).reshape(9, h, w))


def poisson_at(lam, pixel, stream, seed):
    """The grain sampler's Poisson(lam[i]) for pixel[i] on stream[i] (1-D arrays of one length)."""
    lam = np.ascontiguousarray(lam, dtype=np.float32).ravel()
    pixel = np.ascontiguousarray(pixel, dtype=np.uint32).ravel()
    stream = np.ascontiguousarray(stream, dtype=np.uint32).ravel()
    out = np.zeros_like(lam)
    lib().oracle_poisson_at(_p(lam, C.c_float), _p(pixel, C.c_uint32), _p(stream, C.c_uint32), len(lam), seed,
                            _p(out, C.c_float))
    return out


def frozen_f64(plan, rgb, counts, target, idx=-1, h=0.0, wt_dir=None, in_dir=None, want_enc=False):
    """Loss (and optionally the encoded output, 3 x H x W) of the float64 frozen-count render with
    quantity idx moved by h (idx >= 40: the pixels along in_dir)."""
    w, hh = plan_size(plan)
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    tgt = np.ascontiguousarray(target, dtype=np.float32)
    cnt = np.ascontiguousarray(counts, dtype=np.float32)
    wd = np.ascontiguousarray(wt_dir if wt_dir is not None else np.zeros(243), dtype=np.float64)
    idr = None if in_dir is None else np.ascontiguousarray(in_dir, dtype=np.float64)
    enc = np.zeros(3 * w * hh) if want_enc else None
    loss = lib().oracle_frozen_f64(C.cast(plan, C.c_void_p), _p(rgb, C.c_float), _p(cnt, C.c_float), _p(tgt, C.c_float),
                                   idx, h, _p(wd, C.c_double), None if idr is None else _p(idr, C.c_double),
                                   None if enc is None else _p(enc, C.c_double))
    return (loss, enc.reshape(3, hh, w)) if want_enc else loss


def frozen_grad(plan, rgb, counts, target, wt_dir, in_dirs):
    """Forward-mode float64 derivatives of the frozen-count loss: (loss, grad[48])."""
    w, hh = plan_size(plan)
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    tgt = np.ascontiguousarray(target, dtype=np.float32)
    cnt = np.ascontiguousarray(counts, dtype=np.float32)
    wd = np.ascontiguousarray(wt_dir, dtype=np.float64)
    dirs = np.ascontiguousarray(in_dirs, dtype=np.float64).reshape(-1, 3 * w * hh)
    assert dirs.shape[0] <= AD_INPUT_DIRS
    grad = np.zeros(AD_QUANTITIES + AD_INPUT_DIRS)
    loss = lib().oracle_frozen_grad(C.cast(plan, C.c_void_p), _p(rgb, C.c_float), _p(cnt, C.c_float), _p(tgt, C.c_float),
                                    _p(wd, C.c_double), _p(dirs, C.c_double), dirs.shape[0], _p(grad, C.c_double))
    return loss, grad


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    lib().oracle_philox(_p(ctr, C.c_uint32), _p(key, C.c_uint32), _p(out, C.c_uint32), len(ctr))
    return out


def poisson(lam, stream, seed):
    lam = np.ascontiguousarray(lam, dtype=np.float32).ravel()
    out = np.zeros_like(lam)
    lib().oracle_poisson(_p(lam, C.c_float), len(lam), stream, seed, _p(out, C.c_float))
    return out


def fir_taps(sigma):
    r = lib().oracle_fir_radius(sigma)
    return np.array([lib().oracle_fir_tap(sigma, i) for i in range(2 * r + 1)])


def yvv(sigma):
    out = (C.c_double * 4)()
    lib().oracle_yvv(sigma, out)
    return tuple(out)


# ---- single-stage probes (known-answer tests) -----------------------------------------------------
def mitchell_sample(table, t0, t1, f32=True):
    """The upsample's bicubic sample of an L x L x 3 table at (t0, t1) (float32 or float64 arithmetic)."""
    table = np.ascontiguousarray(table, dtype=np.float64)
    L = table.shape[0]
    out = np.zeros(3)
    lib().oracle_mitchell_sample(_p(table, C.c_double), L, t0, t1, int(f32), _p(out, C.c_double))
    return out


def curve_lookup(x, inv, y, v):
    """The film density-curve lookup (clamped piecewise linear, float32)."""
    x, inv, y = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, inv, y))
    return float(lib().oracle_curve_lookup(_p(x, C.c_float), _p(inv, C.c_float), _p(y, C.c_float), len(x), v))


def paper_lookup(plan, c, v):
    return float(lib().oracle_paper_lookup(C.cast(plan, C.c_void_p), c, v))


def boost(values, ev):
    """The highlight boost of a plane of exposure values (float32), in place on a copy."""
    h = np.array(values, dtype=np.float32)
    lib().oracle_boost(_p(h, C.c_float), h.size, ev)
    return h


def encode16(v, mode, active=True):
    """The output encode of one value: transfer function `mode` (when active), then 16 bits."""
    return int(lib().oracle_encode16(v, mode, int(active)))


def cam16_forward(plan, xyz, f32=False):
    """(J', M', h) of an XYZ triple under the plan's viewing conditions."""
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    out = np.zeros(3)
    lib().oracle_cam16_forward(C.cast(plan, C.c_void_p), _p(x, C.c_double), int(f32), _p(out, C.c_double))
    return out


def cam16_inverse(plan, Jp, Mp, h):
    out = np.zeros(3)
    lib().oracle_cam16_inverse(C.cast(plan, C.c_void_p), Jp, Mp, h, _p(out, C.c_double))
    return out


def knee(d, tau, lim, pw):
    return lib().oracle_knee(d, tau, lim, pw)


def cos(x):
    """The film side's binary64 cosine (Cody-Waite reduction by pi/2, Taylor series)."""
    return lib().oracle_cos(float(x))


def lognormal(m, s, u1, u2):
    """The micro-structure lognormal draw from the uniforms u1, u2 (binary32)."""
    return float(lib().oracle_lognormal(m, s, u1, u2))


def micro_field(n, sigma, seed):
    """The micro-structure field of pixels 0..n-1: (3, n) float32, lognormal(1, sigma) on streams 100 + c."""
    out = np.zeros(3 * n, dtype=np.float32)
    lib().oracle_micro_field(n, sigma, seed, _p(out, C.c_float))
    return out.reshape(3, n)
