"""The differentiable render on the GPU (spkx_diff_*, hand-written reverse pass) against the
float64 frozen-count reference of the oracle (forward-mode dual numbers of the same chain).

Bars: the encoded output within 3e-5 of float64 (about 2 LSB of the 16-bit encode), and every
gradient -- the 39 scalar quantities, the enlarger weights along a random direction, and the
input pixels along two random directions -- within a relative 1e-3 of the float64 derivative
(with an absolute floor of 1e-6 of the largest derivative for quantities whose true derivative is
near zero).  The grain counts are the ones each side's binary32 develop draws (bit-identical).
"""
import numpy as np
import pytest

import spektralab_engine_amd as spk
from spektralab_engine_amd import synth

from conftest import delta

pytestmark = pytest.mark.gpu

W, H = 96, 64  # the size of the cases that name none
MAPS = {"film_gamma": [1.05, 0.97, 1.0], "film_offset": [0.01, -0.02, 0.0],
        "paper_gamma": [0.98, 1.02, 1.0], "paper_offset": [-0.01, 0.0, 0.015]}
CASES = {
    # FIR below sigma 16.5: two FIR components and one recursive
    "fir_iir": (delta(grain_active=True, halation_active=True, film_format_mm=4.0),
                dict(MAPS, halation_mixture=[[1.4, 0.6], [6.1, 0.3], [20.0, 0.1]], halation_scatter=False,
                     fir_crossover=16.5)),
    # the default crossover (recursive halation), glare and exposure compensation
    "iir_glare": (delta(grain_active=True, halation_active=True, glare_active=True, exposure_compensation_ev=0.4),
                  dict(MAPS, halation_mixture=[[3.4, 0.7], [9.0, 0.3]], halation_scatter=False)),
    # no grain (every quantity's move re-renders without redrawing counts), FIR and recursive halation
    "halation": (delta(grain_active=False, halation_active=True, film_format_mm=4.0),
                 dict(MAPS, halation_mixture=[[1.4, 0.6], [6.1, 0.3], [20.0, 0.1]], halation_scatter=False,
                      fir_crossover=16.5)),
    # no grain, no halation
    "plain": (delta(grain_active=False, halation_active=False), dict(MAPS)),
    # a separable FIR component (sigma 6: radius 18, above the 2-D tile's 8) at a width where the
    # reverse pass fuses both sigma blurs with the dot product (k_fir_v2 + k_fir_h_dot)
    "fir_separable": (delta(grain_active=True, halation_active=True, film_format_mm=10.0),
                      dict(MAPS, halation_mixture=[[6.0, 0.6], [1.4, 0.3], [20.0, 0.1]], halation_scatter=False,
                           fir_crossover=16.5), (256, 160)),
    # the stock's own halation: per-channel sigmas (no mixture override)
    "per_channel": (delta(grain_active=False, halation_active=True, film_format_mm=4.0),
                    dict(MAPS, halation_scatter=False)),
}
GRAIN = {"fir_iir", "iir_glare", "fir_separable"}


def size_of(name):
    c = CASES[name]
    return c[2] if len(c) > 2 else (W, H)


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, engine, oracle):
    import torch
    dl, ov = CASES[request.param][:2]
    w, h = size_of(request.param)
    img = synth.make_frame(w, h, 11)
    s = engine.open(img, dl)
    s.set_overrides(ov)
    d = spk.Diff(s, slots=2)
    plan = d.plan()
    print(f"[diff case {request.param}] {d.info()}", flush=True)
    x = torch.from_numpy(img).cuda()
    yield request.param, s, d, plan, img, x, oracle
    d.close()
    s.close()


def _enc_ref(o, plan, img, counts, idx=-1, h=0.0, wt_dir=None):
    _, enc = o.frozen_f64(plan, img, counts, np.zeros(img.shape, np.float32), idx, h, wt_dir, want_enc=True)
    return enc.transpose(1, 2, 0)


def test_forward_matches_float64(case):
    import torch
    name, s, d, plan, img, x, o = case
    counts = o.grain_counts(plan, img)
    enc = spk.diff_render(d, x).detach().cpu().numpy()
    ref = _enc_ref(o, plan, img, counts)
    err = np.abs(enc - ref).max()
    print(f"[{name}] forward max |enc - f64| = {err:.3g}", flush=True)
    assert err <= 3e-5
    torch.cuda.synchronize()


def test_theta_moves_like_the_plan(case):
    """A quantity moved through theta renders as the float64 reference moved by the same amount.
    With grain on, a forward pass draws its counts at its own densities, so only the print-side
    quantities are moved there (the film side's moves change the draws)."""
    name, s, d, plan, img, x, o = case
    counts = o.grain_counts(plan, img)
    th0 = d.defaults()
    checks = [(31, 0.05), (33, 0.03), (36, -0.02), (38, 0.2)]
    if name not in GRAIN:
        checks += [(0, 0.05), (29, 0.02), (26, -0.03)]
    if name == "halation":
        checks += [(17, 0.5), (8, 0.05), (2, 0.03), (6, -0.02), (22, 1.0)]
    import torch
    for idx, hh in checks:
        th = th0.copy()
        th[idx] += hh
        enc = spk.diff_render(d, x, torch.from_numpy(th)).detach().cpu().numpy()
        ref = _enc_ref(o, plan, img, counts, idx, hh)
        err = np.abs(enc - ref).max()
        print(f"[{name}] theta[{idx}] += {hh}: max err {err:.3g}", flush=True)
        assert err <= 3e-5, (idx, err)


def test_gradients_match_forward_mode(case):
    import torch
    name, s, d, plan, img, x, o = case
    counts = o.grain_counts(plan, img)
    rng = np.random.default_rng(5)
    # the target: a render at quantities perturbed by N(0, 0.05) relative (absolute for the zero
    # offsets), with its own grain draw
    th_t = d.defaults()
    th_t = th_t * (1.0 + 0.05 * rng.normal(size=th_t.size)) + 0.01 * rng.normal(size=th_t.size) * (th_t == 0)
    th_t[39] = 0.0
    target = spk.diff_render(d, x, torch.from_numpy(th_t)).detach().cpu().numpy()
    wt_dir = rng.normal(0, 1, 243) * np.abs(o.field(plan, "enl_wt64", 243)).mean()
    in_dirs = rng.normal(0, 1, (2,) + img.shape) * img[None] * 0.1
    loss_ref, g = o.frozen_grad(plan, img, counts, target, wt_dir, in_dirs)

    xg = x.clone().requires_grad_(True)
    theta = torch.from_numpy(d.defaults()).requires_grad_(True)
    enc = spk.diff_render(d, xg, theta, slot=1)
    loss = ((enc - torch.from_numpy(target).cuda()) ** 2).mean()
    loss.backward()
    gt = theta.grad.numpy()
    gx = xg.grad.detach().cpu().numpy().astype(np.float64)
    assert float(loss) == pytest.approx(loss_ref, rel=1e-4)
    gpu = np.concatenate([gt[:39], [gt[40:] @ wt_dir], [(gx * in_dirs[j]).sum() for j in range(2)]])
    ref = np.concatenate([g[:40], g[40:42]])
    scale = np.abs(ref).max()
    rel = np.abs(gpu - ref) / np.maximum(np.abs(ref), 1e-6 * scale)
    for q in range(len(ref)):
        if ref[q] != 0.0 or gpu[q] != 0.0:
            print(f"[{name}] q{q:2d} gpu {gpu[q]: .6e} f64 {ref[q]: .6e} rel {rel[q]:.2e}", flush=True)
    print(f"[{name}] norm-wise relative error {np.linalg.norm(gpu - ref) / np.linalg.norm(ref):.2e}", flush=True)
    bad = [(q, gpu[q], ref[q]) for q in range(len(ref)) if abs(gpu[q] - ref[q]) > 1e-3 * abs(ref[q]) + 1e-6 * scale]
    assert not bad, bad
    assert gt[39] == 0.0


def test_filter_gradients_match_forward_mode(case):
    """The CMY filtration as quantities: theta's enlarger weights from the filters
    (spk.filter_theta); the gradient with respect to each filter against the float64 forward-mode
    derivative of the loss along that filter's column of d wt / d filters."""
    import torch
    name, s, d, plan, img, x, o = case
    counts = o.grain_counts(plan, img)
    wt0, jac, f0 = d.filter_weights()
    assert np.array_equal(wt0, o.field(plan, "enl_wt64", 243))
    rng = np.random.default_rng(9)
    th_t = d.defaults()
    th_t[spk.Diff.WT:] *= 1.0 + 0.05 * rng.normal(size=243)
    target = spk.diff_render(d, x, torch.from_numpy(th_t)).detach().cpu().numpy()
    f = torch.from_numpy(f0.copy()).requires_grad_(True)
    theta = spk.filter_theta(d, torch.from_numpy(d.defaults()), f)
    enc = spk.diff_render(d, x, theta, slot=1)
    loss = ((enc - torch.from_numpy(target).cuda()) ** 2).mean()
    loss.backward()
    gf = f.grad.numpy()
    none = np.zeros((1,) + img.shape)
    ref = np.array([o.frozen_grad(plan, img, counts, target, jac[:, k], none)[1][39] for k in range(5)])
    print(f"[{name}] filter gradients gpu {gf} f64 {ref}", flush=True)
    scale = np.abs(ref).max()
    assert np.all(np.abs(gf - ref) <= 1e-3 * np.abs(ref) + 1e-6 * scale), (gf, ref)


def test_refusals(engine):
    img = synth.make_frame(32, 24, 1)
    with engine.open(img, delta(dir_couplers_active=True, grain_active=False, halation_active=False)) as s:
        with pytest.raises(spk.SpkError, match="couplers"):
            spk.Diff(s)
    with engine.open(img, delta(grain_active=False, halation_active=False)) as s:
        s.set_overrides({"spectral_precision": "fp16"})
        with pytest.raises(spk.SpkError, match="spectral_precision"):
            spk.Diff(s)
        s.set_overrides({"spectral_precision": "fp32"})
        with pytest.raises(spk.SpkError):
            spk.Diff(s, slots=0)
        with spk.Diff(s) as d:
            import torch
            x = torch.zeros((24, 32, 3), device="cuda")
            with pytest.raises(spk.SpkError):  # no forward pass in the slot yet
                d.backward(0, x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr())
            with pytest.raises(spk.SpkError):
                d.forward(5, x.data_ptr(), x.data_ptr())
