"""The float64 differentiable reference (oracle/dual.h + the frozen-count render of oracle/chain.cpp)
against central finite differences of the same float64 function.

Forward-mode derivatives are exact to rounding; the finite differences (step 1e-6) agree to their
own truncation and to the kinks of the piecewise-linear curves, which is what the tolerances allow.
The mixture sigmas stay clear of the FIR radius steps (3 sigma + 1/2 crossing an integer), where
the render itself is discontinuous in sigma.
"""
import numpy as np
import pytest

import spektralab_engine_amd as spk
from spektralab_engine_amd import synth

W, H = 40, 28


@pytest.fixture(scope="module")
def case(resources):
    from oracle import oracle as o
    delta = dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK, auto_exposure=False,
                 dir_couplers_active=False, glare_active=False, grain_active=True, halation_active=True,
                 film_format_mm=4.0)
    ov = {"halation_mixture": [[1.4, 0.6], [6.1, 0.3], [20.0, 0.1]], "halation_scatter": False,
          "fir_crossover": 16.5, "film_gamma": [1.05, 0.97, 1.0], "film_offset": [0.01, -0.02, 0.0],
          "paper_gamma": [0.98, 1.02, 1.0], "paper_offset": [-0.01, 0.0, 0.015]}
    plan = spk.host_plan(resources, delta, W, H, film_seed=0x4ABEA221, overrides=ov)
    img = synth.make_frame(W, H, 3)
    counts = o.grain_counts(plan, img)
    _, enc = o.frozen_f64(plan, img, counts, np.zeros((H, W, 3), np.float32), want_enc=True)
    rng = np.random.default_rng(0)
    target = (enc.transpose(1, 2, 0) + rng.normal(0, 0.02, (H, W, 3))).astype(np.float32)
    wt_dir = rng.normal(0, 1, 243) * np.abs(o.field(plan, "enl_wt64", 243)).mean()
    in_dirs = rng.normal(0, 1, (2, H, W, 3)) * img[None] * 0.1
    return o, plan, img, counts, target, wt_dir, in_dirs


def test_counts_are_the_renders_draws(case):
    o, plan, img, counts, *_ = case
    assert counts.shape == (H, W, 9) and np.all(counts >= 0) and np.all(counts == np.floor(counts))
    # the frozen-count float64 render is the binary32 render up to rounding
    _, enc = o.frozen_f64(plan, img, counts, np.zeros((H, W, 3), np.float32), want_enc=True)
    out = o.render(plan, img)
    assert np.abs(np.floor(np.clip(enc.transpose(1, 2, 0), 0, 1) * 65535 + 0.5) - out[..., :3]).max() <= 2


def test_forward_mode_matches_finite_differences(case):
    o, plan, img, counts, target, wt_dir, in_dirs = case
    loss, g = o.frozen_grad(plan, img, counts, target, wt_dir, in_dirs)
    assert loss == pytest.approx(o.frozen_f64(plan, img, counts, target), rel=1e-12)
    scale = np.abs(g).max()
    bad = []
    for q in list(range(o.AD_QUANTITIES)) + [40, 41]:
        d = in_dirs[q - 40] if q >= 40 else None
        # two steps: a kink of a piecewise-linear curve within one step spoils that difference only
        fds = [(o.frozen_f64(plan, img, counts, target, q, h, wt_dir, d) -
                o.frozen_f64(plan, img, counts, target, q, -h, wt_dir, d)) / (2 * h) for h in (1e-6, 1e-7)]
        fd = min(fds, key=lambda v: abs(v - g[q]))
        if abs(fd - g[q]) > 1e-3 * max(abs(g[q]), 1e-3 * scale):
            bad.append((q, g[q], fds))
    assert not bad, bad
