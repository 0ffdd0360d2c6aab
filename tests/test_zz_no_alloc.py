"""The render that allocates nothing (spkx_render_device_static) against spkx_render_device.

Twin sessions are opened on the same frame with the same parameters, so both draw the same seeds:
the static render must equal the pooled one bit for bit (same kernels, same plan).  Two static
calls run in a row, the second on a smaller frame, so the second sees the first one's leftovers in
the shared scratch region.  The engine's memory pool must not grow during a static call.
"""
import numpy as np
import pytest

import spektralab_engine_amd as spk
from conftest import GRAIN_HALATION, MIX3_OVERRIDES, delta
from spektralab_engine_amd import synth

pytestmark = pytest.mark.gpu


def _pool(engine):
    return engine.memory_report()["pool"]


def _render(engine, frame, params, overrides, static, reprint_after=False):
    import torch
    h, w = frame.shape[:2]
    with engine.open(frame, params) as s:
        if overrides:
            s.set_overrides(overrides)
        out = torch.zeros((h, w, 4), dtype=torch.int16, device="cuda")
        if static:
            before = _pool(engine)
            s.render_device_static(out.data_ptr(), w)
            after = _pool(engine)
            # no block of the engine's pool was taken for the static render
            assert after["live_bytes"] == before["live_bytes"]
            assert after["frame_high_water_bytes"] == before["frame_high_water_bytes"]
        else:
            s.render_device(out.data_ptr(), w)
        torch.cuda.synchronize()
        first = out.cpu().numpy().view(np.uint16).copy()
        second = None
        if reprint_after:
            # a negative cached by the pooled render, reprinted by both entry points
            s.render_device(out.data_ptr(), w)
            out.zero_()
            if static:
                s.render_device_static(out.data_ptr(), w, reprint=True)
            else:
                s.render_device(out.data_ptr(), w, reprint=True)
            torch.cuda.synchronize()
            second = out.cpu().numpy().view(np.uint16).copy()
        return first, second


CASES = [
    ("grain_halation_600x400", 600, 400, GRAIN_HALATION, MIX3_OVERRIDES),
    ("grain_halation_320x200", 320, 200, GRAIN_HALATION, MIX3_OVERRIDES),
]


def test_static_render_matches_pooled_render_twice(engine):
    print(f"static render cases: {[c[0] for c in CASES]}", flush=True)
    for name, w, h, params, overrides in CASES:
        print(f"case {name}", flush=True)
        frame = synth.make_frame(w, h, 3)
        ref, _ = _render(engine, frame, params, overrides, static=False)
        got, _ = _render(engine, frame, params, overrides, static=True)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), f"{name}: static render differs from spkx_render_device"
        assert (got[..., 3] == 65535).all()


def test_static_reprint_with_grain_and_glare(engine):
    frame = synth.make_frame(400, 300, 5)
    params = delta(grain_active=True, halation_active=True, glare_active=True)
    ref1, ref2 = _render(engine, frame, params, None, static=False, reprint_after=True)
    got1, got2 = _render(engine, frame, params, None, static=True, reprint_after=True)
    assert np.array_equal(got1, ref1)
    assert np.array_equal(got2, ref2)


def test_static_render_refuses_a_frame_beyond_the_region(engine):
    import torch
    w, h = 3200, 2400  # the grain + 3-Gaussian halation chain needs about 1 GB of scratch at this size
    params = dict(GRAIN_HALATION, film_format_mm=135.0)  # keeps the pitch above the dye-cloud limit
    frame = np.full((h, w, 3), 0.18, np.float32)
    with engine.open(frame, params) as s:
        s.set_overrides(MIX3_OVERRIDES)
        out = torch.zeros((h, w, 4), dtype=torch.int16, device="cuda")
        with pytest.raises(spk.SpkError) as e:
            s.render_device_static(out.data_ptr(), w)
        assert e.value.status == -1  # SPK_ERR_INVALID_ARG
        assert "static" in e.value.message
        torch.cuda.synchronize()
        assert int(out.abs().sum()) == 0  # nothing was enqueued
