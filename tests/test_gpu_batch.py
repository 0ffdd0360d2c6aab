"""Batched rendering (spkx_batch_*): every frame of a batch equals the render its own session would
make, seeds included.

A fresh session's first render draws the first value of its seed sequence; a batch run draws, per
frame, the next value of that frame's session sequence.  So the batch's first run over fresh
sessions must reproduce fresh sessions' first renders, and its second run their second renders.
Film side bit-exact (same kernels, no contraction); print within +-1 LSB (the batched print
kernel may contract differently from the per-frame one).
"""
import numpy as np
import pytest
import torch

import spektralab_engine_amd as spk
from conftest import PLAIN, GRAIN_HALATION, MIX3_OVERRIDES
from spektralab_engine_amd import synth

pytestmark = pytest.mark.gpu

SIZES = [(300, 200), (257, 129), (64, 64), (1024, 600), (17, 301)]


def frames(n=len(SIZES)):
    out = []
    for i, (w, h) in enumerate(SIZES[:n]):
        params = {"print_exposure": 0.6 + 0.3 * i, "m_filter_shift": 0.05 * (i - 2), "y_filter_shift": -0.04 * i}
        mix = [[2.0 * (0.6 + 0.2 * i), 0.6], [8.0 * (1.4 - 0.15 * i), 0.3], [32.0 * (0.5 + 0.3 * i), 0.1]]
        out.append((synth.make_frame(w, h, 40 + i), params, mix))
    return out


def open_all(engine, fr, delta, fp16_index=None):
    ss = []
    for i, (img, params, mix) in enumerate(fr):
        d = dict(delta)
        d.update(params)
        s = engine.open(img, d)
        ov = dict(MIX3_OVERRIDES, halation_mixture=mix, fir_crossover=16.5)
        if i == fp16_index:
            ov["spectral_precision"] = "fp16"
        s.set_overrides(ov)
        ss.append(s)
    return ss


def device_outputs(fr):
    return [torch.zeros((img.shape[0], img.shape[1] + 3, 4), dtype=torch.int16, device="cuda") for img, _, _ in fr]


def host(t, w):
    return t.cpu().numpy().view(np.uint16)[:, :w]


def assert_close(a, b, what):
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    assert d.max() <= 1, f"{what}: max {d.max()} LSB"


@pytest.mark.parametrize("schedule", ["phased", "streams"])
@pytest.mark.parametrize("delta", [PLAIN, GRAIN_HALATION], ids=["grain_off", "grain_on"])
def test_batch_matches_per_session_renders(engine, delta, schedule):
    fr = frames()
    batch_sessions = open_all(engine, fr, delta)
    ref_sessions = open_all(engine, fr, delta)
    outs = device_outputs(fr)
    try:
        with spk.Batch(engine, batch_sessions, [o.data_ptr() for o in outs], [o.shape[1] for o in outs],
                       options={"autotune": False, "schedule": schedule, "streams": 2, "grid_per_cu": 4,
                                "block": 256}) as b:
            for run in range(2):
                b.run()
                b.sync()
                for i, (img, _, _) in enumerate(fr):
                    want, _ = ref_sessions[i].render("full")
                    assert_close(host(outs[i], img.shape[1]), want, f"run {run} frame {i}")
    finally:
        for s in batch_sessions + ref_sessions:
            s.close()


def test_batch_autotune_and_fp16_frame(engine):
    fr = frames(4)
    ss = open_all(engine, fr, PLAIN, fp16_index=1)
    refs = open_all(engine, fr, PLAIN, fp16_index=1)
    outs = device_outputs(fr)
    try:
        with spk.Batch(engine, ss, [o.data_ptr() for o in outs], [o.shape[1] for o in outs]) as b:
            info = b.info()
            assert info["frames"] == 4 and info["pixels"] == sum(img.shape[0] * img.shape[1] for img, _, _ in fr)
            assert len(info["tuning"]) >= 5 and all(t["ms"] > 0 for t in info["tuning"])
            best = min(t["ms"] for t in info["tuning"])
            chosen = [t for t in info["tuning"] if all(t[k] == info["config"][k] for k in info["config"])]
            assert {t["schedule"] for t in info["tuning"]} == {"phased"}
            assert chosen and min(t["ms"] for t in chosen) == best
            b.run()
            b.sync()
            for i, (img, _, _) in enumerate(fr):
                want, _ = refs[i].render("full")
                assert_close(host(outs[i], img.shape[1]), want, f"frame {i}")
    finally:
        for s in ss + refs:
            s.close()


def test_batch_autotune_without_fp16_takes_the_streams_schedule(engine):
    """With no FP16-spectral frame the tuned schedule is "streams" (only that schedule is timed),
    and every frame equals its own session's render, two runs in a row."""
    fr = frames(4)
    ss = open_all(engine, fr, GRAIN_HALATION)
    refs = open_all(engine, fr, GRAIN_HALATION)
    outs = device_outputs(fr)
    try:
        with spk.Batch(engine, ss, [o.data_ptr() for o in outs], [o.shape[1] for o in outs]) as b:
            info = b.info()
            assert info["config"]["schedule"] == "streams"
            assert {t["schedule"] for t in info["tuning"]} == {"streams"} and len(info["tuning"]) >= 5
            for run in range(2):
                b.run()
                b.sync()
                for i, (img, _, _) in enumerate(fr):
                    want, _ = refs[i].render("full")
                    assert_close(host(outs[i], img.shape[1]), want, f"run {run} frame {i}")
    finally:
        for s in ss + refs:
            s.close()


@pytest.mark.skip(reason="the streams schedule inside an autotuning sequence, with an FP16-spectral frame, "
                         "ended in a host segmentation fault whose cause is not established; that combination "
                         "is switched off (not a tuning candidate; explicit streams + FP16 is refused)")
def test_batch_autotune_over_both_schedules_with_fp16_frame(engine):
    pass


def test_batch_streams_schedule_refuses_fp16_sessions(engine):
    fr = frames(2)
    ss = open_all(engine, fr, PLAIN, fp16_index=0)
    outs = device_outputs(fr)
    try:
        with pytest.raises(spk.SpkError) as e:
            spk.Batch(engine, ss, [o.data_ptr() for o in outs], [o.shape[1] for o in outs],
                      options={"autotune": False, "schedule": "streams"})
        assert "FP16" in e.value.message
    finally:
        for s in ss:
            s.close()


def test_batch_rejects_bad_arguments(engine):
    img = synth.make_frame(32, 32, 1)
    with engine.open(img, PLAIN) as s:
        out = torch.zeros((32, 32, 4), dtype=torch.int16, device="cuda")
        with pytest.raises(spk.SpkError):
            spk.Batch(engine, [s], [out.data_ptr()], [16])  # stride below the width
        with pytest.raises(spk.SpkError):
            spk.Batch(engine, [s], [out.data_ptr()], [32], options={"streams": 9})


# Frames at pitches that switch the grain models on (film_format_mm 4): the single-layer model, dye
# clouds on one plane and on all nine, and dye clouds with the micro-structure field.
GRAIN_MODELS = [((512, 384), dict(grain_sublayers_active=False)), ((1024, 64), {}), ((4096, 64), {}),
                ((9000, 24), {})]


@pytest.mark.parametrize("schedule", ["phased", "streams"])
def test_batch_grain_models_match_per_session_renders(engine, schedule):
    """The batch's dye-cloud planes, micro-structure field (its seed read from the batch's device
    seed table) and single-layer draws equal each frame's own session render, two runs in a row."""
    base = dict(PLAIN, grain_active=True, film_format_mm=4.0)
    fr, deltas = [], []
    for i, ((w, h), extra) in enumerate(GRAIN_MODELS):
        fr.append((synth.make_frame(w, h, 90 + i), {}, None))
        deltas.append(dict(base, **extra))
    ss = [engine.open(img, d) for (img, _, _), d in zip(fr, deltas)]
    refs = [engine.open(img, d) for (img, _, _), d in zip(fr, deltas)]
    outs = device_outputs(fr)
    try:
        with spk.Batch(engine, ss, [o.data_ptr() for o in outs], [o.shape[1] for o in outs],
                       options={"autotune": False, "schedule": schedule, "streams": 2, "grid_per_cu": 4,
                                "block": 256}) as b:
            for run in range(2):
                b.run()
                b.sync()
                for i, (img, _, _) in enumerate(fr):
                    want, _ = refs[i].render("full")
                    assert_close(host(outs[i], img.shape[1]), want, f"run {run} frame {i}")
    finally:
        for s in ss + refs:
            s.close()
