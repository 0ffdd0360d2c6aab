"""Session behaviour the API document fixes, on the GPU: the seed sequence and its order, progress
records, cancellation and the parameter layers (reference docs/API_SPEC.md sections 2.3, 2.8 and
2.9; the reference's own checks: tests/session_check.py:170-207 and 281-307)."""
import threading

import numpy as np
import pytest

import spektralab_engine_amd as spk
from conftest import PLAIN, delta
from spektralab_engine_amd import synth

pytestmark = pytest.mark.gpu

# splitmix64 from 0x243F6A8885A308D3: the first three seeds after a (re)build (API_SPEC.md 2.8)
SEEDS = [0x4ABEA221, 0x23148989, 0x609DFE03]


def seeds_of(oracle, s):
    plan = s.last_plan()
    return int(oracle.field(plan, "grain_seed")[0]), int(oracle.field(plan, "glare_seed")[0])


def test_seed_order_render_then_reprint(engine, oracle):
    """Open, render("live"): draw 1 for grain, draw 2 for glare; reprint("live"): draw 3 for glare,
    grain unchanged; a rebuild restarts the sequence, a live field does not."""
    img = synth.make_frame(320, 200, 21)
    with engine.open(img, delta(grain_active=True, halation_active=False, glare_active=True)) as s:
        s.render("live")
        assert seeds_of(oracle, s) == (SEEDS[0], SEEDS[1])
        _, f = s.render("live", reprint=True)
        assert f["reprint"] and seeds_of(oracle, s) == (SEEDS[0], SEEDS[2])
        s.set_params({"print_exposure": 1.3})  # live field: same pipeline, the sequence goes on
        s.render("live", reprint=True)
        assert seeds_of(oracle, s)[1] not in SEEDS
        s.set_params({"glare_amount": 2.0})  # print-layer rebuild: the sequence restarts
        _, f = s.render("live", reprint=True)
        assert f["reprint"] and seeds_of(oracle, s) == (SEEDS[0], SEEDS[0])


def test_seeds_without_grain_or_glare(engine, oracle):
    """No grain: no film draw, so the print seed is the first draw (the film seed 0x5EED reaches no
    kernel and the plan's grain fields stay zero); no glare: print seed 0 and no print draw."""
    img = synth.make_frame(160, 120, 22)
    with engine.open(img, delta(grain_active=False, halation_active=False, glare_active=True)) as s:
        s.render("full")
        assert seeds_of(oracle, s) == (0, SEEDS[0])
        s.render("full")
        assert seeds_of(oracle, s) == (0, SEEDS[1])
    with engine.open(img, delta(grain_active=True, halation_active=False, glare_active=False)) as s:
        s.render("full")
        assert seeds_of(oracle, s) == (SEEDS[0], 0)


def test_pure_function_without_grain_and_glare(engine):
    """Grain and glare off: render and reprint give byte-identical pictures, and so do two sessions."""
    img = synth.make_frame(240, 160, 23)
    with engine.open(img, PLAIN) as a, engine.open(img, PLAIN) as b:
        pa, _ = a.render("full")
        pr, _ = a.render("full", reprint=True)
        pb, _ = b.render("full")
        assert np.array_equal(pa, pr) and np.array_equal(pa, pb)


def test_progress_record(engine):
    img = synth.make_frame(200, 140, 24)
    with engine.open(img, PLAIN) as s:
        p0 = s.progress()
        assert p0["progress_id"] == "" and p0["effect_versions"] == "" and p0["done"] is False
        _, f = s.render("full")
        p = s.progress()
        assert p["progress_id"] == f["progress_id"] and p["progress_id"].startswith("p")
        assert p["done"] is True and p["pct"] == 100 and p["cancelled"] is False
        assert p["auto_exposure_ev"] is None and p["node_times"] == {} and p["memory_pressure_critical"] is False
        assert "halation=" in p["effect_versions"] and p["stage"] == "scanning.print"
        assert s.progress(f["progress_id"])["done"] is True
        with pytest.raises(spk.SpkError) as e:
            s.progress("p999999999")
        assert e.value.status == -2 and "unknown progress id" in e.value.message
        # a reprint of the cached negative passes one node only
        _, f2 = s.render("full", reprint=True)
        assert f2["progress_id"] != f["progress_id"] and s.progress()["stage"] == "scanning.print"


def test_cancel_before_a_render_is_lost(engine):
    """A cancel marks the current render; the next render resets the record, so a cancel that
    arrived before it is lost and the render completes.  A cancel naming another id does nothing."""
    img = synth.make_frame(200, 140, 25)
    with engine.open(img, PLAIN) as s:
        _, f = s.render("full")
        s.cancel("p-not-this-one")
        assert s.progress()["cancelled"] is False
        s.cancel(f["progress_id"])
        assert s.progress()["cancelled"] is True
        px, f2 = s.render("full")
        p = s.progress()
        assert p["cancelled"] is False and p["done"] is True and px.shape == (140, 200, 4)


def test_cancel_from_another_thread(engine):
    """spk_cancel is callable while a render holds the session lock: the render either completes or
    returns SPK_ERR_CANCELLED with nothing cached of its own (API_SPEC.md 2.9)."""
    img = synth.make_frame(1600, 1200, 26)
    with engine.open(img, PLAIN) as s:
        result = {}

        def work():
            try:
                result["px"] = s.render("full")
            except spk.SpkError as e:
                result["err"] = e

        t = threading.Thread(target=work)
        t.start()
        while t.is_alive():
            s.cancel()
            t.join(0.0005)
        if "err" in result:
            assert result["err"].status == -7
            with pytest.raises(spk.SpkError):
                s.negative("full")  # a cancelled render caches no negative of its own
        else:
            assert result["px"][0].shape == (1200, 1600, 4)
        px, _ = s.render("full")  # the session renders normally afterwards
        assert px.shape == (1200, 1600, 4)


def test_layers_and_invalidation(engine):
    """A shoot-layer change drops every cached negative (reprint develops again); a print-layer
    change keeps them (reprint reuses); {} changes nothing (API_SPEC.md 2.3)."""
    img = synth.make_frame(240, 160, 27)
    with engine.open(img, PLAIN) as s:
        s.render("full")
        r = s.set_params({"print_exposure": 1.2})
        assert r["invalidated"] == "print"
        _, f = s.render("full", reprint=True)
        assert f["reprint"] and f["negative_was_cached"]
        r = s.set_params({"exposure_compensation_ev": 0.5})
        assert r["invalidated"] == "shoot"
        _, f = s.render("full", reprint=True)
        assert not f["reprint"]
        r = s.set_params({})
        assert r["invalidated"] == "print" and len(r["params"]) == 48
