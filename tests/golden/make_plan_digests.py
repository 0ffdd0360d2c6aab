"""Writes plan_digests.json: the SHA-256 of the spkx_plan bytes of the benchmark configurations on
the synthetic resources (spektralab_engine_amd.synth at L = 192).

The digests pin the float64 set-up: any change to what a render is told to compute (a constant,
a table, a narrowing) changes a digest.  They were recorded when the set-up was rewritten into
fx_model / film_model / cam16 / print_model / plan, after checking that every plan byte (header
and the used part of every table) equals the previous set-up's for these configurations and 19
others covering every stage.

    python tests/golden/make_plan_digests.py
"""
import hashlib
import json
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

import spektralab_engine_amd as spk  # noqa: E402
from spektralab_engine_amd import synth  # noqa: E402
from conftest import PLAIN, GRAIN_HALATION, MIX3_OVERRIDES, delta  # noqa: E402

SEED = 0x4ABEA221  # the first film seed of a fresh pipeline

# name -> (delta, overrides, width, height)
CASES = {
    "plain_512x512": (PLAIN, None, 512, 512),
    "grain_halation_600x400": (GRAIN_HALATION, MIX3_OVERRIDES, 600, 400),
    "grain_halation_6000x4000": (dict(GRAIN_HALATION, film_format_mm=35.0), dict(MIX3_OVERRIDES, fir_crossover=16.5), 6000, 4000),
    "diff_2048x1365": (delta(grain_active=True, halation_active=True, film_format_mm=35.0),
                          dict(MIX3_OVERRIDES, fir_crossover=16.5), 2048, 1365),
}


def digests(resources):
    out = {}
    for name, (d, ov, w, h) in CASES.items():
        plan = spk.host_plan(resources, d, w, h, SEED, ov)
        out[name] = hashlib.sha256(bytes(plan)).hexdigest()
    return out


if __name__ == "__main__":
    with tempfile.TemporaryDirectory() as tmp:
        res = synth.make_resources(Path(tmp) / "synth", 192)
        (HERE / "plan_digests.json").write_text(json.dumps(digests(res), indent=1, sort_keys=True) + "\n")
