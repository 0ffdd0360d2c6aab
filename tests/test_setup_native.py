"""The engine's host set-up functions through a native known-answer program (CPU, no GPU):
tests/native/setup_kats.cpp is compiled against the set-up sources of the library and run on the
synthetic resources folder.  What it checks is stated in that file."""
import os
import subprocess
from pathlib import Path

import pytest

from spektralab_engine_amd import synth

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "spektralab_engine_amd" / "csrc" / "host"
SOURCES = ["json", "params", "resources", "fx_model", "film_model", "cam16", "print_model", "plan"]


@pytest.fixture(scope="module")
def kats_binary(tmp_path_factory):
    out = tmp_path_factory.mktemp("kats") / "setup_kats"
    cmd = ["g++", "-std=c++20", "-O1", "-ffp-contract=off", "-fopenmp", "-I", str(ROOT / "include"), "-I", str(HOST),
           str(ROOT / "tests" / "native" / "setup_kats.cpp")] + [str(HOST / f"{s}.cpp") for s in SOURCES] + \
          ["-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def test_setup_known_answers(kats_binary, resources):
    r = subprocess.run([str(kats_binary), str(resources), synth.PAPER_STOCK], capture_output=True, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
