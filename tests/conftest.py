"""Shared fixtures: the synthetic resources folder, the library, the oracle, a GPU engine."""
from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

from spektralab_engine_amd import synth  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def resources(tmp_path_factory):
    d = Path(os.environ.get("SPK_SYNTH_DIR") or tmp_path_factory.mktemp("spk_synth_L192"))
    return synth.make_resources(d, 192)


@pytest.fixture(scope="session")
def lib():
    import spektralab_engine_amd as spk
    return spk.load()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


# The render configurations: only stages the benchmark workloads use are on.
BASE_DELTA = dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK, auto_exposure=False,
                  dir_couplers_active=False, glare_active=False)


def delta(**kw):
    d = dict(BASE_DELTA)
    d.update(kw)
    return d


PLAIN = delta(grain_active=False, halation_active=False)
GRAIN_HALATION = delta(grain_active=True, halation_active=True)
MIX3_OVERRIDES = {"halation_mixture": [[2.0, 0.6], [8.0, 0.3], [32.0, 0.1]], "halation_scatter": False}


@pytest.fixture(scope="session")
def engine(resources):
    import spektralab_engine_amd as spk
    e = spk.Engine(resources)
    yield e
    e.close()
