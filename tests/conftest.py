import importlib
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")


def load_oracle():
    spec = importlib.util.spec_from_file_location("es_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def orc():
    return load_oracle()


@pytest.fixture(scope="session")
def es():
    return importlib.import_module("encode-score_amd")


@pytest.fixture(scope="session")
def gpu(es):
    es.init(0)
    return es


def staircase(h, w, step_px=5, base=100, vertical=False):
    """A ramp of 1-code steps `step_px` wide: flat (masked) everywhere with neighbouring values in every window."""
    x = np.arange(w)[None, :].repeat(h, 0) if not vertical else np.arange(h)[:, None].repeat(w, 1)
    return (base + x // step_px).astype(np.uint16)


def banded_mosaic(es, w, h, seed, frame=0, **kw):
    d = es.SynthDesc(kind=1, seed=seed, full_w=w, full_h=h, w=w, h=h, depth=10, grain_q8=256, **kw)
    return es.synth_frame_host(d, frame)
