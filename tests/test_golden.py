"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py
from the FP64 oracle): the oracle must keep reproducing them (CPU), and the
HIP path must match them to the plain-FP64 parity bar (GPU)."""

import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["zero_field_n16", "alc_n192", "tf_n64", "fit_n64", "dissipative_n8"]


def _case(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    builder, kw = mod.CASES[name]
    data = np.load(os.path.join(HERE, f"{name}.npz"))
    assert str(data["builder"]) == builder.__name__
    return builder(**kw), data["results"]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    from oracle import pipeline
    (systems, batch), ref = _case(name)
    res = pipeline.run(systems, batch)
    assert np.abs(res - ref).max() <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(gpu_lib, name):
    from muspinsim_amd.engine import SpinEngine
    (systems, batch), ref = _case(name)
    res = SpinEngine("cuda:0").run(systems, batch)
    assert np.abs(res - ref).max() <= 1e-10
