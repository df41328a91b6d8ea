"""Input files run end to end on the GPU: MuSpinInput -> MuSpinConfig ->
ExperimentRunner.run against the FP64 oracle on the same instances (and the
results function applied on the host), against closed forms for a bare
muon, and FittingRunner(MuSpinInput) recovering a relaxation rate and a
results-function scale from synthetic data.

Tolerance: 1e-10 absolute on the polarisation (the parity bar of the
engine); fitted values within 1e-3 relative of the truth on noiseless data.
"""

import glob
import os

import numpy as np
import pytest

from oracle import pipeline as opipe

PARITY = 1e-10
INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")
FILES = sorted(os.path.basename(p)[:-3] for p in glob.glob(os.path.join(INPUTS, "*.in")))


def _runner(name, **kw):
    from muspinsim_amd.experiment import ExperimentRunner
    from muspinsim_amd.input import MuSpinInput
    with open(os.path.join(INPUTS, name + ".in")) as fh:
        return ExperimentRunner(MuSpinInput(fh), **kw)


def _oracle_results(er):
    cfg = er.config
    systems, batch = er.instance_batch()
    orc = opipe.run(systems, batch)
    shape = cfg.results.shape
    if cfg.x_axis == "t":
        ref = orc.reshape(shape)
    else:
        ref = orc.reshape(shape + (orc.shape[1],))
        ref = ref.mean(axis=-1) if cfg.time_averaged else ref[..., 0]
    return er.apply_results_function(ref, {})


@pytest.mark.gpu
@pytest.mark.parametrize("name", FILES)
def test_input_file_matches_oracle(gpu_lib, name):
    print("input file", name, flush=True)
    er = _runner(name)
    res = er.run()
    ref = _oracle_results(er)
    assert res.shape == ref.shape == er.config.results.shape
    assert np.abs(res - ref).max() <= PARITY


@pytest.mark.gpu
@pytest.mark.parametrize("name,f", [
    ("muon_precession", lambda t: 0.5 * np.cos(4 * np.pi * t)),
    ("muon_dissipation", lambda t: 0.5 * np.cos(4 * np.pi * t) * np.exp(-5.0 * t)),
    ("muon_local_zeeman", lambda t: 0.5 * np.cos(40 * np.pi * t)),
])
def test_input_file_closed_forms(gpu_lib, name, f):
    er = _runner(name)
    res = er.run()
    assert np.abs(res - f(np.asarray(er.config.x_axis_values))).max() <= PARITY


def _decay_input(tmp_path, method, g_true=3.0, extra=""):
    t = np.linspace(0, 2.0, 40)
    y = 0.5 * np.cos(2 * np.pi * t) * np.exp(-g_true * t)
    path = tmp_path / "decay.dat"
    np.savetxt(path, np.stack([t, y], axis=1))
    return (f"name\n    fit_decay\nspins\n    mu\nfield\n    1.0/muon_gyr\n"
            f'fitting_data\n    load("{path}")\nfitting_variables\n    g 0.5 0.0 inf\n'
            f"fitting_method\n    {method}\nfitting_tolerance\n    1e-8\n"
            f"dissipation 1\n    g\n{extra}")


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["nelder-mead", "lbfgs", "least-squares"])
def test_fitting_runner_from_input(gpu_lib, tmp_path, method):
    from io import StringIO
    from muspinsim_amd.fitting import FittingRunner
    from muspinsim_amd.input import MuSpinInput
    inp = MuSpinInput(StringIO(_decay_input(tmp_path, method)))
    assert inp.fitting_info["fit"] and not inp.fitting_info["single_simulation"]
    fr = FittingRunner(inp)
    sol = fr.run()
    assert abs(sol.x[0] - 3.0) <= 3e-3
    fr.write_report(path=str(tmp_path))
    report = open(tmp_path / "fit_decay_fit_report.txt").read()
    assert "Values found for fitting variables" in report and "g = " in report
    fr.write_data(path=str(tmp_path))
    d = np.loadtxt(tmp_path / "fit_decay.dat")
    # the file's own time axis (the default range(0, 10, 101)) with the fitted g
    t = np.linspace(0, 10, 101)
    assert np.allclose(d[:, 0], t)
    assert np.abs(d[:, 1] - 0.5 * np.cos(2 * np.pi * t) * np.exp(-sol.x[0] * t)).max() <= 1e-9


@pytest.mark.gpu
def test_fitting_single_simulation_results_function(gpu_lib, tmp_path):
    """A variable used only by the results function: one simulation serves
    every fitting step."""
    from io import StringIO
    from muspinsim_amd.fitting import FittingRunner
    from muspinsim_amd.input import MuSpinInput
    t = np.linspace(0, 1.0, 30)
    data = np.stack([t, 1.7 * 0.5 * np.cos(2 * np.pi * t)], axis=1)
    path = tmp_path / "scaled.dat"
    np.savetxt(path, data)
    text = (f"spins\n    mu\nfield\n    1.0/muon_gyr\nfitting_data\n    load(\"{path}\")\n"
            "fitting_variables\n    A 1.0 0.0 5.0\nfitting_method\n    least-squares\n"
            "fitting_tolerance\n    1e-10\nresults_function\n    A*y\n")
    inp = MuSpinInput(StringIO(text))
    assert inp.fitting_info["single_simulation"]
    fr = FittingRunner(inp)
    sol = fr.run()
    assert fr.n_evaluations == 1
    assert abs(sol.x[0] - 1.7) <= 1e-6
