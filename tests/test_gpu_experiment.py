"""ExperimentRunner / MuSpinConfig on the GPU: the known answers of the
reference's experiment tests (muspinsim/tests/test_experiment.py:13-291),
the batched run() against the per-snapshot run_single loop (the reference's
own algorithm through the operator classes) and against the FP64 oracle.

Tolerances: closed forms 1e-12; batched vs per-snapshot and vs oracle 1e-10.
"""

import numpy as np
import pytest
import scipy.constants as cnst

from oracle import pipeline as opipe

pytestmark = pytest.mark.gpu

PARITY = 1e-10


def _runner(system, **kw):
    from muspinsim_amd.config import MuSpinConfig
    from muspinsim_amd.experiment import ExperimentRunner
    return ExperimentRunner(MuSpinConfig(system, **kw))


def test_create_hz_and_rho0(gpu_lib):
    from muspinsim_amd.constants import ELEC_GAMMA as ge, MU_GAMMA as gmu
    from muspinsim_amd.hamiltonian import Hamiltonian
    from muspinsim_amd.spinsys import MuonSpinSystem
    s = MuonSpinSystem(["mu"])
    s.add_zeeman_term(0, [0, 0, 1])
    er = _runner(s)
    er.B = [0, 0, 1.0]
    assert np.abs(er.Hsys.matrix.toarray() - np.diag([0.5 * gmu, -0.5 * gmu])).max() < 1e-12
    assert np.abs(er.Hz.matrix.toarray() - er.Hsys.matrix.toarray()).max() < 1e-12
    er = _runner(MuonSpinSystem(["e", "mu"]))
    er.B = [0, 0, 1.0]
    assert er.system.elec_indices == {0} and er.system.muon_index == 1
    ref = np.diag([0.5 * (ge + gmu), 0.5 * (ge - gmu), 0.5 * (-ge + gmu), 0.5 * (-ge - gmu)])
    assert np.abs(er.Hz.matrix.toarray() - ref).max() < 1e-9
    assert isinstance(er.Hz, Hamiltonian)
    assert abs(er.rho0.expectation(er.system.operator({1: "x"})) - 0.5) < 1e-15
    assert abs(er.rho0.expectation(er.system.operator({0: "z"}))) < 1e-15


def test_rho0_thermal(gpu_lib):
    from muspinsim_amd.constants import ELEC_GAMMA as ge
    from muspinsim_amd.spinsys import MuonSpinSystem
    er = _runner(MuonSpinSystem(["e", "mu"]))
    ref = np.array([[1, 1, 0, 0], [1, 1, 0, 0], [0, 0, 1, 1], [0, 0, 1, 1]]) * 0.25
    assert np.abs(er.rho0.matrix.toarray() - ref).max() < 1e-15
    T = 100
    er.B = [0, 0, 2.0e-6 * cnst.k * T / (ge * cnst.h)]
    er.T = T
    er.p = [0, 0, 1.0]
    Z = np.exp([-1, 1])
    Z /= np.sum(Z)
    assert np.abs(np.diag(er.rho0.matrix.toarray()) - [Z[0], 0, Z[1], 0]).max() < 1e-14


def test_run_known_answers(gpu_lib):
    from muspinsim_amd.constants import MU_GAMMA as gmu, MU_TAU as tau
    from muspinsim_amd.spinsys import MuonSpinSystem
    t = np.linspace(0, 10, 100)
    res = _runner(MuonSpinSystem(["e", "mu"]), time=t).run()
    assert res.shape == (100,) and np.abs(res - 0.5).max() < 1e-14
    s = MuonSpinSystem(["e", "mu"])
    s.add_zeeman_term(1, [0, 0, 1.0 / gmu])
    er = _runner(s, time=t)
    res = er.run()
    assert np.abs(res - 0.5 * np.cos(2 * np.pi * er.config.x_axis_values)).max() < 1e-12
    for kw in ("field", "intrinsic_field"):
        res = _runner(s, time=t, y_axis="integral", x_axis=kw, **{kw: [0.0, 1.0]}).run()
        assert res.shape == (2,)
        assert abs(res[0] - 0.5 / (1.0 + 4 * np.pi ** 2 * tau ** 2)) < 1e-13


def test_results_function_and_file_ranges(gpu_lib):
    from muspinsim_amd.config import MuSpinConfig
    from muspinsim_amd.constants import MU_GAMMA as gmu
    from muspinsim_amd.experiment import ExperimentRunner
    from muspinsim_amd.spinsys import MuonSpinSystem
    t = np.linspace(0, 10, 100)
    res = _runner(MuonSpinSystem(["e", "mu"]), time=t,
                  results_function=lambda x, y: 2 * y).run()
    assert np.abs(res - 1.0).max() < 1e-14
    s = MuonSpinSystem(["e", "mu"])
    s.add_zeeman_term(1, [0, 0, 1.0 / gmu])
    res = _runner(s, time=t, results_function=lambda x, y: 2 * y / np.e).run()
    assert np.abs(res - np.cos(2 * np.pi * t) / np.e).max() < 1e-12
    cfg = MuSpinConfig(MuonSpinSystem(["e", "mu"]), time=t, results_function=lambda x, y: 2 * y)
    res = ExperimentRunner(cfg, variables={"g": 1.0}).run()
    assert np.abs(res - 0.5).max() < 1e-14
    res = _runner(MuonSpinSystem(["e", "mu"]), time=t, field=[0.0, 10.0],
                  temperature=[np.inf, 0.0]).run()
    assert res.shape == (2, 2, 100)


def _powder_system():
    from muspinsim_amd.spinsys import MuonSpinSystem
    s = MuonSpinSystem(["mu", "e", ("H", 1)])
    s.add_hyperfine_term(0, np.array([[250.0, 10, 0], [10, 260, 5], [0, 5, 240]]))
    s.add_hyperfine_term(2, np.diag([5.0, 6.0, 4.0]))
    s.add_dipolar_term(0, 2, [0.9, -0.5, 1.1])
    return s


@pytest.mark.parametrize("case", ["zf_powder", "lf_file_range", "finite_T", "alc_integral",
                                  "dissipative", "time_averaged"])
def test_batched_run_matches_snapshots_and_oracle(gpu_lib, case):
    from muspinsim_amd.orientations import zcw
    from muspinsim_amd.spinsys import MuonSpinSystem
    s = _powder_system()
    t = np.linspace(0, 5, 64)
    ori = zcw(13)
    kw = dict(time=t, orientation=ori, polarization="longitudinal")
    if case == "lf_file_range":
        kw.update(field=[0.0, 0.01])
    elif case == "finite_T":
        kw.update(field=[[0.0, 0.002, 0.01]], temperature=2.0)
    elif case == "alc_integral":
        kw.update(field=np.linspace(0.0, 0.2, 5), x_axis="field", y_axis="integral")
    elif case == "dissipative":
        kw.update(field=0.005, dissipation={1: 20.0}, orientation=ori[:5])
    elif case == "time_averaged":
        kw.update(field=[0.0, 0.01, 0.02], x_axis="field", average_axes=["orientation", "time"])
    er = _runner(s, **kw)
    res = er.run()
    cfg = er.config
    ref = np.zeros(cfg.results.shape)
    cfg.results = ref
    for k in range(len(cfg)):
        snap = cfg[k]
        cfg.store_time_slice(snap.id, er.run_single(snap))
    assert np.abs(res - cfg.results).max() <= PARITY
    systems, batch = er.instance_batch()
    orc = opipe.run(systems, batch)
    assert np.abs(res.reshape(-1) - _fold(orc, cfg).reshape(-1)).max() <= PARITY


def _fold(orc, cfg):
    """Oracle (n_slots, width) -> results shape of the configuration."""
    shape = cfg.results.shape
    if cfg.x_axis == "t":
        return orc.reshape(shape)
    out = orc.reshape(shape + (orc.shape[1],))
    return out.mean(axis=-1) if cfg.time_averaged else out[..., 0]
