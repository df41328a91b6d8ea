"""Celio's Trotter method on the GPU (msp_celio_evolve) against the FP64
oracle (oracle/celio.py, itself pinned to the reference's known answers in
tests/test_celio_host.py), the reference's known answers and closed forms
(tests/test_celio.py, tests/test_spinsys.py:192-300,
tests/test_experiment.py:340-470), and the error cases of the reference.

The random-phase states are seeded (muspinsim_amd.celio.seed) and the oracle
draws the same phases, so the comparison is deterministic.  Tolerance
1e-10 on the polarisation; the factors come from msp_eigh instead of a Pade
expm, which accounts for ~1e-14 per factor."""

import os
from io import StringIO

import numpy as np
import pytest

from oracle import celio as oc
from oracle import physics

PARITY = 1e-10
HERE = os.path.dirname(os.path.abspath(__file__))
SZ = np.array([[1, 0], [0, -1]], dtype=complex)


def _oracle_terms(system):
    out = []
    for t in system.terms:
        idx = [int(i) for i in np.atleast_1d(t.indices)]
        T = np.asarray(t.tensor, dtype=float)
        out.append(("linear", idx[0], idx[0], T) if len(idx) == 1 else
                   ("bilinear", idx[0], idx[1], T))
    return out


@pytest.mark.gpu
def test_trotter_factors_known_answers(gpu_lib):
    from muspinsim_amd.spinsys import SpinSystem
    ssys = SpinSystem(["mu", "e"], celio_k=10)
    ssys.add_linear_term(0, [1, 0, 1])
    ssys.add_bilinear_term(0, 1, np.eye(3))
    f = ssys.hamiltonian._calc_trotter_evol_op_contribs(1, True)
    assert len(f) == 2
    assert np.allclose(f[0].matrix, [[0.902916796 - 0.30392537j, -0.30392537j],
                                     [-0.30392537j, 0.902916796 + 0.30392537j]], atol=1e-8)
    assert f[0].other_dim == 2 and np.array_equal(f[0].indices, [0, 1, 2, 3])
    full = ssys.hamiltonian._calc_trotter_evol_op_contribs(1, False)
    E = (full[0] @ full[1]).toarray()
    ref = [[0.84425599 - 0.44143085j, -0.09276182 - 0.01469203j, 0.04521741 - 0.28549151j, 0],
           [0, 0.89336999 - 0.15115734j, -0.049114 - 0.29027351j, -0.0475444 - 0.30018354j],
           [-0.0475444 - 0.30018354j, 0.13640963 - 0.26088945j, 0.80293516 + 0.41982569j, 0],
           [0, 0.04521741 - 0.28549151j, -0.09276182 - 0.01469203j, 0.93934479 + 0.15893624j]]
    assert np.allclose(E, ref, atol=1e-8)
    e = SpinSystem(["mu", "e"], celio_k=10)
    e.add_linear_term(0, [1, 0, 0])
    c = e.hamiltonian._calc_trotter_evol_op_contribs(1, False)
    cs, sn = 0.95105652, -0.30901699j
    assert np.allclose(c[0].toarray(), [[cs, 0, sn, 0], [0, cs, 0, sn], [sn, 0, cs, 0],
                                        [0, sn, 0, cs]])


@pytest.mark.gpu
def test_closed_forms(gpu_lib):
    from muspinsim_amd.spinop import DensityOperator
    from muspinsim_amd.spinsys import MuonSpinSystem, SpinSystem
    t = np.linspace(0, 1, 100)
    s = MuonSpinSystem(["mu", "e"], celio_k=10)
    s.add_linear_term(0, [1, 0, 0])
    r = s.hamiltonian.fast_evolve(s.sigma_mu([0, 0, 1]), t, 10)
    assert np.abs(r - 0.5 * np.cos(2 * np.pi * t)).max() <= PARITY
    e = SpinSystem(["e"], celio_k=10)
    e.add_linear_term(0, [1, 0, 0])
    ev = e.hamiltonian.evolve(DensityOperator.from_vectors(), t, [e.operator({0: "z"})])
    assert np.abs(ev[:, 0] - 0.5 * np.cos(2 * np.pi * t)).max() <= PARITY
    with pytest.raises(ValueError):
        e.hamiltonian.evolve(DensityOperator.from_vectors(), t, [])
    with pytest.raises(NotImplementedError):
        e.hamiltonian.integrate_decaying(DensityOperator.from_vectors(), 1.0,
                                         e.operator({0: "z"}))


def _system():
    from muspinsim_amd.spinsys import MuonSpinSystem
    rng = np.random.default_rng(4)
    s = MuonSpinSystem(["mu", ("H", 2), "e"], celio_k=7)
    A = rng.normal(0, 3.0, (3, 3))
    s.add_hyperfine_term(0, 80 * np.eye(3) + A + A.T)
    s.add_dipolar_term(0, 1, [0.9, -0.4, 1.2])
    s.add_quadrupolar_term(1, np.diag([1.0, -0.4, -0.6]) * 0.3)
    s.add_hyperfine_term(1, np.eye(3) * 2.5)
    s.add_zeeman_term(0, [0.0, 0.001, 0.002])
    return s


@pytest.mark.gpu
def test_fast_evolve_matches_oracle(gpu_lib):
    from muspinsim_amd import celio
    s = _system()
    t = np.linspace(0, 0.4, 30)
    sig = s.sigma_mu([0.6, 0.0, 0.8]).toarray()
    celio.seed(21)
    got = s.hamiltonian.fast_evolve(sig, t, 5)
    ref = oc.fast_evolve(s.Is, _oracle_terms(s), 7, t, sig,
                         np.random.default_rng(21).random((5, 6)))
    assert np.abs(got - ref).max() <= PARITY


@pytest.mark.gpu
def test_evolve_matches_oracle(gpu_lib):
    from muspinsim_amd.spinop import DensityOperator
    s = _system()
    t = np.linspace(0, 0.4, 25)
    rho0 = DensityOperator.from_vectors(s.Is, [[0, 0, 1.0], [1.0, 0, 0], [0, 0, 1.0]],
                                        [0, 0.3, 0.0])
    op = s.muon_operator([0, 0, 1.0])
    got = s.hamiltonian.evolve(rho0, t, [op])[:, 0]
    ref = oc.evolve(s.Is, _oracle_terms(s), 7, t, rho0.matrix.toarray()
                    if hasattr(rho0.matrix, "toarray") else np.asarray(rho0.matrix),
                    op.matrix.toarray() if hasattr(op.matrix, "toarray") else op.matrix)
    assert np.abs(got - ref).max() <= PARITY


def _run_text(text):
    from muspinsim_amd.experiment import ExperimentRunner
    from muspinsim_amd.input import MuSpinInput
    er = ExperimentRunner(MuSpinInput(StringIO(text)))
    return er, er.run()


@pytest.mark.gpu
def test_evolve_general_state_and_spin1_operator(gpu_lib):
    """A density matrix that is not a product of one-spin states (through
    its eigenvectors) and an operator on the spin-1 nucleus."""
    s = _system()
    t = np.linspace(0, 0.3, 12)
    rng = np.random.default_rng(8)
    G = rng.normal(size=(12, 12)) + 1j * rng.normal(size=(12, 12))
    rho = G @ G.conj().T
    rho /= np.trace(rho)
    from muspinsim_amd.spinop import DensityOperator
    rho0 = DensityOperator(rho, dim=(2, 3, 2))
    op = s.operator({1: "z"})
    got = s.hamiltonian.evolve(rho0, t, [op])[:, 0]
    ref = oc.evolve(s.Is, _oracle_terms(s), 7, t, rho,
                    op.matrix.toarray() if hasattr(op.matrix, "toarray") else op.matrix)
    assert np.abs(got - ref).max() <= PARITY


@pytest.mark.gpu
def test_experiment_celio_cases(gpu_lib):
    """The reference's tests/test_experiment.py:340-470 cases."""
    with pytest.raises(ValueError):                                   # no terms
        _run_text("spins\n    e mu\ntime\n    range(0, 10)\ncelio\n    10\n")
    with pytest.raises(NotImplementedError):                          # dissipation
        _run_text("spins\n    e mu\ntime\n    range(0, 10)\nzeeman 2\n    0 0 1.0/muon_gyr\n"
                  "celio\n    10\ndissipation 1\n    0.5\n")
    er, res = _run_text("spins\n    e mu\ntime\n    range(0, 10)\nzeeman 2\n"
                        "    0 0 1.0/muon_gyr\ncelio\n    10\n")
    t = np.asarray(er.config.x_axis_values)
    assert np.abs(res - 0.5 * np.cos(2 * np.pi * t)).max() <= PARITY
    with pytest.raises(ValueError):                                   # muon not first
        _run_text("spins\n    e mu\ntime\n    range(0, 10)\nzeeman 2\n    0 0 1.0/muon_gyr\n"
                  "celio\n    10 8\n")
    er, res = _run_text("spins\n    mu e\ntime\n    range(0, 10)\nzeeman 1\n"
                        "    0 0 1.0/muon_gyr\ncelio\n    10 8\n")
    assert er.config.celio_k == 10 and er.config.celio_averages == 8
    assert np.abs(res - 0.5 * np.cos(2 * np.pi * t)).max() <= PARITY
    with pytest.raises(ValueError):                                   # averages need T = inf
        _run_text("spins\n    mu e\ntime\n    range(0, 10)\nzeeman 2\n    0 0 1.0/muon_gyr\n"
                  "celio\n    10 8\ntemperature\n    1.0\n")


def _experiment_vs_oracle(text, seed):
    """ExperimentRunner with Celio against the oracle over the same
    configurations and seeded phases (zero field, T = inf)."""
    from muspinsim_amd import celio
    from muspinsim_amd.experiment import ExperimentRunner
    from muspinsim_amd.input import MuSpinInput
    er = ExperimentRunner(MuSpinInput(StringIO(text)))
    cfg = er.config
    celio.seed(seed)
    res = er.run()
    rng = np.random.default_rng(seed)
    s = er.system
    t = np.asarray(cfg[0].t, dtype=float)
    half = int(np.prod(s.dimension)) // 2
    ref = np.zeros_like(res)
    for k in range(len(cfg)):
        snap = cfg[k]
        w = er.load_config(snap)
        sig = physics.sigma_mu(er.p)
        terms = _oracle_terms(s)
        if np.any(er.B != 0):
            terms += [("linear", i, i, er.B * g) for i, g in enumerate(s.gammas)]
        row = oc.fast_evolve(s.Is, terms, cfg.celio_k, t, sig,
                             rng.random((cfg.celio_averages, half))) * w
        ref += row / cfg.avg_N
    return res, ref


@pytest.mark.gpu
def test_experiment_n8192_matches_oracle(gpu_lib):
    """mu + 4 V (n = 8192, one state per workgroup in LDS), polarisation
    averaged over three directions, two random states each."""
    with open(os.path.join(HERE, "golden", "celio", "mu_4V.in")) as fh:
        text = fh.read()
    res, ref = _experiment_vs_oracle(text, 5)
    print("n=8192 max dev", np.abs(res - ref).max(), flush=True)
    assert np.abs(res - ref).max() <= PARITY


@pytest.mark.gpu
def test_experiment_n16384_global_states(gpu_lib):
    """mu + 4 V + e (n = 16384): states kept in the workspace instead of LDS,
    with a field (Zeeman factors on every spin)."""
    with open(os.path.join(HERE, "golden", "celio", "mu_4V.in")) as fh:
        text = fh.read()
    text = text.replace("spins\n    mu V V V V", "spins\n    mu V V V V e")
    text = text.replace("celio\n    6 2", "celio\n    3 1\nhyperfine 1 6\n    90 0 0\n    0 90 0\n"
                        "    0 0 95\nfield\n    0.002")
    text = text.replace("range(0, 1.0, 5)", "range(0, 0.2, 3)")
    res, ref = _experiment_vs_oracle(text, 9)
    print("n=16384 max dev", np.abs(res - ref).max(), flush=True)
    assert np.abs(res - ref).max() <= PARITY
