"""Drop-in classes on the GPU: Hamiltonian (evolve, integrate_decaying,
fast_evolve, diag), SpinSystem / MuonSpinSystem (term builders and the
GPU-assembled Hamiltonian), checked against the closed forms of the
reference's tests (muspinsim/tests/test_hamiltonian.py:10-128,
tests/test_spinsys.py:136-180,364-413) and the FP64 oracle.

Tolerances: closed forms 1e-12; oracle parity 1e-10 (P in [-0.5, 0.5]) or
1e-12 relative for matrices.
"""

import numpy as np
import pytest

from oracle import dynamics, physics

pytestmark = pytest.mark.gpu


def _herm(rng, n):
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return (X + X.conj().T) / 2


def test_hamiltonian_creation_and_diag(gpu_lib):
    from muspinsim_amd.hamiltonian import Hamiltonian
    from muspinsim_amd.spinop import SpinOperator
    H = Hamiltonian(np.array([[1, 0], [0, -1]]))
    assert H.dimension == (2,)
    with pytest.raises(ValueError):
        Hamiltonian(np.array([[1, 1], [0, 1]]))
    H = Hamiltonian(SpinOperator.from_axes().matrix)
    E, U = H.diag()
    assert np.abs(E - [-0.5, 0.5]).max() < 1e-15
    # columns are (1, -1)/sqrt 2 and (1, 1)/sqrt 2 up to a phase each (the
    # phase convention of LAPACK is not reproduced; observables do not see it)
    Vexp = np.array([[1.0, 1.0], [-1.0, 1.0]]) / 2 ** 0.5
    assert np.abs(np.abs(U.conj().T @ Vexp) - np.eye(2)).max() < 1e-14
    assert np.abs(H.basis_change(U).matrix - np.diag(E)).max() < 1e-14


def test_hamiltonian_closed_forms(gpu_lib):
    from muspinsim_amd.hamiltonian import Hamiltonian
    from muspinsim_amd.spinop import DensityOperator
    from muspinsim_amd.spinsys import MuonSpinSystem, SpinSystem
    ss = SpinSystem(["e"])
    ss.add_linear_term(0, [1, 0, 0])          # precession about x at 1 MHz
    H = ss.hamiltonian
    assert isinstance(H, Hamiltonian)
    rho0 = DensityOperator.from_vectors()    # along z
    t = np.linspace(0, 1, 100)
    ev = H.evolve(rho0, t, ss.operator({0: "z"}))
    assert ev.shape == (100, 1)
    assert np.abs(ev[:, 0] - 0.5 * np.cos(2 * np.pi * t)).max() < 1e-12
    avg = H.integrate_decaying(rho0, 1.0, ss.operator({0: "z"}))
    assert abs(avg[0] - 0.5 / (1.0 + 4 * np.pi ** 2)) < 1e-13
    states = H.evolve(rho0, np.linspace(0, 1, 4))
    assert len(states) == 4
    for k, tt in enumerate(np.linspace(0, 1, 4)):
        c, s = np.cos(np.pi * tt), np.sin(np.pi * tt)
        ref = np.array([[c * c, 1j * c * s], [-1j * c * s, s * s]])
        assert np.abs(states[k].matrix.toarray() - ref).max() < 1e-12
    ms = MuonSpinSystem(["mu", "e"])
    ms.add_linear_term(0, [1, 0, 0])
    ev = ms.hamiltonian.fast_evolve(ms.sigma_mu([0, 0, 1]), t, 2)
    assert ev.shape == (100,)
    assert np.abs(ev - 0.5 * np.cos(2 * np.pi * t)).max() < 1e-12


def test_hamiltonian_random_vs_oracle(gpu_lib):
    from muspinsim_amd.hamiltonian import Hamiltonian
    from muspinsim_amd.spinop import DensityOperator, SpinOperator
    rng = np.random.default_rng(21)
    for Is in ([0.5, 0.5, 0.5], [0.5, 1.0, 0.5, 0.5]):
        dims = tuple(int(2 * I + 1) for I in Is)
        n = int(np.prod(dims))
        Hm = _herm(rng, n) * 5
        H = Hamiltonian(Hm, dim=dims)
        rho = DensityOperator.from_vectors(Is, rng.normal(size=(len(Is), 3)),
                                           list(rng.uniform(0, 1, len(Is))))
        ops = [SpinOperator.from_axes(Is, "x" + "0" * (len(Is) - 1)),
               SpinOperator.from_axes(Is, "z" * len(Is))]
        t = np.linspace(0, 3, 57)
        ev = H.evolve(rho, t, ops)
        ref = dynamics.evolve(Hm, rho.dense(), t, [o.dense() for o in ops])
        assert np.abs(ev - ref).max() <= 1e-10
        iv = H.integrate_decaying(rho, 2.19703, ops)
        ref = dynamics.integrate_decaying(Hm, rho.dense(), 2.19703, [o.dense() for o in ops])
        assert np.abs(iv - ref).max() <= 1e-12
        d = n // 2
        sig = physics.sigma_mu([0.3, -0.5, 0.81])
        fe = H.fast_evolve(sig, t, d)
        ref = dynamics.fast_evolve(Hm, sig, t, d)
        assert np.abs(fe - ref).max() <= 1e-10


def test_hamiltonian_argument_errors(gpu_lib):
    from muspinsim_amd.hamiltonian import Hamiltonian
    from muspinsim_amd.spinop import DensityOperator, SpinOperator
    H = Hamiltonian(SpinOperator.from_axes(0.5, "z").matrix)
    rho = DensityOperator.from_vectors()
    sx = SpinOperator.from_axes()
    with pytest.raises(TypeError):
        H.evolve(sx, np.linspace(0, 1, 3), sx)
    with pytest.raises(ValueError):
        H.evolve(rho, np.zeros((2, 2)), sx)
    with pytest.raises(ValueError):
        H.evolve(rho, np.linspace(0, 1, 3), [np.eye(2)])
    with pytest.raises(ValueError):
        H.integrate_decaying(rho, 0.0, sx)
    with pytest.raises(ValueError):
        H.integrate_decaying(rho, 1.0, [])
    with pytest.raises(TypeError):
        H.integrate_decaying(sx, 1.0, sx)
    with pytest.raises(ValueError):
        H.fast_evolve(np.eye(2), np.array([0.0, 1.0]), 3)


def test_spinsystem_hamiltonian_vs_oracle(gpu_lib):
    """Zeeman + dipolar + quadrupolar + hyperfine terms assembled on the GPU."""
    from muspinsim_amd.spinsys import MuonSpinSystem
    ss = MuonSpinSystem(["mu", "e", ("H", 1), ("N", 14)])
    ss.add_zeeman_term(0, [0.01, 0.02, 0.5])
    ss.add_hyperfine_term(0, np.diag([300.0, 310.0, 290.0]))
    ss.add_hyperfine_term(2, 5.0 * np.eye(3))
    ss.add_dipolar_term(0, 2, [0.5, 1.0, -1.2])
    efg = np.array([[1.0, 0.2, 0.0], [0.2, -0.4, 0.1], [0.0, 0.1, -0.6]])
    ss.add_quadrupolar_term(3, efg)
    H = ss.hamiltonian.matrix
    H = H.toarray() if hasattr(H, "toarray") else H
    Is = list(ss.Is)
    g = list(ss.gammas)
    terms = [("linear", 0, 0, np.array([0.01, 0.02, 0.5]) * g[0]),
             ("bilinear", 0, 1, np.diag([300.0, 310.0, 290.0])),
             ("bilinear", 2, 1, 5.0 * np.eye(3)),
             ("bilinear", 0, 2, physics.dipolar_tensor(g[0], g[2], [0.5, 1.0, -1.2])),
             ("bilinear", 3, 3, physics.quadrupolar_tensor(ss.Q(3), 1.0, efg))]
    ref = physics.hamiltonian(Is, terms)
    assert np.abs(H - ref).max() <= 1e-12 * np.abs(ref).max()
    # the host operator algebra of every term sums to the same matrix
    host = sum(t.matrix.toarray() for t in ss.terms)
    assert np.abs(host - ref).max() <= 1e-12 * np.abs(ref).max()


def test_spinsystem_known_hamiltonian(gpu_lib):
    """Exact 4 x 4 matrix of mu-e with a bilinear term S_mu . diag(1, 2, 3) . S_e
    (a bilinear product of spin-1/2 matrices)."""
    from muspinsim_amd.spinsys import MuonSpinSystem
    ss = MuonSpinSystem(["mu", "e"])
    ss.add_bilinear_term(0, 1, np.diag([1.0, 2.0, 3.0]))
    H = ss.hamiltonian.matrix
    H = H.toarray() if hasattr(H, "toarray") else H
    ref = np.array([[0.75, 0, 0, -0.25], [0, -0.75, 0.75, 0], [0, 0.75, -0.75, 0],
                    [-0.25, 0, 0, 0.75]])
    assert np.abs(H - ref).max() < 1e-15
