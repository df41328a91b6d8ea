"""Pins the oracle where the reference has no literal: identities the
mathematics fixes, chosen so that a dropped term, a wrong sign or a
transposed operand fails.

- evolve against brute-force propagation exp(-2 pi i H t) rho0 exp(+...)
  (scipy.linalg.expm), for random n <= 8 systems;
- fast_evolve (cosine sum) against evolve at T = inf, muon first;
- integrate_decaying against numerical quadrature of exp(-t/tau) evolve(t);
- Lindblad evolve against expm(2 pi L t) on vec(rho0);
- ExperimentRunner-level closed forms (tests/test_experiment.py:117-202,
  581-640) through oracle.pipeline.
"""

import types

import numpy as np
import pytest
import scipy.constants as cnst
from scipy.linalg import expm

from oracle import dynamics, orient, physics, pipeline, spin


def _random_system(rng, Is):
    terms = []
    n = len(Is)
    for i in range(n):
        terms.append(("linear", i, i, rng.normal(size=3)))
        for j in range(i + 1, n):
            terms.append(("bilinear", i, j, rng.normal(size=(3, 3))))
        if Is[i] >= 1:
            Q = rng.normal(size=(3, 3))
            terms.append(("bilinear", i, i, Q + Q.T))
    return terms


@pytest.mark.parametrize("Is", [[0.5, 0.5], [0.5, 1.0], [0.5, 0.5, 0.5]])
def test_evolve_matches_expm(Is):
    rng = np.random.default_rng(11)
    H = physics.hamiltonian(Is, _random_system(rng, Is))
    r0 = physics.rho0(Is, [1.0] * len(Is), 0, [0.1, 0.2, 0.3], [0.3, 0.1, 0.9], 0.02)
    O = physics.muon_operator(Is, 0, [0.2, -0.5, 0.8])
    times = np.linspace(0, 2.0, 9)
    ev = dynamics.evolve(H, r0, times, [O])[:, 0]
    for k, t in enumerate(times):
        Ut = expm(-2j * np.pi * H * t)
        ref = np.trace(Ut @ r0 @ Ut.conj().T @ O)
        assert abs(ev[k] - ref) < 1e-12


def test_fast_evolve_equals_evolve_at_infinite_T():
    rng = np.random.default_rng(5)
    Is = [0.5, 0.5, 1.0]
    H = physics.hamiltonian(Is, _random_system(rng, Is))
    p = np.array([0.3, -0.4, 0.866])
    p /= np.linalg.norm(p)
    times = np.linspace(0, 3, 50)
    r0 = physics.rho0(Is, [1.0] * 3, 0, [0, 0, 0], p, np.inf)
    ev = np.real(dynamics.evolve(H, r0, times, [physics.muon_operator(Is, 0, p)])[:, 0])
    fe = dynamics.fast_evolve(H, physics.sigma_mu(p), times, 6)
    fe_py = dynamics.cosine_sum_py(times, 6, *_fast_AW(H, p, 6)) * 0.5
    assert np.max(np.abs(ev - fe)) < 1e-13
    assert np.max(np.abs(fe_py - fe)) < 1e-13


def _fast_AW(H, p, d):
    E, U = np.linalg.eigh(H)
    A = np.abs(U.conj().T @ np.kron(physics.sigma_mu(p), np.eye(d)) @ U) ** 2
    return A, 2 * np.pi * np.subtract.outer(E, E)


def test_integral_matches_quadrature():
    rng = np.random.default_rng(3)
    Is = [0.5, 0.5]
    H = 0.3 * physics.hamiltonian(Is, _random_system(rng, Is))
    r0 = physics.rho0(Is, [1.0, 1.0], 0, [0, 0, 0], [0, 0, 1], np.inf)
    O = physics.muon_operator(Is, 0, [0, 0, 1])
    tau = 0.7
    t = np.linspace(0, 40 * tau, 200001)
    f = dynamics.evolve(H, r0, t, [O])[:, 0] * np.exp(-t / tau)
    quad = np.trapezoid(f, t)
    val = dynamics.integrate_decaying(H, r0, tau, [O])[0]
    assert abs(val - quad) < 1e-7


def test_lindblad_matches_expm():
    rng = np.random.default_rng(9)
    Is = [0.5, 0.5]
    H = physics.hamiltonian(Is, _random_system(rng, Is))
    dops = physics.dissipation_operators(Is, [1.0, -2.0], [0.2, 0.1, 0.5], 1e-6, {1: 0.7})
    L = physics.lindbladian(H, dops)
    r0 = physics.rho0(Is, [1.0, 1.0], 0, [0, 0, 0], [1, 0, 0], np.inf)
    O = physics.muon_operator(Is, 0, [1, 0, 0])
    times = np.linspace(0, 1.5, 7)
    ev = dynamics.lindblad_evolve(L, r0, times, [O])[:, 0]
    for k, t in enumerate(times):
        rt = (expm(2 * np.pi * L * t) @ r0.reshape(-1)).reshape(4, 4)
        assert abs(ev[k] - np.trace(rt @ O)) < 1e-11


def _sys(Is, gammas, terms, muon_index, dissip=None):
    return types.SimpleNamespace(Is=Is, gammas=gammas, terms=terms,
                                 muon_index=muon_index, dissip=dissip or {})


def _batch(y, times, B, p, T, q=None, Bint=None):
    return types.SimpleNamespace(
        y=y, times=np.asarray(times), q=np.array([q if q is not None else [1.0, 0, 0, 0]]),
        B=np.array([B], dtype=float), Bint=np.array([Bint if Bint is not None else [0, 0, 0]], dtype=float),
        p=np.array([p], dtype=float), T=np.array([T], dtype=float), w=np.array([1.0]),
        slot=np.array([0]), sys_id=np.array([0]), n_slots=1, avg_N=1)


def test_pipeline_empty_and_precession():
    """tests/test_experiment.py:117-151: empty (e, mu) -> 0.5; Zeeman on the
    muon with B = 1/gamma_mu -> 0.5 cos(2 pi t)."""
    ge, gmu = physics.ELEC_GAMMA, physics.MU_GAMMA
    t = np.linspace(0, 10, 100)
    s = _sys([0.5, 0.5], [ge, gmu], [], 1)
    res = pipeline.run([s], _batch("asymmetry", t, [0, 0, 0], [1, 0, 0], np.inf))
    assert np.allclose(res, 0.5)
    s = _sys([0.5, 0.5], [ge, gmu], [("linear", 1, 1, np.array([0, 0, 1.0 / gmu]) * gmu)], 1)
    res = pipeline.run([s], _batch("asymmetry", t, [0, 0, 0], [1, 0, 0], np.inf))
    assert np.allclose(res[0], 0.5 * np.cos(2 * np.pi * t))
    # muon first -> fast path; same answer
    s = _sys([0.5, 0.5], [gmu, ge], [("linear", 0, 0, np.array([0, 0, 1.0 / gmu]) * gmu)], 0)
    res = pipeline.run([s], _batch("asymmetry", t, [0, 0, 0], [1, 0, 0], np.inf))
    assert np.allclose(res[0], 0.5 * np.cos(2 * np.pi * t))


def test_pipeline_alc_integral():
    """tests/test_experiment.py:153-202: integral = 0.5 / (1 + 4 pi^2 tau^2)
    for a 1 MHz local muon Zeeman term, transverse polarisation, on both the
    field and the intrinsic-field axis.  A field b along z adds gamma_mu b to
    the precession frequency, so the integral is 0.5 / (1 + 4 pi^2 tau^2 f^2)
    with f = 1 + gamma_mu b: an applied field reaches the muon rotated into
    the crystallite frame (experiment.py:405-406), an intrinsic one is not
    rotated (experiment.py:407), so under a tilted orientation only the
    intrinsic field keeps the closed form."""
    ge, gmu = physics.ELEC_GAMMA, physics.MU_GAMMA
    tau = physics.MU_TAU
    s = _sys([0.5, 0.5], [ge, gmu], [("linear", 1, 1, np.array([0, 0, 1.0 / gmu]) * gmu)], 1)

    def closed(b):
        f = 1.0 + gmu * b
        return 0.5 / (1.0 + 4 * np.pi ** 2 * tau ** 2 * f ** 2)

    ident = [1.0, 0, 0, 0]
    tilt = np.array([np.cos(0.35), np.sin(0.35), 0.0, 0.0])      # rotation about x
    for b in (0.0, 0.004, -0.011, 0.05):
        bz = [0, 0, b]
        res = pipeline.run([s], _batch("integral", [np.inf], bz, [1, 0, 0], np.inf, q=ident))
        assert np.isclose(res[0, 0], closed(b), rtol=1e-12, atol=0)
        for q in (ident, tilt):
            res = pipeline.run([s], _batch("integral", [np.inf], [0, 0, 0], [1, 0, 0], np.inf,
                                           q=q, Bint=bz))
            # the polarisation is rotated with the crystal too; along x it is
            # still transverse to z for a rotation about x
            assert np.isclose(res[0, 0], closed(b), rtol=1e-12, atol=0)
        if b != 0.0:
            res = pipeline.run([s], _batch("integral", [np.inf], bz, [1, 0, 0], np.inf, q=tilt))
            assert not np.isclose(res[0, 0], closed(b), rtol=1e-6, atol=0)


def test_pipeline_dissipation():
    """tests/test_experiment.py:581-640: 0.5 exp(-g t); thermal equilibrium
    0.5 (1 - Z) / (1 + Z) at long times."""
    g = 1.0
    gmu = physics.MU_GAMMA
    t = np.linspace(0, 10, 101)
    s = _sys([0.5], [gmu], [], 0, {0: g})
    res = pipeline.run([s], _batch("asymmetry", t, [0, 0, 0], [1, 0, 0], np.inf))
    assert np.allclose(res[0], 0.5 * np.exp(-g * t))
    T = 0.1
    B = [0, 0, -1.0]
    Br = np.array(B)
    times = np.array([0.0, 20.0])
    H = physics.zeeman([0.5], [gmu], Br)
    dops = physics.dissipation_operators([0.5], [gmu], Br, T, {0: g})
    L = physics.lindbladian(H, dops)
    r0 = physics.rho0([0.5], [gmu], 0, Br, [1, 0, 0], T)
    res = dynamics.lindblad_evolve(L, r0, times, [spin.s_z(0.5)])
    Z = np.exp(-cnst.h * gmu * 1e6 / (cnst.k * T))
    assert np.isclose(np.real(res[-1, 0]), 0.5 * (1 - Z) / (1 + Z))


def test_pipeline_rotation_field_vs_intrinsic():
    """tests/test_experiment.py:293-338: a rotated external field differs from
    an unrotated intrinsic one for a non-trivial orientation."""
    ge, gmu = physics.ELEC_GAMMA, physics.MU_GAMMA
    s = _sys([0.5, 0.5], [ge, gmu], [("linear", 1, 1, np.array([0, 0, 1.0 / gmu]) * gmu)], 1)
    q, _ = orient.orientation_entry(orient.zcw_angles(13)[3])
    t = np.linspace(0, 10, 100)
    a = pipeline.run([s], _batch("asymmetry", t, [10, 0, 0], [1, 0, 0], np.inf, q=q))
    b = pipeline.run([s], _batch("asymmetry", t, [0, 0, 0], [1, 0, 0], np.inf, q=q, Bint=[10, 0, 0]))
    assert not np.allclose(a, b)
