"""Celio's method on the CPU: the oracle against the reference's known
answers (tests/test_celio.py:21-157, tests/test_spinsys.py:192-300) and
against exact evolution, and the host side of the GPU path (contribution
splitting, fiber tables, descriptors, initial states) run through a numpy
emulation of msp_celio_evolve's contract against the oracle.

Tolerances: known answers to 1e-8 (the printed digits of the reference);
emulation vs oracle 1e-12; Trotter vs exact evolution within the k-step
splitting error (checked to shrink as 1/k)."""

import types

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import celio as oc
from oracle import dynamics, physics

SZ = np.array([[1, 0], [0, -1]], dtype=complex)


def test_oracle_contributions_known_answers():
    Is = [0.5, 0.5, 1.0]
    terms = [("linear", 0, 0, [1, 0, 0]),
             ("bilinear", 0, 1, np.array([[0, 0, 1]] * 3, float)),
             ("bilinear", 0, 2, np.array([[0, 0, -1]] * 3, float))]
    cs = oc.contributions(Is, terms)
    assert [c[1] for c in cs] == [(0,), (0, 1), (0, 2)]
    assert np.allclose(cs[0][0], [[0, 0.5], [0.5, 0]])
    assert np.allclose(cs[1][0], [[0.25, 0, 0.25 - 0.25j, 0], [0, -0.25, 0, -0.25 + 0.25j],
                                  [0.25 + 0.25j, 0, -0.25, 0], [0, -0.25 - 0.25j, 0, 0.25]])
    h = 0.5
    assert np.allclose(cs[2][0], [[-h, 0, 0, -h + h * 1j, 0, 0], [0] * 6,
                                  [0, 0, h, 0, 0, h - h * 1j],
                                  [-h - h * 1j, 0, 0, h, 0, 0], [0] * 6,
                                  [0, 0, h + h * 1j, 0, 0, -h]])


def test_oracle_trotter_factors_known_answers():
    f = oc.trotter_factors(oc.contributions([0.5, 0.5], [("linear", 0, 0, [1, 0, 0])]), 1, 10)
    full = oc.full_operator([2, 2], f[0][1], f[0][0])
    c, s = 0.95105652, -0.30901699j
    assert np.allclose(full, [[c, 0, s, 0], [0, c, 0, s], [s, 0, c, 0], [0, s, 0, c]])
    assert np.allclose(f[0][0], [[c, s], [s, c]])
    terms = [("linear", 0, 0, [1, 0, 1]), ("bilinear", 0, 1, np.eye(3))]
    f = oc.trotter_factors(oc.contributions([0.5, 0.5], terms), 1, 10)
    E = oc.full_operator([2, 2], *f[0][::-1]) @ oc.full_operator([2, 2], *f[1][::-1])
    ref = [[0.84425599 - 0.44143085j, -0.09276182 - 0.01469203j, 0.04521741 - 0.28549151j, 0],
           [0, 0.89336999 - 0.15115734j, -0.049114 - 0.29027351j, -0.0475444 - 0.30018354j],
           [-0.0475444 - 0.30018354j, 0.13640963 - 0.26088945j, 0.80293516 + 0.41982569j, 0],
           [0, 0.04521741 - 0.28549151j, -0.09276182 - 0.01469203j, 0.93934479 + 0.15893624j]]
    assert np.allclose(E, ref, atol=1e-8)
    assert np.allclose(f[0][0], [[0.902916796 - 0.30392537j, -0.30392537j],
                                 [-0.30392537j, 0.902916796 + 0.30392537j]], atol=1e-8)


def test_oracle_closed_forms():
    t = np.linspace(0, 1, 100)
    rng = np.random.default_rng(0)
    r = oc.fast_evolve([0.5, 0.5], [("linear", 0, 0, [1, 0, 0])], 10, t, SZ, rng.random((4, 2)))
    assert np.abs(r - 0.5 * np.cos(2 * np.pi * t)).max() <= 1e-12
    rho0 = np.diag([1.0, 0.0]).astype(complex)
    ev = oc.evolve([0.5], [("linear", 0, 0, [1, 0, 0])], 10, t, rho0, SZ / 2)
    assert np.abs(ev - 0.5 * np.cos(2 * np.pi * t)).max() <= 1e-12


def _random_system(seed=3):
    rng = np.random.default_rng(seed)
    Is = [0.5, 1.0, 0.5]
    A = rng.normal(0, 2.0, (3, 3))
    Q = rng.normal(0, 1.0, (3, 3))
    terms = [("bilinear", 0, 1, A + A.T), ("bilinear", 0, 2, np.diag([3.0, 3.0, 4.0])),
             ("bilinear", 1, 1, Q + Q.T), ("linear", 0, 0, [0.3, 0.0, 1.1]),
             ("bilinear", 2, 1, np.eye(3) * 0.7)]
    return Is, terms


def test_oracle_converges_to_exact_evolution():
    """T = inf random-state average with all n/2 basis phases replaced by the
    exact trace: the density mode at growing k approaches exact evolution
    (error ~ 1/k)."""
    Is, terms = _random_system()
    t = np.linspace(0, 2.0, 21)
    H = physics.hamiltonian(Is, terms)
    rho0 = physics.rho0(Is, [1, 1, 1], 0, np.zeros(3), [0, 0, 1.0], np.inf)
    O = physics.muon_operator(Is, 0, [0, 0, 1.0])
    exact = np.real(dynamics.evolve(H, rho0, t, [O])[:, 0])
    errs = []
    for k in (20, 80, 320):
        tro = np.real(oc.evolve(Is, terms, k, t, rho0, O))
        errs.append(np.abs(tro - exact).max())
    # first-order splitting: the error falls about 4x per 4x in k
    assert errs[1] < errs[0] / 3.0 and errs[2] < errs[1] / 3.0 and errs[2] < 2e-3


# ------------------------------------------------------------------ host side
class _T:
    def __init__(self, idx, tensor):
        self.indices = np.array(idx)
        self.tensor = np.asarray(tensor, dtype=float)


def _as_terms(terms):
    return [_T([i] if k == "linear" else [i, j], T) for k, i, j, T in terms]


def _emulate(job, ntimes):
    """numpy restatement of msp_celio_evolve's contract on the job's arrays,
    with the factors computed by scipy's expm."""
    n = job.n
    mats = np.zeros(job._nmat, dtype=complex)
    for off, M in job._fixed:
        mats[off:off + M.size] = M.reshape(-1)
    for off, H, sign in job._H:
        mats[off:off + H.size] = expm(sign * 2j * np.pi * H * job.dt / job.k).reshape(-1)
    for off, frame, parts in job._composites:
        # fused run: product of the parts, each as a full-space factor
        # restricted to the frame (built by the oracle's own embedding)
        fd = [job.dims[q] for q in frame]
        F = np.eye(int(np.prod(fd)), dtype=complex)
        for po, sp, m in parts:
            c = mats[po:po + m * m].reshape(m, m)
            E = oc.full_operator(fd, tuple(frame.index(q) for q in sp), c)
            F = E @ F
        mats[off:off + F.size] = F.reshape(-1)
    tab = np.concatenate(job._tables)
    ivec = np.concatenate(job._init_vec)
    out = np.zeros((job.nstates, ntimes))
    dims = job.dims
    for s in range(job.nstates):
        cfg = job._state_cfg[s]
        if job.mode == 0:
            mu = ivec[job._init_idx[s]:job._init_idx[s] + 2]
            psi = np.kron(mu, np.exp(2j * np.pi * job._phases[s])) / np.sqrt(n // 2)
        else:
            psi = np.ones(1, dtype=complex)
            for q, off in enumerate(job._init_idx[s]):
                psi = np.kron(psi, ivec[off:off + dims[q]])
        first, count = job._cfg_steps[cfg]
        m, mo, fo, eo = job._meas[cfg]
        for t in range(ntimes):
            O = mats[mo:mo + m * m].reshape(m, m)
            fib, el = tab[fo:fo + n // m], tab[eo:eo + m]
            X = psi[fib[:, None] + el[None, :]]
            out[s, t] = np.real(np.einsum("fi,ij,fj->", X.conj(), O, X))
            if t == ntimes - 1:
                break
            for _ in range(job.k):
                for m2, mo2, fo2, eo2 in job._desc[first:first + count]:
                    M = mats[mo2:mo2 + m2 * m2].reshape(m2, m2)
                    fib, el = tab[fo2:fo2 + n // m2], tab[eo2:eo2 + m2]
                    ix = fib[:, None] + el[None, :]
                    psi[ix] = psi[ix] @ M.T
    return out


def test_job_random_states_match_oracle():
    from muspinsim_amd import celio
    Is, terms = _random_system()
    t = np.linspace(0, 1.5, 12)
    sig = np.array([[0, 1], [1, 0]], dtype=complex)
    celio.seed(11)
    job = celio.CelioJob(Is, t[1] - t[0], 7)
    c = job.add_config(celio.h_contribs(Is, _as_terms(terms)), sig, 0)
    job.add_random_states(c, celio.plus_state(sig), 3)
    got = 0.5 * _emulate(job, t.size).mean(axis=0)
    ref = oc.fast_evolve(Is, terms, 7, t, sig, np.random.default_rng(11).random((3, 6)))
    assert np.abs(got - ref).max() <= 1e-12


def test_job_product_states_match_oracle():
    """Density mode: rho0 = pure muon (x) thermal factors as weighted product
    states evolved with the daggered factors, against Tr(U^H rho0 U O)."""
    from muspinsim_amd import celio
    Is, terms = _random_system()
    t = np.linspace(0, 1.5, 10)
    B = np.array([0.01, 0.0, 0.02])
    g = [physics.MU_GAMMA, 10.0, -28024.9514242]
    rho_f = physics.rho0_factors(Is, g, 0, B, [0.6, 0.0, 0.8], 0.02)
    rho0 = physics.rho0(Is, g, 0, B, [0.6, 0.0, 0.8], 0.02)
    O = physics.muon_operator(Is, 0, [0.6, 0.0, 0.8])
    facs = []
    for r in rho_f:
        w, V = np.linalg.eigh(r)
        facs.append((V, np.where(w > 1e-14, w, 0.0)))
    job = celio.CelioJob(Is, t[1] - t[0], 5)
    S = [physics.spin.single(0.5, a) for a in "xyz"]
    c = job.add_config(celio.h_contribs(Is, _as_terms(terms)),
                       sum(p * s for p, s in zip([0.6, 0.0, 0.8], S)), 0, dagger=True)
    wts = job.add_product_states(c, facs)
    got = np.asarray(wts) @ _emulate(job, t.size)
    ref = np.real(oc.evolve(Is, terms, 5, t, rho0, O))
    assert np.abs(got - ref).max() <= 1e-12


def test_job_muon_not_first_and_spin_order():
    """A muon in the middle and a two-spin term with decreasing indices."""
    from muspinsim_amd import celio
    Is = [0.5, 0.5, 1.0]
    terms = [("bilinear", 1, 0, np.diag([1.0, 2.0, 3.0])), ("bilinear", 2, 1, np.eye(3) * 0.4),
             ("linear", 1, 1, [0.0, 0.2, 0.5])]
    t = np.linspace(0, 1.0, 8)
    rho0 = physics.rho0(Is, [1, 1, 1], 1, np.zeros(3), [1.0, 0, 0], np.inf)
    O = physics.muon_operator(Is, 1, [1.0, 0, 0])
    job = celio.CelioJob(Is, t[1] - t[0], 4)
    S = [physics.spin.single(0.5, a) for a in "xyz"]
    c = job.add_config(celio.h_contribs(Is, _as_terms(terms)), S[0], 1, dagger=True)
    facs = [(np.eye(2), np.full(2, 0.5)), (np.array([[1], [1]]) / np.sqrt(2), np.ones(1)),
            (np.eye(3), np.full(3, 1 / 3))]
    wts = job.add_product_states(c, facs)
    got = np.asarray(wts) @ _emulate(job, t.size)
    ref = np.real(oc.evolve(Is, terms, 4, t, rho0, O))
    assert np.abs(got - ref).max() <= 1e-12


def test_embedding_matches_full_operator():
    """The index map that embeds a fused part on its frame equals the
    oracle's (c on the part's spins) (x) identity, for parts in any order."""
    from muspinsim_amd.celio import _embedding
    rng = np.random.default_rng(2)
    dims = [2, 3, 2, 4]
    for frame, part in [((0, 1), (0,)), ((0, 1), (1,)), ((1, 0), (0,)), ((2, 3), (3, 2)),
                        ((3, 1), (1, 3)), ((0, 1), (0, 1)), ((0, 1), (1, 0))]:
        mp = int(np.prod([dims[q] for q in part]))
        c = rng.normal(size=(mp, mp)) + 1j * rng.normal(size=(mp, mp))
        ia, ib, ra, rb, m2 = _embedding(dims, list(frame), list(part))
        fd = [dims[q] for q in frame]
        m = int(np.prod(fd))
        E = np.zeros((m, m), dtype=complex)
        E[ia, ib] = c[ra, rb]
        ref = oc.full_operator(fd, tuple(frame.index(q) for q in part), c)
        assert m2 == mp and np.array_equal(E, ref)


def test_system_split_matches_grouping_of_full_term_list():
    """Zeeman terms merged per configuration give the same contributions as
    grouping the term list system + Zeeman (celio.py:88-100)."""
    from muspinsim_amd import celio
    Is, terms = _random_system()
    terms = terms + [("linear", 2, 2, [0.1, 0.2, 0.3])]           # spin 2 ends on a linear run
    z = [np.array([0.3, -0.2, 0.5]) * (k + 1) for k in range(3)]
    split = celio.SystemSplit(Is, _as_terms(terms))
    got = split.contribs(z)
    ref = celio.h_contribs(Is, _as_terms(terms + [("linear", i, i, z[i]) for i in range(3)]))
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        assert a.spin_order == b.spin_order and np.allclose(a.matrix, b.matrix, atol=1e-14)
    assert len(split.contribs(None)) == len(celio.h_contribs(Is, _as_terms(terms)))
