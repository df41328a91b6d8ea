"""CPU build of the device eigensolver algorithms (the same team functions the
HIP kernels run, compiled with g++ as a one-thread team) checked against
numpy.linalg.eigh / eigvalsh.

The workspace is poisoned (NaN / -1) before each solve, as reused device
memory holds stale values, so any read-before-write shows up here.
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "muspinsim_amd", "csrc")


@pytest.fixture(scope="module")
def hc(tmp_path_factory):
    out = tmp_path_factory.mktemp("hc") / "libhostcheck.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-o", str(out),
                    os.path.join(CSRC, "host_check.cpp")], check=True)
    return ctypes.CDLL(str(out))


def _herm(rng, n):
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return np.ascontiguousarray((X + X.conj().T) / 2)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _check(H, E, U, tol):
    n = H.shape[0]
    Er = np.linalg.eigvalsh(H)
    scale = max(1.0, np.abs(Er).max())
    assert np.abs(E - Er).max() <= tol * n * scale
    assert np.abs(H @ U - U * E).max() <= tol * n * scale
    assert np.abs(U.conj().T @ U - np.eye(n)).max() <= tol * n


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16, 17, 33, 64])
def test_jacobi(hc, n):
    H = _herm(np.random.default_rng(n), n)
    E = np.zeros(n)
    U = np.zeros((n, n), complex)
    sweeps = hc.hc_jacobi(n, _p(H), _p(E), _p(U), 40)
    assert sweeps < 40
    _check(H, E, U, 1e-14)


@pytest.mark.parametrize("n", [2, 3, 5, 8, 13, 64, 100, 192])
def test_divide_and_conquer_tridiagonal(hc, n):
    rng = np.random.default_rng(100 + n)
    d, e = rng.normal(size=n), rng.normal(size=n - 1)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    dd = d.copy()
    Z = np.zeros((n, n))
    hc.hc_stedc(n, _p(dd), _p(e), _p(Z))
    assert np.abs(dd - np.linalg.eigvalsh(T)).max() < 1e-14 * n * max(1, np.abs(dd).max())
    assert np.abs(T @ Z - Z * dd).max() < 1e-14 * n * max(1, np.abs(dd).max())
    assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-14 * n


def test_divide_and_conquer_deflation(hc):
    """Exact degeneracies (deflation of equal poles and of zero couplings)."""
    n = 40
    d = np.repeat([1.0, 2.0, 3.0, 4.0], 10)
    e = np.zeros(n - 1)
    e[::3] = 1e-3
    e[5] = 0.0
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    dd = d.copy()
    Z = np.zeros((n, n))
    hc.hc_stedc(n, _p(dd), _p(e), _p(Z))
    assert np.abs(dd - np.linalg.eigvalsh(T)).max() < 1e-13
    assert np.abs(T @ Z - Z * dd).max() < 1e-13
    assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-13


@pytest.mark.parametrize("n,cw,nb", [(2, 4, 4), (8, 4, 3), (17, 5, 4), (65, 16, 16), (192, 16, 16)])
def test_full_path_blocked(hc, n, cw, nb):
    """Householder tridiagonalisation + divide and conquer + blocked WY
    back-transformation."""
    H = _herm(np.random.default_rng(200 + n), n)
    E = np.zeros(n)
    U = np.zeros((n, n), complex)
    hc.hc_eigh_dc_blocked(n, _p(H), _p(E), _p(U), cw, nb)
    _check(H, E, U, 1e-14)


@pytest.mark.parametrize("n,nb", [(3, 1), (5, 2), (17, 4), (65, 16), (192, 8)])
def test_blocked_tridiagonalisation_matches_unblocked(hc, n, nb):
    H = _herm(np.random.default_rng(300 + n), n)
    A1, A2 = H.copy(), H.copy()
    d1, e1, t1 = np.zeros(n), np.zeros(n), np.zeros(n, complex)
    d2, e2, t2 = np.zeros(n), np.zeros(n), np.zeros(n, complex)
    hc.hc_tridiag(n, _p(A1), _p(d1), _p(e1), _p(t1))
    hc.hc_tridiag_blocked(n, _p(A2), _p(d2), _p(e2), _p(t2), nb)
    assert np.abs(d1 - d2).max() < 1e-12 * n
    assert np.abs(e1[: n - 1] - e2[: n - 1]).max() < 1e-12 * n
    T = np.diag(d2) + np.diag(e2[: n - 1], 1) + np.diag(e2[: n - 1], -1)
    assert np.abs(np.linalg.eigvalsh(T) - np.linalg.eigvalsh(H)).max() < 1e-13 * n


def test_physical_hamiltonians(hc):
    """Zero-field degenerate spectrum (n = 16) and a 3 T ALC Hamiltonian
    (n = 192, 5 orders of magnitude between couplings)."""
    from muspinsim_amd import workloads as W
    from oracle import physics
    s1, _ = W.c1_zero_field_powder(nori=1)
    s = s1[0]
    H = np.ascontiguousarray(physics.hamiltonian(s.Is, s.terms))
    for fn in ("jac", "dc"):
        E = np.zeros(16)
        U = np.zeros((16, 16), complex)
        if fn == "jac":
            hc.hc_jacobi(16, _p(H), _p(E), _p(U), 40)
        else:
            hc.hc_eigh_dc_blocked(16, _p(H), _p(E), _p(U), 4, 4)
        _check(H, E, U, 1e-14)
    s = W.c2_system()
    H = np.ascontiguousarray(physics.hamiltonian(s.Is, s.terms) +
                             physics.zeeman(s.Is, s.gammas, [0.3, 0.2, 3.0]))
    E = np.zeros(192)
    U = np.zeros((192, 192), complex)
    hc.hc_eigh_dc_blocked(192, _p(H), _p(E), _p(U), 16, 16)
    _check(H, E, U, 1e-14)


def _modes_ref(L, r, u):
    """lambda_m and amp_m = (u^T V)_m (V^{-1} r)_m from numpy.linalg.eig (the
    reference's Lindbladian.evolve, lindbladian.py:84-93); amp does not depend
    on the scaling of V's columns."""
    lam, V = np.linalg.eig(L)
    amp = (u @ V) * np.linalg.solve(V, r)
    order = np.lexsort((lam.imag, lam.real))
    return lam[order], amp[order]


@pytest.mark.parametrize("n", [4, 9, 16, 36, 64])
def test_mode_amplitudes_with_and_without_schur_vectors(hc, n):
    """Non-Hermitian eigendecomposition: the amplitudes from the explicit
    Schur vectors Z (hc_eig) and from r and u carried through the reduction
    instead of Z (hc_eig_vec, the device's LDS form without eigenvectors)
    both match numpy.linalg.eig + solve."""
    rng = np.random.default_rng(70 + n)
    L = np.ascontiguousarray(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    r = np.ascontiguousarray(rng.normal(size=n) + 1j * rng.normal(size=n))
    u = np.ascontiguousarray(rng.normal(size=n) + 1j * rng.normal(size=n))
    lam_ref, amp_ref = _modes_ref(L, r, u)
    scale = np.abs(amp_ref).max()
    for vec in (False, True):
        lam = np.zeros(n, complex)
        c = np.zeros(n, complex)
        o = np.zeros(n, complex)
        if vec:
            info = hc.hc_eig_vec(n, _p(L), _p(lam), _p(r), _p(u), _p(c), _p(o), 30)
        else:
            V = np.zeros((n, n), complex)
            info = hc.hc_eig(n, _p(L), _p(lam), _p(V), _p(r), _p(u), _p(c), _p(o), 30)
        assert info == 0
        order = np.lexsort((lam.imag, lam.real))
        assert np.abs(lam[order] - lam_ref).max() <= 1e-11 * np.abs(lam_ref).max()
        assert np.abs((c * o)[order] - amp_ref).max() <= 1e-10 * scale
