"""Closed- and open-system dynamics (oracle).

Restates:
  muspinsim/hamiltonian.py:40-117    Hamiltonian.evolve
  muspinsim/hamiltonian.py:119-164   Hamiltonian.integrate_decaying
  muspinsim/hamiltonian.py:166-217   Hamiltonian.fast_evolve
  muspinsim/cython/parallel.pyx:16-69  parallel_fast_time_evolve (cosine sum)
  muspinsim/lindbladian.py:43-111    Lindbladian.evolve
  muspinsim/lindbladian.py:113-173   Lindbladian.integrate_decaying

All arithmetic is numpy FP64 / complex128 with LAPACK eigh (zheevd), eig
(zgeev) and solve (zgesv), the same drivers the reference calls.
"""

import ctypes
import os

import numpy as np

from . import spin


def eigh(H):
    """np.linalg.eigh of the dense matrix (Hermitian.diag, spinop.py:51-82)."""
    return np.linalg.eigh(H)


def evolve(H, rho0, times, operators):
    """Expectation values <O>(t), shape (T, n_ops) complex (hamiltonian.py:40-107).

    Vectorised over times; the arithmetic per (t, i, j) is the reference's:
    exp(-2 pi i (E_i - E_j) t) * rho'_ij * O'_ji, summed over i, j."""
    times = np.asarray(times, dtype=float)
    evals, evecs = eigh(H)
    rho = spin.basis_change(rho0, evecs)
    opsT = np.array([spin.basis_change(o, evecs).T for o in operators])
    ll = -2.0j * np.pi * (evals[:, None] - evals[None, :])
    out = np.empty((len(times), len(operators)), dtype=complex)
    for k, t in enumerate(times):
        r = np.exp(ll * t) * rho
        out[k] = np.sum(r[None, :, :] * opsT, axis=(1, 2))
    return out


def evolve_density(H, rho0, times):
    """Density matrices rho(t) = U (e^{ll t} * rho') U^+ (hamiltonian.py:108-116)."""
    times = np.asarray(times, dtype=float)
    evals, evecs = eigh(H)
    rho = spin.basis_change(rho0, evecs)
    ll = -2.0j * np.pi * (evals[:, None] - evals[None, :])
    out = []
    for t in times:
        r = evecs @ (np.exp(ll * t) * rho) @ evecs.conj().T
        out.append(r / np.trace(r))
    return out


def integrate_decaying(H, rho0, tau, operators):
    """sum_ij rho'_ij O'_ji / (1/tau + 2 pi i (E_i - E_j))  (hamiltonian.py:150-162)."""
    evals, evecs = eigh(H)
    rho = spin.basis_change(rho0, evecs)
    ll = 2.0j * np.pi * (evals[:, None] - evals[None, :])
    intops = np.array([(-spin.basis_change(o, evecs) / (ll - 1.0 / tau)).T for o in operators])
    return np.sum(rho[None, :, :] * intops, axis=(1, 2))


def cosine_sum_py(times, d, A, W):
    """Pure-numpy form of parallel_fast_time_evolve (cython/parallel.pyx:51-67):
    (1/d)[sum_j A_jj / 2 + sum_{k<j} A_jk cos(W_jk t)]."""
    times = np.asarray(times, dtype=float)
    n = A.shape[0]
    jj, kk = np.tril_indices(n, -1)
    a = A[jj, kk]
    w = W[jj, kk]
    diag = 0.5 * np.sum(np.diag(A))
    out = np.empty(len(times))
    for i, t in enumerate(times):
        out[i] = (diag + np.sum(a * np.cos(w * t))) / d
    return out


_LIB = None


def _cosine_lib():
    """The C/OpenMP restatement of the Cython kernel (oracle/cosine_sum.c),
    built by oracle/Makefile; None when it has not been built."""
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(__file__), "libmsp_oracle.so")
        if not os.path.exists(path):
            return None
        lib = ctypes.CDLL(path)
        lib.oracle_cosine_sum.restype = None
        lib.oracle_cosine_sum.argtypes = [
            ctypes.c_void_p, ctypes.c_long, ctypes.c_long, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        _LIB = lib
    return _LIB


def cosine_sum(times, d, A, W, threads=1):
    """parallel_fast_time_evolve through the compiled C restatement when built,
    else the numpy form (identical arithmetic, slower)."""
    lib = _cosine_lib()
    if lib is None:
        return cosine_sum_py(times, d, A, W)
    times = np.ascontiguousarray(times, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    W = np.ascontiguousarray(W, dtype=np.float64)
    out = np.zeros(len(times))
    lib.oracle_cosine_sum(times.ctypes.data, len(times), A.shape[0], A.ctypes.data,
                          W.ctypes.data, out.ctypes.data, int(threads))
    return out * (1.0 / d)


def fast_evolve(H, sigma_mu, times, other_dimension, threads=1):
    """T -> inf muon polarisation, muon first (hamiltonian.py:166-217)."""
    evals, evecs = eigh(H)
    sig = np.kron(sigma_mu, np.eye(other_dimension))
    A = evecs.conj().T @ (sig @ evecs)
    A = np.abs(A) ** 2
    W = 2 * np.pi * np.subtract.outer(evals, evals)
    return cosine_sum(times, other_dimension, A, W, threads) * 0.5


def lindblad_evolve(L, rho0, times, operators):
    """Mode expansion of exp(2 pi L t) vec(rho0) (lindbladian.py:76-111)."""
    times = np.asarray(times, dtype=float)
    evals, revecs = np.linalg.eig(L)
    r0 = np.linalg.solve(revecs, rho0.reshape(-1))
    opsT = np.array([o.T.reshape(-1) @ revecs for o in operators])
    rho = np.exp(2.0 * np.pi * evals[None, :] * times[:, None]) * r0[None, :]
    if len(operators) > 0:
        return np.sum(opsT[None, :, :] * rho[:, None, :], axis=-1)
    n = rho0.shape[0]
    return [(revecs @ r).reshape(n, n) for r in rho]


def lindblad_integrate_decaying(L, rho0, tau, operators):
    """sum_k o_k c_k / (1/tau - 2 pi lambda_k) (lindbladian.py:140-173)."""
    evals, revecs = np.linalg.eig(L)
    r0 = np.linalg.solve(revecs, rho0.reshape(-1))
    intops = np.array([(o.T.reshape(-1) @ revecs) / (1.0 / tau - 2.0 * np.pi * evals)
                       for o in operators])
    return np.sum(r0[None, :] * intops, axis=1)
