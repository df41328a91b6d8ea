"""Celio's Trotter method (oracle; test infrastructure only).

Restates muspinsim/celio.py:54-476 (Hamiltonian contributions, Trotter
factors, density-matrix evolve, random-state fast_evolve) and the loop of
cpp/celio.cpp:23-68 with the kernels of cpp/parallel.cpp:130-266, as dense
numpy on small systems.

Contributions (celio.py:73-143): for each spin i in order, the terms whose
first index is i, grouped into runs of consecutive terms with the same
index tuple; each run is one m x m matrix on its spins (a quadrupolar (i, i)
run acts on spin i alone), built here in the order of the term's indices.
Trotter factor c = expm(-2 pi i H dt / k) (celio.py:167-170).

fast_evolve (celio.py:295-476): psi0 = mu_psi (x) exp(2 pi i phi) / sqrt(n/2)
with mu_psi the +1 eigenvector of sigma_mu and phases phi in [0, 1) given
by the caller (the reference draws them with np.random); each time step
measures <psi| sigma_mu (x) 1 |psi> and then applies c_1, ..., c_m, k times.
The result is 0.5 x the average over the phase draws.

evolve (celio.py:210-293): rho(t_i) = U_i^H rho0 U_i with U_i = E^i and
E = (c_1 c_2 ... c_m)^k, and P = Tr(rho O).
"""

import itertools

import numpy as np
from scipy.linalg import expm

from . import spin

_XYZ = "xyz"


def term_spins(kind, i, j):
    """Index tuple of a term as the reference groups it."""
    return (i,) if kind == "linear" else (i, j)


def small_matrix(Is, kind, i, j, tensor):
    """Operator of one term on its own spins, in the order of its indices."""
    if kind == "linear":
        v = np.asarray(tensor, dtype=float)
        return sum(v[a] * spin.single(Is[i], _XYZ[a]) for a in range(3))
    T = np.asarray(tensor, dtype=float)
    if i == j:
        return sum(T[a, b] * (spin.single(Is[i], _XYZ[a]) @ spin.single(Is[i], _XYZ[b]))
                   for a in range(3) for b in range(3))
    return sum(T[a, b] * np.kron(spin.single(Is[i], _XYZ[a]), spin.single(Is[j], _XYZ[b]))
               for a in range(3) for b in range(3))


def contributions(Is, terms):
    """[(H (m x m), spins)] in Trotter order; terms are (kind, i, j, tensor)
    with kind 'linear' or 'bilinear' (scalar terms do not evolve a state)."""
    out = []
    terms = [t for t in terms if t[0] != "scalar"]
    for i in range(len(Is)):
        mine = [t for t in terms if t[1] == i]
        for key, run in itertools.groupby(mine, key=lambda t: term_spins(t[0], t[1], t[2])):
            run = list(run)
            spins = (key[0],) if len(key) == 2 and key[0] == key[1] else tuple(key)
            out.append((sum(small_matrix(Is, *t) for t in run), spins))
    return out


def trotter_factors(contribs, dt, k):
    return [(expm(-2j * np.pi * H * dt / k), spins) for H, spins in contribs]


def apply_local(psi, dims, spins, c):
    """(c on `spins`) (x) identity, applied to the state vector psi."""
    t = psi.reshape(dims)
    others = [a for a in range(len(dims)) if a not in spins]
    order = list(spins) + others
    t = np.transpose(t, order).reshape(int(np.prod([dims[a] for a in spins])), -1)
    t = c @ t
    t = t.reshape([dims[a] for a in order])
    return np.transpose(t, np.argsort(order)).reshape(-1)


def full_operator(dims, spins, c):
    """The same factor as a full n x n matrix."""
    n = int(np.prod(dims))
    return np.stack([apply_local(np.eye(n, dtype=complex)[:, a], dims, spins, c)
                     for a in range(n)], axis=1)


def plus_state(sigma):
    """+1 eigenvector of a 2 x 2 Pauli combination."""
    w, v = np.linalg.eigh(sigma)
    return v[:, np.argmax(w)]


def fast_evolve(Is, terms, k, times, sigma, phases):
    """Random-state Celio evolution; phases: (averages, n/2) in [0, 1)."""
    times = np.asarray(times, dtype=float)
    dims = [int(round(2 * I + 1)) for I in Is]
    n = int(np.prod(dims))
    half = n // 2
    facs = trotter_factors(contributions(Is, terms), times[1] - times[0], k)
    mu = plus_state(sigma)
    op = np.kron(sigma, np.eye(half))
    out = np.zeros(len(times))
    for ph in np.atleast_2d(phases):
        psi = np.kron(mu, np.exp(2j * np.pi * ph)) / np.sqrt(half)
        for it in range(len(times)):
            out[it] += np.real(np.vdot(psi, op @ psi))
            if it == len(times) - 1:
                break
            for _ in range(k):
                for c, sp in facs:
                    psi = apply_local(psi, dims, sp, c)
    return 0.5 * out / len(np.atleast_2d(phases))


def evolve(Is, terms, k, times, rho0, operator):
    """Density-matrix Celio evolution: Tr(U_i^H rho0 U_i O)."""
    times = np.asarray(times, dtype=float)
    dims = [int(round(2 * I + 1)) for I in Is]
    n = int(np.prod(dims))
    facs = trotter_factors(contributions(Is, terms), times[1] - times[0], k)
    E = np.eye(n, dtype=complex)
    for c, sp in facs:
        E = E @ full_operator(dims, sp, c)
    E = np.linalg.matrix_power(E, k)
    U = np.eye(n, dtype=complex)
    out = np.zeros(len(times), dtype=complex)
    for it in range(len(times)):
        out[it] = np.trace(U.conj().T @ rho0 @ U @ operator)
        U = U @ E
    return out


def validate(terms, times):
    """celio parameters (validation.py:48-71)."""
    times = np.asarray(times, dtype=float)
    if len([t for t in terms if t[0] != "scalar"]) == 0:
        raise ValueError("No interaction terms to evolve")
    if times[0] != 0:
        raise ValueError("Cannot use Celio's method with a non-zero start time")
    d = np.diff(times)
    if not np.isclose(d, d[0]).all():
        raise ValueError("Cannot use Celio's method with uneven spacing between times")
