"""Physical model -> dense operators (oracle).

Restates:
  muspinsim/constants.py:12-20        MU_TAU, ELEC_GAMMA, MU_GAMMA, EFG_2_MHZ
  muspinsim/spinsys.py:34-69          InteractionTerm._recalc_operator
  muspinsim/spinsys.py:377-442        dipolar and quadrupolar tensors
  muspinsim/spinsys.py:613-626        SpinSystem.hamiltonian (sum of terms)
  muspinsim/experiment.py:169-236     ExperimentRunner.rho0
  muspinsim/experiment.py:238-269     ExperimentRunner.Hz
  muspinsim/experiment.py:271-325     ExperimentRunner.dissipation_operators
  muspinsim/lindbladian.py:19-41      Lindbladian.from_hamiltonian / add_dissipative_term
  muspinsim/spinsys.py:707-760        muon_operator / sigma_mu

A system is described by plain data: ``Is`` (spin quantum numbers in kron
order), ``gammas`` (MHz/T) and ``terms``, a list of (kind, i, j, tensor) with
kind "scalar" (tensor: float), "linear" (tensor: 3-vector, on spin i) or
"bilinear" (tensor: 3x3 on spins i, j; i == j means the same-spin product
S_i^a S_i^b, as for quadrupolar terms).
"""

import functools

import numpy as np
import scipy.constants as cnst

from . import spin


@functools.lru_cache(maxsize=64)
def _spin_ops(Is):
    """Full-space S_i^a for every spin, built once per spin list like the
    reference's cached single-spin operators (experiment.py:105-111)."""
    ops = np.array([[spin.spin_op(list(Is), i, a) for a in _XYZ] for i in range(len(Is))])
    ops.setflags(write=False)
    return ops

ELEC_GAMMA = -28024.9514242                # constants.py:12
MU_GAMMA = -(ELEC_GAMMA / 206.7669883)     # constants.py:13
MU_TAU = 2.19703                           # constants.py:14
EFG_2_MHZ = (                              # constants.py:20-25
    cnst.physical_constants["atomic unit of electric field gradient"][0]
    * cnst.e * 1e-37 / cnst.h
)

_XYZ = "xyz"


def dipolar_tensor(g_i, g_j, r):
    """Dipolar coupling tensor in MHz for vector r in Angstrom (spinsys.py:400-409)."""
    r = np.asarray(r, dtype=float)
    rnorm = np.linalg.norm(r)
    D = -(np.eye(3) - 3.0 / rnorm ** 2.0 * r[:, None] * r[None, :])
    dij = -(cnst.mu_0 * cnst.hbar * (g_i * g_j * 1e6)) / (2 * (rnorm * 1e-10) ** 3)
    return D * dij


def quadrupolar_tensor(Q, I, efg):
    """Quadrupolar tensor in MHz from an EFG tensor (spinsys.py:433-438)."""
    if I == 0.5:
        raise ValueError("Can not set up quadrupolar coupling for spin 1/2 particle")
    return EFG_2_MHZ * Q / (2 * I * (2 * I - 1)) * np.asarray(efg, dtype=float)


def term_matrix(Is, kind, i, j, tensor):
    """Dense operator of one interaction term (spinsys.py:34-69)."""
    n = int(np.prod([int(round(2 * I + 1)) for I in Is]))
    if kind == "scalar":
        return float(tensor) * np.eye(n, dtype=complex)
    if kind == "linear":
        v = np.asarray(tensor, dtype=float)
        return sum(v[a] * spin.spin_op(Is, i, _XYZ[a]) for a in range(3))
    if kind == "bilinear":
        M = np.asarray(tensor, dtype=float)
        out = np.zeros((n, n), dtype=complex)
        for a in range(3):
            for b in range(3):
                if i == j:
                    op = spin.embed(Is, {i: spin.single(Is[i], _XYZ[a]) @ spin.single(Is[i], _XYZ[b])})
                else:
                    op = spin.embed(Is, {i: spin.single(Is[i], _XYZ[a]),
                                         j: spin.single(Is[j], _XYZ[b])})
                out += M[a, b] * op
        return out
    raise ValueError(f"unknown term kind {kind}")


def hamiltonian(Is, terms):
    """Sum of all term operators (SpinSystem.hamiltonian, spinsys.py:613-626)."""
    n = int(np.prod([int(round(2 * I + 1)) for I in Is]))
    H = np.zeros((n, n), dtype=complex)
    for kind, i, j, tensor in terms:
        H += term_matrix(Is, kind, i, j, tensor)
    return H


def zeeman(Is, gammas, B):
    """Sum_i gamma_i B.S_i (ExperimentRunner.Hz, experiment.py:238-254)."""
    B = np.asarray(B, dtype=float)
    ops = _spin_ops(tuple(Is))
    n = ops.shape[-1]
    H = np.zeros((n, n), dtype=complex)
    for i, g in enumerate(gammas):
        for a in range(3):
            H += ops[i, a] * (B[a] * g)
    return H


def rho0(Is, gammas, muon_index, B, p, T):
    """Initial density matrix: muon pure along p, every other spin thermal under
    its own Zeeman Hamiltonian at temperature T (experiment.py:169-236)."""
    B = np.asarray(B, dtype=float)
    rhos = []
    for i, I in enumerate(Is):
        if i == muon_index:
            r = spin.pure_state(I, p)
        else:
            Hz = sum(B[a] * spin.single(I, _XYZ[a]) for a in range(3))
            evals, evecs = np.linalg.eigh(Hz)
            E = evals * 1e6 * gammas[i]
            if T > 0:
                Z = np.exp(-cnst.h * E / (cnst.k * T))
            else:
                Z = np.where(E == np.amin(E), 1.0, 0.0)
            if np.sum(Z) > 0:
                Z = Z / np.sum(Z)
            else:
                Z = np.ones(len(E)) / len(E)
            r = np.sum(evecs[:, None, :] * evecs[None, :, :].conj() * Z[None, None, :], axis=-1)
            r = r / np.trace(r)
        rhos.append(r)
    return spin.kron_all(rhos)


def rho0_factors(Is, gammas, muon_index, B, p, T):
    """Same state as rho0 but returned as the list of per-spin factors."""
    out = []
    for i in range(len(Is)):
        sub = rho0([Is[i]], [gammas[i]], 0 if i == muon_index else -1, B, p, T)
        out.append(sub)
    return out


def muon_operator(Is, muon_index, v):
    """v.S_mu in the full space (spinsys.py:707-732)."""
    ops = _spin_ops(tuple(Is))
    return sum(v[a] * ops[muon_index, a] for a in range(3))


def sigma_mu(v):
    """v.sigma (Pauli, 2x2) (spinsys.py:734-760)."""
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sy = np.array([[0, -1j], [1j, 0]], dtype=complex)
    sz = np.array([[1, 0], [0, -1]], dtype=complex)
    return v[0] * sx + v[1] * sy + v[2] * sz


def get_xy(z):
    """Two axes orthogonal to z (utils.py:71-93)."""
    zn = np.linalg.norm(z)
    if zn == 0:
        raise ValueError("Can not find X and Y for null Z vector")
    z = np.array(z, dtype=float) / zn
    if z[0] == 0 and z[1] == 0:
        x = np.array([1.0, 0, 0])
        y = np.array([0, z[2], 0])
    else:
        x = np.array([z[1], -z[0], 0.0]) / (z[0] ** 2 + z[1] ** 2) ** 0.5
        y = np.array([z[1] * x[2] - z[2] * x[1],
                      z[2] * x[0] - z[0] * x[2],
                      z[0] * x[1] - z[1] * x[0]])
    return x, y


def dissipation_operators(Is, gammas, B, T, dissip):
    """Jump operators S+/S- in the frame perpendicular to B, with rates
    (a/pi)Z/(1+Z) and (a/pi)/(1+Z) (experiment.py:271-325).

    dissip: {spin index: a (MHz)}.  Returns [(A, gamma), ...]."""
    B = np.asarray(B, dtype=float)
    Bn = np.linalg.norm(B)
    g = np.asarray(gammas, dtype=float)
    if T > 0:
        Zu = np.exp(-cnst.h * g * Bn * 1e6 / (cnst.k * T))
    else:
        Zu = g * 0.0
    if Bn == 0:
        x, y = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    else:
        x, y = get_xy(B / Bn)
    out = []
    for i, a in dissip.items():
        ox = sum(spin.spin_op(Is, i, _XYZ[k]) * x[k] for k in range(3))
        oy = sum(spin.spin_op(Is, i, _XYZ[k]) * y[k] for k in range(3))
        out.append((ox + 1.0j * oy, 1 / np.pi * a * Zu[i] / (1 + Zu[i])))
        out.append((ox - 1.0j * oy, 1 / np.pi * a / (1 + Zu[i])))
    return out


def lindbladian(H, dissipators):
    """L = -i[H, .] + sum_k g_k (A rho A^+ - 1/2 {A^+A, rho}) on row-major
    vec(rho) (lindbladian.py:19-41)."""
    L = -1.0j * spin.commutator(H)
    for A, gam in dissipators:
        AA = A.conj().T @ A
        L = L + gam * (spin.bracket(A) - 0.5 * spin.anticommutator(AA))
    return L
