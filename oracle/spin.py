"""Spin matrices and Kronecker-structured operators (oracle).

Restates muspinsim/spinop.py:14-41 (single-spin matrices), :266-295 (kron),
:330-355 (basis change), :455-530 (DensityOperator.from_vectors) and
:615-715 (superoperators), as plain dense complex128 numpy functions.
"""

import numpy as np


def mvals(I):
    """m values I, I-1, ..., -I (spinop.py:14-15)."""
    return np.linspace(I, -I, int(round(2 * I + 1)))


def s_plus(I):
    """Raising operator, <m_k|S+|m_k+1> = sqrt((k+1)(2I-k)) (spinop.py:18-19)."""
    m = mvals(I)
    return 2 * np.diag(0.5 * (np.cumsum(2 * m)[:-1] ** 0.5), k=1) + 0.0j


def s_minus(I):
    """spinop.py:22-23."""
    return s_plus(I).T


def s_x(I):
    """spinop.py:26-28."""
    o = s_plus(I)
    return 0.5 * (o + o.T)


def s_y(I):
    """spinop.py:31-33."""
    o = s_plus(I)
    return 0.5j * (o.T - o)


def s_z(I):
    """spinop.py:36-37."""
    return np.diag(mvals(I)) + 0.0j


def s_0(I):
    """spinop.py:40-41."""
    return np.eye(len(mvals(I))) + 0.0j


_AXES = {"x": s_x, "y": s_y, "z": s_z, "+": s_plus, "-": s_minus, "0": s_0}


def single(I, axis):
    """One spin matrix by axis letter (SpinOperator.from_axes, spinop.py:359-411)."""
    return _AXES[axis](I)


def kron_all(mats):
    """Left-to-right Kronecker product: first factor is the slowest index
    (spinop.py:407-409, spinsys.py:587-590)."""
    M = mats[0]
    for m in mats[1:]:
        M = np.kron(M, m)
    return M


def embed(Is, ops):
    """Full-space operator from {spin index: d x d matrix}, identity elsewhere
    (SpinSystem.operator, spinsys.py:545-592)."""
    mats = [ops.get(k, s_0(I)) for k, I in enumerate(Is)]
    return kron_all(mats)


def spin_op(Is, i, axis):
    """S_i^axis in the full space of spins Is."""
    return embed(Is, {i: single(Is[i], axis)})


def basis_change(M, U):
    """U^dagger M U (Operator.basis_change, spinop.py:330-355)."""
    return U.conj().T @ M @ U


def pure_state(I, vec):
    """|psi><psi| for psi the top eigenvector of vec.S
    (DensityOperator.from_vectors with gamma=0, spinop.py:489-530)."""
    S = [s_x(I), s_y(I), s_z(I)]
    o = sum(S[k] * vec[k] for k in range(3))
    evals, evecs = np.linalg.eigh(o)
    psi = evecs[:, np.argmax(evals)]
    m = psi[:, None] * psi[None, :].conj()
    return m / np.trace(m)


def from_vectors(Is, vectors, gammas):
    """DensityOperator.from_vectors with decoherence gamma (spinop.py:455-530)."""
    mats = []
    for I, vec, g in zip(Is, vectors, gammas):
        S = [s_x(I), s_y(I), s_z(I)]
        o = sum(S[k] * vec[k] for k in range(3))
        evals, evecs = np.linalg.eigh(o)
        psi = evecs[:, np.argmax(evals)]
        m = psi[:, None] * psi[None, :].conj()
        m = m * ((1 - g) * np.ones(m.shape) + g * np.eye(m.shape[0]))
        mats.append(m)
    M = kron_all(mats)
    return M / np.trace(M)


# Superoperators on row-major vec(rho) (spinop.py:615-715)
def left_mult(O):
    """L rho = O rho  (spinop.py:615-635)."""
    return np.kron(O, np.eye(O.shape[0]))


def right_mult(O):
    """L rho = rho O  (spinop.py:637-657)."""
    return np.kron(np.eye(O.shape[0]), O.T)


def commutator(O):
    """spinop.py:659-673."""
    return left_mult(O) - right_mult(O)


def anticommutator(O):
    """spinop.py:675-689."""
    return left_mult(O) + right_mult(O)


def bracket(O):
    """L rho = O rho O^dagger  (spinop.py:691-715)."""
    return np.kron(O, O.conj())
