"""Lindbladian of the drop-in surface (muspinsim/lindbladian.py:17-173).

A SuperOperator on row-major vec(rho).  Construction (from_hamiltonian,
add_dissipative_term) and the evolution methods run on the GPU: the
Liouvillian through msp_liouvillian, its eigendecomposition and the mode
amplitudes through msp_lindblad_modes, the time series and the decaying
integral through msp_mode_sum / msp_mode_integral / msp_mode_states.
"""

import numpy as np
import torch
from scipy import sparse

from . import gpu
from .hamiltonian import Hamiltonian
from .spinop import DensityOperator, SpinOperator, SuperOperator
from .validation import validate_evolve_params, validate_integrate_decaying_params


def _liouvillian_dev(H, dissipators):
    """Dense L on the host, built by msp_liouvillian from H and [(A, gamma)]."""
    eng = gpu.engine()
    Hm = gpu.dense(H)
    n = Hm.shape[0]
    if dissipators:
        A = np.array([gpu.dense(a) for a, _ in dissipators])[None]
        g = np.array([float(np.real(gm)) for _, gm in dissipators])[None]
        if np.any(np.array([np.imag(gm) for _, gm in dissipators]) != 0):
            raise ValueError("dissipation rates must be real")
        Lm = eng.liouvillian(gpu.to_dev(Hm[None]), gpu.to_dev(A),
                             gpu.to_dev(g, torch.float64))
    else:
        Lm = eng.liouvillian(gpu.to_dev(Hm[None]))
    return Lm[0].cpu().numpy().reshape(n * n, n * n)


class Lindbladian(SuperOperator):
    @classmethod
    def from_hamiltonian(cls, H, dissipators=()):
        """L = -i[H, .] plus one Lindblad term per (A, gamma) in dissipators
        (lindbladian.py:18-33).  ValueError if H is not a Hamiltonian."""
        if not isinstance(H, Hamiltonian):
            raise ValueError("Must use Hamiltonian to create Lindbladian")
        dissipators = list(dissipators)
        for A, _ in dissipators:
            if A.dimension != H.dimension:
                raise ValueError("Invalid dissipation operator for this Lindbladian")
        L = _liouvillian_dev(H, dissipators)
        return cls(L, tuple(H.dimension) * 2)

    def add_dissipative_term(self, A, gamma=1.0):
        """L += gamma (A rho A^H - 1/2 {A^H A, rho}) (lindbladian.py:35-41)."""
        n = int(np.prod(A.dimension))
        if tuple(A.dimension) * 2 != tuple(self.dimension):
            raise ValueError("Invalid dissipation operator for this Lindbladian")
        Ld = _liouvillian_dev(np.zeros((n, n), dtype=complex), [(A, gamma)])
        M = self._matrix + Ld
        self._matrix = sparse.csr_matrix(M) if self._sparse else np.asarray(M)

    def _modes(self, rho0, operators, want_states=False):
        eng = gpu.engine()
        Lm = gpu.to_dev(gpu.dense(self)[None])
        r = gpu.to_dev(gpu.dense(rho0).reshape(1, -1))
        u = None
        if operators:
            u = gpu.to_dev(np.array([gpu.dense(o).T.reshape(-1) for o in operators])[None])
        return eng.lindblad_modes(Lm, r, u, want_c=want_states, want_V=want_states)

    def _check_dims(self, rho0, operators):
        dim = tuple(rho0.dimension)
        if tuple(self.dimension) != dim * 2:
            raise ValueError("Incompatible rho0 dimension")
        if any(tuple(self.dimension) != tuple(o.dimension) * 2 for o in operators):
            raise ValueError("Incompatible measure operator dimension")

    def evolve(self, rho0, times, operators=()):
        """Expectation values (T, len(operators)) complex128 of vec(rho(t)) =
        exp(2 pi L t) vec(rho0), t in microseconds; with no operators the
        list of DensityOperators rho(t) (lindbladian.py:43-111)."""
        times = np.array(times)
        if isinstance(operators, SpinOperator):
            operators = [operators]
        operators = list(operators)
        validate_evolve_params(rho0, times, operators)
        self._check_dims(rho0, operators)
        eng = gpu.engine()
        t = gpu.to_dev(np.asarray(times, dtype=float), torch.float64)
        if operators:
            lam, amp, _, _ = self._modes(rho0, operators)
            if times.shape[0] == 0:
                return np.zeros((0, len(operators)), dtype=complex)
            return eng.mode_sum(lam, amp, t)[0].cpu().numpy().T.copy()
        lam, _, c, V = self._modes(rho0, [], want_states=True)
        dim = rho0.dimension
        n = int(np.prod(dim))
        if times.shape[0] == 0:
            return []
        states = eng.mode_states(lam[0], c[0], V[0], t).cpu().numpy()
        return [DensityOperator(s.reshape(n, n), dim) for s in states]

    def integrate_decaying(self, rho0, tau, operators):
        """Integral over t in [0, inf) of exp(-t/tau) <O>(t) per operator
        (lindbladian.py:113-173)."""
        if isinstance(operators, SpinOperator):
            operators = [operators]
        validate_integrate_decaying_params(rho0, tau, operators)
        operators = list(operators)
        self._check_dims(rho0, operators)
        lam, amp, _, _ = self._modes(rho0, operators)
        return gpu.engine().mode_integral(lam, amp, tau)[0].cpu().numpy()
