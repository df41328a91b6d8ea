"""Hamiltonian of the drop-in surface (muspinsim/hamiltonian.py:19-217).

Same constructor, methods, argument meaning, return types and exceptions as
the reference; the eigendecomposition, basis changes, time sums and
integrals run on the GPU (libmuspin_hip.so through gpu.py).
"""

import numpy as np

from . import gpu
from .spinop import DensityOperator, Hermitian, Operator, SpinOperator
from .validation import (validate_evolve_params, validate_integrate_decaying_params,
                         validate_times)


class Hamiltonian(Operator, Hermitian):
    def __init__(self, matrix, dim=None, use_sparse=True):
        """Hermitian operator; ValueError if the matrix is not square or not
        Hermitian within 1e-6 (spinop.py:47-49,86)."""
        super().__init__(matrix, dim, use_sparse=use_sparse)

    @classmethod
    def from_spin_operator(cls, spinop):
        return cls(spinop.matrix, spinop.dimension)

    def evolve(self, rho0, times, operators=None):
        """Expectation values (T, len(operators)) complex128 of the operators
        on rho(t) = exp(-2 pi i H t) rho0 exp(2 pi i H t), t in microseconds
        and H in MHz; with no operators, the list of DensityOperators rho(t)
        (hamiltonian.py:40-117)."""
        if operators is None:
            operators = []
        times = np.array(times)
        if isinstance(operators, SpinOperator):
            operators = [operators]
        validate_evolve_params(rho0, times, operators)
        if len(operators) > 0:
            if times.shape[0] == 0:
                return np.zeros((0, len(operators)), dtype=complex)
            return gpu.expectations(self, rho0, operators, times)
        dim = rho0.dimension
        return [DensityOperator(m, dim) for m in gpu.evolve_states(self, rho0, times)]

    def integrate_decaying(self, rho0, tau, operators):
        """Integral over t in [0, inf) of exp(-t/tau) <O>(t), one complex value
        per operator (hamiltonian.py:119-164)."""
        if isinstance(operators, SpinOperator):
            operators = [operators]
        validate_integrate_decaying_params(rho0, tau, operators)
        return gpu.integrals(self, rho0, operators, tau)

    def fast_evolve(self, sigma_mu, times, other_dimension):
        """Muon polarisation P(t) (real, (T,)) along sigma_mu for a muon that
        is the first spin of the system and every other spin at infinite
        temperature (hamiltonian.py:166-217)."""
        times = np.array(times)
        validate_times(times)
        if times.shape[0] == 0:
            return np.zeros(0)
        return gpu.fast_evolve(self, sigma_mu, times, other_dimension)
