"""Spin systems and interaction terms of the drop-in surface
(muspinsim/spinsys.py:22-760).

A system is a list of spins (muon, electron, nuclei) in Kronecker order --
the first spin is the slowest index (spinsys.py:587-590) -- and a list of
terms T . S_i (linear) or S_i . M . S_j (bilinear).  The term formulas
(Zeeman, dipolar, quadrupolar, hyperfine) follow spinsys.py:354-442,664-705.
`SpinSystem.hamiltonian` is assembled on the GPU (msp_build_h) from the same
plain-data description (`to_desc`) the batched engine uses.
"""

import copy
import logging
from numbers import Number

import numpy as np
import scipy.constants as cnst

from .constants import EFG_2_MHZ, gyromagnetic_ratio, quadrupole_moment, spin
from .engine import SystemDesc
from .spinop import SpinOperator


class InteractionTerm:
    """Tensor T contracted with the spin vectors of `indices`
    (spinsys.py:22-92)."""

    def __init__(self, spinsys, indices=(), tensor=0, label=None):
        self._spinsys = spinsys
        self._indices = np.array(indices, dtype=int)
        self._tensor = np.array(tensor)
        self._label = "Term" if label is None else label
        if self._tensor.ndim != len(self._indices) or any(s != 3 for s in self._tensor.shape):
            raise ValueError("Tensor is not fit for the number of spin indices")

    def clone(self):
        ans = copy.copy(self)
        ans._tensor = self._tensor.copy()
        ans._indices = self._indices.copy()
        return ans

    @property
    def label(self):
        return self._label

    @property
    def indices(self):
        return tuple(self._indices)

    @property
    def tensor(self):
        return np.array(self._tensor)

    @property
    def operator(self):
        """SpinOperator sum_{a,b,...} T_ab... S_i^a S_j^b ... (host operator
        algebra, used by tests and by the operator API)."""
        ss = self._spinsys
        ops = [[ss.operator({i: a}) for a in "xyz"] for i in self._indices]
        if len(ops) == 0:
            return ss.operator({}) * float(self._tensor)
        total = None
        for idx in np.ndindex(*self._tensor.shape):
            c = self._tensor[idx]
            if c == 0:
                continue
            term = ops[0][idx[0]]
            for k in range(1, len(idx)):
                term = term * ops[k][idx[k]]
            term = term * c
            total = term if total is None else total + term
        if total is None:
            total = ss.operator({}) * 0.0
        return total

    @property
    def matrix(self):
        return self.operator.matrix

    def __repr__(self):
        return self._label


class SingleTerm(InteractionTerm):
    def __init__(self, spinsys, i, vector, label="Single"):
        super().__init__(spinsys, [i], vector, label)

    @property
    def i(self):
        return self._indices[0]

    def rotate(self, rotmat):
        R = np.array(rotmat)
        return SingleTerm(self._spinsys, self.i, R.dot(self._tensor), self._label)

    def __repr__(self):
        return "{0} {{ S_{1} * {2} }}".format(self._label, self.i, self._tensor)


class DoubleTerm(InteractionTerm):
    def __init__(self, spinsys, i, j, matrix, label="Double"):
        super().__init__(spinsys, [i, j], matrix, label)

    @property
    def i(self):
        return self._indices[0]

    @property
    def j(self):
        return self._indices[1]

    def rotate(self, rotmat):
        R = np.array(rotmat)
        return DoubleTerm(self._spinsys, self.i, self.j, R.dot(self._tensor.dot(R.T)),
                          self._label)

    def __repr__(self):
        return "{0} {{ S_{1} * [{2}] * S_{3} }}".format(self._label, self.i, self._tensor,
                                                       self.j)


class DissipationTerm:
    """Jump operator with a rate in MHz (spinsys.py:143-158)."""

    def __init__(self, operator, gamma=0.0):
        self._op = operator
        self._g = gamma

    @property
    def operator(self):
        return self._op

    @property
    def gamma(self):
        return self._g

    @property
    def tuple(self):
        return (self._op, self._g)


class SpinSystem:
    def __init__(self, spins=None, celio_k=0):
        """Spins as 'e', 'mu', an element symbol, or (symbol, mass number)
        (spinsys.py:161-216).  celio_k > 0: the hamiltonian is a
        CelioHamiltonian with that Trotter factor."""
        if int(celio_k) < 0:
            raise ValueError("celio_k must be a non-negative integer")
        spins = [] if spins is None else spins
        gammas, Qs, Is = [], [], []
        for s in spins:
            el, iso = s if isinstance(s, tuple) else (s, None)
            gammas.append(gyromagnetic_ratio(el, iso))
            Qs.append(quadrupole_moment(el, iso))
            Is.append(spin(el, iso))
        self._spins = list(spins)
        self._gammas = np.array(gammas, dtype=float)
        self._Qs = np.array(Qs, dtype=float)
        self._Is = np.array(Is, dtype=float)
        self._dim = tuple((2 * self._Is + 1).astype(int))
        self._operators = [{a: SpinOperator.from_axes(I, a) for a in "xyz+-0"} for I in self._Is]
        self._terms = []
        self._dissip_terms = []
        self._celio_k = int(celio_k)

    def clone(self):
        ans = copy.copy(self)
        ans._spins = list(self._spins)
        ans._terms = [t.clone() for t in self._terms]
        for t in ans._terms:
            t._spinsys = ans
        ans._dissip_terms = list(self._dissip_terms)
        return ans

    @property
    def spins(self):
        return list(self._spins)

    @property
    def celio_k(self):
        return self._celio_k

    @property
    def gammas(self):
        return self._gammas.copy()

    @property
    def Qs(self):
        return self._Qs.copy()

    @property
    def Is(self):
        return self._Is.copy()

    @property
    def dimension(self):
        return self._dim

    @property
    def is_dissipative(self):
        return any(t.gamma != 0.0 for t in self._dissip_terms)

    @property
    def terms(self):
        return list(self._terms)

    def _check_index(self, i, name="i"):
        if i < 0 or i >= len(self._spins):
            raise ValueError(f"Invalid index {name}")

    def add_term(self, indices, tensor, label="Term"):
        for i in indices:
            self._check_index(i)
        term = InteractionTerm(self, indices, np.array(tensor), label=label)
        self._terms.append(term)
        return term

    def add_linear_term(self, i, vector, label="Single"):
        self._check_index(i)
        term = SingleTerm(self, i, np.array(vector), label=label)
        self._terms.append(term)
        return term

    def add_bilinear_term(self, i, j, matrix, label="Double"):
        self._check_index(i)
        self._check_index(j, "j")
        term = DoubleTerm(self, i, j, np.array(matrix), label=label)
        self._terms.append(term)
        return term

    def add_zeeman_term(self, i, B):
        """gamma_i B . S_i, B in tesla (a number is taken along z)
        (spinsys.py:354-375)."""
        if isinstance(B, Number):
            B = [0, 0, B]
        logging.info("Adding Zeeman term to spin %s", i + 1)
        return self.add_linear_term(i, np.array(B, dtype=float) * self.gamma(i), "Zeeman")

    def add_dipolar_term(self, i, j, r):
        """Point-dipole coupling from the vector r (angstrom) joining spins i
        and j: D = -(mu0 hbar gamma_i gamma_j 1e6 / (2 |r|^3)) (1 - 3 r^ r^)
        in MHz (spinsys.py:377-413)."""
        if i == j:
            raise ValueError("Can not set up dipolar coupling with itself")
        r = np.array(r, dtype=float)
        rn = np.linalg.norm(r)
        scale = -(cnst.mu_0 * cnst.hbar * (self.gamma(i) * self.gamma(j) * 1e6)) / (
            2 * (rn * 1e-10) ** 3)
        D = -(np.eye(3) - 3.0 * np.outer(r, r) / rn ** 2) * scale
        logging.info("Adding dipolar term to spins %s-%s", i + 1, j + 1)
        return self.add_bilinear_term(i, j, D, "Dipolar")

    def add_quadrupolar_term(self, i, EFG):
        """EFG_2_MHZ Q / (2 I (2 I - 1)) EFG between S_i and itself, I >= 1
        (spinsys.py:415-442)."""
        I = self.I(i)
        if I == 0.5:
            raise ValueError("Can not set up quadrupolar coupling for spin 1/2 particle")
        Qt = EFG_2_MHZ * self.Q(i) / (2 * I * (2 * I - 1)) * np.array(EFG, dtype=float)
        logging.info("Adding quadrupolar term to spin %s", i + 1)
        return self.add_bilinear_term(i, i, Qt, "Quadrupolar")

    def remove_term(self, term):
        self._terms.remove(term)

    def clear_terms(self):
        self._terms = []

    def add_dissipative_term(self, op, d=0.0):
        term = DissipationTerm(op, d)
        self._dissip_terms.append(term)
        return term

    def remove_dissipative_term(self, term):
        self._dissip_terms.remove(term)

    def clear_dissipative_terms(self):
        self._dissip_terms = []

    def gamma(self, i):
        return self._gammas[i]

    def Q(self, i):
        return self._Qs[i]

    def I(self, i):
        return self._Is[i]

    def operator(self, terms=None, include_only_given=False):
        """Kronecker product of one-spin operators: `terms` maps spin index
        to 'x', 'y', 'z', '+', '-', '0' or a list of them (their product);
        absent spins contribute the identity unless include_only_given
        (spinsys.py:545-592)."""
        terms = {} if terms is None else terms
        factors = []
        for i in range(len(self)):
            sym = terms.get(i, None if include_only_given else "0")
            if not sym:
                continue
            if isinstance(sym, list):
                f = self._operators[i][sym[0]]
                for s in sym[1:]:
                    f = f * self._operators[i][s]
            else:
                f = self._operators[i][sym]
            factors.append(f)
        M = factors[0]
        for f in factors[1:]:
            M = M.kron(f)
        return M

    def rotate(self, rotmat=np.eye(3)):
        """Copy of the system with every linear / bilinear tensor rotated
        (spinsys.py:594-611)."""
        ans = self.clone()
        try:
            ans._terms = [t.rotate(rotmat) for t in self._terms]
        except AttributeError as exc:
            raise RuntimeError(
                "Can only rotate SpinSystems containing Single or Double terms") from exc
        for t in ans._terms:
            t._spinsys = ans
        return ans

    def to_desc(self):
        """Plain-data description (SystemDesc) consumed by the engine."""
        terms = []
        for t in self._terms:
            idx = t.indices
            if len(idx) == 0:
                terms.append(("scalar", 0, 0, np.array([float(t.tensor)])))
            elif len(idx) == 1:
                terms.append(("linear", idx[0], idx[0], t.tensor.astype(float)))
            elif len(idx) == 2:
                terms.append(("bilinear", idx[0], idx[1], t.tensor.astype(float)))
            else:
                raise NotImplementedError("terms in more than two spins are not supported "
                                          "by the GPU Hamiltonian builder")
        mu = getattr(self, "_mu_i", 0)
        return SystemDesc(Is=list(self._Is), gammas=list(self._gammas), terms=terms,
                          muon_index=mu)

    @property
    def hamiltonian(self):
        """Hamiltonian (MHz) of all terms, built on the GPU, or the
        CelioHamiltonian of the terms when celio_k > 0 (spinsys.py:613-626)."""
        if self._celio_k:
            from .celio import CelioHamiltonian
            return CelioHamiltonian(self._terms, self._celio_k, self)
        from . import gpu
        from .hamiltonian import Hamiltonian
        return Hamiltonian(gpu.build_h(self.to_desc()), dim=self.dimension)

    @property
    def lindbladian(self):
        if self._celio_k:
            raise NotImplementedError("Dissipation is not supported when using Celio's method")
        from .lindbladian import Lindbladian
        return Lindbladian.from_hamiltonian(self.hamiltonian,
                                            [t.tuple for t in self._dissip_terms])

    def __len__(self):
        return len(self._gammas)


class MuonSpinSystem(SpinSystem):
    def __init__(self, spins=("mu", "e"), celio_k=0):
        """A spin system with exactly one muon (spinsys.py:640-662)."""
        super().__init__(list(spins), celio_k)
        if self._spins.count("mu") != 1:
            raise ValueError("Spins passed to MuonSpinSystem must contain exactly one muon")
        self._mu_i = self._spins.index("mu")
        self._e_i = {i for i, s in enumerate(self._spins) if s == "e"}
        self._mu_ops = None

    @property
    def muon_index(self):
        return self._mu_i

    @property
    def elec_indices(self):
        return self._e_i

    def add_hyperfine_term(self, i, A, j=None):
        """S_i . A . S_e with A in MHz; j must be an electron (the only one
        when omitted) and i must not be (spinsys.py:664-705)."""
        elec = self.elec_indices
        if j is None:
            if len(elec) > 1:
                raise ValueError("Must specify an electron index in system with multiple "
                                 "electrons")
            j = list(elec)[0]
        elif j not in elec:
            raise ValueError("Second index in hyperfine coupling must refer to an electron")
        if i in elec:
            raise ValueError("First index in hyperfine coupling must not refer to an electron")
        logging.info("Adding hyperfine term to spins %s-%s", i + 1, j + 1)
        return self.add_bilinear_term(i, j, A, "Hyperfine")

    def muon_operator(self, v):
        """v . S_mu on the whole system (spinsys.py:707-732)."""
        if len(v) != 3:
            raise ValueError("Vector passed to muon_operator must be three dimensional")
        if self._mu_ops is None:
            self._mu_ops = [self.operator({self._mu_i: a}) for a in "xyz"]
        op = self._mu_ops[0] * v[0]
        for k in (1, 2):
            op = op + self._mu_ops[k] * v[k]
        return op

    def sigma_mu(self, v):
        """v . (sigma_x, sigma_y, sigma_z), 2 x 2 (spinsys.py:734-760)."""
        if len(v) != 3:
            raise ValueError("Vector passed to muon_operator must be three dimensional")
        from scipy import sparse
        sx = np.array([[0, 1], [1, 0]], dtype=complex)
        sy = np.array([[0, -1j], [1j, 0]], dtype=complex)
        sz = np.array([[1, 0], [0, -1]], dtype=complex)
        return sparse.csr_matrix(v[0] * sx + v[1] * sy + v[2] * sz)
