"""Operator algebra of the drop-in surface (muspinsim/spinop.py).

Operators are built on the host as scipy CSR (or dense) complex matrices,
exactly as the reference's API constructs them; the only numerical kernel of
this layer, Hermitian.diag (spinop.py:51-82), runs on the GPU through
libmuspin_hip.so.
"""

from numbers import Number

import numpy as np
from scipy import sparse


def _mvals(I):
    return np.linspace(I, -I, int(round(2 * I + 1)))


def spin_matrices(I):
    """Dense (Sx, Sy, Sz) for spin I with m = I, ..., -I (spinop.py:14-41)."""
    d = int(round(2 * I + 1))
    k = np.arange(d - 1)
    sp = np.sqrt((k + 1.0) * (d - 1.0 - k))
    Sp = np.zeros((d, d), dtype=complex)
    Sp[k, k + 1] = sp
    Sx = 0.5 * (Sp + Sp.T)
    Sy = 0.5j * (Sp.T - Sp)
    Sz = np.diag(_mvals(I)).astype(complex)
    return Sx, Sy, Sz


def _single(I, axis):
    Sx, Sy, Sz = spin_matrices(I)
    if axis == "x":
        return Sx
    if axis == "y":
        return Sy
    if axis == "z":
        return Sz
    if axis == "+":
        return Sx + 1j * Sy
    if axis == "-":
        return Sx - 1j * Sy
    if axis == "0":
        return np.eye(len(Sz), dtype=complex)
    raise ValueError(f"{axis} is not a valid spin axis")


class Hermitian:
    """Mixin for Hermitian operators (spinop.py:44-82)."""

    def __init__(self):
        if not self.is_hermitian:
            raise ValueError("Operator must be hermitian")

    def diag(self):
        """(eigenvalues ascending, eigenvector columns), computed on the GPU and
        cached until the matrix changes (spinop.py:51-82)."""
        try:
            dd = self._diagdata
            if dd["matrix"].shape == self._matrix.shape and (
                    abs(dd["matrix"] - self._matrix).max() == 0):
                return dd["eigh"]
        except AttributeError:
            pass
        from . import gpu
        M = self._matrix.toarray() if sparse.issparse(self._matrix) else np.asarray(self._matrix)
        eigh = gpu.eigh(M)
        self._diagdata = {"matrix": self._matrix.copy(), "eigh": eigh}
        return eigh


class Operator:
    """Spin operator with dimensions (spinop.py:85-355)."""

    def __init__(self, matrix, dim=None, herm_tol=1e-6, use_sparse=True):
        self._sparse = use_sparse
        if use_sparse:
            self._matrix = sparse.csr_matrix(matrix, dtype=complex)
        else:
            self._matrix = np.array(matrix, dtype=complex)
        if not matrix.shape[0] == matrix.shape[1]:
            raise ValueError("Matrix passed to Operator must be square")
        if dim is None:
            dim = (matrix.shape[0],)
        elif np.prod(dim) != matrix.shape[0]:
            raise ValueError("Dimensions are not compatible with matrix")
        self._dim = tuple(int(x) for x in dim)
        self._htol = herm_tol
        super().__init__()

    def clone(self):
        ans = self.__class__.__new__(self.__class__)
        ans.__dict__ = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in self.__dict__.items()}
        ans.__dict__.pop("_diagdata", None)
        return ans

    @property
    def Is(self):
        return tuple((d - 1) / 2.0 for d in self._dim)

    @property
    def dimension(self):
        return self._dim

    @property
    def N(self):
        return int(np.prod(self._dim))

    @property
    def matrix(self):
        return self._matrix

    @property
    def is_hermitian(self):
        d = self._matrix - self._matrix.conjugate().T
        return abs(d).max() < self._htol if d.shape[0] else True

    @property
    def is_sparse(self):
        return self._sparse

    def dense(self):
        return self._matrix.toarray() if self._sparse else np.asarray(self._matrix)

    def dagger(self):
        ans = self.__class__.__new__(self.__class__)
        ans._dim = tuple(self._dim)
        ans._matrix = self._matrix.conjugate().T
        if self._sparse:
            ans._matrix = sparse.csr_matrix(ans._matrix)
        ans._sparse = self._sparse
        ans._htol = getattr(self, "_htol", 1e-6)
        return ans

    def _eye(self, x):
        n = self._matrix.shape[0]
        return sparse.eye(n, format="csr") * x if self._sparse else np.eye(n) * x

    def __add__(self, x):
        if isinstance(x, Operator):
            if self.dimension != x.dimension:
                raise ArithmeticError("Can not add to Operators with different dimensions")
            ans = self.clone()
            ans._matrix = ans._matrix + x._matrix
            return ans
        if isinstance(x, Number):
            ans = self.clone()
            ans._matrix = ans._matrix + self._eye(x)
            return ans
        raise TypeError("Unsupported operation for Operator")

    def __sub__(self, x):
        if isinstance(x, Operator):
            if self.dimension != x.dimension:
                raise ArithmeticError("Can not subtract to Operators with different dimensions")
            ans = self.clone()
            ans._matrix = ans._matrix - x._matrix
            return ans
        if isinstance(x, Number):
            return self + (-x)
        raise TypeError("Unsupported operation for Operator")

    def __mul__(self, x):
        if isinstance(x, Operator):
            if self.dimension != x.dimension:
                raise ArithmeticError("Can not multiply to Operators with different dimensions")
            ans = self.clone()
            ans._matrix = ans._matrix @ x._matrix
            if self._sparse:
                ans._matrix = sparse.csr_matrix(ans._matrix)
            return ans
        if isinstance(x, Number):
            ans = self.clone()
            ans._matrix = ans._matrix * x
            return ans
        raise TypeError("Unsupported operation for Operator")

    def __rmul__(self, x):
        if isinstance(x, Number):
            return self.__mul__(x)
        if isinstance(x, Operator):
            return x * self
        raise TypeError("Unsupported operation for Operator")

    def __truediv__(self, x):
        if isinstance(x, Number):
            return self * (1.0 / x)
        raise TypeError("Unsupported operation for Operator")

    def __eq__(self, x):
        if not isinstance(x, Operator) or self.dimension != x.dimension:
            return False
        return abs(self._matrix - x._matrix).max() == 0 if self._matrix.shape[0] else True

    __hash__ = None

    def kron(self, x):
        if not isinstance(x, Operator):
            raise ValueError("Can only perform Kronecker product with another Operator")
        ans = self.__class__.__new__(self.__class__)
        ans._dim = self._dim + x._dim
        ans._sparse = self._sparse
        ans._htol = self._htol
        if self._sparse:
            ans._matrix = sparse.kron(self._matrix, x._matrix, format="csr")
        else:
            ans._matrix = np.kron(self._matrix, x._matrix)
        return ans

    def hilbert_schmidt(self, x):
        if not isinstance(x, Operator):
            raise ValueError("Can only perform Hilbert-Schmidt product with another Operator")
        if not x.dimension == self.dimension:
            raise ValueError("Operators must have the same dimension to perform "
                             "Hilbert-Schmidt product")
        return complex((self._matrix.conjugate().T @ x._matrix).diagonal().sum())

    def basis_change(self, basis, use_sparse=False):
        """U^dagger M U (spinop.py:330-355); the GPU path uses msp_gemm_cn, this
        host form serves the operator API."""
        ans = self.clone()
        x = np.asarray(basis).T.conjugate()
        m = x @ (ans._matrix @ np.asarray(basis))
        ans._matrix = sparse.csr_matrix(m) if use_sparse else np.asarray(m)
        ans._sparse = use_sparse
        return ans


class SpinOperator(Operator):
    @classmethod
    def from_axes(cls, Is=0.5, axes="x", use_sparse=True):
        """spinop.py:359-411."""
        if isinstance(Is, Number):
            Is = [Is]
        if len(Is) != len(axes) or len(Is) == 0:
            raise ValueError("Arrays of moments and axes must have same length > 0")
        dim = tuple(int(round(2 * I + 1)) for I in Is)
        mats = []
        for I, axis in zip(Is, axes):
            if I % 0.5 or I < 0.5:
                raise ValueError(f"{I} is not a valid spin value")
            if axis not in "xyz+-0":
                raise ValueError(f"{axis} is not a valid spin axis")
            mats.append(_single(I, axis))
        M = mats[0]
        for m in mats[1:]:
            M = np.kron(M, m)
        return cls(M, dim=dim, use_sparse=use_sparse)


class DensityOperator(Operator):
    def __init__(self, matrix, dim=None, use_sparse=True):
        """Trace-normalised Hermitian state (spinop.py:414-453)."""
        super().__init__(matrix, dim, use_sparse=use_sparse)
        tr = self._matrix.diagonal().sum()
        if tr == 0:
            raise ValueError("Can not define a DensityOperator with zero trace")
        self.normalize()
        if not self.is_hermitian:
            raise ValueError("DensityOperator must be hermitian!")

    @classmethod
    def from_vectors(cls, Is=0.5, vectors=(0, 0, 1), gammas=0, use_sparse=True):
        """Kron of per-spin pure states along vectors with decoherence gamma
        (spinop.py:455-530)."""
        if isinstance(Is, Number):
            Is = [Is]
        if len(np.array(vectors).shape) == 1:
            vectors = [vectors]
        if isinstance(gammas, Number):
            gammas = [gammas]
        if len(Is) != len(vectors) or len(Is) != len(gammas) or len(Is) == 0:
            raise ValueError("Arrays of moments, axes and gammas must have same length > 0")
        dim = tuple(int(round(2 * I + 1)) for I in Is)
        mats = []
        for I, vec, g in zip(Is, vectors, gammas):
            if I % 0.5 or I < 0.5:
                raise ValueError(f"{I} is not a valid spin value")
            if not len(vec) == 3:
                raise ValueError(f"{vec} is not a valid 3D vector")
            if g < 0 or g > 1:
                raise ValueError(f"{g} is not a valid gamma value")
            S = spin_matrices(I)
            o = sum(S[i] * vec[i] for i in range(3))
            ev, evec = np.linalg.eigh(o)
            psi = evec[:, np.argmax(ev)]
            m = psi[:, None] * psi[None, :].conj()
            m = m * ((1 - g) * np.ones(m.shape) + g * np.eye(m.shape[0]))
            mats.append(m)
        M = mats[0]
        for m in mats[1:]:
            M = np.kron(M, m)
        return cls(M, dim=dim, use_sparse=use_sparse)

    @property
    def trace(self):
        return self._matrix.diagonal().sum()

    def normalize(self):
        t = self.trace
        if self._sparse:
            self._matrix = sparse.csr_matrix(self._matrix.multiply(1 / t))
        else:
            self._matrix = np.multiply(self._matrix, 1 / t)

    def partial_trace(self, trace_dim=None):
        """spinop.py:543-571."""
        trace_dim = [] if trace_dim is None else trace_dim
        dim = list(self._dim)
        tdim = list(sorted(trace_dim))
        m = self.dense().reshape(dim + dim)
        while tdim:
            td = tdim.pop(-1)
            m = np.trace(m, axis1=td, axis2=td + len(dim))
            dim.pop(td)
        n = int(np.prod(dim))
        return DensityOperator(m.reshape(n, n), dim, use_sparse=self._sparse)

    def expectation(self, operator):
        """spinop.py:573-600."""
        if not isinstance(operator, SpinOperator):
            raise TypeError("Argument must be a SpinOperator")
        if not operator.dimension == self.dimension:
            raise ValueError("SpinOperator and DensityOperator do not have compatible dimensions")
        a = operator.dense()
        return np.sum(a * self.dense().T)


class SuperOperator(Operator):
    def __init__(self, matrix, dim=None):
        """Liouville-space operator on row-major vec(rho) (spinop.py:603-613)."""
        n = matrix.shape[0] ** 0.5
        if int(n) != n:
            raise ValueError("Matrix rank should be a perfect square")
        if dim is None:
            dim = (int(n), int(n))
        super().__init__(matrix, dim)

    @classmethod
    def left_multiplier(cls, operator):
        m = operator.matrix
        return cls(sparse.kron(m, sparse.eye(m.shape[0]), format="csr"), operator.dimension * 2)

    @classmethod
    def right_multiplier(cls, operator):
        m = operator.matrix
        return cls(sparse.kron(sparse.eye(m.shape[0]), m.T, format="csr"), operator.dimension * 2)

    @classmethod
    def commutator(cls, operator):
        return cls.left_multiplier(operator) - cls.right_multiplier(operator)

    @classmethod
    def anticommutator(cls, operator):
        return cls.left_multiplier(operator) + cls.right_multiplier(operator)

    @classmethod
    def bracket(cls, operator):
        m = operator.matrix
        return cls(sparse.kron(m, m.conjugate(), format="csr"), operator.dimension * 2)

    def __add__(self, x):
        if isinstance(x, (SuperOperator, Number)):
            return super().__add__(x)
        raise TypeError("Unsupported operation for SuperOperator")

    def __sub__(self, x):
        if isinstance(x, (SuperOperator, Number)):
            return super().__sub__(x)
        raise TypeError("Unsupported operation for SuperOperator")

    def __mul__(self, x):
        if isinstance(x, (SuperOperator, Number)):
            return super().__mul__(x)
        if isinstance(x, Operator):
            m = x.dense()
            s = m.shape
            return Operator((self._matrix @ m.reshape(-1)).reshape(s), x.dimension)
        raise TypeError("Unsupported operation for Operator")
