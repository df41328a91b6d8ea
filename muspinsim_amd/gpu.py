"""Process-wide default engine and the single-Hamiltonian operations behind
the drop-in classes (Hermitian.diag, Hamiltonian.evolve / integrate_decaying /
fast_evolve, SpinSystem.hamiltonian).  Every numerical step is a call into
libmuspin_hip.so; torch only holds device memory and reshapes it.
"""

import numpy as np
import torch
from scipy import sparse

from . import _native as N

_ENGINE = None


def engine():
    """The default SpinEngine on cuda:0 (raises if no HIP device or library)."""
    global _ENGINE
    if _ENGINE is None:
        if not torch.cuda.is_available():
            raise RuntimeError("muspinsim_amd needs an MI355X (HIP device); none is visible")
        from .engine import SpinEngine
        _ENGINE = SpinEngine("cuda:0")
    return _ENGINE


def dense(M):
    """Dense complex128 ndarray of a matrix, sparse matrix or Operator."""
    if hasattr(M, "matrix"):
        M = M.matrix
    if sparse.issparse(M):
        M = M.toarray()
    return np.asarray(M, dtype=complex)


def to_dev(a, dtype=torch.complex128):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(engine().device)


def _eigh_dev(H):
    """E (1, n), U (1, n, n) on the device for one Hermitian matrix."""
    eng = engine()
    E, U, info = eng.eigh(to_dev(dense(H)[None]))
    if int(info.max().item()) != 0:
        raise np.linalg.LinAlgError("Eigenvalues did not converge")
    return E, U


def eigh(M):
    """numpy.linalg.eigh contract (ascending E, eigenvector columns) for one
    dense Hermitian matrix, computed by msp_eigh."""
    E, U = _eigh_dev(M)
    return E[0].cpu().numpy(), U[0].cpu().numpy()


def _rep(X, count):
    """X (1, ...) repeated along the leading axis to (count, ...)."""
    return X.expand((count,) + tuple(X.shape[1:])).contiguous() if X.shape[0] != count else X


def _basis_change_dev(mats, U):
    """U^H M U for a stack of host matrices (k, n, n), U (1, n, n) on the device:
    M U = (M^H)^H U and U^H (M U) are both msp_gemm_cn products."""
    eng = engine()
    k = mats.shape[0]
    Uk = _rep(U, k)
    MU = eng.gemm_cn(to_dev(np.conj(np.swapaxes(mats, 1, 2))), Uk)
    return eng.gemm_cn(Uk, MU)


def expectations(H, rho0, ops, times):
    """(T, k) complex: sum_ij exp(-2 pi i (E_i - E_j) t) rho'_ij O'_ji for each
    operator, the evolve contract of hamiltonian.py:40-107."""
    eng = engine()
    E, U = _eigh_dev(H)
    k = len(ops)
    rho = _basis_change_dev(dense(rho0)[None], U)
    O = _basis_change_dev(np.array([dense(o) for o in ops]), U)
    C = eng.form_c(_rep(rho, k), O, False)
    t = to_dev(np.asarray(times, dtype=float), torch.float64)
    out = eng.timesum(_rep(E, k), C, t, False, None, True)
    return out.cpu().numpy().T.copy()


def integrals(H, rho0, ops, tau):
    """(k,) complex: sum_ij rho'_ij O'_ji / (1/tau + 2 pi i (E_i - E_j)), the
    integrate_decaying contract of hamiltonian.py:119-164."""
    eng = engine()
    E, U = _eigh_dev(H)
    k = len(ops)
    rho = _basis_change_dev(dense(rho0)[None], U)
    O = _basis_change_dev(np.array([dense(o) for o in ops]), U)
    C = eng.form_c(_rep(rho, k), O, False)
    out = eng.integral(_rep(E, k), C, False, float(tau), 1.0, None, True)
    return out.cpu().numpy()


def fast_evolve(H, sigma, times, other_dimension):
    """(T,) real: (1/(4d)) sum_jk |U^H (sigma (x) 1_d) U|^2_jk cos(2 pi (E_j - E_k) t),
    the fast_evolve contract of hamiltonian.py:166-217 (its cosine sum over
    the lower triangle times 1/2, cython/parallel.pyx:59-67)."""
    eng = engine()
    E, U = _eigh_dev(H)
    d = int(other_dimension)
    n = U.shape[-1]
    s = dense(sigma)
    if s.shape != (2, 2) or 2 * d != n:
        raise ValueError("sigma_mu must be 2 x 2 and 2 * other_dimension must equal the "
                         "Hamiltonian's dimension")
    F = to_dev(np.concatenate([s.reshape(-1), np.eye(d, dtype=complex).reshape(-1)])[None])
    Y = eng.apply_kron([2, d], F, U)
    O = eng.gemm_cn(U, Y)
    C = eng.form_c(None, O, True, 1.0 / (4.0 * d), 0.0)
    t = to_dev(np.asarray(times, dtype=float), torch.float64)
    return eng.timesum(E, C, t, True).cpu().numpy()[0]


def evolve_states(H, rho0, times):
    """rho(t) = U (exp(-2 pi i (E_i - E_j) t) rho'_ij) U^H for every t, each a
    dense (n, n) array (hamiltonian.py:108-116)."""
    eng = engine()
    E, U = _eigh_dev(H)
    rho = _basis_change_dev(dense(rho0)[None], U)
    t = to_dev(np.asarray(times, dtype=float), torch.float64)
    X = eng.phase_evolve(E, rho, t)                       # (T, n, n), Hermitian each
    T = X.shape[0]
    Uh = _rep(U.conj().transpose(-1, -2).contiguous(), T)  # A = U^H
    XA = eng.gemm_cn(X, Uh)                                # X^H A = X A
    R = eng.gemm_cn(Uh, XA)                                # A^H X A = U X U^H
    return list(R.cpu().numpy())


def build_h(system, B=None):
    """Dense Hamiltonian of a SystemDesc built by msp_build_h: the sum of its
    terms, plus the Zeeman term of field B [T] when given."""
    from .engine import pack_systems
    eng = engine()
    L = N.lib()
    sysarr, termarr = pack_systems([system])
    dsys, dterms = eng._bytes_view(sysarr), eng._bytes_view(termarr)
    n = system.n
    sid = to_dev(np.zeros(1, np.int32), torch.int32)
    H = torch.empty((1, n, n), dtype=torch.complex128, device=eng.device)
    if B is None:
        N.check(L.msp_build_h(n, 1, dsys.data_ptr(), dterms.data_ptr(), sid.data_ptr(), None,
                              None, 0, H.data_ptr(), eng.stream()))
    else:
        Br = to_dev(np.asarray(B, dtype=float).reshape(1, 3), torch.float64)
        N.check(L.msp_build_h(n, 1, dsys.data_ptr(), dterms.data_ptr(), sid.data_ptr(),
                              Br.data_ptr(), None, 1, H.data_ptr(), eng.stream()))
    return H[0].cpu().numpy()
