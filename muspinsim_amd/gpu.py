"""Process-wide default engine and single-matrix helpers used by the
drop-in classes (Hermitian.diag, Hamiltonian.evolve, ...)."""

import numpy as np
import torch

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


def to_dev(a, dtype=torch.complex128):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(engine().device)


def eigh(M):
    """numpy.linalg.eigh contract (ascending E, eigenvector columns) for one
    dense Hermitian matrix, computed by msp_eigh."""
    eng = engine()
    H = to_dev(np.asarray(M, dtype=complex)[None])
    E, U, info = eng.eigh(H)
    if int(info.max().item()) != 0:
        raise np.linalg.LinAlgError("Eigenvalues did not converge")
    return E[0].cpu().numpy(), U[0].cpu().numpy()
