"""Mixed-precision path on the GPU: stage checks against numpy's FP64 eigh.

Bars: the FP32 stage (msp_eigh_mixed: FP32 tridiagonalisation in LDS, FP64
divide and conquer and back-transformation on its reflectors) is backward
stable in FP32, so its eigenvalues and residuals are bounded by
5e-5 ||H|| (n eps_32 ||H|| with n <= 192, eps_32 = 6e-8) and the columns are
orthonormal to 1e-5 (the reflectors carry FP32 rounding).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _herm(rng, n, batch, scale=1.0):
    X = rng.normal(size=(batch, n, n)) + 1j * rng.normal(size=(batch, n, n))
    return np.ascontiguousarray(scale * (X + np.conj(np.swapaxes(X, 1, 2))) / 2)


def _dev(a, dt=torch.complex128):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()


def eigh_mixed(lib, H):
    from muspinsim_amd import _native as N
    batch, n, _ = H.shape
    dH = _dev(H)
    E = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    U = torch.empty((batch, n, n), dtype=torch.complex128, device="cuda")
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    wsb = lib.msp_eigh_mixed_workspace(n, batch)
    work = torch.empty((max(wsb, 8) // 8,), dtype=torch.float64, device="cuda")
    N.check(lib.msp_eigh_mixed(n, batch, dH.data_ptr(), E.data_ptr(), U.data_ptr(), info.data_ptr(),
                               work.data_ptr(), wsb, torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert np.array_equal(dH.cpu().numpy(), H), "msp_eigh_mixed must preserve H"
    return E.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("n,batch", [(2, 3), (5, 4), (33, 5), (64, 3), (65, 6), (100, 9),
                                     (128, 3), (130, 4), (191, 3), (192, 11)])
def test_eigh_mixed_fp32_stage(gpu_lib, n, batch):
    rng = np.random.default_rng(1000 + n)
    H = _herm(rng, n, batch, scale=30.0)
    E, U, info = eigh_mixed(gpu_lib, H)
    assert np.all(info == 0)
    for b in range(batch):
        Er = np.linalg.eigvalsh(H[b])
        nrm = np.abs(Er).max()
        assert np.all(np.diff(E[b]) >= 0)
        assert np.abs(E[b] - Er).max() <= 5e-5 * nrm
        assert np.abs(H[b] @ U[b] - U[b] * E[b]).max() <= 5e-5 * nrm
        assert np.abs(U[b].conj().T @ U[b] - np.eye(n)).max() <= 1e-5


def test_eigh_mixed_alc_hamiltonians(gpu_lib):
    """The n = 192 ALC Hamiltonians (electron Zeeman up to 42 GHz dominates
    ||H||), full matrices built on the host by the oracle."""
    from muspinsim_amd import workloads as W
    from oracle import physics, pipeline
    systems, batch = W.c2_alc_sweep(nfields=5, nori=3)
    s = systems[0]
    Hsys = physics.hamiltonian(s.Is, s.terms)
    Hs = []
    for k in range(len(batch)):
        Br, _, _ = pipeline.load_config(batch.q[k], batch.B[k], batch.Bint[k], batch.p[k],
                                        batch.T[k], 0)
        Hs.append(Hsys + physics.zeeman(s.Is, s.gammas, Br))
    H = np.array(Hs)
    E, U, info = eigh_mixed(gpu_lib, H)
    assert np.all(info == 0)
    for b in range(len(H)):
        Er = np.linalg.eigvalsh(H[b])
        nrm = np.abs(Er).max()
        assert np.abs(E[b] - Er).max() <= 5e-5 * nrm
        assert np.abs(H[b] @ U[b] - U[b] * E[b]).max() <= 5e-5 * nrm
