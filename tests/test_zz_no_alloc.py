"""msp_eigh_noalloc (workspace in the library's static device scratch, no
allocation, every launch on the caller's stream) against msp_eigh with a
caller-supplied workspace, through the C ABI.

Bar: the same kernels run on the same inputs, so eigenvalues and vectors
must agree to the eigensolver's backward-error bound n 1e-13 ||H|| (they are
expected bit-identical); both are also checked against numpy's eigvalsh.
Two calls in a row, the second smaller, so the second call sees the first
call's leftovers in the shared scratch.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _herm(rng, n, batch):
    X = rng.normal(size=(batch, n, n)) + 1j * rng.normal(size=(batch, n, n))
    # row-major copy: the sum with a transposed view may come out in another
    # memory order, and the C ABI reads row-major matrices
    return np.ascontiguousarray((X + np.conj(np.swapaxes(X, 1, 2))) / 2)


def _eigh(L, n, H, noalloc, method):
    batch = H.shape[0]
    dH = torch.as_tensor(H, dtype=torch.complex128).cuda()
    E = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    U = torch.empty((batch, n, n), dtype=torch.complex128, device="cuda")
    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    from muspinsim_amd import _native as N
    if noalloc:
        N.check(L.msp_eigh_noalloc(n, batch, dH.data_ptr(), E.data_ptr(), U.data_ptr(),
                                   info.data_ptr(), method, st))
    else:
        wsb = L.msp_eigh_workspace(n, batch, method)
        work = torch.empty((max(wsb, 8) // 8,), dtype=torch.float64, device="cuda")
        N.check(L.msp_eigh(n, batch, dH.data_ptr(), E.data_ptr(), U.data_ptr(), info.data_ptr(),
                           method, work.data_ptr(), wsb, st))
    torch.cuda.synchronize()
    return E.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("method", [2, 0])
def test_noalloc_matches_workspace_entry(gpu_lib, method):
    from muspinsim_amd import _native as N
    L = gpu_lib
    rng = np.random.default_rng(4242 + method)
    # method 2 (divide and conquer): 30 x n = 192 then 7 x n = 100;
    # method 0 (auto): n = 64 takes the Jacobi path, then n = 96 D&C
    sizes = [(192, 30), (100, 7)] if method == 2 else [(64, 12), (96, 9)]
    for n, batch in sizes:
        print(f"no-alloc eigh: method {method} n {n} batch {batch}", flush=True)
        H = _herm(rng, n, batch)
        E0, U0, i0 = _eigh(L, n, H, False, method)
        E1, U1, i1 = _eigh(L, n, H, True, method)
        assert np.all(i0 == 0) and np.all(i1 == 0)
        nrm = np.abs(np.linalg.eigvalsh(H)).max()
        tol = 1e-13 * n * nrm
        assert np.abs(E1 - E0).max() <= tol
        assert np.abs(E1 - np.linalg.eigvalsh(H)).max() <= tol
        for b in range(batch):
            assert np.abs(H[b] @ U1[b] - U1[b] * E1[b]).max() <= tol
            assert np.abs(U1[b].conj().T @ U1[b] - np.eye(n)).max() <= 1e-13 * n
            # eigenvector phases are free; compare |U0^H U1| on the identity
            ov = np.abs(U0[b].conj().T @ U1[b])
            assert np.abs(np.diag(ov) - 1).max() <= 1e-10
    # beyond the static scratch: a clear error, nothing launched
    big = 1 + (N.MSP_NOALLOC_SCRATCH_BYTES // L.msp_eigh_workspace(192, 1, 2))
    rc = L.msp_eigh_noalloc(192, int(big), 1, 1, 1, 1, 2, None)
    assert rc == N.MSP_ERR_VALUE
    assert "static scratch" in N.last_error()
