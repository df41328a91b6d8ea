"""Parity of the HIP path with the FP64 oracle (runs on an MI355X).

Tolerances: the plain FP64 path must reproduce the oracle's P(t) / integral
to max |dP| <= 1e-10 (absolute, P in [-0.5, 0.5]); stage-level checks use
the tolerances written next to each assertion.
"""

import numpy as np
import pytest
import torch

from oracle import dynamics, physics
from oracle import pipeline as opipe

pytestmark = pytest.mark.gpu

PARITY = 1e-10


def _herm(rng, n, scale=1.0):
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return scale * (X + X.conj().T) / 2


@pytest.fixture(scope="module")
def eng(gpu_lib):
    from muspinsim_amd.engine import SpinEngine
    return SpinEngine("cuda:0")


def _dev(a, dt=torch.complex128):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt).cuda()


@pytest.mark.parametrize("n,method", [(1, 1), (2, 1), (3, 1), (8, 1), (16, 1), (33, 1), (64, 1),
                                      (2, 2), (5, 2), (16, 2), (65, 2), (128, 2), (192, 2),
                                      (161, 2), (256, 2), (300, 2), (513, 2), (600, 2), (767, 2), (1000, 2),
                                      (1100, 2),
                                      (34, 3), (47, 3), (100, 3), (129, 3), (192, 3), (255, 3),
                                      (256, 3)])
def test_eigh_random(eng, n, method):
    rng = np.random.default_rng(n)
    Hs = np.array([_herm(rng, n) for _ in range(3)])
    E, U, info = eng.eigh(_dev(Hs), method=method)
    torch.cuda.synchronize()
    E, U = E.cpu().numpy(), U.cpu().numpy()
    assert np.all(info.cpu().numpy() == 0)
    for k in range(3):
        Er = np.linalg.eigvalsh(Hs[k])
        nrm = max(1.0, np.abs(Er).max())
        # backward-stable eigensolver: errors of order n eps ||H||
        assert np.abs(E[k] - Er).max() <= 1e-13 * n * nrm
        assert np.abs(Hs[k] @ U[k] - U[k] * E[k]).max() <= 1e-13 * n * nrm
        assert np.abs(U[k].conj().T @ U[k] - np.eye(n)).max() <= 1e-13 * n
        assert np.all(np.diff(E[k]) >= 0)


def test_eigh_degenerate_and_zero(eng):
    """Exactly degenerate spectra (zero field, isotropic hyperfine) and the
    zero matrix."""
    Is = [0.5, 0.5, 0.5, 0.5]
    H = physics.hamiltonian(Is, [("bilinear", 0, 1, np.eye(3) * 100)])
    for method in (1, 2):
        E, U, info = eng.eigh(_dev(np.array([H, np.zeros_like(H)])), method=method)
        E, U = E.cpu().numpy(), U.cpu().numpy()
        assert np.abs(E[0] - np.linalg.eigvalsh(H)).max() < 1e-12
        assert np.abs(H @ U[0] - U[0] * E[0]).max() < 1e-12
        assert np.abs(U[1].conj().T @ U[1] - np.eye(16)).max() < 1e-14
        assert np.all(E[1] == 0)


def test_eigh_two_stage_structured(eng):
    """Two-stage reduction (method 3) on matrices whose reflectors have zero
    tails (tau = 0, length-one reflectors): a degenerate spin Hamiltonian, the
    zero matrix, a diagonal and a tridiagonal matrix; and its size limits."""
    from muspinsim_amd import _native as N
    Is = [0.5] * 6
    H = physics.hamiltonian(Is, [("bilinear", 0, k, np.eye(3) * 50.0 * k) for k in range(1, 6)])
    rng = np.random.default_rng(5)
    D = np.diag(rng.normal(size=64)).astype(complex)
    T = np.diag(rng.normal(size=64)) + np.diag(rng.normal(size=63) + 1j * rng.normal(size=63), -1)
    T = T + np.tril(T, -1).conj().T
    mats = np.array([H, np.zeros_like(H), D, T])
    E, U, info = eng.eigh(_dev(mats), method=3)
    E, U = E.cpu().numpy(), U.cpu().numpy()
    assert np.all(info.cpu().numpy() == 0)
    for k, M in enumerate(mats):
        Er = np.linalg.eigvalsh(M)
        nrm = max(1.0, np.abs(Er).max())
        assert np.abs(E[k] - Er).max() <= 1e-12 * nrm
        assert np.abs(M @ U[k] - U[k] * E[k]).max() <= 1e-12 * nrm
        assert np.abs(U[k].conj().T @ U[k] - np.eye(64)).max() <= 1e-13
    assert np.all(E[1] == 0)
    for n in (33, 257):
        with pytest.raises(NotImplementedError, match="two-stage"):
            eng.eigh(_dev(np.zeros((1, n, n))), method=3)


def _build(eng, systems, batch, idx):
    from muspinsim_amd import _native as N
    from muspinsim_amd.engine import pack_systems
    L = N.lib()
    sysarr, termarr = pack_systems(systems)
    dsys, dterms = eng._bytes_view(sysarr), eng._bytes_view(termarr)
    n = systems[0].n
    nb = len(idx)
    Br = torch.empty((nb, 3), dtype=torch.float64, device="cuda")
    pr = torch.empty_like(Br)
    q, B, Bi, p = (_dev(getattr(batch, a)[idx], torch.float64) for a in ("q", "B", "Bint", "p"))
    sid = _dev(batch.sys_id[idx].astype(np.int32), torch.int32)
    N.check(L.msp_rotate_fields(nb, q.data_ptr(), B.data_ptr(), Bi.data_ptr(), p.data_ptr(),
                                Br.data_ptr(), pr.data_ptr(), eng.stream()))
    H = torch.empty((nb, n, n), dtype=torch.complex128, device="cuda")
    N.check(L.msp_build_h(n, nb, dsys.data_ptr(), dterms.data_ptr(), sid.data_ptr(), Br.data_ptr(),
                          None, 1, H.data_ptr(), eng.stream()))
    torch.cuda.synchronize()
    return H.cpu().numpy(), Br.cpu().numpy(), pr.cpu().numpy()


def test_build_h_matches_oracle(eng):
    from muspinsim_amd import workloads as W
    for systems, batch in (W.c1_zero_field_powder(nori=6), W.c2_alc_sweep(nfields=3, nori=2)):
        idx = np.arange(len(batch))
        H, Br, pr = _build(eng, systems, batch, idx)
        s = systems[0]
        for k in idx:
            Brh, prh, _ = opipe.load_config(batch.q[k], batch.B[k], batch.Bint[k], batch.p[k],
                                            batch.T[k], 0)
            assert np.abs(Br[k] - Brh).max() < 1e-14 * max(1, np.abs(Brh).max())
            assert np.abs(pr[k] - prh).max() < 1e-15
            Href = physics.hamiltonian(s.Is, s.terms) + physics.zeeman(s.Is, s.gammas, Brh)
            assert np.abs(H[k] - Href).max() <= 1e-12 * np.abs(Href).max()


def _rand_system(rng, Is):
    from muspinsim_amd.engine import SystemDesc
    terms = [("linear", 0, 0, rng.normal(size=3))]
    for i in range(len(Is)):
        for j in range(i, len(Is)):
            if i == j and Is[i] < 1:
                continue
            M = rng.normal(size=(3, 3)) * 10
            terms.append(("bilinear", i, j, (M + M.T) / 2 if i == j else M))
    return SystemDesc(Is=list(Is), gammas=list(rng.normal(size=len(Is)) * 100), terms=terms)


@pytest.mark.parametrize("Is", [[0.5, 0.5, 1.0], [0.5, 3.5, 0.5], [0.5, 8.5]])
def test_build_h_from_base(eng, Is):
    """H = H_sys (built once, zeeman = 0) + Zeeman per instance equals the
    direct build; d = 18 (I = 17/2) exercises the per-element fallback."""
    from muspinsim_amd import _native as N
    from muspinsim_amd.engine import pack_systems
    L = N.lib()
    rng = np.random.default_rng(len(Is) + int(2 * max(Is)))
    systems = [_rand_system(rng, Is), _rand_system(rng, Is)]
    n = systems[0].n
    sysarr, termarr = pack_systems(systems)
    dsys, dterms = eng._bytes_view(sysarr), eng._bytes_view(termarr)
    nb = 5
    sid = _dev(np.array([0, 1, 1, 0, 1], np.int32), torch.int32)
    Br = _dev(rng.normal(size=(nb, 3)), torch.float64)
    base = torch.empty((2, n, n), dtype=torch.complex128, device="cuda")
    ids = _dev(np.arange(2, dtype=np.int32), torch.int32)
    N.check(L.msp_build_h(n, 2, dsys.data_ptr(), dterms.data_ptr(), ids.data_ptr(), None, None, 0,
                          base.data_ptr(), eng.stream()))
    H1 = torch.empty((nb, n, n), dtype=torch.complex128, device="cuda")
    H2 = torch.empty_like(H1)
    N.check(L.msp_build_h(n, nb, dsys.data_ptr(), dterms.data_ptr(), sid.data_ptr(), Br.data_ptr(),
                          None, 1, H1.data_ptr(), eng.stream()))
    N.check(L.msp_build_h(n, nb, dsys.data_ptr(), None, sid.data_ptr(), Br.data_ptr(),
                          base.data_ptr(), 1, H2.data_ptr(), eng.stream()))
    torch.cuda.synchronize()
    H1, H2, Brh = H1.cpu().numpy(), H2.cpu().numpy(), Br.cpu().numpy()
    for k, s in enumerate([0, 1, 1, 0, 1]):
        S = systems[s]
        ref = physics.hamiltonian(S.Is, S.terms) + physics.zeeman(S.Is, S.gammas, Brh[k])
        assert np.abs(H1[k] - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.abs(H2[k] - ref).max() <= 1e-12 * np.abs(ref).max()
    # msp_build_h_lower: the same lower triangles (both kernels), the strict
    # upper triangle untouched
    low = np.tril(np.ones((n, n), bool))
    for base_ptr, terms_ptr, full in ((None, dterms.data_ptr(), H1),
                                      (base.data_ptr(), None, H2)):
        H3 = torch.full((nb, n, n), 7.0 + 7.0j, dtype=torch.complex128, device="cuda")
        N.check(L.msp_build_h_lower(n, nb, dsys.data_ptr(), terms_ptr, sid.data_ptr(),
                                    Br.data_ptr(), base_ptr, 1, H3.data_ptr(), eng.stream()))
        H3 = H3.cpu().numpy()
        assert np.array_equal(H3[:, low], full[:, low])
        assert np.all(H3[:, ~low] == 7.0 + 7.0j)


def test_gemm_cn(eng):
    rng = np.random.default_rng(7)
    for n in (1, 5, 32, 47, 192):
        A = rng.normal(size=(2, n, n)) + 1j * rng.normal(size=(2, n, n))
        B = rng.normal(size=(2, n, n)) + 1j * rng.normal(size=(2, n, n))
        C = eng.gemm_cn(_dev(A), _dev(B)).cpu().numpy()
        ref = np.conj(np.swapaxes(A, 1, 2)) @ B
        assert np.abs(C - ref).max() <= 1e-13 * n


@pytest.mark.parametrize("n", [16, 100, 192])
def test_upper_block_gemm_and_observable_integral(eng, n):
    """msp_gemm_cn_upper writes the blocks on and above the 64 x 64 block
    diagonal bit-identically to msp_gemm_cn and leaves the others untouched;
    msp_integral_obs on those blocks equals form_c + integral on the full
    Hermitian O = U^H (M U) (numpy reference of the integral too)."""
    from muspinsim_amd import _native as N
    rng = np.random.default_rng(n)
    nb = 3
    _, U = np.linalg.eigh(np.stack([_herm(rng, n) for _ in range(nb)]))
    M = np.stack([_herm(rng, n) for _ in range(nb)])
    Y = M @ U
    E = np.sort(rng.normal(size=(nb, n)) * 30, axis=1)
    w = np.array([0.5, 1.0, 2.0])
    dU, dY = _dev(U), _dev(Y)
    full = eng.gemm_cn(dU, dY)
    up = torch.full_like(full, 7.0)
    st = eng.stream()
    N.check(eng.lib.msp_gemm_cn_upper(n, nb, dU.data_ptr(), dY.data_ptr(), up.data_ptr(), st))
    f, u = full.cpu().numpy(), up.cpu().numpy()
    blk = np.arange(n) // 64
    mask = blk[:, None] <= blk[None, :]
    assert np.array_equal(u[:, mask], f[:, mask])
    assert np.all(u[:, ~mask] == 7.0)
    dE, dw = _dev(E, torch.float64), _dev(w, torch.float64)   # kept alive across the calls
    for alpha, beta in ((0.25, 0.0), (2.0 / n, 1.0 / n)):
        C = eng.form_c(None, full, True, alpha, beta)
        ref = eng.integral(dE, C, True, 2.197, 1 / 2.197, dw).cpu().numpy()
        out = torch.empty((nb,), dtype=torch.float64, device="cuda")
        N.check(eng.lib.msp_integral_obs(n, nb, dE.data_ptr(), up.data_ptr(), alpha, beta, 2.197,
                                         1 / 2.197, dw.data_ptr(), out.data_ptr(), st))
        got = out.cpu().numpy()
        for b in range(nb):
            c = alpha * np.abs(f[b]) ** 2 + beta * np.diag(np.diag(f[b]).real)
            npref = w[b] / 2.197 * np.real(np.sum(c / (1 / 2.197 + 2j * np.pi *
                                                        (E[b][:, None] - E[b][None, :]))))
            scale = np.abs(c).sum()
            assert abs(got[b] - ref[b]) <= 1e-13 * scale
            assert abs(got[b] - npref) <= 1e-12 * scale


@pytest.mark.parametrize("dims,mu", [((2, 3, 2), 0), ((3, 2, 2, 4), 1), ((4, 2, 2, 6), 2),
                                     ((2, 96), 0), ((12, 2, 8), 1)])
def test_muon_mixing_upper_block_gemm(eng, dims, mu):
    """msp_gemm_muon_upper equals msp_apply_muon_op + msp_gemm_cn_upper on the
    upper blocks, and U^H (f v.S_mu (x) 1) U built with numpy, for a spin-1/2
    muon at any position and both the Pauli (f = 2) and spin (f = 1) forms."""
    from muspinsim_amd import _native as N
    n = int(np.prod(dims))
    rng = np.random.default_rng(n + mu)
    nb = 4
    _, U = np.linalg.eigh(np.stack([_herm(rng, n) for _ in range(nb)]))
    vec = rng.normal(size=(nb, 3))
    sysarr = np.zeros(2, dtype=N.SYSTEM_DTYPE)
    for k in range(2):
        sysarr[k]["n_spins"] = len(dims)
        sysarr[k]["n"] = n
        sysarr[k]["muon_index"] = mu
        sysarr[k]["dims"][: len(dims)] = dims
    dsys = eng._bytes_view(sysarr)
    sid = _dev(np.array([1, 0, 1, 0], dtype=np.int32), torch.int32)
    dU, dv = _dev(U), _dev(vec, torch.float64)
    st = eng.stream()
    sx = np.array([[0, 0.5], [0.5, 0]]); sy = np.array([[0, -0.5j], [0.5j, 0]])
    sz = np.diag([0.5, -0.5])
    blk = np.arange(n) // 64
    mask = blk[:, None] <= blk[None, :]
    for pauli in (1, 0):
        Y = torch.empty_like(dU)
        ref = torch.empty_like(dU)
        got = torch.full_like(dU, 7.0)
        N.check(eng.lib.msp_apply_muon_op(n, nb, dsys.data_ptr(), sid.data_ptr(), dv.data_ptr(),
                                          pauli, dU.data_ptr(), Y.data_ptr(), st))
        N.check(eng.lib.msp_gemm_cn_upper(n, nb, dU.data_ptr(), Y.data_ptr(), ref.data_ptr(), st))
        N.check(eng.lib.msp_gemm_muon_upper(n, nb, dsys.data_ptr(), sid.data_ptr(), dv.data_ptr(),
                                            pauli, dU.data_ptr(), got.data_ptr(), st))
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.all(g[:, ~mask] == 7.0)
        f = 2.0 if pauli else 1.0
        for b in range(nb):
            m = f * (vec[b, 0] * sx + vec[b, 1] * sy + vec[b, 2] * sz)
            op = np.kron(np.kron(np.eye(int(np.prod(dims[:mu]))), m),
                         np.eye(int(np.prod(dims[mu + 1:]))))
            npref = U[b].conj().T @ op @ U[b]
            assert np.abs(g[b][mask] - r[b][mask]).max() <= 1e-13 * f * np.abs(vec[b]).sum()
            assert np.abs(g[b][mask] - npref[mask]).max() <= 1e-12 * f * np.abs(vec[b]).sum()


def test_muon_mixing_gemm_rejects_other_multiplicity(eng):
    """A system whose muon position holds a dimension other than 2 gets NaN
    blocks from msp_gemm_muon_upper (the mix is defined for spin 1/2 only)."""
    from muspinsim_amd import _native as N
    dims, n, nb = (3, 4), 12, 2
    sysarr = np.zeros(1, dtype=N.SYSTEM_DTYPE)
    sysarr[0]["n_spins"] = 2
    sysarr[0]["n"] = n
    sysarr[0]["muon_index"] = 0
    sysarr[0]["dims"][:2] = dims
    dsys = eng._bytes_view(sysarr)
    U = _dev(np.stack([np.eye(n)] * nb).astype(complex))
    dv = _dev(np.ones((nb, 3)), torch.float64)
    O = torch.zeros_like(U)
    N.check(eng.lib.msp_gemm_muon_upper(n, nb, dsys.data_ptr(), None, dv.data_ptr(), 1,
                                        U.data_ptr(), O.data_ptr(), eng.stream()))
    assert torch.isnan(O.cpu()).all()


def test_xcd_tile_order_batch11(eng):
    """Batch 11 at n = 192: the first 8 instances take the XCD-grouped tile
    order and the last 3 the plain tail order, in one launch each of
    msp_gemm_cn_upper, msp_gemm_muon_upper, msp_timesum and the divide and
    conquer eigh (MFMA back-transformation); all checked against numpy."""
    from muspinsim_amd import _native as N
    n, nb = 192, 11
    rng = np.random.default_rng(1911)
    Hs = np.stack([_herm(rng, n, 20.0) for _ in range(nb)])
    E, U, info = eng.eigh(_dev(Hs), method=2)
    E, Ud = E.cpu().numpy(), U
    Un = U.cpu().numpy()
    assert np.all(info.cpu().numpy() == 0)
    for b in range(nb):
        Er = np.linalg.eigvalsh(Hs[b])
        nrm = np.abs(Er).max()
        assert np.abs(E[b] - Er).max() <= 1e-13 * n * nrm
        assert np.abs(Hs[b] @ Un[b] - Un[b] * E[b]).max() <= 1e-13 * n * nrm
        assert np.abs(Un[b].conj().T @ Un[b] - np.eye(n)).max() <= 1e-13 * n
    st = eng.stream()
    blk = np.arange(n) // 64
    mask = blk[:, None] <= blk[None, :]
    # upper-block GEMM of U^H (M U)
    M = np.stack([_herm(rng, n) for _ in range(nb)])
    Y = _dev(M @ Un)
    up = torch.full_like(Ud, 7.0)
    N.check(eng.lib.msp_gemm_cn_upper(n, nb, Ud.data_ptr(), Y.data_ptr(), up.data_ptr(), st))
    u = up.cpu().numpy()
    for b in range(nb):
        ref = Un[b].conj().T @ M[b] @ Un[b]
        assert np.abs(u[b][mask] - ref[mask]).max() <= 1e-12 * np.abs(M[b]).max() * n ** 0.5
        assert np.all(u[b][~mask] == 7.0)
    # muon-mixing upper GEMM, muon first of (2, 96)
    sysarr = np.zeros(1, dtype=N.SYSTEM_DTYPE)
    sysarr[0]["n_spins"] = 2
    sysarr[0]["n"] = n
    sysarr[0]["muon_index"] = 0
    sysarr[0]["dims"][:2] = (2, 96)
    dsys = eng._bytes_view(sysarr)
    vec = rng.normal(size=(nb, 3))
    dv = _dev(vec, torch.float64)
    got = torch.full_like(Ud, 7.0)
    N.check(eng.lib.msp_gemm_muon_upper(n, nb, dsys.data_ptr(), None, dv.data_ptr(), 1,
                                        Ud.data_ptr(), got.data_ptr(), st))
    g = got.cpu().numpy()
    sx = np.array([[0, 1], [1, 0]]); sy = np.array([[0, -1j], [1j, 0]]); sz = np.diag([1.0, -1])
    for b in range(nb):
        op = np.kron(vec[b, 0] * sx + vec[b, 1] * sy + vec[b, 2] * sz, np.eye(96))
        ref = Un[b].conj().T @ op @ Un[b]
        assert np.abs(g[b][mask] - ref[mask]).max() <= 1e-12 * np.abs(vec[b]).sum()
        assert np.all(g[b][~mask] == 7.0)
    # time sum (complex C), 3 time tiles
    T = 150
    C = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
    t = np.linspace(0, 10, T)
    out = eng.timesum(_dev(E, torch.float64), _dev(C), _dev(t, torch.float64), False,
                      None, True).cpu().numpy()
    for b in range(nb):
        ph = np.exp(-2j * np.pi * E[b][None, :] * t[:, None])           # (T, n)
        ref = np.einsum("ti,ij,tj->t", ph, C[b], ph.conj())
        assert np.abs(out[b] - ref).max() <= 1e-12 * np.abs(C[b]).sum()


def test_timesum_and_integral(eng):
    rng = np.random.default_rng(3)
    for n, T in ((4, 7), (16, 130), (70, 65)):
        E = np.sort(rng.normal(size=(2, n)) * 50)
        Cc = rng.normal(size=(2, n, n)) + 1j * rng.normal(size=(2, n, n))
        t = np.linspace(0, 10, T)
        w = np.array([0.7, 1.3])
        out = eng.timesum(_dev(E, torch.float64), _dev(Cc), _dev(t, torch.float64), False,
                          _dev(w, torch.float64)).cpu().numpy()
        outc = eng.timesum(_dev(E, torch.float64), _dev(Cc), _dev(t, torch.float64), False,
                           None, True).cpu().numpy()
        for b in range(2):
            ll = -2j * np.pi * (E[b][:, None] - E[b][None, :])
            ref = np.array([np.sum(np.exp(ll * tt) * Cc[b]) for tt in t])
            assert np.abs(outc[b] - ref).max() <= 1e-12 * np.abs(Cc[b]).sum()
            assert np.abs(out[b] - w[b] * ref.real).max() <= 1e-12 * np.abs(Cc[b]).sum()
        integ = eng.integral(_dev(E, torch.float64), _dev(Cc), False, 2.0, 0.5, None,
                             True).cpu().numpy()
        for b in range(2):
            ref = np.sum(Cc[b] / (0.5 + 2j * np.pi * (E[b][:, None] - E[b][None, :]))) * 0.5
            assert abs(integ[b] - ref) <= 1e-13 * np.abs(Cc[b]).sum()


def _end_to_end(systems, batch, eng, tol=PARITY):
    res = eng.run(systems, batch)
    ref = opipe.run(systems, batch)
    err = np.abs(res - ref).max()
    assert err <= tol, f"max |dP| = {err:.3e}"
    return err


def test_c1_full_parity(eng):
    """Zero-field powder, n=16, 100 orientations x 1000 times (fast path)."""
    from muspinsim_amd import workloads as W
    systems, batch = W.c1_zero_field_powder()
    _end_to_end(systems, batch, eng)


def test_c2_subset_parity(eng):
    """ALC integral sweep, n=192, a 12 field x 5 orientation subset."""
    from muspinsim_amd import workloads as W
    systems, batch = W.c2_alc_sweep(nfields=12, nori=5)
    _end_to_end(systems, batch, eng)


def test_c3_reduced_parity(eng):
    """n=512 transverse field, 3 orientations x 300 times in 0-20 us."""
    from muspinsim_amd import workloads as W
    systems, batch = W.c3_tf_stress(nori=3, ntimes=300)
    _end_to_end(systems, batch, eng)


def test_c5_reduced_parity(eng):
    """n=64 (Jacobi bucket), 16 orientations x 500 times."""
    from muspinsim_amd import workloads as W
    systems, batch = W.c5_fit(nori=16, ntimes=500)
    _end_to_end(systems, batch, eng)


def test_c4_small_buckets_parity(eng):
    """Uneven batch: the first 60 instances (n = 4 .. 256) in one call."""
    from muspinsim_amd import workloads as W
    systems, batch = W.c4_uneven_batch(ninst=60)
    keep = [k for k in range(60) if systems[k].n <= 256]
    from muspinsim_amd.engine import InstanceBatch
    sub = InstanceBatch(q=batch.q[keep], B=batch.B[keep], Bint=batch.Bint[keep], p=batch.p[keep],
                        T=batch.T[keep], w=batch.w[keep], slot=batch.slot[keep],
                        sys_id=batch.sys_id[keep], times=batch.times, y=batch.y,
                        n_slots=1, avg_N=len(keep))
    _end_to_end(systems, sub, eng)


def _c4_subset(picks):
    from muspinsim_amd import workloads as W
    from muspinsim_amd.engine import InstanceBatch
    systems, batch = W.c4_uneven_batch()
    keep = np.array(picks)
    sub = InstanceBatch(q=batch.q[keep], B=batch.B[keep], Bint=batch.Bint[keep], p=batch.p[keep],
                        T=batch.T[keep], w=batch.w[keep], slot=batch.slot[keep],
                        sys_id=batch.sys_id[keep], times=batch.times, y=batch.y,
                        n_slots=1, avg_N=len(keep))
    return systems, sub


def test_c4_large_buckets_parity(eng):
    """Uneven batch, end to end through H build -> eigh -> basis change ->
    time sum: two instances of the full 4096-instance batch from each of
    the n = 384, 512, 768 and 1024 buckets (panel pair in the workspace,
    8-wave MFMA back-transformation above n = 512), all in one call."""
    from muspinsim_amd import workloads as W
    systems, _ = W.c4_uneven_batch()
    ns = np.array([s.n for s in systems])
    picks = []
    for n in (384, 512, 768, 1024):
        picks += list(np.nonzero(ns == n)[0][:2])
    systems, sub = _c4_subset(picks)
    _end_to_end(systems, sub, eng)


def test_general_route_finite_temperature(eng):
    """Finite T in a field: rho0 = kron of thermal states, second basis change
    (experiment.py:169-236, hamiltonian.py:40-107); muon not first."""
    from muspinsim_amd.engine import InstanceBatch, SystemDesc
    ge, gmu = physics.ELEC_GAMMA, physics.MU_GAMMA
    s = SystemDesc([0.5, 0.5, 1.0], [ge, gmu, 3.0777],
                   [("bilinear", 1, 0, np.eye(3) * 50 + np.diag([1, 2, -3.0])),
                    ("bilinear", 2, 0, np.eye(3) * 5),
                    ("bilinear", 2, 2, np.diag([0.3, 0.2, -0.5]))], muon_index=1)
    n = 4
    q = np.array([[1.0, 0, 0, 0], [0.9, 0.1, 0.3, 0.2], [0.5, 0.5, 0.5, 0.5], [0.2, 0.9, 0.1, 0.3]])
    q /= np.linalg.norm(q, axis=1)[:, None]
    for y, times in (("asymmetry", np.linspace(0, 5, 77)), ("integral", np.zeros(1))):
        batch = InstanceBatch(q=q, B=np.tile([0.01, 0.0, 0.05], (n, 1)), Bint=np.zeros((n, 3)),
                              p=np.tile([1.0, 0, 0], (n, 1)), T=np.array([0.5, 1.0, 2.0, 5.0]),
                              w=np.ones(n), slot=np.array([0, 1, 0, 1]),
                              sys_id=np.zeros(n, dtype=np.int64), times=times, y=y,
                              n_slots=2, avg_N=2)
        _end_to_end([s], batch, eng)


@pytest.mark.parametrize("T", [np.inf, 20.0])
def test_multi_chunk_matches_single_chunk(gpu_lib, T):
    """The chunk loop (several launches per dimension group, last chunk
    partial) gives the same results as one launch per group, on a batch of
    mixed Hilbert dimensions; T = 20 K takes the general (thermal) route."""
    from muspinsim_amd import workloads as W
    from muspinsim_amd.engine import SpinEngine
    systems, batch = W.c4_uneven_batch(seed=9, ninst=40, ntimes=31, max_dim=48)
    batch.T = np.full(batch.w.size, T)
    one = SpinEngine("cuda:0").run(systems, batch)
    many = SpinEngine("cuda:0", max_chunk=3).run(systems, batch)
    assert np.abs(one - many).max() <= 1e-13
    ref = opipe.run(systems, batch)
    assert np.abs(one - ref).max() <= 1e-10
