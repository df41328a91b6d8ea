"""Batched GPU pipeline for the spin-dynamics hot path.

One call of :meth:`SpinEngine.run` carries a whole set of instances
(orientation x field x temperature snapshots) through

  (a) msp_rotate_fields + msp_build_h   B' = q.B + B_int, H = H_sys + Zeeman(B')
  (b) msp_eigh                          Jacobi in LDS (n <= 64) or Householder + D&C
  (c) msp_apply_muon_op + msp_gemm_cn   O' = U^H (p'.S_mu) U   (and rho' when needed)
  (d) msp_form_c + msp_timesum / msp_integral, then msp_accumulate
      (weights, 1/avg_N and the results array, in instance order)

which is ExperimentRunner.run / run_single (muspinsim/experiment.py:358-498)
for the non-dissipative routes:
  - y = asymmetry, T_inf speedup, muon first  -> Hamiltonian.fast_evolve
    (hamiltonian.py:166-217): C = |U^H (sigma.p' (x) 1) U|^2 / (4 d)
  - y = asymmetry otherwise                   -> Hamiltonian.evolve (hamiltonian.py:40-107)
  - y = integral                              -> integrate_decaying / MU_TAU
    (hamiltonian.py:119-164, experiment.py:492-496)
When every non-muon spin is at infinite temperature (T = inf or B' = 0) the
initial state is rho0 = (1 + 2 p'.S_mu)/n, so rho'_ij O'_ji is the real
symmetric (2/n)|O'_ij|^2 + (1/n) delta_ij O'_ii and no second basis change
is needed; otherwise rho0 = kron of per-spin states is applied through
msp_apply_kron and a second GEMM.

Device memory comes from torch (allocated once per chunk shape and reused);
every kernel runs on torch's current stream of the engine's device.
"""

from dataclasses import dataclass, field

import numpy as np
import scipy.constants as cnst
import torch

from . import _native as N
from .constants import MU_TAU


@dataclass
class SystemDesc:
    """A spin system as plain data (spins in Kronecker order)."""

    Is: list
    gammas: list
    terms: list                      # (kind, i, j, tensor), kind in scalar/linear/bilinear
    muon_index: int = 0
    dissip: dict = field(default_factory=dict)

    @property
    def dims(self):
        return [int(round(2 * I + 1)) for I in self.Is]

    @property
    def n(self):
        return int(np.prod(self.dims))


@dataclass
class InstanceBatch:
    """Per-instance arrays of one run (see oracle.pipeline for the same layout)."""

    q: np.ndarray        # (N, 4) stored (conjugated) orientation quaternions
    B: np.ndarray        # (N, 3) applied field, lab frame [T]
    Bint: np.ndarray     # (N, 3) intrinsic field, not rotated [T]
    p: np.ndarray        # (N, 3) muon polarisation, lab frame
    T: np.ndarray        # (N,) temperature [K]
    w: np.ndarray        # (N,) orientation weight
    slot: np.ndarray     # (N,) flat index into the results array
    sys_id: np.ndarray   # (N,) index into the system list
    times: np.ndarray    # (T,) microseconds (ignored for integrals)
    y: str               # "asymmetry" or "integral"
    n_slots: int
    avg_N: int

    def __len__(self):
        return len(self.w)


_KIND = {"scalar": 0, "linear": 1, "bilinear": 2}


def pack_systems(systems):
    """Systems and their terms as the C structs MspSystem / MspTerm."""
    sysarr = np.zeros(len(systems), dtype=N.SYSTEM_DTYPE)
    terms = []
    for k, s in enumerate(systems):
        if len(s.Is) > N.MSP_MAX_SPINS:
            raise NotImplementedError(f"at most {N.MSP_MAX_SPINS} spins per system")
        sysarr[k]["n_spins"] = len(s.Is)
        sysarr[k]["n"] = s.n
        sysarr[k]["muon_index"] = s.muon_index
        sysarr[k]["term_offset"] = len(terms)
        sysarr[k]["n_terms"] = len(s.terms)
        sysarr[k]["dims"][: len(s.Is)] = s.dims
        sysarr[k]["gammas"][: len(s.Is)] = s.gammas
        for kind, i, j, tensor in s.terms:
            t = np.zeros(1, dtype=N.TERM_DTYPE)[0]
            t["kind"] = _KIND[kind]
            t["i"] = i
            t["j"] = j
            flat = np.atleast_1d(np.asarray(tensor, dtype=float)).reshape(-1)
            t["t"][: flat.size] = flat
            terms.append(t)
    termarr = np.array(terms, dtype=N.TERM_DTYPE) if terms else np.zeros(1, dtype=N.TERM_DTYPE)
    return sysarr, termarr


def rotate_host(q, v):
    """q v q* for arrays of quaternions (N, 4) and vectors (N, 3)."""
    w = q[:, :1]
    u = q[:, 1:]
    c = np.cross(u, v)
    return v + 2.0 * (w * c + np.cross(u, c))


class SpinEngine:
    """Runs instance batches through the HIP pipeline on one device."""

    def __init__(self, device="cuda:0", method=N.MSP_EIGH_AUTO, chunk_bytes=6 << 30,
                 max_chunk=8192, base_bytes=1 << 30):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("SpinEngine needs a HIP device (torch 'cuda' device on ROCm)")
        self.lib = N.lib()
        self.method = method
        self.chunk_bytes = chunk_bytes
        self.max_chunk = max_chunk
        self.base_bytes = base_bytes

    # ------------------------------------------------------------------ utils
    def stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _dev(self, a, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(self.device)

    def _bytes_view(self, structarr):
        raw = np.frombuffer(np.ascontiguousarray(structarr).tobytes(), dtype=np.uint8)
        return torch.from_numpy(raw.copy()).to(self.device)

    def per_instance_bytes(self, n, complex_rho):
        mats = 5 if complex_rho else 4
        ws = N.load_library().msp_eigh_workspace(n, 1, self.method)
        return mats * n * n * 16 + ws + n * 8 * 4

    def chunk_size(self, n, complex_rho):
        per = self.per_instance_bytes(n, complex_rho)
        return int(max(1, min(self.max_chunk, self.chunk_bytes // per)))

    # ---------------------------------------------------------------- stages
    def eigh(self, H, method=None):
        """E (B, n) ascending and U (B, n, n) of a batch of Hermitian matrices
        (complex128, destroyed)."""
        B, n, _ = H.shape
        m = self.method if method is None else method
        E = torch.empty((B, n), dtype=torch.float64, device=self.device)
        U = torch.empty((B, n, n), dtype=torch.complex128, device=self.device)
        info = torch.empty((B,), dtype=torch.int32, device=self.device)
        wsb = self.lib.msp_eigh_workspace(n, B, m)
        work = torch.empty((max(wsb, 8) // 8,), dtype=torch.float64, device=self.device)
        N.check(self.lib.msp_eigh(n, B, H.data_ptr(), E.data_ptr(), U.data_ptr(), info.data_ptr(),
                                  m, work.data_ptr(), wsb, self.stream()))
        return E, U, info

    def gemm_cn(self, A, Bm):
        C = torch.empty_like(A)
        N.check(self.lib.msp_gemm_cn(A.shape[-1], A.shape[0], A.data_ptr(), Bm.data_ptr(),
                                     C.data_ptr(), self.stream()))
        return C

    def form_c(self, rho, obs, real_mode, alpha=0.0, beta=0.0):
        B, n, _ = obs.shape
        if real_mode:
            C = torch.empty((B, n, n), dtype=torch.float64, device=self.device)
        else:
            C = torch.empty((B, n, n), dtype=torch.complex128, device=self.device)
        N.check(self.lib.msp_form_c(n, B, None if rho is None else rho.data_ptr(), obs.data_ptr(),
                                    int(real_mode), float(alpha), float(beta), C.data_ptr(),
                                    self.stream()))
        return C

    def timesum(self, E, C, times, real_mode, w=None, out_complex=False):
        B, n = E.shape
        T = times.shape[0]
        if out_complex:
            out = torch.empty((B, T), dtype=torch.complex128, device=self.device)
        else:
            out = torch.empty((B, T), dtype=torch.float64, device=self.device)
        N.check(self.lib.msp_timesum(n, B, T, E.data_ptr(), C.data_ptr(), int(real_mode),
                                     times.data_ptr(), None if w is None else w.data_ptr(),
                                     int(out_complex), out.data_ptr(), self.stream()))
        return out

    def integral(self, E, C, real_mode, tau, out_scale=1.0, w=None, out_complex=False):
        B, n = E.shape
        if out_complex:
            out = torch.empty((B,), dtype=torch.complex128, device=self.device)
        else:
            out = torch.empty((B,), dtype=torch.float64, device=self.device)
        N.check(self.lib.msp_integral(n, B, E.data_ptr(), C.data_ptr(), int(real_mode), float(tau),
                                      float(out_scale), None if w is None else w.data_ptr(),
                                      int(out_complex), out.data_ptr(), self.stream()))
        return out

    def phase_evolve(self, E, rho, times):
        """(T, n, n): exp(-2 pi i (E_i - E_j) t) rho_ij for one instance."""
        n = E.shape[-1]
        T = times.shape[0]
        out = torch.empty((T, n, n), dtype=torch.complex128, device=self.device)
        N.check(self.lib.msp_phase_evolve(n, T, E.data_ptr(), rho.data_ptr(), times.data_ptr(),
                                          out.data_ptr(), self.stream()))
        return out

    def apply_kron(self, dims, F, U):
        B, n, _ = U.shape
        Y = torch.empty_like(U)
        tmp = torch.empty_like(U)
        hd = (ctypes_int_array(dims))
        N.check(self.lib.msp_apply_kron(n, B, len(dims), hd, F.data_ptr(), U.data_ptr(),
                                        Y.data_ptr(), tmp.data_ptr(), self.stream()))
        return Y

    # ------------------------------------------------------------------- run
    def run(self, systems, batch, return_rows=False):
        """Results array (n_slots, width) for a whole instance batch; width is
        len(times) for asymmetry and 1 for integrals."""
        plan = self.prepare(systems, batch)
        res = self.execute(plan)
        out = res.cpu().numpy()
        if return_rows:
            return out, plan.rows.cpu().numpy()
        return out

    def prepare(self, systems, batch):
        """Route instances (experiment.py:413-431), upload every per-instance
        array and allocate the chunk buffers; nothing is computed."""
        for s in systems:
            if s.dissip:
                raise NotImplementedError(
                    "dissipative (Lindblad) instances are not handled by SpinEngine.run")
        nb = len(batch)
        width = 1 if batch.y == "integral" else len(batch.times)
        plan = _Plan()
        plan.batch = batch
        plan.systems = systems
        plan.width = width
        plan.results = torch.zeros((batch.n_slots, width), dtype=torch.float64, device=self.device)
        plan.rows = torch.zeros((max(nb, 1), width), dtype=torch.float64, device=self.device)
        plan.groups = []
        if nb == 0:
            return plan
        Br_h = rotate_host(batch.q, batch.B) + batch.Bint
        Bn = np.linalg.norm(Br_h, axis=1)
        Tarr = np.asarray(batch.T, dtype=float)
        with np.errstate(divide="ignore", invalid="ignore"):
            check = (cnst.e * (cnst.hbar ** 2) * Bn) / (2 * cnst.m_p * cnst.k * Tarr)
        iso = check == 0                       # every non-muon spin at infinite temperature
        mu_first = np.array([systems[int(k)].muon_index == 0 for k in batch.sys_id])
        fast = iso & mu_first & (batch.y == "asymmetry")
        ns = np.array([systems[int(k)].n for k in batch.sys_id])
        sysarr, termarr = pack_systems(systems)
        plan.d_sys = self._bytes_view(sysarr)
        plan.d_terms = self._bytes_view(termarr)
        plan.times = self._dev(batch.times if batch.y == "asymmetry" else np.zeros(1))
        for n in np.unique(ns):
            n = int(n)
            for route in ("fast", "iso", "general"):
                if route == "fast":
                    sel = (ns == n) & fast
                elif route == "iso":
                    sel = (ns == n) & iso & ~fast
                else:
                    sel = (ns == n) & ~iso
                idx = np.nonzero(sel)[0]
                if idx.size == 0:
                    continue
                if route == "general":
                    dims0 = systems[int(batch.sys_id[idx[0]])].dims
                    for k in idx:
                        if systems[int(batch.sys_id[k])].dims != dims0:
                            raise NotImplementedError(
                                "finite-temperature instances of one dimension must share spin dims")
                cs = self.chunk_size(n, route == "general")
                g = _Group()
                g.n, g.route = n, route
                g.cs = min(cs, idx.size)
                g.idx = idx
                g.d_idx = self._dev(idx.astype(np.int64), torch.int64)
                g.q = self._dev(batch.q[idx])
                g.B = self._dev(batch.B[idx])
                g.Bint = self._dev(batch.Bint[idx])
                g.p = self._dev(batch.p[idx])
                g.w = self._dev(batch.w[idx])
                g.sid = self._dev(batch.sys_id[idx].astype(np.int32), torch.int32)
                g.F = None
                if route == "general":
                    g.F = self._dev(rho0_factors_host(systems, batch, idx, Br_h), torch.complex128)
                    g.dims = systems[int(batch.sys_id[idx[0]])].dims
                if route == "fast":
                    s0 = systems[int(batch.sys_id[idx[0]])]
                    g.d_other = n // s0.dims[s0.muon_index]
                # systems shared by many instances get their field-independent
                # H_sys built once per execute; H = H_sys + Zeeman per instance
                uniq, local = np.unique(batch.sys_id[idx], return_inverse=True)
                g.base = None
                if idx.size >= 2 * uniq.size and uniq.size * n * n * 16 <= self.base_bytes:
                    g.base = torch.empty((uniq.size, n, n), dtype=torch.complex128, device=self.device)
                    g.base_sys = self._bytes_view(sysarr[uniq])
                    g.base_ids = self._dev(np.arange(uniq.size, dtype=np.int32), torch.int32)
                    g.sid_local = self._dev(local.astype(np.int32), torch.int32)
                self._alloc(g)
                plan.groups.append(g)
        order = np.argsort(batch.slot, kind="stable")
        counts = np.bincount(batch.slot, minlength=batch.n_slots)
        ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        plan.d_ptr = self._dev(ptr, torch.int32)
        plan.d_order = self._dev(order.astype(np.int32), torch.int32)
        return plan

    def _alloc(self, g):
        n, cs, dev = g.n, g.cs, self.device
        g.Br = torch.empty((cs, 3), dtype=torch.float64, device=dev)
        g.pr = torch.empty((cs, 3), dtype=torch.float64, device=dev)
        g.H = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
        g.U = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
        g.O = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
        g.E = torch.empty((cs, n), dtype=torch.float64, device=dev)
        g.info = torch.empty((cs,), dtype=torch.int32, device=dev)
        g.wsb = self.lib.msp_eigh_workspace(n, cs, self.method)
        g.work = torch.empty((max(g.wsb, 8) // 8,), dtype=torch.float64, device=dev)
        real = g.route != "general"
        g.C = torch.empty((cs, n, n), dtype=torch.float64 if real else torch.complex128, device=dev)
        if not real:
            g.R = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
            g.T2 = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)

    def execute(self, plan):
        """Launch every stage for every chunk; returns the device results
        tensor (n_slots, width).  Asynchronous on the current stream."""
        L = self.lib
        st = self.stream()
        batch = plan.batch
        plan.results.zero_()
        for g in plan.groups:
            n = g.n
            if g.base is not None:
                N.check(L.msp_build_h(n, g.base.shape[0], g.base_sys.data_ptr(),
                                      plan.d_terms.data_ptr(), g.base_ids.data_ptr(), None, None, 0,
                                      g.base.data_ptr(), st))
            for c0 in range(0, g.idx.size, g.cs):
                nb = min(g.cs, g.idx.size - c0)
                sl = slice(c0, c0 + nb)

                def p(t):
                    return t[sl].data_ptr()

                N.check(L.msp_rotate_fields(nb, p(g.q), p(g.B), p(g.Bint), p(g.p),
                                            g.Br.data_ptr(), g.pr.data_ptr(), st))
                if g.base is not None:
                    N.check(L.msp_build_h(n, nb, g.base_sys.data_ptr(), plan.d_terms.data_ptr(),
                                          p(g.sid_local), g.Br.data_ptr(), g.base.data_ptr(), 1,
                                          g.H.data_ptr(), st))
                else:
                    N.check(L.msp_build_h(n, nb, plan.d_sys.data_ptr(), plan.d_terms.data_ptr(),
                                          p(g.sid), g.Br.data_ptr(), None, 1, g.H.data_ptr(), st))
                N.check(L.msp_eigh(n, nb, g.H.data_ptr(), g.E.data_ptr(), g.U.data_ptr(),
                                   g.info.data_ptr(), self.method, g.work.data_ptr(), g.wsb, st))
                Y = g.H   # destroyed by the eigensolver; reused as scratch
                pauli = 1 if g.route == "fast" else 0
                N.check(L.msp_apply_muon_op(n, nb, plan.d_sys.data_ptr(), p(g.sid),
                                            g.pr.data_ptr(), pauli, g.U.data_ptr(), Y.data_ptr(), st))
                N.check(L.msp_gemm_cn(n, nb, g.U.data_ptr(), Y.data_ptr(), g.O.data_ptr(), st))
                if g.route == "fast":
                    N.check(L.msp_form_c(n, nb, None, g.O.data_ptr(), 1, 1.0 / (4.0 * g.d_other),
                                         0.0, g.C.data_ptr(), st))
                    real = 1
                elif g.route == "iso":
                    N.check(L.msp_form_c(n, nb, None, g.O.data_ptr(), 1, 2.0 / n, 1.0 / n,
                                         g.C.data_ptr(), st))
                    real = 1
                else:
                    hd = ctypes_int_array(g.dims)
                    N.check(L.msp_apply_kron(n, nb, len(g.dims), hd, p(g.F), g.U.data_ptr(),
                                             Y.data_ptr(), g.T2.data_ptr(), st))
                    N.check(L.msp_gemm_cn(n, nb, g.U.data_ptr(), Y.data_ptr(), g.R.data_ptr(), st))
                    N.check(L.msp_form_c(n, nb, g.R.data_ptr(), g.O.data_ptr(), 0, 0.0, 0.0,
                                         g.C.data_ptr(), st))
                    real = 0
                out = torch.empty((nb, plan.width), dtype=torch.float64, device=self.device)
                if batch.y == "integral":
                    N.check(L.msp_integral(n, nb, g.E.data_ptr(), g.C.data_ptr(), real, MU_TAU,
                                           1.0 / MU_TAU, p(g.w), 0, out.data_ptr(), st))
                else:
                    N.check(L.msp_timesum(n, nb, plan.width, g.E.data_ptr(), g.C.data_ptr(), real,
                                          plan.times.data_ptr(), p(g.w), 0, out.data_ptr(), st))
                plan.rows.index_copy_(0, g.d_idx[sl], out)
        N.check(L.msp_accumulate(batch.n_slots, plan.width, plan.d_ptr.data_ptr(),
                                 plan.d_order.data_ptr(), plan.rows.data_ptr(), 1.0 / batch.avg_N,
                                 plan.results.data_ptr(), st))
        return plan.results


class _Plan:
    pass


class _Group:
    pass


def ctypes_int_array(vals):
    import ctypes
    arr = (ctypes.c_int32 * len(vals))(*[int(v) for v in vals])
    return ctypes.cast(arr, ctypes.c_void_p)


def rho0_factors_host(systems, batch, part, Br_h):
    """Per-spin initial states (experiment.py:169-236) as kron factors, packed
    per instance: muon pure along p', others thermal under their Zeeman term."""
    s = systems[int(batch.sys_id[part[0]])]
    pr = rotate_host(batch.q[part], batch.p[part])
    out = []
    for k, I in enumerate(s.Is):
        d = int(round(2 * I + 1))
        from .spinop import spin_matrices
        Sx, Sy, Sz = spin_matrices(I)
        if k == s.muon_index:
            o = (pr[:, 0, None, None] * Sx + pr[:, 1, None, None] * Sy + pr[:, 2, None, None] * Sz)
            ev, vec = np.linalg.eigh(o)
            psi = vec[np.arange(len(part)), :, np.argmax(ev, axis=1)]
            m = psi[:, :, None] * psi[:, None, :].conj()
            m = m / np.trace(m, axis1=1, axis2=2)[:, None, None]
        else:
            Bk = Br_h[part]
            Hz = Bk[:, 0, None, None] * Sx + Bk[:, 1, None, None] * Sy + Bk[:, 2, None, None] * Sz
            ev, vec = np.linalg.eigh(Hz)
            E = ev * 1e6 * s.gammas[k]
            T = np.asarray(batch.T[part], dtype=float)[:, None]
            with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
                Z = np.where(T > 0, np.exp(-cnst.h * E / (cnst.k * np.where(T > 0, T, 1.0))),
                             (E == E.min(axis=1, keepdims=True)).astype(float))
            tot = Z.sum(axis=1, keepdims=True)
            Z = np.where(tot > 0, Z / np.where(tot > 0, tot, 1.0), 1.0 / d)
            m = np.einsum("bim,bjm,bm->bij", vec, vec.conj(), Z)
            m = m / np.trace(m, axis1=1, axis2=2)[:, None, None]
        out.append(m.reshape(len(part), d * d))
    return np.concatenate(out, axis=1)
