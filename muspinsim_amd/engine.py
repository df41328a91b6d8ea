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


def _rowmajor(*ts):
    """The tensors as row-major (contiguous) device buffers: the C ABI reads
    matrices through data_ptr() as dense row-major arrays, so a strided view
    (a transpose, a slice) is copied first.  None passes through."""
    out = tuple(t if t is None or t.is_contiguous() else t.contiguous() for t in ts)
    return out if len(out) > 1 else out[0]


class SpinEngine:
    """Runs instance batches through the HIP pipeline on one device."""

    # A chunk's buffers take up to 32 GiB of the 288 GiB of HBM: at n = 192
    # that is 8192 instances, 16 full residency waves of the two-per-CU
    # tridiagonalisation, so partial last waves and launch gaps are rare.
    # precision "fp64": every stage in FP64 (the drop-in, max|dP| <= 1e-10
    # against the FP64 oracle); "mixed": the time sums of the real routes run
    # on FP32 matrix cores with FP64 phase reduction and FP64 sums over rows
    # (msp_timesum_mixed, max|dP| of order 1e-7), everything else FP64.
    PRECISIONS = ("fp64", "mixed")

    def __init__(self, device="cuda:0", method=N.MSP_EIGH_AUTO, chunk_bytes=32 << 30,
                 max_chunk=None, base_bytes=1 << 30, precision="fp64", group_streams=4,
                 tuning="default"):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("SpinEngine needs a HIP device (torch 'cuda' device on ROCm)")
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {self.PRECISIONS}")
        self.lib = N.lib()
        self.method = method
        # launch choices measured by muspinsim_amd.autotune on this kind of
        # device: "default" reads muspinsim_amd/tuning.json when it exists,
        # None uses the built-in choices, a dict or a path is used as given
        from . import autotune
        if isinstance(tuning, dict):
            self.tuning = tuning
        elif tuning == "default":
            self.tuning = autotune.load()
        elif tuning:
            self.tuning = autotune.load(tuning)
        else:
            self.tuning = None
        t = self.tuning or {}
        if max_chunk is None:
            max_chunk = int(t.get("max_chunk", 8192))
        self.mixed_min_n = int(t.get("mixed_timesum_min_n", 48))
        self.mixed_min_work = float(t.get("mixed_timesum_min_work", 1e9))
        # FP64 time sum on matrix cores (plain precision) from this n on
        self.mfma64_min_n = int(t.get("fp64_timesum_mfma_min_n", 48))
        self.precision = precision
        self.group_streams = int(group_streams)
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

    def method_for(self, n, count):
        """Eigensolver for `count` instances of dimension n: the engine's
        method when one was given, else the tuned choice (AUTO without one)."""
        if self.method != N.MSP_EIGH_AUTO:
            return self.method
        from . import autotune
        m = autotune.eigh_method(self.tuning, n, count)
        if m == N.MSP_EIGH_JACOBI and n > 64:
            m = N.MSP_EIGH_AUTO
        return m

    def per_instance_bytes(self, n, complex_rho, method=None):
        mats = 5 if complex_rho else 4
        m = self.method if method is None else method
        ws = N.load_library().msp_eigh_workspace(n, 1, m)
        return mats * n * n * 16 + ws + n * 8 * 4

    def chunk_size(self, n, complex_rho, method=None):
        per = self.per_instance_bytes(n, complex_rho, method)
        return int(max(1, min(self.max_chunk, self.chunk_bytes // per)))

    # ---------------------------------------------------------------- stages
    def _dc_lower(self, n, method=None):
        """Whether msp_eigh with this method reads only the lower triangle of
        an n x n input (the D&C path up to n = 1024)."""
        m = self.method if method is None else method
        if m == N.MSP_EIGH_AUTO:
            m = N.MSP_EIGH_JACOBI if n <= 64 else N.MSP_EIGH_DC
        return m == N.MSP_EIGH_DC and n <= 1024

    def eigh(self, H, method=None):
        """E (B, n) ascending and U (B, n, n) of a batch of Hermitian matrices
        (complex128, destroyed)."""
        H = _rowmajor(H)
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
        A, Bm = _rowmajor(A, Bm)
        C = torch.empty_like(A)
        N.check(self.lib.msp_gemm_cn(A.shape[-1], A.shape[0], A.data_ptr(), Bm.data_ptr(),
                                     C.data_ptr(), self.stream()))
        return C

    def form_c(self, rho, obs, real_mode, alpha=0.0, beta=0.0):
        rho, obs = _rowmajor(rho, obs)
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
        E, C, times, w = _rowmajor(E, C, times, w)
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
        E, C, w = _rowmajor(E, C, w)
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
        E, rho, times = _rowmajor(E, rho, times)
        n = E.shape[-1]
        T = times.shape[0]
        out = torch.empty((T, n, n), dtype=torch.complex128, device=self.device)
        N.check(self.lib.msp_phase_evolve(n, T, E.data_ptr(), rho.data_ptr(), times.data_ptr(),
                                          out.data_ptr(), self.stream()))
        return out

    def apply_kron(self, dims, F, U):
        F, U = _rowmajor(F, U)
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
        self.check_converged(plan)
        if return_rows:
            return out, plan.rows.cpu().numpy()
        return out

    @staticmethod
    def check_converged(plan):
        """LinAlgError if any eigensolver of the last execute did not converge
        (synchronises with the device)."""
        if plan.checks and int(torch.stack(plan.checks).max().item()) != 0:
            raise np.linalg.LinAlgError("Eigenvalues did not converge")

    def prepare(self, systems, batch):
        """Route instances (experiment.py:413-431), upload every per-instance
        array and allocate the chunk buffers; nothing is computed."""
        nb = len(batch)
        width = 1 if batch.y == "integral" else len(batch.times)
        plan = _Plan()
        plan.batch = batch
        plan.systems = systems
        plan.width = width
        plan.results = torch.zeros((batch.n_slots, width), dtype=torch.float64, device=self.device)
        plan.rows = torch.zeros((max(nb, 1), width), dtype=torch.float64, device=self.device)
        plan.groups = []
        plan.checks = []
        if nb == 0:
            return plan
        Br_h = rotate_host(batch.q, batch.B) + batch.Bint
        Bn = np.linalg.norm(Br_h, axis=1)
        Tarr = np.asarray(batch.T, dtype=float)
        with np.errstate(divide="ignore", invalid="ignore"):
            check = (cnst.e * (cnst.hbar ** 2) * Bn) / (2 * cnst.m_p * cnst.k * Tarr)
        iso = check == 0                       # every non-muon spin at infinite temperature
        mu_first = np.array([systems[int(k)].muon_index == 0 for k in batch.sys_id])
        lind = np.array([bool(systems[int(k)].dissip) for k in batch.sys_id])
        fast = iso & mu_first & (batch.y == "asymmetry") & ~lind
        iso = iso & ~lind
        ns = np.array([systems[int(k)].n for k in batch.sys_id])
        sysarr, termarr = pack_systems(systems)
        plan.d_sys = self._bytes_view(sysarr)
        plan.d_terms = self._bytes_view(termarr)
        plan.times = self._dev(batch.times if batch.y == "asymmetry" else np.zeros(1))
        for n in np.unique(ns):
            n = int(n)
            for route in ("fast", "iso", "general", "lindblad"):
                if route == "fast":
                    sel = (ns == n) & fast
                elif route == "iso":
                    sel = (ns == n) & iso & ~fast
                elif route == "general":
                    sel = (ns == n) & ~iso & ~lind
                else:
                    sel = (ns == n) & lind
                idx = np.nonzero(sel)[0]
                if idx.size == 0:
                    continue
                if route in ("general", "lindblad"):
                    dims0 = systems[int(batch.sys_id[idx[0]])].dims
                    for k in idx:
                        if systems[int(batch.sys_id[k])].dims != dims0:
                            raise NotImplementedError(
                                "finite-temperature instances of one dimension must share spin dims")
                if route == "lindblad":
                    if n * n > 1024:
                        raise NotImplementedError(
                            "dissipative systems are limited to n^2 <= 1024 (Liouville space)")
                    cs = int(max(1, min(self.max_chunk, self.chunk_bytes //
                                        (3 * n ** 4 * 16 + 8 * n * n * 16))))
                    method = N.MSP_EIGH_AUTO
                else:
                    method = self.method_for(n, idx.size)
                    cs = self.chunk_size(n, route == "general", method)
                g = _Group()
                g.method = method
                g.n, g.route, g.width = n, route, width
                g.cost = float(idx.size) * float(n) ** 3
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
                if route == "lindblad":
                    r, u, A, rates = lindblad_inputs_host(systems, batch, idx, Br_h)
                    g.r = self._dev(r, torch.complex128)
                    g.u = self._dev(u, torch.complex128)
                    g.A = self._dev(A, torch.complex128)
                    g.rates = self._dev(rates)
                    g.n_jump = A.shape[1]
                if route == "fast":
                    s0 = systems[int(batch.sys_id[idx[0]])]
                    g.d_other = n // s0.dims[s0.muon_index]
                g.mu_half = all(systems[int(k)].dims[systems[int(k)].muon_index] == 2
                                for k in np.unique(batch.sys_id[idx]))
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
        if g.route == "lindblad":
            Nl = n * n
            g.Br = torch.empty((cs, 3), dtype=torch.float64, device=dev)
            g.pr = torch.empty((cs, 3), dtype=torch.float64, device=dev)
            g.H = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
            g.L = torch.empty((cs, Nl, Nl), dtype=torch.complex128, device=dev)
            g.lam = torch.empty((cs, Nl), dtype=torch.complex128, device=dev)
            g.amp = torch.empty((cs, 1, Nl), dtype=torch.complex128, device=dev)
            g.info = torch.empty((cs,), dtype=torch.int32, device=dev)
            g.wsb = self.lib.msp_lindblad_workspace(Nl, cs)
            g.work = torch.empty((max(g.wsb, 8) // 8,), dtype=torch.float64, device=dev)
            return
        g.Br = torch.empty((cs, 3), dtype=torch.float64, device=dev)
        g.pr = torch.empty((cs, 3), dtype=torch.float64, device=dev)
        g.H = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
        g.U = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
        g.O = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
        g.E = torch.empty((cs, n), dtype=torch.float64, device=dev)
        g.info = torch.empty((cs,), dtype=torch.int32, device=dev)
        g.wsb = self.lib.msp_eigh_workspace(n, cs, g.method)
        g.work = torch.empty((max(g.wsb, 8) // 8,), dtype=torch.float64, device=dev)
        real = g.route != "general"
        g.C = torch.empty((cs, n, n), dtype=torch.float64 if real else torch.complex128, device=dev)
        g.tswork, g.tswsb, g.tsmode = None, 0, None
        # The real routes' time sum runs on matrix cores from a phase table:
        # FP32 products for precision="mixed", FP64 for "fp64".  Below the
        # tuned dimension (48 without tuning) the time sum takes a few
        # microseconds either way and the phase-table launch costs more, and
        # below about 1e9 multiply-adds per launch the phase-table launch is
        # not repaid either (few instances, short time axes).
        if real and g.width > 1 and float(n) * n * g.width * cs >= self.mixed_min_work:
            if self.precision == "mixed" and n >= self.mixed_min_n:
                g.tsmode = "mixed"
                g.tswsb = self.lib.msp_timesum_mixed_workspace(n, cs, g.width)
            elif self.precision == "fp64" and n >= self.mfma64_min_n:
                g.tsmode = "mfma64"
                g.tswsb = self.lib.msp_timesum_mfma64_workspace(n, cs, g.width)
            if g.tsmode is not None:
                g.tswork = torch.empty((max(g.tswsb, 8) // 8,), dtype=torch.float64, device=dev)
        if not real:
            g.R = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)
            g.T2 = torch.empty((cs, n, n), dtype=torch.complex128, device=dev)

    def execute_graph(self, plan):
        """execute() replayed from a captured graph.  The first call runs the
        plan once eagerly (kernel attributes, the side stream, buffers), then
        captures one execute into a torch CUDA graph on a stream of default
        priority; every call replays it.  Returns the results tensor, which
        each replay overwrites; plan.checks refresh with it."""
        if getattr(plan, "graph", None) is None:
            # captured on one stream: capturing the multi-stream form of a
            # several-group plan ended in a host crash inside capture_end, so
            # graphs keep the single-stream launch order
            self.execute(plan, concurrent=False)
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            cap = torch.cuda.Stream(self.device)
            cap.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.graph(g, stream=cap, capture_error_mode="thread_local"):
                res = self.execute(plan, concurrent=False)
            torch.cuda.current_stream(self.device).wait_stream(cap)
            plan.graph, plan.graph_results = g, res
        plan.graph.replay()
        return plan.graph_results

    def _stream_pool(self):
        """Streams of default priority on which dimension groups run side by
        side (created once per engine)."""
        if getattr(self, "_pool", None) is None:
            self._pool = [torch.cuda.Stream(self.device) for _ in range(self.group_streams)]
        return self._pool

    def execute(self, plan, concurrent=True):
        """Launch every stage for every chunk; returns the device results
        tensor (n_slots, width).  Asynchronous on the current stream.

        Dimension groups are independent until the final accumulation, so a
        plan with several groups (uneven batches) runs them on a pool of
        streams forked from and joined back to the current stream, the
        costliest group first: one group's one-workgroup-per-instance
        launches (a few hundred workgroups at the largest n) leave CUs free
        for the others."""
        L = self.lib
        main = torch.cuda.current_stream(self.device)
        st = main.cuda_stream
        batch = plan.batch
        plan.results.zero_()
        plan.checks = []
        groups = plan.groups
        if concurrent and self.group_streams > 1 and len(groups) > 1:
            pool = self._stream_pool()
            order = sorted(range(len(groups)), key=lambda i: -groups[i].cost)
            used = pool[: min(len(pool), len(groups))]
            for s in used:
                s.wait_stream(main)
            for k, gi in enumerate(order):
                with torch.cuda.stream(used[k % len(used)]):
                    self._execute_group(plan, groups[gi])
            for s in used:
                main.wait_stream(s)
        else:
            for g in groups:
                self._execute_group(plan, g)
        N.check(L.msp_accumulate(batch.n_slots, plan.width, plan.d_ptr.data_ptr(),
                                 plan.d_order.data_ptr(), plan.rows.data_ptr(), 1.0 / batch.avg_N,
                                 plan.results.data_ptr(), st))
        return plan.results

    def _execute_group(self, plan, g):
        """Every chunk of one dimension group, on the current stream."""
        L = self.lib
        st = self.stream()
        batch = plan.batch
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
            # the D&C eigensolver (n <= 1024) reads the lower triangle only
            build = (L.msp_build_h_lower if g.route != "lindblad" and self._dc_lower(n, g.method)
                     else L.msp_build_h)
            if g.base is not None:
                N.check(build(n, nb, g.base_sys.data_ptr(), plan.d_terms.data_ptr(),
                              p(g.sid_local), g.Br.data_ptr(), g.base.data_ptr(), 1,
                              g.H.data_ptr(), st))
            else:
                N.check(build(n, nb, plan.d_sys.data_ptr(), plan.d_terms.data_ptr(),
                              p(g.sid), g.Br.data_ptr(), None, 1, g.H.data_ptr(), st))
            if g.route == "lindblad":
                self._execute_lindblad(plan, g, sl, nb, st)
                continue
            N.check(L.msp_eigh(n, nb, g.H.data_ptr(), g.E.data_ptr(), g.U.data_ptr(),
                               g.info.data_ptr(), g.method, g.work.data_ptr(), g.wsb, st))
            plan.checks.append(g.info[:nb].amax())
            Y = g.H   # destroyed by the eigensolver; reused as scratch
            pauli = 1 if g.route == "fast" else 0
            out = torch.empty((nb, plan.width), dtype=torch.float64, device=self.device)
            if batch.y == "integral" and g.route in ("fast", "iso"):
                # O is Hermitian and the integral weights |O_ij|^2 by an
                # even kernel: the upper block triangle of O suffices
                if g.mu_half:
                    N.check(L.msp_gemm_muon_upper(n, nb, plan.d_sys.data_ptr(), p(g.sid),
                                                  g.pr.data_ptr(), pauli, g.U.data_ptr(),
                                                  g.O.data_ptr(), st))
                else:
                    N.check(L.msp_apply_muon_op(n, nb, plan.d_sys.data_ptr(), p(g.sid),
                                                g.pr.data_ptr(), pauli, g.U.data_ptr(),
                                                Y.data_ptr(), st))
                    N.check(L.msp_gemm_cn_upper(n, nb, g.U.data_ptr(), Y.data_ptr(),
                                                g.O.data_ptr(), st))
                alpha, beta = ((1.0 / (4.0 * g.d_other), 0.0) if g.route == "fast"
                               else (2.0 / n, 1.0 / n))
                N.check(L.msp_integral_obs(n, nb, g.E.data_ptr(), g.O.data_ptr(), alpha, beta,
                                           MU_TAU, 1.0 / MU_TAU, p(g.w), out.data_ptr(), st))
                plan.rows.index_copy_(0, g.d_idx[sl], out)
                continue
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
            if batch.y == "integral":
                N.check(L.msp_integral(n, nb, g.E.data_ptr(), g.C.data_ptr(), real, MU_TAU,
                                       1.0 / MU_TAU, p(g.w), 0, out.data_ptr(), st))
            elif g.tsmode is not None:
                fn = L.msp_timesum_mixed if g.tsmode == "mixed" else L.msp_timesum_mfma64
                N.check(fn(n, nb, plan.width, g.E.data_ptr(), g.C.data_ptr(),
                           plan.times.data_ptr(), p(g.w), out.data_ptr(),
                           g.tswork.data_ptr(), g.tswsb, st))
            else:
                N.check(L.msp_timesum(n, nb, plan.width, g.E.data_ptr(), g.C.data_ptr(), real,
                                      plan.times.data_ptr(), p(g.w), 0, out.data_ptr(), st))
            plan.rows.index_copy_(0, g.d_idx[sl], out)


    def _execute_lindblad(self, plan, g, sl, nb, st):
        """Liouvillian -> modes -> mode sum / integral for one chunk of a
        dissipative group (lindbladian.py:43-173 via experiment.py:489-496)."""
        L = self.lib
        n = g.n
        Nl = n * n

        def p(t):
            return t[sl].data_ptr()

        N.check(L.msp_liouvillian(n, nb, g.H.data_ptr(), g.n_jump, p(g.A), p(g.rates),
                                  g.L.data_ptr(), st))
        N.check(L.msp_lindblad_modes(Nl, nb, g.L.data_ptr(), p(g.r), 1, p(g.u), g.lam.data_ptr(),
                                     g.amp.data_ptr(), None, None, g.info.data_ptr(),
                                     g.work.data_ptr(), g.wsb, st))
        plan.checks.append(g.info[:nb].amax())
        out = torch.empty((nb, plan.width), dtype=torch.float64, device=self.device)
        if plan.batch.y == "integral":
            N.check(L.msp_mode_integral(Nl, nb, 1, g.lam.data_ptr(), g.amp.data_ptr(), MU_TAU,
                                        1.0 / MU_TAU, p(g.w), 0, out.data_ptr(), st))
        else:
            N.check(L.msp_mode_sum(Nl, nb, 1, plan.width, g.lam.data_ptr(), g.amp.data_ptr(),
                                   plan.times.data_ptr(), p(g.w), 0, out.data_ptr(), st))
        plan.rows.index_copy_(0, g.d_idx[sl], out)

    # ------------------------------------------------------------- Lindblad
    def liouvillian(self, H, A=None, rates=None):
        """L (B, n^2, n^2) of Hamiltonians H (B, n, n) and jump operators
        A (B, k, n, n) with rates (B, k)."""
        H, A, rates = _rowmajor(H, A, rates)
        B, n, _ = H.shape
        k = 0 if A is None else A.shape[1]
        Lm = torch.empty((B, n * n, n * n), dtype=torch.complex128, device=self.device)
        N.check(self.lib.msp_liouvillian(n, B, H.data_ptr(), k, None if A is None else A.data_ptr(),
                                         None if rates is None else rates.data_ptr(),
                                         Lm.data_ptr(), self.stream()))
        return Lm

    def lindblad_modes(self, Lm, r, u=None, want_c=False, want_V=False):
        """lambda (B, N), amp (B, k, N), c (B, N) or None, V (B, N, N) or None;
        Lm is destroyed.  Raises LinAlgError if a QR iteration did not converge."""
        Lm, r, u = _rowmajor(Lm, r, u)
        B, Nl, _ = Lm.shape
        k = 0 if u is None else u.shape[1]
        dev = self.device
        lam = torch.empty((B, Nl), dtype=torch.complex128, device=dev)
        amp = torch.empty((B, max(k, 1), Nl), dtype=torch.complex128, device=dev)
        c = torch.empty((B, Nl), dtype=torch.complex128, device=dev) if want_c else None
        V = torch.empty((B, Nl, Nl), dtype=torch.complex128, device=dev) if want_V else None
        info = torch.empty((B,), dtype=torch.int32, device=dev)
        wsb = self.lib.msp_lindblad_workspace(Nl, B)
        work = torch.empty((max(wsb, 8) // 8,), dtype=torch.float64, device=dev)
        N.check(self.lib.msp_lindblad_modes(
            Nl, B, Lm.data_ptr(), r.data_ptr(), k, None if u is None else u.data_ptr(),
            lam.data_ptr(), amp.data_ptr(), None if c is None else c.data_ptr(),
            None if V is None else V.data_ptr(), info.data_ptr(), work.data_ptr(), wsb,
            self.stream()))
        if int(info.max().item()) != 0:
            raise np.linalg.LinAlgError("Eigenvalues did not converge")
        return lam, amp, c, V

    def mode_sum(self, lam, amp, times):
        """(B, k, T) complex: sum_m amp_km exp(2 pi lambda_m t)."""
        B, k, Nl = amp.shape
        T = times.shape[0]
        out = torch.empty((B, k, T), dtype=torch.complex128, device=self.device)
        N.check(self.lib.msp_mode_sum(Nl, B, k, T, lam.data_ptr(), amp.data_ptr(),
                                      times.data_ptr(), None, 1, out.data_ptr(), self.stream()))
        return out

    def mode_integral(self, lam, amp, tau):
        """(B, k) complex: sum_m amp_km / (1/tau - 2 pi lambda_m)."""
        B, k, Nl = amp.shape
        out = torch.empty((B, k), dtype=torch.complex128, device=self.device)
        N.check(self.lib.msp_mode_integral(Nl, B, k, lam.data_ptr(), amp.data_ptr(), float(tau),
                                           1.0, None, 1, out.data_ptr(), self.stream()))
        return out

    def mode_states(self, lam, c, V, times):
        """(T, N) complex: V (c exp(2 pi lambda t)) for one instance."""
        Nl = lam.shape[-1]
        T = times.shape[0]
        out = torch.empty((T, Nl), dtype=torch.complex128, device=self.device)
        N.check(self.lib.msp_mode_states(Nl, T, lam.data_ptr(), c.data_ptr(), V.data_ptr(),
                                         times.data_ptr(), out.data_ptr(), self.stream()))
        return out


class _Plan:
    pass


class _Group:
    pass


def ctypes_int_array(vals):
    import ctypes
    arr = (ctypes.c_int32 * len(vals))(*[int(v) for v in vals])
    return ctypes.cast(arr, ctypes.c_void_p)


def _full_spin_ops(Is, i):
    """(3, n, n) dense S_x, S_y, S_z of spin i on the whole system."""
    from .spinop import spin_matrices
    dims = [int(round(2 * I + 1)) for I in Is]
    out = []
    for S in spin_matrices(Is[i]):
        M = np.ones((1, 1), dtype=complex)
        for k, d in enumerate(dims):
            M = np.kron(M, S if k == i else np.eye(d))
        out.append(M)
    return np.array(out)


def _perp_axes(Bv):
    """(x, y) perpendicular to each field vector; (x^, y^) for B = 0
    (experiment.py:296-301, utils.py:71-93)."""
    nb = len(Bv)
    x = np.tile([1.0, 0.0, 0.0], (nb, 1))
    y = np.tile([0.0, 1.0, 0.0], (nb, 1))
    Bn = np.linalg.norm(Bv, axis=1)
    for k in np.nonzero(Bn != 0)[0]:
        z = Bv[k] / Bn[k]
        if z[0] == 0 and z[1] == 0:
            x[k] = [1.0, 0.0, 0.0]
            y[k] = [0.0, z[2], 0.0]
        else:
            xx = np.array([z[1], -z[0], 0.0]) / np.hypot(z[0], z[1])
            x[k] = xx
            y[k] = np.cross(z, xx)
    return x, y


def lindblad_inputs_host(systems, batch, part, Br_h):
    """Per-instance inputs of the dissipative route: vec(rho0) (B, N),
    vec(O^T) of O = p'.S_mu (B, 1, N), jump operators S_i.x +- i S_i.y
    (B, 2 k, n, n) and their rates (a/pi) Z/(1+Z), (a/pi)/(1+Z),
    Z = exp(-h gamma_i |B| 1e6 / (k T)) (experiment.py:169-236,271-325)."""
    s = systems[int(batch.sys_id[part[0]])]
    n = s.n
    nb = len(part)
    F = rho0_factors_host(systems, batch, part, Br_h)
    rho = np.ones((nb, 1, 1), dtype=complex)
    off = 0
    for I in s.Is:
        d = int(round(2 * I + 1))
        f = F[:, off:off + d * d].reshape(nb, d, d)
        off += d * d
        m = rho.shape[1]
        rho = np.einsum("bij,bkl->bikjl", rho, f).reshape(nb, m * d, m * d)
    pr = rotate_host(batch.q[part], batch.p[part])
    Smu = _full_spin_ops(s.Is, s.muon_index)
    O = np.einsum("ba,aij->bij", pr, Smu)
    u = np.swapaxes(O, 1, 2).reshape(nb, 1, n * n)
    Bv = Br_h[part]
    Bn = np.linalg.norm(Bv, axis=1)
    x, y = _perp_axes(Bv)
    T = np.asarray(batch.T[part], dtype=float)
    members = [systems[int(j)] for j in batch.sys_id[part]]
    keys = sorted(s.dissip)
    if any(sorted(m.dissip) != keys for m in members):
        raise NotImplementedError("dissipative instances of one dimension must dissipate on "
                                  "the same spins")
    gam = np.array([m.gammas for m in members], dtype=float)
    A, rates = [], []
    for i in keys:
        a = np.array([m.dissip[i] for m in members], dtype=float)
        Si = _full_spin_ops(s.Is, i)
        ox = np.einsum("ba,aij->bij", x, Si)
        oy = np.einsum("ba,aij->bij", y, Si)
        with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
            Zu = np.where(T > 0, np.exp(-cnst.h * gam[:, i] * Bn * 1e6 /
                                        (cnst.k * np.where(T > 0, T, 1.0))), 0.0)
        A.append(ox + 1.0j * oy)
        rates.append(a / np.pi * Zu / (1 + Zu))
        A.append(ox - 1.0j * oy)
        rates.append(a / np.pi / (1 + Zu))
    A = np.stack(A, axis=1)
    rates = np.stack(rates, axis=1)
    return rho.reshape(nb, n * n), u, A, rates


def rho0_factors_host(systems, batch, part, Br_h):
    """Per-spin initial states (experiment.py:169-236) as kron factors, packed
    per instance: muon pure along p', others thermal under their Zeeman term."""
    s = systems[int(batch.sys_id[part[0]])]
    gam = np.array([systems[int(j)].gammas for j in batch.sys_id[part]], dtype=float)
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
            E = ev * 1e6 * gam[:, k, None]
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
