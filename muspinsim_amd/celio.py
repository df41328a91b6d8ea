"""Celio's Trotter method (muspinsim/celio.py:1-486) on the GPU.

The Hamiltonian is split into contributions H_1..H_m, one per run of
consecutive terms with the same index tuple among the terms whose first
index is spin i, for i in spin order (celio.py:73-143); each acts on one or
two spins.  A time step dt is k Trotter steps, each applying
c_j = exp(-2 pi i H_j dt / k) for j = 1..m in order.

  fast_evolve (celio_averages > 0, T = inf, muon first): random-phase states
  psi0 = mu (x) exp(2 pi i phi) / sqrt(n/2) (celio.py:295-476); the result is
  half the mean of <psi| sigma_mu (x) 1 |psi> over the draws.
  evolve (celio_averages = 0): rho(t_i) = U_i^H rho0 U_i with
  U_i = ((c_1 ... c_m)^k)^i (celio.py:210-293).  rho0 is written as a
  weighted sum of product states (or of its eigenvectors), each evolved as a
  state vector by the daggered factors, so no n x n matrix is formed.

Every state vector is evolved by msp_celio_evolve: one workgroup per state,
the state in LDS (n <= 8192), the whole time loop in one launch.  The
factors are exponentials of the small contribution matrices computed by
msp_eigh (c = U exp(-i theta E) U^H); the index tables of each factor are
built here.  A two-spin term whose indices are in decreasing order is
applied with its spins in that order (the reference builds the matrix in
increasing spin order but applies it in term order; for equal dimensions and
a symmetric tensor both agree).

The random phases come from this module's generator: seed(s) makes runs
repeatable (the reference draws them from np.random).
"""

import itertools
from dataclasses import dataclass
from typing import List

import numpy as np
import torch

from . import _native as N
from .spinop import spin_matrices

_RNG = np.random.default_rng()


def seed(s=None):
    """Reseed the generator of the random-phase states."""
    global _RNG
    _RNG = np.random.default_rng(s)


@dataclass
class CelioHContrib:
    """One contribution: `matrix` on the spins spin_order[:len(spin_order) -
    n_other], `other_dimension` the product of the other spins' dimensions
    (celio.py:29-51)."""

    matrix: np.ndarray
    other_dimension: int
    spin_order: List[int]
    spin_dimensions: List[int]


def _term_parts(term):
    idx = tuple(int(i) for i in np.atleast_1d(term.indices))
    return idx, np.asarray(term.tensor, dtype=float)


def _small_matrix(Is, idx, T):
    """Operator of one term on its own spins (in the order of its indices)."""
    if len(idx) == 1:
        S = spin_matrices(Is[idx[0]])
        return sum(T[a] * S[a] for a in range(3))
    Si = spin_matrices(Is[idx[0]])
    if idx[0] == idx[1]:
        return sum(T[a, b] * (Si[a] @ Si[b]) for a in range(3) for b in range(3))
    Sj = spin_matrices(Is[idx[1]])
    return sum(T[a, b] * np.kron(Si[a], Sj[b]) for a in range(3) for b in range(3))


def h_contribs(Is, terms):
    """CelioHContrib list of a term list (celio.py:73-143)."""
    dims = [int(round(2 * I + 1)) for I in Is]
    parts = [_term_parts(t) for t in terms]
    parts = [(idx, T) for idx, T in parts if len(idx) > 0]
    out = []
    for i in range(len(Is)):
        mine = [p for p in parts if p[0][0] == i]
        for idx, run in itertools.groupby(mine, key=lambda p: p[0]):
            run = list(run)
            spins = [idx[0]] if len(idx) == 2 and idx[0] == idx[1] else list(idx)
            others = [q for q in range(len(Is)) if q not in spins]
            H = sum(_small_matrix(Is, idx, T) for _, T in run)
            order = spins + others
            out.append(CelioHContrib(H, int(np.prod([dims[q] for q in others])), order,
                                     [dims[q] for q in order]))
    return out


class SystemSplit:
    """The contributions of a system's terms, computed once, with the Zeeman
    terms of a field merged in per configuration: the Zeeman term of spin i
    comes after spin i's own terms, so it joins their last run when that run
    is a linear term on spin i and otherwise follows it (the grouping of
    celio.py:88-100 applied to the term list system + Zeeman)."""

    def __init__(self, Is, terms):
        self.Is = list(Is)
        self.dims = [int(round(2 * I + 1)) for I in Is]
        parts = [_term_parts(t) for t in terms]
        self._n_terms = len([p for p in parts if len(p[0]) > 0])
        self._per_spin = []
        for i in range(len(Is)):
            mine = [p for p in parts if len(p[0]) > 0 and p[0][0] == i]
            runs = []
            for idx, run in itertools.groupby(mine, key=lambda p: p[0]):
                runs.append((idx, sum(_small_matrix(self.Is, idx, T) for _, T in run)))
            self._per_spin.append(runs)
        self._S = [spin_matrices(I) for I in self.Is]

    @property
    def has_terms(self):
        return self._n_terms > 0

    def contribs(self, zeeman=None):
        """CelioHContrib list; zeeman: per-spin field vectors gamma_i B (MHz)
        or None (no field)."""
        out = []
        for i, runs in enumerate(self._per_spin):
            runs = list(runs)
            if zeeman is not None:
                Hz = sum(zeeman[i][a] * self._S[i][a] for a in range(3))
                if runs and runs[-1][0] == (i,):
                    runs[-1] = ((i,), runs[-1][1] + Hz)
                else:
                    runs.append(((i,), Hz))
            for idx, H in runs:
                spins = [idx[0]] if len(idx) == 2 and idx[0] == idx[1] else list(idx)
                others = [q for q in range(len(self.Is)) if q not in spins]
                order = spins + others
                out.append(CelioHContrib(H, int(np.prod([self.dims[q] for q in others])), order,
                                         [self.dims[q] for q in order]))
        return out


def validate(terms, times):
    """validation.py:48-71."""
    times = np.asarray(times, dtype=float)
    if len([t for t in terms if len(np.atleast_1d(t.indices)) > 0]) == 0:
        raise ValueError("No interaction terms to evolve")
    if times[0] != 0:
        raise ValueError("Cannot use Celio's method with a non-zero start time")
    d = np.diff(times)
    if d.size and not np.isclose(d, d[0]).all():
        raise ValueError("Cannot use Celio's method with uneven spacing between times")


def _strides(dims):
    return [int(np.prod(dims[q + 1:])) for q in range(len(dims))]


def _fiber_tables(dims, spins):
    """(fiber base offsets over the other spins, element offsets over `spins`
    in their order, first slowest)."""
    st = _strides(dims)
    others = [q for q in range(len(dims)) if q not in spins]
    base = np.zeros(1, dtype=np.int64)
    for q in others:
        base = (base[:, None] + np.arange(dims[q])[None, :] * st[q]).reshape(-1)
    el = np.zeros(1, dtype=np.int64)
    for q in spins:
        el = (el[:, None] + np.arange(dims[q])[None, :] * st[q]).reshape(-1)
    return base.astype(np.int32), el.astype(np.int32)


def _n_involved(c):
    """Number of spins a contribution acts on (the leading spin_order entries
    whose dimensions multiply to its matrix size)."""
    m, acc, q = c.matrix.shape[0], 1, 0
    while acc < m:
        acc *= c.spin_dimensions[q]
        q += 1
    return max(q, 1)


def _fuse(contribs):
    """Runs of consecutive contributions whose spin sets nest, as
    [(frame spins, [(H, spins)])]: applying the run's factors one after the
    other equals applying their product on the frame (the largest set)."""
    groups = []
    for c in contribs:
        sp = tuple(c.spin_order[:_n_involved(c)])
        if groups:
            frame, parts = groups[-1]
            if set(sp) <= set(frame):
                parts.append((c.matrix, sp))
                continue
            if set(frame) < set(sp):
                groups[-1] = (sp, parts + [(c.matrix, sp)])
                continue
        groups.append((sp, [(c.matrix, sp)]))
    return groups


def _embedding(dims, frame, part):
    """Nonzeros of (factor on `part`) (x) identity on the other frame spins,
    in the frame's index order: E[ia, ib] = c[ra, rb]; also the part size."""
    fd = [dims[q] for q in frame]
    m = int(np.prod(fd))
    digits = np.array(np.unravel_index(np.arange(m), fd)).T          # (m, len(frame))
    pos = [frame.index(q) for q in part]
    rest = [a for a in range(len(frame)) if a not in pos]
    pd = [dims[q] for q in part]
    sub = np.ravel_multi_index(tuple(digits[:, pos].T), pd) if pos else np.zeros(m, int)
    key = (np.ravel_multi_index(tuple(digits[:, rest].T), [fd[a] for a in rest])
           if rest else np.zeros(m, int))
    A, B = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    keep = key[A] == key[B]
    ia, ib = A[keep], B[keep]
    return ia, ib, sub[ia], sub[ib], int(np.prod(pd))


def _exp_factors(eng, Hs, signs, theta):
    """exp(sign i theta H) for a batch of Hermitian matrices, by msp_eigh:
    U diag(exp(sign i theta E)) U^H (device tensor)."""
    E, U, info = eng.eigh(Hs.clone())
    if int(info.max().item()) != 0:
        raise np.linalg.LinAlgError("Eigenvalues did not converge")
    ph = torch.polar(torch.ones_like(E), signs[:, None] * theta * E)
    # U D U^H = A^H B with A = (U D)^H and B = U^H, on msp_gemm_cn
    A = (U * ph[:, None, :]).conj().transpose(-1, -2).contiguous()
    B = U.conj().transpose(-1, -2).contiguous()
    return eng.gemm_cn(A, B)


class CelioJob:
    """State vectors of one system evolved together: configurations (each a
    factor sequence and a measured one-spin operator) and their states."""

    def __init__(self, Is, dt, k):
        self.Is = list(Is)
        self.dims = [int(round(2 * I + 1)) for I in Is]
        self.n = int(np.prod(self.dims))
        self.dt = float(dt)
        self.k = int(k)
        self._tables = []
        self._table_at = {}
        self._ntab = 0
        self._H = []            # (H, sign) to exponentiate, in mats order
        self._fixed = []        # (offset, matrix) stored as given
        self._composites = []   # (offset, frame spins, [(part offset, part spins, m)])
        self._mat_key = {}
        self._nmat = 0
        self._desc = []
        self._cfg_steps = []
        self._meas = []
        self._state_cfg = []
        self._init_vec = []
        self._nvec = 0
        self._init_idx = []
        self._phases = []
        self.mode = None

    def _table(self, spins):
        key = tuple(spins)
        if key not in self._table_at:
            fib, el = _fiber_tables(self.dims, list(spins))
            self._table_at[key] = (self._ntab, self._ntab + fib.size)
            self._tables += [fib, el]
            self._ntab += fib.size + el.size
        return self._table_at[key]

    def _matrix(self, H, sign=None):
        """Offset of exp(sign 2 pi i H dt/k) (sign given) or of H itself."""
        H = np.ascontiguousarray(H, dtype=complex)
        key = (sign, H.shape, H.tobytes())
        if key not in self._mat_key:
            self._mat_key[key] = self._nmat
            if sign is None:
                self._fixed.append((self._nmat, H))
            else:
                self._H.append((self._nmat, H, sign))
            self._nmat += H.size
        return self._mat_key[key]

    def _composite(self, frame, parts):
        """Offset of the product of consecutive factors (parts applied in
        order) embedded on the spins `frame`."""
        key = ("composite", tuple(frame), tuple((o, tuple(sp)) for o, sp, _ in parts))
        if key not in self._mat_key:
            m = int(np.prod([self.dims[q] for q in frame]))
            self._mat_key[key] = self._nmat
            self._composites.append((self._nmat, tuple(frame), parts))
            self._nmat += m * m
        return self._mat_key[key]

    def add_config(self, contribs, meas_matrix, meas_spin, dagger=False):
        """contribs: CelioHContrib list in Trotter order; the measured operator
        acts on spin meas_spin.  dagger: evolve with c_j^H (density mode)."""
        first = len(self._desc)
        sign = +1 if dagger else -1
        for frame, parts in _fuse(contribs):
            m = int(np.prod([self.dims[q] for q in frame]))
            if m > 64:
                raise NotImplementedError("Celio factors act on at most 64 states (two spins "
                                          "with dimensions up to 8 x 8)")
            fib, el = self._table(frame)
            if len(parts) == 1:
                off = self._matrix(parts[0][0], sign)
            else:
                off = self._composite(frame, [(self._matrix(H, sign), sp, H.shape[0])
                                              for H, sp in parts])
            self._desc.append((m, off, fib, el))
        self._cfg_steps.append((first, len(self._desc) - first))
        fib, el = self._table([meas_spin])
        mm = np.asarray(meas_matrix, dtype=complex)
        d = self.dims[meas_spin]
        if mm.shape != (d, d) or not 2 <= d <= 10:
            raise NotImplementedError("Celio evolution measures operators on one spin of "
                                      "dimension 2 to 10")
        self._meas.append((mm.shape[0], self._matrix(mm), fib, el))
        return len(self._cfg_steps) - 1

    def _vec(self, v):
        v = np.asarray(v, dtype=complex).reshape(-1)
        off = self._nvec
        self._init_vec.append(v)
        self._nvec += v.size
        return off

    def add_random_states(self, cfg, mu, count):
        """`count` random-phase states of configuration cfg (muon first)."""
        if self.mode not in (None, 0):
            raise ValueError("one job holds one kind of initial state")
        self.mode = 0
        off = self._vec(mu)
        half = self.n // 2
        for _ in range(count):
            self._state_cfg.append(cfg)
            self._init_idx.append(off)
            self._phases.append(_RNG.random(half))

    def add_product_states(self, cfg, factors):
        """Product states of configuration cfg: factors[q] = (vectors (d_q, r_q)
        columns, weights (r_q,)); returns the weight of each state."""
        if self.mode not in (None, 1):
            raise ValueError("one job holds one kind of initial state")
        self.mode = 1
        offs = []
        for V, w in factors:
            V = np.asarray(V, dtype=complex)
            offs.append([(self._vec(V[:, r]), float(w[r])) for r in range(V.shape[1])])
        weights = []
        for combo in itertools.product(*offs):
            wt = float(np.prod([c[1] for c in combo]))
            if wt == 0.0:
                continue
            self._state_cfg.append(cfg)
            self._init_idx.append([c[0] for c in combo])
            weights.append(wt)
        return weights

    def add_vector_states(self, cfg, vectors, weights):
        """States given as the columns of `vectors` (n, r) with `weights`;
        returns the weights kept (nonzero)."""
        if self.mode not in (None, 2):
            raise ValueError("one job holds one kind of initial state")
        self.mode = 2
        kept = []
        V = np.asarray(vectors, dtype=complex)
        for r in range(V.shape[1]):
            if weights[r] == 0.0:
                continue
            self._state_cfg.append(cfg)
            self._init_idx.append(self._vec(V[:, r]))
            kept.append(float(weights[r]))
        return kept

    @property
    def nstates(self):
        return len(self._state_cfg)

    def prepare(self, engine=None):
        """Upload the tables, descriptors and initial-state data and compute
        the factor matrices on the device; returns the launch arguments."""
        from . import gpu
        eng = engine or gpu.engine()
        dev = eng.device
        mats = torch.zeros((max(self._nmat, 1),), dtype=torch.complex128, device=dev)
        for off, M in self._fixed:
            mats[off:off + M.size] = torch.as_tensor(M.reshape(-1), device=dev)
        by_m = {}
        for off, H, sign in self._H:
            by_m.setdefault(H.shape[0], []).append((off, H, sign))
        theta = 2.0 * np.pi * self.dt / self.k
        for m, items in by_m.items():
            Hs = torch.as_tensor(np.stack([H for _, H, _ in items]), device=dev)
            sg = torch.as_tensor([s for _, _, s in items], dtype=torch.float64, device=dev)
            C = _exp_factors(eng, Hs, sg, theta)
            idx = torch.as_tensor(np.array([off for off, _, _ in items])[:, None] +
                                  np.arange(m * m)[None, :], device=dev)
            mats[idx.reshape(-1)] = C.reshape(-1)
        self._compose(mats, dev, eng)
        tables = np.concatenate(self._tables).astype(np.int32)
        for (m, _, fib, el) in self._desc + self._meas:
            if fib + self.n // m > tables.size or el + m > tables.size:
                raise RuntimeError("Celio index tables are inconsistent")

        def i32(a):
            return torch.as_tensor(np.array(a, dtype=np.int32).reshape(-1), device=dev)
        prep = {
            "eng": eng, "mats": mats, "tables": torch.as_tensor(tables, device=dev),
            "desc": i32(self._desc), "meas": i32(self._meas), "steps": i32(self._cfg_steps),
            "state_cfg": i32(self._state_cfg),
            "init_vec": torch.as_tensor(np.concatenate(self._init_vec), device=dev),
            "init_idx": i32(self._init_idx),
            "phases": (torch.as_tensor(np.stack(self._phases), device=dev)
                       if self.mode == 0 else None),
        }
        wsb = N.load_library().msp_celio_workspace(self.n, self.nstates)
        prep["work"] = (torch.empty((max(wsb, 8) // 8,), dtype=torch.float64, device=dev)
                        if wsb else None)
        prep["wsb"] = int(wsb)
        return prep

    def _compose(self, mats, dev, eng):
        """Products of the fused factors, batched by pattern (frame and part
        spins): F = E_last ... E_1 with E_j the j-th factor on the frame
        (E F = (E^H)^H F on msp_gemm_cn)."""
        pats = {}
        for off, frame, parts in self._composites:
            key = (frame, tuple(tuple(sp) for _, sp, _ in parts))
            pats.setdefault(key, []).append((off, [o for o, _, _ in parts]))
        for (frame, pspins), items in pats.items():
            m = int(np.prod([self.dims[q] for q in frame]))
            F = None
            for j, sp in enumerate(pspins):
                ia, ib, ra, rb, mp = _embedding(self.dims, frame, sp)
                src = torch.as_tensor(np.array([po[j] for _, po in items])[:, None] +
                                      (ra * mp + rb)[None, :], device=dev)
                E = torch.zeros((len(items), m * m), dtype=torch.complex128, device=dev)
                E[:, torch.as_tensor(ia * m + ib, device=dev)] = mats[src]
                E = E.reshape(len(items), m, m)
                F = E if F is None else eng.gemm_cn(E.conj().transpose(-1, -2).contiguous(), F)
            dst = torch.as_tensor(np.array([off for off, _ in items])[:, None] +
                                  np.arange(m * m)[None, :], device=dev)
            mats[dst.reshape(-1)] = F.reshape(-1)

    def launch(self, prep, ntimes, out=None):
        """One msp_celio_evolve over every state; (nstates, ntimes) device
        tensor of Re <psi_s(t)| O |psi_s(t)>."""
        from .engine import ctypes_int_array
        eng = prep["eng"]
        if out is None:
            out = torch.empty((self.nstates, ntimes), dtype=torch.float64, device=eng.device)
        if self.nstates == 0 or ntimes == 0:
            return out
        ph = prep["phases"]
        wk = prep["work"]
        N.check(eng.lib.msp_celio_evolve(
            self.n, len(self.dims), ctypes_int_array(self.dims), self.nstates, int(ntimes), self.k,
            prep["state_cfg"].data_ptr(), prep["steps"].data_ptr(), prep["desc"].data_ptr(),
            prep["meas"].data_ptr(), torch.view_as_real(prep["mats"]).data_ptr(),
            prep["tables"].data_ptr(), int(self.mode),
            torch.view_as_real(prep["init_vec"]).data_ptr(), prep["init_idx"].data_ptr(),
            None if ph is None else ph.data_ptr(), out.data_ptr(),
            None if wk is None else wk.data_ptr(), prep["wsb"], eng.stream()))
        return out

    def flops_per_state(self, ntimes):
        """Real flops of one state's evolution (complex multiply-add = 8):
        8 n m per factor application, k applications of each per step."""
        per_cfg = []
        for first, count in self._cfg_steps:
            per_cfg.append(sum(8.0 * self.n * self._desc[first + c][0] for c in range(count)))
        per_state = np.array([per_cfg[c] for c in self._state_cfg])
        return float(per_state.mean()) * self.k * max(ntimes - 1, 0)

    def run(self, ntimes, engine=None):
        """(nstates, ntimes) Re <psi_s(t)| O |psi_s(t)> on the GPU."""
        if self.nstates == 0 or ntimes == 0:
            return np.zeros((self.nstates, ntimes))
        prep = self.prepare(engine)
        return self.launch(prep, ntimes).cpu().numpy()


@dataclass
class EvolveContrib:
    """A Trotter factor as the native evolution takes it (cpp/celio.cpp:4-10):
    the factor on its spins, the other spins' dimension, and the index map
    of the state ordered with the factor's spins first."""

    matrix: np.ndarray
    other_dim: int
    indices: np.ndarray


def plus_state(sigma):
    """+1 eigenvector of a 2 x 2 Pauli combination (celio.py:380-383)."""
    w, v = np.linalg.eigh(np.asarray(sigma, dtype=complex))
    return v[:, int(np.argmax(w))]


def _single_spin_form(O, dims):
    """(spin, d x d matrix A) with O = A on that spin (x) identity, or None."""
    O = np.asarray(O, dtype=complex)
    n = O.shape[0]
    scale = max(1.0, np.abs(O).max())
    for q, d in enumerate(dims):
        other = n // d
        T = O.reshape(dims + dims)
        perm = [q] + [a for a in range(len(dims)) if a != q]
        T = np.transpose(T, perm + [len(dims) + a for a in perm]).reshape(d, other, d, other)
        A = np.einsum("iaja->ij", T) / other
        ref = np.einsum("ij,ab->iajb", A, np.eye(other))
        if np.abs(ref - T).max() <= 1e-12 * scale:
            return q, A
    return None


def _product_form(rho, dims):
    """Per-spin factors (vectors, weights) of a product density matrix, or None."""
    rho = np.asarray(rho, dtype=complex)
    n = rho.shape[0]
    T = rho.reshape(dims + dims)
    facs, full = [], np.ones((1, 1), dtype=complex)
    ns = len(dims)
    for q in range(ns):
        keep = [q, ns + q]
        letters = "abcdefghijklmnopqrstuvwxyz"
        src = "".join(letters[a] if a < ns else letters[a - ns] for a in range(2 * ns))
        src = list(src)
        src[q], src[ns + q] = "Y", "Z"
        r = np.einsum("".join(src) + "->YZ", T)
        facs.append(r)
        full = np.kron(full, r)
    if np.abs(full - rho).max() > 1e-12 * max(1.0, np.abs(rho).max()):
        return None
    out = []
    for r in facs:
        w, V = np.linalg.eigh(r)
        w = np.where(w > 1e-14 * max(w.max(), 1e-300), w, 0.0)
        out.append((V, w))
    return out


class CelioHamiltonian:
    def __init__(self, terms, k, spinsys):
        """Terms of a spin system evolved by Celio's method with Trotter
        factor k (celio.py:54-71)."""
        self._terms = list(terms)
        self._k = int(k)
        self._spinsys = spinsys

    def __add__(self, x):
        return CelioHamiltonian(self._terms + x._terms, self._k, self._spinsys)

    @property
    def terms(self):
        return list(self._terms)

    def _calc_H_contribs(self):
        return h_contribs(self._spinsys.Is, self._terms)

    def _calc_trotter_evol_op_contribs(self, time_step, cpp=True):
        """Trotter factors c_j = exp(-2 pi i H_j dt / k) (celio.py:145-208),
        computed by msp_eigh: with cpp, EvolveContrib objects (the m x m
        matrix, the other spins' dimension, and the index map that orders the
        state with the factor's spins first); otherwise each factor as an
        n x n sparse matrix on the whole system."""
        from scipy import sparse
        from . import gpu
        eng = gpu.engine()
        dims = [int(round(2 * I + 1)) for I in self._spinsys.Is]
        n = int(np.prod(dims))
        theta = 2.0 * np.pi * time_step / self._k
        out = []
        for c in h_contribs(self._spinsys.Is, self._terms):
            Hs = torch.as_tensor(c.matrix[None], dtype=torch.complex128, device=eng.device)
            sg = torch.full((1,), -1.0, dtype=torch.float64, device=eng.device)
            C = _exp_factors(eng, Hs, sg, theta)[0].cpu().numpy()
            idx = np.transpose(np.arange(n, dtype=np.uint64).reshape(dims),
                               axes=c.spin_order).flatten()
            if cpp:
                out.append(EvolveContrib(C, int(c.other_dimension), idx))
            else:
                K = sparse.kron(sparse.csr_matrix(C), sparse.identity(c.other_dimension),
                                format="coo")
                out.append(sparse.csr_matrix((K.data, (idx[K.row].astype(np.int64),
                                                       idx[K.col].astype(np.int64))),
                                             shape=(n, n)))
        return out

    def evolve(self, rho0, times, operators):
        """Expectation values (len(times), len(operators)) of operators acting
        on one spin each, from rho(t_i) = U_i^H rho0 U_i (celio.py:210-293)."""
        times = np.array(times, dtype=float)
        if not operators:
            raise ValueError("At least one SpinOperator must be present in 'operators'")
        validate(self._terms, times)
        dims = [int(round(2 * I + 1)) for I in self._spinsys.Is]
        rho = np.asarray(rho0.matrix.toarray() if hasattr(rho0.matrix, "toarray")
                         else rho0.matrix, dtype=complex)
        dt = times[1] - times[0] if times.size > 1 else 0.0
        job = CelioJob(self._spinsys.Is, dt, self._k)
        contribs = h_contribs(self._spinsys.Is, self._terms)
        prod = _product_form(rho, dims)
        ops = []
        for o in operators:
            M = o.matrix.toarray() if hasattr(o.matrix, "toarray") else np.asarray(o.matrix)
            form = _single_spin_form(M, dims)
            if form is None:
                raise NotImplementedError("Celio evolution measures operators on one spin")
            ops.append(form)
        if prod is None:
            # a general density matrix: its eigenvectors, weighted by the
            # eigenvalues (msp_eigh)
            from . import gpu
            w, V = gpu.eigh(rho)
            w = np.where(w > 1e-14 * max(w.max(), 1e-300), w, 0.0)
        weights = []
        for q, A in ops:
            cfg = job.add_config(contribs, A, q, dagger=True)
            if prod is not None:
                weights.append(job.add_product_states(cfg, prod))
            else:
                weights.append(job.add_vector_states(cfg, V, w))
        out = job.run(times.size)
        res = np.zeros((times.size, len(ops)), dtype=complex)
        s0 = 0
        for j, w in enumerate(weights):
            res[:, j] = np.asarray(w) @ out[s0:s0 + len(w)]
            s0 += len(w)
        return res

    def fast_evolve(self, sigma_mu, times, averages, cpp=True):
        """Random-state evolution of the muon polarisation (celio.py:295-476)."""
        times = np.array(times, dtype=float)
        validate(self._terms, times)
        if averages <= 0:
            raise ValueError("averages must be a positive integer")
        if getattr(self._spinsys, "muon_index", 0) != 0:
            raise ValueError("Muon must be the first spin in the system in order to use the "
                             "fast Celio method")
        sig = sigma_mu.toarray() if hasattr(sigma_mu, "toarray") else np.asarray(sigma_mu)
        dt = times[1] - times[0] if times.size > 1 else 0.0
        job = CelioJob(self._spinsys.Is, dt, self._k)
        cfg = job.add_config(h_contribs(self._spinsys.Is, self._terms), sig, 0)
        job.add_random_states(cfg, plus_state(sig), int(averages))
        out = job.run(times.size)
        return 0.5 * out.mean(axis=0)

    def integrate_decaying(self, rho0, tau, operators):
        raise NotImplementedError("integrate_decaying is not implemented for Celio's method")
