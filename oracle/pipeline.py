"""Per-instance routing and powder accumulation (oracle).

Restates ExperimentRunner.load_config / run_single / run
(muspinsim/experiment.py:358-498) and MuSpinConfig.store_time_slice
(muspinsim/simconfig.py:347-368) for a batch of instances described by plain
arrays.

A system is an object with attributes ``Is``, ``gammas``, ``terms``,
``muon_index`` and ``dissip`` ({spin index: rate in MHz}); an instance batch
has per-instance arrays ``q`` (N, 4) stored (conjugated) quaternions, ``B``
(N, 3) lab field in T, ``Bint`` (N, 3) intrinsic field, ``p`` (N, 3) muon
polarisation, ``T`` (N,) temperature in K, ``w`` (N,) orientation weight,
``slot`` (N,) flat output index and ``sys_id`` (N,) system index, plus
``times`` (T,), ``y`` ("asymmetry" or "integral"), ``n_slots`` and ``avg_N``.
"""

import numpy as np
import scipy.constants as cnst

from . import dynamics, orient, physics


def load_config(q, B, Bint, p, T, muon_index):
    """B' = q.B + B_intr, p' = q.p, fast-path flag (experiment.py:384-432)."""
    Br = orient.rotate(q, B) + np.asarray(Bint, dtype=float)
    pr = orient.rotate(q, p)
    Bn = np.linalg.norm(Br)
    with np.errstate(divide="ignore", invalid="ignore"):
        check = (cnst.e * (cnst.hbar ** 2) * Bn) / (2 * cnst.m_p * cnst.k * T)
    fast = bool(check == 0) and muon_index == 0
    return Br, pr, fast


def run_single(system, y, times, Br, pr, T, w, fast, Hsys=None):
    """One instance: Htot -> fast_evolve / evolve / Lindblad / integral,
    Re(data) * w (experiment.py:434-498).  Hsys, when given, is the
    field-independent part, built once per system as ExperimentRunner does
    (experiment.py:119)."""
    Is = system.Is
    if Hsys is None:
        Hsys = physics.hamiltonian(Is, system.terms)
    H = Hsys + physics.zeeman(Is, system.gammas, Br)
    dissip = getattr(system, "dissip", {}) or {}
    pop = physics.muon_operator(Is, system.muon_index, pr)
    if dissip:
        dops = physics.dissipation_operators(Is, system.gammas, Br, T, dissip)
        L = physics.lindbladian(H, dops)
        r0 = physics.rho0(Is, system.gammas, system.muon_index, Br, pr, T)
        if y == "asymmetry":
            data = dynamics.lindblad_evolve(L, r0, times, [pop])[:, 0]
        else:
            data = dynamics.lindblad_integrate_decaying(L, r0, physics.MU_TAU, [pop])[0] / physics.MU_TAU
        return np.real(data) * w
    if y == "asymmetry":
        if fast:
            dims = [int(round(2 * I + 1)) for I in Is]
            d = int(np.prod(dims)) // dims[system.muon_index]
            data = dynamics.fast_evolve(H, physics.sigma_mu(pr), times, d)
        else:
            r0 = physics.rho0(Is, system.gammas, system.muon_index, Br, pr, T)
            data = dynamics.evolve(H, r0, times, [pop])[:, 0]
    else:
        r0 = physics.rho0(Is, system.gammas, system.muon_index, Br, pr, T)
        data = dynamics.integrate_decaying(H, r0, physics.MU_TAU, [pop])[0] / physics.MU_TAU
    return np.real(data) * w


def instance_rows(systems, batch, indices=None):
    """Weighted per-instance results Re(P) * w, shape (len(indices), T or 1)."""
    if indices is None:
        indices = range(len(batch.w))
    rows = []
    hsys = {}
    for k in indices:
        sid = int(batch.sys_id[k])
        s = systems[sid]
        if sid not in hsys:
            hsys[sid] = physics.hamiltonian(s.Is, s.terms)
        Br, pr, fast = load_config(batch.q[k], batch.B[k], batch.Bint[k], batch.p[k],
                                   batch.T[k], s.muon_index)
        r = run_single(s, batch.y, batch.times, Br, pr, batch.T[k], batch.w[k], fast, hsys[sid])
        rows.append(np.atleast_1d(r))
    return np.array(rows)


def accumulate(rows, slots, n_slots, avg_N, width):
    """results[slot] += row / avg_N in instance order (simconfig.py:347-368)."""
    out = np.zeros((n_slots, width))
    for r, s in zip(rows, slots):
        out[int(s)] += r / avg_N
    return out


def run(systems, batch):
    """Whole batch -> results (n_slots, T) or (n_slots, 1) for integrals."""
    rows = instance_rows(systems, batch)
    width = rows.shape[1]
    return accumulate(rows, batch.slot, batch.n_slots, batch.avg_N, width)
