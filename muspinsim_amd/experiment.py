"""ExperimentRunner of the drop-in surface (muspinsim/experiment.py:21-498).

`run()` carries every configuration of a MuSpinConfig through the batched
GPU engine in one pass (H build, eigensolver or Lindbladian modes, basis
change, time sum or integral, powder accumulation); `load_config` /
`run_single` and the cached properties (B, p, T, rho0, Hz, Htot,
dissipation_operators, p_operator) evaluate one configuration through the
Hamiltonian / Lindbladian classes, as the reference does per snapshot.
"""

import logging

import numpy as np
import scipy.constants as cnst

from .config import MuSpinConfig
from .constants import MU_TAU
from .engine import InstanceBatch, SpinEngine
from .hamiltonian import Hamiltonian
from .lindbladian import Lindbladian
from .orientations import rotate
from .spinop import DensityOperator, SpinOperator, spin_matrices


def _perp(z):
    """Two axes orthogonal to z (utils.py:71-93)."""
    zn = np.linalg.norm(z)
    if zn == 0:
        raise ValueError("Can not find X and Y for null Z vector")
    z = np.asarray(z, dtype=float) / zn
    if z[0] == 0 and z[1] == 0:
        return np.array([1.0, 0, 0]), np.array([0, z[2], 0])
    x = np.array([z[1], -z[0], 0.0]) / np.hypot(z[0], z[1])
    return x, np.cross(z, x)


class _Term:
    """A linear term as Celio's splitting reads it (indices and tensor)."""

    def __init__(self, indices, tensor):
        self.indices = np.array(indices, dtype=int)
        self.tensor = np.asarray(tensor, dtype=float)


def celio_plus(op):
    """Eigenvector of the largest eigenvalue of a Hermitian matrix."""
    w, v = np.linalg.eigh(np.asarray(op, dtype=complex))
    return v[:, int(np.argmax(w))]


def fold_results(cfg, out):
    """Engine output (n_slots, width) -> the configuration's results shape:
    slots x time when time is the x axis, else one value per slot (the
    time average when time is an averaged axis)."""
    out = np.asarray(out)
    shape = cfg.results.shape
    if cfg.x_axis == "t":
        if cfg.time_averaged:
            raise NotImplementedError("time can not be both the x axis and averaged")
        return out.reshape(shape)
    res = out.reshape(shape + (out.shape[1],))
    return np.mean(res, axis=-1) if cfg.time_averaged else res[..., 0]


def unfold_gradient(cfg, grad, width):
    """Adjoint of fold_results: dL/d(results) -> dL/d(engine output)."""
    grad = np.asarray(grad, dtype=float)
    if cfg.x_axis == "t":
        return grad.reshape(-1, width)
    g = grad.reshape(-1, 1)
    if cfg.time_averaged:
        return np.repeat(g / width, width, axis=1)
    out = np.zeros((g.shape[0], width))
    out[:, :1] = g
    return out


class ExperimentRunner:
    def __init__(self, config, variables=None, use_experimental_x_axis=True):
        """config: a MuSpinConfig, or a MuSpinInput evaluated with `variables`
        (experiment.py:25-70); variables: values of fitting variables (when
        given, run() leaves the results function to the caller);
        use_experimental_x_axis: for a fitting input, run on the x values of
        the fitting data rather than the file's x axis."""
        from .input import MuSpinInput
        if isinstance(config, MuSpinInput):
            config = MuSpinConfig(config.evaluate(**(variables or {})),
                                  use_experimental_x_axis=use_experimental_x_axis)
        if not isinstance(config, MuSpinConfig):
            raise TypeError("ExperimentRunner needs a MuSpinConfig or a MuSpinInput")
        self._config = config
        self._variables = {} if variables is None else dict(variables)
        self._system = config.system
        self._B = np.zeros(3)
        self._p = np.array([1.0, 0, 0])
        self._T = np.inf
        self._Hsys = None
        self._rho0 = None
        self._Hz = None
        self._dops = None
        self._T_inf_speedup = False
        self._engine = None

    # ---------------------------------------------------------- parameters
    @property
    def config(self):
        return self._config

    @property
    def system(self):
        return self._system

    @property
    def B(self):
        return self._B

    @B.setter
    def B(self, x):
        x = np.array(x, dtype=float)
        if (x != self._B).any():
            self._B = x
            self._rho0 = None
            self._Hz = None
            self._dops = None

    @property
    def p(self):
        return self._p

    @p.setter
    def p(self, x):
        x = np.array(x, dtype=float)
        if (x != self._p).any():
            self._p = x
            self._rho0 = None

    @property
    def T(self):
        return self._T

    @T.setter
    def T(self, x):
        if x != self._T:
            self._T = x
            self._rho0 = None
            self._dops = None

    # ---------------------------------------------------------- operators
    @property
    def rho0(self):
        """Muon pure along p, every other spin thermal under its own Zeeman
        term at T (experiment.py:169-236)."""
        if self._rho0 is None:
            T, B = self._T, self._B
            if np.isclose(np.linalg.norm(B), 0.0) and T < np.inf:
                logging.warning("initial density matrix at zero field and finite temperature "
                                "uses an approximation that can fail")
            rho = None
            for i in range(len(self._system)):
                I = self._system.I(i)
                if i == self._system.muon_index:
                    r = DensityOperator.from_vectors(I, self._p, 0)
                else:
                    Hz = sum(B[j] * SpinOperator.from_axes(I, a, use_sparse=False).matrix
                             for j, a in enumerate("xyz"))
                    E, V = np.linalg.eigh(Hz)
                    E = E * 1e6 * self._system.gamma(i)
                    if T > 0:
                        Z = np.exp(-cnst.h * E / (cnst.k * T))
                    else:
                        Z = np.where(E == np.min(E), 1.0, 0.0)
                    Z = Z / np.sum(Z) if np.sum(Z) > 0 else np.ones(len(E)) / len(E)
                    r = DensityOperator((V * Z) @ V.conj().T)
                rho = r if rho is None else rho.kron(r)
            self._rho0 = rho
        return self._rho0

    @property
    def Hsys(self):
        if self._Hsys is None:
            self._Hsys = self._system.hamiltonian
        return self._Hsys

    @property
    def Hz(self):
        """Zeeman Hamiltonian sum_i gamma_i B.S_i (experiment.py:238-269),
        built on the GPU."""
        if self._Hz is None:
            from . import gpu
            from .engine import SystemDesc
            s = self._system
            bare = SystemDesc(Is=list(s.Is), gammas=list(s.gammas), terms=[],
                              muon_index=s.muon_index)
            self._Hz = Hamiltonian(gpu.build_h(bare, self._B), dim=s.dimension)
        return self._Hz

    @property
    def dissipation_operators(self):
        """[(S_i.x + i S_i.y, (a/pi) Z/(1+Z)), (S_i.x - i S_i.y, (a/pi)/(1+Z))]
        for each dissipative spin, x, y perpendicular to B
        (experiment.py:271-325)."""
        if self._dops is None:
            T = self._T
            Bn = np.linalg.norm(self._B)
            g = self._system.gammas
            if T > 0:
                Zu = np.exp(-cnst.h * g * Bn * 1e6 / (cnst.k * T))
            else:
                Zu = g * 0.0
            if Bn == 0:
                x, y = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
            else:
                x, y = _perp(self._B / Bn)
            self._dops = []
            for i, a in self._config.dissipation_terms.items():
                ops = [self._system.operator({i: e}) for e in "xyz"]
                ox = ops[0] * x[0] + ops[1] * x[1] + ops[2] * x[2]
                oy = ops[0] * y[0] + ops[1] * y[1] + ops[2] * y[2]
                self._dops.append((ox + oy * 1.0j, 1 / np.pi * a * Zu[i] / (1 + Zu[i])))
                self._dops.append((ox - oy * 1.0j, 1 / np.pi * a / (1 + Zu[i])))
        return self._dops

    @property
    def Htot(self):
        H = self.Hsys + self.Hz
        H = Hamiltonian(H.matrix, dim=H.dimension)
        if len(self._config.dissipation_terms) > 0:
            H = Lindbladian.from_hamiltonian(H, self.dissipation_operators)
        return H

    @property
    def p_operator(self):
        return self._system.muon_operator(self._p)

    # ---------------------------------------------------------- snapshots
    def load_config(self, cfg_snap):
        """B' = q.B + B_intr, p' = q.p, T; returns the orientation weight
        (experiment.py:384-432)."""
        q, w = cfg_snap.orient
        self.B = rotate(q, cfg_snap.B) + np.asarray(cfg_snap.intrinsic_B, dtype=float)
        self.p = rotate(q, cfg_snap.mupol)
        self.T = cfg_snap.T
        Bn = np.linalg.norm(self.B)
        with np.errstate(divide="ignore", invalid="ignore"):
            check = (cnst.e * (cnst.hbar ** 2) * Bn) / (2 * cnst.m_p * cnst.k * self.T)
        self._T_inf_speedup = bool(check == 0) and self._system.muon_index == 0
        return w

    def run_single(self, cfg_snap):
        """One configuration through the operator classes (experiment.py:434-498)."""
        if self._config.celio_k > 0:
            return self._celio_rows([cfg_snap])[0]
        w = self.load_config(cfg_snap)
        H = self.Htot
        if cfg_snap.y == "asymmetry":
            t = np.array(cfg_snap.t, dtype=float)
            if self._T_inf_speedup and not isinstance(H, Lindbladian):
                d = int(np.prod(self._system.dimension)) // self._system.dimension[
                    self._system.muon_index]
                data = H.fast_evolve(self._system.sigma_mu(self.p), t, d)
            else:
                data = H.evolve(self.rho0, t, operators=[self.p_operator])[:, 0]
        else:
            data = H.integrate_decaying(self.rho0, MU_TAU, operators=[self.p_operator])[0] / MU_TAU
        return np.real(data) * w

    # ---------------------------------------------------------- batched run
    def instance_batch(self):
        """The whole configuration product as one InstanceBatch (config
        order) plus the system description, for SpinEngine."""
        cfg = self._config
        snaps = cfg[0:len(cfg)]
        n = len(snaps)
        q = np.array([s.orient[0] for s in snaps]).reshape(n, 4)
        w = np.array([s.orient[1] for s in snaps], dtype=float)
        B = np.array([s.B for s in snaps], dtype=float).reshape(n, 3)
        Bint = np.array([s.intrinsic_B for s in snaps], dtype=float).reshape(n, 3)
        p = np.array([s.mupol for s in snaps], dtype=float).reshape(n, 3)
        T = np.array([s.T for s in snaps], dtype=float)
        shape = cfg.results.shape
        time_is_x = cfg.x_axis == "t"
        if cfg.y_axis == "integral":
            times = np.zeros(1)
        else:
            times = np.atleast_1d(np.array(snaps[0].t, dtype=float))
        if time_is_x:
            slot = np.array([np.ravel_multi_index(tuple(s.id[0]), shape[:-1]) if shape[:-1] else 0
                             for s in snaps], dtype=np.int64)
            n_slots = int(np.prod(shape[:-1])) if shape[:-1] else 1
        else:
            slot = np.array([np.ravel_multi_index(cfg.result_index(s.id), shape) for s in snaps],
                            dtype=np.int64)
            n_slots = int(np.prod(shape))
        desc = self._system.to_desc()
        desc.dissip = dict(cfg.dissipation_terms)
        batch = InstanceBatch(q=q, B=B, Bint=Bint, p=p, T=T, w=w, slot=slot,
                              sys_id=np.zeros(n, dtype=np.int64), times=times, y=cfg.y_axis,
                              n_slots=n_slots, avg_N=cfg.avg_N)
        return [desc], batch

    # ---------------------------------------------------------- Celio
    def _celio_rows(self, snaps):
        """Weighted time series of configurations under Celio's method
        (experiment.py:238-269,451-472): random-phase states when
        celio_averages > 0 (T = inf, muon first), else the density matrix as
        a weighted sum of product states; all configurations in one launch."""
        job, combine, ntimes = self._celio_job(snaps)
        out = job.run(ntimes)
        rows, s0 = [], 0
        for wts, w in combine:
            rows.append((wts @ out[s0:s0 + len(wts)]) * w)
            s0 += len(wts)
        return np.array(rows)

    def _celio_job(self, snaps):
        """(CelioJob of the configurations, [(state weights, orientation
        weight)] per configuration, number of time points)."""
        from . import celio
        cfg = self._config
        s = self._system
        if cfg.dissipation_terms:
            raise NotImplementedError("Dissipation is not supported when using Celio's method")
        if cfg.y_axis == "integral":
            raise NotImplementedError("integrate_decaying is not implemented for Celio's method")
        times = np.atleast_1d(np.array(snaps[0].t, dtype=float))
        avg = cfg.celio_averages
        dt = times[1] - times[0] if times.size > 1 else 0.0
        job = celio.CelioJob(s.Is, dt, cfg.celio_k)
        split = celio.SystemSplit(s.Is, s.terms)
        Sm = spin_matrices(0.5)
        combine = []
        for snap in snaps:
            w = self.load_config(snap)
            field = not np.array_equal(self._B, [0, 0, 0])
            if not field and not split.has_terms:
                raise ValueError("No interaction terms to evolve")
            celio.validate([_Term([0], np.zeros(3))], times)
            contribs = split.contribs([self._B * g for g in s.gammas] if field else None)
            if avg > 0:
                if self._T != np.inf:
                    raise ValueError("The fast version of Celio's method requires T -> inf. "
                                     "Either remove the number of averages or don't use celio.")
                if s.muon_index != 0:
                    raise ValueError("Muon must be the first spin in the system in order to "
                                     "use the fast Celio method")
                sig = sum(self._p[a] * 2.0 * Sm[a] for a in range(3))
                c = job.add_config(contribs, sig, 0)
                job.add_random_states(c, celio.plus_state(sig), avg)
                combine.append((np.full(avg, 0.5 / avg), w))
            else:
                c = job.add_config(contribs, sum(self._p[a] * Sm[a] for a in range(3)),
                                   s.muon_index, dagger=True)
                combine.append((np.array(job.add_product_states(c, self._rho0_factors())), w))
        return job, combine, times.size

    def _rho0_factors(self):
        """Per-spin (eigenvectors, weights) of the initial state: the muon
        pure along p, the others thermal under their Zeeman term (the factors
        of the rho0 property)."""
        out = []
        for i in range(len(self._system)):
            I = self._system.I(i)
            S = spin_matrices(I)
            if i == self._system.muon_index:
                out.append((celio_plus(sum(self._p[a] * S[a] for a in range(3)))[:, None],
                            np.ones(1)))
                continue
            Hz = sum(self._B[a] * S[a] for a in range(3))
            E, V = np.linalg.eigh(Hz)
            E = E * 1e6 * self._system.gamma(i)
            if self._T > 0:
                Z = np.exp(-cnst.h * E / (cnst.k * self._T))
            else:
                Z = np.where(E == np.min(E), 1.0, 0.0)
            Z = Z / np.sum(Z) if np.sum(Z) > 0 else np.ones(len(E)) / len(E)
            out.append((V, Z))
        return out

    def run(self):
        """Every configuration on the GPU; results (file ranges..., x), with
        the results function applied when no fitting variables are set
        (experiment.py:358-382)."""
        cfg = self._config
        if cfg.celio_k > 0:
            snaps = cfg[0:len(cfg)]
            rows = self._celio_rows(snaps)
            cfg.results = np.zeros(cfg.results.shape)
            scalar = cfg.x_axis != "t" and not cfg.time_averaged
            for snap, row in zip(snaps, rows):
                cfg.store_time_slice(snap.id, float(row[0]) if scalar else row)
            res = cfg.results
            if not self._variables:
                res = self.apply_results_function(res, {})
            cfg.results = res
            return cfg.results
        systems, batch = self.instance_batch()
        if self._engine is None:
            self._engine = SpinEngine("cuda:0")
        out = self._engine.run(systems, batch)          # (n_slots, width)
        res = fold_results(cfg, out)
        if not self._variables:
            res = self.apply_results_function(res, {})
        cfg.results = res
        return cfg.results

    def apply_results_function(self, results, variables):
        """The configuration's results function of x (the x axis values), y
        (the results) and the fitting variables, in the results' shape
        (experiment.py:347-356).  It is a callable or a results_function
        keyword of an input file; a scalar value fills the whole shape."""
        f = self._config.results_function
        if f is None:
            return np.asarray(results)
        x = self._config.x_axis_values
        if hasattr(f, "evaluate"):
            out = f.evaluate(**variables, x=x, y=results)
        else:
            out = f(x=x, y=results, **variables)
        out = np.asarray(out)
        shape = np.shape(results)
        if out.size == int(np.prod(shape)):
            return out.reshape(shape)
        return np.array(np.broadcast_to(out.reshape(-1)[0] if out.size == 1 else out, shape))
