"""FittingRunner of the drop-in surface (muspinsim/fitting.py:15-179).

The reference evaluates its input file with the current variable values and
runs an ExperimentRunner per objective evaluation.  Here either an input
file (FittingRunner(MuSpinInput), as in the reference) or a callable builds
the MuSpinConfig from the experimental x values and the variables, and
every evaluation runs the whole configuration in one batched GPU pass.

Objective and methods follow fitting.py:67-179:
  'nelder-mead'   scipy.optimize.minimize(method='nelder-mead')
  'lbfgs'         scipy.optimize.minimize(method='L-BFGS-B'), finite differences
  'least-squares' scipy.optimize.least_squares, 'lm' unconstrained, 'trf' with bounds
with the mean absolute deviation mean|y - y_target| as the scalar objective
(fitting.py:168-179) or the residual vector y - y_target for least squares.
objective='chi2' selects sum((y - y_target) / sigma)^2 instead.
"""

import os
from numbers import Number

import numpy as np
from scipy.optimize import least_squares, minimize

from .config import MuSpinConfig
from .experiment import ExperimentRunner

_METHODS = ("nelder-mead", "lbfgs", "least-squares")


class FittingRunner:
    def __init__(self, make_config, variables=None, data=None, method="nelder-mead",
                 rtol=1e-3, objective="mean_abs", sigma=1.0, single_simulation=False,
                 jac=None):
        """make_config(x, **values) -> MuSpinConfig, where x holds the data's x
        values (the experimental x axis, simconfig.py:183-231); variables:
        {name: value} or {name: (value, lower, upper)}; data: (N, 2) array of
        (x, y_target).

        Drop-in form: FittingRunner(MuSpinInput) takes the variables, data,
        method, tolerance and single-simulation flag from the input file's
        fitting keywords (fitting.py:15-64).

        jac: None (finite differences inside scipy, as the reference) or
        'adjoint': with method 'lbfgs' and an input file whose fitting
        variables enter only Hamiltonian couplings and the results function,
        each L-BFGS-B step costs one forward and one adjoint pass instead of
        (variables + 1) simulations."""
        from .input import MuSpinInput
        self._input = None
        if isinstance(make_config, MuSpinInput):
            inp = make_config
            if len(inp.variables) == 0:
                raise ValueError("MuSpinInput passed to FittingRunner has no variables to fit")
            info = inp.fitting_info
            self._input = inp
            variables = {n: (v.value,) + tuple(v.bounds) for n, v in inp.variables.items()}
            data, method, rtol = info["data"], info["method"], info["rtol"]
            single_simulation = info["single_simulation"]

            def make_config(x, **values):
                return MuSpinConfig(inp.evaluate(**values))
        if variables is None or len(variables) == 0:
            raise ValueError("FittingRunner needs at least one variable to fit")
        if method not in _METHODS:
            raise ValueError(f"fitting method must be one of {_METHODS}")
        if objective not in ("mean_abs", "chi2"):
            raise ValueError("objective must be 'mean_abs' or 'chi2'")
        if jac not in (None, "adjoint"):
            raise ValueError("jac must be None or 'adjoint'")
        if jac == "adjoint" and (self._input is None or method != "lbfgs"):
            raise ValueError("jac='adjoint' needs a MuSpinInput and method 'lbfgs'")
        self._jac = jac
        data = np.asarray(data, dtype=float)
        if data.ndim != 2 or data.shape[1] < 2:
            raise ValueError("fitting data must be an (N, 2) array of x, y")
        self._make_config = make_config
        self._xdata = data[:, 0]
        self._ytarg = data[:, 1]
        self._xnames = tuple(sorted(variables))          # order matters (fitting.py:45)
        vals, bounds = [], []
        for nm in self._xnames:
            v = variables[nm]
            if isinstance(v, Number):
                vals.append(float(v))
                bounds.append((-np.inf, np.inf))
            else:
                vals.append(float(v[0]))
                lo = -np.inf if len(v) < 2 or v[1] is None else float(v[1])
                hi = np.inf if len(v) < 3 or v[2] is None else float(v[2])
                bounds.append((lo, hi))
        self._x = np.array(vals)
        self._xbounds = np.array(bounds)
        self._constrained = bool(np.any(np.isfinite(self._xbounds)))
        self._method = method
        self._rtol = float(rtol)
        self._objective = objective
        self._sigma = sigma
        self._single = single_simulation
        self._cached = None
        self._runner = None
        self._sol = None
        self.n_evaluations = 0

    @property
    def variable_names(self):
        return self._xnames

    @property
    def solution(self):
        return self._sol

    def simulate(self, x):
        """Results (with the results function applied) for variable values x."""
        vardict = dict(zip(self._xnames, np.atleast_1d(x)))
        if self._single and self._cached is not None:
            y = self._cached
        else:
            cfg = self._make_config(self._xdata, **vardict)
            self._runner = ExperimentRunner(cfg, variables=vardict)
            y = self._runner.run()
            self.n_evaluations += 1
            if self._single:
                self._cached = y
        y = self._runner.apply_results_function(y, vardict)
        return np.asarray(y).reshape(-1)

    def residuals(self, x):
        return self.simulate(x) - self._ytarg

    def objective(self, x):
        r = self.residuals(x)
        if self._objective == "chi2":
            return float(np.sum((r / self._sigma) ** 2))
        return float(np.average(np.abs(r)))

    def run(self):
        """Optimise and return the scipy OptimizeResult (fitting.py:67-124)."""
        if self._method == "least-squares":
            self._sol = least_squares(self.residuals, self._x,
                                      method="trf" if self._constrained else "lm",
                                      gtol=self._rtol,
                                      bounds=(self._xbounds[:, 0], self._xbounds[:, 1]))
        elif self._jac == "adjoint":
            adj = InputAdjoint(self._input, self._xnames, self._x)
            bounds = None if not self._constrained else [tuple(b) for b in self._xbounds]

            def fg(x):
                self.n_evaluations += 1
                return adj.value_and_grad(x, self._ytarg, self._objective, self._sigma)[:2]
            self._sol = minimize(fg, self._x, jac=True, method="L-BFGS-B", tol=self._rtol,
                                 bounds=bounds)
        else:
            method = {"nelder-mead": "nelder-mead", "lbfgs": "L-BFGS-B"}[self._method]
            bounds = None if not self._constrained else [tuple(b) for b in self._xbounds]
            self._sol = minimize(self.objective, self._x, method=method, tol=self._rtol,
                                 bounds=bounds)
        self._x = np.array(self._sol.x)
        return self._sol

    def _final_runner(self, use_experimental_x_axis):
        """A run at the solution, results function applied."""
        vardict = dict(zip(self._xnames, self._x))
        if self._input is not None:
            runner = ExperimentRunner(self._input, variables=vardict,
                                      use_experimental_x_axis=use_experimental_x_axis)
        else:
            runner = ExperimentRunner(self._make_config(self._xdata, **vardict),
                                      variables=vardict)
        y = runner.run()
        runner.config.results = runner.apply_results_function(y, vardict)
        return runner

    def write_data(self, name=None, path="."):
        """Save the fitted model as .dat files: on the input file's own x axis
        when it has one, else on the data's x values (fitting.py:181-203)."""
        if self._sol is None:
            raise RuntimeError("run() the fit before writing its data")
        self._final_runner(use_experimental_x_axis=False).config.save_output(name=name, path=path)

    def write_report(self, fname=None, path="./"):
        """Human-readable summary of the optimisation (fitting.py:205-249)."""
        if self._sol is None:
            raise RuntimeError("run() the fit before writing its report")
        sol = self._sol
        variables = dict(zip(self._xnames, sol.x))
        if self._input is not None:
            name = MuSpinConfig(self._input.evaluate(**variables)).name
        else:
            name = self._make_config(self._xdata, **variables).name
        fname = fname if fname is not None else name + "_fit_report.txt"
        err = sol.fun
        if isinstance(err, np.ndarray):
            err = float(np.average(np.abs(err)))
        lines = [f"Fitting process for {name} completed",
                 f"Success achieved: {sol.success}"]
        if not sol.success:
            lines.append(f"   Message: {sol.message}")
        lines.append(f"Final absolute error <|f-f_targ|>: {err}")
        n_sim = sol.nfev
        if self._single:
            n_sim = 1
            lines.append(f"Number of 'results_function' evaluations: {sol.nfev}")
        lines.append(f"Number of simulations: {n_sim}")
        if self._method != "least-squares":
            lines.append(f"Number of iterations: {sol.nit}")
        lines += ["", "=" * 20, "", "Values found for fitting variables:", ""]
        lines += [f"\t{k} = {v}" for k, v in variables.items()]
        with open(os.path.join(path, fname), "w", encoding="utf-8") as fh:
            fh.write("\n".join(lines) + "\n")


# keywords whose expressions may hold fitting variables for the adjoint
_ADJOINT_KEYWORDS = ("zeeman", "dipolar", "hyperfine", "quadrupolar", "results_function")


class InputAdjoint:
    """Objective and adjoint gradient for an input file whose fitting
    variables enter only the Hamiltonian couplings (zeeman, dipolar,
    hyperfine, quadrupolar) and the results function, on the fast route
    (T = inf or B = 0, muon first, asymmetry, no dissipation).

    dH/dtheta_k: the coupling tensors are central differences of the
    evaluated input (step 1e-6 max(1, |theta_k|)); the Hamiltonian is linear
    in its tensors, so these terms give dH exactly up to the O(h^2) error of
    the tensor difference.  The results function y = f(x, P, theta) acts
    elementwise on the results array; its partial derivatives are central
    differences on the host (exact 1 and 0 for the default f = y)."""

    def __init__(self, inp, names, theta0, engine=None):
        from .adjoint import PolarisationModel
        from .input.input import _uses_any
        names = tuple(names)
        used = set(names)
        for name, kw in inp._keywords.items():
            for k in (kw.values() if isinstance(kw, dict) else [kw]):
                if name not in _ADJOINT_KEYWORDS and _uses_any(k, used):
                    raise ValueError("the adjoint gradient needs the fitting variables to enter "
                                     f"only couplings or the results function; '{name}' uses "
                                     "one")
        if "dissipation" in inp._keywords:
            raise ValueError("the adjoint gradient does not cover dissipation")
        self._inp = inp
        self._names = names
        self._runner = ExperimentRunner(inp, variables=dict(zip(names, theta0)))
        self._cfg = self._runner.config
        _, batch = self._runner.instance_batch()
        self.model = PolarisationModel(self._desc, self._dterms, batch, engine=engine)
        rf = self._cfg.results_function
        vals = getattr(rf, "_values", None)
        self._identity = rf is None or (vals is not None and len(vals) == 1 and
                                        len(vals[0]) == 1 and
                                        getattr(vals[0][0], "source", None) == "y")

    def _desc(self, theta):
        values = dict(zip(self._names, np.asarray(theta, dtype=float)))
        return MuSpinConfig(self._inp.evaluate(**values)).system.to_desc()

    def _dterms(self, theta):
        theta = np.asarray(theta, dtype=float)
        out = []
        for k in range(theta.size):
            h = 1e-6 * max(1.0, abs(theta[k]))
            tp, tm = theta.copy(), theta.copy()
            tp[k] += h
            tm[k] -= h
            terms = []
            for (kind, i, j, a), (_, _, _, b) in zip(self._desc(tp).terms, self._desc(tm).terms):
                d = (np.asarray(a, dtype=float) - np.asarray(b, dtype=float)) / (2 * h)
                if np.any(d != 0):
                    terms.append((kind, i, j, d))
            out.append(terms or [("scalar", 0, 0, np.zeros(1))])
        return out

    def _rf(self, P, theta):
        return self._runner.apply_results_function(P, dict(zip(self._names, theta)))

    def _rf_partials(self, P, theta):
        """(dy/dP elementwise, [dy/dtheta_k elementwise])."""
        if self._identity:
            return np.ones_like(P), [np.zeros_like(P) for _ in theta]
        hP = 1e-6 * np.maximum(1.0, np.abs(P))
        dP = (self._rf(P + hP, theta) - self._rf(P - hP, theta)) / (2 * hP)
        dth = []
        for k in range(theta.size):
            h = 1e-6 * max(1.0, abs(theta[k]))
            tp, tm = theta.copy(), theta.copy()
            tp[k] += h
            tm[k] -= h
            dth.append((self._rf(P, tp) - self._rf(P, tm)) / (2 * h))
        return dP, dth

    def value_and_grad(self, theta, ytarg, objective="mean_abs", sigma=1.0):
        """(objective, gradient, y) at theta; objective 'mean_abs' is
        mean|y - y_target| (its gradient is the subgradient sign(r)/N) and
        'chi2' is sum((y - y_target) / sigma)^2."""
        from .experiment import fold_results, unfold_gradient
        theta = np.asarray(theta, dtype=float)
        out = self.model.forward(theta).cpu().numpy()
        P = fold_results(self._cfg, out)
        y = self._rf(P, theta)
        r = y.reshape(-1) - np.asarray(ytarg, dtype=float).reshape(-1)
        if objective == "chi2":
            L = float(np.sum((r / sigma) ** 2))
            dLdy = 2.0 * r / np.asarray(sigma, dtype=float) ** 2
        else:
            L = float(np.average(np.abs(r)))
            dLdy = np.sign(r) / r.size
        dLdy = dLdy.reshape(P.shape)
        dP, dth = self._rf_partials(P, theta)
        gP = unfold_gradient(self._cfg, dLdy * dP, out.shape[1])
        g = self.model.vjp(theta, gP) + np.array([np.sum(dLdy * d) for d in dth])
        return L, g, y
