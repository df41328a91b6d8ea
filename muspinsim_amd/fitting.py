"""FittingRunner of the drop-in surface (muspinsim/fitting.py:15-179).

The reference evaluates its input file with the current variable values and
runs an ExperimentRunner per objective evaluation; here a callable builds
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

from numbers import Number

import numpy as np
from scipy.optimize import least_squares, minimize

from .experiment import ExperimentRunner

_METHODS = ("nelder-mead", "lbfgs", "least-squares")


class FittingRunner:
    def __init__(self, make_config, variables, data, method="nelder-mead", rtol=1e-3,
                 objective="mean_abs", sigma=1.0, single_simulation=False):
        """make_config(x, **values) -> MuSpinConfig, where x holds the data's x
        values (the experimental x axis, simconfig.py:183-231); variables:
        {name: value} or {name: (value, lower, upper)}; data: (N, 2) array of
        (x, y_target)."""
        if len(variables) == 0:
            raise ValueError("FittingRunner needs at least one variable to fit")
        if method not in _METHODS:
            raise ValueError(f"fitting method must be one of {_METHODS}")
        if objective not in ("mean_abs", "chi2"):
            raise ValueError("objective must be 'mean_abs' or 'chi2'")
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
        f = self._runner.config.results_function
        if f is not None:
            y = np.array(f(x=self._runner.config.x_axis_values, y=y, **vardict)).reshape(y.shape)
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
        else:
            method = {"nelder-mead": "nelder-mead", "lbfgs": "L-BFGS-B"}[self._method]
            bounds = None if not self._constrained else [tuple(b) for b in self._xbounds]
            self._sol = minimize(self.objective, self._x, method=method, tol=self._rtol,
                                 bounds=bounds)
        self._x = np.array(self._sol.x)
        return self._sol
