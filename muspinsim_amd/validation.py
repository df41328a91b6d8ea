"""Argument checks of the evolution entry points, raising the same exception
types as the reference does before any arithmetic (muspinsim/validation.py:8-101):

  TypeError   times not an ndarray, rho0 not a DensityOperator
  ValueError  times not 1-D, operators not SpinOperators, tau not a real
              number > 0, no operators for an integral
"""

from numbers import Number

import numpy as np

from .spinop import DensityOperator, SpinOperator


def validate_times(times):
    if not isinstance(times, np.ndarray):
        raise TypeError("times must be an array of values in microseconds")
    if times.ndim != 1:
        raise ValueError("times must be a one-dimensional array of values in microseconds")


def validate_evolve_params(rho0, times, operators):
    if not isinstance(rho0, DensityOperator):
        raise TypeError("rho0 must be a DensityOperator")
    validate_times(times)
    if not all(isinstance(o, SpinOperator) for o in operators):
        raise ValueError("operators must be a SpinOperator or a list of SpinOperators")


def validate_integrate_decaying_params(rho0, tau, operators):
    if not isinstance(rho0, DensityOperator):
        raise TypeError("rho0 must be a DensityOperator")
    if not (isinstance(tau, Number) and np.isreal(tau) and tau > 0):
        raise ValueError("tau must be a real number > 0")
    if not operators:
        raise ValueError("integrate_decaying needs at least one SpinOperator in operators")
    if not all(isinstance(o, SpinOperator) for o in operators):
        raise ValueError("operators must be a SpinOperator or a list of SpinOperators")
