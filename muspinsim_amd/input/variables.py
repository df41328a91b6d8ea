"""Fitting variables of an input file (muspinsim/input/variables.py:4-47)."""

import numpy as np


class FittingVariable:
    """A named fitting variable with a starting value and bounds
    [min_value, max_value]; an empty range or a start outside it raises
    ValueError with every problem listed."""

    def __init__(self, name, value=0.0, min_value=-np.inf, max_value=np.inf):
        self._name = name
        self._value = float(value)
        self._min = float(min_value)
        self._max = float(max_value)
        problems = []
        if self._max <= self._min:
            problems.append(f"Variable {name} has invalid range: (max value {self._max} cannot "
                            f"be less than or equal to min value {self._min})\n")
        if self._value > self._max:
            problems.append(f"Variable {name} has invalid starting value: (starting value "
                            f"{self._value} cannot be greater than max value {self._max})")
        if self._value < self._min:
            problems.append(f"Variable {name} has invalid starting value: (starting value "
                            f"{self._value} cannot be less than min value {self._min})")
        if problems:
            raise ValueError("".join(problems))

    @property
    def name(self):
        return self._name

    @property
    def value(self):
        return self._value

    @property
    def bounds(self):
        return (self._min, self._max)

    def __repr__(self):
        return f"FittingVariable({self._name!r}, {self._value}, {self._min}, {self._max})"
