"""Simulation configuration: the parameter product of one run
(muspinsim/simconfig.py:82-647), built from Python values.

Every parameter is a list of entries, as the lines of an input-file keyword
block: one entry is a constant, several entries a range.  A range is the
x axis, an averaged axis, or otherwise a "file range" (one output file per
value in the reference).  Entries:

  field, intrinsic_field   a number (tesla along z) or a 3-vector; a flat
                           sequence of numbers is a range of fields along z,
                           so a single vector is written [[bx, by, bz]]
  polarization             "longitudinal" (z), "transverse" (x) or a 3-vector
                           (normalised); a flat sequence is one vector
  time                     microseconds, one entry per time point
  orientation              (theta, phi) polar, (a, b, c) Euler or (a, b, c, w)
                           Euler + weight (euler_mode 'zyz' or 'zxz'); a flat
                           sequence is one orientation
  temperature              kelvin (inf allowed), one entry per value

A bare number or string is a one-entry list.  The configurations are the product file ranges x averaged axes x
x axis in that order (simconfig.py:300-316); results have shape
(file ranges..., x) and each configuration adds slice / avg_N
(store_time_slice, simconfig.py:347-368).
"""

from collections import OrderedDict, namedtuple
from itertools import product
from numbers import Number

import numpy as np

from .orientations import normalised_weights, stored_orientations

# keyword name -> snapshot field (simconfig.py:23-31)
KEYWORD_FIELDS = OrderedDict([
    ("polarization", "mupol"),
    ("field", "B"),
    ("intrinsic_field", "intrinsic_B"),
    ("time", "t"),
    ("orientation", "orient"),
    ("temperature", "T"),
])

ConfigSnapshot = namedtuple("ConfigSnapshot", ["id", "y"] + list(KEYWORD_FIELDS.values()))


class MuSpinConfigError(Exception):
    pass


def _flat_numbers(v):
    """True for a flat sequence of numbers (one vector entry)."""
    if isinstance(v, (str, Number)):
        return False
    try:
        return all(isinstance(x, Number) for x in v)
    except TypeError:
        return False


class MuSpinConfig:
    def __init__(self, system, name="muspinsim", field=0.0, intrinsic_field=0.0,
                 polarization="transverse", time=None, orientation=((0.0, 0.0, 0.0),),
                 temperature=np.inf, x_axis="time", y_axis="asymmetry",
                 average_axes=("orientation",), dissipation=None, euler_mode="zyz",
                 results_function=None):
        """system: a MuonSpinSystem with its coupling terms; dissipation:
        {spin index: rate in MHz}; results_function: optional callable
        f(x=..., y=..., **variables) applied to the results by run()."""
        from .spinsys import MuonSpinSystem
        if not isinstance(system, MuonSpinSystem):
            raise MuSpinConfigError("the system of a configuration must be a MuonSpinSystem")
        if y_axis not in ("asymmetry", "integral"):
            raise MuSpinConfigError("y_axis must be 'asymmetry' or 'integral'")
        if x_axis not in KEYWORD_FIELDS:
            raise MuSpinConfigError("Invalid x axis name")
        avg = [a for a in (average_axes or ()) if str(a).lower() != "none"]
        for a in avg:
            if a not in KEYWORD_FIELDS:
                raise MuSpinConfigError("Invalid average axis name")
        self._name = name
        self._system = system
        self._y_axis = y_axis
        self._results_function = results_function
        self._dissip_terms = {int(k): float(v) for k, v in (dissipation or {}).items()}
        for i in self._dissip_terms:
            if i < 0 or i >= len(system):
                raise MuSpinConfigError("Out of range index for dissipation")
        if time is None:
            time = np.linspace(0, 10, 101)
        raw = {"polarization": polarization, "field": field, "intrinsic_field": intrinsic_field,
               "time": time, "orientation": orientation, "temperature": temperature}
        self._constants = {}
        self._file_ranges = OrderedDict()
        self._avg_ranges = OrderedDict((KEYWORD_FIELDS[a], None) for a in avg)
        self._x_range = OrderedDict([(KEYWORD_FIELDS[x_axis], None)])
        self._time_isavg = "t" in self._avg_ranges
        self._time_N = 0
        for kw, cname in KEYWORD_FIELDS.items():
            v = self._validate(cname, raw[kw], euler_mode)
            if cname == "t":
                if y_axis == "integral":
                    v = [np.inf]
                self._time_N = len(v)
            elif cname == "orient":
                q, w = v
                w = normalised_weights(w, "orient" in self._avg_ranges)
                v = [(q[k], w[k]) for k in range(len(w))]
            if len(v) > 1:
                if cname in self._x_range:
                    self._x_range[cname] = v
                elif cname in self._avg_ranges:
                    self._avg_ranges[cname] = v
                else:
                    self._file_ranges[cname] = v
            else:
                self._constants[cname] = v[0]
        if y_axis == "integral" and "t" in self._x_range:
            raise MuSpinConfigError("Can not use time as X axis when evaluating integral of signal")
        if list(self._x_range.values())[0] is None:
            raise MuSpinConfigError("Specified x axis is not a range")
        if "t" in self._file_ranges:
            raise MuSpinConfigError("time can only be the x axis, an averaged axis or a constant")
        self._avg_ranges = OrderedDict((k, v) for k, v in self._avg_ranges.items() if v is not None)
        shape = [len(v) for v in self._file_ranges.values()] + [len(v) for v in
                                                                self._x_range.values()]
        self._results = np.zeros(shape)

        def configs(od):
            axes = [[slice(None)] if k == "t" else np.arange(len(v)) for k, v in od.items()]
            return list(product(*axes))

        self._configurations = list(product(configs(self._file_ranges),
                                            configs(self._avg_ranges), configs(self._x_range)))
        self._avg_N = len(configs(self._avg_ranges))

    # ---------------------------------------------------------- validation
    @staticmethod
    def _validate(cname, value, mode):
        if cname in ("B", "intrinsic_B"):
            # entries are numbers (tesla along z) or 3-vectors; a flat
            # sequence is a list of number entries
            if isinstance(value, Number):
                value = [value]
            out = []
            for e in value:
                e = np.atleast_1d(np.asarray(e, dtype=float))
                if e.size == 1:
                    e = np.array([0.0, 0.0, e[0]])
                elif e.size != 3:
                    raise MuSpinConfigError("Invalid magnetic field value")
                out.append(e)
            return out
        if cname == "mupol":
            if isinstance(value, str) or _flat_numbers(value):
                value = [value]
            out = []
            for e in value:
                if isinstance(e, str):
                    if e == "longitudinal":
                        e = [0.0, 0.0, 1.0]
                    elif e == "transverse":
                        e = [1.0, 0.0, 0.0]
                    else:
                        raise MuSpinConfigError("Invalid muon polarization direction")
                e = np.array(e, dtype=float)
                if e.shape != (3,) or np.linalg.norm(e) == 0:
                    raise MuSpinConfigError("Invalid muon polarization direction")
                out.append(e / np.linalg.norm(e))
            return out
        if cname in ("t", "T"):
            return [float(x) for x in np.atleast_1d(np.asarray(value, dtype=float))]
        if cname == "orient":
            if _flat_numbers(value):
                value = [value]
            return stored_orientations(value, mode)
        raise MuSpinConfigError(f"unknown parameter {cname}")

    # ---------------------------------------------------------- properties
    @property
    def name(self):
        return self._name

    @property
    def system(self):
        return self._system

    @property
    def constants(self):
        return dict(self._constants)

    @property
    def dissipation_terms(self):
        return dict(self._dissip_terms)

    @property
    def results(self):
        return np.array(self._results)

    @results.setter
    def results(self, r):
        r = np.array(r)
        if r.shape != self._results.shape:
            raise ValueError("Trying to set an invalid results array")
        self._results = r

    @property
    def x_axis(self):
        return list(self._x_range.keys())[0]

    @property
    def x_axis_values(self):
        v = list(self._x_range.values())[0]
        if self.x_axis == "orient":
            return np.array([q for q, _ in v])
        return np.array(v)

    @property
    def y_axis(self):
        return self._y_axis

    @property
    def results_function(self):
        return self._results_function

    @property
    def avg_N(self):
        return self._avg_N

    @property
    def time_averaged(self):
        return self._time_isavg

    @property
    def celio_k(self):
        return 0

    def __len__(self):
        return len(self._configurations)

    def _snapshot(self, cfg):
        fc, ac, xc = cfg

        def pick(inds, od):
            return {k: v[inds[i]] for i, (k, v) in enumerate(od.items())}

        return ConfigSnapshot(id=cfg, y=self._y_axis, **self._constants,
                              **pick(fc, self._file_ranges), **pick(ac, self._avg_ranges),
                              **pick(xc, self._x_range))

    def __getitem__(self, i):
        if isinstance(i, (int, np.integer)):
            return self._snapshot(self._configurations[i])
        if isinstance(i, slice):
            return [self._snapshot(c) for c in self._configurations[i]]
        raise TypeError(f"Indices must be integer or slices, not {type(i)}")

    def result_index(self, config_id):
        """Index of the results array a configuration contributes to."""
        fc, _, xc = config_id
        return tuple(list(fc) + list(xc))

    def store_time_slice(self, config_id, tslice):
        """results[file, x] += slice / avg_N (simconfig.py:347-368)."""
        if np.ndim(tslice) > 0 and len(tslice) != self._time_N:
            raise ValueError("Time slice has invalid length")
        if self._time_isavg:
            tslice = np.average(tslice)
        self._results[self.result_index(config_id)] += tslice / self._avg_N
