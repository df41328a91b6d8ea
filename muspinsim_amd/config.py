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

import datetime
import os
import re
from collections import OrderedDict, namedtuple
from itertools import product
from numbers import Number

import numpy as np

from .orientations import euler_zyz, normalised_weights, stored_orientations

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


# coupling keyword -> (value shape, label) (simconfig.py:34-40,615-641)
_COUPLINGS = {
    "zeeman": ((3,), "Zeeman"),
    "dipolar": ((3,), "dipolar"),
    "hyperfine": ((3, 3), "hyperfine"),
    "quadrupolar": ((3, 3), "quadrupolar"),
    "dissipation": ((1,), "dissipation"),
}


def _spin_name(s):
    """'2H' -> ('H', 2); other names unchanged (simconfig.py:528-537)."""
    m = re.match("([0-9]+)([A-Z][a-z]*|e)", s)
    return (m.group(2), int(m.group(1))) if m else str(s)


def _celio_pair(words):
    """(k, averages) of the celio keyword (simconfig.py:539-569)."""
    out = []
    for what, w in zip(("k", "averages"), words):
        try:
            v = int(w)
        except ValueError:
            raise MuSpinConfigError(
                f"Value of {what} for Celio's method must be an integer") from None
        if v < 0:
            raise MuSpinConfigError(
                f"Value of {what} for Celio's method must a postive integer or 0")
        out.append(v)
    out += [0] * (2 - len(out))
    return out[0], out[1]


def _entries(value):
    """Entries of an evaluated keyword as flat float arrays."""
    return [np.asarray(e, dtype=float).reshape(-1) for e in value]


def _args_from_input(params, use_experimental_x_axis=True):
    """Constructor arguments from MuSpinInput.evaluate's dictionary
    (simconfig.py:82-321): name, spins and isotopes, axes, parameter ranges,
    the fitting x axis, and the system built from the coupling blocks
    (indices 1-based in the file)."""
    from .spinsys import MuonSpinSystem
    names = params["name"].value
    if any(len(row) > 1 for row in names):
        raise MuSpinConfigError("Name must be a word without spaces")
    x_kw = params["x_axis"].value[0][0]
    if x_kw not in KEYWORD_FIELDS:
        raise MuSpinConfigError("Invalid x axis name in input file")
    avg = []
    for a in np.asarray(params["average_axes"].value).reshape(-1):
        if a.lower() == "none":
            continue
        if a not in KEYWORD_FIELDS:
            raise MuSpinConfigError("Invalid average axis name in input file")
        avg.append(str(a))
    for kw in KEYWORD_FIELDS:
        if kw not in params:
            raise MuSpinConfigError(
                f"Invalid params object passed to MuSpinConfig: missing {kw}")
    y_axis = str(params["y_axis"].value[0][0])
    vals = {kw: list(params[kw].value) for kw in KEYWORD_FIELDS}
    vals["time"] = []
    for e in _entries(params["time"].value):
        if e.size != 1:
            raise MuSpinConfigError("Invalid line in time range")
        vals["time"].append(float(e[0]))
    vals["temperature"] = [float(e[0]) for e in _entries(params["temperature"].value)]
    for kw in ("field", "intrinsic_field"):
        vals[kw] = _entries(vals[kw])
        for e in vals[kw]:
            if e.size not in (1, 3):
                raise MuSpinConfigError(
                    "Invalid magnetic field value" if kw == "field" else
                    "Invalid intrinsic magnetic field value")
    vals["polarization"] = _entries(vals["polarization"])
    vals["orientation"] = _entries(vals["orientation"])

    finfo = params["fitting_info"]
    if finfo["fit"]:
        n = {kw: (1 if kw == "time" and y_axis == "integral" else len(v))
             for kw, v in vals.items()}
        if any(n[kw] > 1 for kw in KEYWORD_FIELDS if kw != x_kw and kw not in avg):
            raise MuSpinConfigError("Can not have file ranges when fitting")
        if use_experimental_x_axis or n[x_kw] <= 1:
            xdata = np.asarray(finfo["data"], dtype=float)[:, 0]
            if x_kw in ("field", "intrinsic_field"):
                vals[x_kw] = [np.array([0.0, 0.0, v]) for v in xdata]
            else:
                vals[x_kw] = list(xdata)

    celio = _celio_pair(params["celio"].value[0])
    system = MuonSpinSystem([_spin_name(s) for s in params["spins"].value[0]], celio[0])
    dissipation = {}
    for c in params["couplings"].values():
        if c.name not in _COUPLINGS:
            raise MuSpinConfigError(
                f"Invalid params object passed to MuSpinConfig: unknown {c.name} coupling")
        idx = [v for v in c.args.values() if v is not None]
        if any(v < 1 or v > len(system) for v in idx):
            raise MuSpinConfigError("Out of range indices for coupling")
        shape, label = _COUPLINGS[c.name]
        v = np.asarray(c.value[0], dtype=float)
        if v.shape != shape:
            raise MuSpinConfigError(f"Invalid shape for {label} coupling term")
        i = c.args["i"] - 1
        j = c.args.get("j")
        j = None if j is None else j - 1
        if c.name == "zeeman":
            system.add_zeeman_term(i, v)
        elif c.name == "dipolar":
            system.add_dipolar_term(i, j, v)
        elif c.name == "hyperfine":
            system.add_hyperfine_term(i, v, j)
        elif c.name == "quadrupolar":
            system.add_quadrupolar_term(i, v)
        else:
            dissipation[i] = float(v[0])

    return dict(system=system, name=str(names[0][0]), field=vals["field"],
                intrinsic_field=vals["intrinsic_field"], polarization=vals["polarization"],
                time=vals["time"], orientation=vals["orientation"],
                temperature=vals["temperature"], x_axis=x_kw, y_axis=y_axis,
                average_axes=avg, dissipation=dissipation,
                euler_mode=params["orientation"].args.get("mode", "zyz"),
                results_function=params.get("results_function"),
                celio=celio)


def _print_value(key, v):
    """A file-range value as the .dat header shows it (simconfig.py:636-647)."""
    if key in ("B", "intrinsic_B"):
        return f"{v} T"
    if key == "orient":
        q, w = v
        a, b, c = euler_zyz(q) * 180 / np.pi
        return f"[ZYZ] a = {a:.1f} deg, b = {b:.1f} deg, c = {c:.1f} deg, weight = {w}"
    return f"{v}"


class MuSpinConfig:
    def __init__(self, system=None, name="muspinsim", field=0.0, intrinsic_field=0.0,
                 polarization="transverse", time=None, orientation=((0.0, 0.0, 0.0),),
                 temperature=np.inf, x_axis="time", y_axis="asymmetry",
                 average_axes=("orientation",), dissipation=None, euler_mode="zyz",
                 results_function=None, celio=(0, 0), params=None,
                 use_experimental_x_axis=True):
        """system: a MuonSpinSystem with its coupling terms; dissipation:
        {spin index: rate in MHz}; results_function: optional callable
        f(x=..., y=..., **variables) applied to the results by run();
        celio: (k, averages) of Celio's method, (0, 0) for exact evolution.

        Drop-in form: MuSpinConfig(params) with the dictionary returned by
        MuSpinInput.evaluate (simconfig.py:82-321); an empty dictionary (or
        no argument) gives an empty configuration.  use_experimental_x_axis
        (fitting only): the x axis becomes the x column of the fitting data.
        """
        from .spinsys import MuonSpinSystem
        if isinstance(system, dict) and params is None:
            params, system = system, None
        self._constants = {}
        self._file_ranges = OrderedDict()
        self._avg_ranges = OrderedDict()
        self._x_range = OrderedDict()
        self._celio = (0, 0)
        if params is not None:
            if len(params) == 0:
                return
            self.__init__(**_args_from_input(params, use_experimental_x_axis))
            return
        if system is None:
            return
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
        if len(celio) != 2 or any(int(c) != c or c < 0 for c in celio):
            raise MuSpinConfigError("celio must be (k, averages), non-negative integers")
        self._celio = (int(celio[0]), int(celio[1]))
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
    def spins(self):
        return list(self._system.spins)

    @property
    def celio_k(self):
        """Trotter steps of Celio's method (0: exact evolution)."""
        return self._celio[0]

    @property
    def celio_averages(self):
        """Random-state averages of Celio's method (0: none)."""
        return self._celio[1]

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

    def save_output(self, name=None, path=".", extension=".dat"):
        """One text file per combination of file-range values, named
        name[_i_j...]extension, with the columns x (|B| for a field axis) and
        result, and a '#' header listing the file-range values
        (simconfig.py:370-429)."""
        from . import __version__
        name = self.name if name is None else name
        head = (f"MUSPINSIM v.{__version__}\nOutput file written on "
                f"{datetime.datetime.now().ctime()}\nParameters used:\n")
        x = self.x_axis_values
        if self.x_axis in ("B", "intrinsic_B"):
            x = np.linalg.norm(x, axis=-1)
        elif self.x_axis == "orient":
            x = np.arange(len(x))
        keys = list(self._file_ranges)
        written = []
        for inds in product(*[range(len(v)) for v in self._file_ranges.values()]):
            header = head + "".join(
                "\t{0:<20} = {1}\n".format(k, _print_value(k, self._file_ranges[k][i]))
                for k, i in zip(keys, inds))
            fname = os.path.join(path, name + "".join(f"_{i}" for i in inds) + extension)
            np.savetxt(fname, np.stack([x, self._results[inds]], axis=1), header=header)
            written.append(fname)
        return written

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
