"""Keyword blocks of the input files (muspinsim/input/keyword.py:49-677).

A keyword block is a header line `name [args...]` followed by indented
lines.  Each keyword class states how its block is read:

  block_size        lines that form one entry (3 for tensors)
  accept_range      whether several entries (a range) are allowed
  accept_as_x       whether the keyword can be the x axis
  default           block text used when the block is empty
  expr_size_bounds  admissible number of expressions per line
  _validators       functions row -> error text ("" when valid)

Plain keywords keep their lines as words; evaluated keywords parse every
word into an expression of the expression language with the keyword's
constants and functions (plus the fitting variables); expand keywords also
spread `range(...)`, `zcw(...)`, `eulrange(...)` and `load(...)` results
into one entry per row (keyword.py:278-321).  The class and attribute
names and the error texts are those of the reference, so input files,
callers and their tests read the same.
"""

import inspect

import numpy as np

from ..constants import MU_GAMMA
from ..orientations import eulrange, zcw
from .larkeval import LarkExpression, lark_tokenize
from .variables import FittingVariable

_math_functions = {
    "sin": np.sin, "cos": np.cos, "tan": np.tan,
    "arcsin": np.arcsin, "arccos": np.arccos, "arctan": np.arctan, "arctan2": np.arctan2,
    "exp": np.exp, "log": np.log, "sqrt": np.sqrt,
}
_math_constants = {"pi": np.pi, "deg": np.pi / 180.0, "e": np.exp(1), "inf": np.inf}
# muon_gyr: gamma_mu in MHz/T; MHz: 1 / (2 gamma_mu) in tesla (keyword.py:35)
_phys_constants = {"muon_gyr": MU_GAMMA, "MHz": 1.0 / (2 * MU_GAMMA)}
_reserved_variables = {"x", "y"}


def _range(x1, x2, n=100):
    """n evenly spaced values from x1 to x2, one per row."""
    return np.linspace(x1, x2, int(n))[:, None]


def _zcw(n):
    return zcw(int(n))


def _eulrange(n):
    return eulrange(int(n))


_pwd_functions = {"zcw": _zcw, "eulrange": _eulrange}


def _deepmap(func, obj):
    if isinstance(obj, (list, tuple, np.ndarray)):
        return [_deepmap(func, o) for o in obj]
    return func(obj)


def _call_validator(v, row):
    return (v.__func__ if isinstance(v, staticmethod) else v)(row)


class MuSpinKeyword:
    """A keyword whose lines are kept as lists of words."""

    name = "keyword"
    block_size = 1
    block_size_bounds = (1, np.inf)
    accept_range = True
    accept_as_x = False
    default = None
    expr_size_bounds = (1, np.inf)
    _validators = []

    def __init__(self, block=None, args=None):
        """block: the indented lines (stripped); args: the words after the
        keyword name on the header line (keyword.py:61-155)."""
        block = [] if block is None else block
        args = [] if args is None else args
        lo, hi = self.block_size_bounds
        if not lo <= self.block_size <= hi:
            raise RuntimeError(f"Invalid block_size for keyword {self.name}")
        self._store_args(args)
        lines = np.array(block)
        try:
            rows = lines.reshape((-1, self.block_size))
        except ValueError:
            raise RuntimeError("Invalid number of entries for block, expected "
                               f"{self.block_size}, got {len(lines)}") from None
        if not self.accept_range and len(rows) > 1:
            raise RuntimeError(f"Can not accept range of values for keyword {self.name}")
        from_default = len(rows) == 0
        if from_default:
            if not self.has_default:
                raise RuntimeError(f"Input is empty and keyword '{self.name}' "
                                   "doesn't have a default value")
            rows = np.array([self.default.split("\n")])
        self._store_values(rows)
        self._validate_values()
        if not from_default:
            self._check_entry_sizes(rows)

    # ------------------------------------------------------------ reading
    def _default_args(self):
        """Signature = the in-line arguments the keyword takes."""
        return {}

    def _store_args(self, args):
        n_params = len(inspect.signature(self._default_args).parameters)
        try:
            self._args = self._default_args(*args)
        except TypeError as exc:
            if args:
                raise RuntimeError(f"Wrong number of in-line arguments given '{' '.join(args)}', "
                                   f"expected {n_params}, got {len(args)}") from exc
            raise RuntimeError(f"This keyword requires {n_params} in-line arguments") from exc
        except ValueError as exc:
            raise RuntimeError(
                f"Error parsing keyword argument(s) '{self.name}': {exc}") from exc

    def _store_values(self, rows):
        vals = []
        for row in rows:
            words = [line.split() for line in row]
            vals.append(words[0] if len(words) == 1 else words)
        self._values = np.array(vals)

    def _validate_values(self):
        msgs = [m for v in self._validators for row in self._values
                for m in [_call_validator(v, row)] if m != ""]
        if msgs:
            raise ValueError("\n".join(msgs))

    def _line_lengths(self):
        """(row, line, number of items) for every line of every entry."""
        out = []
        for r, row in enumerate(self._values):
            lines = [row] if self.block_size == 1 else row
            out += [(r, k, len(line)) for k, line in enumerate(lines)]
        return out

    def _check_entry_sizes(self, rows):
        lo, hi = self.expr_size_bounds
        for r, k, n in self._line_lengths():
            if n < lo or n > hi:
                expected = f"between {lo} and {hi}" if lo != hi else lo
                raise RuntimeError(f"Incorrect number of args for entry '{rows[r][k]}', "
                                   f"expected {expected}, got {n}")

    # ------------------------------------------------------------ access
    @property
    def arguments(self):
        return dict(self._args)

    @property
    def id(self):
        return self.name

    @property
    def has_default(self):
        return self.default is not None

    def evaluate(self):
        return self._values.copy()

    def __len__(self):
        return len(self._values)


class MuSpinEvaluateKeyword(MuSpinKeyword):
    """A keyword whose words are expressions (keyword.py:223-275)."""

    name = "evaluate_keyword"
    _functions = dict(_math_functions)
    _constants = dict(_math_constants)
    # names accepted by the expressions without a value at parse time
    _special_variables = None

    def __init__(self, block=None, args=None, variables=None):
        """variables: names of the fitting variables the expressions may use."""
        names = set(variables) if variables is not None else set()
        if self._special_variables:
            names |= set(self._special_variables)
        clash = (set(self._constants) | set(self._functions)) & names
        if clash:
            raise ValueError(f"Variable names '{clash}' conflict with existing constants")
        self._variables = list(set(self._constants) | names)
        super().__init__(block, args)

    def _store_values(self, rows):
        vals = []
        for row in rows:
            exprs = [[LarkExpression(t, variables=self._variables, functions=self._functions)
                      for t in lark_tokenize(line)] for line in row]
            vals.append(exprs[0] if len(exprs) == 1 else exprs)
        self._values = vals

    def evaluate(self, **variables):
        values = {**variables, **self._constants}
        return np.array(_deepmap(lambda e: e.evaluate(**values), self._values))


class MuSpinExpandKeyword(MuSpinEvaluateKeyword):
    """An evaluated keyword whose one-expression lines may produce many rows
    (keyword.py:278-321)."""

    name = "expand_keyword"
    block_size_bounds = (1, 1)
    _functions = {**_math_functions, "range": _range}

    def _reshape_value(self, evaluated_line, line_length):
        """Rows of one evaluated line: a line of numbers is one row; a single
        expression that returned a vector is one row and one that returned a
        table gives one row per table row."""
        if evaluated_line.ndim == 1:
            return [evaluated_line]
        if line_length == 1 and evaluated_line.ndim == 2:
            return [evaluated_line[0]]
        if line_length == 1 and evaluated_line.ndim == 3:
            return list(evaluated_line[0])
        raise RuntimeError(f"Unable to evaluate expression for keyword {self.name}")

    def evaluate(self, **variables):
        if self.block_size != 1:
            raise RuntimeError("MuSpinExpandKeyword can not have block_size > 1")
        values = {**variables, **self._constants}
        rows = []
        for line in self._values:
            got = [e.evaluate(**values) for e in line]
            try:
                arr = np.array(got)
            except ValueError:
                raise RuntimeError(
                    f"Unable to evaluate expression for keyword {self.name}") from None
            rows += self._reshape_value(arr, len(line))
        return rows


class MuSpinCouplingKeyword(MuSpinEvaluateKeyword):
    """A term of the spin Hamiltonian: indices in-line, never a range
    (keyword.py:328-352)."""

    name = "coupling_keyword"
    block_size = 1
    default = "0 0 0"

    def _default_args(self, i=None, j=None):
        return {"i": None if i is None else int(i), "j": None if j is None else int(j)}

    @property
    def accept_range(self):
        return False

    @property
    def id(self):
        i, j = self._args.get("i"), self._args.get("j")
        suffix = "" if i is None else f"_{i}" + ("" if j is None else f"_{j}")
        return self.name + suffix


# ---------------------------------------------------------------- validators
def _check_x_axis(row):
    if row[0] not in InputKeywords or not InputKeywords[row[0]].accept_as_x:
        accepts = [k for k, c in InputKeywords.items() if c.accept_as_x]
        return f"Invalid value {row[0]}, accepts {accepts}"
    return ""


def _check_y_axis(row):
    if len(row) != 1 or row[0] not in ("asymmetry", "integral"):
        return f"Invalid value '{row}', accepts ['asymmetry', 'integral']"
    return ""


def _check_average_axes(row):
    bad = [w for w in row if w not in InputKeywords and w.lower() != "none"]
    if bad:
        return f"Invalid value(s) '{bad}': accepts {list(InputKeywords)} or none"
    return ""


def _check_fitting_method(row):
    if row[0].lower() not in KWFittingMethod.ACCEPTED_FITTING_METHODS:
        return (f"Invalid value '{row[0].lower()}', accepted values "
                f"{KWFittingMethod.ACCEPTED_FITTING_METHODS}")
    return ""


def _check_fitting_tolerance(row):
    if not float(row[0]):
        return f"Invalid value '{row[0]}', accepts only single float value"
    return ""


def _check_not_constant(var):
    if var.name in _math_constants or var.name in _phys_constants:
        return f"Invalid value '{var.name}': variable name conflicts with a constant"
    return ""


def _check_not_reserved(var):
    if var.name in _reserved_variables:
        return (f"Invalid value '{var.name}': variable name conflicts with a reserved "
                "variable name")
    return ""


# ---------------------------------------------------------------- keywords
class KWName(MuSpinKeyword):
    name = "name"
    accept_range = False
    default = "muspinsim"


class KWSpins(MuSpinKeyword):
    name = "spins"
    accept_range = False
    default = "mu e"


class KWCelio(MuSpinKeyword):
    name = "celio"
    expr_size_bounds = (1, 2)
    accept_range = False
    default = "0 0"


class KWPolarization(MuSpinExpandKeyword):
    name = "polarization"
    expr_size_bounds = (1, 3)
    default = "transverse"
    _constants = {**_math_constants, "longitudinal": np.array([0, 0, 1.0]),
                  "transverse": np.array([1.0, 0, 0])}
    _functions = dict(_math_functions)


class KWField(MuSpinExpandKeyword):
    name = "field"
    accept_as_x = True
    default = "0.0"
    _constants = {**_math_constants, **_phys_constants}


class KWIntrinsicField(MuSpinExpandKeyword):
    name = "intrinsic_field"
    accept_as_x = True
    default = "0.0"
    _constants = {**_math_constants, **_phys_constants}


class KWTime(MuSpinExpandKeyword):
    name = "time"
    accept_as_x = True
    default = "range(0, 10, 101)"


class KWXAxis(MuSpinKeyword):
    name = "x_axis"
    accept_range = False
    default = "time"
    _validators = [_check_x_axis]


class KWYAxis(MuSpinKeyword):
    name = "y_axis"
    accept_range = False
    default = "asymmetry"
    _validators = [_check_y_axis]


class KWAverageAxes(MuSpinKeyword):
    name = "average_axes"
    default = "orientation"
    _validators = [_check_average_axes]


class KWOrientation(MuSpinExpandKeyword):
    name = "orientation"
    expr_size_bounds = (1, 4)
    default = "0 0 0"
    _functions = {**_math_functions, **_pwd_functions}

    def _default_args(self, mode="zyz"):
        return {"mode": mode}


class KWTemperature(MuSpinExpandKeyword):
    name = "temperature"
    accept_as_x = True
    default = "inf"


class KWZeeman(MuSpinCouplingKeyword):
    name = "zeeman"
    expr_size_bounds = (3, 3)
    _constants = {**_math_constants, **_phys_constants}

    def _default_args(self, i):
        return {"i": int(i)}


class KWDipolar(MuSpinCouplingKeyword):
    name = "dipolar"
    expr_size_bounds = (3, 3)

    def _default_args(self, i, j):
        return {"i": int(i), "j": int(j)}


class KWHyperfine(MuSpinCouplingKeyword):
    name = "hyperfine"
    block_size = 3
    expr_size_bounds = (3, 3)

    def _default_args(self, i, j=None):
        return {"i": int(i), "j": None if j is None else int(j)}


class KWQuadrupolar(MuSpinCouplingKeyword):
    name = "quadrupolar"
    block_size = 3
    expr_size_bounds = (3, 3)

    def _default_args(self, i):
        return {"i": int(i)}


class KWDissipation(MuSpinCouplingKeyword):
    name = "dissipation"
    expr_size_bounds = (1, 1)

    def _default_args(self, i):
        return {"i": int(i)}


class KWFittingVariables(MuSpinKeyword):
    """One variable per line: `name [start [min [max]]]`, the numbers being
    expressions of the constants (keyword.py:538-583)."""

    name = "fitting_variables"
    default = ""
    _constants = {**_math_constants, **_phys_constants}
    _validators = [_check_not_constant, _check_not_reserved]

    def _store_values(self, rows):
        self._values = []
        consts = list(self._constants)
        for row in rows:
            parts = row[0].split(maxsplit=1)
            if not parts:
                return
            nums = []
            if len(parts) == 2:
                nums = [LarkExpression(t, variables=consts).evaluate(**self._constants)
                        for t in lark_tokenize(parts[1])]
            if len(nums) > 3:
                raise RuntimeError(f"Incorrect number of args for entry '{row[0]}', "
                                   f"expected between 1 and 4, got {len(nums) + 1}")
            self._values.append(FittingVariable(parts[0], *nums))

    def _check_entry_sizes(self, rows):
        pass

    def evaluate(self):
        return list(self._values)


class KWFittingData(MuSpinExpandKeyword):
    name = "fitting_data"
    expr_size_bounds = (1, 2)
    default = ""
    _constants = {}
    _functions = {"load": np.loadtxt}


class KWFittingMethod(MuSpinKeyword):
    ACCEPTED_FITTING_METHODS = ["nelder-mead", "lbfgs", "least-squares"]
    name = "fitting_method"
    accept_range = False
    default = "nelder-mead"
    _validators = [_check_fitting_method]


class KWFittingTolerance(MuSpinKeyword):
    name = "fitting_tolerance"
    accept_range = False
    default = "1e-3"
    _validators = [_check_fitting_tolerance]


class KWResultsFunction(MuSpinExpandKeyword):
    """Function of the results y, the x axis values x and the fitting
    variables, applied after the run; kept unevaluated until then."""

    name = "results_function"
    accept_range = False
    default = "y"
    _constants = {**_math_constants, **_phys_constants}
    _special_variables = ["x", "y"]

    def _reshape_value(self, evaluated_line, line_length):
        # the value keeps the full shape of the results
        return [evaluated_line[0]]


class KWExperiment(MuSpinKeyword):
    name = "experiment"
    accept_range = False
    default = ""


# Every KW class by keyword name, in the order of the class names: the
# order in which evaluated inputs list their couplings (and so the order in
# which terms are added to a spin system)
InputKeywords = {c.name: c for c in sorted(
    (c for n, c in dict(globals()).items() if n.startswith("KW") and inspect.isclass(c)),
    key=lambda c: c.__name__)}
