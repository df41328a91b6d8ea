"""Input-file reader (muspinsim/input/input.py:20-342).

A file is a sequence of keyword blocks: an unindented header line
`keyword [args...]` and the indented lines under it, all indented alike;
'#' starts a comment.  MuSpinInput parses every block into its keyword
class, collects the fitting keywords apart, applies an `experiment`
preset, and `evaluate(**variables)` returns the dictionary that
MuSpinConfig consumes:

  result[keyword]              MuSpinInputValue(name, args, value) for
                               every non-coupling keyword (its default
                               when absent)
  result["couplings"][id]      one MuSpinInputValue per coupling block,
                               id = "zeeman_1", "hyperfine_1_2", ...
  result["results_function"]   the keyword object, evaluated after the run
  result["fitting_info"]       fit, data, method, rtol, single_simulation,
                               evaluate_x_axis
"""

from collections import namedtuple

import numpy as np

from .keyword import InputKeywords, MuSpinCouplingKeyword, MuSpinEvaluateKeyword
from .larkeval import LarkExpressionError


class MuSpinInputError(Exception):
    """Errors found while reading an input file."""


MuSpinInputValue = namedtuple("MuSpinInputValue", ["name", "args", "value"])

# Blocks set by `experiment <kind>`, replacing the file's own blocks of the
# same name (input.py:29-48)
EXPERIMENT_PRESETS = {
    "alc": {"polarization": ["longitudinal"], "y_axis": ["integral"], "x_axis": ["field"]},
    "zero_field": {"field": ["0.0"], "polarization": ["transverse"], "x_axis": ["time"],
                   "y_axis": ["asymmetry"]},
}

_FITTING_KEYWORDS = ("fitting_data", "fitting_tolerance", "fitting_variables",
                     "fitting_method")


def write_error(keyword, block_line_num, err):
    return (f"Error occurred when parsing keyword '{keyword}' "
            f"(block starting at line {block_line_num}):\n{err}")


def _make_blocks(file_stream):
    """({header: [stripped body lines]}, {header: 1-based line of the header})
    (input.py:58-99)."""
    blocks, first_line = {}, {}
    header, indent = None, None
    for num, raw in enumerate(file_stream.readlines(), start=1):
        text = raw.split("#", 1)[0]
        body = text.strip()
        if not body:
            continue
        lead = text[:len(text) - len(text.lstrip())]
        if lead:
            if indent is None:
                indent = lead
            elif lead != indent:
                raise MuSpinInputError("Invalid indent in input file")
            if header is None:
                raise MuSpinInputError("Badly formatted input file")
            blocks[header].append(body)
        else:
            if body in blocks:
                raise MuSpinInputError(f"Redefinition of '{body}' found in input file")
            header, indent = body, None
            blocks[header] = []
            first_line[header] = num
    return blocks, first_line


def _uses_any(kw, names):
    """True when an expression of the keyword names one of `names`."""
    if not names:
        return False
    stack = [kw._values]
    while stack:
        v = stack.pop()
        if isinstance(v, (list, tuple, np.ndarray)):
            stack.extend(v)
        elif hasattr(v, "variables") and v.variables & names:
            return True
    return False


class MuSpinInput:
    def __init__(self, file_stream=None):
        """Read an input file from an open text stream (a file or StringIO)."""
        self._keywords = {}
        self._variables = {}
        self._fitting_info = {
            "fit": False, "data": None, "method": None, "rtol": None, "function": None,
            # True while no simulation parameter depends on a fitting
            # variable: one simulation then serves every fitting step
            "single_simulation": True,
            "evaluate_x_axis": False,
        }
        if file_stream is None:
            return
        blocks, lines = _make_blocks(file_stream)
        errors = self._load_fitting_kw(blocks, lines)
        if errors:
            raise MuSpinInputError(
                "Found {0} Error(s) whilst trying to parse fitting keywords: \n\n{1}".format(
                    len(errors), "\n\n".join(errors)))
        errors = self._apply_experiment(blocks, lines)
        names = set(self._variables)
        for header, block in blocks.items():
            name, *args = header.split()
            try:
                cls = InputKeywords.get(name)
                if cls is None:
                    raise RuntimeError(f"Invalid keyword '{name}' found in input file")
                if issubclass(cls, MuSpinEvaluateKeyword):
                    kw = cls(block, args=args, variables=self._variables)
                    if name != "results_function" and _uses_any(kw, names):
                        self._fitting_info["single_simulation"] = False
                else:
                    kw = cls(block, args=args)
                if kw.id != name:
                    self._keywords.setdefault(name, {})[kw.id] = kw
                else:
                    self._keywords[name] = kw
            except (ValueError, LarkExpressionError, RuntimeError) as exc:
                errors.append(write_error(name, lines.get(header, "?"), exc))
        if errors:
            raise MuSpinInputError(
                "Found {0} error(s) whilst trying to parse keywords: \n\n{1}".format(
                    len(errors), "\n\n".join(errors)))

    def _apply_experiment(self, blocks, lines):
        if "experiment" not in blocks:
            return []
        block = blocks.pop("experiment")
        try:
            words = InputKeywords["experiment"](block).evaluate()[0]
        except RuntimeError as exc:
            return [write_error("experiment", lines["experiment"], exc)]
        kind = words[0] if len(words) else ""
        if kind not in EXPERIMENT_PRESETS:
            return [write_error("experiment", lines["experiment"],
                                f"Invalid experiment type '{kind}' defined, possible types "
                                f"include {list(EXPERIMENT_PRESETS)}")]
        blocks.update({k: list(v) for k, v in EXPERIMENT_PRESETS[kind].items()})
        return []

    def _load_fitting_kw(self, blocks, lines):
        """Fitting variables, data, tolerance and method (input.py:275-342);
        returns the error texts."""
        if "fitting_variables" in blocks:
            try:
                kw = InputKeywords["fitting_variables"](blocks.pop("fitting_variables"))
                self._variables = {v.name: v for v in kw.evaluate()}
            except (RuntimeError, ValueError, LarkExpressionError) as exc:
                return [write_error("fitting_variables", lines["fitting_variables"], exc)]
        if not self._variables:
            return []
        info = self._fitting_info
        info["fit"] = True
        errors = []
        if "fitting_data" not in blocks:
            errors.append(write_error("fitting_variables", lines["fitting_variables"],
                                      "Fitting variables defined without defining any "
                                      "data to fit"))
        else:
            try:
                kw = InputKeywords["fitting_data"](blocks.pop("fitting_data"))
                info["data"] = np.array(kw.evaluate())
            except (RuntimeError, ValueError, LarkExpressionError, IOError) as exc:
                errors.append(write_error("fitting_data", lines["fitting_data"], exc))
        for key, field, conv in (("fitting_tolerance", "rtol", float),
                                 ("fitting_method", "method", str)):
            try:
                kw = InputKeywords[key](blocks.pop(key, []))
                info[field] = conv(kw.evaluate()[0][0])
            except (RuntimeError, ValueError) as exc:
                errors.append(write_error(key, lines.get(key, "?"), exc))
        return errors

    @property
    def variables(self):
        return dict(self._variables)

    @property
    def fitting_info(self):
        return dict(self._fitting_info)

    def evaluate(self, **variables):
        """Every keyword's value for the given fitting-variable values
        (input.py:208-273)."""
        result = {"couplings": {}, "fitting_info": self.fitting_info}
        for name, cls in InputKeywords.items():
            if issubclass(cls, MuSpinCouplingKeyword):
                for kwid, kw in self._keywords.get(name, {}).items():
                    result["couplings"][kwid] = MuSpinInputValue(
                        name, kw.arguments, kw.evaluate(**variables))
            elif name in _FITTING_KEYWORDS:
                continue
            elif name == "results_function":
                result[name] = (self._keywords[name] if name in self._keywords
                                else cls(variables=variables))
            elif name in self._keywords:
                kw = self._keywords[name]
                if name in ("x_axis", "time"):
                    # an x axis the user gave is the one of the final output
                    result["fitting_info"]["evaluate_x_axis"] = True
                v = variables if issubclass(cls, MuSpinEvaluateKeyword) else {}
                result[name] = MuSpinInputValue(name, kw.arguments, kw.evaluate(**v))
            elif cls.default is not None:
                kw = cls()
                result[name] = MuSpinInputValue(name, kw.arguments, np.array(kw.evaluate()))
        return result
