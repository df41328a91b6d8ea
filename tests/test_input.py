"""Input-file front end on the CPU: the expression language, keyword blocks,
the reader and the configuration built from it.  The expected values and
error texts are the known answers of the reference's tests/test_lark.py,
tests/test_input.py and tests/test_config.py; the operator precedences and
number forms are pinned against Python's own arithmetic on the same text."""

import os
from io import StringIO

import numpy as np
import pytest

from muspinsim_amd.config import MuSpinConfig, MuSpinConfigError
from muspinsim_amd.input import MuSpinInput, MuSpinInputError
from muspinsim_amd.input.keyword import (InputKeywords, KWFittingMethod, MuSpinCouplingKeyword,
                                         MuSpinEvaluateKeyword, MuSpinExpandKeyword,
                                         MuSpinKeyword)
from muspinsim_amd.input.larkeval import LarkExpression, LarkExpressionError, lark_tokenize
from muspinsim_amd.orientations import from_euler, qmul, axis_angle, conj

REF_EXAMPLES = "/root/reference/examples"


def _double(x):
    return 2 * x


# ------------------------------------------------------------------ expressions
@pytest.mark.parametrize("src,funcs,variables,out", [
    ("3+2*6^2/9", {}, {}, 11),
    ("double(2)", {"double": _double}, {}, 4),
    ("double(3+2*6^2/9)", {"double": _double}, {}, 22),
    ("double(3+2* 6^2/9) - double(2 *5)", {"double": _double}, {}, 2),
    ("x+1", {}, {"x": 5}, 6),
    ("double(x)", {"double": _double}, {"x": 2}, 4),
    ("x+y+1", {}, {"x": 2, "y": 5}, 8),
])
def test_expression_values(src, funcs, variables, out):
    e = LarkExpression(src, functions=funcs, variables=list(variables))
    assert e.evaluate(**variables) == out


@pytest.mark.parametrize("src,funcs,variables,message", [
    ("notafunc(123)", {"double": _double}, {}, "Invalid function: 'notafunc()', valid "
                                               "functions are ['double()']"),
    ("3*", {}, {}, "Invalid characters in LarkExpression: Unexpected token"),
    ("", {}, {}, "Empty String"),
    ("x+1", {}, {"x": None}, "Some necessary variable(s) {'x'} have not been defined "
                             "when evaluating LarkExpression"),
    ("x+1", {}, {"y": 2}, "Invalid variable/constant: 'x', valid variables/constants "
                          "are ['y']"),
])
def test_expression_errors(src, funcs, variables, message):
    with pytest.raises(LarkExpressionError) as err:
        e = LarkExpression(src, functions=funcs, variables=list(variables))
        e.evaluate(**{k: v for k, v in variables.items() if v is not None})
    assert message in str(err.value)


def test_expression_extra_variable():
    with pytest.raises(LarkExpressionError) as err:
        LarkExpression("x+1", variables="x").evaluate(x=1, y=1)
    assert str(err.value) == ("Some invalid variable(s) {'y'} have been defined when "
                              "evaluating LarkExpression")


def test_tokenize():
    toks = lark_tokenize("3+0.4 2.3 sin(x) atan2(3, 4)")
    assert len(toks) == 4
    ex = [LarkExpression(t, variables="x", functions={"sin": np.sin, "atan2": np.arctan2})
          for t in toks]
    assert ex[0].evaluate() == 3.4 and ex[1].evaluate() == 2.3
    assert abs(ex[2].evaluate(x=np.pi / 2) - 1.0) < 1e-15
    assert abs(ex[3].evaluate() - np.arctan2(3.0, 4.0)) < 1e-15


@pytest.mark.parametrize("src,py", [
    ("2^3^2", "(2**3)**2"),          # '^' is left associative
    ("-2^2", "(-2)**2"),             # unary minus binds tighter than '^'
    ("2^-1", "2**-1"),
    ("--3", "3"),
    ("1-2-3", "(1-2)-3"),
    ("8/4/2", "(8/4)/2"),
    ("2*3^2", "2*(3**2)"),
    ("1.27e-2+.5+5.+1E3", "1.27e-2+.5+5.+1E3"),
    ("(1+2)*(3-4)/5", "(1+2)*(3-4)/5"),
])
def test_expression_precedence(src, py):
    assert LarkExpression(src).evaluate() == eval(py)


def test_expression_strings_and_numbers_are_floats():
    assert LarkExpression('"a b\\"c"').evaluate() == 'a b\\"c'
    assert isinstance(LarkExpression("3").evaluate(), float)


@pytest.mark.parametrize("line,err", [("1 $ 2", LarkExpressionError),
                                      ("(1 + 2", RuntimeError),
                                      ("2x", RuntimeError)])
def test_tokenize_errors(line, err):
    with pytest.raises(err):
        lark_tokenize(line)


# ------------------------------------------------------------------ keywords
def test_plain_keyword_and_defaults():
    kw = MuSpinKeyword(["a b c", "d e f"])
    assert (kw.evaluate() == [["a", "b", "c"], ["d", "e", "f"]]).all() and len(kw) == 2

    class DefKeyword(MuSpinKeyword):
        default = "1"
    assert DefKeyword().evaluate()[0][0] == "1"
    with pytest.raises(RuntimeError) as err:
        MuSpinKeyword()
    assert str(err.value) == "Input is empty and keyword 'keyword' doesn't have a default value"


def test_evaluate_and_expand_keywords():
    assert (MuSpinEvaluateKeyword(["exp(0) 1+1 2^2"]).evaluate()[0] == [1, 2, 4]).all()
    assert len(MuSpinExpandKeyword(["range(0, 1)", "5.0 2.0"]).evaluate()) == 101

    class RepeatKW(MuSpinExpandKeyword):
        _functions = {"repeat3": lambda x: [x, x, x]}
    assert (RepeatKW(["repeat3(1)"]).evaluate()[0] == [1, 1, 1]).all()


def test_keyword_argument_errors():
    with pytest.raises(RuntimeError) as err:
        MuSpinKeyword([], args=["a"])
    assert str(err.value) == "Wrong number of in-line arguments given 'a', expected 0, got 1"
    with pytest.raises(RuntimeError) as err:
        MuSpinCouplingKeyword([], args=["a"])
    assert str(err.value) == ("Error parsing keyword argument(s) 'coupling_keyword': invalid "
                              "literal for int() with base 10: 'a'")


@pytest.mark.parametrize("kw,args,block,out", [
    ("name", [], ["othername"], "othername"),
    ("name", [], [], "muspinsim"),
    ("spins", [], ["F mu F"], np.array(["F", "mu", "F"])),
    ("spins", [], [], np.array(["mu", "e"])),
    ("polarization", [], ["0 1 1"], np.array([0, 1, 1])),
    ("polarization", [], [], np.array([1, 0, 0])),
    ("polarization", [], ["longitudinal"], np.array([0, 0, 1.0])),
    ("polarization", [], ["transverse"], np.array([1.0, 0, 0])),
    ("field", [], ["range(0, 20, 21)"], np.arange(21)[:, None]),
    ("field", [], [], np.array([0])),
    ("intrinsic_field", [], ["range(0, 20, 21)"], np.arange(21)[:, None]),
    ("intrinsic_field", [], [], np.array([0])),
    ("time", [], [], np.linspace(0, 10, 101)[:, None]),
    ("time", [], ["range(0, 10, 5)"], np.array([0, 2.5, 5, 7.5, 10])[:, None]),
    ("time", [], ["10", "20", "30", "range(0, 10, 5)"],
     np.array([10, 20, 30, 0, 2.5, 5, 7.5, 10])[:, None]),
    ("y_axis", [], ["asymmetry"], np.array(["asymmetry"])),
    ("y_axis", [], [], np.array(["asymmetry"])),
    ("orientation", [], [], np.array([0, 0, 0])),
    ("zeeman", ["1"], ["0 0 1"], np.array([0, 0, 1])),
    ("zeeman", ["1"], [], np.array([0, 0, 0])),
    ("hyperfine", ["1"], ["1 0 0", "0 1 0", "0 0 1"], np.eye(3)),
    ("hyperfine", ["1"], [], np.zeros((3, 3))),
    ("dipolar", ["1", "2"], ["0 0 1"], np.array([0, 0, 1])),
    ("dipolar", ["1", "2"], [], np.array([0, 0, 0])),
    ("fitting_method", [], ["nelder-mead"], np.array(["nelder-mead"])),
    ("fitting_method", [], ["lbfgs"], np.array(["lbfgs"])),
    ("fitting_method", [], ["least-squares"], np.array(["least-squares"])),
    ("celio", [], ["1"], np.array(["1"])),
    ("celio", [], ["1 2"], np.array(["1", "2"])),
    ("celio", [], [], np.array(["0", "0"])),
])
def test_keyword_values(kw, args, block, out):
    k = InputKeywords[kw](block, args=args)
    assert np.all(np.asarray(k.evaluate()) == out)
    assert k.id == kw + "".join(f"_{a}" for a in args)


def test_keyword_constants_and_special_forms():
    # MHz = 1 / (2 gamma_mu) tesla: 500 MHz -> 1.84449016 T
    assert np.isclose(InputKeywords["field"](["500*MHz"]).evaluate()[0][0], 1.84449016)
    assert np.isclose(InputKeywords["intrinsic_field"](["500*MHz"]).evaluate()[0][0],
                      1.84449016)
    assert len(InputKeywords["orientation"](["zcw(20)"]).evaluate()) >= 20
    assert len(InputKeywords["orientation"](["eulrange(5)"]).evaluate()) == 125
    q = InputKeywords["quadrupolar"](["1 0 0", "0 1 0", "0 0 cos(1)^2+sin(1)^2"], args=["1"])
    assert np.allclose(q.evaluate()[0], np.eye(3)) and q.id == "quadrupolar_1"
    assert InputKeywords["hyperfine"](["1 0 0", "0 1 0", "0 0 1"], args=["1", "3"]).id == \
        "hyperfine_1_3"


def test_keyword_value_errors():
    with pytest.raises(ValueError) as err:
        InputKeywords["y_axis"](["something"], args=[])
    assert str(err.value) == "Invalid value '['something']', accepts ['asymmetry', 'integral']"
    with pytest.raises(ValueError) as err:
        InputKeywords["fitting_method"](["something"], args=[])
    assert str(err.value) == ("Invalid value 'something', accepted values "
                              f"{KWFittingMethod.ACCEPTED_FITTING_METHODS}")
    with pytest.raises(RuntimeError) as err:
        InputKeywords["celio"](["1 2 3"], args=[])
    assert str(err.value) == ("Incorrect number of args for entry '1 2 3', expected between "
                              "1 and 2, got 3")
    with pytest.raises(RuntimeError):
        InputKeywords["zeeman"](["1 0"], args=["1"])           # 3 expressions needed
    with pytest.raises(RuntimeError):
        InputKeywords["hyperfine"](["1 0 0", "0 1", "0 0 1"], args=["1"])
    with pytest.raises(RuntimeError):
        InputKeywords["zeeman"](["1 0 0", "0 1 0"], args=["1"])  # couplings take no range
    with pytest.raises(RuntimeError):
        InputKeywords["hyperfine"](["1 0 0", "0 1 0"], args=["1"])  # 3 lines per entry
    with pytest.raises(ValueError):
        InputKeywords["x_axis"](["orientation"])                 # not an x axis keyword


# ------------------------------------------------------------------ reader
def _inp(text):
    return MuSpinInput(StringIO(text))


def test_input_valid_and_results_function():
    e = _inp("\nname\n    test_1\nspins\n    mu H\nzeeman 1\n    1 0 0\n").evaluate()
    assert e["name"].value[0][0] == "test_1"
    assert (e["spins"].value[0] == ["mu", "H"]).all()
    assert (e["couplings"]["zeeman_1"].value[0] == [1, 0, 0]).all()
    np.testing.assert_allclose(e["results_function"]._values[0][0].evaluate(y=np.arange(10)),
                               np.arange(10))
    e = _inp("spins\n    mu H\nresults_function\n    2*y+sin(x)\n").evaluate()
    f = e["results_function"]._values[0][0]
    np.testing.assert_allclose(f.evaluate(y=np.arange(5), x=np.arange(5)),
                               2 * np.arange(5) + np.sin(np.arange(5)))


@pytest.mark.parametrize("text,message", [
    ("\n    name\n        test_1\nspins\n    mu H\n", "Badly formatted input file"),
    ("\nname\n    test_1\nhyperfine 1\n    1 0 0\n     2 0 0\n    3 0 0\n",
     "Invalid indent in input file"),
    ("\nname\n    test_1\nspins\n    mu H\nnotakeyword 1\n    1 0 0\n",
     "Found 1 error(s) whilst trying to parse keywords: \n\nError occurred when parsing "
     "keyword 'notakeyword' (block starting at line 6):\nInvalid keyword 'notakeyword' found "
     "in input file"),
    ("zeeman 1\n    1 0 0\nzeeman 1\n    0 1 0\n", "Redefinition of 'zeeman 1' found in input file"),
    ("hyperfine 2\n    1 0 0\n    0 1 0\n    0 0 1\nhyperfine 2\n    1 0 0\n    0 1 0\n"
     "    0 0 1\n", "Redefinition of 'hyperfine 2' found in input file"),
    ("name\n    a\nspins\n    mu e\nname\n    b\n", "Redefinition of 'name' found in input file"),
    ("experiment\n    nmr\n", "Invalid experiment type 'nmr' defined, possible types include "
                              "['alc', 'zero_field']"),
])
def test_input_errors(text, message):
    with pytest.raises(MuSpinInputError) as err:
        _inp(text).evaluate()
    assert message in str(err.value)


def test_comments_and_line_numbers():
    text = "# header comment\nname   # trailing\n    c1 # x\n\nspins\n\tmu e\nzeeman 4\n    0 0 1\n"
    e = _inp(text).evaluate()
    assert e["name"].value[0][0] == "c1" and "zeeman_4" in e["couplings"]
    with pytest.raises(MuSpinInputError) as err:
        _inp("# c\n\nbogus\n    1\n")
    assert "(block starting at line 3)" in str(err.value)


def test_experiment_presets():
    e = _inp("spins\n    mu e\nexperiment\n    alc\nx_axis\n    time\n").evaluate()
    assert (e["polarization"].value[0] == [0, 0, 1]).all()
    assert e["y_axis"].value[0][0] == "integral" and e["x_axis"].value[0][0] == "field"
    e = _inp("spins\n    mu e\nfield\n    3\nexperiment\n    zero_field\n").evaluate()
    assert e["field"].value[0][0] == 0.0 and e["x_axis"].value[0][0] == "time"
    assert (e["polarization"].value[0] == [1, 0, 0]).all()


_FIT_HEAD = "fitting_variables\n    A 1.0 0.0 2.0\nfitting_data\n    0  0.0\n    1  1.0\n" \
            "    2  4.0\n    3  9.0\n"


def test_input_fitting():
    i1 = _inp(_FIT_HEAD + "field\n    2*A\nzeeman 1\n    A A 0\n")
    info = i1.fitting_info
    assert info["fit"] and (info["data"] == [[0, 0], [1, 1], [2, 4], [3, 9]]).all()
    e = i1.evaluate(A=2.0)
    assert e["field"].value[0][0] == 4.0
    assert (e["couplings"]["zeeman_1"].value[0] == [2, 2, 0]).all()
    assert i1.variables["A"].value == 1.0 and i1.variables["A"].bounds == (0.0, 2.0)
    assert not info["single_simulation"]
    assert info["method"] == "nelder-mead" and info["rtol"] == 1e-3


@pytest.mark.parametrize("body,single", [
    ("field\n    2\nzeeman 1\n    1 1 0\nresults_function\n    2*A\n", True),
    ("field\n    2*A\nzeeman 1\n    1 1 0\nresults_function\n    2*A\n", False),
    ("field\n    2/muon_gyr\nzeeman 1\n    1 1 0\nresults_function\n    2*A\n", True),
    # a variable on a later line of a range also needs one simulation per step
    ("field\n    1\n    2*A\n", False),
])
def test_single_simulation_detection(body, single):
    assert _inp(_FIT_HEAD + body).fitting_info["single_simulation"] is single


def test_fitting_data_from_file(tmp_path):
    data = np.stack([np.linspace(0, 1, 10), np.linspace(0, 1, 10) ** 2], axis=1)
    path = tmp_path / "d.dat"
    np.savetxt(path, data)
    info = _inp(f'fitting_variables\n    A\nfitting_data\n    load("{path}")\n'
                "fitting_method\n    nelder-mead\n").fitting_info
    assert info["fit"] and np.allclose(info["data"], data)
    assert info["method"] == "nelder-mead" and abs(info["rtol"] - 1e-3) < 1e-15


@pytest.mark.parametrize("text,message", [
    ("fitting_variables\n    A 1.0 0.0 5.0\n",
     "Found 1 Error(s) whilst trying to parse fitting keywords: \n\nError occurred when "
     "parsing keyword 'fitting_variables' (block starting at line 1):\nFitting variables "
     "defined without defining any data to fit"),
    ("fitting_variables\n    A 1.0 0.0 -5.0\nfitting_data\n    0 1\n",
     "Variable A has invalid range: (max value -5.0 cannot be less than or equal to min value "
     "0.0)\nVariable A has invalid starting value: (starting value 1.0 cannot be greater than "
     "max value -5.0)"),
    ("fitting_variables\n    MHz 1.0 0.0 5.0\nfitting_data\n    0 1\n",
     "Invalid value 'MHz': variable name conflicts with a constant"),
    ("fitting_variables\n    x 1.0 0.0 5.0\nfitting_data\n    0 1\n",
     "Invalid value 'x': variable name conflicts with a reserved variable name"),
    ("fitting_variables\n    y 1.0 0.0 5.0\nfitting_data\n    0 1\n",
     "Invalid value 'y': variable name conflicts with a reserved variable name"),
])
def test_fitting_errors(text, message):
    with pytest.raises(MuSpinInputError) as err:
        _inp(text)
    assert message in str(err.value)


# ------------------------------------------------------------------ config
_CONFIG_TEXT = """
spins
    mu 2H e
field
    range(1, 11, 2)
intrinsic_field
    range(1, 21, 2)
temperature
    range(0, 10, 2)
time
    range(0, 10, 21)
orientation
    0 0
    0 180
polarization
    0 1 1
zeeman 1
    1 0 0
dipolar 1 2
    0 1 0
hyperfine 1 3
    10 2 2
    2 10 2
    2 2 10
quadrupolar 2
    -2 0 0
    0  1 0
    0  0 1
dissipation 2
    0.1
"""


def test_config_from_input():
    cfg = MuSpinConfig(_inp(_CONFIG_TEXT).evaluate())
    assert cfg.name == "muspinsim" and len(cfg) == 16
    assert cfg.results.shape == (2, 2, 2, 21)
    c0 = cfg[0]
    assert (c0.B == [0, 0, 1]).all() and (c0.intrinsic_B == [0, 0, 1]).all() and c0.T == 0
    assert (np.array(c0.t) == np.linspace(0, 10, 21)).all()
    assert np.isclose(np.linalg.norm(c0.mupol), 1.0)
    for c in cfg:
        cfg.store_time_slice(c.id, np.ones(len(c.t)))
    assert (cfg.results == 1).all()
    assert cfg.system.spins == ["mu", ("H", 2), "e"]
    assert len(cfg.system._terms) == 4 and cfg.dissipation_terms == {1: 0.1}
    assert set(cfg._file_ranges) == {"B", "intrinsic_B", "T"}
    assert "t" in cfg._x_range and "orient" in cfg._avg_ranges
    a = np.arange(5)
    np.testing.assert_allclose(cfg.results_function.evaluate(x=2 * a, y=a)[0], a)


@pytest.mark.parametrize("text,err", [
    ("spins\n    mu e\nzeeman 3\n    0 0 1\n", MuSpinConfigError),     # index out of range
    ("y_axis\n    integral\n", MuSpinConfigError),                     # time x axis, integral
    ("spins\n    mu e\ncelio\n    10 -8\n", MuSpinConfigError),
    ("spins\n    mu e\ncelio\n    -1 0\n", MuSpinConfigError),
    ("x_axis\n    field\n", MuSpinConfigError),                          # x axis not a range
])
def test_config_errors(text, err):
    with pytest.raises(err):
        MuSpinConfig(_inp(text).evaluate())
    with pytest.raises(MuSpinInputError):
        _inp("spins\n    mu e\nzeeman 1\n    2 0\n")


@pytest.mark.parametrize("block,k,avg", [("", 0, 0), ("celio\n    10\n", 10, 0),
                                          ("celio\n    0 0\n", 0, 0),
                                          ("celio\n    10 8\n", 10, 8)])
def test_config_celio(block, k, avg):
    cfg = MuSpinConfig(_inp("spins\n    mu e\n" + block).evaluate())
    assert cfg.celio_k == k and cfg.celio_averages == avg


def _q_zyz(a, b, c, mid=(0, 1, 0)):
    """Conjugate of Qz(c) Qmid(b) Qz(a), composed from axis-angle factors."""
    return conj(qmul(qmul(axis_angle((0, 0, 1), c), axis_angle(mid, b)),
                     axis_angle((0, 0, 1), a)))


def _same_rotation(q1, q2):
    return min(np.abs(q1 - q2).max(), np.abs(q1 + q2).max()) < 1e-12


def test_config_orientations():
    cfg = MuSpinConfig(_inp("orientation\n    0  0  0  1.0\n    0.5*pi 0 0 2.0\n").evaluate())
    orange = cfg._avg_ranges["orient"]
    assert sum(w for _, w in orange) == 2.0
    assert np.allclose(orange[1][0], [2 ** -0.5, 0, 0, -(2 ** -0.5)])
    rng = np.linspace(0, np.pi, 4)
    angles = np.array(np.meshgrid(rng, rng, rng)).reshape((3, -1)).T
    block = "\n".join("\t{0} {1} {2}".format(*a) for a in angles)
    for mode, mid in (("", (0, 1, 0)), (" zxz", (1, 0, 0))):
        cfg = MuSpinConfig(_inp("orientation" + mode + "\n" + block).evaluate())
        for (a, b, c), (q, _) in zip(angles, cfg._avg_ranges["orient"]):
            assert _same_rotation(q, _q_zyz(a, b, c, mid))
    pairs = np.array(np.meshgrid(rng, rng)).reshape((2, -1)).T
    block = "\n".join("\t{0} {1}".format(*a) for a in pairs)
    cfg = MuSpinConfig(_inp("orientation\n" + block).evaluate())
    for (theta, phi), (q, _) in zip(pairs, cfg._avg_ranges["orient"]):
        assert _same_rotation(q, _q_zyz(phi, theta, phi))


def test_config_fitting_axes():
    head = "fitting_variables\n    A\nfitting_data\n    0   0.5\n    1   0.2\n"
    cfg = MuSpinConfig(_inp(head).evaluate(A=0.0))
    assert (np.array(cfg._x_range["t"]) == [0, 1]).all()
    with pytest.raises(MuSpinConfigError):
        MuSpinConfig(_inp(head + "average_axes\n    none\norientation\n    0  0\n    0  1\n")
                     .evaluate(A=0.0))
    for kw, key in (("field", "B"), ("intrinsic_field", "intrinsic_B")):
        cfg = MuSpinConfig(_inp(f"y_axis\n    integral\nx_axis\n    {kw}\n" + head)
                           .evaluate(A=0.0))
        assert (np.array(cfg._x_range[key]) == [[0, 0, 0], [0, 0, 1]]).all()
    # the file's own x axis is kept when asked, for the final output
    cfg = MuSpinConfig(_inp(head + "time\n    range(0, 2, 5)\n").evaluate(A=0.0),
                       use_experimental_x_axis=False)
    assert len(cfg._x_range["t"]) == 5


def test_save_output(tmp_path):
    cfg = MuSpinConfig(_inp("name\n    out\nspins\n    mu e\nfield\n    0\n    1\n"
                            "time\n    range(0, 1, 4)\n").evaluate())
    cfg.results = np.arange(8.0).reshape(2, 4)
    files = cfg.save_output(path=str(tmp_path))
    assert [os.path.basename(f) for f in files] == ["out_0.dat", "out_1.dat"]
    d = np.loadtxt(files[1])
    assert np.allclose(d, np.stack([np.linspace(0, 1, 4), [4, 5, 6, 7]], axis=1))
    head = open(files[1]).read()
    assert head.startswith("# MUSPINSIM v.") and "B" in head and "1.] T" in head
    cfg = MuSpinConfig(_inp("spins\n    mu e\nfield\n    range(0, 2, 3)\nx_axis\n    field\n"
                            "y_axis\n    integral\n").evaluate())
    cfg.results = np.array([1.0, 2.0, 3.0])
    d = np.loadtxt(cfg.save_output(name="alc", path=str(tmp_path))[0])
    assert np.allclose(d, [[0, 1], [1, 2], [2, 3]])


# ------------------------------------------------------------------ reference examples
def _example_inputs():
    if not os.path.isdir(REF_EXAMPLES):
        return []
    out = []
    for root, _, files in os.walk(REF_EXAMPLES):
        out += [os.path.join(root, f) for f in sorted(files) if f.endswith(".in")]
    return sorted(out)


@pytest.mark.skipif(not _example_inputs(), reason="the reference's example inputs are not "
                                                  "present on this machine")
@pytest.mark.parametrize("path", _example_inputs(), ids=os.path.basename)
def test_reference_examples_parse(path, monkeypatch):
    """Every example input of the reference reads and gives a configuration
    (load() paths are relative to the example's folder)."""
    monkeypatch.chdir(os.path.dirname(path))
    with open(path) as fh:
        inp = MuSpinInput(fh)
    values = {n: v.value for n, v in inp.variables.items()}
    cfg = MuSpinConfig(inp.evaluate(**values))
    assert len(cfg) >= 1 and cfg.results.size >= 1


# ------------------------------------------------------------------ input files -> oracle
GOLDEN_INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")

CLOSED_FORMS = {
    "muon_precession": lambda t: 0.5 * np.cos(4 * np.pi * t),
    "muon_dissipation": lambda t: 0.5 * np.cos(4 * np.pi * t) * np.exp(-5.0 * t),
    "muon_local_zeeman": lambda t: 0.5 * np.cos(40 * np.pi * t),
}


@pytest.mark.parametrize("name", sorted(CLOSED_FORMS))
def test_input_file_closed_forms_cpu(name):
    """A bare muon read from an input file: field, dissipation and local
    Zeeman blocks give the textbook polarisation through the oracle."""
    from oracle import pipeline as opipe
    from muspinsim_amd.experiment import ExperimentRunner
    with open(os.path.join(GOLDEN_INPUTS, name + ".in")) as fh:
        er = ExperimentRunner(MuSpinInput(fh))
    systems, batch = er.instance_batch()
    t = np.asarray(er.config.x_axis_values)
    assert np.abs(opipe.run(systems, batch)[0] - CLOSED_FORMS[name](t)).max() <= 1e-12


# ------------------------------------------------------------------ utils surface
def test_utils_known_answers():
    """The reference's tests/test_utils.py known answers."""
    import copy
    from muspinsim_amd.utils import (Clonable, Quaternion, deepmap, eulrange_gen, get_xy,
                                     quat_from_polar, zcw_gen)

    class TC(Clonable):
        def __init__(self):
            self.data = {"x": [1, 2, 3]}
    tc = TC()
    tc2 = tc.clone()
    tc2.data["x"][0] = 2
    assert tc.data["x"][0] == 1
    assert deepmap(lambda x: x ** 2, [[1, 2, 3], [4, 5], [6, [7, 8]]]) == \
        [[1, 4, 9], [16, 25], [36, [49, 64]]]
    z1 = quat_from_polar(np.pi / 4.0, 0).rotate([0, 0, 1])
    assert np.allclose(z1, [2 ** -0.5, 0, 2 ** -0.5])
    th, ph = 0.6 * np.pi, 0.4 * np.pi
    q2 = quat_from_polar(th, ph)
    st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    assert np.allclose(q2.rotate([0, 0, 1]), [st * cp, st * sp, ct])
    assert np.allclose(q2.conjugate().rotate([0, 0, 1]), [-st * cp, st * sp, ct])
    f = 3.0 * np.cos(zcw_gen(1000)[:, 0]) ** 2 - 1.0
    assert abs(np.average(f)) < 1e-5
    ow = eulrange_gen(10)
    A = np.array([[1, 2, 3], [2, 4, 4], [3, 4, -3]])
    Asum = sum(Quaternion.from_euler_angles(a, b, c).rotation_matrix() @ A @
               Quaternion.from_euler_angles(a, b, c).rotation_matrix().T * w
               for a, b, c, w in ow) / np.sum(ow[:, -1])
    assert np.sum(np.abs(Asum - np.eye(3) * np.trace(A) / 3)) < 1.0
    x, y = get_xy(np.array([0, 0, 1]))
    assert np.allclose(x, [1, 0, 0]) and np.allclose(y, [0, 1, 0])
    x, y = get_xy(np.array([1.0, 1.0, 0.5]))
    assert np.allclose(np.cross(x, y), np.array([1.0, 1.0, 0.5]) / 1.5)
    q = Quaternion.from_euler_angles(0.3, 1.1, -0.4)
    assert np.allclose(Quaternion.from_euler_angles(*q.euler_angles()).rotation_matrix(),
                       q.rotation_matrix())


def test_package_surface():
    import muspinsim_amd as m
    from muspinsim_amd.simconfig import MuSpinConfig as SC
    assert m.MuSpinConfig is SC and m.MuSpinInput.__name__ == "MuSpinInput"
    assert m.ExperimentRunner.__name__ == "ExperimentRunner"
    assert m.FittingRunner.__name__ == "FittingRunner" and m.constants.MU_TAU == 2.19703
    assert isinstance(m.__version__, str)
