"""Expression language of the input files (muspinsim/input/larkeval.py:9-230).

The reference parses expressions with a lark LALR grammar; lark is not part
of this engine, so this module holds a recursive-descent parser of the same
language:

    sum     := product (('+' | '-') product)*          left associative
    product := power (('*' | '/') power)*              left associative
    power   := atom ('^' atom)*                        LEFT associative: 2^3^2 = 64
    atom    := NUMBER | '-' atom | NAME | STRING | '(' sum ')'
             | NAME '(' [sum (',' sum)*] ')'

NUMBER is an integer, decimal or exponent literal (1, 1., .5, 1.27e-2) and
always evaluates to a float; NAME is [A-Za-z_][A-Za-z0-9_]*; STRING is a
double-quoted literal whose value is the text between the quotes.  Unary
minus binds tighter than '^' (-2^2 = 4, 2^-1 = 0.5), as in the grammar at
larkeval.py:24-33.  Spaces and tabs between tokens are ignored.

Errors fall in the reference's two classes: a character that starts no
token ("unexpected character") and a token that cannot continue the
expression, including a premature end ("unexpected token").  The first one
met left to right wins, as with a lazily lexed LALR parse; lark_tokenize
merges whitespace-separated words only across the second kind
(larkeval.py:45-64).
"""

import re

__all__ = ["LarkExpression", "LarkExpressionError", "lark_tokenize"]


class LarkExpressionError(Exception):
    pass


class _ParseError(Exception):
    """kind: 'token' (a valid token out of place, or the end) or 'char'."""

    def __init__(self, kind, message):
        super().__init__(message)
        self.kind = kind


_LEX = re.compile(r"""
    (?P<ws>[ \t]+)
  | (?P<num>(?:[0-9]+\.[0-9]*|\.[0-9]+|[0-9]+)(?:[eE][+-]?[0-9]+)?)
  | (?P<name>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<str>"(?:[^"\\\n]|\\.)*")
  | (?P<op>[-+*/^(),])
""", re.X)

_END = ("end", "$END", None)


def _lex(text):
    """Tokens (kind, text, column); an unlexable character becomes a final
    ('bad', char, column) token so the parser meets it in order."""
    out = []
    pos = 0
    while pos < len(text):
        m = _LEX.match(text, pos)
        if m is None:
            out.append(("bad", text[pos], pos + 1))
            return out
        if m.lastgroup != "ws":
            out.append((m.lastgroup, m.group(), pos + 1))
        pos = m.end()
    out.append(_END)
    return out


class _Parser:
    def __init__(self, text):
        self.toks = _lex(text)
        self.k = 0

    def peek(self):
        t = self.toks[self.k]
        if t[0] == "bad":
            raise _ParseError("char", f"Unexpected character '{t[1]}' at column {t[2]}")
        return t

    def fail(self, expected):
        t = self.peek()
        where = "at the end" if t[0] == "end" else f"at column {t[2]}"
        raise _ParseError("token", f"Unexpected token '{t[1]}' {where}. Expected {expected}")

    def take(self, kind, text=None):
        t = self.peek()
        if t[0] != kind or (text is not None and t[1] != text):
            self.fail(repr(text) if text else kind)
        self.k += 1
        return t

    def is_op(self, *ops):
        t = self.peek()
        return t[0] == "op" and t[1] in ops

    def parse(self):
        node = self.sum()
        if self.peek()[0] != "end":
            self.fail("an operator or the end")
        return node

    def sum(self):
        node = self.product()
        while self.is_op("+", "-"):
            op = "add" if self.take("op")[1] == "+" else "sub"
            node = (op, node, self.product())
        return node

    def product(self):
        node = self.power()
        while self.is_op("*", "/"):
            op = "mul" if self.take("op")[1] == "*" else "div"
            node = (op, node, self.power())
        return node

    def power(self):
        node = self.atom()
        while self.is_op("^"):
            self.take("op")
            node = ("pow", node, self.atom())
        return node

    def atom(self):
        t = self.peek()
        if t[0] == "num":
            self.k += 1
            return ("num", float(t[1]))
        if t[0] == "str":
            self.k += 1
            return ("str", t[1][1:-1])
        if t[0] == "name":
            self.k += 1
            if self.is_op("("):
                self.take("op")
                args = []
                if not self.is_op(")"):
                    args.append(self.sum())
                    while self.is_op(","):
                        self.take("op")
                        args.append(self.sum())
                self.take("op", ")")
                return ("fun", t[1], args)
            return ("var", t[1])
        if t[0] == "op" and t[1] == "-":
            self.k += 1
            return ("neg", self.atom())
        if t[0] == "op" and t[1] == "(":
            self.k += 1
            node = self.sum()
            self.take("op", ")")
            return node
        self.fail("a number, name, string, '-' or '('")


def _parse(text):
    return _Parser(text).parse()


def _collect(node, names, calls):
    kind = node[0]
    if kind == "var":
        names.add(node[1])
    elif kind == "fun":
        calls.add(node[1])
        for a in node[2]:
            _collect(a, names, calls)
    elif kind == "neg":
        _collect(node[1], names, calls)
    elif kind in ("add", "sub", "mul", "div", "pow"):
        _collect(node[1], names, calls)
        _collect(node[2], names, calls)


_BINARY = {
    "add": lambda a, b: a + b,
    "sub": lambda a, b: a - b,
    "mul": lambda a, b: a * b,
    "div": lambda a, b: a / b,
    "pow": lambda a, b: a ** b,
}


def lark_tokenize(line):
    """Split a line into the expressions it holds: whitespace-separated words,
    joined (without the space) while the text so far stops early or has a
    token out of place, e.g. 'atan2(3, 4)' is one expression
    (larkeval.py:45-64).  An unknown character raises LarkExpressionError; a
    line that ends inside an expression raises RuntimeError."""
    words = line.split()
    exprs = []
    k = 0
    while k < len(words):
        text = words[k]
        k += 1
        while True:
            try:
                _parse(text)
                break
            except _ParseError as exc:
                if exc.kind == "char":
                    raise LarkExpressionError(f"Could not parse input: {exc}") from None
                if k == len(words):
                    raise RuntimeError(
                        f"Line can not be tokenized into valid expressions: {exc}") from None
                text += words[k]
                k += 1
        exprs.append(text)
    return exprs


_UNSET = object()


class LarkExpression:
    def __init__(self, source, variables=None, functions=None):
        """Parse `source`, accepting only the names in `variables` (a name or a
        list of names) and the functions in `functions` {name: callable}
        (larkeval.py:72-138).  An expression without variables is evaluated
        at once."""
        if variables is None:
            variables = []
        if functions is None:
            functions = {}
        self._source = source
        if source == "":
            raise LarkExpressionError("Empty String")
        try:
            self._tree = _parse(source)
        except _ParseError as exc:
            raise LarkExpressionError(f"Invalid characters in LarkExpression: {exc}") from None
        names, calls = set(), set()
        _collect(self._tree, names, calls)
        self._variables = names
        self._functions = calls
        self._all_variables = {variables} if isinstance(variables, str) else set(variables)
        for v in sorted(names - self._all_variables):
            raise LarkExpressionError(
                "Invalid variable/constant: '{0}', valid variables/constants are ['{1}']".format(
                    v, "', '".join(sorted(self._all_variables))))
        for f in sorted(calls - set(functions)):
            raise LarkExpressionError(
                "Invalid function: '{0}()', valid functions are ['{1}()']".format(
                    f, "()', '".join(functions.keys())))
        self._function_bodies = {f: functions[f] for f in calls}
        self._store_eval = _UNSET
        if not names:
            self._store_eval = self._eval(self._tree, {})

    def _eval(self, node, values):
        kind = node[0]
        if kind in ("num", "str"):
            return node[1]
        if kind == "var":
            return values[node[1]]
        if kind == "neg":
            return -self._eval(node[1], values)
        if kind == "fun":
            return self._function_bodies[node[1]](*[self._eval(a, values) for a in node[2]])
        return _BINARY[kind](self._eval(node[1], values), self._eval(node[2], values))

    @property
    def source(self):
        return self._source

    @property
    def functions(self):
        return self._functions

    @property
    def variables(self):
        return self._variables

    def evaluate(self, **variables):
        """Value of the expression; every variable it uses must be given and
        every given name must be an admissible variable (larkeval.py:205-230)."""
        given = set(variables)
        missing = self._variables - given
        if missing:
            raise LarkExpressionError(
                f"Some necessary variable(s) {missing} have not been defined when "
                "evaluating LarkExpression")
        extra = given - self._all_variables
        if extra:
            raise LarkExpressionError(
                f"Some invalid variable(s) {extra} have been defined when evaluating "
                "LarkExpression")
        if self._store_eval is not _UNSET:
            return self._store_eval
        return self._eval(self._tree, variables)

    def __repr__(self):
        return f"LarkExpression({self._source!r})"
