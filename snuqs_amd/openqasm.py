"""OpenQASM 3 programs -> gate operations (src/braket/snuqs/openqasm/ of the
reference).

The reference interprets OpenQASM through an ANTLR-generated parser
(openqasm/parser/, interpreter.py, program_context.py); this module is a
hand-written recursive-descent reader of the subset that Braket circuits are
serialised to, producing the engine's gate objects directly:

* declarations: ``qubit[n] q;``, ``qubit q;``, ``qreg q[n];``, physical qubits
  ``$k``, ``bit[n] b;`` / ``creg``, ``input <type> name;`` (values from the
  program's ``inputs``), ``const``/classical scalars with an initialiser;
* gate calls with broadcasting over registers (interpreter.py:364-400), the
  Braket gate names (gate_operations.py:202-239; ``cnot``, ``U``,
  ``gphase``) plus ``cv``, ``ecr``, ``gpi``, ``gpi2``, ``ms``, ``prx``;
* modifiers ``ctrl @``, ``ctrl(k) @``, ``negctrl @``, ``inv @``, ``pow(x) @``
  in any combination, on built-in and user-defined gates;
* ``gate name(params) a, b { ... }`` definitions, expanded at the call site;
* ``gphase(x);`` and controlled phases (``ctrl @ gphase(x) a``);
* ``#pragma braket unitary(<matrix>) qubits``; ``#pragma braket result``
  state_vector / probability / amplitude / expectation / variance of tensor
  products of ``x y z h i`` and ``hermitian(<matrix>)`` factors (``all`` or
  per-factor qubits), and
  ``adjoint_gradient expectation(<sum of c * products>) <inputs | all>``:
  inputs are carried as dual numbers through the angle expressions, so each
  gate records d angle / d input and the engine's adjoint sweep gives the
  gradient by the chain rule;
* ``measure`` in its assignment, bare and ``->`` forms; ``barrier`` and
  ``box { }`` (verbatim blocks) are transparent.

Modifier semantics follow the OpenQASM 3 definition rather than the
reference's shortcuts: ``pow(k) @ g`` for a user-defined ``g`` repeats (k
integer) or takes the principal power of (k fractional) the whole body, where
the reference raises each body statement to the power separately
(_helpers/quantum.py:93-109; wrong when the statements do not commute), and
``negctrl @ gphase`` is the phase on the |0> branch, which the reference
rejects (_helpers/quantum.py:57-62).  Integer powers, inverses and controls of
built-in gates become the gate classes' ``power`` / ``ctrl_modifiers`` exactly
as program_context.py:745-776 passes them; fractional powers are computed here
(principal branch, eigenphases in (-pi, pi]) and applied as ``Unitary``.

Classical control flow, subroutines, ``reset``, noise and mid-circuit
measurement feedback are outside the state-vector path and raise
``NotImplementedError``.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Any, Optional

import numpy as np

from ._C.operation import gate_operations as g
from .gate_operations import BRAKET_GATES

_TOKEN = re.compile(r"""
    (?P<ws>\s+)
  | (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?(?:im)?)
  | (?P<phys>\$\d+)
  | (?P<str>"[^"]*")
  | (?P<id>[A-Za-z_\u0080-\uffff][A-Za-z0-9_\u0080-\uffff]*)
  | (?P<op>\*\*|->|[@()\[\]{},;:=+\-*/%])
""", re.X)

_CONSTANTS = {"pi": math.pi, "π": math.pi, "tau": 2 * math.pi, "τ": 2 * math.pi,
              "euler": math.e, "ℇ": math.e}
_FUNCTIONS = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "arcsin": np.arcsin,
              "arccos": np.arccos, "arctan": np.arctan, "exp": np.exp, "log": np.log,
              "sqrt": np.sqrt, "abs": abs, "ceiling": math.ceil, "floor": math.floor,
              "mod": lambda a, b: a % b, "pow": lambda a, b: a ** b}
_TYPES = {"float", "int", "uint", "angle", "bool", "complex", "bit", "duration"}

# Gates outside the reference's table (gate_operations.py:238 lists CV, ECR and
# PRx as removed): matrices of the Braket SDK's definitions.
_S2 = 1 / math.sqrt(2)


def _gpi(phi):
    return np.array([[0, np.exp(-1j * phi)], [np.exp(1j * phi), 0]])


def _gpi2(phi):
    return _S2 * np.array([[1, -1j * np.exp(-1j * phi)], [-1j * np.exp(1j * phi), 1]])


def _prx(theta, phi):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -1j * np.exp(-1j * phi) * s], [-1j * np.exp(1j * phi) * s, c]])


def _ms(phi0, phi1, theta=math.pi / 2):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    e = lambda x: np.exp(1j * x)  # noqa: E731
    return np.array([[c, 0, 0, -1j * e(-(phi0 + phi1)) * s],
                     [0, c, -1j * e(-(phi0 - phi1)) * s, 0],
                     [0, -1j * e(phi0 - phi1) * s, c, 0],
                     [-1j * e(phi0 + phi1) * s, 0, 0, c]])


_ECR = _S2 * np.array([[0, 0, 1, 1j], [0, 0, 1j, 1], [1, -1j, 0, 0], [-1j, 1, 0, 0]])
_MATRIX_GATES = {  # name -> (qubits, accepted parameter counts, matrix function)
    "gpi": (1, (1,), _gpi), "gpi2": (1, (1,), _gpi2), "prx": (1, (2,), _prx),
    "ms": (2, (2, 3), _ms), "ecr": (2, (0,), lambda: _ECR),
}
_NAMED = {k: v for k, v in BRAKET_GATES.items() if k not in ("unitary", "gphase")}
# one-parameter gates exp(-i t A / 2)-style, for which G(t)^k = G(k t)
_EXP_FAMILY = {"rx", "ry", "rz", "phaseshift", "cphaseshift", "cphaseshift00", "cphaseshift01",
               "cphaseshift10", "xx", "yy", "zz", "xy"}


def _fraction(p) -> Fraction:
    if isinstance(p, Dual):
        p = p.v
    return Fraction(p) if isinstance(p, (int, Fraction)) else Fraction(float(p))


class Dual:
    """A real value with its partial derivatives with respect to the program's
    inputs (forward-mode differentiation of the angle expressions)."""
    __slots__ = ("v", "d")

    def __init__(self, v, d=None):
        self.v, self.d = float(v), dict(d or {})

    @staticmethod
    def _lift(x):
        return x if isinstance(x, Dual) else Dual(x)

    def _comb(self, o, dv_self, dv_o, v):
        o = Dual._lift(o)
        d = {k: dv_self * x for k, x in self.d.items()}
        for k, x in o.d.items():
            d[k] = d.get(k, 0.0) + dv_o * x
        return Dual(v, d)

    def __add__(self, o):
        return self._comb(o, 1.0, 1.0, self.v + float(Dual._lift(o).v))

    __radd__ = __add__

    def __sub__(self, o):
        return self._comb(o, 1.0, -1.0, self.v - Dual._lift(o).v)

    def __rsub__(self, o):
        return Dual._lift(o) - self

    def __mul__(self, o):
        o = Dual._lift(o)
        return self._comb(o, o.v, self.v, self.v * o.v)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = Dual._lift(o)
        return self._comb(o, 1.0 / o.v, -self.v / (o.v * o.v), self.v / o.v)

    def __rtruediv__(self, o):
        return Dual._lift(o) / self

    def __pow__(self, o):
        o = Dual._lift(o)
        v = self.v ** o.v
        if self.v != 0.0:
            dbase = o.v * self.v ** (o.v - 1)
        elif o.v > 1.0 or o.v == 0.0:
            dbase = 0.0  # d/dx x^e at x = 0: 0 for e > 1; x^0 is constant
        elif o.v == 1.0:
            dbase = 1.0
        else:
            dbase = math.inf  # e < 1: the slope at 0 is unbounded
        dlog = math.log(self.v) if self.v > 0 else 0.0
        return self._comb(o, dbase, v * dlog, v)

    def __rpow__(self, o):
        return Dual._lift(o) ** self

    def __neg__(self):
        return Dual(-self.v, {k: -x for k, x in self.d.items()})

    def __mod__(self, o):
        # a % b = a - b floor(a/b) (Python's floored remainder), so
        # d(a % b) = da - floor(a/b) db away from the jumps
        o = Dual._lift(o)
        return self._comb(o, 1.0, -math.floor(self.v / o.v), self.v % o.v)

    def __rmod__(self, o):
        return Dual._lift(o) % self

    def __float__(self):
        return self.v

    def __int__(self):
        return int(self.v)


_DERIV = {"sin": math.cos, "cos": lambda x: -math.sin(x), "tan": lambda x: 1 / math.cos(x) ** 2,
          "arcsin": lambda x: 1 / math.sqrt(1 - x * x), "arccos": lambda x: -1 / math.sqrt(1 - x * x),
          "arctan": lambda x: 1 / (1 + x * x), "exp": math.exp, "log": lambda x: 1 / x,
          "sqrt": lambda x: 0.5 / math.sqrt(x)}


def _apply_function(name, args):
    f = _FUNCTIONS[name]
    if not any(isinstance(a, Dual) for a in args):
        return f(*[float(a) for a in args])
    if name == "pow" and len(args) == 2:
        return Dual._lift(args[0]) ** args[1]
    if name == "mod" and len(args) == 2:
        return Dual._lift(args[0]) % args[1]
    if len(args) != 1:
        raise ValueError(f"{name} takes one argument")
    x = Dual._lift(args[0])
    if name == "abs":
        d = math.copysign(1.0, x.v) if x.v else 0.0
    elif name in ("ceiling", "floor"):
        d = 0.0  # piecewise constant
    else:
        d = _DERIV[name](x.v)
    return Dual(f(x.v), {k: d * dx for k, dx in x.d.items()})


def unitary_power(m: np.ndarray, p: float) -> np.ndarray:
    """Principal power of a unitary: eigenphases taken in (-pi, pi]."""
    from scipy.linalg import schur

    t, z = schur(np.asarray(m, dtype=np.complex128), output="complex")
    ph = np.angle(np.diag(t))
    ph[np.abs(ph + math.pi) < 1e-12] = math.pi
    return (z * np.exp(1j * p * ph)) @ z.conj().T


class QasmError(ValueError):
    """Malformed or unsupported OpenQASM text."""


@dataclass
class Circuit:
    """Result of reading a program: gates, result types and measured qubits
    (the fields openqasm/circuit.py:38-92 of the reference keeps)."""
    instructions: list = field(default_factory=list)
    results: list = field(default_factory=list)
    measured_qubits: list = field(default_factory=list)
    qubit_set: set = field(default_factory=set)
    # per instruction: [(angle index, {input: d angle / d input})], or a set of
    # input names for a gate whose dependence on them is not differentiable
    # here (fractional powers, matrix gates)
    input_deps: list = field(default_factory=list)
    inputs: list = field(default_factory=list)

    def add_instruction(self, op, deps=None) -> None:
        self.instructions.append(op)
        self.input_deps.append(deps if deps is not None else [])
        self.qubit_set |= set(op.targets)


# --------------------------------------------------------------------------
# Tokens and expressions
# --------------------------------------------------------------------------
def _strip_comments(src: str) -> str:
    src = re.sub(r"/\*.*?\*/", lambda m: "\n" * m.group(0).count("\n"), src, flags=re.S)
    return re.sub(r"//[^\n]*", "", src)


def _tokenize(text: str, line: int) -> list:
    out, pos = [], 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise QasmError(f"line {line}: unexpected character {text[pos]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind == "ws":
            line += m.group(0).count("\n")
            continue
        out.append((kind, m.group(0), line))
    return out


class _Tokens:
    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self, k=0):
        j = self.i + k
        return self.t[j] if j < len(self.t) else ("eof", "", self.t[-1][2] if self.t else 0)

    def next(self):
        tok = self.peek()
        self.i += 1
        return tok

    def at(self, value, k=0) -> bool:
        return self.peek(k)[1] == value and self.peek(k)[0] in ("op", "id")

    def accept(self, value) -> bool:
        if self.at(value):
            self.i += 1
            return True
        return False

    def expect(self, value):
        tok = self.next()
        if tok[1] != value:
            raise QasmError(f"line {tok[2]}: expected {value!r}, found {tok[1]!r}")
        return tok

    def ident(self) -> str:
        tok = self.next()
        if tok[0] != "id":
            raise QasmError(f"line {tok[2]}: expected a name, found {tok[1]!r}")
        return tok[1]

    def done(self) -> bool:
        return self.i >= len(self.t)


def _parse_expr(ts: _Tokens):
    return _parse_sum(ts)


def _parse_sum(ts):
    e = _parse_product(ts)
    while ts.at("+") or ts.at("-"):
        op = ts.next()[1]
        e = ("bin", op, e, _parse_product(ts))
    return e


def _parse_product(ts):
    e = _parse_unary(ts)
    while ts.at("*") or ts.at("/") or ts.at("%"):
        op = ts.next()[1]
        e = ("bin", op, e, _parse_unary(ts))
    return e


def _parse_unary(ts):
    if ts.accept("-"):
        return ("neg", _parse_unary(ts))
    if ts.accept("+"):
        return _parse_unary(ts)
    return _parse_power(ts)


def _parse_power(ts):
    base = _parse_atom(ts)
    if ts.accept("**"):
        return ("bin", "**", base, _parse_unary(ts))  # right-associative
    return base


def _parse_atom(ts):
    kind, text, line = ts.next()
    if kind == "num":
        if text.endswith("im"):
            return ("num", complex(0, float(text[:-2])))
        v = float(text)
        return ("num", int(text) if re.fullmatch(r"\d+", text) else v)
    if kind == "id":
        if ts.at("("):
            ts.next()
            args = []
            if not ts.at(")"):
                args.append(_parse_expr(ts))
                while ts.accept(","):
                    args.append(_parse_expr(ts))
            ts.expect(")")
            return ("call", text, args, line)
        return ("id", text, line)
    if text == "(":
        e = _parse_expr(ts)
        ts.expect(")")
        return e
    if text == "[":
        items = []
        if not ts.at("]"):
            items.append(_parse_expr(ts))
            while ts.accept(","):
                items.append(_parse_expr(ts))
        ts.expect("]")
        return ("list", items)
    raise QasmError(f"line {line}: unexpected {text!r} in an expression")


def _eval(e, env):
    k = e[0]
    if k == "num":
        return e[1]
    if k == "id":
        if e[1] in env:
            return env[e[1]]
        if e[1] in _CONSTANTS:
            return _CONSTANTS[e[1]]
        raise QasmError(f"line {e[2]}: {e[1]!r} is not defined")
    if k == "neg":
        return -_eval(e[1], env)
    if k == "bin":
        a, b = _eval(e[2], env), _eval(e[3], env)
        op = e[1]
        if op == "+":
            return a + b
        if op == "-":
            return a - b
        if op == "*":
            return a * b
        if isinstance(a, Fraction) and isinstance(b, Dual):
            a = float(a)
        if isinstance(b, Fraction) and isinstance(a, Dual):
            b = float(b)
        if op == "/":
            if isinstance(a, int) and isinstance(b, int) and not isinstance(a, bool):
                return Fraction(a, b) if a % b else a // b
            return a / b
        if op == "%":
            return a % b
        return a ** b
    if k == "call":
        if e[1] not in _FUNCTIONS:
            raise QasmError(f"line {e[3]}: unknown function {e[1]!r}")
        return _apply_function(e[1], [_eval(a, env) for a in e[2]])
    if k == "list":
        return [_eval(x, env) for x in e[1]]
    raise QasmError(f"bad expression {e!r}")


def _num(v) -> float:
    if isinstance(v, Dual):
        return v.v
    if isinstance(v, complex):
        if v.imag:
            raise QasmError(f"a real number is needed, got {v}")
        v = v.real
    return float(v)


# --------------------------------------------------------------------------
# Statements
# --------------------------------------------------------------------------
def _skip_type(ts: _Tokens) -> None:
    ts.next()
    if ts.at("["):
        depth = 0
        while True:
            tok = ts.next()
            if tok[0] == "eof":
                raise QasmError("unterminated type designator")
            depth += tok[1] == "["
            depth -= tok[1] == "]"
            if depth == 0:
                break


def _parse_qarg(ts: _Tokens):
    kind, text, line = ts.next()
    if kind == "phys":
        return ("phys", int(text[1:]), line)
    if kind != "id":
        raise QasmError(f"line {line}: expected a qubit, found {text!r}")
    if ts.accept("["):
        lo = _parse_expr(ts)
        if ts.accept(":"):
            hi = _parse_expr(ts)
            ts.expect("]")
            return ("slice", text, lo, hi, line)
        ts.expect("]")
        return ("index", text, lo, line)
    return ("name", text, line)


def _parse_qargs(ts: _Tokens, end=(";",)) -> list:
    out = []
    if ts.done() or any(ts.at(x) for x in end):
        return out
    out.append(_parse_qarg(ts))
    while ts.accept(","):
        out.append(_parse_qarg(ts))
    return out


def _parse_modifiers(ts: _Tokens) -> list:
    mods = []
    while True:
        name = ts.peek()[1]
        if ts.peek()[0] == "id" and name in ("ctrl", "negctrl", "inv", "pow"):
            # a modifier is followed by '@' (possibly after an argument)
            j = 1
            if ts.at("(", 1):
                depth = 0
                while True:
                    tok = ts.peek(j)
                    if tok[0] == "eof":
                        break
                    depth += tok[1] == "("
                    depth -= tok[1] == ")"
                    j += 1
                    if depth == 0:
                        break
            if not ts.at("@", j):
                return mods
            ts.next()
            arg = None
            if ts.accept("("):
                arg = _parse_expr(ts)
                ts.expect(")")
            ts.expect("@")
            mods.append((name, arg))
        else:
            return mods


def _parse_statements(ts: _Tokens, in_body=False) -> list:
    out = []
    while not ts.done():
        if in_body and ts.at("}"):
            return out
        out.append(_parse_statement(ts, in_body))
    if in_body:
        raise QasmError("missing '}'")
    return out


def _parse_block(ts: _Tokens) -> list:
    ts.expect("{")
    body = _parse_statements(ts, in_body=True)
    ts.expect("}")
    return body


def _parse_statement(ts: _Tokens, in_body: bool):
    kind, text, line = ts.peek()
    if kind == "pragma":
        ts.next()
        return ("pragma", text, line)
    if kind != "id":
        raise QasmError(f"line {line}: unexpected {text!r}")
    if text == "OPENQASM":
        ts.next()
        while not ts.accept(";"):
            ts.next()
        return ("nop",)
    if text == "include":
        ts.next()
        ts.next()
        ts.expect(";")
        return ("nop",)
    if text in ("qubit", "qreg"):
        ts.next()
        size = None
        if text == "qubit" and ts.accept("["):
            size = _parse_expr(ts)
            ts.expect("]")
        name = ts.ident()
        if text == "qreg" and ts.accept("["):
            size = _parse_expr(ts)
            ts.expect("]")
        ts.expect(";")
        return ("qubit", name, size, line)
    if text in ("bit", "creg"):
        ts.next()
        if text == "bit" and ts.at("["):
            ts.next()
            _parse_expr(ts)
            ts.expect("]")
        name = ts.ident()
        if text == "creg" and ts.accept("["):
            _parse_expr(ts)
            ts.expect("]")
        if ts.accept("="):
            ts.expect("measure")
            q = _parse_qargs(ts)
            ts.expect(";")
            return ("measure", q, line)
        ts.expect(";")
        return ("nop",)
    if text == "input":
        ts.next()
        _skip_type(ts)
        name = ts.ident()
        ts.expect(";")
        return ("input", name, line)
    if text == "const" or (text in _TYPES and (ts.at("[", 1) or ts.peek(1)[0] == "id")):
        if text == "const":
            ts.next()
        _skip_type(ts)
        name = ts.ident()
        if not ts.accept("="):
            raise QasmError(f"line {line}: classical variable {name!r} needs an initial value")
        e = _parse_expr(ts)
        ts.expect(";")
        return ("const", name, e, line)
    if text == "gate":
        ts.next()
        name = ts.ident()
        params = []
        if ts.accept("("):
            if not ts.at(")"):
                params.append(ts.ident())
                while ts.accept(","):
                    params.append(ts.ident())
            ts.expect(")")
        qargs = [ts.ident()]
        while ts.accept(","):
            qargs.append(ts.ident())
        body = _parse_block(ts)
        return ("gatedef", name, params, qargs, body, line)
    if text == "measure":
        ts.next()
        q = _parse_qargs(ts, end=(";", "->"))
        if ts.accept("->"):
            _parse_qarg(ts)
        ts.expect(";")
        return ("measure", q, line)
    if text == "barrier":
        ts.next()
        while not ts.accept(";"):
            ts.next()
        return ("nop",)
    if text == "box":
        ts.next()
        if ts.at("["):
            while not ts.accept("]"):
                ts.next()
        return ("box", _parse_block(ts))
    if ts.at("=", 1) or (ts.at("[", 1) and _assignment_ahead(ts)):
        # b = measure q;  b[0] = measure q[0];
        ts.next()
        if ts.accept("["):
            while not ts.accept("]"):
                if ts.next()[0] == "eof":
                    raise QasmError(f"line {line}: missing ']'")
        ts.expect("=")
        if not ts.accept("measure"):
            raise NotImplementedError(f"line {line}: classical assignments are not supported")
        q = _parse_qargs(ts)
        ts.expect(";")
        return ("measure", q, line)
    if text in ("reset", "if", "for", "while", "def", "return", "defcal", "cal", "delay",
                "switch", "let", "extern", "end"):
        raise NotImplementedError(f"line {line}: '{text}' is outside the state-vector subset "
                                  "this engine reads")
    mods = _parse_modifiers(ts)
    name = ts.ident()
    args = []
    if ts.accept("("):
        if not ts.at(")"):
            args.append(_parse_expr(ts))
            while ts.accept(","):
                args.append(_parse_expr(ts))
        ts.expect(")")
    qargs = _parse_qargs(ts)
    ts.expect(";")
    return ("call", mods, name, args, qargs, line)


def _assignment_ahead(ts: _Tokens) -> bool:
    j, depth = 1, 0
    while True:
        tok = ts.peek(j)
        if tok[0] == "eof" or tok[1] == ";":
            return False
        depth += tok[1] == "["
        depth -= tok[1] == "]"
        j += 1
        if depth == 0:
            return ts.at("=", j)


# --------------------------------------------------------------------------
# Interpreter
# --------------------------------------------------------------------------
class _Interpreter:
    def __init__(self, inputs: Optional[dict]):
        self.inputs = dict(inputs or {})
        self.env: dict[str, Any] = {}
        self.regs: dict[str, tuple[int, int]] = {}
        self.nalloc = 0
        self.gates: dict[str, tuple] = {}
        self.circuit = Circuit()
        self.physical = False

    # -- qubits ---------------------------------------------------------
    def _resolve(self, qa, scope) -> list:
        kind = qa[0]
        if kind == "phys":
            if self.regs:
                raise QasmError(f"line {qa[2]}: physical and declared qubits cannot be mixed")
            self.physical = True
            return [qa[1]]
        name, line = qa[1], qa[-1]
        if scope is not None:
            if kind != "name" or name not in scope:
                raise QasmError(f"line {line}: {name!r} is not a qubit argument of this gate")
            return [scope[name]]
        if name not in self.regs:
            raise QasmError(f"line {line}: qubit register {name!r} is not declared")
        start, size = self.regs[name]
        if kind == "name":
            return list(range(start, start + size))
        if kind == "index":
            i = int(_eval(qa[2], self.env))
            if i < 0:
                i += size
            if not 0 <= i < size:
                raise QasmError(f"line {line}: index {i} is out of range for {name}[{size}]")
            return [start + i]
        lo, hi = int(_eval(qa[2], self.env)), int(_eval(qa[3], self.env))
        if not (0 <= lo <= hi < size):
            raise QasmError(f"line {line}: range {lo}:{hi} is out of range for {name}[{size}]")
        return list(range(start + lo, start + hi + 1))

    # -- statements -----------------------------------------------------
    def run(self, stmts) -> None:
        for s in stmts:
            self.statement(s)

    def statement(self, s) -> None:
        k = s[0]
        if k == "nop":
            return
        if k == "qubit":
            if self.physical:
                raise QasmError(f"line {s[3]}: physical and declared qubits cannot be mixed")
            size = 1 if s[2] is None else int(_eval(s[2], self.env))
            if s[1] in self.regs:
                raise QasmError(f"line {s[3]}: qubit register {s[1]!r} is declared twice")
            self.regs[s[1]] = (self.nalloc, size)
            self.nalloc += size
        elif k == "input":
            if s[1] not in self.inputs:
                raise QasmError(f"line {s[2]}: no value was given for input {s[1]!r}")
            self.env[s[1]] = Dual(self.inputs[s[1]], {s[1]: 1.0})
            self.circuit.inputs.append(s[1])
        elif k == "const":
            self.env[s[1]] = _eval(s[2], self.env)
        elif k == "gatedef":
            _, name, params, qargs, body, line = s
            if name in self.gates:
                raise QasmError(f"line {line}: gate {name!r} is defined twice")
            self.gates[name] = (params, qargs, body)
        elif k == "measure":
            for qa in s[1]:
                for q in self._resolve(qa, None):
                    if q in self.circuit.measured_qubits:
                        raise ValueError(f"Qubit {q} is already measured or captured.")
                    self.circuit.measured_qubits.append(q)
                    self.circuit.qubit_set.add(q)
        elif k == "box":
            self.run(s[1])
        elif k == "pragma":
            self.pragma(s[1], s[2])
        elif k == "call":
            self.call(s, None, {})
        else:
            raise QasmError(f"unsupported statement {k}")

    def call(self, s, scope, env) -> None:
        _, mods, name, args, qargs, line = s
        ev = {**self.env, **env}
        params = [_eval(a, ev) for a in args]
        groups = [self._resolve(qa, scope) for qa in qargs]
        lens = {len(q) for q in groups if len(q) > 1}
        if len(lens) > 1:
            raise ValueError("Qubit registers must all be the same length.")
        reps = lens.pop() if lens else 1
        for i in range(reps):
            qubits = [q[i] if len(q) > 1 else q[0] for q in groups]
            self.apply(name, params, qubits, mods, ev, line)

    def apply(self, name, params, qubits, mods, ev, line) -> None:
        """One gate with modifiers; controls take the leading qubits in
        modifier order (program_context.py:765-775)."""
        ctrl, power = [], Fraction(1)
        for mname, arg in mods:
            if mname in ("ctrl", "negctrl"):
                n = 1 if arg is None else int(_eval(arg, ev))
                ctrl += [1 if mname == "ctrl" else 0] * n
            elif mname == "inv":
                power = -power
            else:
                power *= _fraction(_eval(arg, ev))
        if len(set(qubits)) != len(qubits):
            raise ValueError(f"line {line}: gate {name} is applied to repeated qubits {qubits}")
        nc = len(ctrl)
        if nc > len(qubits):
            raise ValueError(f"line {line}: {nc} control(s) need more than {len(qubits)} qubit(s)")
        cq, tq = qubits[:nc], qubits[nc:]
        self.emit(name, list(params), cq, ctrl, tq, power, line)

    def emit(self, name, raw, cq, ctrl, tq, power: Fraction, line) -> None:
        """raw: parameter values, Dual where they depend on inputs."""
        if power == 0:
            return
        integer = power.denominator == 1
        params = [_num(p) for p in raw]
        deps = [p.d if isinstance(p, Dual) else {} for p in raw]
        used = {k for d in deps for k in d}
        if name in self.gates:
            fparams, fq, body = self.gates[name]
            if len(params) != len(fparams) or len(tq) != len(fq):
                raise ValueError(f"line {line}: gate {name} takes {len(fparams)} parameter(s) "
                                 f"and {len(fq)} qubit(s)")
            if integer:
                seq = self._expand(name, raw, tq)
                if power < 0:
                    seq = [(n, p, q, c, -pw) for (n, p, q, c, pw) in reversed(seq)]
                for _ in range(abs(int(power))):
                    for (n, p, q, c, pw) in seq:
                        self.emit(n, p, cq + q[:len(c)], ctrl + c, q[len(c):], pw, line)
            else:
                m = self._body_matrix(name, params, tq)
                self._unitary(cq + tq, unitary_power(m, float(power)), ctrl, used)
            return
        if name == "gphase":  # on qubits without controls it is still a global phase
            theta = params[0] if len(params) == 1 else None
            if theta is None:
                raise ValueError(f"line {line}: gphase takes one parameter")
            if integer:
                phase = theta * int(power)
            else:
                phase = float(np.angle(unitary_power(np.array([[np.exp(1j * theta)]]),
                                                     float(power))[0, 0]))
            if ctrl:
                d = [(0, {k: x * int(power) for k, x in deps[0].items()})] if integer else used
                self._add(g.GPhase(cq, phase, ctrl_modifiers=ctrl), d if deps[0] else None)
            else:
                self._add(g.GPhase([], phase))  # global: no effect on expectations
            return
        if name == "cv":
            self.emit("v", raw, cq + tq[:1], ctrl + [1], tq[1:], power, line)
            return
        if name in _MATRIX_GATES:
            nq, npar, fn = _MATRIX_GATES[name]
            if len(tq) != nq or len(params) not in npar:
                raise ValueError(f"line {line}: gate {name} takes {nq} qubit(s) and "
                                 f"{' or '.join(map(str, npar))} parameter(s)")
            m = fn(*params)
            self._unitary(cq + tq, m if power == 1 else unitary_power(m, float(power)), ctrl,
                          used)
            return
        cls = _NAMED.get(name)
        if cls is None:
            raise QasmError(f"line {line}: gate {name!r} is not defined")
        try:
            if integer and power != 1 and name in _EXP_FAMILY:
                # G(t)^k = G(k t): keeps the gate differentiable
                k = int(power)
                params = [k * x for x in params]
                deps = [{n: k * x for n, x in d.items()} for d in deps]
                power = Fraction(1)
            if integer:
                op = cls(cq + tq, *params, ctrl_modifiers=ctrl, power=int(power))
                if used and power != 1:
                    self._add(op, used)
                else:
                    self._add(op, [(a, d) for a, d in enumerate(deps) if d])
            else:
                base = np.array(cls(list(range(len(tq))), *params).matrix())
                self._unitary(cq + tq, unitary_power(base, float(power)), ctrl, used)
        except (TypeError, ValueError) as e:
            raise ValueError(f"line {line}: {name}: {e}") from None

    def _expand(self, name, params, tq) -> list:
        """Body of a user gate as (name, params, qubits, ctrl, power) tuples,
        nested definitions expanded, controls leading in `qubits`."""
        fparams, fq, body = self.gates[name]
        env = dict(zip(fparams, params))
        scope = dict(zip(fq, tq))
        out = []
        rec = _Recorder(self, out)
        for s in body:
            if s[0] != "call":
                raise NotImplementedError(f"only gate calls may appear in the body of {name}")
            rec.call(s, scope, env)
        return out

    def _body_matrix(self, name, params, tq) -> np.ndarray:
        from .transpile import _apply_to_block

        local = list(range(len(tq)))
        seq = self._expand(name, params, local)
        sub = _Interpreter(None)
        sub.gates = self.gates
        for (n, p, q, c, pw) in seq:
            sub.emit(n, p, q[:len(c)], c, q[len(c):], pw, 0)
        m = np.eye(1 << len(tq), dtype=np.complex128)
        for op in sub.circuit.instructions:
            t = list(op.targets)
            om = np.array(op.matrix())
            m = om[0, 0] * m if not t else _apply_to_block(m, local, om, t)
        return m

    def _unitary(self, qubits, m, ctrl, used=None) -> None:
        self._add(g.Unitary(list(qubits), np.asarray(m, dtype=np.complex128),
                            ctrl_modifiers=list(ctrl)), set(used) if used else None)

    def _add(self, op, deps=None) -> None:
        self.circuit.add_instruction(op, deps)

    # -- pragmas --------------------------------------------------------
    def pragma(self, text: str, line: int) -> None:
        body = text.strip()[len("#pragma"):].strip()
        if not body.startswith("braket"):
            raise NotImplementedError(f"line {line}: unknown pragma {text.strip()!r}")
        body = body[len("braket"):].strip()
        ts = _Tokens(_tokenize(body, line))
        word = ts.ident()
        if word == "verbatim":
            return
        if word == "unitary":
            ts.expect("(")
            m = np.array(_eval(_parse_expr(ts), self.env), dtype=np.complex128)
            ts.expect(")")
            qubits = [q for qa in _parse_qargs(ts, end=("eof",)) for q in self._resolve(qa, None)]
            if m.ndim != 2 or m.shape != (1 << len(qubits), 1 << len(qubits)):
                raise ValueError(f"line {line}: unitary matrix of shape {m.shape} does not fit "
                                 f"{len(qubits)} qubit(s)")
            if not np.allclose(m @ m.conj().T, np.eye(m.shape[0]), atol=1e-7):
                raise ValueError(f"line {line}: the matrix is not unitary")
            self._unitary(qubits, m, [])
            return
        if word == "result":
            self.circuit.results.append(self._result(ts, line))
            return
        if word == "noise":
            raise NotImplementedError(f"line {line}: noise channels need a density-matrix "
                                      "simulator")
        raise NotImplementedError(f"line {line}: unknown braket pragma {word!r}")

    def _result(self, ts: _Tokens, line: int) -> dict:
        kind = ts.ident()
        if kind == "state_vector":
            return {"type": "statevector"}
        if kind == "probability":
            if ts.done() or ts.at("all"):
                return {"type": "probability"}
            q = [x for qa in _parse_qargs(ts, end=("eof",)) for x in self._resolve(qa, None)]
            self.circuit.qubit_set |= set(q)
            return {"type": "probability", "targets": q}
        if kind == "amplitude":
            states = []
            while not ts.done():
                tok = ts.next()
                if tok[0] == "str":
                    states.append(tok[1].strip('"'))
                elif tok[1] != ",":
                    raise QasmError(f"line {line}: amplitude takes quoted bitstrings")
            return {"type": "amplitude", "states": states}
        if kind == "adjoint_gradient":
            return self._adjoint_result(ts, line)
        if kind in ("expectation", "variance", "sample"):
            letters, targets = [], []
            while True:
                obs = ts.ident().lower()
                if obs == "hermitian":
                    ts.expect("(")
                    m = hermitian_matrix(_eval(_parse_expr(ts), self.env))
                    ts.expect(")")
                    q = [x for qa in _parse_qargs(ts, end=("@",)) for x in self._resolve(qa, None)]
                    if (1 << len(q)) != m.shape[0]:
                        raise ValueError(f"line {line}: a {m.shape[0]}x{m.shape[0]} Hermitian "
                                         f"needs {m.shape[0].bit_length() - 1} qubit(s)")
                    letters.append(m)
                    targets += q
                    if not ts.accept("@"):
                        break
                    continue
                if obs not in ("x", "y", "z", "h", "i"):
                    raise NotImplementedError(f"line {line}: observable {obs!r} is not supported "
                                              "(Pauli products with h and i factors and Hermitian "
                                              "matrices are)")
                if ts.accept("all"):
                    if letters:
                        raise QasmError(f"line {line}: 'all' cannot be part of a product")
                    return {"type": kind, "observable": [obs]}
                ts.expect("(")
                q = [x for qa in _parse_qargs(ts, end=(")",)) for x in self._resolve(qa, None)]
                ts.expect(")")
                letters += [obs] * len(q)
                targets += q
                if not ts.accept("@"):
                    break
            if not ts.done():
                raise QasmError(f"line {line}: unexpected {ts.peek()[1]!r} after the observable")
            self.circuit.qubit_set |= set(targets)
            return {"type": kind, "observable": letters, "targets": targets}
        raise NotImplementedError(f"line {line}: result type {kind!r} is not supported")


def _adjoint_result(self, ts: _Tokens, line: int) -> dict:
    """adjoint_gradient expectation(<sum of coefficient * Pauli products>)
    <input names | all>"""
    ts.expect("expectation")
    ts.expect("(")
    terms, touched = [], set()
    while True:
        sign = 1.0
        while ts.at("+") or ts.at("-"):
            if ts.next()[1] == "-":
                sign = -sign
        coef = 1.0
        if ts.peek()[0] == "num":
            coef = float(ts.next()[1])
            ts.expect("*")
        letters, qubits = [], []
        while True:
            obs = ts.ident().lower()
            if obs not in ("x", "y", "z", "h", "i"):
                raise NotImplementedError(f"line {line}: observable {obs!r} is not supported "
                                          "(Pauli products with h and i factors are)")
            ts.expect("(")
            q = [x for qa in _parse_qargs(ts, end=(")",)) for x in self._resolve(qa, None)]
            ts.expect(")")
            letters += [obs] * len(q)
            qubits += q
            if not ts.accept("@"):
                break
        terms.append((sign * coef, letters, qubits))
        touched |= set(qubits)
        if ts.at(")"):
            break
        if not (ts.at("+") or ts.at("-")):
            raise QasmError(f"line {line}: expected '+', '-' or ')' in the observable")
    ts.expect(")")
    if ts.accept("all"):
        names = None
    else:
        names = [ts.ident()]
        while ts.accept(","):
            names.append(ts.ident())
        unknown = [n for n in names if n not in self.circuit.inputs]
        if unknown:
            raise QasmError(f"line {line}: {unknown} are not inputs of the program")
    if not ts.done():
        raise QasmError(f"line {line}: unexpected {ts.peek()[1]!r} after the parameters")
    self.circuit.qubit_set |= touched
    return {"type": "adjoint_gradient", "observable": terms, "parameters": names}


_Interpreter._adjoint_result = _adjoint_result


class _Recorder(_Interpreter):
    """Collects a gate body's calls instead of emitting gate objects."""

    def __init__(self, parent: _Interpreter, out: list):
        self.__dict__.update(parent.__dict__)
        self.out = out

    def apply(self, name, params, qubits, mods, ev, line) -> None:
        ctrl, power = [], Fraction(1)
        for mname, arg in mods:
            if mname in ("ctrl", "negctrl"):
                n = 1 if arg is None else int(_eval(arg, ev))
                ctrl += [1 if mname == "ctrl" else 0] * n
            elif mname == "inv":
                power = -power
            else:
                power *= _fraction(_eval(arg, ev))
        self.out.append((name, list(params), list(qubits), ctrl, power))


def parse(source: str, inputs: Optional[dict] = None) -> Circuit:
    """Read an OpenQASM 3 program into gates, result types and measured qubits."""
    src = _strip_comments(source)
    toks = []
    for ln, text in enumerate(src.split("\n"), start=1):
        if text.strip().startswith("#pragma"):
            toks.append(("pragma", text, ln))
        else:
            toks += _tokenize(text, ln)
    interp = _Interpreter(inputs)
    interp.run(_parse_statements(_Tokens(toks)))
    return interp.circuit


_PAULI_FACTORS = {
    "x": (lambda q: [g.Hadamard(q)], (1.0, -1.0)),
    "y": (lambda q: [g.PauliZ(q), g.S(q), g.Hadamard(q)], (1.0, -1.0)),
    "z": (lambda q: [], (1.0, -1.0)),
    "h": (lambda q: [g.RotY(q, -math.pi / 4)], (1.0, -1.0)),
    "i": (lambda q: [], (1.0, 1.0)),
}


def hermitian_matrix(m) -> np.ndarray:
    """A Hermitian observable's matrix from a complex array or the JAQCD
    [[[re, im], ...], ...] form; checked square, power-of-two, Hermitian."""
    a = np.asarray(m)
    if a.ndim == 3 and a.shape[-1] == 2 and not np.iscomplexobj(a):
        a = a[..., 0] + 1j * a[..., 1]
    a = np.asarray(a, dtype=np.complex128)
    d = a.shape[0] if a.ndim == 2 else 0
    if a.ndim != 2 or a.shape != (d, d) or d < 2 or d & (d - 1):
        raise ValueError(f"a Hermitian observable needs a square 2^k matrix, got {a.shape}")
    if not np.allclose(a, a.conj().T, atol=1e-8):
        raise ValueError("the observable matrix is not Hermitian")
    return a


def observable_factor(o):
    """(qubits, rotation builder, eigenvalues) of one observable factor: the
    gates that map its eigenbasis to the computational basis and the
    eigenvalue of each outcome (observables.py:136-305: Pauli X/Y/Z, H, I and
    Hermitian's eigendecomposition)."""
    if isinstance(o, str):
        f = _PAULI_FACTORS.get(o.lower())
        if f is None:
            raise NotImplementedError(f"observable {o!r} is not supported")
        make, eig = f
        return 1, (lambda qs, make=make: make([qs[0]])), np.array(eig)
    m = hermitian_matrix(o)
    lam, v = np.linalg.eigh(m)
    k = m.shape[0].bit_length() - 1
    d = v.conj().T
    return k, (lambda qs, d=d: [g.Unitary(list(qs), d)]), lam


def basis_rotations(results: list, qubits) -> list:
    """Gates that rotate each observable's eigenbasis to Z before sampling
    (openqasm/circuit.py:95-143): x -> H; y -> Z, S, H; h -> RotY(-pi/4);
    a Hermitian -> the adjoint of its eigenvector matrix.  A qubit measured
    in two different bases is an error."""
    per, out = {}, []
    for r in results:
        if r.get("type") not in ("expectation", "variance", "sample"):
            continue
        obs, targets = list(r["observable"]), r.get("targets")
        if not targets:
            obs, targets = obs * len(qubits), list(qubits)
        k = 0
        for o in obs:
            nq, make, _ = observable_factor(o)
            qs = tuple(targets[k:k + nq])
            k += nq
            key = o.lower() if isinstance(o, str) else hermitian_matrix(o).tobytes()
            for q in qs:
                prev = per.get(q)
                if prev is not None and prev != (key, qs):
                    raise ValueError("Conflicting result types applied to a single qubit")
            if all(per.get(q) == (key, qs) for q in qs):
                continue
            for q in qs:
                per[q] = (key, qs)
            out += make(list(qs))
    return out
