"""`operator(expr)`: an operator from textbook notation, the `@operator` macro of the
reference (src/operators/operator_macro.jl:1-134).

Julia rewrites the expression at parse time; here the notation comes as a string whose
free names are looked up in `names` / keyword arguments, e.g.

    operator("∇² + k² * f", {"k²": 1.7})
    operator("∇ ⋅ (κ * ∇)", κ=[3.0, 0.5])
    operator("ν * ∇² - c ⋅ ∇", ν=0.1, c=[1.0, 0.5])

Recognised (operator_macro.jl:37-43, 52-80, 83-134):
  ∇², Δ           Laplacian()
  f, I            Identity()
  ∂(i)            Partial(1, i);  ∂(i, j) MixedPartial(i, j);  ∂²(i) Partial(2, i)
  ∇ ⋅ ∇           Laplacian()
  ∇ ⋅ (κ * ∇)     κ * Laplacian() for a scalar κ, Σᵢ κ[i] * Partial(2, i) for a vector
  c ⋅ ∇           Σᵢ c[i] * Partial(1, i)
  + - * / ^       on operators and scalars, as in the operator algebra
and everything else is a scalar: numbers, names, `name[i]` (1-based, as in Julia),
`[a, b]` vectors, and calls of functions given in `names` (sqrt, exp, log, sin, cos, tan,
abs and π are predefined). `⋅` (or `·`) binds like `*`; `^` is a power; a number written
against a name (`2k`) multiplies, as in Julia.
The expression is parsed with Python's `ast` module and walked here; nothing in it is
executed as Python code.
"""
import ast
import math
import numbers

import numpy as np

from .errors import ArgumentError
from .operators import AbstractOperator, Laplacian, Identity, Partial, MixedPartial

_LAP, _NABLA, _D, _D2 = "__lap__", "__nabla__", "__d__", "__d2__"
_DEFAULTS = {"sqrt": math.sqrt, "exp": math.exp, "log": math.log, "sin": math.sin, "cos": math.cos,
             "tan": math.tan, "abs": abs, "π": math.pi, "pi": math.pi}
_NUMBER_CHARS = "0123456789."


def _translate(src):
    """Notation -> Python source; every free name is replaced by a placeholder __nK__."""
    out, names, i, n = [], {}, 0, len(src)

    def place(name):
        if name not in names:
            names[name] = f"__n{len(names)}__"
        return names[name]

    while i < n:
        ch = src[i]
        nxt = src[i + 1] if i + 1 < n else ""
        if ch == "∇":
            out.append(f" {_LAP} " if nxt == "²" else f" {_NABLA} ")
            i += 2 if nxt == "²" else 1
        elif ch == "∂":
            out.append(f" {_D2} " if nxt == "²" else f" {_D} ")
            i += 2 if nxt == "²" else 1
        elif ch in "⋅·":
            out.append(" @ ")
            i += 1
        elif ch == "^":
            out.append("**")
            i += 1
        elif ch.isalpha() or ch == "_":
            j = i
            while j < n and (src[j].isalnum() or src[j] == "_"):
                j += 1
            out.append(place(src[i:j]))
            i = j
        elif ch.isdigit() or (ch == "." and nxt.isdigit()):
            j = i
            while j < n and src[j] in _NUMBER_CHARS:
                j += 1
            if j < n and src[j] in "eE" and j + 1 < n and (src[j + 1].isdigit() or
                                                       (src[j + 1] in "+-" and src[j + 2:j + 3].isdigit())):
                j += 2
                while j < n and src[j].isdigit():
                    j += 1
            out.append(src[i:j])
            if j < n and (src[j].isalpha() or src[j] in "∇∂("):
                out.append("*")  # juxtaposition: 2k, 2∇², 3(…)
            i = j
        else:
            out.append(ch)
            i += 1
    return "".join(out), {v: k for k, v in names.items()}


class _Eval:
    def __init__(self, names, scope):
        self.names, self.scope = names, scope

    # ------------------------------------------------------------ leaves
    def name(self, node):
        ident = node.id
        if ident == _LAP:
            return Laplacian()
        if ident in (_NABLA, _D, _D2):
            sym = {_NABLA: "∇", _D: "∂", _D2: "∂²"}[ident]
            raise ArgumentError(f"{sym} must appear as " + ("∇ ⋅ ∇, ∇ ⋅ (κ * ∇) or c ⋅ ∇" if ident == _NABLA
                                                             else f"{sym}(dim)"))
        orig = self.names[ident]
        if orig in ("f", "I"):
            return Identity()
        if orig == "Δ":
            return Laplacian()
        if orig in self.scope:
            v = self.scope[orig]
            return np.asarray(v, dtype=np.float64) if isinstance(v, (list, tuple)) else v
        if orig in _DEFAULTS:
            return _DEFAULTS[orig]
        raise ArgumentError(f"operator: the name {orig!r} is not given")

    def ev(self, node):
        if isinstance(node, ast.Expression):
            return self.ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, numbers.Number):
            return node.value
        if isinstance(node, ast.Name):
            return self.name(node)
        if isinstance(node, (ast.List, ast.Tuple)):
            return np.array([self.scalar(self.ev(e)) for e in node.elts], dtype=np.float64)
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = self.ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.BinOp):
            return self.binop(node)
        if isinstance(node, ast.Call):
            return self.call(node)
        if isinstance(node, ast.Subscript):
            v, k = self.ev(node.value), self.scalar(self.ev(node.slice))
            if isinstance(v, AbstractOperator) or int(k) != k or k < 1:
                raise ArgumentError("operator: indexing takes a vector and a 1-based integer")
            return np.asarray(v)[int(k) - 1]
        raise ArgumentError(f"operator: unsupported notation {ast.dump(node)[:60]}")

    @staticmethod
    def scalar(v):
        if isinstance(v, AbstractOperator):
            raise ArgumentError(f"operator: {v!r} used where a scalar is expected")
        return v

    # ------------------------------------------------------------ calls
    def call(self, node):
        if node.keywords:
            raise ArgumentError("operator: keyword arguments are not supported")
        f = node.func
        if isinstance(f, ast.Name) and f.id in (_D, _D2):
            args = [int(self.scalar(self.ev(a))) for a in node.args]
            if f.id == _D:
                if len(args) == 1:
                    return Partial(1, args[0])
                if len(args) == 2:
                    return MixedPartial(args[0], args[1])
                raise ArgumentError(f"∂ expects 1 or 2 arguments, got {len(args)}")
            if len(args) != 1:
                raise ArgumentError(f"∂²(dim) expects exactly 1 argument, got {len(args)}")
            return Partial(2, args[0])
        fn = self.ev(f)
        if not callable(fn) or isinstance(fn, AbstractOperator):
            raise ArgumentError("operator: only functions given in names can be called")
        return fn(*[self.scalar(self.ev(a)) for a in node.args])

    # ------------------------------------------------------------ operators
    def factors(self, node):
        """Flatten a * b * c."""
        if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Mult):
            return self.factors(node.left) + self.factors(node.right)
        return [node]

    @staticmethod
    def is_nabla(node):
        return isinstance(node, ast.Name) and node.id == _NABLA

    def dot(self, node):
        lhs, rhs = node.left, node.right
        if self.is_nabla(lhs) and self.is_nabla(rhs):
            return Laplacian()
        if self.is_nabla(lhs):
            fs = self.factors(rhs)
            if any(self.is_nabla(x) for x in fs):
                coeff = [x for x in fs if not self.is_nabla(x)]
                if len(coeff) != len(fs) - 1:
                    raise ArgumentError("operator: ∇ ⋅ (κ * ∇) takes one ∇ on the right")
                k = 1.0
                for x in coeff:
                    k = k * self.scalar(self.ev(x))
                return expand_div_grad(k)
        if self.is_nabla(rhs):
            return expand_dot_grad(self.scalar(self.ev(lhs)))
        a, b = self.scalar(self.ev(lhs)), self.scalar(self.ev(rhs))
        return np.dot(a, b)

    def binop(self, node):
        if isinstance(node.op, ast.MatMult):
            return self.dot(node)
        a, b = self.ev(node.left), self.ev(node.right)
        op = node.op
        if isinstance(op, ast.Pow):
            return self.scalar(a) ** self.scalar(b)
        try:
            if isinstance(op, ast.Add):
                r = a + b
            elif isinstance(op, ast.Sub):
                r = a - b
            elif isinstance(op, ast.Mult):
                r = a * b
            elif isinstance(op, ast.Div):
                if isinstance(b, AbstractOperator):
                    raise TypeError
                r = a / b
            else:
                raise TypeError
        except TypeError:
            raise ArgumentError(f"operator: cannot apply {type(op).__name__} to {a!r} and {b!r}") from None
        if r is NotImplemented:
            raise ArgumentError(f"operator: cannot apply {type(op).__name__} to {a!r} and {b!r}")
        return r


def expand_dot_grad(c):
    """c ⋅ ∇ = Σᵢ c[i] ∂/∂xᵢ (_expand_dot_grad, operator_macro.jl:112-117)."""
    c = np.atleast_1d(np.asarray(c, dtype=np.float64))
    if c.ndim != 1 or c.size == 0:
        raise ArgumentError("c ⋅ ∇ takes a vector c")
    op = None
    for i, ci in enumerate(c):
        term = float(ci) * Partial(1, i + 1)
        op = term if op is None else op + term
    return op


def expand_div_grad(kappa):
    """∇ ⋅ (κ∇): κ ∇² for a scalar, Σᵢ κ[i] ∂²/∂xᵢ² for a vector (operator_macro.jl:119-134)."""
    if isinstance(kappa, numbers.Number) or np.ndim(kappa) == 0:
        return float(kappa) * Laplacian()
    k = np.asarray(kappa, dtype=np.float64).ravel()
    op = None
    for i, ki in enumerate(k):
        term = float(ki) * Partial(2, i + 1)
        op = term if op is None else op + term
    return op


def operator(expr, names=None, **kw):
    """The operator written in `expr` (the `@operator` macro, operator_macro.jl:25-27)."""
    if not isinstance(expr, str):
        raise ArgumentError("operator takes the notation as a string")
    scope = dict(names or {})
    scope.update(kw)
    src, placeholders = _translate(expr)
    try:
        tree = ast.parse(src.strip(), mode="eval")
    except SyntaxError as e:
        raise ArgumentError(f"operator: cannot parse {expr!r} ({e.msg})") from None
    out = _Eval(placeholders, scope).ev(tree)
    if not isinstance(out, AbstractOperator):
        raise ArgumentError(f"operator: {expr!r} is a scalar, not an operator")
    return out
