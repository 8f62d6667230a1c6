"""Closures outside the traced forms, compiled for the GPU at problem set-up.

The reference accepts any Julia closure for D(x, y, t, u, p) and S(x, y, t, u, p)
(problem.jl:118-163, 425-440) and Julia compiles it into fvm_eqs!.  The forms in
forms.py cover the reference's own tests and tutorials; a closure outside them
(exp(-u), sin(u), a quotient, a u-dependent exponent, x-dependent nonlinear terms,
np.where selections ...) is traced here with `Expr` stand-ins for x, y, t and u into
an expression DAG, which is emitted as HIP source:

    fvm_jit_D(x, y, t, u), fvm_jit_dD (its u-derivative, by forward differentiation
    of the DAG), fvm_jit_S(x, y, t, u), fvm_jit_dS

The library compiles that source with hipRTC into the tile, fix-up and VJP kernels
(FVM_FLUX_JIT, fvm_jit.hip), so the JVP and VJP of such problems come from the
generated derivatives.  The binary operations keep the closure's evaluation order
(left to right, as Python evaluates it), so sums and products round as the closure
does on the host; elementary functions are the device math library's (within an ulp
or two of the host's).  Constants are written as hexadecimal floats (exact).
Branching in Python on u (`if u > 0:`) cannot be traced and raises
UnsupportedClosure; np.where, np.minimum / np.maximum and comparisons can.
"""
import numbers

import numpy as np

from .forms import UnsupportedClosure

_UNARY = {
    np.exp: "exp", np.log: "log", np.sin: "sin", np.cos: "cos", np.tan: "tan", np.sqrt: "sqrt",
    np.absolute: "abs", np.fabs: "abs", np.tanh: "tanh", np.sinh: "sinh", np.cosh: "cosh",
    np.arctan: "atan", np.arcsin: "asin", np.arccos: "acos", np.log1p: "log1p", np.expm1: "expm1",
    np.log10: "log10", np.log2: "log2", np.exp2: "exp2", np.square: "square", np.negative: "neg",
    np.positive: "pos", np.reciprocal: "recip", np.cbrt: "cbrt", np.arcsinh: "asinh",
    np.arccosh: "acosh", np.arctanh: "atanh",
}
_BINARY = {
    np.add: "add", np.subtract: "sub", np.multiply: "mul", np.true_divide: "div", np.power: "pow",
    np.minimum: "min", np.maximum: "max", np.fmin: "min", np.fmax: "max", np.less: "lt",
    np.less_equal: "le", np.greater: "gt", np.greater_equal: "ge", np.arctan2: "atan2",
    np.hypot: "hypot",
}
_CMP = {"lt": "<", "le": "<=", "gt": ">", "ge": ">="}


class Expr:
    """A node of the traced expression DAG: op with argument nodes, or a leaf
    ('var' with a name, 'const' with a float)."""

    __array_priority__ = 1000
    __slots__ = ("op", "args", "value")

    def __init__(self, op, args=(), value=None):
        self.op, self.args, self.value = op, tuple(args), value

    @staticmethod
    def lift(v):
        if isinstance(v, Expr):
            return v
        if isinstance(v, (bool, np.bool_)):
            return Expr("const", value=float(v))
        if isinstance(v, (numbers.Real, np.generic)) or (isinstance(v, np.ndarray) and v.ndim == 0):
            return Expr("const", value=float(v))
        raise UnsupportedClosure(f"a closure compiled for the device may only combine scalars, got {type(v).__name__}")

    def _b(self, op, o, rev=False):
        try:
            o = Expr.lift(o)
        except UnsupportedClosure:
            return NotImplemented
        return Expr(op, (o, self) if rev else (self, o))

    def __add__(self, o):
        return self._b("add", o)

    def __radd__(self, o):
        return self._b("add", o, True)

    def __sub__(self, o):
        return self._b("sub", o)

    def __rsub__(self, o):
        return self._b("sub", o, True)

    def __mul__(self, o):
        return self._b("mul", o)

    def __rmul__(self, o):
        return self._b("mul", o, True)

    def __truediv__(self, o):
        return self._b("div", o)

    def __rtruediv__(self, o):
        return self._b("div", o, True)

    def __pow__(self, o):
        return self._b("pow", o)

    def __rpow__(self, o):
        return self._b("pow", o, True)

    def __neg__(self):
        return Expr("neg", (self,))

    def __pos__(self):
        return self

    def __abs__(self):
        return Expr("abs", (self,))

    def __lt__(self, o):
        return self._b("lt", o)

    def __le__(self, o):
        return self._b("le", o)

    def __gt__(self, o):
        return self._b("gt", o)

    def __ge__(self, o):
        return self._b("ge", o)

    def __bool__(self):
        raise UnsupportedClosure("branching in Python on a solution-dependent quantity (use np.where)")

    def __float__(self):
        raise UnsupportedClosure("float() / math-module functions of a solution-dependent quantity "
                                 "(use the numpy functions)")

    __int__ = __index__ = __float__

    def __array_ufunc__(self, ufunc, method, *inputs, **kw):
        if method != "__call__" or kw.get("out") is not None:
            return NotImplemented
        if ufunc in _UNARY and len(inputs) == 1:
            op = _UNARY[ufunc]
            a = Expr.lift(inputs[0])
            if op == "pos":
                return a
            if op == "square":
                return Expr("mul", (a, a))
            if op == "recip":
                return Expr("div", (Expr("const", value=1.0), a))
            return Expr(op, (a,))
        if ufunc in _BINARY and len(inputs) == 2:
            return Expr(_BINARY[ufunc], (Expr.lift(inputs[0]), Expr.lift(inputs[1])))
        raise UnsupportedClosure(f"numpy.{ufunc.__name__} cannot be compiled for the device")

    def __array_function__(self, func, types, args, kwargs):
        if func is np.where and len(args) == 3 and not kwargs:
            return Expr("where", tuple(Expr.lift(a) for a in args))
        if func is np.clip and len(args) == 3 and not kwargs:
            a, lo, hi = (Expr.lift(v) for v in args)
            return Expr("min", (Expr("max", (a, lo)), hi))
        raise UnsupportedClosure(f"numpy.{func.__name__} cannot be compiled for the device")


def _var(name):
    return Expr("var", value=name)


def trace(fn):
    """The DAG of fn(x, y, t, u) (a closure already bound to its parameters)."""
    try:
        r = fn(_var("x"), _var("y"), _var("t"), _var("u"))
    except UnsupportedClosure:
        raise
    except Exception as err:  # noqa: BLE001 - any failure means the closure cannot be compiled
        raise UnsupportedClosure(f"compiling the closure failed: {type(err).__name__}: {err}") from err
    if isinstance(r, np.ndarray) and r.ndim == 0:
        r = r[()]
    return Expr.lift(r)


def _lit(v):
    """A double literal that reads back exactly (hexadecimal float), parenthesised
    when negative."""
    v = float(v)
    if v != v:
        return "__builtin_nan(\"\")"
    if v in (float("inf"), float("-inf")):
        return "__builtin_inf()" if v > 0 else "(-__builtin_inf())"
    if v == 0:
        return "0.0" if np.copysign(1.0, v) > 0 else "(-0.0)"
    return v.hex() if v > 0 else f"({v.hex()})"


class _Emitter:
    """Straight-line code for the value and u-derivative of every node (forward
    differentiation), each node once."""

    def __init__(self):
        self.lines = []
        self.names = {}
        self.k = 0

    def node(self, e):
        key = id(e)
        if key in self.names:
            return self.names[key]
        if e.op == "var":
            res = (e.value, "1.0" if e.value == "u" else "0.0")
        elif e.op == "const":
            res = (_lit(e.value), "0.0")
        else:
            args = [self.node(a) for a in e.args]
            v, d = self._op(e, args)
            n = self.k
            self.k += 1
            self.lines.append(f"  const double v{n} = {v};")
            self.lines.append(f"  const double d{n} = {d};")
            res = (f"v{n}", f"d{n}")
        self.names[key] = res
        return res

    @staticmethod
    def _op(e, args):
        op = e.op
        (a, da) = args[0]
        if len(args) > 1:
            (b, db) = args[1]
        if op == "add":
            return f"{a} + {b}", f"{da} + {db}"
        if op == "sub":
            return f"{a} - {b}", f"{da} - {db}"
        if op == "mul":
            return f"{a} * {b}", f"{da} * {b} + {a} * {db}"
        if op == "div":
            return f"{a} / {b}", f"({da} * {b} - {a} * {db}) / ({b} * {b})"
        if op == "neg":
            return f"-{a}", f"-{da}"
        if op == "pow":
            cst = e.args[1].op == "const"
            if cst and e.args[1].value == 2.0:
                return f"{a} * {a}", f"2.0 * {a} * {da}"
            if cst:
                return f"pow({a}, {b})", f"{b} * pow({a}, {b} - 1.0) * {da}"
            return (f"pow({a}, {b})",
                    f"pow({a}, {b}) * ({db} * log({a}) + {b} * {da} / {a})")
        if op in ("lt", "le", "gt", "ge"):
            return f"(({a} {_CMP[op]} {b}) ? 1.0 : 0.0)", "0.0"
        if op == "where":
            (c, _), (b2, db2) = args[1], args[2]
            return f"({a} != 0.0 ? {c} : {b2})", f"({a} != 0.0 ? {args[1][1]} : {db2})"
        if op == "min":
            return f"fmin({a}, {b})", f"({a} <= {b} ? {da} : {db})"
        if op == "max":
            return f"fmax({a}, {b})", f"({a} >= {b} ? {da} : {db})"
        if op == "atan2":
            return f"atan2({a}, {b})", f"({b} * {da} - {a} * {db}) / ({a} * {a} + {b} * {b})"
        if op == "hypot":
            return f"hypot({a}, {b})", f"({a} * {da} + {b} * {db}) / hypot({a}, {b})"
        unary = {
            "exp": ("exp({a})", "exp({a}) * {da}"),
            "log": ("log({a})", "{da} / {a}"),
            "log10": ("log10({a})", "{da} / ({a} * 0x1.26bb1bbb55516p+1)"),
            "log2": ("log2({a})", "{da} / ({a} * 0x1.62e42fefa39efp-1)"),
            "exp2": ("exp2({a})", "exp2({a}) * 0x1.62e42fefa39efp-1 * {da}"),
            "sin": ("sin({a})", "cos({a}) * {da}"),
            "cos": ("cos({a})", "-sin({a}) * {da}"),
            "tan": ("tan({a})", "{da} / (cos({a}) * cos({a}))"),
            "sqrt": ("sqrt({a})", "{da} / (2.0 * sqrt({a}))"),
            "cbrt": ("cbrt({a})", "{da} / (3.0 * cbrt({a}) * cbrt({a}))"),
            "abs": ("fabs({a})", "({a} > 0.0 ? {da} : ({a} < 0.0 ? -{da} : 0.0))"),
            "tanh": ("tanh({a})", "(1.0 - tanh({a}) * tanh({a})) * {da}"),
            "sinh": ("sinh({a})", "cosh({a}) * {da}"),
            "cosh": ("cosh({a})", "sinh({a}) * {da}"),
            "atan": ("atan({a})", "{da} / (1.0 + {a} * {a})"),
            "asin": ("asin({a})", "{da} / sqrt(1.0 - {a} * {a})"),
            "acos": ("acos({a})", "-{da} / sqrt(1.0 - {a} * {a})"),
            "asinh": ("asinh({a})", "{da} / sqrt({a} * {a} + 1.0)"),
            "acosh": ("acosh({a})", "{da} / sqrt({a} * {a} - 1.0)"),
            "atanh": ("atanh({a})", "{da} / (1.0 - {a} * {a})"),
            "log1p": ("log1p({a})", "{da} / (1.0 + {a})"),
            "expm1": ("expm1({a})", "exp({a}) * {da}"),
        }
        if op in unary:
            v, d = unary[op]
            return v.format(a=a, da=da), d.format(a=a, da=da)
        raise UnsupportedClosure(f"operation {op} cannot be compiled for the device")


def emit(name, e):
    """HIP source of fvm_jit_<name> and fvm_jit_d<name> for the DAG e."""
    em = _Emitter()
    v, d = em.node(e)
    body = "\n".join(em.lines)
    sig = "(double x, double y, double t, double u)"
    unused = "  (void)x; (void)y; (void)t; (void)u;\n"
    return (f"__device__ __forceinline__ void fvm_jit_{name}_vd{sig[:-1]}, double* val, double* der) {{\n"
            f"{unused}{body}\n  *val = {v};\n  *der = {d};\n}}\n"
            f"__device__ double fvm_jit_{name}{sig} {{\n  double v, d;\n"
            f"  fvm_jit_{name}_vd(x, y, t, u, &v, &d);\n  return v;\n}}\n"
            f"__device__ double fvm_jit_d{name}{sig} {{\n  double v, d;\n"
            f"  fvm_jit_{name}_vd(x, y, t, u, &v, &d);\n  return d;\n}}\n")


def source_of(D, S):
    """HIP source defining fvm_jit_D, dD, S, dS for the closures D(x, y, t, u) and
    S(x, y, t, u) (both already bound to their parameters)."""
    return ("// generated from the problem's diffusion and source closures\n"
            + emit("D", trace(D)) + emit("S", trace(S)))


def evaluate(e, x, y, t, u):
    """Value of the DAG at one point in double arithmetic on the host, in the order the
    generated code evaluates it (a check of the trace, used by the tests)."""
    memo = {}

    def ev(n):
        k = id(n)
        if k in memo:
            return memo[k]
        if n.op == "var":
            r = {"x": x, "y": y, "t": t, "u": u}[n.value]
        elif n.op == "const":
            r = n.value
        else:
            a = [ev(m) for m in n.args]
            op = n.op
            f = {"add": lambda: a[0] + a[1], "sub": lambda: a[0] - a[1], "mul": lambda: a[0] * a[1],
                 "div": lambda: a[0] / a[1], "neg": lambda: -a[0],
                 "pow": lambda: a[0] * a[0] if (n.args[1].op == "const" and n.args[1].value == 2.0)
                 else float(np.power(a[0], a[1])),
                 "lt": lambda: float(a[0] < a[1]), "le": lambda: float(a[0] <= a[1]),
                 "gt": lambda: float(a[0] > a[1]), "ge": lambda: float(a[0] >= a[1]),
                 "where": lambda: a[1] if a[0] != 0.0 else a[2], "min": lambda: min(a[0], a[1]),
                 "max": lambda: max(a[0], a[1]), "abs": lambda: abs(a[0]),
                 "atan2": lambda: float(np.arctan2(a[0], a[1])), "hypot": lambda: float(np.hypot(a[0], a[1]))}
            if op in f:
                r = f[op]()
            else:
                r = float(getattr(np, {"atan": "arctan", "asin": "arcsin", "acos": "arccos", "asinh": "arcsinh",
                                       "acosh": "arccosh", "atanh": "arctanh"}.get(op, op))(a[0]))
        memo[k] = r
        return r

    return ev(e)
