"""Mapping of the user's closures onto the kernels' closed forms.

The reference accepts arbitrary Julia closures for the diffusion / flux /
source / condition functions (problem.jl:118-163, conditions.jl:16-20).  Device
kernels cannot call back into the host, so each closure is traced once, at
problem set-up, by calling it with symbolic stand-ins for the solution-dependent
arguments (u, t, and alpha/beta/gamma) and NumPy arrays for the coordinates.  The
trace must land on one of the forms the kernels implement:

  diffusion D(x, y, t, u, p): polynomial of degree <= 3 in u with constant
      coefficients (FVM_FLUX_D_POLY_U), or any function of (x, y) alone
      (FVM_FLUX_D_EDGE: evaluated at every CV-edge midpoint and boundary
      half-edge point), or a `NodalDiffusion` field (FVM_FLUX_D_NODAL);
  flux q(x, y, t, a, b, g, p) -> (qx, qy): -D(x, y) (a, b);
  source S(x, y, t, u, p): s(x, y) + cubic polynomial in u + c t;
  condition a(x, y, t, u, p): g(x, y) + cu u + ct t.

Anything else raises `UnsupportedClosure` with the reason: there is no silent
CPU fallback.
"""
import numbers
from collections import defaultdict

import numpy as np


class UnsupportedClosure(NotImplementedError):
    pass


class NodalDiffusion:
    """A per-node diffusivity field D_v = exp(log_d[v]), interpolated piecewise
    linearly inside each triangle (what a reference diffusion function that
    point-locates (x, y) and calls pl_interpolate on exp.(log_d) evaluates)."""

    def __init__(self, log_d):
        self.log_d = np.ascontiguousarray(log_d, np.float64)

    def __call__(self, x, y, t, u, p):
        raise UnsupportedClosure("NodalDiffusion is evaluated by the device kernels only")


class Sym:
    """Polynomial in symbolic variables with coefficients that are scalars or
    arrays over the evaluation points: {exponent tuple: coefficient}."""

    __array_priority__ = 1000

    def __init__(self, nvars, terms=None):
        self.nvars = nvars
        self.terms = dict(terms or {})

    @classmethod
    def var(cls, nvars, k):
        e = [0] * nvars
        e[k] = 1
        return cls(nvars, {tuple(e): 1.0})

    @classmethod
    def const(cls, nvars, c):
        return cls(nvars, {(0,) * nvars: c})

    def _lift(self, o):
        if isinstance(o, Sym):
            return o
        if isinstance(o, (numbers.Number, np.ndarray, np.generic)):
            return Sym.const(self.nvars, o)
        return NotImplemented

    def __add__(self, o):
        o = self._lift(o)
        if o is NotImplemented:
            return o
        t = dict(self.terms)
        for e, c in o.terms.items():
            t[e] = t[e] + c if e in t else c
        return Sym(self.nvars, t)

    __radd__ = __add__

    def __neg__(self):
        return Sym(self.nvars, {e: -c for e, c in self.terms.items()})

    def __pos__(self):
        return self

    def __sub__(self, o):
        o = self._lift(o)
        if o is NotImplemented:
            return o
        return self + (-o)

    def __rsub__(self, o):
        o = self._lift(o)
        if o is NotImplemented:
            return o
        return o + (-self)

    def __mul__(self, o):
        o = self._lift(o)
        if o is NotImplemented:
            return o
        t = {}
        for e1, c1 in self.terms.items():
            for e2, c2 in o.terms.items():
                e = tuple(a + b for a, b in zip(e1, e2))
                t[e] = t[e] + c1 * c2 if e in t else c1 * c2
        return Sym(self.nvars, t)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, Sym):
            if not o.is_const():
                raise UnsupportedClosure("division by a solution-dependent quantity")
            o = o.const_value()
        return Sym(self.nvars, {e: c / o for e, c in self.terms.items()})

    def __rtruediv__(self, o):
        if not self.is_const():
            raise UnsupportedClosure("division by a solution-dependent quantity")
        return o / self.const_value()

    def __pow__(self, k):
        if isinstance(k, Sym):
            raise UnsupportedClosure("solution-dependent exponent")
        if int(k) != k or k < 0:
            if self.is_const():
                return Sym.const(self.nvars, self.const_value() ** k)
            raise UnsupportedClosure("non-integer power of a solution-dependent quantity")
        r = Sym.const(self.nvars, 1.0)
        for _ in range(int(k)):
            r = r * self
        return r

    def __bool__(self):
        raise UnsupportedClosure("branching on a solution-dependent quantity")

    def _cmp(self, o):
        raise UnsupportedClosure("comparison of a solution-dependent quantity")

    __lt__ = __le__ = __gt__ = __ge__ = _cmp

    def __array_ufunc__(self, ufunc, method, *inputs, **kw):
        if method != "__call__":
            return NotImplemented
        vals = []
        for v in inputs:
            if isinstance(v, Sym):
                if not v.is_const():
                    if ufunc in (np.add, np.subtract, np.multiply, np.negative, np.true_divide):
                        break
                    raise UnsupportedClosure(f"{ufunc.__name__} of a solution-dependent quantity")
                vals.append(v.const_value())
            else:
                vals.append(v)
        else:
            return Sym.const(self.nvars, ufunc(*vals, **kw))
        a = inputs[0]
        b = inputs[1] if len(inputs) > 1 else None
        if ufunc is np.add:
            return (a if isinstance(a, Sym) else Sym.const(self.nvars, a)) + b
        if ufunc is np.subtract:
            return (a if isinstance(a, Sym) else Sym.const(self.nvars, a)) - b
        if ufunc is np.multiply:
            return (a if isinstance(a, Sym) else Sym.const(self.nvars, a)) * b
        if ufunc is np.negative:
            return -a
        if ufunc is np.true_divide:
            return (a if isinstance(a, Sym) else Sym.const(self.nvars, a)) / b
        return NotImplemented

    def is_const(self):
        return all(not any(e) for e, c in self.terms.items() if not _is_zero(c))

    def const_value(self):
        return self.terms.get((0,) * self.nvars, 0.0)

    def coef(self, e):
        return self.terms.get(tuple(e), 0.0)

    def nonzero_exponents(self):
        return [e for e, c in self.terms.items() if not _is_zero(c)]


def _is_zero(c):
    return np.all(np.asarray(c) == 0)


def _const_over(c, n):
    """c as a float if it is the same at all n points, else None."""
    a = np.broadcast_to(np.asarray(c, np.float64), (n,)) if n else np.asarray([], np.float64)
    if n == 0:
        return 0.0
    if np.all(a == a[0]):
        return float(a[0])
    return None


def _trace(fn, nvars, x, y, extra_args, sym_positions):
    """Call fn(x, y, *args) with Sym variables at sym_positions of args."""
    args = list(extra_args)
    for k, pos in enumerate(sym_positions):
        args[pos] = Sym.var(nvars, k)
    try:
        r = fn(x, y, *args)
    except UnsupportedClosure:
        raise
    except Exception as err:  # noqa: BLE001 - any failure means the form is not traceable
        raise UnsupportedClosure(f"tracing the closure failed: {type(err).__name__}: {err}") from err
    return r


def _as_sym(r, nvars):
    return r if isinstance(r, Sym) else Sym.const(nvars, r)


# ---------------------------------------------------------------------------
def classify_condition(fn, x, y):
    """a(x, y, t, u) = g(x, y) + cu u + ct t over the points (x, y).

    Returns (g[len(x)], cu, ct)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = len(x)
    if n == 0:
        return np.zeros(0), 0.0, 0.0
    r = _as_sym(_trace(fn, 2, x, y, [None, None], [0, 1]), 2)  # vars: t, u
    for e in r.nonzero_exponents():
        if e not in ((0, 0), (1, 0), (0, 1)):
            raise UnsupportedClosure("condition functions must be affine in t and u "
                                     "(a = g(x, y) + cu*u + ct*t)")
    cu = _const_over(r.coef((0, 1)), n)
    ct = _const_over(r.coef((1, 0)), n)
    if cu is None or ct is None:
        raise UnsupportedClosure("the u and t coefficients of a condition must not depend on x, y")
    g = np.broadcast_to(np.asarray(r.coef((0, 0)), np.float64), (n,)).copy()
    return g, cu, ct


def classify_source(fn, x, y):
    """S(x, y, t, u) = s(x, y) + s1 u + s2 u^2 + s3 u^3 + st t.

    Returns (s_node or None, [s0, s1, s2, s3], st)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = len(x)
    r = _as_sym(_trace(fn, 2, x, y, [None, None], [0, 1]), 2)  # vars: t, u
    poly = [0.0, 0.0, 0.0, 0.0]
    st = 0.0
    for e in r.nonzero_exponents():
        kt, ku = e
        if kt == 0 and ku == 0:
            continue
        if kt == 1 and ku == 0:
            st = _const_over(r.coef(e), n)
            if st is None:
                raise UnsupportedClosure("the t coefficient of the source must not depend on x, y")
        elif kt == 0 and ku <= 3:
            c = _const_over(r.coef(e), n)
            if c is None:
                raise UnsupportedClosure("the u coefficients of the source must not depend on x, y")
            poly[ku] = c
        else:
            raise UnsupportedClosure("source must be s(x, y) + cubic polynomial in u + c*t")
    c0 = r.coef((0, 0))
    s0 = _const_over(c0, n)
    if s0 is not None:
        poly[0] = s0
        return None, poly, st
    return np.broadcast_to(np.asarray(c0, np.float64), (n,)).copy(), poly, st


def classify_diffusion(fn):
    """Trace D(x, y, t, u, p) (already bound to p).  Returns ('poly', [d0..d3]) for
    a constant-coefficient cubic in u, or ('xy', None) when D depends on (x, y)
    only (the caller evaluates it at the CV-edge points)."""
    probe_x = np.array([0.1234, 0.7421, 0.3317, 0.9012])
    probe_y = np.array([0.5531, 0.0713, 0.8822, 0.2417])
    r = _as_sym(_trace(fn, 2, probe_x, probe_y, [None, None], [0, 1]), 2)  # vars: t, u
    poly = [0.0, 0.0, 0.0, 0.0]
    depends_xy = False
    for e in r.nonzero_exponents():
        kt, ku = e
        if kt != 0:
            raise UnsupportedClosure("time-dependent diffusion is not supported by the device kernels")
        if ku > 3:
            raise UnsupportedClosure("diffusion must be a polynomial of degree <= 3 in u")
        c = _const_over(r.coef(e), len(probe_x))
        if c is None:
            if ku != 0:
                raise UnsupportedClosure("u coefficients of the diffusion must not depend on x, y")
            depends_xy = True
        else:
            poly[ku] = c
    if depends_xy:
        if any(poly[1:]):
            raise UnsupportedClosure("diffusion mixing (x, y) and u dependence is not supported")
        return "xy", None
    return "poly", poly


def eval_xy(fn, x, y):
    """Evaluate a t/u-independent closure D(x, y) at the points (vectorised when
    the closure accepts arrays, pointwise otherwise)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    try:
        v = fn(x, y, 0.0, 0.0)
        v = np.broadcast_to(np.asarray(v, np.float64), x.shape).copy()
        return v
    except Exception:  # noqa: BLE001 - fall back to a pointwise loop
        return np.array([float(fn(float(a), float(b), 0.0, 0.0)) for a, b in zip(x, y)])


def classify_flux(q):
    """Trace q(x, y, t, a, b, g, p) -> (qx, qy) (already bound to p): must be
    -D(x, y) (a, b).  Returns D as a function of (x, y) (arrays)."""
    probe_x = np.array([0.1234, 0.7421, 0.3317])
    probe_y = np.array([0.5531, 0.0713, 0.8822])
    r = _trace(q, 4, probe_x, probe_y, [0.0, None, None, None], [1, 2, 3])  # vars: a, b, g
    if not isinstance(r, (tuple, list)) or len(r) != 2:
        raise UnsupportedClosure("the flux function must return a pair (qx, qy)")
    qx, qy = (_as_sym(v, 4) for v in r)
    ok = all(e == (1, 0, 0, 0) for e in qx.nonzero_exponents()) and all(
        e == (0, 1, 0, 0) for e in qy.nonzero_exponents())
    if not ok:
        raise UnsupportedClosure("flux functions must have the form q = -D(x, y) * (alpha, beta); "
                                 "give u-dependent diffusion as diffusion_function")
    if not np.allclose(np.asarray(qx.coef((1, 0, 0, 0))), np.asarray(qy.coef((0, 1, 0, 0))),
                       rtol=0, atol=0):
        raise UnsupportedClosure("flux functions must use the same D for both components")

    def D(x, y):
        rr = q(np.asarray(x, np.float64), np.asarray(y, np.float64), 0.0,
               Sym.var(4, 0), Sym.var(4, 1), Sym.var(4, 2))
        c = _as_sym(rr[0], 4).coef((1, 0, 0, 0))
        return -np.broadcast_to(np.asarray(c, np.float64), np.shape(x))

    return D


def group_indices(fidx_of_locations):
    """{fidx: array of location indices} for non-negative fidx values."""
    g = defaultdict(list)
    for k, f in enumerate(fidx_of_locations):
        if f >= 0:
            g[int(f)].append(k)
    return {f: np.asarray(v, np.int64) for f, v in g.items()}
