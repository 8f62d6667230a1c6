"""Mapping of the user's closures onto the kernels' closed forms.

The reference accepts arbitrary Julia closures for the diffusion / flux /
source / condition functions (problem.jl:118-163, conditions.jl:16-20).  Device
kernels cannot call back into the host, so each closure is traced once, at
problem set-up, by calling it with symbolic stand-ins for the solution-dependent
arguments (u, t, and alpha/beta/gamma) and NumPy arrays for the coordinates.  The
trace must land on one of the forms the kernels implement:

  diffusion D(x, y, t, u, p): polynomial of degree <= 3 in u with constant
      coefficients (FVM_FLUX_D_POLY_U), d0 + c |u|^p (FVM_FLUX_D_POWER, recognised
      by evaluation), or any function of (x, y) alone
      (FVM_FLUX_D_EDGE: evaluated at every CV-edge midpoint and boundary
      half-edge point), or a `NodalDiffusion` field (FVM_FLUX_D_NODAL);
  flux q(x, y, t, a, b, g, p) -> (qx, qy): -D(x, y) (a, b) (FVM_FLUX_D_EDGE), or
      any q linear in (a, b, g) with coefficients of (x, y) (FVM_FLUX_LINEAR);
  source S(x, y, t, u, p): s(x, y) + cubic polynomial in u + c t;
  condition a(x, y, t, u, p): g(x, y) + cu u + ct t.

For an FVMSystem (u, alpha, beta, gamma are tuples over the variables):

  flux q(x, y, t, a, b, g, p) -> (qx, qy): -sum_w K[w] (a[w], b[w]), K constant;
  source S(x, y, t, u, p): s(x, y) + sum of monomials c prod_w u[w]^e_w (e_w <= 3) + c t;
  condition: g(x, y) + sum_w cu[w] u[w] + ct t, or c0 + ... + g(x, y) exp(lam t).

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
    """Polynomial in symbolic variables, times exponentials exp(rate t), with
    coefficients that are scalars or arrays over the evaluation points:
    {(exponent tuple, rate): coefficient}.  `tvar` is the index of the variable
    that stands for t (None: exp of t is not traceable)."""

    __array_priority__ = 1000

    def __init__(self, nvars, terms=None, tvar=None):
        self.nvars = nvars
        self.tvar = tvar
        self.terms = dict(terms or {})

    @classmethod
    def var(cls, nvars, k, tvar=None):
        e = [0] * nvars
        e[k] = 1
        return cls(nvars, {(tuple(e), 0.0): 1.0}, tvar)

    @classmethod
    def const(cls, nvars, c, tvar=None):
        return cls(nvars, {((0,) * nvars, 0.0): c}, tvar)

    def _new(self, terms, o=None):
        tv = self.tvar if self.tvar is not None else getattr(o, "tvar", None)
        return Sym(self.nvars, terms, tv)

    def _lift(self, o):
        if isinstance(o, Sym):
            return o
        if isinstance(o, (numbers.Number, np.ndarray, np.generic)):
            return Sym.const(self.nvars, o, self.tvar)
        return NotImplemented

    def __add__(self, o):
        o = self._lift(o)
        if o is NotImplemented:
            return o
        t = dict(self.terms)
        for k, c in o.terms.items():
            t[k] = t[k] + c if k in t else c
        return self._new(t, o)

    __radd__ = __add__

    def __neg__(self):
        return self._new({k: -c for k, c in self.terms.items()})

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
        for (e1, r1), c1 in self.terms.items():
            for (e2, r2), c2 in o.terms.items():
                k = (tuple(a + b for a, b in zip(e1, e2)), r1 + r2)
                t[k] = t[k] + c1 * c2 if k in t else c1 * c2
        return self._new(t, o)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, Sym):
            if not o.is_const():
                raise UnsupportedClosure("division by a solution-dependent quantity")
            o = o.const_value()
        return self._new({k: c / o for k, c in self.terms.items()})

    def __rtruediv__(self, o):
        if not self.is_const():
            raise UnsupportedClosure("division by a solution-dependent quantity")
        return o / self.const_value()

    def __pow__(self, k):
        if isinstance(k, Sym):
            raise UnsupportedClosure("solution-dependent exponent")
        if int(k) != k or k < 0:
            if self.is_const():
                return Sym.const(self.nvars, self.const_value() ** k, self.tvar)
            raise UnsupportedClosure("non-integer power of a solution-dependent quantity")
        r = Sym.const(self.nvars, 1.0, self.tvar)
        for _ in range(int(k)):
            r = r * self
        return r

    def __bool__(self):
        raise UnsupportedClosure("branching on a solution-dependent quantity")

    def _cmp(self, o):
        raise UnsupportedClosure("comparison of a solution-dependent quantity")

    __lt__ = __le__ = __gt__ = __ge__ = _cmp

    def _exp(self):
        """exp(c(x, y) + lam t) = exp(c) exp(lam t) for a constant lam."""
        zero = (0,) * self.nvars
        c, lam = 0.0, 0.0
        for (e, r), v in self.terms.items():
            if _is_zero(v):
                continue
            if r != 0.0:
                raise UnsupportedClosure("exp of an exponential")
            if not any(e):
                c = v
            elif self.tvar is not None and e[self.tvar] == 1 and sum(e) == 1:
                a = np.asarray(v, np.float64).ravel()
                if not np.all(a == a[0]):
                    raise UnsupportedClosure("exp(lam t) with lam depending on x, y")
                lam = float(a[0])
            else:
                raise UnsupportedClosure("exp of a solution-dependent quantity")
        return Sym(self.nvars, {(zero, lam): np.exp(c)}, self.tvar)

    def __array_ufunc__(self, ufunc, method, *inputs, **kw):
        if method != "__call__":
            return NotImplemented
        if ufunc is np.exp and len(inputs) == 1 and isinstance(inputs[0], Sym) and not inputs[0].is_const():
            return inputs[0]._exp()
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
            return Sym.const(self.nvars, ufunc(*vals, **kw), self.tvar)
        a = inputs[0]
        b = inputs[1] if len(inputs) > 1 else None
        lift = (lambda v: v if isinstance(v, Sym) else Sym.const(self.nvars, v, self.tvar))
        if ufunc is np.add:
            return lift(a) + b
        if ufunc is np.subtract:
            return lift(a) - b
        if ufunc is np.multiply:
            return lift(a) * b
        if ufunc is np.negative:
            return -a
        if ufunc is np.true_divide:
            return lift(a) / b
        return NotImplemented

    def is_const(self):
        return all(not any(e) and r == 0.0 for (e, r), c in self.terms.items() if not _is_zero(c))

    def const_value(self):
        return self.terms.get(((0,) * self.nvars, 0.0), 0.0)

    def coef(self, e, rate=0.0):
        return self.terms.get((tuple(e), float(rate)), 0.0)

    def nonzero_exponents(self):
        """Exponent tuples of the nonzero terms; a term with an exp(rate t) factor
        shows as its exponents followed by the rate."""
        return [e if r == 0.0 else e + (r,) for (e, r), c in self.terms.items() if not _is_zero(c)]


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


def _trace(fn, nvars, x, y, extra_args, sym_positions, tvar=None):
    """Call fn(x, y, *args) with Sym variables at sym_positions of args (variable
    tvar, if given, stands for t: exp of it is traced)."""
    args = list(extra_args)
    for k, pos in enumerate(sym_positions):
        args[pos] = Sym.var(nvars, k, tvar)
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
    """a(x, y, t, u) over the points (x, y): g(x, y) + cu u + ct t, or
    c0 + cu u + ct t + g(x, y) exp(lam t).  Returns dict(g, c0, cu, ct, lam)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = len(x)
    if n == 0:
        return dict(g=np.zeros(0), c0=0.0, cu=0.0, ct=0.0, lam=0.0)
    r = _as_sym(_trace(fn, 2, x, y, [None, None], [0, 1], tvar=0), 2)  # vars: t, u
    out = _condition_from_sym(r, 2, n)
    out["cu"] = out["cu"][0]
    return out


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


def _power_law(fn):
    """D(x, y, t, u) = d0 + c |u|^p with p > 0, recognised by evaluation (the symbolic
    stand-in has no abs or non-integer powers): D must not depend on (x, y, t), and
    d0 = D(0), c = D(1) - d0, p = log2((D(2) - d0) / c) must reproduce D at every
    probe, negative u included, to 1e-13.  Returns (d0, c, p) or None."""
    us = (0.0, 1.0, 2.0, 0.5, 3.7, 0.013, -1.3, -2.0, -0.25)
    pts = ((0.37, 0.61, 0.0), (-1.9, 2.3, 0.0), (0.37, 0.61, 1.7))

    def fits(p):
        return all(abs(v - (d0 + c * abs(u) ** p)) <= 1e-13 * max(abs(v), abs(d0), abs(c))
                   for u, v in zip(us, vals))

    try:
        vals = []
        for u in us:
            with np.errstate(all="ignore"):
                row = [fn(np.float64(x), np.float64(y), np.float64(t), np.float64(u)) for x, y, t in pts]
            row = [complex(v) for v in row]
            if any(v.imag != 0 or not np.isfinite(v.real) for v in row) or len({v.real for v in row}) != 1:
                return None
            vals.append(row[0].real)
        d0, c = vals[0], vals[1] - vals[0]
        if c == 0.0 or not (vals[2] - d0) / c > 0:
            return None
        p = float(np.log2((vals[2] - d0) / c))
        if not p > 0 or not fits(p):
            return None
        pr = round(p, 10)  # the exponent as written (log2 leaves an ulp or two)
        return d0, c, pr if fits(pr) else p
    except Exception:  # noqa: BLE001 - not this form
        return None


def _poly_holds(fn, poly, x, y):
    """Whether D(x, y, 0, u) = d0 + d1 u + d2 u^2 + d3 u^3 at every point (x, y), for a
    few values of u, to rounding: a trace at the probe points cannot see x, y
    dependence that is constant there (a D that changes only in part of the domain)."""
    for uk in (0.0, 0.75, -1.25, 2.5):
        v = np.broadcast_to(np.asarray(fn(x, y, 0.0, uk), np.float64), x.shape)
        terms = np.abs([poly[0], poly[1] * uk, poly[2] * uk ** 2, poly[3] * uk ** 3])
        ref = poly[0] + uk * (poly[1] + uk * (poly[2] + uk * poly[3]))
        if not np.all(np.abs(v - ref) <= 1e-13 * (terms.sum() + np.abs(v))):
            return False
    return True


def classify_diffusion(fn, x=None, y=None):
    """Trace D(x, y, t, u, p) (already bound to p).  Returns ('poly', [d0..d3]) for
    a constant-coefficient cubic in u, ('power', (d0, c, p)) for d0 + c |u|^p, or
    ('xy', None) when D depends on (x, y) only (the caller evaluates it at the
    CV-edge points).  The trace runs at fixed probe points; given the mesh's points
    (x, y), a constant-coefficient result is checked there too, and D found to vary
    over them is ('xy', None) if it does not depend on u, else unsupported (a
    closure mixing (x, y) and u goes to the compiled path)."""
    probe_x = np.array([0.1234, 0.7421, 0.3317, 0.9012])
    probe_y = np.array([0.5531, 0.0713, 0.8822, 0.2417])
    try:
        r = _as_sym(_trace(fn, 2, probe_x, probe_y, [None, None], [0, 1]), 2)  # vars: t, u
    except UnsupportedClosure:
        pw = _power_law(fn)
        if pw is None:
            raise
        return "power", pw
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
    if x is not None and len(x) and not _poly_holds(fn, poly, np.asarray(x, np.float64), np.asarray(y, np.float64)):
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
    """Trace q(x, y, t, a, b, g, p) -> (qx, qy) (already bound to p).  It must be
    linear in (alpha, beta, gamma), with coefficients that may depend on (x, y) but
    not on t.  Returns ("xy", D) when q = -D(x, y) (alpha, beta) (FVM_FLUX_D_EDGE,
    D a function of (x, y) arrays), otherwise ("linear", A) with A(x, y) -> [6, n]
    the coefficients of q = (a1 alpha + a2 beta + a3 gamma, a4 alpha + a5 beta +
    a6 gamma) (FVM_FLUX_LINEAR: advection-diffusion, anisotropic diffusion)."""
    probe_x = np.array([0.1234, 0.7421, 0.3317])
    probe_y = np.array([0.5531, 0.0713, 0.8822])
    lin = ((1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0))  # alpha, beta, gamma

    def coefs(x, y, t):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        r = _trace(q, 4, x, y, [t, None, None, None], [1, 2, 3])  # vars: a, b, g
        if not isinstance(r, (tuple, list)) or len(r) != 2:
            raise UnsupportedClosure("the flux function must return a pair (qx, qy)")
        out = np.zeros((6, len(x)))
        for c, comp in enumerate(_as_sym(v, 4) for v in r):
            for e in comp.nonzero_exponents():
                if e not in lin:
                    raise UnsupportedClosure(
                        "flux functions must be linear in (alpha, beta, gamma), e.g. "
                        "q = -D(x, y) (alpha, beta) + v(x, y) u; give u-dependent diffusion as "
                        "diffusion_function")
            for k, e in enumerate(lin):
                out[3 * c + k] = np.broadcast_to(np.asarray(comp.coef(e), np.float64), x.shape)
        return out

    A0 = coefs(probe_x, probe_y, 0.0)
    if not np.array_equal(A0, coefs(probe_x, probe_y, 1.37)):
        raise UnsupportedClosure("time-dependent flux functions are not supported by the device kernels")
    if not A0[[1, 2, 3, 5]].any() and np.array_equal(A0[0], A0[4]):
        return "xy", lambda x, y: -coefs(x, y, 0.0)[0]
    return "linear", lambda x, y: coefs(x, y, 0.0)


def group_indices(fidx_of_locations):
    """{fidx: array of location indices} for non-negative fidx values."""
    g = defaultdict(list)
    for k, f in enumerate(fidx_of_locations):
        if f >= 0:
            g[int(f)].append(k)
    return {f: np.asarray(v, np.int64) for f, v in g.items()}


# ---------------------------------------------------------------------------
# FVMSystem closures
# ---------------------------------------------------------------------------
def _call(fn, *args):
    try:
        return fn(*args)
    except UnsupportedClosure:
        raise
    except Exception as err:  # noqa: BLE001 - any failure means the form is not traceable
        raise UnsupportedClosure(f"tracing the closure failed: {type(err).__name__}: {err}") from err


def _uniform(c, n, what):
    v = _const_over(c, n)
    if v is None:
        raise UnsupportedClosure(f"{what} must not depend on x, y")
    return v


def _unit(nvars, k):
    e = [0] * nvars
    e[k] = 1
    return tuple(e)


_PROBE_X = np.array([0.1234, 0.7421, 0.3317, 0.9012])
_PROBE_Y = np.array([0.5531, 0.0713, 0.8822, 0.2417])


def classify_system_flux(q, nv):
    """Trace one variable's flux q(x, y, t, alpha, beta, gamma) (bound to p) of an
    nv-variable system: it must be -sum_w K[w] (alpha[w], beta[w]) with constant K.
    Returns K as a list of nv floats."""
    nvars = 1 + 3 * nv  # t, alpha_w, beta_w, gamma_w
    t = Sym.var(nvars, 0, 0)
    al = tuple(Sym.var(nvars, 1 + w, 0) for w in range(nv))
    be = tuple(Sym.var(nvars, 1 + nv + w, 0) for w in range(nv))
    ga = tuple(Sym.var(nvars, 1 + 2 * nv + w, 0) for w in range(nv))
    r = _call(q, _PROBE_X, _PROBE_Y, t, al, be, ga)
    if not isinstance(r, (tuple, list)) or len(r) != 2:
        raise UnsupportedClosure("the flux function must return a pair (qx, qy)")
    qx, qy = (v if isinstance(v, Sym) else Sym.const(nvars, v) for v in r)
    n = len(_PROBE_X)
    kx, ky = [0.0] * nv, [0.0] * nv
    msg = ("system flux functions must have the form q = -sum_w K[w] (alpha[w], beta[w]) "
           "with constant K")
    for comp, K, off in ((qx, kx, 1), (qy, ky, 1 + nv)):
        for e in comp.nonzero_exponents():
            w = next((w for w in range(nv) if e == _unit(nvars, off + w)), None)
            if w is None:
                raise UnsupportedClosure(msg)
            K[w] = -_uniform(comp.coef(e), n, "the flux coefficients")
    if kx != ky:
        raise UnsupportedClosure(msg + " (the same K for both components)")
    return kx


def classify_system_source(S, nv, x, y):
    """Trace one variable's source S(x, y, t, u) (bound to p) over the points (x, y).
    Returns (s_node or None, [(exponents, coef)], st)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = len(x)
    nvars = 1 + nv  # t, u_w
    t = Sym.var(nvars, 0, 0)
    u = tuple(Sym.var(nvars, 1 + w, 0) for w in range(nv))
    r = _call(S, x, y, t, u)
    r = r if isinstance(r, Sym) else Sym.const(nvars, r)
    terms, st, s_node = [], 0.0, None
    for e in r.nonzero_exponents():
        if len(e) != nvars:
            raise UnsupportedClosure("time-exponential sources are not supported")
        if e[0] == 1 and not any(e[1:]):
            st = _uniform(r.coef(e), n, "the t coefficient of the source")
        elif e[0] != 0:
            raise UnsupportedClosure("the source must be s(x, y) + polynomial in u + c*t")
        elif not any(e[1:]):
            c = _const_over(r.coef(e), n)
            if c is None:
                s_node = np.broadcast_to(np.asarray(r.coef(e), np.float64), (n,)).copy()
            else:
                terms.append(((0,) * nv, c))
        else:
            if max(e[1:]) > 3:
                raise UnsupportedClosure("source exponents above 3 are not supported")
            terms.append((tuple(e[1:]), _uniform(r.coef(e), n, "the u coefficients of the source")))
    return s_node, terms, st


def _condition_from_sym(r, nvars, n):
    """g, c0, cu (one per u variable), ct, lam of a traced condition with variables
    (t, u_0, ...): affine in t and u plus at most one g(x, y) exp(lam t) term."""
    zero = (0,) * nvars
    cu, ct, lam = [0.0] * (nvars - 1), 0.0, 0.0
    g_exp = None
    for k, c in r.terms.items():
        e, rate = k
        if _is_zero(c):
            continue
        if rate != 0.0:
            if e != zero or g_exp is not None:
                raise UnsupportedClosure("conditions may hold one term g(x, y) exp(lam t)")
            lam, g_exp = rate, c
        elif e == zero:
            continue
        elif e == _unit(nvars, 0):
            ct = _uniform(c, n, "the t coefficient of a condition")
        else:
            w = next((w for w in range(nvars - 1) if e == _unit(nvars, 1 + w)), None)
            if w is None:
                raise UnsupportedClosure("conditions must be affine in t and u")
            cu[w] = _uniform(c, n, "the u coefficients of a condition")
    c_const = r.coef(zero)
    if g_exp is None:
        g = np.broadcast_to(np.asarray(c_const, np.float64), (n,)).copy()
        return dict(g=g, c0=0.0, cu=cu, ct=ct, lam=0.0)
    c0 = _uniform(c_const, n, "with a g(x, y) exp(lam t) term, the constant part of a condition")
    g = np.broadcast_to(np.asarray(g_exp, np.float64), (n,)).copy()
    return dict(g=g, c0=c0, cu=cu, ct=ct, lam=lam)


def classify_system_condition(fn, nv, x, y):
    """Trace a condition a(x, y, t, u) (bound to p) of an nv-variable system over
    the points (x, y): g + sum_w cu[w] u[w] + ct t, where g is g(x, y), or
    c0 + g(x, y) exp(lam t).  Returns dict(g, c0, cu, ct, lam)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = len(x)
    nvars = 1 + nv
    if n == 0:
        return dict(g=np.zeros(0), c0=0.0, cu=[0.0] * nv, ct=0.0, lam=0.0)
    t = Sym.var(nvars, 0, 0)
    u = tuple(Sym.var(nvars, 1 + w, 0) for w in range(nv))
    r = _call(fn, x, y, t, u)
    r = r if isinstance(r, Sym) else Sym.const(nvars, r)
    return _condition_from_sym(r, nvars, n)
