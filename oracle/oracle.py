"""ctypes wrapper of libfvm_oracle.so (test infrastructure only).

Functions mirror the reference files they restate:
  geometry            geometry.jl:99-169
  merge_conditions    conditions.jl:486-552 (pure Python restatement)
  OracleProblem.rhs   main_equations.jl:28-81 and the equations/*.jl passes
  OracleProblem.rhs_cs  same, complex-step (Jacobian-vector products)
  template_assemble   specific_problems/abstract_templates.jl:73-325 and the
                      Poisson / MeanExitTime / Diffusion constructors
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ORC_PROPS = 24
FLUX_POLY_U, FLUX_NODAL, FLUX_DXY = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i8p = ctypes.POINTER(ctypes.c_int8)

FLUX_FN = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                           ctypes.c_void_p, _dp, _dp)
SOURCE_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, ctypes.c_int64, ctypes.c_void_p)
COND_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                           ctypes.c_double, ctypes.c_double, ctypes.c_void_p)
DXY_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p)
SYS_FLUX_FN = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                               ctypes.c_void_p, _dp, _dp)
SYS_SOURCE_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, _dp, ctypes.c_void_p)
SYS_COND_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, _dp, ctypes.c_void_p)


class _Mesh(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_int64), ("n_tri", ctypes.c_int64),
                ("n_bedges", ctypes.c_int64), ("x", _dp), ("y", _dp), ("tri", _i32p),
                ("props", _dp), ("cv", _dp), ("bedges", _i32p), ("badj", _i32p),
                ("is_vertex", _u8p)]


class _Cond(ctypes.Structure):
    _fields_ = [("node_kind", _i8p), ("node_fidx", _i32p), ("bedge_neumann_fidx", _i32p)]


class _Phys(ctypes.Structure):
    _fields_ = [("flux", ctypes.c_void_p), ("source", ctypes.c_void_p),
                ("cond", ctypes.c_void_p), ("ctx", ctypes.c_void_p)]


class _System(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int32), ("node_kind", _i8p), ("node_fidx", _i32p),
                ("bedge_neumann_fidx", _i32p), ("flux", ctypes.c_void_p), ("source", ctypes.c_void_p),
                ("cond", ctypes.c_void_p), ("ctx", ctypes.c_void_p)]


class _Builtin(ctypes.Structure):
    _fields_ = [("flux_kind", ctypes.c_int), ("dpoly", ctypes.c_double * 4),
                ("dscale", ctypes.c_double), ("logD", _dp), ("props", _dp), ("tri", _i32p),
                ("spoly", ctypes.c_double * 4), ("s_node", _dp), ("n_fn", ctypes.c_int32),
                ("fn_c0", _dp), ("fn_cu", _dp), ("fn_ct", _dp), ("fn_cx", _dp), ("fn_cy", _dp)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libfvm_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        L = ctypes.CDLL(path)
        L.orc_geometry.argtypes = [ctypes.c_int64, _dp, _dp, ctypes.c_int64, _i32p, _dp, _dp]
        L.orc_geometry.restype = None
        for name in ("orc_fvm_eqs_serial",):
            f = getattr(L, name)
            f.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Cond), ctypes.POINTER(_Phys),
                          _dp, ctypes.c_double, _dp]
            f.restype = None
        L.orc_fvm_eqs_parallel.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Cond),
                                           ctypes.POINTER(_Phys), _dp, ctypes.c_double, _dp,
                                           ctypes.c_int, _dp]
        L.orc_fvm_eqs_parallel.restype = None
        L.orc_fvm_scale.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Cond),
                                    ctypes.POINTER(_Phys), _dp, ctypes.c_double, _dp]
        L.orc_fvm_scale.restype = None
        L.orc_update_dirichlet.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Cond),
                                           ctypes.POINTER(_Phys), ctypes.c_double, _dp]
        L.orc_update_dirichlet.restype = None
        for name in ("orc_fvm_eqs_system", "orc_fvm_scale_system"):
            f = getattr(L, name)
            f.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_System), _dp, ctypes.c_double, _dp]
            f.restype = None
        L.orc_update_dirichlet_system.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_System),
                                                  ctypes.c_double, _dp]
        L.orc_update_dirichlet_system.restype = None
        L.orc_fvm_eqs_serial_cs.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Cond),
                                            ctypes.POINTER(_Builtin), _dp, _dp, ctypes.c_double, _dp]
        L.orc_fvm_eqs_serial_cs.restype = None
        L.orc_template_assemble.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Cond),
                                            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, _dp,
                                            _dp, _dp, _dp, _dp, _i64p, _i64p, _dp, _dp]
        L.orc_template_assemble.restype = ctypes.c_int64
        _LIB = L
    return _LIB


def _p(a, typ):
    return a.ctypes.data_as(typ) if a is not None else ctypes.cast(None, typ)


def geometry(x, y, tri):
    """FVMGeometry(tri) (geometry.jl:99-169): (cv_volumes[N], props[T, 24])."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    tri = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
    cv = np.zeros(len(x))
    props = np.zeros((len(tri), ORC_PROPS))
    lib().orc_geometry(len(x), _p(x, _dp), _p(y, _dp), len(tri), _p(tri, _i32p), _p(cv, _dp),
                       _p(props, _dp))
    return cv, props


# Reference ConditionType order (conditions.jl:36-41)
NEUMANN, DUDT, DIRICHLET, CONSTRAINED = "Neumann", "Dudt", "Dirichlet", "Constrained"


def merge_conditions(bedges, bseg, seg_types, n_internal_functions=0, internal_dirichlet=None,
                     internal_dudt=None):
    """prepare_conditions + merge_conditions! (conditions.jl:486-552).

    Returns dicts (neumann_edges, constrained_edges, dirichlet_nodes, dudt_nodes)
    mapping to 0-based function indices into (internal..., boundary...).  Boundary
    segments are visited in ascending order (the reference iterates a Dict of ghost
    vertices; the order only matters for a node shared by two Dirichlet segments
    with different functions: parity unpinned there)."""
    dirichlet = dict(internal_dirichlet or {})
    dudt = dict(internal_dudt or {})
    neumann, constrained = {}, {}
    bedges = np.asarray(bedges).reshape(-1, 2)
    for s in sorted(set(int(v) for v in np.asarray(bseg))):
        cond = seg_types[s]
        fidx = s + n_internal_functions
        for e in np.nonzero(np.asarray(bseg) == s)[0]:
            u, v = int(bedges[e, 0]), int(bedges[e, 1])
            if cond == NEUMANN:
                neumann[(u, v)] = fidx
            elif cond == CONSTRAINED:
                constrained[(u, v)] = fidx
            elif cond == DIRICHLET:
                dirichlet[u] = fidx
                dirichlet[v] = fidx
            else:
                dudt[u] = fidx
                dudt[v] = fidx
    return neumann, constrained, dirichlet, dudt


def flatten_conditions(n_points, bedges, neumann, dirichlet, dudt):
    """Per-node kind (0 none, 1 Dirichlet, 2 Dudt; Dirichlet first as in
    source_contributions.jl:5-12), per-node fidx, per-boundary-edge Neumann fidx."""
    kind = np.zeros(n_points, np.int8)
    fidx = np.full(n_points, -1, np.int32)
    for i, f in dudt.items():
        kind[i], fidx[i] = 2, f
    for i, f in dirichlet.items():
        kind[i], fidx[i] = 1, f
    bedges = np.asarray(bedges).reshape(-1, 2)
    bn = np.full(len(bedges), -1, np.int32)
    for e, (u, v) in enumerate(bedges):
        f = neumann.get((int(u), int(v)))
        if f is not None:
            bn[e] = f
    return kind, fidx, bn


class OracleProblem:
    """A mesh + flattened conditions + closures, evaluated by the C restatement.

    physics: either dict(builtin=...) with keys of orc_builtin (flux_kind, dpoly,
    dscale, logD, spoly, s_node, fn=(c0, cu, ct[, cx, cy]) arrays) or dict of
    Python callables flux(x,y,t,a,b,g,tri)->(qx,qy), source(x,y,t,u,node)->S,
    cond(fidx,x,y,t,u)->a."""

    def __init__(self, x, y, tri, bedges, badj, node_kind, node_fidx, bedge_neumann_fidx,
                 physics, cv=None, props=None, is_vertex=None):
        self.x = np.ascontiguousarray(x, np.float64)
        self.y = np.ascontiguousarray(y, np.float64)
        self.tri = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
        self.bedges = np.ascontiguousarray(bedges, np.int32).reshape(-1, 2)
        self.badj = np.ascontiguousarray(badj, np.int32)
        N = len(self.x)
        if cv is None or props is None:
            cv, props = geometry(self.x, self.y, self.tri)
        self.cv, self.props = np.ascontiguousarray(cv), np.ascontiguousarray(props)
        if is_vertex is None:
            is_vertex = np.zeros(N, np.uint8)
            is_vertex[self.tri.ravel()] = 1
        self.is_vertex = np.ascontiguousarray(is_vertex, np.uint8)
        self.kind = np.ascontiguousarray(node_kind, np.int8)
        self.fidx = np.ascontiguousarray(node_fidx, np.int32)
        self.bneu = np.ascontiguousarray(bedge_neumann_fidx, np.int32)
        self._mesh = _Mesh(N, len(self.tri), len(self.bedges), _p(self.x, _dp), _p(self.y, _dp),
                           _p(self.tri, _i32p), _p(self.props, _dp), _p(self.cv, _dp),
                           _p(self.bedges, _i32p), _p(self.badj, _i32p), _p(self.is_vertex, _u8p))
        self._cond = _Cond(_p(self.kind, _i8p), _p(self.fidx, _i32p), _p(self.bneu, _i32p))
        self._keep = []
        self._builtin = None
        if "builtin" in physics:
            self._set_builtin(physics["builtin"])
        else:
            self._set_callbacks(physics)

    def _set_builtin(self, b):
        L = lib()
        B = _Builtin()
        B.flux_kind = int(b.get("flux_kind", FLUX_POLY_U))
        for k, v in enumerate((list(b.get("dpoly", [0.0])) + [0.0] * 4)[:4]):
            B.dpoly[k] = v
        B.dscale = float(b.get("dscale", 1.0))
        for k, v in enumerate((list(b.get("spoly", [0.0])) + [0.0] * 4)[:4]):
            B.spoly[k] = v
        arrs = {}
        for name in ("logD", "s_node"):
            if b.get(name) is not None:
                arrs[name] = np.ascontiguousarray(b[name], np.float64)
                setattr(B, name, _p(arrs[name], _dp))
        B.props = _p(self.props, _dp)
        B.tri = _p(self.tri, _i32p)
        fn = b.get("fn", ([0.0], [0.0], [0.0]))
        names = ("fn_c0", "fn_cu", "fn_ct", "fn_cx", "fn_cy")
        for k, arr in enumerate(fn):
            a = np.ascontiguousarray(arr, np.float64)
            arrs[names[k]] = a
            setattr(B, names[k], _p(a, _dp))
        B.n_fn = len(arrs["fn_c0"])
        self._keep.append(arrs)
        self._builtin = B
        self._phys = _Phys(ctypes.cast(L.orc_builtin_flux, ctypes.c_void_p).value,
                           ctypes.cast(L.orc_builtin_source, ctypes.c_void_p).value,
                           ctypes.cast(L.orc_builtin_cond, ctypes.c_void_p).value,
                           ctypes.cast(ctypes.pointer(B), ctypes.c_void_p).value)

    def _set_callbacks(self, ph):
        flux, source, cond = ph["flux"], ph.get("source"), ph.get("cond")

        def _flux(x, y, t, a, b, g, tri, ctx, qx, qy):
            fx, fy = flux(x, y, t, a, b, g, tri)
            qx[0] = fx
            qy[0] = fy

        def _source(x, y, t, u, node, ctx):
            return float(source(x, y, t, u, node)) if source is not None else 0.0

        def _cond(f, x, y, t, u, ctx):
            return float(cond(f, x, y, t, u)) if cond is not None else 0.0

        cbs = (FLUX_FN(_flux), SOURCE_FN(_source), COND_FN(_cond))
        self._keep.append(cbs)
        self._phys = _Phys(ctypes.cast(cbs[0], ctypes.c_void_p).value,
                           ctypes.cast(cbs[1], ctypes.c_void_p).value,
                           ctypes.cast(cbs[2], ctypes.c_void_p).value, None)

    def rhs(self, u, t=0.0, nthreads=0):
        """fvm_eqs!(du, u, p, t): serial if nthreads == 0, else the threaded form."""
        u = np.ascontiguousarray(u, np.float64)
        du = np.zeros(len(self.x))
        L = lib()
        if nthreads <= 0:
            L.orc_fvm_eqs_serial(ctypes.byref(self._mesh), ctypes.byref(self._cond),
                                 ctypes.byref(self._phys), _p(u, _dp), t, _p(du, _dp))
        else:
            scratch = np.empty(len(self.x) * nthreads)
            L.orc_fvm_eqs_parallel(ctypes.byref(self._mesh), ctypes.byref(self._cond),
                                   ctypes.byref(self._phys), _p(u, _dp), t, _p(du, _dp),
                                   nthreads, _p(scratch, _dp))
        return du

    def scale(self, u, t=0.0):
        """Per-node sum |Q| / V + |S| (the magnitude the parity bound is relative to)."""
        u = np.ascontiguousarray(u, np.float64)
        sc = np.zeros(len(self.x))
        lib().orc_fvm_scale(ctypes.byref(self._mesh), ctypes.byref(self._cond),
                            ctypes.byref(self._phys), _p(u, _dp), t, _p(sc, _dp))
        return sc

    def rhs_cs(self, u, logD=None, t=0.0):
        """Complex-step fvm_eqs! with the built-in closures: complex u (and logD)."""
        if self._builtin is None:
            raise ValueError("complex-step evaluation needs built-in closures")
        u = np.ascontiguousarray(u, np.complex128)
        lg = np.ascontiguousarray(logD if logD is not None else np.zeros(len(self.x)), np.complex128)
        du = np.zeros(len(self.x), np.complex128)
        lib().orc_fvm_eqs_serial_cs(ctypes.byref(self._mesh), ctypes.byref(self._cond),
                                    ctypes.byref(self._builtin),
                                    u.view(np.float64).ctypes.data_as(_dp),
                                    lg.view(np.float64).ctypes.data_as(_dp), t,
                                    du.view(np.float64).ctypes.data_as(_dp))
        return du

    def update_dirichlet(self, u, t=0.0):
        u = np.array(u, np.float64)
        lib().orc_update_dirichlet(ctypes.byref(self._mesh), ctypes.byref(self._cond),
                                   ctypes.byref(self._phys), t, _p(u, _dp))
        return u

    def template_assemble(self, kind, D, f_node=None, neu_a=None, dir_val=None, dudt_val=None,
                          r_node=None):
        """COO triplets + b of the template operator (kind 0 Poisson / Laplace,
        1 MeanExitTime, 2 Diffusion / linear reaction-diffusion with the diagonal
        r_node).  D: a float (constant), the string "exp_sincos:<scale>", or a
        Python callable D(x, y)."""
        L = lib()
        keep = []
        if isinstance(D, (int, float)):
            c = ctypes.c_double(float(D))
            fn, ctx = ctypes.cast(L.orc_dxy_const, ctypes.c_void_p).value, ctypes.addressof(c)
            keep.append(c)
        elif isinstance(D, str) and D.startswith("exp_sincos"):
            c = ctypes.c_double(float(D.split(":")[1]) if ":" in D else 1.0)
            fn, ctx = ctypes.cast(L.orc_dxy_exp_sincos, ctypes.c_void_p).value, ctypes.addressof(c)
            keep.append(c)
        else:
            cb = DXY_FN(lambda x, y, ctx: float(D(x, y)))
            keep.append(cb)
            fn, ctx = ctypes.cast(cb, ctypes.c_void_p).value, None
        N = len(self.x)
        arr = lambda a, n: np.ascontiguousarray(a if a is not None else np.zeros(n), np.float64)
        f_node, neu_a = arr(f_node, N), arr(neu_a, 2 * len(self.bedges))
        dir_val, dudt_val = arr(dir_val, N), arr(dudt_val, N)
        r_arr = None if r_node is None else arr(r_node, N)
        args = (ctypes.byref(self._mesh), ctypes.byref(self._cond), kind, fn, ctx,
                _p(f_node, _dp), _p(neu_a, _dp), _p(dir_val, _dp), _p(dudt_val, _dp),
                None if r_arr is None else _p(r_arr, _dp))
        n = L.orc_template_assemble(*args, None, None, None, None)
        rows = np.empty(n, np.int64)
        cols = np.empty(n, np.int64)
        vals = np.empty(n)
        b = np.empty(N)
        L.orc_template_assemble(*args, _p(rows, _i64p), _p(cols, _i64p), _p(vals, _dp), _p(b, _dp))
        return rows, cols, vals, b


class OracleSystem:
    """An FVMSystem (problem.jl:196-413) evaluated by the C restatement.

    node_kind / node_fidx: [nv, n_points]; bedge_neumann_fidx: [nv, n_bedges] (function
    indices into the merged list); flux(x, y, t, a, b, g) -> (qx[nv], qy[nv]) with
    a, b, g arrays over the variables; source(var, x, y, t, u) and cond(fidx, x, y, t, u)
    with u the vector of every variable's value.  States are [n_points, nv] (C order),
    the reference's du[var, i] layout."""

    def __init__(self, x, y, tri, bedges, badj, node_kind, node_fidx, bedge_neumann_fidx, flux,
                 source, cond):
        base = OracleProblem(x, y, tri, bedges, badj, np.zeros(len(x), np.int8),
                             np.full(len(x), -1, np.int32), np.full(len(np.asarray(bedges).reshape(-1, 2)), -1, np.int32),
                             dict(flux=lambda *a: (0.0, 0.0)))
        self.base = base
        self.nv = nv = int(np.asarray(node_kind).shape[0])
        if nv > 8:
            raise ValueError("at most 8 variables")
        self.kind = np.ascontiguousarray(node_kind, np.int8)
        self.fidx = np.ascontiguousarray(node_fidx, np.int32)
        self.bneu = np.ascontiguousarray(bedge_neumann_fidx, np.int32)

        def _flux(x, y, t, a, b, g, ctx, qx, qy):
            av = np.ctypeslib.as_array(a, (nv,)).copy()
            bv = np.ctypeslib.as_array(b, (nv,)).copy()
            gv = np.ctypeslib.as_array(g, (nv,)).copy()
            fx, fy = flux(x, y, t, av, bv, gv)
            for v in range(nv):
                qx[v] = fx[v]
                qy[v] = fy[v]

        def _source(var, x, y, t, u, ctx):
            return float(source(var, x, y, t, np.ctypeslib.as_array(u, (nv,)).copy()))

        def _cond(f, x, y, t, u, ctx):
            return float(cond(f, x, y, t, np.ctypeslib.as_array(u, (nv,)).copy()))

        self._cbs = (SYS_FLUX_FN(_flux), SYS_SOURCE_FN(_source), SYS_COND_FN(_cond))
        self._sys = _System(nv, _p(self.kind, _i8p), _p(self.fidx, _i32p), _p(self.bneu, _i32p),
                            ctypes.cast(self._cbs[0], ctypes.c_void_p).value,
                            ctypes.cast(self._cbs[1], ctypes.c_void_p).value,
                            ctypes.cast(self._cbs[2], ctypes.c_void_p).value, None)

    def rhs(self, u, t=0.0):
        u = np.ascontiguousarray(u, np.float64).reshape(-1, self.nv)
        du = np.zeros_like(u)
        lib().orc_fvm_eqs_system(ctypes.byref(self.base._mesh), ctypes.byref(self._sys), _p(u, _dp), t,
                                 _p(du, _dp))
        return du

    def scale(self, u, t=0.0):
        u = np.ascontiguousarray(u, np.float64).reshape(-1, self.nv)
        sc = np.zeros_like(u)
        lib().orc_fvm_scale_system(ctypes.byref(self.base._mesh), ctypes.byref(self._sys), _p(u, _dp), t,
                                   _p(sc, _dp))
        return sc

    def update_dirichlet(self, u, t=0.0):
        u = np.array(u, np.float64, order="C").reshape(-1, self.nv)
        lib().orc_update_dirichlet_system(ctypes.byref(self.base._mesh), ctypes.byref(self._sys), t,
                                          _p(u, _dp))
        return u


def directed_edge_triangles(tri):
    """{(a, b): triangle index} for every directed edge a -> b of the counter-clockwise
    triangles (what get_adjacent(tri, a, b) answers with the third vertex)."""
    m = {}
    for T, (a, b, c) in enumerate(np.asarray(tri).reshape(-1, 3)):
        a, b, c = int(a), int(b), int(c)
        m[(a, b)] = T
        m[(b, c)] = T
        m[(c, a)] = T
    return m


def compute_flux(x, y, tri, props, edge_tri, u, i, j, t, flux):
    """compute_flux (problem.jl:458-487) restated for one edge.  u [n_points, nv] (nv
    columns; a 1-D u is one variable); flux(x, y, t, alpha, beta, gamma, T) with arrays
    over the variables (T: the triangle used) returns (qx, qy) arrays.  Returns q . n
    per variable."""
    tri = np.asarray(tri).reshape(-1, 3)
    u = np.asarray(u, np.float64)
    u2 = u[:, None] if u.ndim == 1 else u
    px, py, qx, qy = x[i], y[i], x[j], y[j]
    ex, ey = qx - px, qy - py
    ell = np.sqrt(ex * ex + ey * ey)
    nx, ny = ey / ell, -ex / ell
    T = edge_tri.get((j, i))  # the triangle in the direction of the normal
    if T is None:
        T = edge_tri[(i, j)]  # boundary edge with the domain on its left
    a, b, c = tri[T]
    s = props[T]
    al = s[0] * u2[a] + s[1] * u2[b] + s[2] * u2[c]
    be = s[3] * u2[a] + s[4] * u2[b] + s[5] * u2[c]
    ga = s[6] * u2[a] + s[7] * u2[b] + s[8] * u2[c]
    fx, fy = flux((px + qx) / 2, (py + qy) / 2, t, al, be, ga, T)
    return nx * np.asarray(fx) + ny * np.asarray(fy)


def pl_interpolate(tri, props, edge_tri, u, T, px, py):
    """pl_interpolate (utils.jl:23-27) restated: the stored triangle holding the given
    rotation (i, j, k), alpha x + beta y + gamma per variable."""
    tri = np.asarray(tri).reshape(-1, 3)
    u = np.asarray(u, np.float64)
    u2 = u[:, None] if u.ndim == 1 else u
    i, j, k = (int(v) for v in T)
    Ti = edge_tri[(i, j)]
    a, b, c = tri[Ti]
    if {int(a), int(b), int(c)} != {i, j, k}:
        raise KeyError((i, j, k))
    s = props[Ti]
    al = s[0] * u2[a] + s[1] * u2[b] + s[2] * u2[c]
    be = s[3] * u2[a] + s[4] * u2[b] + s[5] * u2[c]
    ga = s[6] * u2[a] + s[7] * u2[b] + s[8] * u2[c]
    return al * px + be * py + ga
