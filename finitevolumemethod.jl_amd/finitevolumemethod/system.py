"""FVMSystem (reference src/problem.jl:196-413): nv coupled FVMProblems on one mesh.

The reference evaluates every variable's flux from one call per CV edge, keeps the
conditions of each variable separately and numbers their functions in one merged
list (merge_problem_conditions, problem.jl:368-400).  Here the closures are traced
once onto the system kernels' forms (forms.classify_system_*) and fvm_eqs! runs in
libfvmhip.so (fvm_system_rhs).  States are the reference's matrices u[var, i]
(shape (nv, n_points)); on the device they are stored column-major, u[var + nv i],
as a Julia Matrix is.
"""
import ctypes

import numpy as np

from . import _lib, forms
from .problem import AbstractFVMProblem, FVMProblem, SteadyFVMProblem, _half_edge_points

FLUX_ERROR = """
The flux function errored when evaluated. Please recheck your specification of the function.
If any of your problems have been defined in terms of a diffusion function D(x, y, t, u, p)
rather than a flux function, then you need to instead provide a flux function q(x, y, t, α, β, γ, p),
recalling the relationship between the two:

    q(x, y, t, α, β, γ, p) = -D(x, y, t, α*x + β*y + γ, p) .* (α, β)

where p is the same argument for both functions.
"""


class InvalidFluxError(TypeError):
    """problem.jl:277-281: a problem of the system has no usable flux function."""

    def __init__(self):
        super().__init__(FLUX_ERROR)


class FVMSystem(AbstractFVMProblem):
    """FVMSystem(prob1, prob2, ..., probN) (problem.jl:359-366).

    Every probi is the FVMProblem of variable i; its closures take u as the tuple of
    all variables and its flux function takes tuples alpha, beta, gamma.  Solutions
    are matrices (nv, n_points), row i holding variable i."""

    def __init__(self, *probs):
        if len(probs) == 0:
            raise ValueError("There must be at least one problem.")
        for p in probs:
            if not isinstance(p, FVMProblem):
                raise TypeError("FVMSystem takes FVMProblems")
        mesh = probs[0].mesh
        if not all(p.mesh is mesh for p in probs):
            raise AssertionError("All problems must have the same mesh.")
        if not all(p.initial_time == probs[0].initial_time for p in probs):
            raise AssertionError("All problems must have the same initial time.")
        if not all(p.final_time == probs[0].final_time for p in probs):
            raise AssertionError("All problems must have the same final time.")
        self.mesh = mesh
        self.problems = tuple(probs)
        self.initial_condition = np.stack([p.initial_condition for p in probs])
        self.initial_time = probs[0].initial_time
        self.final_time = probs[0].final_time
        self.conditions = tuple(p.conditions for p in probs)
        num = [len(p.conditions.functions) for p in probs]
        self.cnum_fncs = tuple(int(v) for v in np.cumsum([0] + num[:-1]))
        self.functions = tuple(f for p in probs for f in p.conditions.functions)
        self.source_functions = tuple((p.source_function, p.source_parameters) for p in probs)
        self.flux_functions = tuple((p.user_flux_function, p.flux_parameters) for p in probs)
        self._check_flux_function()
        self._device = None

    @property
    def n_vars(self):
        return len(self.problems)

    def _check_flux_function(self):
        """problem.jl:283-302: evaluate the fluxes once at the first triangle's centroid."""
        if any(q is None for q, _ in self.flux_functions):
            raise InvalidFluxError()
        tri = self.mesh.triangulation
        i, j, k = (int(v) for v in tri.triangles[0])
        P = self.mesh.props_array[0]
        cx = (tri.x[i] + tri.x[j] + tri.x[k]) / 3
        cy = (tri.y[i] + tri.y[j] + tri.y[k]) / 3
        u = self.initial_condition
        al = tuple(P[0] * u[v, i] + P[1] * u[v, j] + P[2] * u[v, k] for v in range(self.n_vars))
        be = tuple(P[3] * u[v, i] + P[4] * u[v, j] + P[5] * u[v, k] for v in range(self.n_vars))
        ga = tuple(P[6] * u[v, i] + P[7] * u[v, j] + P[8] * u[v, k] for v in range(self.n_vars))
        try:
            self.eval_flux_function(cx, cy, self.initial_time, al, be, ga)
        except (TypeError, IndexError) as err:
            raise InvalidFluxError() from err

    # problem.jl:304-354 accessors
    def map_fidx(self, fidx, var):
        return fidx + self.cnum_fncs[var]

    def get_conditions(self, var):
        return self.conditions[var]

    def eval_condition_fnc(self, fidx, var, x, y, t, u):
        return self.functions[self.map_fidx(fidx, var)](x, y, t, u)

    def eval_source_fnc(self, var, x, y, t, u):
        S, p = self.source_functions[var]
        return S(x, y, t, u, p)

    def eval_flux_function(self, x, y, t, a, b, g):
        """Every variable's (qx, qy) (eval_all_fncs_in_tuple)."""
        return tuple(q(x, y, t, a, b, g, p) for q, p in self.flux_functions)

    def has_dirichlet_nodes(self, var=None):
        if var is not None:
            return self.conditions[var].has_dirichlet_nodes()
        return any(c.has_dirichlet_nodes() for c in self.conditions)

    def get_point(self, i):
        return self.mesh.get_point(i)

    def get_volume(self, i):
        return self.mesh.get_volume(i)

    def device_context(self):
        """The traced, device-resident form of this system (built once)."""
        if self._device is None:
            self._device = SystemContext(self.mesh, system_physics(self))
        return self._device

    # column-major device layout u[var + nv i] <-> matrices (nv, n_points)
    def flat(self, U):
        U = np.asarray(U, np.float64)
        if U.shape != self.initial_condition.shape:
            raise ValueError(f"expected a state of shape {self.initial_condition.shape}")
        return np.ascontiguousarray(U.T).ravel()

    def unflat(self, u):
        return np.ascontiguousarray(np.asarray(u).reshape(-1, self.n_vars).T)

    def __repr__(self):
        return (f"FVMSystem with {self.n_vars} equations and time span "
                f"({self.initial_time}, {self.final_time})")


def system_physics(sys):
    """Trace an FVMSystem's closures onto the fvm_system_desc forms."""
    tri = sys.mesh.triangulation
    nv = sys.n_vars
    N = tri.num_points()
    B = len(tri.boundary_edges)
    K = np.zeros((nv, nv))
    for v, (q, qp) in enumerate(sys.flux_functions):
        K[v] = forms.classify_system_flux(lambda x, y, t, a, b, g, q=q, qp=qp: q(x, y, t, a, b, g, qp), nv)
    terms, s_node, st = [], None, np.zeros(nv)
    for v, (S, Sp) in enumerate(sys.source_functions):
        sn, tv, st[v] = forms.classify_system_source(lambda x, y, t, u, S=S, Sp=Sp: S(x, y, t, u, Sp),
                                                      nv, tri.x, tri.y)
        terms += [(v, e, c) for e, c in tv]
        if sn is not None:
            if s_node is None:
                s_node = np.zeros((nv, N))
            s_node[v] = sn
    if len(terms) > 24:
        raise forms.UnsupportedClosure("system sources may hold at most 24 monomials in total")
    nfn = len(sys.functions)
    kind = np.zeros((nv, N), np.int8)
    fidx = np.full((nv, N), -1, np.int32)
    bneu = np.full((nv, B), -1, np.int32)
    c0, ct, lam = np.zeros(max(nfn, 1)), np.zeros(max(nfn, 1)), np.zeros(max(nfn, 1))
    cu = np.zeros((max(nfn, 1), nv))
    node_g = np.zeros((nv, N))
    bedge_g = np.zeros((nv, 2 * B))
    mx, my = _half_edge_points(tri)
    seen = {}

    def record(f, r):
        key = (r["c0"], tuple(r["cu"]), r["ct"], r["lam"])
        if f in seen and seen[f] != key:
            raise forms.UnsupportedClosure(f"condition function {f} traced differently at two places")
        seen[f] = key
        c0[f], cu[f], ct[f], lam[f] = r["c0"], r["cu"], r["ct"], r["lam"]

    for v, conds in enumerate(sys.conditions):
        k_v, f_v, b_v = conds.flatten(tri)
        off = sys.cnum_fncs[v]
        kind[v] = k_v
        fidx[v] = np.where(k_v != 0, f_v + off, -1)
        bneu[v] = np.where(b_v >= 0, b_v + off, -1)
        for f, nodes in forms.group_indices(np.where(k_v != 0, f_v, -1)).items():
            r = forms.classify_system_condition(conds.functions[f], nv, tri.x[nodes], tri.y[nodes])
            record(f + off, r)
            node_g[v, nodes] = r["g"]
        for f, edges in forms.group_indices(b_v).items():
            pts = np.stack([2 * edges, 2 * edges + 1], 1).ravel()
            r = forms.classify_system_condition(conds.functions[f], nv, mx[pts], my[pts])
            record(f + off, r)
            bedge_g[v, pts] = r["g"]
    return dict(n_vars=nv, flux_k=K, terms=terms, s_node=s_node, st=st, node_kind=kind,
                node_fidx=fidx, bedge_neumann_fidx=bneu, n_functions=nfn, fn_c0=c0, fn_cu=cu,
                fn_ct=ct, fn_lam=lam, node_g=node_g if np.any(node_g) else None,
                bedge_g=bedge_g if np.any(bedge_g) else None)


def system_desc(ph):
    """fvm_system_desc of traced system physics; returns (desc, arrays it points into)."""
    keep = []

    def arr(a, dt=np.float64):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a

    nv = ph["n_vars"]
    d = _lib.SystemDesc()
    d.n_vars = nv
    d.flux_k = _lib.ptr(arr(ph["flux_k"]))
    terms = ph["terms"]
    d.n_terms = len(terms)
    if terms:
        d.term_var = _lib.ptr(arr([v for v, _, _ in terms], np.int32), _lib._i32p)
        d.term_exp = _lib.ptr(arr([e for _, e, _ in terms], np.int32), _lib._i32p)
        d.term_coef = _lib.ptr(arr([c for _, _, c in terms]))
    d.s_node = _lib.ptr(arr(ph["s_node"]))
    d.st = _lib.ptr(arr(ph["st"]))
    d.node_kind = _lib.ptr(arr(ph["node_kind"], np.int8), _lib._i8p)
    d.node_fidx = _lib.ptr(arr(ph["node_fidx"], np.int32), _lib._i32p)
    d.bedge_neumann_fidx = _lib.ptr(arr(ph["bedge_neumann_fidx"], np.int32), _lib._i32p)
    d.n_functions = ph["n_functions"]
    d.fn_c0 = _lib.ptr(arr(ph["fn_c0"]))
    d.fn_cu = _lib.ptr(arr(ph["fn_cu"]))
    d.fn_ct = _lib.ptr(arr(ph["fn_ct"]))
    d.fn_lam = _lib.ptr(arr(ph["fn_lam"]))
    d.node_g = _lib.ptr(arr(ph["node_g"]))
    d.bedge_g = _lib.ptr(arr(ph["bedge_g"]))
    return d, keep


class SystemContext:
    """An fvm_system: mesh + traced system physics on the GPU."""

    def __init__(self, mesh, ph):
        self.mesh = mesh
        self.n_vars = ph["n_vars"]
        self.n = mesh.triangulation.num_points() * self.n_vars
        d, self._keep = system_desc(ph)
        h = ctypes.c_void_p()
        _lib.check(_lib.hip_lib().fvm_system_create(mesh.handle, ctypes.byref(d), ctypes.byref(h)),
                   "fvm_system_create")
        self._handle = h.value
        self.physics = ph

    @property
    def handle(self):
        return self._handle

    def rhs(self, u, du, t=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_system_rhs(self._handle, u.ptr, du.ptr, float(t), stream),
                   "fvm_system_rhs")

    def apply_dirichlet(self, u, t=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_system_dirichlet_apply(self._handle, u.ptr, float(t), stream),
                   "fvm_system_dirichlet_apply")

    def jvp(self, u, v, jv, t=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_system_jvp(self._handle, u.ptr, v.ptr, jv.ptr, float(t), stream),
                   "fvm_system_jvp")

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and _lib._hip is not None:
            _lib._hip.fvm_system_destroy(h)
            self._handle = None


def is_system(prob):
    """problem.jl:445."""
    if isinstance(prob, SteadyFVMProblem):
        prob = prob.problem
    return isinstance(prob, FVMSystem)
