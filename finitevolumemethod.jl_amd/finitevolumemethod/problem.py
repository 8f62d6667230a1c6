"""FVMProblem and SteadyFVMProblem (reference src/problem.jl:96-194).

The problem keeps the user's closures (for host-side helpers such as
compute_flux) and, on first use, traces them onto the kernels' forms and builds
the device context (`DeviceContext`) that fvm_eqs! calls.
"""
import ctypes

import numpy as np

from . import _lib, forms
from .conditions import Conditions, InternalConditions


def construct_flux_function(q, D, Dp):
    """problem.jl:425-440: q, or (x, y, t, a, b, g, p) -> -D(x, y, t, a x + b y + g, Dp) (a, b)."""
    if q is not None:
        return q

    def flux(x, y, t, a, b, g, p):
        u = a * x + b * y + g
        Dval = D(x, y, t, u, Dp)
        return (-Dval * a, -Dval * b)

    flux.diffusion_function = D
    flux.diffusion_parameters = Dp
    return flux


def _zero_source(x, y, t, u, p):
    """zero(eltype(u)) (problem.jl:123); u is a tuple for the problems of an FVMSystem."""
    return 0.0 if isinstance(u, (tuple, list)) else 0.0 * u


class AbstractFVMProblem:
    pass


class FVMProblem(AbstractFVMProblem):
    """FVMProblem(mesh, boundary_conditions[, internal_conditions]; diffusion_function,
    diffusion_parameters, source_function, source_parameters, flux_function,
    flux_parameters, initial_condition, initial_time=0.0, final_time)."""

    def __init__(self, mesh, boundary_conditions, internal_conditions=None, *,
                 diffusion_function=None, diffusion_parameters=None, source_function=None,
                 source_parameters=None, flux_function=None, flux_parameters=None,
                 initial_condition, initial_time=0.0, final_time, conditions=None):
        ic = np.asarray(initial_condition, np.float64)
        if len(ic) != mesh.triangulation.num_points():
            raise AssertionError("The initial condition must have the same number of elements as the "
                                 "number of nodes in the mesh (including nodes that aren't vertices "
                                 "in the mesh itself).")
        self.mesh = mesh
        if conditions is None:
            conditions = Conditions(mesh, boundary_conditions,
                                    internal_conditions if internal_conditions is not None
                                    else InternalConditions())
        self.conditions = conditions
        self.diffusion_function = diffusion_function
        self.diffusion_parameters = diffusion_parameters
        self.flux_function = construct_flux_function(flux_function, diffusion_function,
                                                     diffusion_parameters)
        self.flux_parameters = flux_parameters
        self.user_flux_function = flux_function
        self.source_function = source_function if source_function is not None else _zero_source
        self.source_parameters = source_parameters
        self.initial_condition = ic
        self.initial_time = float(initial_time)
        self.final_time = float(final_time)
        self._device = None

    # problem.jl:1-23 accessors
    def get_point(self, i):
        return self.mesh.get_point(i)

    def get_volume(self, i):
        return self.mesh.get_volume(i)

    def eval_flux_function(self, x, y, t, a, b, g):
        return self.flux_function(x, y, t, a, b, g, self.flux_parameters)

    def eval_source_fnc(self, x, y, t, u):
        return self.source_function(x, y, t, u, self.source_parameters)

    def has_dirichlet_nodes(self):
        return self.conditions.has_dirichlet_nodes()

    def device_context(self):
        """The traced, device-resident form of this problem (built once)."""
        if self._device is None:
            self._device = DeviceContext.from_problem(self)
        return self._device

    def __repr__(self):
        nv = self.mesh.triangulation.num_solid_vertices()
        return f"FVMProblem with {nv} nodes and time span ({self.initial_time}, {self.final_time})"


class SteadyFVMProblem(AbstractFVMProblem):
    """problem.jl:174-194: a problem to be solved for its steady state."""

    def __init__(self, prob):
        self.problem = prob
        self.mesh = prob.mesh

    def __repr__(self):
        return f"SteadyFVMProblem with {self.mesh.triangulation.num_solid_vertices()} nodes"


def _half_edge_points(tri):
    """Boundary half-edge points m_i, m_j of every boundary edge (control_volumes.jl:41-56),
    interleaved [2*B]."""
    be = tri.boundary_edges
    px, py = tri.x[be[:, 0]], tri.y[be[:, 0]]
    qx, qy = tri.x[be[:, 1]], tri.y[be[:, 1]]
    mijx, mijy = (px + qx) / 2, (py + qy) / 2
    mx = np.empty(2 * len(be))
    my = np.empty(2 * len(be))
    mx[0::2], my[0::2] = (px + mijx) / 2, (py + mijy) / 2
    mx[1::2], my[1::2] = (qx + mijx) / 2, (qy + mijy) / 2
    return mx, my


def physics_desc(phys_kwargs):
    """fvm_physics_desc of traced physics (see physics_of); returns (desc, arrays the
    desc points into, which must outlive its use)."""
    keep = []

    def arr(a, dt=np.float64):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a

    d = _lib.PhysicsDesc()
    d.flux_kind = phys_kwargs["flux_kind"]
    for k, v in enumerate(phys_kwargs.get("dpoly", [0.0] * 4)):
        d.dpoly[k] = v
    for k, v in enumerate(phys_kwargs.get("spoly", [0.0] * 4)):
        d.spoly[k] = v
    d.st = phys_kwargs.get("st", 0.0)
    d.d_edge = _lib.ptr(arr(phys_kwargs.get("d_edge")))
    d.d_bedge = _lib.ptr(arr(phys_kwargs.get("d_bedge")))
    d.log_d = _lib.ptr(arr(phys_kwargs.get("log_d")))
    d.s_node = _lib.ptr(arr(phys_kwargs.get("s_node")))
    d.node_kind = _lib.ptr(arr(phys_kwargs["node_kind"], np.int8), _lib._i8p)
    d.node_fidx = _lib.ptr(arr(phys_kwargs["node_fidx"], np.int32), _lib._i32p)
    d.bedge_neumann_fidx = _lib.ptr(arr(phys_kwargs["bedge_neumann_fidx"], np.int32), _lib._i32p)
    fn = phys_kwargs["fn"]
    d.n_functions = len(fn[0])
    d.fn_c0, d.fn_cu, d.fn_ct = (_lib.ptr(arr(v)) for v in fn[:3])
    d.fn_cx = _lib.ptr(arr(fn[3]) if len(fn) > 3 else None)
    d.fn_cy = _lib.ptr(arr(fn[4]) if len(fn) > 4 else None)
    d.fn_lam = _lib.ptr(arr(fn[5]) if len(fn) > 5 else None)
    d.node_g = _lib.ptr(arr(phys_kwargs.get("node_g")))
    d.bedge_g = _lib.ptr(arr(phys_kwargs.get("bedge_g")))
    src = phys_kwargs.get("jit_source")
    if src is not None:
        b = ctypes.create_string_buffer(src.encode())
        keep.append(b)
        d.jit_source = ctypes.cast(b, ctypes.c_char_p)
    return d, keep


class DeviceContext:
    """An fvm_ctx: mesh + traced physics + conditions on the GPU."""

    def __init__(self, mesh, phys_kwargs):
        self.mesh = mesh
        self.n = mesh.triangulation.num_points()
        self.flux_kind = phys_kwargs["flux_kind"]
        d, self._keep = physics_desc(phys_kwargs)
        h = ctypes.c_void_p()
        _lib.check(_lib.hip_lib().fvm_ctx_create(mesh.handle, ctypes.byref(d), ctypes.byref(h)),
                   "fvm_ctx_create")
        self._handle = h.value
        self.physics = phys_kwargs

    @classmethod
    def from_problem(cls, prob):
        return cls(prob.mesh, physics_of(prob))

    @property
    def handle(self):
        return self._handle

    def rhs(self, u, du, t=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_rhs(self._handle, u.ptr, du.ptr, float(t), stream), "fvm_rhs")

    def jvp(self, u, v, jv, t=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_jvp(self._handle, u.ptr, v.ptr, jv.ptr, float(t), stream), "fvm_jvp")

    def vjp(self, u, w, ubar, logd_bar=None, t=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_vjp(self._handle, u.ptr, w.ptr, ubar.ptr,
                                          logd_bar.ptr if logd_bar is not None else None, float(t),
                                          stream), "fvm_vjp")

    def apply_dirichlet(self, u, t=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_dirichlet_apply(self._handle, u.ptr, float(t), stream),
                   "fvm_dirichlet_apply")

    def set_log_d(self, d_log_d):
        _lib.check(_lib.hip_lib().fvm_ctx_set_log_d(self._handle, d_log_d.ptr), "fvm_ctx_set_log_d")
        self._log_d_ref = d_log_d

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and _lib._hip is not None:
            _lib._hip.fvm_ctx_destroy(h)
            self._handle = None


def conditions_physics(conds, tri):
    """Flattened conditions plus the traced condition functions
    (a = c0 + cu u + ct t + g, or + g exp(lam t), g per location)."""
    kind, fidx, bneu = conds.flatten(tri)
    nfn = len(conds.functions)
    c0 = np.zeros(max(nfn, 1))
    cu = np.zeros(max(nfn, 1))
    ct = np.zeros(max(nfn, 1))
    lam = np.zeros(max(nfn, 1))
    node_g = np.zeros(tri.num_points())
    bedge_g = np.zeros(2 * len(tri.boundary_edges))
    any_node_g = any_bedge_g = False
    seen = {}
    node_groups = forms.group_indices(np.where(kind != 0, fidx, -1))
    mx, my = _half_edge_points(tri)
    edge_groups = forms.group_indices(bneu)

    def record(f, r):
        key = (r["c0"], r["cu"], r["ct"], r["lam"])
        if f in seen and seen[f] != key:
            raise forms.UnsupportedClosure(f"condition function {f} traced differently at two places")
        seen[f] = key
        c0[f], cu[f], ct[f], lam[f] = key

    for f, nodes in node_groups.items():
        r = forms.classify_condition(conds.functions[f], tri.x[nodes], tri.y[nodes])
        record(f, r)
        node_g[nodes] = r["g"]
        any_node_g = any_node_g or bool(np.any(r["g"] != 0))
    for f, edges in edge_groups.items():
        pts = np.stack([2 * edges, 2 * edges + 1], 1).ravel()
        r = forms.classify_condition(conds.functions[f], mx[pts], my[pts])
        record(f, r)
        bedge_g[pts] = r["g"]
        any_bedge_g = any_bedge_g or bool(np.any(r["g"] != 0))
    zeros = np.zeros(max(nfn, 1))
    return dict(node_kind=kind, node_fidx=fidx, bedge_neumann_fidx=bneu,
                fn=(c0, cu, ct, zeros, zeros, lam), node_g=node_g if any_node_g else None,
                bedge_g=bedge_g if any_bedge_g else None)


def _jit_physics(prob):
    """FVM_FLUX_JIT physics: the diffusion and source closures compiled for the device
    (jit.py), for closures outside the traced forms."""
    from . import jit

    D, Dp = prob.diffusion_function, prob.diffusion_parameters
    S, Sp = prob.source_function, prob.source_parameters
    return dict(flux_kind=_lib.FLUX_JIT,
                jit_source=jit.source_of(lambda x, y, t, u: D(x, y, t, u, Dp),
                                         lambda x, y, t, u: S(x, y, t, u, Sp)))


def physics_of(prob):
    """Trace an FVMProblem's closures onto the device forms (see forms.py); a diffusion
    or source closure outside them is compiled for the device instead (jit.py)."""
    mesh = prob.mesh
    tri = mesh.triangulation
    ph = conditions_physics(prob.conditions, tri)
    D = prob.diffusion_function
    if (prob.user_flux_function is None and D is not None and not isinstance(D, forms.NodalDiffusion)):
        Dp, S, Sp = prob.diffusion_parameters, prob.source_function, prob.source_parameters
        try:
            forms.classify_diffusion(lambda x, y, t, u: D(x, y, t, u, Dp), tri.x, tri.y)
            forms.classify_source(lambda x, y, t, u: S(x, y, t, u, Sp), tri.x, tri.y)
        except forms.UnsupportedClosure:
            ph.update(_jit_physics(prob))
            return ph
    if prob.user_flux_function is None and isinstance(D, forms.NodalDiffusion):
        ph.update(flux_kind=_lib.FLUX_D_NODAL, log_d=D.log_d)
    else:
        if prob.user_flux_function is not None:
            q = prob.user_flux_function
            kind, F = forms.classify_flux(lambda x, y, t, a, b, g: q(x, y, t, a, b, g, prob.flux_parameters))
            poly = None
            Dfun = lambda x, y, t, u: F(x, y)  # noqa: E731
        else:
            if D is None:
                raise ValueError("FVMProblem needs a diffusion_function or a flux_function")
            Dp = prob.diffusion_parameters
            Dfun = lambda x, y, t, u: D(x, y, t, u, Dp)  # noqa: E731
            kind, poly = forms.classify_diffusion(Dfun, tri.x, tri.y)
        if kind == "poly":
            ph.update(flux_kind=_lib.FLUX_D_POLY_U, dpoly=poly)
        elif kind == "power":
            ph.update(flux_kind=_lib.FLUX_D_POWER, dpoly=[poly[0], poly[1], poly[2], 0.0])
        elif kind == "linear":
            # a1..a6 per CV edge [T][3][6] and per boundary half-edge point [2B][6]
            P = mesh.props_array
            ex = P[:, [9, 11, 13]].ravel()
            ey = P[:, [10, 12, 14]].ravel()
            mx, my = _half_edge_points(tri)
            ph.update(flux_kind=_lib.FLUX_LINEAR, d_edge=F(ex, ey).T.ravel(), d_bedge=F(mx, my).T.ravel(),
                      a_xy=F)
        else:
            P = mesh.props_array
            ex = P[:, [9, 11, 13]].ravel()
            ey = P[:, [10, 12, 14]].ravel()
            mx, my = _half_edge_points(tri)
            ph.update(flux_kind=_lib.FLUX_D_EDGE, d_edge=forms.eval_xy(Dfun, ex, ey),
                      d_bedge=forms.eval_xy(Dfun, mx, my), d_xy=Dfun)
    S = prob.source_function
    Sp = prob.source_parameters
    s_node, spoly, st = forms.classify_source(lambda x, y, t, u: S(x, y, t, u, Sp), tri.x, tri.y)
    ph.update(s_node=s_node, spoly=spoly, st=st)
    return ph
