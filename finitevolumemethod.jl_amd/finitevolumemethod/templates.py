"""Linear templates (reference src/specific_problems/*.jl) assembled and solved on
the GPU.

PoissonsEquation, LaplacesEquation and MeanExitTimeProblem are steady linear
systems A u = b (the reference hands LinearProblem(sparse(A), b) to LinearSolve);
here A and b are assembled by fvm_template_create in the CSR pattern of
jacobian_sparsity and solved by fvm_template_solve (Jacobi-preconditioned CG on
the volume-scaled operator, FP64 or FP32-inner iterative refinement, or
BiCGStab).  DiffusionEquation and LinearReactionDiffusionEquation are linear ODEs
du/dt = A u + b (the reference's MatrixOperator of the augmented [A b; 0 0]
system) integrated on the device with fvm_template_apply as the right-hand side.

Closures are evaluated on the host where the reference evaluates them: D(x, y, p)
at the CV-edge midpoints and boundary half-edge points, conditions a(x, y, nothing,
nothing, p) at the nodes and half-edge points, sources f(x, y, p) at the nodes.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .conditions import Conditions, InternalConditions
from .problem import _half_edge_points


@dataclass
class CG:
    """FP64 Jacobi-preconditioned CG on M = -diag(V) A (free rows and columns)."""
    check_every: int = 32


@dataclass
class MixedCG:
    """FP32 Jacobi-PCG inside FP64 iterative refinement: the residual b - A x, the
    solution and the convergence test stay FP64.

    update="reliable" (default): one CG sequence; each time its residual has dropped
    by inner_rtol (0: 0.1) the FP32 correction is added to x and the FP32 residual
    replaced by the FP64 one, keeping the search direction.  update="restart": each
    refinement step is a fresh FP32 CG solve that reduces the residual by
    inner_rtol (0: 1e-4)."""
    update: str = "reliable"
    inner_rtol: float = 0.0
    check_every: int = 32

    def __post_init__(self):
        if self.update not in ("reliable", "restart"):
            raise ValueError(f"MixedCG update must be 'reliable' or 'restart', not {self.update!r}")


@dataclass
class BiCGStab:
    """FP64 BiCGStab on M with Jacobi right preconditioning."""
    check_every: int = 32


_METHOD = {CG: _lib.SOLVE_CG, MixedCG: _lib.SOLVE_CG_MIXED, BiCGStab: _lib.SOLVE_BICGSTAB}


@dataclass
class LinearSolution:
    """What solve(::LinearProblem) returns: u and a retcode, plus solver statistics."""
    u: np.ndarray
    retcode: str = "Success"
    resid: float = 0.0
    iters: int = 0
    stats: dict = field(default_factory=dict)


def _eval_static(f, xs, ys):
    """a(x, y, nothing, nothing) of a ParametrisedFunction (vectorised when it can be)."""
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    if len(xs) == 0:
        return np.zeros(0)
    try:
        v = np.broadcast_to(np.asarray(f(xs, ys, None, None), np.float64), xs.shape).copy()
        return v
    except Exception:  # noqa: BLE001 - fall back to a pointwise loop
        return np.array([float(f(float(x), float(y), None, None)) for x, y in zip(xs, ys)])


def _eval_xyp(fn, p, xs, ys):
    """fn(x, y, p) at the points (vectorised when it can be)."""
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    if len(xs) == 0:
        return np.zeros(0)
    try:
        return np.broadcast_to(np.asarray(fn(xs, ys, p), np.float64), xs.shape).copy()
    except Exception:  # noqa: BLE001
        return np.array([float(fn(float(x), float(y), p)) for x, y in zip(xs, ys)])


_SPMV_FORMATS = {"auto": 0, "csr": 1, "ell": 2}


class DeviceTemplate:
    """An fvm_template: A and b assembled on the GPU."""

    def __init__(self, mesh, conditions, kind, diffusion_function, diffusion_parameters,
                 f_node=None, r_node=None, condition_values=True):
        tri = mesh.triangulation
        node_kind, fidx, bneu = conditions.flatten(tri)
        P = mesh.props_array
        ex = P[:, [9, 11, 13]].ravel()
        ey = P[:, [10, 12, 14]].ravel()
        d_edge = _eval_xyp(diffusion_function, diffusion_parameters, ex, ey)
        mx, my = _half_edge_points(tri)
        d_bedge = _eval_xyp(diffusion_function, diffusion_parameters, mx, my)
        n = tri.num_points()
        B = len(tri.boundary_edges)
        funcs = conditions.functions
        neu_a = np.zeros(2 * B)
        cond_val = np.zeros(n)
        if condition_values:  # b only: the operator A does not depend on them
            for f in np.unique(bneu[bneu >= 0]):
                edges = np.nonzero(bneu == f)[0]
                pts = np.stack([2 * edges, 2 * edges + 1], 1).ravel()
                neu_a[pts] = _eval_static(funcs[f], mx[pts], my[pts])
            for f in np.unique(fidx[node_kind != 0]):
                nodes = np.nonzero((node_kind != 0) & (fidx == f))[0]
                cond_val[nodes] = _eval_static(funcs[f], tri.x[nodes], tri.y[nodes])
        keep = []

        def arr(a, dt=np.float64, p=_lib._dp):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return _lib.ptr(a, p)

        d = _lib.TemplateDesc()
        d.kind = kind
        d.d_edge = arr(d_edge)
        d.d_bedge = arr(d_bedge)
        d.node_kind = arr(node_kind, np.int8, _lib._i8p)
        d.bedge_neumann_fidx = arr(bneu, np.int32, _lib._i32p)
        d.f_node = arr(f_node)
        d.r_node = arr(r_node)
        d.neu_a = arr(neu_a)
        d.dir_val = arr(np.where(node_kind == _lib.NODE_DIRICHLET, cond_val, 0.0))
        d.dudt_val = arr(np.where(node_kind == _lib.NODE_DUDT, cond_val, 0.0))
        h = ctypes.c_void_p()
        _lib.check(_lib.hip_lib().fvm_template_create(mesh.handle, ctypes.byref(d), ctypes.byref(h)),
                   "fvm_template_create")
        self._handle = h.value
        self._mesh = mesh  # the template reads the mesh's device geometry
        self.kind = kind
        self.n = n
        self.node_kind = node_kind
        self.cond_val = cond_val
        nn, nnz = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.hip_lib().fvm_template_info(self._handle, ctypes.byref(nn), ctypes.byref(nnz)),
                   "fvm_template_info")
        self.nnz = nnz.value
        self._csr = None

    @classmethod
    def from_csr(cls, mesh, row_ptr, col, val, b, fixed):
        """A x = b from host CSR arrays (fvm_template_create_csr): rows with fixed != 0
        are x_i = b_i; the solvers and products of an assembled template apply."""
        self = cls.__new__(cls)
        rp = np.ascontiguousarray(row_ptr, np.int32)
        cl = np.ascontiguousarray(col, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        bb = np.ascontiguousarray(b, np.float64)
        fx = np.ascontiguousarray(fixed, np.uint8)
        n = len(rp) - 1
        h = ctypes.c_void_p()
        _lib.check(_lib.hip_lib().fvm_template_create_csr(
            mesh.handle, n, _lib.ptr(rp, _lib._i32p), _lib.ptr(cl, _lib._i32p), _lib.ptr(v), _lib.ptr(bb),
            fx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(h)), "fvm_template_create_csr")
        self._handle = h.value
        self._mesh = mesh
        self.kind = _lib.TEMPLATE_POISSON
        self.n = n
        self.node_kind = np.where(fx != 0, _lib.NODE_DIRICHLET, 0).astype(np.int8)
        self.cond_val = np.where(fx != 0, bb, 0.0)
        self.nnz = int(rp[-1])
        self._csr = None
        return self

    @property
    def handle(self):
        return self._handle

    def download(self):
        """(row_ptr, col, val, b) of the assembled operator (host arrays)."""
        if self._csr is None:
            rp = np.empty(self.n + 1, np.int32)
            col = np.empty(self.nnz, np.int32)
            val = np.empty(self.nnz)
            b = np.empty(self.n)
            _lib.check(_lib.hip_lib().fvm_template_download(
                self._handle, _lib.ptr(rp, _lib._i32p), _lib.ptr(col, _lib._i32p), _lib.ptr(val),
                _lib.ptr(b)), "fvm_template_download")
            self._csr = (rp, col, val, b)
        return self._csr

    def apply(self, u, du, stream=None):
        _lib.check(_lib.hip_lib().fvm_template_apply(self._handle, u.ptr, du.ptr, stream),
                   "fvm_template_apply")

    def spmv(self, which, x, y, stream=None):
        _lib.check(_lib.hip_lib().fvm_template_spmv(self._handle, int(which), x.ptr, y.ptr, stream),
                   "fvm_template_spmv")

    def ell_info(self, index_bytes=False):
        """(unrolled slots of the sliced ELL copy of the operator (its longest row's
        off-diagonals, 4..10), off-diagonals past those slots[, bytes of each entry's
        column index: 2 for 16-bit offsets, 4 for columns])."""
        w, o, ib = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        _lib.check(_lib.hip_lib().fvm_template_ell(self._handle, ctypes.byref(w), ctypes.byref(o),
                                                   ctypes.byref(ib)), "fvm_template_ell")
        return (w.value, o.value, ib.value) if index_bytes else (w.value, o.value)

    def residual(self, x, stream=None):
        r = ctypes.c_double()
        _lib.check(_lib.hip_lib().fvm_template_residual(self._handle, x.ptr, ctypes.byref(r), stream),
                   "fvm_template_residual")
        return r.value

    def solve(self, alg=None, *, rtol=1e-10, maxiter=100000, x=None, max_outer=0, stream=None, spmv_format="auto"):
        """Solve A x = b on the device.  Returns (x DeviceArray, fvm_solve_stats dict).
        spmv_format: the form of the scaled operator in the CG direction kernel,
        "auto" (= "ell"), "ell" or "csr"."""
        alg = alg if alg is not None else CG()
        if type(alg) not in _METHOD:
            raise TypeError(f"unsupported linear solver {alg!r}; use CG(), MixedCG() or BiCGStab()")
        o = _lib.SolveOpts()
        o.method = _METHOD[type(alg)]
        if isinstance(alg, MixedCG) and alg.update == "reliable":
            o.method = _lib.SOLVE_CG_RELIABLE
        o.maxiter = int(maxiter)
        o.check_every = int(alg.check_every)
        o.max_outer = int(max_outer)
        o.rtol = float(rtol)
        o.inner_rtol = float(getattr(alg, "inner_rtol", 0.0))
        o.spmv_format = _SPMV_FORMATS[spmv_format]
        if x is None:
            x = _lib.DeviceArray(self.n)
            o.use_x0 = 0
        else:
            o.use_x0 = 1
        st = _lib.SolveStats()
        _lib.check(_lib.hip_lib().fvm_template_solve(self._handle, ctypes.byref(o), x.ptr,
                                                     ctypes.byref(st), stream), "fvm_template_solve")
        stats = dict(converged=bool(st.converged), iterations=st.iterations, outer=st.outer,
                     flag=st.flag, residual=st.residual, seconds=st.seconds)
        return x, stats

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and _lib._hip is not None:
            _lib._hip.fvm_template_destroy(h)
            self._handle = None


class AbstractFVMTemplate:
    """abstract_templates.jl:1-60: solve(prob, args...) forwards to prob.problem."""

    def _init_common(self, mesh, BCs, ICs):
        self.mesh = mesh
        self.conditions = Conditions(mesh, BCs, ICs if ICs is not None else InternalConditions())

    @property
    def A(self):
        """The assembled operator as a scipy CSR matrix (downloaded on first use)."""
        import scipy.sparse as sp

        rp, col, val, _ = self.problem.download()
        n = self.problem.n
        return sp.csr_matrix((val, col, rp), shape=(n, n))

    @property
    def b(self):
        return self.problem.download()[3]

    def solve(self, alg=None, **kwargs):
        raise NotImplementedError


class _SteadyTemplate(AbstractFVMTemplate):
    def solve(self, alg=None, *, rtol=1e-10, maxiter=100000, max_outer=0, spmv_format="auto"):
        """LinearProblem(A, b) solved on the device to |b - A u| <= rtol |b|;
        returns a LinearSolution."""
        x, st = self.problem.solve(alg, rtol=rtol, maxiter=maxiter, max_outer=max_outer, spmv_format=spmv_format)
        return LinearSolution(u=x.to_host(), retcode="Success" if st["converged"] else "MaxIters",
                              resid=st["residual"], iters=st["iterations"], stats=st)


class PoissonsEquation(_SteadyTemplate):
    """div[D(x) grad u] = f(x) (poissons_equation.jl:58-88).  Dudt conditions are
    rejected, as at :69-70."""

    def __init__(self, mesh, BCs, ICs=None, *, diffusion_function=None, diffusion_parameters=None,
                 source_function=None, source_parameters=None, **kwargs):
        if source_function is None:
            raise TypeError("PoissonsEquation needs a source_function")
        self._init_common(mesh, BCs, ICs)
        if self.conditions.has_dudt_nodes():
            raise ValueError("PoissonsEquation does not support Dudt nodes.")
        self.diffusion_function = diffusion_function if diffusion_function is not None else _one
        self.diffusion_parameters = diffusion_parameters
        self.source_function = source_function
        self.source_parameters = source_parameters
        tri = mesh.triangulation
        f_node = _eval_xyp(source_function, source_parameters, tri.x, tri.y)
        self.problem = DeviceTemplate(mesh, self.conditions, _lib.TEMPLATE_POISSON,
                                      self.diffusion_function, diffusion_parameters, f_node=f_node)

    def __repr__(self):
        return f"PoissonsEquation with {self.mesh.triangulation.num_solid_vertices()} nodes"


class LaplacesEquation(_SteadyTemplate):
    """div[D(x) grad u] = 0 (laplaces_equation.jl:51-78)."""

    def __init__(self, mesh, BCs, ICs=None, *, diffusion_function=None, diffusion_parameters=None,
                 **kwargs):
        self._init_common(mesh, BCs, ICs)
        if self.conditions.has_dudt_nodes():
            raise ValueError("LaplacesEquation does not support Dudt nodes.")
        self.diffusion_function = diffusion_function if diffusion_function is not None else _one
        self.diffusion_parameters = diffusion_parameters
        self.problem = DeviceTemplate(mesh, self.conditions, _lib.TEMPLATE_POISSON,
                                      self.diffusion_function, diffusion_parameters)

    def __repr__(self):
        return f"LaplacesEquation with {self.mesh.triangulation.num_solid_vertices()} nodes"


class MeanExitTimeProblem(_SteadyTemplate):
    """div[D(x) grad T] = -1 with T = 0 on absorbing boundaries (mean_exit_time.jl:57-94).
    Dudt conditions and Constrained edges are rejected, as at :66-69."""

    def __init__(self, mesh, BCs, ICs=None, *, diffusion_function, diffusion_parameters=None, **kwargs):
        self._init_common(mesh, BCs, ICs)
        if self.conditions.has_dudt_nodes():
            raise ValueError("MeanExitTimeProblem does not support Dudt nodes.")
        if self.conditions.has_constrained_edges():
            raise ValueError("MeanExitTimeProblem does not support Constrained edges.")
        self.diffusion_function = diffusion_function
        self.diffusion_parameters = diffusion_parameters
        self.problem = DeviceTemplate(mesh, self.conditions, _lib.TEMPLATE_MEAN_EXIT_TIME,
                                      diffusion_function, diffusion_parameters)

    def __repr__(self):
        return f"MeanExitTimeProblem with {self.mesh.triangulation.num_solid_vertices()} nodes"


class _OdeTemplate(AbstractFVMTemplate):
    def _init_ode(self, initial_condition, initial_time, final_time):
        ic = np.array(initial_condition, np.float64)
        self.initial_condition = np.asarray(initial_condition, np.float64)
        self.initial_time = float(initial_time)
        self.final_time = float(final_time)
        # apply_dirichlet_conditions! (abstract_templates.jl:109-117) on the augmented state
        p = self.problem
        dn = p.node_kind == _lib.NODE_DIRICHLET
        ic[dn] = p.cond_val[dn]
        self._u0 = np.concatenate([ic, [1.0]])

    def solve(self, alg=None, *, saveat=None, abstol=1e-6, reltol=1e-3, dt=None, maxiters=10**7,
              save_start=True, callback=None, **kwargs):
        """ODEProblem(MatrixOperator([A b; 0 0]), [u0; 1]) integrated on the device;
        the states keep the augmented last entry (always 1), as in the reference.
        callback(u, t) runs after every accepted step on the device state (its first n
        entries), as a DiscreteCallback(true, ...) handed to the ODEProblem would."""
        from .solve import Tsit5, integrate

        p = self.problem
        n = p.n

        def rhs(u, du, t):
            # [A b; 0 0] [u; 1]: the trailing entry of the augmented state is 1 and
            # its derivative 0; the device keeps only the first n entries
            p.apply(u, du)

        sol = integrate(rhs, self._u0[:n], self.initial_time, self.final_time,
                        alg if alg is not None else Tsit5(), saveat=saveat, abstol=abstol,
                        reltol=reltol, dt=dt, maxiters=maxiters, save_start=save_start,
                        user_callback=callback, jvp=lambda u, v, out, t: p.spmv(0, v, out),
                        op=(_lib.OP_TEMPLATE, p.handle, False), fsal=callback is None, **kwargs)
        sol.u = [np.concatenate([s, [1.0]]) for s in sol.u]
        return sol


class DiffusionEquation(_OdeTemplate):
    """du/dt = div[D(x) grad u] (diffusion_equation.jl:69-101)."""

    def __init__(self, mesh, BCs, ICs=None, *, diffusion_function, diffusion_parameters=None,
                 initial_condition, initial_time=0.0, final_time, **kwargs):
        self._init_common(mesh, BCs, ICs)
        self.diffusion_function = diffusion_function
        self.diffusion_parameters = diffusion_parameters
        self.problem = DeviceTemplate(mesh, self.conditions, _lib.TEMPLATE_DIFFUSION,
                                      diffusion_function, diffusion_parameters)
        self._init_ode(initial_condition, initial_time, final_time)

    def __repr__(self):
        return f"DiffusionEquation with {self.mesh.triangulation.num_solid_vertices()} nodes and time span ({self.initial_time}, {self.final_time})"


class LinearReactionDiffusionEquation(_OdeTemplate):
    """du/dt = div[D(x) grad u] + f(x) u (linear_reaction_diffusion_equations.jl:76-125)."""

    def __init__(self, mesh, BCs, ICs=None, *, diffusion_function, diffusion_parameters=None,
                 source_function, source_parameters=None, initial_condition, initial_time=0.0,
                 final_time, **kwargs):
        self._init_common(mesh, BCs, ICs)
        self.diffusion_function = diffusion_function
        self.diffusion_parameters = diffusion_parameters
        self.source_function = source_function
        self.source_parameters = source_parameters
        tri = mesh.triangulation
        r_node = _eval_xyp(source_function, source_parameters, tri.x, tri.y)
        self.problem = DeviceTemplate(mesh, self.conditions, _lib.TEMPLATE_DIFFUSION,
                                      diffusion_function, diffusion_parameters, r_node=r_node)
        self._init_ode(initial_condition, initial_time, final_time)

    def __repr__(self):
        return f"LinearReactionDiffusionEquation with {self.mesh.triangulation.num_solid_vertices()} nodes and time span ({self.initial_time}, {self.final_time})"


def _one(x, y, p):
    return 1.0
