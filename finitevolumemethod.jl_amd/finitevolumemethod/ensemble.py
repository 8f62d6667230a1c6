"""Ensembles of FVMProblems advanced together on the GPU.

The reference runs an ensemble through SciML's EnsembleProblem, which calls
solve (solve.jl:222-256) once per member.  EnsembleProblem here packs every
member's mesh and closures into one device mesh and context (fvm_batch_create),
so each right-hand side evaluation of the whole ensemble is one launch of the
RHS kernels, and explicit steps are replayed from a captured HIP graph.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib, forms
from .problem import FVMProblem, SteadyFVMProblem, physics_desc, physics_of
from .solve import Euler, ODESolution


def explicit_dt(prob, safety=0.9):
    """A stable forward-Euler step for a problem whose flux is linear in u:
    safety * 2 / rho, with rho the Gershgorin bound max_i sum_j |J_ij| of the
    Jacobian of fvm_eqs!, assembled as the DiffusionEquation template operator
    (abstract_templates.jl:73-267) with D evaluated at t = 0, u = 0."""
    from .templates import DeviceTemplate

    D = prob.diffusion_function
    if D is None or prob.user_flux_function is not None:
        raise ValueError("explicit_dt needs a problem given by a diffusion_function")
    Dp = prob.diffusion_parameters
    tmpl = DeviceTemplate(prob.mesh, prob.conditions, _lib.TEMPLATE_DIFFUSION,
                          lambda x, y, p: D(x, y, 0.0, 0.0 * x, Dp), None, condition_values=False)
    rp, col, val, _ = tmpl.download()
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rho = np.bincount(rows, weights=np.abs(val), minlength=len(rp) - 1).max()
    return safety * 2.0 / rho


@dataclass
class EnsembleSolution:
    """One ODESolution per member (u at the initial and final times) plus timings."""
    u: list = field(default_factory=list)
    stats: dict = field(default_factory=dict)

    def __getitem__(self, k):
        return self.u[k]

    def __len__(self):
        return len(self.u)


class EnsembleProblem:
    """EnsembleProblem(prob; prob_func, trajectories) over FVMProblems, FVMSystems or
    SteadyFVMProblems, or EnsembleProblem(list_of_problems).  prob_func(prob, i, repeat)
    returns member i (1-based i, as SciML's); trajectories may also be given to solve."""

    def __init__(self, prob, prob_func=None, trajectories=None):
        from .system import FVMSystem

        self._prob, self._prob_func = prob, prob_func
        if isinstance(prob, (list, tuple)):
            self.problems = list(prob)
        else:
            self._make(trajectories or 1)
        if not self.problems or not all(isinstance(p, (FVMProblem, FVMSystem, SteadyFVMProblem))
                                        for p in self.problems):
            raise TypeError("EnsembleProblem members must be FVMProblems, FVMSystems or SteadyFVMProblems")
        self._batch = None

    def _make(self, trajectories):
        f = self._prob_func if self._prob_func is not None else (lambda p, i, repeat: p)
        self.problems = [f(self._prob, i + 1, 1) for i in range(int(trajectories))]

    def __len__(self):
        return len(self.problems)

    def device_batch(self, dt=None, tile_triangles=0, autotune=False, graph_steps=0):
        """The packed device ensemble (an fvm_batch)."""
        return DeviceBatch(self.problems, dt, tile_triangles, autotune, graph_steps)

    def solve(self, alg=None, *, dt=None, autotune=True, graph_steps=100, trajectories=None, **kwargs):
        """Euler() (the default): every member advanced together in the packed device
        batch from its initial_time to its final_time with its own dt (a scalar, a
        list, or None for explicit_dt); every member must take the same number of
        steps; the solutions hold the initial and final states.  Any other algorithm
        (Tsit5(), TRBDF2(), ...), systems, steady problems, and members whose flux function
        is a linear form other than -D(x, y) grad u (advection; not packed): the members are
        solved one after the other on the device, each with its own adaptive steps or
        Newton iterations (SciML's EnsembleSerial), keyword arguments passed to solve.
        Returns an EnsembleSolution."""
        from .solve import solve as solve_one
        from .system import FVMSystem

        if trajectories is not None and not isinstance(self._prob, (list, tuple)):
            self._make(trajectories)
        steady = any(isinstance(p, SteadyFVMProblem) for p in self.problems)
        alg = alg if alg is not None else (None if steady else Euler())
        if (steady or not isinstance(alg, Euler) or any(isinstance(p, FVMSystem) for p in self.problems)
                or any(_linear_flux(p) for p in self.problems)):
            tm = _lib.Timer()
            tm.start()
            sols = [solve_one(p, alg, dt=dt, **kwargs) for p in self.problems]
            return EnsembleSolution(u=sols, stats=dict(ms=tm.stop(), mode="serial"))
        P = len(self.problems)
        if dt is None:
            dt = alg.dt
        if dt is None:
            dts = [explicit_dt(p) for p in self.problems]
        else:
            dts = list(np.broadcast_to(np.asarray(dt, np.float64), (P,)))
        spans = [p.final_time - p.initial_time for p in self.problems]
        steps = {int(round(s / h)) for s, h in zip(spans, dts)}
        if len(steps) != 1 or any(abs(s - h * next(iter(steps))) > 1e-9 * max(1.0, s) for s, h in zip(spans, dts)):
            raise ValueError("every member must reach its final_time in the same number of Euler "
                             "steps: choose final_time = initial_time + n dt per member")
        nsteps = steps.pop()
        t0 = {p.initial_time for p in self.problems}
        if len(t0) > 1 and any(_time_dependent(p) for p in self.problems):
            raise ValueError("the packed Euler path steps every member from one initial_time: members "
                             "with time-dependent sources or conditions need the same initial_time")
        b = DeviceBatch(self.problems, dts, autotune=autotune,
                        graph_steps=min(graph_steps, nsteps) if graph_steps else 0)
        u = b.pack([p.initial_condition for p in self.problems])
        d_u = _lib.DeviceArray.from_host(u)
        tm = _lib.Timer()
        tm.start()
        b.euler(d_u, nsteps, t0.pop() if len(t0) == 1 else 0.0)
        ms = tm.stop()
        finals = b.unpack(d_u.to_host())
        sols = []
        for p, uf in zip(self.problems, finals):
            s = ODESolution()
            s.t = [p.initial_time, p.final_time]
            s.u = [np.asarray(p.initial_condition, np.float64), uf]
            s.retcode = "Success"
            sols.append(s)
        return EnsembleSolution(u=sols, stats=dict(steps=nsteps, ms=ms, dt=dts,
                                                   tile_triangles=b.info.tile_triangles))


def _time_dependent(p):
    """Whether an FVMProblem's traced source or condition functions depend on t."""
    ph = physics_of(p)
    fn = ph["fn"]
    return bool(ph.get("st", 0.0) != 0.0 or np.any(np.asarray(fn[2]) != 0.0)
                or np.any(np.asarray(fn[5]) != 0.0))


def _linear_flux(p):
    """Whether an FVMProblem's flux function traces to the general linear form."""
    q = getattr(p, "user_flux_function", None)
    if q is None:
        return False
    fp = p.flux_parameters
    return forms.classify_flux(lambda x, y, t, a, b, g: q(x, y, t, a, b, g, fp))[0] == "linear"


def _uniform_form(ph, tri):
    """Members with different constant D (traced as the polynomial D(u) = d0) are put
    in the per-CV-edge form, which every member can share."""
    dp = list(ph.get("dpoly", [0.0] * 4))
    if ph["flux_kind"] == _lib.FLUX_D_POLY_U and all(v == 0.0 for v in dp[1:]):
        ph = dict(ph)
        ph["flux_kind"] = _lib.FLUX_D_EDGE
        ph["d_edge"] = np.full(3 * len(tri.triangles), dp[0])
        ph["d_bedge"] = np.full(2 * len(tri.boundary_edges), dp[0])
    return ph


class DeviceBatch:
    """An fvm_batch: the members packed on the device."""

    def __init__(self, problems, dt=None, tile_triangles=0, autotune=False, graph_steps=0, fused=True):
        P = len(problems)
        meshes = (_lib.MeshDesc * P)()
        phys = (_lib.PhysicsDesc * P)()
        self._keep = []
        for k, prob in enumerate(problems):
            tri = prob.mesh.triangulation
            meshes[k] = _lib.MeshDesc(len(tri.x), len(tri.triangles), len(tri.boundary_edges),
                                      _lib.ptr(tri.x), _lib.ptr(tri.y),
                                      _lib.ptr(tri.triangles, _lib._i32p),
                                      _lib.ptr(tri.boundary_edges, _lib._i32p), 0, 0)
            d, keep = physics_desc(_uniform_form(physics_of(prob), tri))
            phys[k] = d
            self._keep.append(keep)
        o = _lib.BatchOpts(int(tile_triangles), 1 if autotune else 0, int(graph_steps), 0 if fused else 1)
        dts = None if dt is None else np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (P,)))
        h = ctypes.c_void_p()
        _lib.check(_lib.hip_lib().fvm_batch_create(P, meshes, phys, _lib.ptr(dts) if dts is not None else None,
                                                   ctypes.byref(o), ctypes.byref(h)), "fvm_batch_create")
        self._handle = h.value
        self._keep = None
        info = _lib.BatchInfo()
        off = np.empty(P + 1, np.int64)
        _lib.check(_lib.hip_lib().fvm_batch_get_info(self._handle, ctypes.byref(info), _lib.ptr(off, _lib._i64p)),
                   "fvm_batch_get_info")
        self.info = info
        self.offsets = off
        self.n = int(off[-1])

    def pack(self, states):
        return np.concatenate([np.asarray(s, np.float64) for s in states])

    def unpack(self, u):
        return [u[self.offsets[k]:self.offsets[k + 1]].copy() for k in range(len(self.offsets) - 1)]

    def rhs(self, u, du, t=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_batch_rhs(self._handle, u.ptr, du.ptr, float(t), stream), "fvm_batch_rhs")

    def euler(self, u, n_steps, t0=0.0, stream=None):
        _lib.check(_lib.hip_lib().fvm_batch_euler(self._handle, u.ptr, int(n_steps), float(t0), stream),
                   "fvm_batch_euler")

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and _lib._hip is not None:
            _lib._hip.fvm_batch_destroy(h)
            self._handle = None
