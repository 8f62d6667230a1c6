"""ctypes binding of libfvmhip.so / libfvmhost.so (the C ABI in include/fvm_hip.h).

This is the host side of the drop-in boundary: the Python mirror of
FiniteVolumeMethod.jl's API calls the HIP kernels only through these entry points.
There is no CPU fallback: when the library or a GPU is missing, calls raise.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i8p = ctypes.POINTER(ctypes.c_int8)
_vp = ctypes.c_void_p


class FVMError(RuntimeError):
    """A failure reported by the HIP library (negative status)."""


class MeshDesc(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_int64), ("n_triangles", ctypes.c_int64),
                ("n_boundary_edges", ctypes.c_int64), ("x", _dp), ("y", _dp),
                ("triangles", _i32p), ("boundary_edges", _i32p),
                ("tile_triangles", ctypes.c_int32), ("tile_max_nodes", ctypes.c_int32)]


class MeshInfo(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_int64), ("n_triangles", ctypes.c_int64),
                ("n_boundary_edges", ctypes.c_int64), ("n_tiles", ctypes.c_int32),
                ("n_fix_nodes", ctypes.c_int32), ("n_partials", ctypes.c_int32),
                ("n_local_nodes", ctypes.c_int64), ("d_cv_volumes", _vp), ("d_x", _vp),
                ("d_y", _vp)]


class PhysicsDesc(ctypes.Structure):
    _fields_ = [("flux_kind", ctypes.c_int32), ("dpoly", ctypes.c_double * 4),
                ("d_edge", _dp), ("d_bedge", _dp), ("log_d", _dp),
                ("spoly", ctypes.c_double * 4), ("st", ctypes.c_double), ("s_node", _dp),
                ("node_kind", _i8p), ("node_fidx", _i32p), ("bedge_neumann_fidx", _i32p),
                ("n_functions", ctypes.c_int32), ("fn_c0", _dp), ("fn_cu", _dp),
                ("fn_ct", _dp), ("fn_cx", _dp), ("fn_cy", _dp), ("node_g", _dp),
                ("bedge_g", _dp), ("fn_lam", _dp), ("jit_source", ctypes.c_char_p)]


class _Tiles(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_int64), ("n_tri", ctypes.c_int64),
                ("n_bedges", ctypes.c_int64), ("n_tiles", ctypes.c_int32),
                ("tile_tris", ctypes.c_int32), ("max_local_nodes", ctypes.c_int32),
                ("n_local", ctypes.c_int64), ("n_fix", ctypes.c_int32),
                ("n_parts", ctypes.c_int32)] + [
        (name, _vp) for name in (
            "perm", "tri_sorted", "tile_tri_start", "tile_node_start", "tile_nodes", "loc",
            "inc_start", "inc", "part_idx", "fix_nodes", "fix_part_start", "fix_bedge_start",
            "fix_bedge", "fix_bvert_start", "fix_bvert", "node_tri_start", "node_tri",
            "badj_sorted", "tile_of_node_excl")]


class TemplateDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("d_edge", _dp), ("d_bedge", _dp),
                ("node_kind", _i8p), ("bedge_neumann_fidx", _i32p), ("f_node", _dp),
                ("r_node", _dp), ("neu_a", _dp), ("dir_val", _dp), ("dudt_val", _dp)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("maxiter", ctypes.c_int32),
                ("check_every", ctypes.c_int32), ("max_outer", ctypes.c_int32),
                ("use_x0", ctypes.c_int32), ("spmv_format", ctypes.c_int32),
                ("rtol", ctypes.c_double), ("inner_rtol", ctypes.c_double)]


class OdeOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("op_kind", ctypes.c_int32), ("op", _vp),
                ("dirichlet", ctypes.c_int32), ("fsal", ctypes.c_int32), ("n", ctypes.c_int64),
                ("t0", ctypes.c_double), ("tf", ctypes.c_double), ("dt", ctypes.c_double),
                ("abstol", ctypes.c_double), ("reltol", ctypes.c_double), ("maxiters", ctypes.c_int64),
                ("n_save", ctypes.c_int64), ("save_t", _dp), ("d_saves", _vp)]


class OdeStats(ctypes.Structure):
    _fields_ = [("nf", ctypes.c_int64), ("naccept", ctypes.c_int64), ("nreject", ctypes.c_int64),
                ("nsaved", ctypes.c_int64), ("retcode", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("t", ctypes.c_double)]


ODE_TSIT5, ODE_EULER, ODE_SSPRK3 = 0, 1, 2
OP_PROBLEM, OP_SYSTEM, OP_TEMPLATE = 0, 1, 2
ODE_RETCODES = {0: "Success", 1: "MaxIters", 2: "Unstable"}


class SolveStats(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("outer", ctypes.c_int32), ("flag", ctypes.c_int32),
                ("residual", ctypes.c_double), ("seconds", ctypes.c_double)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("tile_triangles", ctypes.c_int32), ("autotune", ctypes.c_int32),
                ("graph_steps", ctypes.c_int32), ("unfused", ctypes.c_int32)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("n_problems", ctypes.c_int32), ("tile_triangles", ctypes.c_int32),
                ("n_points", ctypes.c_int64), ("n_triangles", ctypes.c_int64),
                ("n_boundary_edges", ctypes.c_int64), ("n_tiles", ctypes.c_int32),
                ("n_fix_nodes", ctypes.c_int32), ("n_candidates", ctypes.c_int32),
                ("candidate_tiles", ctypes.c_int32 * 4), ("candidate_ms", ctypes.c_float * 4)]


class SystemDesc(ctypes.Structure):
    _fields_ = [("n_vars", ctypes.c_int32), ("flux_k", _dp), ("n_terms", ctypes.c_int32),
                ("term_var", _i32p), ("term_exp", _i32p), ("term_coef", _dp), ("s_node", _dp),
                ("st", _dp), ("node_kind", _i8p), ("node_fidx", _i32p),
                ("bedge_neumann_fidx", _i32p), ("n_functions", ctypes.c_int32), ("fn_c0", _dp),
                ("fn_cu", _dp), ("fn_ct", _dp), ("fn_cx", _dp), ("fn_cy", _dp), ("fn_lam", _dp),
                ("node_g", _dp), ("bedge_g", _dp)]


FLUX_D_POLY_U, FLUX_D_NODAL, FLUX_D_EDGE, FLUX_LINEAR, FLUX_D_POWER, FLUX_JIT = 0, 1, 2, 3, 4, 5
TEMPLATE_POISSON, TEMPLATE_MEAN_EXIT_TIME, TEMPLATE_DIFFUSION = 0, 1, 2
SOLVE_CG, SOLVE_CG_MIXED, SOLVE_BICGSTAB, SOLVE_CG_RELIABLE = 0, 1, 2, 3
NODE_NONE, NODE_DIRICHLET, NODE_DUDT = 0, 1, 2

_lock = threading.Lock()
_host = None
_hip = None


def _setup_host(L):
    L.fvmh_last_error.restype = ctypes.c_char_p
    L.fvmh_rectangle_sizes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, _i64p,
                                       _i64p, _i32p]
    L.fvmh_rectangle.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                         ctypes.c_double, ctypes.c_uint64, _dp, _dp,
                                                         _i32p, _i32p, _i32p]
    L.fvmh_count_non_delaunay.argtypes = [ctypes.c_int64, _dp, _dp, ctypes.c_int64, _i32p,
                                          ctypes.c_double]
    L.fvmh_count_non_delaunay.restype = ctypes.c_int64
    L.fvmh_boundary_adjacent.argtypes = [ctypes.c_int64, _i32p, ctypes.c_int64, _i32p, _i32p]
    L.fvmh_tiles_build.argtypes = [ctypes.c_int64, _dp, _dp, ctypes.c_int64, _i32p,
                                   ctypes.c_int64, _i32p, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.POINTER(_Tiles))]
    L.fvmh_tiles_free.argtypes = [ctypes.POINTER(_Tiles)]
    L.fvmh_tiles_free.restype = None
    L.fvmh_csr_pattern.argtypes = [ctypes.c_int64, ctypes.c_int64, _i32p, _i64p, _i32p, _i64p]
    L.fvmh_colour_d2.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p]


def host_lib():
    """libfvmhost.so: mesh generation and preprocessing (no GPU needed)."""
    global _host
    with _lock:
        if _host is None:
            path = os.path.join(LIB_DIR, "libfvmhost.so")
            if not os.path.exists(path):
                raise FVMError(f"{path} is missing; build with __graft_entry__.build()")
            L = ctypes.CDLL(path)
            _setup_host(L)
            _host = L
    return _host


# (name, argtypes) of every entry point of include/fvm_hip.h
HIP_API = {
    "fvm_last_error": [ctypes.c_char_p, ctypes.c_size_t],
    "fvm_build_info": [ctypes.c_char_p, ctypes.c_size_t],
    "fvm_device_count": [_i32p],
    "fvm_device_synchronize": [],
    "fvm_alloc": [ctypes.POINTER(_vp), ctypes.c_size_t],
    "fvm_free": [_vp],
    "fvm_memcpy_h2d": [_vp, _vp, ctypes.c_size_t, _vp],
    "fvm_memcpy_d2h": [_vp, _vp, ctypes.c_size_t, _vp],
    "fvm_memcpy_d2d": [_vp, _vp, ctypes.c_size_t, _vp],
    "fvm_memset": [_vp, ctypes.c_int, ctypes.c_size_t, _vp],
    "fvm_stream_create": [ctypes.POINTER(_vp)],
    "fvm_stream_destroy": [_vp],
    "fvm_stream_synchronize": [_vp],
    "fvm_event_create": [ctypes.POINTER(_vp)],
    "fvm_event_destroy": [_vp],
    "fvm_event_record": [_vp, _vp],
    "fvm_event_elapsed_ms": [ctypes.POINTER(ctypes.c_float), _vp, _vp],
    "fvm_mesh_create": [ctypes.POINTER(MeshDesc), ctypes.POINTER(_vp)],
    "fvm_mesh_destroy": [_vp],
    "fvm_mesh_get_info": [_vp, ctypes.POINTER(MeshInfo)],
    "fvm_mesh_download_geometry": [_vp, _dp, _dp],
    "fvm_mesh_rebuilt_geometry": [_vp, _dp],
    "fvm_ctx_create": [_vp, ctypes.POINTER(PhysicsDesc), ctypes.POINTER(_vp)],
    "fvm_ctx_destroy": [_vp],
    "fvm_ctx_set_log_d": [_vp, _vp],
    "fvm_rhs": [_vp, _vp, _vp, ctypes.c_double, _vp],
    "fvm_rhs_timed": [_vp, _vp, _vp, ctypes.c_double, _vp, ctypes.POINTER(ctypes.c_float),
                      ctypes.POINTER(ctypes.c_float)],
    "fvm_dirichlet_apply": [_vp, _vp, ctypes.c_double, _vp],
    "fvm_jvp": [_vp, _vp, _vp, _vp, ctypes.c_double, _vp],
    "fvm_vjp": [_vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp],
    "fvm_template_create": [_vp, ctypes.POINTER(TemplateDesc), ctypes.POINTER(_vp)],
    "fvm_template_create_csr": [_vp, ctypes.c_int64, _i32p, _i32p, _dp, _dp, ctypes.POINTER(ctypes.c_uint8),
                                ctypes.POINTER(_vp)],
    "fvm_template_destroy": [_vp],
    "fvm_template_info": [_vp, _i64p, _i64p],
    "fvm_template_ell": [_vp, ctypes.POINTER(ctypes.c_int32), _i64p, ctypes.POINTER(ctypes.c_int32)],
    "fvm_template_download": [_vp, _i32p, _i32p, _dp, _dp],
    "fvm_template_apply": [_vp, _vp, _vp, _vp],
    "fvm_template_spmv": [_vp, ctypes.c_int32, _vp, _vp, _vp],
    "fvm_template_residual": [_vp, _vp, _dp, _vp],
    "fvm_template_solve": [_vp, ctypes.POINTER(SolveOpts), _vp, ctypes.POINTER(SolveStats), _vp],
    "fvm_ode_solve": [ctypes.POINTER(OdeOpts), _vp, ctypes.POINTER(OdeStats), _vp],
    "fvm_locate_points": [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp],
    "fvm_arnoldi_step": [ctypes.c_int32, _vp, ctypes.c_int64, _vp, ctypes.c_double, ctypes.c_double,
                         ctypes.c_double, ctypes.c_int32, ctypes.POINTER(_vp), _vp, _dp, _vp],
    "fvm_batch_create": [ctypes.c_int32, ctypes.POINTER(MeshDesc), ctypes.POINTER(PhysicsDesc), _dp,
                         ctypes.POINTER(BatchOpts), ctypes.POINTER(_vp)],
    "fvm_batch_destroy": [_vp],
    "fvm_batch_get_info": [_vp, ctypes.POINTER(BatchInfo), _i64p],
    "fvm_batch_mesh": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    "fvm_batch_rhs": [_vp, _vp, _vp, ctypes.c_double, _vp],
    "fvm_batch_euler": [_vp, _vp, ctypes.c_int32, ctypes.c_double, _vp],
    "fvm_system_create": [_vp, ctypes.POINTER(SystemDesc), ctypes.POINTER(_vp)],
    "fvm_system_destroy": [_vp],
    "fvm_system_rhs": [_vp, _vp, _vp, ctypes.c_double, _vp],
    "fvm_system_dirichlet_apply": [_vp, _vp, ctypes.c_double, _vp],
    "fvm_system_jvp": [_vp, _vp, _vp, _vp, ctypes.c_double, _vp],
    "fvm_system_compute_flux": [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_double, _vp, _vp],
    "fvm_compute_flux": [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp],
    "fvm_pl_interpolate": [_vp, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp],
    "fvm_vec_lincomb": [ctypes.c_int64, _vp, ctypes.c_int, _dp, ctypes.POINTER(_vp), _vp],
    "fvm_vec_mul": [ctypes.c_int64, _vp, _vp, _vp, _vp],
    "fvm_vec_dot": [ctypes.c_int64, _vp, _vp, _dp, _vp],
    "fvm_vec_mdot": [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_vp), _vp, _dp, _vp],
    "fvm_vec_mdot_noalloc": [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_vp), _vp, _vp, _vp],
    "fvm_vec_maxpy": [ctypes.c_int64, _vp, ctypes.c_int32, _dp, ctypes.POINTER(_vp), _vp],
    "fvm_trbdf2_affine": [_vp, _vp, _vp, _vp, _vp],
    "fvm_arnoldi_step_scaled": [ctypes.c_int32, _vp, ctypes.c_int64, _vp, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, _vp, _vp, ctypes.c_int32, ctypes.POINTER(_vp), _vp, _dp, _vp],
    "fvm_jit_check": [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_size_t)],
    "fvm_vec_lincomb_err_norm": [ctypes.c_int64, ctypes.c_int, _dp, ctypes.POINTER(_vp), _vp, _vp,
                                 ctypes.c_double, ctypes.c_double, _dp, _vp],
    "fvm_vec_err_norm": [ctypes.c_int64, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _dp,
                         _vp],
}


def hip_lib():
    """libfvmhip.so, loaded once.  Raises FVMError when it is absent."""
    global _hip
    with _lock:
        if _hip is None:
            path = os.path.join(LIB_DIR, "libfvmhip.so")
            if not os.path.exists(path):
                raise FVMError(f"{path} is missing; build with __graft_entry__.build()")
            L = ctypes.CDLL(path)
            for name, args in HIP_API.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = ctypes.c_int
            _hip = L
    return _hip


def build_info():
    """The library's build configuration string (fvm_build_info)."""
    buf = ctypes.create_string_buffer(256)
    check(hip_lib().fvm_build_info(buf, 256), "fvm_build_info")
    return buf.value.decode()


def last_error():
    buf = ctypes.create_string_buffer(1024)
    hip_lib().fvm_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status, what=""):
    if status != 0:
        raise FVMError(f"{what}: {last_error()} (status {status})")


_gpu_ok = None


def require_gpu():
    """Fail loudly when no GPU is visible: the product has no CPU path."""
    global _gpu_ok
    if _gpu_ok is None:
        n = ctypes.c_int32(0)
        st = hip_lib().fvm_device_count(ctypes.byref(n))
        _gpu_ok = st == 0 and n.value > 0
    if not _gpu_ok:
        raise FVMError("no HIP device available: the finite-volume kernels need an MI355X GPU")


def ptr(a, typ=_dp):
    return a.ctypes.data_as(typ) if a is not None else ctypes.cast(None, typ)


class DeviceArray:
    """A 1-D device buffer allocated by fvm_alloc (freed with the object)."""

    __slots__ = ("n", "dtype", "ptr", "__weakref__")

    def __init__(self, n, dtype=np.float64):
        require_gpu()
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = _vp()
        check(hip_lib().fvm_alloc(ctypes.byref(p), max(1, self.n * self.dtype.itemsize)),
              "fvm_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else np.asarray(a).dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, n, dtype=np.float64):
        d = cls(n, dtype)
        d.fill_zero()
        return d

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, self.dtype)
        if a.size != self.n:
            raise ValueError(f"size mismatch: device {self.n}, host {a.size}")
        check(hip_lib().fvm_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes, stream), "fvm_memcpy_h2d")

    def to_host(self, stream=None):
        out = np.empty(self.n, self.dtype)
        check(hip_lib().fvm_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream),
              "fvm_memcpy_d2h")
        return out

    def copy_from(self, other, stream=None):
        if other.n != self.n or other.dtype != self.dtype:
            raise ValueError("copy_from: shape/dtype mismatch")
        check(hip_lib().fvm_memcpy_d2d(self.ptr, other.ptr, self.nbytes, stream), "fvm_memcpy_d2d")

    def fill_zero(self, stream=None):
        check(hip_lib().fvm_memset(self.ptr, 0, self.nbytes, stream), "fvm_memset")

    def copy(self):
        d = DeviceArray(self.n, self.dtype)
        d.copy_from(self)
        return d

    def __len__(self):
        return self.n

    def __array__(self, dtype=None, copy=None):
        a = self.to_host()
        return a.astype(dtype) if dtype is not None else a

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p and _hip is not None:
            try:
                _hip.fvm_free(p)
            except Exception:
                pass
            self.ptr = None

    def __repr__(self):
        return f"DeviceArray(n={self.n}, dtype={self.dtype})"


def synchronize():
    check(hip_lib().fvm_device_synchronize(), "fvm_device_synchronize")


def lincomb(out, coefs, xs, stream=None):
    """out = sum coefs[k] * xs[k] on the device (1..8 terms)."""
    n = len(xs)
    c = (ctypes.c_double * n)(*[float(v) for v in coefs])
    p = (_vp * n)(*[x.ptr for x in xs])
    check(hip_lib().fvm_vec_lincomb(out.n, out.ptr, n, c, p, stream), "fvm_vec_lincomb")


def mul(out, x, y, stream=None):
    """out = x * y elementwise on the device."""
    check(hip_lib().fvm_vec_mul(len(out), out.ptr, x.ptr, y.ptr, stream), "fvm_vec_mul")


def dot(x, y, stream=None):
    r = ctypes.c_double(0)
    check(hip_lib().fvm_vec_dot(x.n, x.ptr, y.ptr, ctypes.byref(r), stream), "fvm_vec_dot")
    return r.value


def mdot(xs, y, stream=None):
    """[x . y for x in xs] (one pass over y)."""
    k = len(xs)
    ptrs = (_vp * k)(*[x.ptr for x in xs])
    out = np.empty(k)
    check(hip_lib().fvm_vec_mdot(y.n, k, ptrs, y.ptr, ptr(out), stream), "fvm_vec_mdot")
    return out


def maxpy(y, coefs, xs, stream=None):
    """y += sum_j coefs[j] xs[j]."""
    k = len(xs)
    ptrs = (_vp * k)(*[x.ptr for x in xs])
    c = np.ascontiguousarray(coefs, np.float64)
    check(hip_lib().fvm_vec_maxpy(y.n, y.ptr, k, ptr(c), ptrs, stream), "fvm_vec_maxpy")


def err_norm(err, u0, u1, abstol, reltol, stream=None):
    r = ctypes.c_double(0)
    check(hip_lib().fvm_vec_err_norm(err.n, err.ptr, u0.ptr, u1.ptr, abstol, reltol,
                                     ctypes.byref(r), stream), "fvm_vec_err_norm")
    return r.value


def lincomb_err_norm(coefs, xs, u0, u1, abstol, reltol, stream=None):
    """err_norm(sum coefs[k] * xs[k], u0, u1, abstol, reltol) without storing the sum."""
    n = len(xs)
    c = (ctypes.c_double * n)(*[float(v) for v in coefs])
    p = (_vp * n)(*[x.ptr for x in xs])
    r = ctypes.c_double(0)
    check(hip_lib().fvm_vec_lincomb_err_norm(u0.n, n, c, p, u0.ptr, u1.ptr, abstol, reltol,
                                             ctypes.byref(r), stream), "fvm_vec_lincomb_err_norm")
    return r.value


class Timer:
    """HIP events around device work on a stream (elapsed milliseconds)."""

    def __init__(self, stream=None):
        L = hip_lib()
        self.stream = stream
        self.a, self.b = _vp(), _vp()
        check(L.fvm_event_create(ctypes.byref(self.a)), "fvm_event_create")
        check(L.fvm_event_create(ctypes.byref(self.b)), "fvm_event_create")

    def start(self):
        check(hip_lib().fvm_event_record(self.a, self.stream), "fvm_event_record")

    def stop(self):
        check(hip_lib().fvm_event_record(self.b, self.stream), "fvm_event_record")
        ms = ctypes.c_float(0)
        check(hip_lib().fvm_event_elapsed_ms(ctypes.byref(ms), self.a, self.b), "fvm_event_elapsed_ms")
        return ms.value

    def __del__(self):
        if _hip is not None:
            for e in (getattr(self, "a", None), getattr(self, "b", None)):
                if e is not None and e.value:
                    _hip.fvm_event_destroy(e)
