"""fvm_eqs! and its parameter object (reference src/equations/main_equations.jl,
src/solve.jl:1-131).

`fvm_eqs(du, u, p, t)` is the public right-hand side.  With DeviceArray
arguments it launches the HIP kernels on the arrays in place; with NumPy arrays it
uploads u, runs the same kernels and copies the result into du (convenient for
tests, not for time loops).
"""
import ctypes
from types import SimpleNamespace

import numpy as np

from . import _lib
from .problem import FVMProblem, SteadyFVMProblem
from .system import FVMSystem


def get_fvm_parameters(prob, parallel=True):
    """solve.jl:36-42.  The reference's `parallel` chooses Julia threads; here the
    work always runs on the GPU and `parallel` is kept for signature parity."""
    if isinstance(prob, SteadyFVMProblem):
        prob = prob.problem
    if isinstance(prob, FVMSystem):
        dn = tuple(np.array(sorted(c.dirichlet_nodes), np.int64) for c in prob.conditions)
    else:
        dn = np.array(sorted(prob.conditions.dirichlet_nodes), np.int64)
    return SimpleNamespace(prob=prob, parallel=bool(parallel), device=prob.device_context(),
                           dirichlet_nodes=dn)


def get_multithreading_parameters(prob):
    """solve.jl:1-27."""
    return get_fvm_parameters(prob, True)


def get_serial_parameters(prob):
    """solve.jl:29-34."""
    return get_fvm_parameters(prob, False)


def _prob_of(p):
    if isinstance(p, (FVMProblem, FVMSystem, SteadyFVMProblem)):
        return p.problem if isinstance(p, SteadyFVMProblem) else p
    return p.prob


def fvm_eqs(du, u, p, t):
    """fvm_eqs!(du, u, p, t) (main_equations.jl:28-35).  Returns du.  For an FVMSystem,
    host states are (nv, n_points) matrices and device states the column-major
    vectors u[var + nv i]."""
    prob = _prob_of(p)
    ctx = prob.device_context()
    if isinstance(u, _lib.DeviceArray) and isinstance(du, _lib.DeviceArray):
        ctx.rhs(u, du, t)
        return du
    if isinstance(u, _lib.DeviceArray) or isinstance(du, _lib.DeviceArray):
        raise TypeError("fvm_eqs: pass both u and du as DeviceArray, or both as host arrays")
    du_host = np.asarray(du)
    sys = isinstance(prob, FVMSystem)
    d_u = _lib.DeviceArray.from_host(prob.flat(u) if sys else np.asarray(u, np.float64))
    d_du = _lib.DeviceArray(du_host.size)
    ctx.rhs(d_u, d_du, t)
    du_host[...] = prob.unflat(d_du.to_host()) if sys else d_du.to_host()
    return du


fvm_eqs_ = fvm_eqs  # the reference's name is fvm_eqs!


def update_dirichlet_nodes(u, t, prob):
    """update_dirichlet_nodes! (dirichlet.jl:78-86) on a device or host vector."""
    prob = _prob_of(prob)
    ctx = prob.device_context()
    if isinstance(u, _lib.DeviceArray):
        ctx.apply_dirichlet(u, t)
        return u
    sys = isinstance(prob, FVMSystem)
    d = _lib.DeviceArray.from_host(prob.flat(u) if sys else np.asarray(u, np.float64))
    ctx.apply_dirichlet(d, t)
    u[...] = prob.unflat(d.to_host()) if sys else d.to_host()
    return u


def jacobian_sparsity(prob):
    """jacobian_sparsity(prob) (solve.jl:44-77): CSR matrix of ones, row i holding
    i and its mesh neighbours.  For an FVMSystem of nv variables (solve.jl:51-53,
    96-131), entry (i nv + l, j nv + k) for every pair of variables l, k of related
    nodes i, j (column-major state)."""
    import scipy.sparse as sp

    if isinstance(prob, SteadyFVMProblem):
        prob = prob.problem
    tri = prob.mesh.triangulation if hasattr(prob, "mesh") else prob
    row_ptr, col = csr_pattern(tri)
    A = sp.csr_matrix((np.ones(len(col)), col, row_ptr), shape=(tri.num_points(),) * 2)
    if isinstance(prob, FVMSystem):
        A = sp.kron(A, np.ones((prob.n_vars, prob.n_vars)), format="csr")
    return A


def csr_pattern(tri):
    L = _lib.host_lib()
    n = tri.num_points()
    nnz = ctypes.c_int64()
    L.fvmh_csr_pattern(n, len(tri.triangles), _lib.ptr(tri.triangles, _lib._i32p), None, None,
                       ctypes.byref(nnz))
    row_ptr = np.empty(n + 1, np.int64)
    col = np.empty(nnz.value, np.int32)
    L.fvmh_csr_pattern(n, len(tri.triangles), _lib.ptr(tri.triangles, _lib._i32p),
                       _lib.ptr(row_ptr, _lib._i64p), _lib.ptr(col, _lib._i32p), ctypes.byref(nnz))
    return row_ptr, col
