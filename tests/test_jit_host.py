"""Closures compiled for the device (jit.py): the traced DAG re-evaluated on the host
equals the closure bit for bit; the generated source, built for the host with g++
(the same text with __device__ dropped), reproduces the closure's value and gives
its u-derivative (against central differences); hipRTC compiles it into the
FVM_FLUX_JIT kernels for gfx950 (fvm_jit_check, no GPU needed)."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from finitevolumemethod import _lib, jit

CLOSURES = {
    "exp": lambda x, y, t, u: np.exp(-u),
    "rational": lambda x, y, t, u: 1.0 / (1.0 + u ** 2),
    "xy_mixed": lambda x, y, t, u: (1.0 + x) * np.exp(-u) + 0.1 * np.sin(y * u) - t * u / 3.0,
    "u_exponent": lambda x, y, t, u: 0.2 + 2.0 ** (u - 0.5) * np.sqrt(1.0 + u * u),
    "where": lambda x, y, t, u: np.where(u > 0.5, u ** 2, 0.25 + 0.0 * u) + np.maximum(u, 0.1) / 7,
    "sin": lambda x, y, t, u: np.sin(u),
    "logistic_ratio": lambda x, y, t, u: u * (1 - u) / (1 + u) + np.tanh(x - y) * np.log1p(u * u),
}


def _points(n=200, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0, 2, n), rng.uniform(-0.9, 1.8, n)


@pytest.mark.parametrize("name", sorted(CLOSURES))
def test_trace_reevaluates_to_the_closure(name):
    """The DAG keeps every operation of the closure in order: re-evaluated on the host
    it gives the closure's value to the last bit or two (Python's and NumPy's pow and
    transcendental functions may differ by an ulp)."""
    f = CLOSURES[name]
    e = jit.trace(f)
    for x, y, t, u in zip(*_points(50)):
        ref = float(f(x, y, t, u))
        assert abs(jit.evaluate(e, x, y, t, u) - ref) <= 4.5e-16 * max(1.0, abs(ref))


@pytest.fixture(scope="module")
def host_build():
    """The generated functions of every closure compiled for the host."""
    src = ["#include <cmath>", "using namespace std;", "#define __device__", "#define __forceinline__ inline"]
    for k, name in enumerate(sorted(CLOSURES)):
        code = jit.emit(f"F{k}", jit.trace(CLOSURES[name]))
        src.append(code.replace("__device__ double fvm_jit_", 'extern "C" double fvm_jit_'))
    d = tempfile.mkdtemp()
    cpp, so = os.path.join(d, "gen.cpp"), os.path.join(d, "gen.so")
    open(cpp, "w").write("\n".join(src))
    subprocess.run(["g++", "-O1", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC", cpp, "-o", so], check=True)
    return ctypes.CDLL(so)


@pytest.mark.parametrize("name", sorted(CLOSURES))
def test_generated_code_value_and_derivative(host_build, name):
    k = sorted(CLOSURES).index(name)
    fv, fd = getattr(host_build, f"fvm_jit_F{k}"), getattr(host_build, f"fvm_jit_dF{k}")
    for g in (fv, fd):
        g.restype = ctypes.c_double
        g.argtypes = [ctypes.c_double] * 4
    f = CLOSURES[name]
    for x, y, t, u in zip(*_points()):
        if name == "where" and abs(u - 0.5) < 1e-3 or name == "where" and abs(u - 0.1) < 1e-3:
            continue  # the derivative jumps there
        ref = float(f(x, y, t, u))
        got = fv(x, y, t, u)
        assert abs(got - ref) <= 4.5e-16 * max(1.0, abs(ref)), (got, ref)
        h = 1e-6
        fdv = (float(f(x, y, t, u + h)) - float(f(x, y, t, u - h))) / (2 * h)
        assert abs(fd(x, y, t, u) - fdv) <= 1e-7 * max(1.0, abs(fdv)), (fd(x, y, t, u), fdv)


def test_hiprtc_compiles_the_closure_kernels():
    src = jit.source_of(CLOSURES["exp"], CLOSURES["sin"])
    n = ctypes.c_size_t(0)
    rc = _lib.hip_lib().fvm_jit_check(src.encode(), 512, ctypes.byref(n))
    assert rc == 0, _lib.last_error()
    assert n.value > 10000
    bad = "__device__ double fvm_jit_D(double x) { return x; }\n"
    assert _lib.hip_lib().fvm_jit_check(bad.encode(), 512, ctypes.byref(n)) != 0
    assert "fvm_jit_D" in _lib.last_error()
