"""CPU oracle for FiniteVolumeMethod.jl's vertex-centred FVM hot path.

TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg compare the HIP product against.  The product never
imports this package.

`libfvm_oracle.so` (built from fvm_oracle.c by oracle/Makefile) restates the
reference's arithmetic in C; this module wraps it with ctypes and adds the pure-
Python restatement of the reference's condition merging (conditions.jl:486-552).

Pinning (what fixes each function independently of this restatement) is listed in
the header of each test module; functions none of them pins say
"parity unpinned" there.
"""
from .oracle import *  # noqa: F401,F403
