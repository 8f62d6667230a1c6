"""MI355X-native vertex-centred finite-volume hot path of FiniteVolumeMethod.jl.

The public names mirror the reference module (src/FiniteVolumeMethod.jl:39-52 and
the template exports): FVMGeometry, FVMProblem, FVMSystem, SteadyFVMProblem,
BoundaryConditions, InternalConditions, Conditions, Neumann, Dudt, Dirichlet,
Constrained, solve, fvm_eqs (fvm_eqs!), compute_flux, pl_interpolate, find_triangle, DiffusionEquation, PoissonsEquation,
LaplacesEquation, MeanExitTimeProblem, LinearReactionDiffusionEquation.  Every numerical kernel runs on the GPU through
libfvmhip.so's C ABI (include/fvm_hip.h).
"""
from ._lib import DeviceArray, FVMError, synchronize  # noqa: F401
from .conditions import (BoundaryConditions, ConditionType, Conditions, Constrained,  # noqa: F401
                         Dirichlet, Dudt, InternalConditions, Neumann, ParametrisedFunction)
from .equations import (fvm_eqs, fvm_eqs_, get_fvm_parameters,  # noqa: F401
                        get_multithreading_parameters, get_serial_parameters, jacobian_sparsity,
                        update_dirichlet_nodes)
from .ensemble import EnsembleProblem, EnsembleSolution, explicit_dt  # noqa: F401
from .forms import NodalDiffusion, UnsupportedClosure  # noqa: F401
from .geometry import FVMGeometry, TriangleProperties  # noqa: F401
from .mesh import Triangulation, jittered_rectangle, triangulate, triangulate_rectangle  # noqa: F401
from .problem import FVMProblem, SteadyFVMProblem, construct_flux_function  # noqa: F401
from .system import FVMSystem, InvalidFluxError, is_system  # noqa: F401
from .queries import compute_flux, find_triangle, pl_interpolate  # noqa: F401
from .rules import NoTangent, Tangent, frule, rrule  # noqa: F401
from .solve import SSPRK3, TRBDF2, Euler, ODESolution, Tsit5, solve  # noqa: F401
from .templates import (AbstractFVMTemplate, BiCGStab, CG, DiffusionEquation,  # noqa: F401
                        LaplacesEquation, LinearReactionDiffusionEquation, LinearSolution,
                        MeanExitTimeProblem, MixedCG, PoissonsEquation)

__version__ = "0.1.0"
