"""Boundary and internal conditions (reference src/conditions.jl).

Mirrors ConditionType, ParametrisedFunction, BoundaryConditions,
InternalConditions and Conditions.  Function indices are 0-based: the merged
function list is (internal functions..., boundary functions...), as in
prepare_conditions (conditions.jl:486-504), and boundary segment k uses index
k + n_internal (conditions.jl:518).
"""
import enum

import numpy as np

from . import _lib


class ConditionType(enum.Enum):
    """conditions.jl:36-41."""
    Neumann = 0
    Dudt = 1
    Dirichlet = 2
    Constrained = 3


Neumann = ConditionType.Neumann
Dudt = ConditionType.Dudt
Dirichlet = ConditionType.Dirichlet
Constrained = ConditionType.Constrained


class ParametrisedFunction:
    """f(x, y, t, u) = fnc(x, y, t, u, parameters) (conditions.jl:16-20)."""

    __slots__ = ("fnc", "parameters")

    def __init__(self, fnc, parameters=None):
        self.fnc = fnc
        self.parameters = parameters

    def __call__(self, *args):
        return self.fnc(*args, self.parameters)


def _as_tuple(v):
    if isinstance(v, (tuple, list)):
        return tuple(v)
    return (v,)


class BoundaryConditions:
    """BoundaryConditions(mesh, functions, types; parameters) (conditions.jl:237-251).

    functions[k] / types[k] apply to boundary segment k of the mesh."""

    def __init__(self, mesh, functions, types, parameters=None):
        single = not isinstance(functions, (tuple, list))
        functions = _as_tuple(functions)
        types = _as_tuple(types)
        if parameters is None:
            parameters = (None,) * len(functions)
        elif single:
            parameters = (parameters,)
        parameters = tuple(parameters)
        nseg = mesh.triangulation.num_boundary_segments()
        if len(functions) != nseg:
            raise AssertionError("The number of boundary conditions must be the same as the number "
                                 "of parts of the mesh's boundary.")
        if len(functions) != len(types):
            raise AssertionError("The number of functions and types must be the same.")
        if not all(isinstance(t, ConditionType) for t in types):
            raise AssertionError("The condition types must be ConditionType instances.")
        if len(parameters) != len(functions):
            raise AssertionError("One parameter per function.")
        self.functions = tuple(ParametrisedFunction(f, p) for f, p in zip(functions, parameters))
        self.condition_types = types

    def __repr__(self):
        n = len(self.functions)
        names = tuple(t.name for t in self.condition_types)
        if n > 1:
            return f"BoundaryConditions with {n} boundary conditions with types {names}"
        return f"BoundaryConditions with {n} boundary condition with type {names[0]}"


class InternalConditions:
    """InternalConditions(functions; dirichlet_nodes, dudt_nodes, parameters)
    (conditions.jl:253-269).  Node -> index into functions (0-based)."""

    def __init__(self, functions=(), dirichlet_nodes=None, dudt_nodes=None, parameters=None):
        single = not isinstance(functions, (tuple, list))
        functions = _as_tuple(functions)
        if parameters is None:
            parameters = (None,) * len(functions)
        elif single:
            parameters = (parameters,)
        parameters = tuple(parameters)
        self.functions = tuple(ParametrisedFunction(f, p) for f, p in zip(functions, parameters))
        self.dirichlet_nodes = dict(dirichlet_nodes or {})
        self.dudt_nodes = dict(dudt_nodes or {})

    def __repr__(self):
        return (f"InternalConditions with {len(self.dirichlet_nodes)} Dirichlet nodes and "
                f"{len(self.dudt_nodes)} Dudt nodes")


class Conditions:
    """Merged conditions (conditions.jl:310-324, 486-552)."""

    def __init__(self, mesh, bc, ic=None):
        ic = ic if ic is not None else InternalConditions()
        tri = mesh.triangulation
        nif = len(ic.functions)
        self.functions = ic.functions + bc.functions
        self.dirichlet_nodes = dict(ic.dirichlet_nodes)
        self.dudt_nodes = dict(ic.dudt_nodes)
        self.neumann_edges = {}
        self.constrained_edges = {}
        for seg in range(tri.num_boundary_segments()):
            cond = bc.condition_types[seg]
            fidx = seg + nif
            for e in np.nonzero(tri.boundary_segments == seg)[0]:
                u, v = int(tri.boundary_edges[e, 0]), int(tri.boundary_edges[e, 1])
                if cond is Neumann:
                    self.neumann_edges[(u, v)] = fidx
                elif cond is Constrained:
                    self.constrained_edges[(u, v)] = fidx
                elif cond is Dirichlet:
                    self.dirichlet_nodes[u] = fidx
                    self.dirichlet_nodes[v] = fidx
                else:
                    self.dudt_nodes[u] = fidx
                    self.dudt_nodes[v] = fidx
        self.n_internal = nif

    # conditions.jl:337-484
    def get_dudt_fidx(self, node):
        return self.dudt_nodes[node]

    def get_neumann_fidx(self, i, j):
        return self.neumann_edges[(i, j)]

    def get_dirichlet_fidx(self, node):
        return self.dirichlet_nodes[node]

    def get_constrained_fidx(self, i, j):
        return self.constrained_edges[(i, j)]

    def eval_condition_fnc(self, fidx, x, y, t, u):
        return self.functions[fidx](x, y, t, u)

    def is_dudt_node(self, node):
        return node in self.dudt_nodes

    def is_neumann_edge(self, i, j):
        return (i, j) in self.neumann_edges

    def is_dirichlet_node(self, node):
        return node in self.dirichlet_nodes

    def is_constrained_edge(self, i, j):
        return (i, j) in self.constrained_edges

    def has_constrained_edges(self):
        return bool(self.constrained_edges)

    def has_neumann_edges(self):
        return bool(self.neumann_edges)

    def has_condition(self, node):
        return self.is_dudt_node(node) or self.is_dirichlet_node(node)

    def has_dirichlet_nodes(self):
        return bool(self.dirichlet_nodes)

    def has_dudt_nodes(self):
        return bool(self.dudt_nodes)

    def get_dirichlet_nodes(self):
        return self.dirichlet_nodes

    def get_dudt_nodes(self):
        return self.dudt_nodes

    def get_neumann_edges(self):
        return self.neumann_edges

    def get_constrained_edges(self):
        return self.constrained_edges

    def flatten(self, tri):
        """Per-node kind (0 none, 1 Dirichlet, 2 Dudt; Dirichlet takes precedence,
        source_contributions.jl:5-12), per-node function index, and the Neumann
        function index of every boundary edge (-1 if not Neumann)."""
        n = tri.num_points()
        kind = np.zeros(n, np.int8)
        fidx = np.full(n, -1, np.int32)
        for i, f in self.dudt_nodes.items():
            kind[i], fidx[i] = _lib.NODE_DUDT, f
        for i, f in self.dirichlet_nodes.items():
            kind[i], fidx[i] = _lib.NODE_DIRICHLET, f
        bn = np.full(len(tri.boundary_edges), -1, np.int32)
        for e, (u, v) in enumerate(tri.boundary_edges):
            f = self.neumann_edges.get((int(u), int(v)))
            if f is not None:
                bn[e] = f
        return kind, fidx, bn

    def __repr__(self):
        return (f"Conditions with\n   {len(self.neumann_edges)} Neumann edges\n"
                f"   {len(self.constrained_edges)} Constrained edges\n"
                f"   {len(self.dirichlet_nodes)} Dirichlet nodes\n   {len(self.dudt_nodes)} Dudt nodes")
