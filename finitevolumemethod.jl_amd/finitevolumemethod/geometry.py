"""FVMGeometry (reference src/geometry.jl:1-169) with device-resident geometry.

`FVMGeometry(tri)` uploads the triangulation, cuts it into triangle tiles and
computes every TriangleProperties field and the control-volume areas on the GPU
(fvm_mesh_create).  The host views (`cv_volumes`, `triangle_props`) are copied
back lazily, in the input triangle order, for inspection and tests.
"""
import ctypes
from collections import namedtuple

import numpy as np

from . import _lib

TriangleProperties = namedtuple(
    "TriangleProperties",
    ["shape_function_coefficients", "cv_edge_midpoints", "cv_edge_normals", "cv_edge_lengths"])
TriangleProperties.__doc__ = """geometry.jl:21-26: s1..s9, the midpoints of the three CV edges
(edges (i,j), (j,k), (k,i) joined to the centroid), their unit normals and lengths."""


class _PropsView:
    """Dict-like view {(i, j, k): TriangleProperties} over the downloaded SoA."""

    def __init__(self, tri, props):
        self._tri = tri
        self._props = props
        self._index = None

    def _idx(self):
        if self._index is None:
            self._index = {tuple(int(v) for v in T): n for n, T in enumerate(self._tri)}
        return self._index

    def __len__(self):
        return len(self._tri)

    def __contains__(self, key):
        return tuple(key) in self._idx()

    def keys(self):
        return self._idx().keys()

    def __getitem__(self, key):
        P = self._props[self._idx()[tuple(key)]]
        return TriangleProperties(tuple(P[0:9]), ((P[9], P[10]), (P[11], P[12]), (P[13], P[14])),
                                  ((P[15], P[16]), (P[17], P[18]), (P[19], P[20])),
                                  (P[21], P[22], P[23]))


class FVMGeometry:
    """Mesh plus control-volume geometry, resident on the GPU.

    tile_triangles / tile_max_nodes tune the triangle tiles the kernels stream
    (0 keeps the defaults: 512 triangles, 512 nodes)."""

    def __init__(self, tri, tile_triangles=0, tile_max_nodes=0):
        _lib.require_gpu()
        self.triangulation = tri
        L = _lib.hip_lib()
        d = _lib.MeshDesc(len(tri.x), len(tri.triangles), len(tri.boundary_edges),
                          _lib.ptr(tri.x), _lib.ptr(tri.y), _lib.ptr(tri.triangles, _lib._i32p),
                          _lib.ptr(tri.boundary_edges, _lib._i32p), int(tile_triangles),
                          int(tile_max_nodes))
        h = ctypes.c_void_p()
        _lib.check(L.fvm_mesh_create(ctypes.byref(d), ctypes.byref(h)), "fvm_mesh_create")
        self._handle = h.value
        info = _lib.MeshInfo()
        _lib.check(L.fvm_mesh_get_info(self._handle, ctypes.byref(info)), "fvm_mesh_get_info")
        self.info = info
        self._cv = None
        self._props = None

    @property
    def handle(self):
        return self._handle

    def _download(self):
        n, t = self.info.n_points, self.info.n_triangles
        cv = np.empty(n)
        props = np.empty((t, 24))
        _lib.check(_lib.hip_lib().fvm_mesh_download_geometry(self._handle, _lib.ptr(cv),
                                                             _lib.ptr(props)),
                   "fvm_mesh_download_geometry")
        self._cv, self._props = cv, props

    @property
    def cv_volumes(self):
        if self._cv is None:
            self._download()
        return self._cv

    @property
    def props_array(self):
        """[T, 24] TriangleProperties fields in the triangulation's triangle order."""
        if self._props is None:
            self._download()
        return self._props

    @property
    def triangle_props(self):
        return _PropsView(self.triangulation.triangles, self.props_array)

    def get_volume(self, i):
        return float(self.cv_volumes[i])

    def get_triangle_props(self, i, j, k):
        return self.triangle_props[(i, j, k)]

    def get_point(self, i):
        return self.triangulation.get_point(i)

    def num_points(self):
        return self.triangulation.num_points()

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and _lib._hip is not None:
            _lib._hip.fvm_mesh_destroy(h)
            self._handle = None

    def __repr__(self):
        # geometry.jl:50-55; each interior edge is shared by two triangles
        tri = self.triangulation
        ne = (3 * tri.num_solid_triangles() + len(tri.boundary_edges)) // 2
        return (f"FVMGeometry with {tri.num_solid_vertices()} control volumes, "
                f"{tri.num_solid_triangles()} triangles, and {ne} edges")
