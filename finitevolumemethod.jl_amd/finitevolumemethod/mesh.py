"""Triangulations consumed by FVMGeometry.

The reference builds its meshes with DelaunayTriangulation.jl (not part of this
package's scope).  `Triangulation` holds what the finite-volume path reads from
one: the points, the counter-clockwise solid triangles, and the boundary edges
(counter-clockwise, interior on the left) with their boundary segment index.
Generators: `triangulate_rectangle` (DelaunayTriangulation.triangulate_rectangle's
layout), `jittered_rectangle` (interior points jittered, Delaunay re-triangulated)
and `triangulate` (Delaunay of a point set with its convex hull as the boundary).
All indices are 0-based: reference node k is node k-1 here.
"""
import ctypes

import numpy as np

from . import _lib


class Triangulation:
    """Points, solid triangles and oriented boundary edges of a planar mesh."""

    def __init__(self, x, y, triangles, boundary_edges, boundary_segments=None):
        self.x = np.ascontiguousarray(x, np.float64)
        self.y = np.ascontiguousarray(y, np.float64)
        self.triangles = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        self.boundary_edges = np.ascontiguousarray(boundary_edges, np.int32).reshape(-1, 2)
        if boundary_segments is None:
            boundary_segments = np.zeros(len(self.boundary_edges), np.int32)
        self.boundary_segments = np.ascontiguousarray(boundary_segments, np.int32)
        if len(self.x) != len(self.y):
            raise ValueError("x and y must have the same length")
        if len(self.boundary_segments) != len(self.boundary_edges):
            raise ValueError("one boundary segment index per boundary edge")
        n = len(self.x)
        if self.triangles.size and (self.triangles.min() < 0 or self.triangles.max() >= n):
            raise ValueError("triangle vertex out of range")
        if self.boundary_edges.size and (self.boundary_edges.min() < 0 or self.boundary_edges.max() >= n):
            raise ValueError("boundary edge vertex out of range")
        self._is_vertex = None

    # -- DelaunayTriangulation-like accessors ---------------------------------
    @property
    def points(self):
        return np.stack([self.x, self.y], axis=1)

    def num_points(self):
        return len(self.x)

    def num_solid_triangles(self):
        return len(self.triangles)

    def num_boundary_segments(self):
        """Number of boundary parts (DelaunayTriangulation's ghost vertices)."""
        return int(self.boundary_segments.max()) + 1 if len(self.boundary_segments) else 0

    def num_solid_vertices(self):
        return int(self.is_vertex().sum())

    def is_vertex(self):
        if self._is_vertex is None:
            v = np.zeros(len(self.x), bool)
            v[self.triangles.ravel()] = True
            self._is_vertex = v
        return self._is_vertex

    def get_point(self, i):
        return float(self.x[i]), float(self.y[i])

    def each_solid_triangle(self):
        return (tuple(int(v) for v in T) for T in self.triangles)

    def boundary_nodes(self):
        return np.unique(self.boundary_edges.ravel())

    def boundary_adjacent(self):
        """Index of the triangle holding each boundary edge i->j (get_adjacent)."""
        badj = np.empty(len(self.boundary_edges), np.int32)
        if len(badj):
            L = _lib.host_lib()
            st = L.fvmh_boundary_adjacent(len(self.triangles), _lib.ptr(self.triangles, _lib._i32p),
                                          len(self.boundary_edges),
                                          _lib.ptr(self.boundary_edges, _lib._i32p),
                                          _lib.ptr(badj, _lib._i32p))
            if st != 0:
                raise ValueError(L.fvmh_last_error().decode())
        return badj

    def count_non_delaunay(self, rel_tol=1e-10):
        L = _lib.host_lib()
        return int(L.fvmh_count_non_delaunay(len(self.x), _lib.ptr(self.x), _lib.ptr(self.y),
                                             len(self.triangles), _lib.ptr(self.triangles, _lib._i32p),
                                             rel_tol))

    def neighbours(self, i):
        rows = np.nonzero((self.triangles == i).any(axis=1))[0]
        nb = np.unique(self.triangles[rows].ravel())
        return nb[nb != i]

    def __repr__(self):
        return (f"Triangulation with {self.num_points()} points, {self.num_solid_triangles()} "
                f"triangles, {len(self.boundary_edges)} boundary edges")


def _rectangle(a, b, c, d, nx, ny, single_boundary, jitter, seed):
    L = _lib.host_lib()
    n, t, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    ns = ctypes.c_int32()
    if L.fvmh_rectangle_sizes(nx, ny, int(single_boundary), ctypes.byref(n), ctypes.byref(t),
                              ctypes.byref(e), ctypes.byref(ns)) != 0:
        raise ValueError(L.fvmh_last_error().decode())
    x = np.empty(n.value)
    y = np.empty(n.value)
    tri = np.empty((t.value, 3), np.int32)
    be = np.empty((e.value, 2), np.int32)
    bs = np.empty(e.value, np.int32)
    st = L.fvmh_rectangle(float(a), float(b), float(c), float(d), int(nx), int(ny),
                          int(bool(single_boundary)), float(jitter), ctypes.c_uint64(int(seed)),
                          _lib.ptr(x), _lib.ptr(y), _lib.ptr(tri, _lib._i32p),
                          _lib.ptr(be, _lib._i32p), _lib.ptr(bs, _lib._i32p))
    if st != 0:
        raise ValueError(L.fvmh_last_error().decode())
    return Triangulation(x, y, tri, be, bs)


def triangulate_rectangle(a, b, c, d, nx, ny, single_boundary=False):
    """DelaunayTriangulation.triangulate_rectangle(a, b, c, d, nx, ny; single_boundary).

    Points x-fastest (node i + nx*j at (x_i, y_j) with Julia's LinRange values);
    each cell split along its BR-TL diagonal; boundary segments bottom, right,
    top, left (one segment when single_boundary)."""
    return _rectangle(a, b, c, d, nx, ny, single_boundary, 0.0, 0)


def jittered_rectangle(a, b, c, d, nx, ny, jitter=0.25, seed=20261202, single_boundary=True):
    """Rectangle grid whose interior points move by U(-jitter*h, jitter*h) per axis
    (SplitMix64(seed), node order), re-triangulated as the Delaunay triangulation."""
    return _rectangle(a, b, c, d, nx, ny, single_boundary, jitter, seed)


def triangulate(x, y, boundary_segments_of=None):
    """Delaunay triangulation of a point set (scipy Qhull) with the convex hull as
    a single boundary; triangles and hull edges are oriented counter-clockwise.
    boundary_segments_of(i, j) -> int may split the hull into segments."""
    from scipy.spatial import Delaunay

    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    dl = Delaunay(np.stack([x, y], 1))
    tri = dl.simplices.astype(np.int32)
    px, py = x[tri], y[tri]
    area2 = (px[:, 1] - px[:, 0]) * (py[:, 2] - py[:, 0]) - (px[:, 2] - px[:, 0]) * (py[:, 1] - py[:, 0])
    flip = area2 < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    keep = np.abs(area2) > 0
    tri = tri[keep]
    # boundary edges: directed edges of triangles whose reverse is absent
    d = {}
    for T in tri:
        for s in range(3):
            d[(int(T[s]), int(T[(s + 1) % 3]))] = True
    be = [(i, j) for (i, j) in d if (j, i) not in d]
    # order as a counter-clockwise cycle
    nxt = {i: j for i, j in be}
    start = min(nxt)
    cyc = [start]
    while nxt[cyc[-1]] != start:
        cyc.append(nxt[cyc[-1]])
    edges = np.array([(cyc[k], cyc[(k + 1) % len(cyc)]) for k in range(len(cyc))], np.int32)
    seg = np.zeros(len(edges), np.int32)
    if boundary_segments_of is not None:
        seg = np.array([boundary_segments_of(int(i), int(j)) for i, j in edges], np.int32)
    return Triangulation(x, y, tri, edges, seg)
