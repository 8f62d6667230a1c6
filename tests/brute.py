"""Brute-force control-volume integration, independent of the oracle's loops.

For every node the control volume is assembled from its quadrilateral pieces
p -> m(p,next) -> c -> m(p,prev) inside each incident triangle (the polygon of
test/test_functions.jl:1-34 get_control_volume), its area is the shoelace sum,
and du is minus the outward flux through the polygon's sides divided by the area
(test_functions.jl:601-646 get_dudt_val).  Shape functions come from solving
[x y 1] [a b g]^T = u (numpy), normals from the polygon's own edges.  Nothing is
shared with the oracle's triangle-edge loop or its precomputed geometry."""
import numpy as np


def _shape(px, py, uu):
    A = np.array([[px[0], py[0], 1.0], [px[1], py[1], 1.0], [px[2], py[2], 1.0]])
    return np.linalg.solve(A, uu)


def _outward(a, b):
    d = np.subtract(b, a)
    L = float(np.hypot(d[0], d[1]))
    return L, (d[1] / L, -d[0] / L), ((a[0] + b[0]) / 2, (a[1] + b[1]) / 2)


def control_volumes(x, y, triangles):
    area = np.zeros(len(x))
    for T in triangles:
        p = [(x[v], y[v]) for v in T]
        c = ((p[0][0] + p[1][0] + p[2][0]) / 3, (p[0][1] + p[1][1] + p[2][1]) / 3)
        for s in range(3):
            a, nx_, pr = p[s], p[(s + 1) % 3], p[(s + 2) % 3]
            poly = [a, ((a[0] + nx_[0]) / 2, (a[1] + nx_[1]) / 2), c,
                    ((a[0] + pr[0]) / 2, (a[1] + pr[1]) / 2)]
            A = 0.0
            for k in range(4):
                (x1, y1), (x2, y2) = poly[k], poly[(k + 1) % 4]
                A += x1 * y2 - x2 * y1
            area[T[s]] += A / 2
    return area


def dudt(x, y, triangles, bedges, badj, kind, fidx, bneu, flux, source, cond, u, t):
    """flux(x, y, t, a, b, g) -> (qx, qy); source(x, y, t, u); cond(f, x, y, t, u)."""
    N = len(x)
    out_flux = np.zeros(N)
    area = control_volumes(x, y, triangles)
    coef = {}
    for n, T in enumerate(triangles):
        px, py = x[T], y[T]
        a, b, g = _shape(px, py, u[T])
        coef[n] = (a, b, g)
        c = (px.sum() / 3, py.sum() / 3)
        for s in range(3):
            node = T[s]
            p = (px[s], py[s])
            mn = ((p[0] + px[(s + 1) % 3]) / 2, (p[1] + py[(s + 1) % 3]) / 2)
            mp = ((p[0] + px[(s + 2) % 3]) / 2, (p[1] + py[(s + 2) % 3]) / 2)
            for (aa, bb) in ((mn, c), (c, mp)):
                L, (nx, ny), (mx, my) = _outward(aa, bb)
                qx, qy = flux(mx, my, t, a, b, g)
                out_flux[node] += (qx * nx + qy * ny) * L
    for e, (i, j) in enumerate(bedges):
        a, b, g = coef[int(badj[e])]
        p, q = (x[i], y[i]), (x[j], y[j])
        m = ((p[0] + q[0]) / 2, (p[1] + q[1]) / 2)
        for node, (aa, bb) in ((i, (p, m)), (j, (m, q))):
            L, (nx, ny), (mx, my) = _outward(aa, bb)
            if bneu[e] >= 0:
                qn = cond(int(bneu[e]), mx, my, t, a * mx + b * my + g)
            else:
                qx, qy = flux(mx, my, t, a, b, g)
                qn = qx * nx + qy * ny
            out_flux[node] += qn * L
    du = np.zeros(N)
    isv = np.zeros(N, bool)
    isv[np.asarray(triangles).ravel()] = True
    for i in range(N):
        if not isv[i] or kind[i] == 1:
            du[i] = 0.0
        elif kind[i] == 2:
            du[i] = cond(int(fidx[i]), x[i], y[i], t, u[i])
        else:
            du[i] = -out_flux[i] / area[i] + source(x[i], y[i], t, u[i])
    return du, area
