// Pointwise queries of a solution on the mesh, batched: compute_flux (the normal
// flux across mesh edges) and pl_interpolate (the piecewise-linear interpolant).
//
// Reference: compute_flux, src/problem.jl:458-487, and pl_interpolate,
// src/utils.jl:23-27 (each takes one edge / one point).  Here one thread answers
// one query; the triangle a query needs is found on the device from the node ->
// triangle incidence of the mesh (find_directed_edge) in place of the reference's
// get_adjacent and triangle_props Dict lookups.  Unknown edges or triangles set a
// flag that the host turns into an error, as the reference's lookups throw.
#include <algorithm>
#include <cmath>

#include "fvm_internal.hpp"

using namespace fvm;

namespace {

struct FluxQuery {
  int64_t n, T, N;
  const int32_t* qi;
  const int32_t* qj;
  const int32_t* node_tri_start;
  const int32_t* node_tri;
  const int32_t* tri;
  const double* x;
  const double* y;
  const double* props;
  const double* u;
  const double* log_d;
  const double* dq;  // FVM_FLUX_D_EDGE: D at each query's edge midpoint; FVM_FLUX_LINEAR: a1..a6
  int32_t kind;
  double d0, d1, d2, d3;
  double* out;
  int32_t* bad;
};

// compute_flux (problem.jl:458-487): the triangle on the right of (i, j) (holding
// the directed edge j -> i), or, when (i, j) is a boundary edge with the domain on
// its left, the triangle holding i -> j; q at the edge midpoint from that
// triangle's shape function; the result q . n with n the clockwise normal of i -> j.
__global__ __launch_bounds__(256) void flux_query_kernel(FluxQuery a) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  const int32_t i = a.qi[q], j = a.qj[q];
  int32_t t = -1;
  if (i >= 0 && i < a.N && j >= 0 && j < a.N) {
    t = find_directed_edge(a.node_tri_start, a.node_tri, a.tri, a.T, j, i);
    if (t < 0) t = find_directed_edge(a.node_tri_start, a.node_tri, a.tri, a.T, i, j);
  }
  if (t < 0) {
    a.out[q] = NAN;
    a.bad[0] = 1;
    return;
  }
  const double px = a.x[i], py = a.y[i], qx = a.x[j], qy = a.y[j];
  const double ex = qx - px, ey = qy - py;
  const double l = sqrt(ex * ex + ey * ey);
  const double nx = ey / l, ny = -ex / l;
  const int32_t va = a.tri[t], vb = a.tri[a.T + t], vc = a.tri[2 * a.T + t];
  double s[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) s[f] = a.props[f * a.T + t];
  const double ua = a.u[va], ub = a.u[vb], uc = a.u[vc];
  const double al = s[0] * ua + s[1] * ub + s[2] * uc;
  const double be = s[3] * ua + s[4] * ub + s[5] * uc;
  const double ga = s[6] * ua + s[7] * ub + s[8] * uc;
  const double mx = (px + qx) / 2, my = (py + qy) / 2;
  if (a.kind == FVM_FLUX_LINEAR) {
    const double* A = a.dq + 6 * q;
    const double fx = A[0] * al + A[1] * be + A[2] * ga, fy = A[3] * al + A[4] * be + A[5] * ga;
    a.out[q] = nx * fx + ny * fy;
    return;
  }
  double D;
  if (a.kind == FVM_FLUX_D_POLY_U) {
    const double um = al * mx + be * my + ga;
    D = a.d0 + um * (a.d1 + um * (a.d2 + um * a.d3));
  } else if (a.kind == FVM_FLUX_D_POWER) {
    const double um = al * mx + be * my + ga;
    D = a.d0 + a.d1 * pow(fabs(um), a.d2);
  } else if (a.kind == FVM_FLUX_D_NODAL) {
    const double Da = exp(a.log_d[va]), Db = exp(a.log_d[vb]), Dc = exp(a.log_d[vc]);
    const double da = s[0] * Da + s[1] * Db + s[2] * Dc;
    const double db = s[3] * Da + s[4] * Db + s[5] * Dc;
    const double dg = s[6] * Da + s[7] * Db + s[8] * Dc;
    D = da * mx + db * my + dg;
  } else {
    D = a.dq[q];
  }
  const double fx = -D * al, fy = -D * be;
  a.out[q] = nx * fx + ny * fy;
}

struct InterpQuery {
  int64_t n, T, N;
  int32_t nv;
  const int32_t* qtri;
  const double* px;
  const double* py;
  const int32_t* node_tri_start;
  const int32_t* node_tri;
  const int32_t* tri;
  const double* props;
  const double* u;
  double* out;
  int32_t* bad;
};

// pl_interpolate (utils.jl:23-27): the shape function of the mesh triangle given as
// (i, j, k) in any rotation (the stored rotation's coefficients, as
// _safe_get_triangle_props finds them), evaluated at (x, y) for every variable.
__global__ __launch_bounds__(256) void interp_query_kernel(InterpQuery a) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  const int32_t i = a.qtri[3 * q], j = a.qtri[3 * q + 1], k = a.qtri[3 * q + 2];
  const bool inside = i >= 0 && i < a.N && j >= 0 && j < a.N;
  const int32_t t = inside ? find_directed_edge(a.node_tri_start, a.node_tri, a.tri, a.T, i, j) : -1;
  bool ok = t >= 0;
  int32_t va = 0, vb = 0, vc = 0;
  if (ok) {
    va = a.tri[t];
    vb = a.tri[a.T + t];
    vc = a.tri[2 * a.T + t];
    ok = (int64_t)va + vb + vc - i - j == k;  // the third vertex
  }
  if (!ok) {
    for (int v = 0; v < a.nv; ++v) a.out[q * a.nv + v] = NAN;
    a.bad[0] = 1;
    return;
  }
  double s[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) s[f] = a.props[f * a.T + t];
  const double x = a.px[q], y = a.py[q];
  for (int v = 0; v < a.nv; ++v) {
    const double ua = a.u[v + (int64_t)a.nv * va], ub = a.u[v + (int64_t)a.nv * vb],
                 uc = a.u[v + (int64_t)a.nv * vc];
    const double al = s[0] * ua + s[1] * ub + s[2] * uc;
    const double be = s[3] * ua + s[4] * ub + s[5] * uc;
    const double ga = s[6] * ua + s[7] * ub + s[8] * uc;
    a.out[q * a.nv + v] = al * x + be * y + ga;
  }
}

// ----------------------------------------------------------------------------
// point location (DelaunayTriangulation's jump_and_march, which the reference's
// tutorials call before pl_interpolate)
// ----------------------------------------------------------------------------
struct LocQuery {
  int64_t n, T, N;
  int32_t max_steps;
  const double* px;
  const double* py;
  const int32_t* start;  // a start vertex per query (may be null: vertex 0)
  const int32_t* node_tri_start;
  const int32_t* node_tri;
  const int32_t* tri;
  const int32_t* perm;  // sorted -> input triangle
  const double* x;
  const double* y;
  int32_t* out;  // input triangle index, -1 outside the mesh, -2 while undecided
  int32_t convex;  // the domain is convex: a walk stopped by the boundary means outside
  int64_t q0;      // first query of this scan launch
};

__device__ __forceinline__ double orient(double ax, double ay, double bx, double by, double px, double py) {
  return (bx - ax) * (py - ay) - (by - ay) * (px - ax);
}

__device__ __forceinline__ bool contains(const LocQuery& a, int32_t t, double px, double py, double* o) {
  const int32_t v[3] = {a.tri[t], a.tri[a.T + t], a.tri[2 * a.T + t]};
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int32_t p0 = v[e], p1 = v[e == 2 ? 0 : e + 1];
    o[e] = orient(a.x[p0], a.y[p0], a.x[p1], a.y[p1], px, py);
  }
  return o[0] >= 0.0 && o[1] >= 0.0 && o[2] >= 0.0;
}

// Visibility walk from a triangle of the start vertex: cross the edge the point
// lies most outside of (or another edge it lies outside of when that one is on the
// boundary).  A walk stopped by the boundary of a convex domain has found the point
// outside (-1): it lies strictly on the outer side of a boundary edge.  Otherwise a
// walk that leaves through the boundary or runs out of steps marks the query
// undecided (-2): a non-convex domain can trap a straight walk.
__global__ __launch_bounds__(256) void locate_walk_kernel(LocQuery a) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  const double px = a.px[q], py = a.py[q];
  int32_t v0 = a.start ? a.start[q] : 0;
  if (v0 < 0 || v0 >= a.N) v0 = 0;
  int32_t k = a.node_tri_start[v0];
  if (k == a.node_tri_start[v0 + 1]) {  // not a vertex of the mesh
    a.out[q] = -2;
    return;
  }
  int32_t t = a.node_tri[k] >> 2;
  for (int step = 0; step < a.max_steps; ++step) {
    double o[3];
    if (contains(a, t, px, py, o)) {
      a.out[q] = a.perm[t];
      return;
    }
    const int32_t v[3] = {a.tri[t], a.tri[a.T + t], a.tri[2 * a.T + t]};
    // the outside edges, most negative first
    int order[3] = {0, 1, 2};
    for (int i = 0; i < 2; ++i)
      for (int j = i + 1; j < 3; ++j)
        if (o[order[j]] < o[order[i]]) {
          const int tmp = order[i];
          order[i] = order[j];
          order[j] = tmp;
        }
    int32_t next = -1;
    for (int i = 0; i < 3 && next < 0; ++i) {
      const int e = order[i];
      if (!(o[e] < 0.0)) break;
      // the neighbour across edge v[e] -> v[e+1] holds the directed edge v[e+1] -> v[e]
      next = find_directed_edge(a.node_tri_start, a.node_tri, a.tri, a.T, v[e == 2 ? 0 : e + 1], v[e]);
    }
    if (next < 0) {
      a.out[q] = a.convex ? -1 : -2;
      return;
    }
    t = next;
  }
  a.out[q] = -2;
}

// Undecided queries: one workgroup scans every triangle; the smallest sorted index
// containing the point wins (an LDS minimum: the same answer on every run), -1 if none.
__global__ __launch_bounds__(256) void locate_scan_kernel(LocQuery a) {
  const int64_t q = a.q0 + blockIdx.x;
  if (a.out[q] != -2) return;
  __shared__ int32_t best;
  if (threadIdx.x == 0) best = 0x7fffffff;
  __syncthreads();
  const double px = a.px[q], py = a.py[q];
  for (int64_t t = threadIdx.x; t < a.T; t += blockDim.x) {
    double o[3];
    if (contains(a, (int32_t)t, px, py, o)) {
      atomicMin(&best, (int32_t)t);
      break;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) a.out[q] = best == 0x7fffffff ? -1 : a.perm[best];
}

}  // namespace


extern "C" {

int fvm_locate_points(const fvm_mesh* m, int64_t n, const double* d_x, const double* d_y, const int32_t* d_start,
                      int32_t* d_tri_out, fvm_stream s) {
  FVM_CHECK(m && n >= 0, "fvm_locate_points: bad argument");
  if (n == 0) return FVM_OK;
  FVM_CHECK(d_x && d_y && d_tri_out, "fvm_locate_points: null argument");
  FVM_CHECK(n <= 0x7fffffff, "fvm_locate_points: at most 2^31 - 1 points per call");
  LocQuery a;
  a.n = n;
  a.T = m->T;
  a.N = m->N;
  a.max_steps = (int32_t)std::min<int64_t>(4096, 8 + 4 * (int64_t)std::sqrt((double)m->T));
  a.px = d_x;
  a.py = d_y;
  a.start = d_start;
  a.node_tri_start = m->node_tri_start;
  a.node_tri = m->node_tri;
  a.tri = m->tri;
  a.perm = m->perm;
  a.x = m->x;
  a.y = m->y;
  a.out = d_tri_out;
  a.convex = m->convex ? 1 : 0;
  a.q0 = 0;
  hipStream_t st = (hipStream_t)s;
  hipLaunchKernelGGL(locate_walk_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
  // one workgroup per query, at most 2^22 workgroups (2^30 work-items) per launch;
  // workgroups of decided queries return at once
  constexpr int64_t kChunk = (int64_t)1 << 22;
  for (int64_t q0 = 0; q0 < n; q0 += kChunk) {
    a.q0 = q0;
    hipLaunchKernelGGL(locate_scan_kernel, dim3((unsigned)std::min(kChunk, n - q0)), dim3(256), 0, st, a);
  }
  FVM_HIP(hipGetLastError());
  FVM_HIP(hipStreamSynchronize(st));
  return FVM_OK;
}

int fvm_compute_flux(fvm_ctx* c, int64_t n, const int32_t* d_i, const int32_t* d_j, const double* d_u,
                     double t, const double* d_dq, double* d_out, fvm_stream s) {
  (void)t;  // the traced flux forms do not depend on t
  FVM_CHECK(c && n >= 0, "fvm_compute_flux: bad argument");
  FVM_CHECK(c->flux_kind != FVM_FLUX_JIT,
            "fvm_compute_flux: not available for closures compiled at run time (FVM_FLUX_JIT)");
  if (n == 0) return FVM_OK;
  FVM_CHECK(d_i && d_j && d_u && d_out, "fvm_compute_flux: null argument");
  FVM_CHECK((c->flux_kind != FVM_FLUX_D_EDGE && c->flux_kind != FVM_FLUX_LINEAR) || d_dq,
            "fvm_compute_flux: FVM_FLUX_D_EDGE / FVM_FLUX_LINEAR need the flux coefficients at the edge "
            "midpoints (d_dq)");
  const fvm_mesh* m = c->m;
  FluxQuery a;
  a.n = n;
  a.T = m->T;
  a.N = m->N;
  a.qi = d_i;
  a.qj = d_j;
  a.node_tri_start = m->node_tri_start;
  a.node_tri = m->node_tri;
  a.tri = m->tri;
  a.x = m->x;
  a.y = m->y;
  a.props = m->props;
  a.u = d_u;
  a.log_d = c->log_d;
  a.dq = d_dq;
  a.kind = c->flux_kind;
  a.d0 = c->dpoly[0];
  a.d1 = c->dpoly[1];
  a.d2 = c->dpoly[2];
  a.d3 = c->dpoly[3];
  a.out = d_out;
  hipStream_t st = (hipStream_t)s;
  return run_flagged(st, "fvm_compute_flux", m->qflag, [&](int32_t* bad) {
    a.bad = bad;
    hipLaunchKernelGGL(flux_query_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
  });
}

int fvm_pl_interpolate(const fvm_mesh* m, int32_t n_vars, int64_t n, const int32_t* d_tri,
                       const double* d_x, const double* d_y, const double* d_u, double* d_out,
                       fvm_stream s) {
  FVM_CHECK(m && n >= 0 && n_vars >= 1, "fvm_pl_interpolate: bad argument");
  if (n == 0) return FVM_OK;
  FVM_CHECK(d_tri && d_x && d_y && d_u && d_out, "fvm_pl_interpolate: null argument");
  InterpQuery a;
  a.n = n;
  a.T = m->T;
  a.N = m->N;
  a.nv = n_vars;
  a.qtri = d_tri;
  a.px = d_x;
  a.py = d_y;
  a.node_tri_start = m->node_tri_start;
  a.node_tri = m->node_tri;
  a.tri = m->tri;
  a.props = m->props;
  a.u = d_u;
  a.out = d_out;
  hipStream_t st = (hipStream_t)s;
  return run_flagged(st, "fvm_pl_interpolate", m->qflag, [&](int32_t* bad) {
    a.bad = bad;
    hipLaunchKernelGGL(interp_query_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
  });
}

}  // extern "C"
