// FVMSystem on the device: fvm_eqs! for nv coupled variables on one mesh.
//
// Reference: src/problem.jl:196-413 (FVMSystem, merge_problem_conditions) and the
// system methods of src/equations/*.jl: shape functions per variable
// (shape_functions.jl:12-19), every variable's flux from one flux evaluation per
// CV edge (individual_flux_contributions.jl:36-50), du[var, i] updates
// (triangle_contributions.jl:18-25, boundary_edge_contributions.jl:15-72),
// per-variable source / condition rules (source_contributions.jl:17-30, 49-68) and
// the variable-by-variable Dirichlet update (dirichlet.jl:46-53).
//
// Device forms of the closures (traced on the host):
//   flux    q_v = -sum_w K[v][w] (alpha_w, beta_w)           (constant K)
//   source  S_v = s_node[v][i] + sum of monomials c prod_w u_w^e_w + st_v t
//   cond    a   = c0 + sum_w cu[w] u_w + ct t + cx x + cy y + g exp(lam t)
// The state is the reference's matrix u[var, i] (column-major: var + nv i).
//
// Layout and scheme follow the scalar tile kernels: one workgroup per tile, the
// triangle records rebuilt from the vertex coordinates, nv fluxes per CV edge in
// LDS, per-node sums in a fixed order, exclusive nodes finished in the tile
// kernel, shared nodes through nv partial slots and a fix-up kernel.  No atomics.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fvm_internal.hpp"

using namespace fvm;

namespace {

constexpr int kMaxVars = 4;
constexpr int kMaxTerms = 24;

struct SysSource {
  int32_t n;
  int32_t var[kMaxTerms];
  int8_t e[kMaxTerms][kMaxVars];
  double c[kMaxTerms];
};

struct SysArgs {
  int32_t nv;
  int32_t n_tiles, n_fix, n_bitems;
  int64_t N, T, B;
  const int32_t* tile_tri_start;
  const int32_t* tile_node_start;
  const int32_t* tile_nodes;
  const uint32_t* loc;
  const int32_t* inc_start;
  const uint16_t* inc;
  const int32_t* fix_nodes;
  const int32_t* fix_meta;  // [4][n_fix]: node, first slot, first boundary item,
                            // slots << 20 | items << 8 | the variables' codes, 2 bits each
  const double* fix_cv;     // [n_fix]
  const int32_t* fix_part_start;
  const int32_t* fix_bedge_start;
  const int32_t* fix_bedge;
  const double* props;
  const int32_t* tri;
  const double* x;
  const double* y;
  const double* cv;
  const int32_t* bedges;
  const int32_t* badj;
  const double* lcv;
  const int32_t* lpart;   // partial slot of a local entry, -1 for an exclusive node
  const int8_t* code;     // [nv][N] 0 none, 1 Dirichlet, 2 Dudt, 3 not a vertex
  const int32_t* fidx;    // [nv][N]
  const int32_t* bneu;    // [nv][B]
  double K[kMaxVars][kMaxVars];
  SysSource S;
  const double* s_node;   // [nv][N] or null
  double st[kMaxVars];
  int32_t n_fn;
  const double* fc0;      // [n_fn]
  const double* fcu;      // [n_fn][nv]
  const double* fct;
  const double* fcx;
  const double* fcy;
  const double* flam;
  const double* node_g;   // [nv][N] or null
  const double* bedge_g;  // [nv][2B] or null
  const double* u;
  const double* v;        // tangent direction (Jacobian-vector products)
  double* out;
  double* partial;        // [n_parts][nv]
  double* bflux;          // [n_bitems][nv]
  double t;
};

// a(x, y, t, u[0..nv)) of condition function f
template <int NV>
__device__ __forceinline__ double cond_eval(const SysArgs& a, int f, double x, double y, double t,
                                            const double* u, double g) {
  double v = a.fc0[f];
#pragma unroll
  for (int w = 0; w < NV; ++w) v = v + a.fcu[f * NV + w] * u[w];
  v = v + a.fct[f] * t;
  v = v + a.fcx[f] * x;
  v = v + a.fcy[f] * y;
  const double lam = a.flam[f];
  return lam != 0.0 ? v + g * exp(lam * t) : v + g;
}

template <int NV>
__device__ __forceinline__ double source_eval(const SysArgs& a, int v, int g, const double* u) {
  double S = a.s_node ? a.s_node[(int64_t)v * a.N + g] : 0.0;
  for (int k = 0; k < a.S.n; ++k) {
    if (a.S.var[k] != v) continue;
    double m = a.S.c[k];
#pragma unroll
    for (int w = 0; w < NV; ++w)
      for (int p = 0; p < a.S.e[k][w]; ++p) m = m * u[w];
    S = S + m;
  }
  if (a.st[v] != 0.0) S = S + a.st[v] * a.t;
  return S;
}

// sum_w dS_v/du_w (u) dv_w, the derivative of variable v's source monomials
template <int NV>
__device__ __forceinline__ double source_tangent(const SysArgs& a, int v, const double* u, const double* dv) {
  double S = 0.0;
  for (int k = 0; k < a.S.n; ++k) {
    if (a.S.var[k] != v) continue;
#pragma unroll
    for (int w = 0; w < NV; ++w) {
      const int ew = a.S.e[k][w];
      if (ew == 0) continue;
      double m = a.S.c[k] * ew;
#pragma unroll
      for (int x = 0; x < NV; ++x) {
        const int ex = x == w ? ew - 1 : a.S.e[k][x];
        for (int p = 0; p < ex; ++p) m = m * u[x];
      }
      S = S + m * dv[w];
    }
  }
  return S;
}

// sum_w da/du_w dv_w of condition function f (the conditions are affine in u)
template <int NV>
__device__ __forceinline__ double cond_tangent(const SysArgs& a, int f, const double* dv) {
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < NV; ++w) r = r + a.fcu[f * NV + w] * dv[w];
  return r;
}

// source_contributions.jl:49-68 for variable v of node g (code c), or (JV) its
// derivative along dv
template <int NV, int JV>
__device__ __forceinline__ double finish_var_c(const SysArgs& a, int v, int g, int c, double acc, double V,
                                               const double* u, const double* dv) {
  if (c == 0) return acc / V + (JV ? source_tangent<NV>(a, v, u, dv) : source_eval<NV>(a, v, g, u));
  if (c == 2) {
    const int f = a.fidx[(int64_t)v * a.N + g];
    if (JV) return cond_tangent<NV>(a, f, dv);
    return cond_eval<NV>(a, f, a.x[g], a.y[g], a.t, u, a.node_g ? a.node_g[(int64_t)v * a.N + g] : 0.0);
  }
  return 0.0;
}

template <int NV, int JV>
__device__ __forceinline__ double finish_var(const SysArgs& a, int v, int g, double acc, double V,
                                             const double* u, const double* dv) {
  return finish_var_c<NV, JV>(a, v, g, a.code[(int64_t)v * a.N + g], acc, V, u, dv);
}

// Boundary half-edge fluxes of every variable at endpoint ep of boundary edge e
// (get_boundary_fluxes for a system, boundary_edge_contributions.jl:51-58), or (JV)
// their derivatives along v: the fluxes are linear in u and the Neumann functions
// affine, so the derivative is the flux of v, or cu . v at the half-edge point.
template <int NV, int JV>
__device__ void bedge_fluxes(const SysArgs& a, int e, int ep, double* Q) {
  const double* src = JV ? a.v : a.u;
  const int i = a.bedges[e], j = a.bedges[a.B + e];
  const double px = a.x[i], py = a.y[i], qx = a.x[j], qy = a.y[j];
  const double lij = sqrt((qx - px) * (qx - px) + (qy - py) * (qy - py));
  const double nx = (qy - py) / lij, ny = -(qx - px) / lij;
  const double mijx = (px + qx) / 2, mijy = (py + qy) / 2;
  const double mx = ep ? (qx + mijx) / 2 : (px + mijx) / 2;
  const double my = ep ? (qy + mijy) / 2 : (py + mijy) / 2;
  const double ell = sqrt((mijx - px) * (mijx - px) + (mijy - py) * (mijy - py));
  const int64_t T = a.badj[e];
  const int va = a.tri[T], vb = a.tri[a.T + T], vc = a.tri[2 * a.T + T];
  double s[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) s[f] = a.props[f * a.T + T];
  double al[NV], be[NV], um[NV];
#pragma unroll
  for (int w = 0; w < NV; ++w) {
    const double ua = src[w + NV * (int64_t)va], ub = src[w + NV * (int64_t)vb], uc = src[w + NV * (int64_t)vc];
    al[w] = s[0] * ua + s[1] * ub + s[2] * uc;
    be[w] = s[3] * ua + s[4] * ub + s[5] * uc;
    const double ga = s[6] * ua + s[7] * ub + s[8] * uc;
    um[w] = al[w] * mx + be[w] * my + ga;
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int nf = a.bneu[(int64_t)v * a.B + e];
    double qn;
    if (nf < 0) {
      double qxv = 0.0, qyv = 0.0;
#pragma unroll
      for (int w = 0; w < NV; ++w)
        if (a.K[v][w] != 0.0) {
          qxv = qxv + a.K[v][w] * al[w];
          qyv = qyv + a.K[v][w] * be[w];
        }
      qn = (-qxv) * nx + (-qyv) * ny;
    } else if (JV) {
      qn = cond_tangent<NV>(a, nf, um);
    } else {
      qn = cond_eval<NV>(a, nf, mx, my, a.t, um, a.bedge_g ? a.bedge_g[(int64_t)v * 2 * a.B + 2 * e + ep] : 0.0);
    }
    Q[v] = qn * ell;
  }
}

// JV = 1: the Jacobian-vector product; s_u then holds the direction v (the fluxes
// are linear), and exclusive nodes read their own u from memory for the source.
template <int NV, int JV, int BLOCK, int TPT>
__global__ __launch_bounds__(BLOCK) void sys_tile_kernel(SysArgs a) {
  constexpr int TILE = BLOCK * TPT;
  constexpr int MAXN = 2 * BLOCK < 1024 ? 2 * BLOCK : 1024;
  constexpr int NS = MAXN / BLOCK;
  __shared__ double s_u[NV][MAXN];
  // the node coordinates are dead once the fluxes are formed: the fluxes reuse their
  // space, one variable at a time (the others wait in registers), so a workgroup's LDS
  // does not grow with NV's flux block and more workgroups fit on a CU
  constexpr int XY = 2 * MAXN, QN = 3 * TILE;
  __shared__ double s_xyq[XY > QN ? XY : QN];
  double* const s_x = s_xyq;
  double* const s_y = s_xyq + MAXN;
  double* const s_q = s_xyq;  // [3 TILE] fluxes of the variable of the current round
  __shared__ uint16_t s_inc[3 * TILE];
  __shared__ int s_is[MAXN + 1];
  const int nbw = (a.n_bitems + BLOCK - 1) / BLOCK;
  const int wid = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  if (wid < nbw) {
    const int k = wid * BLOCK + tid;
    if (k < a.n_bitems) {
      const int w = a.fix_bedge[k];
      double Q[NV];
      bedge_fluxes<NV, JV>(a, w >> 1, w & 1, Q);
#pragma unroll
      for (int v = 0; v < NV; ++v) a.bflux[(int64_t)k * NV + v] = Q[v];
    }
    return;
  }
  const int tile = wid - nbw;
  if (tile >= a.n_tiles) return;
  const int t0 = a.tile_tri_start[tile], nt = a.tile_tri_start[tile + 1] - t0;
  const int n0 = a.tile_node_start[tile], nn = a.tile_node_start[tile + 1] - n0;
  const int ib0 = 3 * t0;
  const double* src = JV ? a.v : a.u;
  // every global load of the tile before the first LDS store, branch-free (lanes
  // past the tile's nodes, entries or triangles read the last valid one and never
  // use it), at 32-bit offsets from bases in scalar registers
  int g[NS], pi[NS], is[NS];
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const uint32_t li = (uint32_t)n0 + umin((uint32_t)(tid + r * BLOCK), (uint32_t)nn - 1u);
    g[r] = ld32(a.tile_nodes, li);
    pi[r] = ld32(a.lpart, li);
    is[r] = ld32(a.inc_start, li);
  }
  uint16_t iv[3 * TPT];
#pragma unroll
  for (int k = 0; k < 3 * TPT; ++k)
    iv[k] = ld32(a.inc, (uint32_t)ib0 + umin((uint32_t)(tid + k * BLOCK), 3u * (uint32_t)nt - 1u));
  uint32_t L[TPT];
#pragma unroll
  for (int r = 0; r < TPT; ++r) L[r] = ld32(a.loc, (uint32_t)t0 + umin((uint32_t)(tid + r * BLOCK), (uint32_t)nt - 1u));
  double xg[NS], yg[NS], ug[NS][NV];
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const uint32_t gi = (uint32_t)g[r];
    xg[r] = ld32(a.x, gi);
    yg[r] = ld32(a.y, gi);
#pragma unroll
    for (int v = 0; v < NV; ++v) ug[r][v] = ld32(src, (uint32_t)v + (uint32_t)NV * gi);
  }
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    if (l < nn) {
      s_is[l] = is[r] - ib0;
      s_x[l] = xg[r];
      s_y[l] = yg[r];
#pragma unroll
      for (int v = 0; v < NV; ++v) s_u[v][l] = ug[r][v];
    }
  }
  if (tid == 0) s_is[nn] = 3 * nt;
#pragma unroll
  for (int k = 0; k < 3 * TPT; ++k) {
    const int idx = tid + k * BLOCK;
    if (idx < 3 * nt) s_inc[idx] = iv[k];
  }
  __syncthreads();
  double Qr[TPT][3][NV];
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
    if (lt < nt) {
      const int ia = L[r] & 1023u, ib = (L[r] >> 10) & 1023u, ic = (L[r] >> 20) & 1023u;
      double P[P_COUNT];
      rebuild_record<(FVM_GEO_LEVEL >= 3 ? FVM_GEO_LEVEL : 3)>(s_x[ia], s_y[ia], s_x[ib], s_y[ib], s_x[ic], s_y[ic], P);
      double al[NV], be[NV];
#pragma unroll
      for (int w = 0; w < NV; ++w) {
        al[w] = P[0] * s_u[w][ia] + P[1] * s_u[w][ib] + P[2] * s_u[w][ic];
        be[w] = P[3] * s_u[w][ia] + P[4] * s_u[w][ib] + P[5] * s_u[w][ic];
      }
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const double nx = P[P_NX + 2 * e], ny = P[P_NX + 1 + 2 * e], l = P[P_L + e];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          double qxv = 0.0, qyv = 0.0;
#pragma unroll
          for (int w = 0; w < NV; ++w)
            if (a.K[v][w] != 0.0) {
              qxv = qxv + a.K[v][w] * al[w];
              qyv = qyv + a.K[v][w] * be[w];
            }
          // (qx nx + qy ny) l with q = -K (alpha, beta)
          Qr[r][e][v] = ((-qxv) * nx + (-qyv) * ny) * l;
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int v = 0; v < NV; ++v) Qr[r][e][v] = 0.0;
    }
  }
  double acc[NS][NV];
#pragma unroll
  for (int r = 0; r < NS; ++r)
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[r][v] = 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    // v = 0: every thread has read the coordinates that s_q overwrites;
    // v > 0: every node has summed the previous variable's fluxes
    __syncthreads();
#pragma unroll
    for (int r = 0; r < TPT; ++r) {
      const int lt = tid + r * BLOCK;
      if (lt < nt) {
#pragma unroll
        for (int e = 0; e < 3; ++e) s_q[3 * lt + e] = Qr[r][e][v];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NS; ++r) {
      const int l = tid + r * BLOCK;
      if (l < nn) {
        const int k1 = s_is[l + 1];
        for (int k = s_is[l]; k < k1; ++k) {
          const int w = s_inc[k];
          const int lt = w >> 2, sl = w & 3;
          acc[r][v] = acc[r][v] + s_q[3 * lt + (sl == 0 ? 2 : sl - 1)] - s_q[3 * lt + sl];
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    if (l < nn) {
      if (pi[r] >= 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) a.partial[(int64_t)pi[r] * NV + v] = acc[r][v];
      } else {
        double u[NV], dv[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          dv[v] = ug[r][v];
          u[v] = JV ? a.u[v + NV * (int64_t)g[r]] : dv[v];
        }
        const double V = a.lcv[n0 + l];
#pragma unroll
        for (int v = 0; v < NV; ++v)
          a.out[v + NV * (int64_t)g[r]] = finish_var<NV, JV>(a, v, g[r], acc[r][v], V, u, dv);
      }
    }
  }
}

template <int NV, int JV>
__global__ __launch_bounds__(256) void sys_fix_kernel(SysArgs a) {
  const int fi = fix_wid() * blockDim.x + threadIdx.x;
  if (fi >= a.n_fix) return;
  // the node's static data in fix order: one coalesced record, gathers of u only
  const int n = a.n_fix;
  const int g = a.fix_meta[fi], p0 = a.fix_meta[n + fi], b0 = a.fix_meta[2 * n + fi];
  const uint32_t m = (uint32_t)a.fix_meta[3 * n + fi];
  const int p1 = p0 + (int)(m >> 20), b1 = b0 + (int)((m >> 8) & 4095u);
  double acc[NV], u[NV], dv[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    acc[v] = 0.0;
    u[v] = a.u[v + NV * (int64_t)g];
    dv[v] = JV ? a.v[v + NV * (int64_t)g] : 0.0;
  }
  for (int p = p0; p < p1; ++p)
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = acc[v] + a.partial[(int64_t)p * NV + v];
  for (int k = b0; k < b1; ++k)
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = acc[v] - a.bflux[(int64_t)k * NV + v];
  const double V = a.fix_cv[fi];
#pragma unroll
  for (int v = 0; v < NV; ++v)
    a.out[v + NV * (int64_t)g] = finish_var_c<NV, JV>(a, v, g, (int)((m >> (2 * v)) & 3u), acc[v], V, u, dv);
}

// out[k] = x[idx[k]]
__global__ __launch_bounds__(256) void sys_gather_kernel(int64_t n, const int32_t* __restrict__ idx,
                                                         const double* __restrict__ x, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = x[idx[k]];
}

// dirichlet.jl:46-53: variable by variable; each thread owns one node, so updating
// its variables in order reproduces the serial loop.  mask[k] has bit v set when
// variable v is Dirichlet at nodes[k] (vertex or not, as the reference's loop over
// the Dirichlet nodes).
template <int NV>
__global__ __launch_bounds__(256) void sys_dirichlet_kernel(SysArgs a, const int32_t* __restrict__ nodes,
                                                            const int32_t* __restrict__ mask, int n,
                                                            double* __restrict__ u) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = nodes[k];
  const int bits = mask[k];
  double uv[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) uv[v] = u[v + NV * (int64_t)g];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (!((bits >> v) & 1)) continue;
    uv[v] = cond_eval<NV>(a, a.fidx[(int64_t)v * a.N + g], a.x[g], a.y[g], a.t, uv,
                          a.node_g ? a.node_g[(int64_t)v * a.N + g] : 0.0);
    u[v + NV * (int64_t)g] = uv[v];
  }
}

// compute_flux for a system (problem.jl:458-487, the is_system branch): every
// variable's q . n across the mesh edge (i, j); out[q * nv + v].
template <int NV>
__global__ __launch_bounds__(256) void sys_flux_query_kernel(SysArgs a, const int32_t* __restrict__ node_tri_start,
                                                             const int32_t* __restrict__ node_tri, int64_t n,
                                                             const int32_t* __restrict__ qi,
                                                             const int32_t* __restrict__ qj,
                                                             double* __restrict__ out, int32_t* __restrict__ bad) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int32_t i = qi[q], j = qj[q];
  int32_t t = -1;
  if (i >= 0 && i < a.N && j >= 0 && j < a.N) {
    t = find_directed_edge(node_tri_start, node_tri, a.tri, a.T, j, i);
    if (t < 0) t = find_directed_edge(node_tri_start, node_tri, a.tri, a.T, i, j);
  }
  if (t < 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) out[q * NV + v] = NAN;
    bad[0] = 1;
    return;
  }
  const double px = a.x[i], py = a.y[i], qx = a.x[j], qy = a.y[j];
  const double ex = qx - px, ey = qy - py;
  const double l = sqrt(ex * ex + ey * ey);
  const double nx = ey / l, ny = -ex / l;
  const int32_t va = a.tri[t], vb = a.tri[a.T + t], vc = a.tri[2 * a.T + t];
  double s[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) s[f] = a.props[f * a.T + t];
  double al[NV], be[NV];
#pragma unroll
  for (int w = 0; w < NV; ++w) {
    const double ua = a.u[w + NV * (int64_t)va], ub = a.u[w + NV * (int64_t)vb], uc = a.u[w + NV * (int64_t)vc];
    al[w] = s[0] * ua + s[1] * ub + s[2] * uc;
    be[w] = s[3] * ua + s[4] * ub + s[5] * uc;
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double qxv = 0.0, qyv = 0.0;
#pragma unroll
    for (int w = 0; w < NV; ++w)
      if (a.K[v][w] != 0.0) {
        qxv = qxv + a.K[v][w] * al[w];
        qyv = qyv + a.K[v][w] * be[w];
      }
    out[q * NV + v] = nx * (-qxv) + ny * (-qyv);
  }
}

template <class T>
int dcopy(T** dst, const T* src, int64_t n) {
  *dst = nullptr;
  FVM_HIP(hipMalloc((void**)dst, sizeof(T) * (size_t)(n > 0 ? n : 1)));
  if (n > 0 && src) FVM_HIP(hipMemcpy(*dst, src, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
  return FVM_OK;
}

}  // namespace

struct fvm_system {
  const fvm_mesh* m = nullptr;
  int32_t nv = 0;
  SysArgs a{};
  std::vector<void*> owned;
  int32_t* dir_nodes = nullptr;
  int32_t* dir_mask = nullptr;  // bit v: variable v is Dirichlet at the node
  int32_t n_dir = 0;
};

namespace {

void free_system(fvm_system* s) {
  if (!s) return;
  for (void* p : s->owned)
    if (p) (void)hipFree(p);
  delete s;
}

template <int NV, int JV>
int launch_sys(const fvm_system* s, const SysArgs& a, hipStream_t st) {
  const int tt = s->m->tile_tris;
  const int nb256 = (a.n_bitems + 255) / 256, nb512 = (a.n_bitems + 511) / 512;
  if (tt <= 256)
    hipLaunchKernelGGL((sys_tile_kernel<NV, JV, 256, 1>), dim3((unsigned)xcd_grid(a.n_tiles + nb256)), dim3(256), 0,
                       st, a);
  else if (tt <= 512)
    hipLaunchKernelGGL((sys_tile_kernel<NV, JV, 256, 2>), dim3((unsigned)xcd_grid(a.n_tiles + nb256)), dim3(256), 0,
                       st, a);
  else
    hipLaunchKernelGGL((sys_tile_kernel<NV, JV, 512, 2>), dim3((unsigned)xcd_grid(a.n_tiles + nb512)), dim3(512), 0,
                       st, a);
  if (a.n_fix > 0)
    hipLaunchKernelGGL((sys_fix_kernel<NV, JV>), dim3((unsigned)fix_grid(a.n_fix)), dim3(256), 0, st, a);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

}  // namespace

extern "C" {

int fvm_system_create(const fvm_mesh* m, const fvm_system_desc* d, fvm_system** out) {
  FVM_CHECK(m && d && out, "fvm_system_create: null argument");
  *out = nullptr;
  const int nv = d->n_vars;
  FVM_CHECK(nv >= 1 && nv <= kMaxVars, "fvm_system_create: n_vars must lie in [1, 4]");
  FVM_CHECK(8 * (int64_t)nv * m->N < ((int64_t)1 << 32),
            "fvm_system_create: the state (n_vars x n_points doubles) must stay below 4 GiB");
  FVM_CHECK(d->flux_k, "fvm_system_create: flux_k required");
  FVM_CHECK(d->n_terms >= 0 && d->n_terms <= kMaxTerms, "fvm_system_create: at most 24 source terms");
  FVM_CHECK(d->n_terms == 0 || (d->term_var && d->term_exp && d->term_coef),
            "fvm_system_create: source terms incomplete");
  FVM_CHECK(d->node_kind && d->node_fidx && (m->B == 0 || d->bedge_neumann_fidx),
            "fvm_system_create: conditions required");
  FVM_CHECK(d->n_functions >= 0 && (d->n_functions == 0 || (d->fn_c0 && d->fn_cu)),
            "fvm_system_create: condition coefficients missing");
  FVM_CHECK(m->tile_tris <= 1024, "fvm_system_create: tiles of at most 1024 triangles");
  const int64_t N = m->N, T = m->T, B = m->B;
  const int nf = d->n_functions;
  fvm_system* s = new fvm_system();
  s->m = m;
  s->nv = nv;
  SysArgs& a = s->a;
  a.nv = nv;
  a.n_tiles = m->n_tiles;
  a.n_fix = m->n_fix;
  a.n_bitems = m->n_bedge_items;
  a.N = N;
  a.T = T;
  a.B = B;
  a.tile_tri_start = m->tile_tri_start;
  a.tile_node_start = m->tile_node_start;
  a.tile_nodes = m->tile_nodes;
  a.loc = m->loc;
  a.inc_start = m->inc_start;
  a.inc = m->inc;
  a.fix_nodes = m->fix_nodes;
  a.fix_part_start = m->fix_part_start;
  a.fix_bedge_start = m->fix_bedge_start;
  a.fix_bedge = m->fix_bedge;
  a.props = m->props;
  a.tri = m->tri;
  a.x = m->x;
  a.y = m->y;
  a.cv = m->cv;
  a.bedges = m->bedges;
  a.badj = m->badj;
  a.lcv = m->lcv;
  for (int v = 0; v < kMaxVars; ++v) {
    a.st[v] = (v < nv && d->st) ? d->st[v] : 0.0;
    for (int w = 0; w < kMaxVars; ++w) a.K[v][w] = (v < nv && w < nv) ? d->flux_k[v * nv + w] : 0.0;
  }
  a.S.n = d->n_terms;
  for (int k = 0; k < d->n_terms; ++k) {
    FVM_CHECK(d->term_var[k] >= 0 && d->term_var[k] < nv, "fvm_system_create: term_var out of range");
    a.S.var[k] = d->term_var[k];
    a.S.c[k] = d->term_coef[k];
    for (int w = 0; w < kMaxVars; ++w) {
      const int e = w < nv ? d->term_exp[k * nv + w] : 0;
      FVM_CHECK(e >= 0 && e <= 3, "fvm_system_create: source exponents must lie in [0, 3]");
      a.S.e[k][w] = (int8_t)e;
    }
  }
  // node codes per variable (3 marks points that are not vertices), Dirichlet node list
  std::vector<int8_t> code((size_t)nv * N);
  std::vector<int32_t> dir, dir_mask;
  for (int64_t g = 0; g < N; ++g) {
    const bool vert = m->h_node_tri_start[g + 1] > m->h_node_tri_start[g];
    int32_t bits = 0;
    for (int v = 0; v < nv; ++v) {
      const int k = d->node_kind[(int64_t)v * N + g];
      if (k < 0 || k > 2) {
        free_system(s);
        set_error("fvm_system_create: node_kind entries must be 0, 1 or 2");
        return FVM_ERR_ARG;
      }
      if (k != 0) {
        const int f = d->node_fidx[(int64_t)v * N + g];
        if (f < 0 || f >= nf) {
          free_system(s);
          set_error("fvm_system_create: node_fidx out of range");
          return FVM_ERR_ARG;
        }
      }
      code[(size_t)v * N + g] = vert ? (int8_t)k : (int8_t)3;
      if (k == 1) bits |= 1 << v;
    }
    if (bits) {
      dir.push_back((int32_t)g);
      dir_mask.push_back(bits);
    }
  }
  for (int64_t e = 0; e < (int64_t)nv * B; ++e) {
    const int f = d->bedge_neumann_fidx[e];
    if (f != -1 && (f < 0 || f >= nf)) {
      free_system(s);
      set_error("fvm_system_create: bedge_neumann_fidx out of range");
      return FVM_ERR_ARG;
    }
  }
  std::vector<double> zeros(std::max<int64_t>(nf, 1), 0.0);
  std::vector<int32_t> lp(m->n_local);
  for (int64_t e = 0; e < m->n_local; ++e) lp[e] = m->h_part_idx[e] >= 0 ? m->h_part_idx[e] : -1;
  int rc = FVM_OK;
  auto up = [&](auto** dst, const auto* src, int64_t n) {
    if (rc) return;
    rc = dcopy(dst, src, n);
    if (!rc) s->owned.push_back((void*)*dst);
  };
  int8_t* d_code = nullptr;
  int32_t *d_fidx = nullptr, *d_bneu = nullptr, *d_lpart = nullptr;
  double *d_c0 = nullptr, *d_cu = nullptr, *d_ct = nullptr, *d_cx = nullptr, *d_cy = nullptr, *d_lam = nullptr;
  double *d_snode = nullptr, *d_ng = nullptr, *d_bg = nullptr, *d_part = nullptr, *d_bflux = nullptr;
  up(&d_code, code.data(), (int64_t)nv * N);
  up(&d_fidx, d->node_fidx, (int64_t)nv * N);
  up(&d_bneu, d->bedge_neumann_fidx, (int64_t)nv * B);
  up(&d_lpart, lp.data(), m->n_local);
  up(&d_c0, nf ? d->fn_c0 : zeros.data(), std::max(nf, 1));
  up(&d_cu, nf ? d->fn_cu : zeros.data(), nf ? (int64_t)nf * nv : 1);
  up(&d_ct, nf && d->fn_ct ? d->fn_ct : zeros.data(), std::max(nf, 1));
  up(&d_cx, nf && d->fn_cx ? d->fn_cx : zeros.data(), std::max(nf, 1));
  up(&d_cy, nf && d->fn_cy ? d->fn_cy : zeros.data(), std::max(nf, 1));
  up(&d_lam, nf && d->fn_lam ? d->fn_lam : zeros.data(), std::max(nf, 1));
  if (d->s_node) up(&d_snode, d->s_node, (int64_t)nv * N);
  if (d->node_g) up(&d_ng, d->node_g, (int64_t)nv * N);
  if (d->bedge_g) up(&d_bg, d->bedge_g, (int64_t)nv * 2 * B);
  up(&d_part, (const double*)nullptr, (int64_t)std::max(m->n_parts, 1) * nv);
  up(&d_bflux, (const double*)nullptr, (int64_t)std::max(m->n_bedge_items, 1) * nv);
  up(&s->dir_nodes, dir.data(), (int64_t)dir.size());
  // per fix node: node, first slot, first boundary item, slots << 20 | items << 8 | codes
  int32_t* d_meta = nullptr;
  double* d_fcv = nullptr;
  {
    const int64_t nf = m->n_fix;
    std::vector<int32_t> hfix(std::max<int64_t>(nf, 1), 0), meta(4 * (size_t)std::max<int64_t>(nf, 1), 0);
    if (!rc && nf > 0 && hipMemcpy(hfix.data(), m->fix_nodes, sizeof(int32_t) * nf, hipMemcpyDeviceToHost) != hipSuccess) {
      set_error("fvm_system_create: reading the fix-up nodes");
      rc = FVM_ERR_HIP;
    }
    for (int64_t f = 0; !rc && f < nf; ++f) {
      const int32_t np = m->h_fix_part_start[f + 1] - m->h_fix_part_start[f];
      const int32_t nb = m->h_fix_bedge_start[f + 1] - m->h_fix_bedge_start[f];
      if (np >= 4096 || nb >= 4096) {
        set_error("fvm_system_create: a node shared by 4096 or more tiles");
        rc = FVM_ERR_ARG;
        break;
      }
      uint32_t codes = 0;
      for (int v = 0; v < nv; ++v) codes |= (uint32_t)(uint8_t)code[(size_t)v * N + hfix[f]] << (2 * v);
      meta[f] = hfix[f];
      meta[nf + f] = m->h_fix_part_start[f];
      meta[2 * nf + f] = m->h_fix_bedge_start[f];
      meta[3 * nf + f] = (int32_t)(((uint32_t)np << 20) | ((uint32_t)nb << 8) | codes);
    }
    up(&d_meta, meta.data(), (int64_t)meta.size());
    up(&d_fcv, (const double*)nullptr, std::max<int64_t>(nf, 1));
    if (!rc && nf > 0) {
      hipLaunchKernelGGL(sys_gather_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, 0, nf, m->fix_nodes,
                         (const double*)m->cv, d_fcv);
      if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        set_error("fvm_system_create: gathering the fix-up volumes");
        rc = FVM_ERR_HIP;
      }
    }
  }
  up(&s->dir_mask, dir_mask.data(), (int64_t)dir_mask.size());
  if (rc) {
    free_system(s);
    return rc;
  }
  s->n_dir = (int32_t)dir.size();
  a.lpart = d_lpart;
  a.code = d_code;
  a.fidx = d_fidx;
  a.bneu = d_bneu;
  a.s_node = d_snode;
  a.n_fn = nf;
  a.fc0 = d_c0;
  a.fcu = d_cu;
  a.fct = d_ct;
  a.fcx = d_cx;
  a.fcy = d_cy;
  a.flam = d_lam;
  a.node_g = d_ng;
  a.bedge_g = d_bg;
  a.partial = d_part;
  a.bflux = d_bflux;
  a.fix_meta = d_meta;
  a.fix_cv = d_fcv;
  *out = s;
  return FVM_OK;
}

int fvm_system_destroy(fvm_system* s) {
  free_system(s);
  return FVM_OK;
}

int fvm_system_rhs(fvm_system* s, const double* d_u, double* d_du, double t, fvm_stream st) {
  FVM_CHECK(s && d_u && d_du, "fvm_system_rhs: null argument");
  FVM_CHECK(d_u != d_du, "fvm_system_rhs: du must not alias u");
  SysArgs a = s->a;
  a.u = d_u;
  a.out = d_du;
  a.t = t;
  hipStream_t hs = (hipStream_t)st;
  switch (s->nv) {
    case 1: return launch_sys<1, 0>(s, a, hs);
    case 2: return launch_sys<2, 0>(s, a, hs);
    case 3: return launch_sys<3, 0>(s, a, hs);
    default: return launch_sys<4, 0>(s, a, hs);
  }
}

int fvm_system_jvp(fvm_system* s, const double* d_u, const double* d_v, double* d_jv, double t,
                   fvm_stream st) {
  FVM_CHECK(s && d_u && d_v && d_jv, "fvm_system_jvp: null argument");
  FVM_CHECK(d_jv != d_u && d_jv != d_v, "fvm_system_jvp: jv must not alias u or v");
  SysArgs a = s->a;
  a.u = d_u;
  a.v = d_v;
  a.out = d_jv;
  a.t = t;
  hipStream_t hs = (hipStream_t)st;
  switch (s->nv) {
    case 1: return launch_sys<1, 1>(s, a, hs);
    case 2: return launch_sys<2, 1>(s, a, hs);
    case 3: return launch_sys<3, 1>(s, a, hs);
    default: return launch_sys<4, 1>(s, a, hs);
  }
}

int fvm_system_compute_flux(fvm_system* s, int64_t n, const int32_t* d_i, const int32_t* d_j,
                            const double* d_u, double t, double* d_out, fvm_stream st) {
  FVM_CHECK(s && n >= 0, "fvm_system_compute_flux: bad argument");
  if (n == 0) return FVM_OK;
  FVM_CHECK(d_i && d_j && d_u && d_out, "fvm_system_compute_flux: null argument");
  SysArgs a = s->a;
  a.u = d_u;
  a.t = t;
  const fvm_mesh* m = s->m;
  hipStream_t hs = (hipStream_t)st;
  const dim3 g((unsigned)((n + 255) / 256));
  return fvm::run_flagged(hs, "fvm_system_compute_flux", m->qflag, [&](int32_t* bad) {
    switch (s->nv) {
      case 1: hipLaunchKernelGGL((sys_flux_query_kernel<1>), g, dim3(256), 0, hs, a, m->node_tri_start, m->node_tri, n, d_i, d_j, d_out, bad); break;
      case 2: hipLaunchKernelGGL((sys_flux_query_kernel<2>), g, dim3(256), 0, hs, a, m->node_tri_start, m->node_tri, n, d_i, d_j, d_out, bad); break;
      case 3: hipLaunchKernelGGL((sys_flux_query_kernel<3>), g, dim3(256), 0, hs, a, m->node_tri_start, m->node_tri, n, d_i, d_j, d_out, bad); break;
      default: hipLaunchKernelGGL((sys_flux_query_kernel<4>), g, dim3(256), 0, hs, a, m->node_tri_start, m->node_tri, n, d_i, d_j, d_out, bad);
    }
  });
}

int fvm_system_dirichlet_apply(fvm_system* s, double* d_u, double t, fvm_stream st) {
  FVM_CHECK(s && d_u, "fvm_system_dirichlet_apply: null argument");
  if (s->n_dir == 0) return FVM_OK;
  SysArgs a = s->a;
  a.t = t;
  const dim3 g((unsigned)((s->n_dir + 255) / 256));
  hipStream_t hs = (hipStream_t)st;
  switch (s->nv) {
    case 1: hipLaunchKernelGGL((sys_dirichlet_kernel<1>), g, dim3(256), 0, hs, a, s->dir_nodes, s->dir_mask, s->n_dir, d_u); break;
    case 2: hipLaunchKernelGGL((sys_dirichlet_kernel<2>), g, dim3(256), 0, hs, a, s->dir_nodes, s->dir_mask, s->n_dir, d_u); break;
    case 3: hipLaunchKernelGGL((sys_dirichlet_kernel<3>), g, dim3(256), 0, hs, a, s->dir_nodes, s->dir_mask, s->n_dir, d_u); break;
    default: hipLaunchKernelGGL((sys_dirichlet_kernel<4>), g, dim3(256), 0, hs, a, s->dir_nodes, s->dir_mask, s->n_dir, d_u);
  }
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

}  // extern "C"
