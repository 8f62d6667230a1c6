// fvm_eqs! on the device: RHS, Jacobian-vector and vector-Jacobian products.
//
// Reference: src/equations/*.jl (main_equations.jl:28-81 is the entry point).
// Layout: triangles are stored in tiles (fvm_host.h).  One workgroup per tile:
//   1. issue the loads of its triangles' geometry (SoA, coalesced) into registers
//      and gather u (and the nodal field) of its local nodes into LDS;
//   2. compute the three CV-edge fluxes of every triangle into LDS;
//   3. every local node sums its incident triangles' contributions in a fixed
//      order; a node owned by this tile alone is finished here (divide by the
//      control volume, add the source, apply the node's condition) and written
//      once; a node shared with other tiles or the boundary writes a partial.
// A second, small kernel finishes the shared nodes: partials in tile order, then
// the boundary-edge fluxes (boundary_edge_contributions.jl), then the node rule
// (source_contributions.jl).  Every output is written by exactly one thread,
// with no atomics, so results are bitwise reproducible run to run.
//
// Arithmetic follows the reference's operation order with FMA contraction off
// (-ffp-contract=off), so each triangle's fluxes equal the CPU restatement's;
// only the order of the per-node sums differs.
#include <algorithm>
#include <vector>

#include "fvm_internal.hpp"

using namespace fvm;

namespace {

// RHS, Jacobian-vector product, or one forward-Euler step of the ensemble path
enum { M_RHS = 0, M_JVP = 1, M_EULER = 2 };

struct KArgs {
  int32_t n_tiles, n_fix;
  int32_t n_bitems;  // boundary work items computed by the tile kernel's leading workgroups
  int64_t T, B;
  const int32_t* tile_tri_start;
  const int32_t* tile_node_start;
  const int32_t* tile_nodes;
  const uint32_t* loc;
  const int32_t* inc_start;
  const uint16_t* inc;
  const int32_t* part_idx;
  const int32_t* fix_nodes;
  const int32_t* fix_part_start;
  const int32_t* fix_bedge_start;
  const int32_t* fix_meta;  // [4][n_fix]: node, first slot, first boundary item, counts and code
  const double* fix_cv;     // [n_fix]
  const int32_t* fix_bedge;
  const int32_t* fix_bvert_start;
  const int32_t* fix_bvert;
  const double* props;
  const int32_t* tri;
  const double* x;
  const double* y;
  const double* cv;
  const int32_t* bedges;
  const int32_t* badj;
  const int8_t* code;  // 0 none, 1 Dirichlet, 2 Dudt, 3 not a vertex
  const int32_t* fidx;
  const int32_t* bneu;
  double d0, d1, d2, d3;
  const double* d_edge;   // [3][T] sorted
  const double* d_bedge;  // [2*B] interleaved
  const double* log_d;
  double s0, s1, s2, s3, st;
  const double* s_node;
  CondTable C;
  const double* u;
  const double* v;  // JVP direction, or VJP cotangent
  double* out;
  double* out2;
  double* partial;
  double* partial2;
  const int32_t* lpart;  // per local entry: partial slot, or -1 - code (exclusive node)
  const double* lcv;     // per local entry: control volume
  double* bflux;         // boundary terms per work item
  double t;
  const double* tile_dt;  // M_EULER: time step of each tile's nodes
  const double* dtv;      // M_EULER: time step per node (fix-up nodes)
};

// How much of the triangle record the tile kernels rebuild from the vertex
// coordinates instead of streaming it (the rebuilt values equal the stored ones
// bit for bit: same operations as the geometry kernel):
//   1: s7..s9 and the CV-edge midpoints (only the forms that need them;
//      FVM_FLUX_D_EDGE reads neither);
//   2: also s1..s6;
//   3: the whole record (also the normals and lengths).
constexpr int kGeo = FVM_GEO_LEVEL;
__host__ __device__ constexpr bool loads_field(int fk, int f) {
  return kGeo >= 3 ? false : (kGeo == 2 ? f >= P_NX : (f < 6 || f >= P_NX));
}
__host__ __device__ constexpr bool rebuilds(int fk) { return kGeo >= 2 || fk != FVM_FLUX_D_EDGE; }

__device__ __forceinline__ double dpoly(const KArgs& a, double u) {
  return a.d0 + u * (a.d1 + u * (a.d2 + u * a.d3));
}
__device__ __forceinline__ double dpoly_du(const KArgs& a, double u) {
  return a.d1 + u * (2.0 * a.d2 + 3.0 * a.d3 * u);
}

// Forms whose D depends on u_s = alpha x_s + beta y_s + gamma: the cubic, or the
// power law D = d0 + d1 |u|^d2 (porous-medium diffusivity, d2 > 0).
__host__ __device__ constexpr bool d_of_u(int fk) { return fk == FVM_FLUX_D_POLY_U || fk == FVM_FLUX_D_POWER; }
template <int FK>
__device__ __forceinline__ double dfun(const KArgs& a, double u) {
  if (FK == FVM_FLUX_D_POWER) return a.d0 + a.d1 * pow(fabs(u), a.d2);
  return dpoly(a, u);
}
// dD/du; for the power law d1 d2 |u|^(d2 - 1) sign(u), taken as 0 at u = 0
template <int FK>
__device__ __forceinline__ double dfun_du(const KArgs& a, double u) {
  if (FK == FVM_FLUX_D_POWER) {
    const double au = fabs(u);
    if (!(au > 0.0)) return 0.0;
    const double d = a.d1 * a.d2 * pow(au, a.d2 - 1.0);
    return u > 0.0 ? d : -d;
  }
  return dpoly_du(a, u);
}

// S(x, y, t, u) of the source term (source_contributions.jl:2-14)
__device__ __forceinline__ double source(const KArgs& a, int g, double u) {
  double S = a.s0 + u * (a.s1 + u * (a.s2 + u * a.s3));
  if (a.s_node) S = a.s_node[g] + S;
  if (a.st != 0.0) S = S + a.st * a.t;
  return S;
}
__device__ __forceinline__ double source_du(const KArgs& a, double u) {
  return a.s1 + u * (2.0 * a.s2 + 3.0 * a.s3 * u);
}

// The three CV-edge fluxes Q_e = (q . n) l of one triangle (get_fluxes,
// triangle_contributions.jl:2-7), or their directional derivatives.
template <int FK, int MODE>
__device__ __forceinline__ void tri_fluxes(const KArgs& a, const double* P, double ua, double ub,
                                           double uc, double va, double vb, double vc, double Da,
                                           double Db, double Dc, const double* Ds, double* Q) {
  const double al = P[0] * ua + P[1] * ub + P[2] * uc;
  const double be = P[3] * ua + P[4] * ub + P[5] * uc;
  double ga = 0.0, dal = 0.0, dbe = 0.0, dga = 0.0, da = 0.0, db = 0.0, dg = 0.0;
  constexpr bool G = d_of_u(FK) || FK == FVM_FLUX_LINEAR;  // gamma enters q
  if (G) ga = P[6] * ua + P[7] * ub + P[8] * uc;
  if (MODE == M_JVP) {
    dal = P[0] * va + P[1] * vb + P[2] * vc;
    dbe = P[3] * va + P[4] * vb + P[5] * vc;
    if (G) dga = P[6] * va + P[7] * vb + P[8] * vc;
  }
  if (FK == FVM_FLUX_D_NODAL) {
    da = P[0] * Da + P[1] * Db + P[2] * Dc;
    db = P[3] * Da + P[4] * Db + P[5] * Dc;
    dg = P[6] * Da + P[7] * Db + P[8] * Dc;
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    if (FK == FVM_FLUX_LINEAR) {
      // (q . n) l with q = A (alpha, beta, gamma): the coefficients come projected
      const double* c = Ds + 3 * e;
      Q[e] = MODE != M_JVP ? c[0] * al + c[1] * be + c[2] * ga : c[0] * dal + c[1] * dbe + c[2] * dga;
      continue;
    }
    const double nx = P[P_NX + 2 * e], ny = P[P_NX + 1 + 2 * e], l = P[P_L + e];
    double D, Dd = 0.0;
    if (d_of_u(FK)) {
      const double x = P[P_MX + 2 * e], y = P[P_MX + 1 + 2 * e];
      const double us = al * x + be * y + ga;
      D = dfun<FK>(a, us);
      if (MODE == M_JVP) Dd = dfun_du<FK>(a, us) * (dal * x + dbe * y + dga);
    } else if (FK == FVM_FLUX_D_NODAL) {
      const double x = P[P_MX + 2 * e], y = P[P_MX + 1 + 2 * e];
      D = da * x + db * y + dg;
    } else {
      D = Ds[e];
    }
    if (MODE != M_JVP) {
      const double qx = -D * al, qy = -D * be;
      const double qn = qx * nx + qy * ny;
      Q[e] = qn * l;
    } else {
      const double qx = -(Dd * al + D * dal), qy = -(Dd * be + D * dbe);
      Q[e] = (qx * nx + qy * ny) * l;
    }
  }
}

struct BGeo {
  double nx, ny, mx[2], my[2], ell;
};

// get_boundary_cv_components (control_volumes.jl:41-56)
__device__ __forceinline__ BGeo bgeo(const KArgs& a, int e) {
  const int i = a.bedges[e], j = a.bedges[a.B + e];
  const double px = a.x[i], py = a.y[i], qx = a.x[j], qy = a.y[j];
  BGeo G;
  const double lij = sqrt((qx - px) * (qx - px) + (qy - py) * (qy - py));
  G.nx = (qy - py) / lij;
  G.ny = -(qx - px) / lij;
  const double mijx = (px + qx) / 2, mijy = (py + qy) / 2;
  G.mx[0] = (px + mijx) / 2;
  G.my[0] = (py + mijy) / 2;
  G.mx[1] = (qx + mijx) / 2;
  G.my[1] = (qy + mijy) / 2;
  G.ell = sqrt((mijx - px) * (mijx - px) + (mijy - py) * (mijy - py));
  return G;
}

// Flux through the boundary CV edge of endpoint ep of boundary edge e, times its
// length (get_boundary_fluxes, boundary_edge_contributions.jl:2-48), or its
// directional derivative.
template <int FK, int MODE>
__device__ double bedge_flux(const KArgs& a, int e, int ep) {
  const BGeo G = bgeo(a, e);
  const int64_t Tt = a.badj[e];
  const int va = a.tri[Tt], vb = a.tri[a.T + Tt], vc = a.tri[2 * a.T + Tt];
  double s[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) s[f] = a.props[f * a.T + Tt];
  const double ua = a.u[va], ub = a.u[vb], uc = a.u[vc];
  const double al = s[0] * ua + s[1] * ub + s[2] * uc;
  const double be = s[3] * ua + s[4] * ub + s[5] * uc;
  const double ga = s[6] * ua + s[7] * ub + s[8] * uc;
  double dal = 0, dbe = 0, dga = 0;
  if (MODE == M_JVP) {
    const double wa = a.v[va], wb = a.v[vb], wc = a.v[vc];
    dal = s[0] * wa + s[1] * wb + s[2] * wc;
    dbe = s[3] * wa + s[4] * wb + s[5] * wc;
    dga = s[6] * wa + s[7] * wb + s[8] * wc;
  }
  const double mx = ep ? G.mx[1] : G.mx[0], my = ep ? G.my[1] : G.my[0];
  const int nf = a.bneu[e];
  if (FK == FVM_FLUX_LINEAR && nf < 0) {
    const double* c = a.d_bedge + 3 * (2 * e + ep);  // projected onto n, times the length
    return MODE != M_JVP ? c[0] * al + c[1] * be + c[2] * ga : c[0] * dal + c[1] * dbe + c[2] * dga;
  }
  double qn;
  if (nf < 0) {
    double D, Dd = 0.0;
    if (d_of_u(FK)) {
      const double um = al * mx + be * my + ga;
      D = dfun<FK>(a, um);
      if (MODE == M_JVP) Dd = dfun_du<FK>(a, um) * (dal * mx + dbe * my + dga);
    } else if (FK == FVM_FLUX_D_NODAL) {
      const double Da = exp(a.log_d[va]), Db = exp(a.log_d[vb]), Dc = exp(a.log_d[vc]);
      const double da = s[0] * Da + s[1] * Db + s[2] * Dc;
      const double db = s[3] * Da + s[4] * Db + s[5] * Dc;
      const double dg = s[6] * Da + s[7] * Db + s[8] * Dc;
      D = da * mx + db * my + dg;
    } else {
      D = a.d_bedge[2 * e + ep];
    }
    if (MODE != M_JVP) {
      const double qx = -D * al, qy = -D * be;
      qn = qx * G.nx + qy * G.ny;
    } else {
      const double qx = -(Dd * al + D * dal), qy = -(Dd * be + D * dbe);
      qn = qx * G.nx + qy * G.ny;
    }
  } else {
    if (MODE != M_JVP)
      qn = eval_cond(a.C, nf, mx, my, a.t, al * mx + be * my + ga,
                     a.C.bedge_g ? a.C.bedge_g[2 * e + ep] : 0.0);
    else
      qn = a.C.cu[nf] * (dal * mx + dbe * my + dga);
  }
  return qn * G.ell;
}

// ---------------------------------------------------------------------------
// RHS / JVP: tile kernel and fix-up kernel
// ---------------------------------------------------------------------------
// Local-node capacity of a tile processed by BLOCK threads: two node slots per
// thread, at most 1024 (local ids are packed in 10 bits).
__host__ __device__ constexpr int node_cap(int block) { return 2 * block < 1024 ? 2 * block : 1024; }

// fvm_eqs_single_source_contribution! (source_contributions.jl:33-47) for the RHS,
// its derivative for the JVP, or (M_EULER) the new state u + dt du followed by the
// Dirichlet callback (dirichlet.jl:78-86) at t + dt; the node's code and control
// volume are already in registers.
template <int MODE>
__device__ __forceinline__ double finish_pre(const KArgs& a, int g, double acc, double ug, double vg,
                                             int c, double V, double dt) {
  if (c == 0) {
    if (MODE == M_RHS) return acc / V + source(a, g, ug);
    if (MODE == M_EULER) return ug + dt * (acc / V + source(a, g, ug));
    return acc / V + source_du(a, ug) * vg;
  }
  if (c == 2) {
    const int f = a.fidx[g];
    if (MODE == M_RHS)
      return eval_cond(a.C, f, a.x[g], a.y[g], a.t, ug, a.C.node_g ? a.C.node_g[g] : 0.0);
    if (MODE == M_EULER)
      return ug + dt * eval_cond(a.C, f, a.x[g], a.y[g], a.t, ug, a.C.node_g ? a.C.node_g[g] : 0.0);
    return a.C.cu[f] * vg;
  }
  if (MODE == M_EULER) {
    if (c == 1)
      return eval_cond(a.C, a.fidx[g], a.x[g], a.y[g], a.t + dt, ug, a.C.node_g ? a.C.node_g[g] : 0.0);
    return ug;
  }
  return 0.0;
}

// One workgroup per tile.  All global loads of the tile (node ids, incidence
// block, partial slots, the triangle records, then the per-node gathers that
// depend on the node ids) are issued before the first wait, so the workgroup
// exposes one memory latency; the rest works out of LDS and registers.
template <int FK, int MODE, int BLOCK, int TPT>
__global__ __launch_bounds__(BLOCK) void tile_kernel(KArgs a) {
  constexpr int TILE = BLOCK * TPT;
  constexpr int MAXN = node_cap(BLOCK);
  constexpr int NS = MAXN / BLOCK;
  __shared__ double s_u[MAXN];
  __shared__ double s_v[MODE == M_JVP ? MAXN : 1];
  __shared__ double s_D[FK == FVM_FLUX_D_NODAL ? MAXN : 1];
  constexpr bool RC = rebuilds(FK);  // part of the triangle record rebuilt from coordinates
  // the node coordinates are dead once the fluxes are formed: the fluxes reuse their
  // space (fewer LDS bytes per workgroup, more workgroups per CU)
  constexpr int XY = RC ? 2 * MAXN : 0;
  __shared__ double s_xyq[XY > 3 * TILE ? XY : 3 * TILE];
  double* const s_x = s_xyq;
  double* const s_y = s_xyq + MAXN;
  double* const s_q = s_xyq;
  __shared__ uint16_t s_inc[3 * TILE];
  __shared__ int s_is[MAXN + 1];
  const int nbw = (a.n_bitems + BLOCK - 1) / BLOCK;
  const int wid = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  if (wid < nbw) {
    // leading workgroups: the boundary-edge fluxes, one thread per (edge, endpoint)
    const int k = wid * BLOCK + tid;
    if (k < a.n_bitems) {
      const int w = a.fix_bedge[k];
      a.bflux[k] = bedge_flux<FK, MODE>(a, w >> 1, w & 1);
    }
    return;
  }
  const int tile = wid - nbw;
  if (tile >= a.n_tiles) return;
  const int t0 = a.tile_tri_start[tile], nt = a.tile_tri_start[tile + 1] - t0;
  const int n0 = a.tile_node_start[tile], nn = a.tile_node_start[tile + 1] - n0;
  const int ib0 = 3 * t0;  // a tile's incidence block is [3 t0, 3 (t0 + nt))
  const double dt = MODE == M_EULER ? a.tile_dt[tile] : 0.0;

  int g[NS], pi[NS], is[NS];
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    g[r] = 0;
    pi[r] = -1;
    is[r] = 0;
    if (l < nn) {
      g[r] = a.tile_nodes[n0 + l];
      pi[r] = a.lpart[n0 + l];
      is[r] = a.inc_start[n0 + l] - ib0;
    }
  }
  uint16_t iv[3 * TPT];
#pragma unroll
  for (int k = 0; k < 3 * TPT; ++k) {
    const int idx = tid + k * BLOCK;
    iv[k] = idx < 3 * nt ? a.inc[ib0 + idx] : (uint16_t)0;
  }
  double P[TPT][P_COUNT];
  double Ds[TPT][FK == FVM_FLUX_LINEAR ? 9 : 3];
  uint32_t L[TPT];
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
    L[r] = 0;
    if (lt < nt) {
      const int64_t s = t0 + lt;
      L[r] = a.loc[s];
#pragma unroll
      for (int f = 0; f < P_COUNT; ++f)
        if (loads_field(FK, f)) P[r][f] = __builtin_nontemporal_load(a.props + f * a.T + s);
      if (FK == FVM_FLUX_D_EDGE) {
#pragma unroll
        for (int e = 0; e < 3; ++e) Ds[r][e] = __builtin_nontemporal_load(a.d_edge + e * a.T + s);
      }
      if (FK == FVM_FLUX_LINEAR) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Ds[r][k] = __builtin_nontemporal_load(a.d_edge + k * a.T + s);
      }
    }
  }
  double ug[NS], vg[NS], Vg[NS], lg[NS], xg[NS], yg[NS];
  int cg[NS];
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    ug[r] = vg[r] = Vg[r] = lg[r] = xg[r] = yg[r] = 0.0;
    cg[r] = 3;
    if (l < nn) {
      ug[r] = a.u[g[r]];
      if (MODE == M_JVP) vg[r] = a.v[g[r]];
      if (FK == FVM_FLUX_D_NODAL) lg[r] = a.log_d[g[r]];
      if (RC) {
        xg[r] = a.x[g[r]];
        yg[r] = a.y[g[r]];
      }
      if (pi[r] < 0) {
        cg[r] = -1 - pi[r];
        Vg[r] = a.lcv[n0 + l];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    if (l < nn) {
      s_u[l] = ug[r];
      if (MODE == M_JVP) s_v[l] = vg[r];
      if (FK == FVM_FLUX_D_NODAL) s_D[l] = exp(lg[r]);
      if (RC) {
        s_x[l] = xg[r];
        s_y[l] = yg[r];
      }
      s_is[l] = is[r];
    }
  }
  if (tid == 0) s_is[nn] = 3 * nt;
#pragma unroll
  for (int k = 0; k < 3 * TPT; ++k) {
    const int idx = tid + k * BLOCK;
    if (idx < 3 * nt) s_inc[idx] = iv[k];
  }
  __syncthreads();
  double Q[TPT][3];
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
    Q[r][0] = Q[r][1] = Q[r][2] = 0.0;
    if (lt < nt) {
      const int ia = L[r] & 1023u, ib = (L[r] >> 10) & 1023u, ic = (L[r] >> 20) & 1023u;
      if (RC) rebuild_record<kGeo>(s_x[ia], s_y[ia], s_x[ib], s_y[ib], s_x[ic], s_y[ic], P[r]);
      tri_fluxes<FK, MODE>(a, P[r], s_u[ia], s_u[ib], s_u[ic],
                           MODE == M_JVP ? s_v[ia] : 0.0, MODE == M_JVP ? s_v[ib] : 0.0,
                           MODE == M_JVP ? s_v[ic] : 0.0,
                           FK == FVM_FLUX_D_NODAL ? s_D[ia] : 0.0,
                           FK == FVM_FLUX_D_NODAL ? s_D[ib] : 0.0,
                           FK == FVM_FLUX_D_NODAL ? s_D[ic] : 0.0, Ds[r], Q[r]);
    }
  }
  if (RC) __syncthreads();  // every thread has read the coordinates that s_q overwrites
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
    if (lt < nt) {
      s_q[3 * lt] = Q[r][0];
      s_q[3 * lt + 1] = Q[r][1];
      s_q[3 * lt + 2] = Q[r][2];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    if (l < nn) {
      double acc = 0.0;
      const int k1 = s_is[l + 1];
      for (int k = s_is[l]; k < k1; ++k) {
        const int w = s_inc[k];
        const int lt = w >> 2, sl = w & 3;
        // update_du! (triangle_contributions.jl:10-15): i += Q3 - Q1, j += Q1 - Q2, k += Q2 - Q3
        acc = acc + s_q[3 * lt + (sl == 0 ? 2 : sl - 1)] - s_q[3 * lt + sl];
      }
      // the node's u (and v) are re-read from LDS and its volume from memory here rather
      // than held in registers across the barriers: fewer VGPRs, more waves per SIMD
      if (pi[r] >= 0)
        a.partial[pi[r]] = acc;
      else
        a.out[g[r]] = finish_pre<MODE>(a, g[r], acc, s_u[l], MODE == M_JVP ? s_v[l] : 0.0, -1 - pi[r],
                                       a.lcv[n0 + l], dt);
    }
  }
}

// Shared nodes: partial sums in tile order, then the boundary-edge fluxes (computed
// by the tile kernel's leading workgroups), then the node rule.  The first four slots are loaded without a loop-carried
// dependency (nearly every node has at most four).
template <int FK, int MODE>
__global__ __launch_bounds__(256) void fix_kernel(KArgs a) {
  const int fi = blockIdx.x * blockDim.x + threadIdx.x;
  if (fi >= a.n_fix) return;
  // the node's static data in fix order: one coalesced record, one gather (u)
  const int n = a.n_fix;
  const int g = a.fix_meta[fi], p0 = a.fix_meta[n + fi], b0 = a.fix_meta[2 * n + fi];
  const uint32_t w = (uint32_t)a.fix_meta[3 * n + fi];
  const int p1 = p0 + (int)(w >> 20), b1 = b0 + (int)((w >> 8) & 4095u), c = (int)(w & 255u);
  const double V = a.fix_cv[fi];
  const double ug = a.u[g];
  const double vg = MODE == M_JVP ? a.v[g] : 0.0;
  double pv[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) pv[k] = p0 + k < p1 ? a.partial[p0 + k] : 0.0;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (p0 + k < p1) acc = acc + pv[k];
  for (int p = p0 + 4; p < p1; ++p) acc = acc + a.partial[p];
  for (int k = b0; k < b1; ++k) acc = acc - a.bflux[k];
  a.out[g] = finish_pre<MODE>(a, g, acc, ug, vg, c, V, MODE == M_EULER ? a.dtv[g] : 0.0);
}

// ---------------------------------------------------------------------------
// VJP
// ---------------------------------------------------------------------------
// w / V at nodes whose du carries the fluxes, 0 elsewhere.
__device__ __forceinline__ double wt_of(const KArgs& a, int g) {
  return a.code[g] == 0 ? a.v[g] / a.cv[g] : 0.0;
}
// Adjoint of the node rule: w_g dS/du (free nodes) or w_g da/du (Dudt nodes).
__device__ __forceinline__ double node_adj(const KArgs& a, int g, double ug) {
  const int c = a.code[g];
  if (c == 0) return a.v[g] * source_du(a, ug);
  if (c == 2) return a.v[g] * a.C.cu[a.fidx[g]];
  return 0.0;
}

// Adjoint of one CV-edge flux Q = -l D (alpha nx + beta ny) with upstream Qbar:
// accumulates alpha/beta/gamma adjoints and the adjoint of D.
template <int FK>
__device__ __forceinline__ void edge_adj(const KArgs& a, double Qbar, double x, double y, double nx,
                                         double ny, double l, double al, double be, double ga,
                                         double D_given, double& abar, double& bbar, double& gbar,
                                         double& Dbar) {
  const double gn = al * nx + be * ny;
  double D, Du = 0.0;
  if (d_of_u(FK)) {
    const double us = al * x + be * y + ga;
    D = dfun<FK>(a, us);
    Du = dfun_du<FK>(a, us);
  } else {
    D = D_given;
  }
  abar += Qbar * (-l * (D * nx + Du * x * gn));
  bbar += Qbar * (-l * (D * ny + Du * y * gn));
  gbar += Qbar * (-l * Du * gn);
  Dbar += Qbar * (-l * gn);
}

// Adjoint of both boundary CV-edge fluxes of boundary edge e with respect to the
// vertex in slot `slot` of its triangle.
template <int FK>
__device__ void bedge_adj(const KArgs& a, int e, int slot, double& ubar, double& lbar) {
  const BGeo G = bgeo(a, e);
  const int64_t Tt = a.badj[e];
  const int vtx[3] = {a.tri[Tt], a.tri[a.T + Tt], a.tri[2 * a.T + Tt]};
  double s[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) s[f] = a.props[f * a.T + Tt];
  const double ua = a.u[vtx[0]], ub = a.u[vtx[1]], uc = a.u[vtx[2]];
  const double al = s[0] * ua + s[1] * ub + s[2] * uc;
  const double be = s[3] * ua + s[4] * ub + s[5] * uc;
  const double ga = s[6] * ua + s[7] * ub + s[8] * uc;
  double Dn[3] = {0, 0, 0}, da = 0, db = 0, dg = 0;
  if (FK == FVM_FLUX_D_NODAL) {
    for (int v = 0; v < 3; ++v) Dn[v] = exp(a.log_d[vtx[v]]);
    da = s[0] * Dn[0] + s[1] * Dn[1] + s[2] * Dn[2];
    db = s[3] * Dn[0] + s[4] * Dn[1] + s[5] * Dn[2];
    dg = s[6] * Dn[0] + s[7] * Dn[1] + s[8] * Dn[2];
  }
  const int nf = a.bneu[e];
  double abar = 0, bbar = 0, gbar = 0, lb = 0;
  for (int ep = 0; ep < 2; ++ep) {
    const int node = a.bedges[ep * a.B + e];
    const double Qbar = -wt_of(a, node);  // du_node -= Q
    const double mx = ep ? G.mx[1] : G.mx[0], my = ep ? G.my[1] : G.my[0];
    if (nf < 0 && FK == FVM_FLUX_LINEAR) {
      const double* c = a.d_bedge + 3 * (2 * e + ep);
      abar += Qbar * c[0];
      bbar += Qbar * c[1];
      gbar += Qbar * c[2];
    } else if (nf < 0) {
      double Dg = 0, Dbar = 0;
      if (FK == FVM_FLUX_D_NODAL) Dg = da * mx + db * my + dg;
      if (FK == FVM_FLUX_D_EDGE) Dg = a.d_bedge[2 * e + ep];
      edge_adj<FK>(a, Qbar, mx, my, G.nx, G.ny, G.ell, al, be, ga, Dg, abar, bbar, gbar, Dbar);
      if (FK == FVM_FLUX_D_NODAL) lb += Dbar * (s[slot] * mx + s[3 + slot] * my + s[6 + slot]);
    } else {
      const double c = Qbar * G.ell * a.C.cu[nf];
      abar += c * mx;
      bbar += c * my;
      gbar += c;
    }
  }
  ubar += s[slot] * abar + s[3 + slot] * bbar + s[6 + slot] * gbar;
  lbar += lb * Dn[slot];
}

template <int FK, int BLOCK, int TPT>
__global__ __launch_bounds__(BLOCK) void vjp_tile_kernel(KArgs a) {
  constexpr int TILE = BLOCK * TPT;
  constexpr int MAXN = node_cap(BLOCK);
  constexpr int CW = FK == FVM_FLUX_D_NODAL ? 6 : 3;  // u-adjoint (+ log D adjoint) per vertex
  __shared__ double s_u[MAXN];
  constexpr bool RC = rebuilds(FK);  // part of the triangle record rebuilt from coordinates
  // w~, D, x and y are dead once the per-vertex adjoints are formed: the adjoints
  // reuse their space (fewer LDS bytes per workgroup, more workgroups per CU)
  constexpr int ND = FK == FVM_FLUX_D_NODAL ? MAXN : 0;
  constexpr int PRE = MAXN + ND + (RC ? 2 * MAXN : 0);
  __shared__ double s_big[PRE > CW * TILE ? PRE : CW * TILE];
  double* const s_w = s_big;
  double* const s_D = s_big + MAXN;
  double* const s_x = s_big + MAXN + ND;
  double* const s_y = s_x + MAXN;
  double* const s_c = s_big;
  const int nbw = (a.n_bitems + BLOCK - 1) / BLOCK;
  const int wid = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  if (wid < nbw) {
    // leading workgroups: adjoints of the boundary-edge fluxes per (edge, triangle vertex)
    const int k = wid * BLOCK + tid;
    if (k < a.n_bitems) {
      const int w = a.fix_bvert[k];
      double ub = 0.0, lb = 0.0;
      bedge_adj<FK>(a, w >> 2, w & 3, ub, lb);
      a.bflux[2 * k] = ub;
      a.bflux[2 * k + 1] = lb;
    }
    return;
  }
  const int tile = wid - nbw;
  if (tile >= a.n_tiles) return;
  const int t0 = a.tile_tri_start[tile], nt = a.tile_tri_start[tile + 1] - t0;
  const int n0 = a.tile_node_start[tile], nn = a.tile_node_start[tile + 1] - n0;
  double P[TPT][P_COUNT];
  double Ds[TPT][FK == FVM_FLUX_LINEAR ? 9 : 3];
  uint32_t L[TPT];
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
    L[r] = 0;
    if (lt < nt) {
      const int64_t s = t0 + lt;
      L[r] = a.loc[s];
#pragma unroll
      for (int f = 0; f < P_COUNT; ++f)
        if (loads_field(FK, f)) P[r][f] = __builtin_nontemporal_load(a.props + f * a.T + s);
      if (FK == FVM_FLUX_D_EDGE) {
#pragma unroll
        for (int e = 0; e < 3; ++e) Ds[r][e] = __builtin_nontemporal_load(a.d_edge + e * a.T + s);
      }
      if (FK == FVM_FLUX_LINEAR) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Ds[r][k] = __builtin_nontemporal_load(a.d_edge + k * a.T + s);
      }
    }
  }
  for (int l = tid; l < nn; l += BLOCK) {
    const int g = a.tile_nodes[n0 + l];
    s_u[l] = a.u[g];
    s_w[l] = wt_of(a, g);
    if (FK == FVM_FLUX_D_NODAL) s_D[l] = exp(a.log_d[g]);
    if (RC) {
      s_x[l] = a.x[g];
      s_y[l] = a.y[g];
    }
  }
  __syncthreads();
  double cr[TPT][CW];
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
#pragma unroll
    for (int v = 0; v < CW; ++v) cr[r][v] = 0.0;
    if (lt < nt) {
      const int id[3] = {(int)(L[r] & 1023u), (int)((L[r] >> 10) & 1023u), (int)((L[r] >> 20) & 1023u)};
      if (RC) rebuild_record<kGeo>(s_x[id[0]], s_y[id[0]], s_x[id[1]], s_y[id[1]], s_x[id[2]], s_y[id[2]], P[r]);
      const double* p = P[r];
      const double ua = s_u[id[0]], ub = s_u[id[1]], uc = s_u[id[2]];
      const double al = p[0] * ua + p[1] * ub + p[2] * uc;
      const double be = p[3] * ua + p[4] * ub + p[5] * uc;
      const double ga = d_of_u(FK) ? p[6] * ua + p[7] * ub + p[8] * uc : 0.0;
      double Dn[3] = {0, 0, 0}, da = 0, db = 0, dg = 0;
      if (FK == FVM_FLUX_D_NODAL) {
        Dn[0] = s_D[id[0]];
        Dn[1] = s_D[id[1]];
        Dn[2] = s_D[id[2]];
        da = p[0] * Dn[0] + p[1] * Dn[1] + p[2] * Dn[2];
        db = p[3] * Dn[0] + p[4] * Dn[1] + p[5] * Dn[2];
        dg = p[6] * Dn[0] + p[7] * Dn[1] + p[8] * Dn[2];
      }
      double abar = 0, bbar = 0, gbar = 0, lbar[3] = {0, 0, 0};
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        // du_e1 -= Q_e, du_e2 += Q_e  ->  Qbar = w~_e2 - w~_e1
        const double Qbar = s_w[id[(e + 1) % 3]] - s_w[id[e]];
        const double nx = p[P_NX + 2 * e], ny = p[P_NX + 1 + 2 * e], l = p[P_L + e];
        double x = 0, y = 0, Dg = 0;
        if (FK != FVM_FLUX_D_EDGE) {
          x = p[P_MX + 2 * e];
          y = p[P_MX + 1 + 2 * e];
        }
        if (FK == FVM_FLUX_D_NODAL) Dg = da * x + db * y + dg;
        if (FK == FVM_FLUX_D_EDGE) Dg = Ds[r][e];
        double Dbar = 0;
        if (FK == FVM_FLUX_LINEAR) {
          abar += Qbar * Ds[r][3 * e];
          bbar += Qbar * Ds[r][3 * e + 1];
          gbar += Qbar * Ds[r][3 * e + 2];
        } else {
          edge_adj<FK>(a, Qbar, x, y, nx, ny, l, al, be, ga, Dg, abar, bbar, gbar, Dbar);
        }
        if (FK == FVM_FLUX_D_NODAL) {
#pragma unroll
          for (int v = 0; v < 3; ++v) lbar[v] += Dbar * (p[v] * x + p[3 + v] * y + p[6 + v]);
        }
      }
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        double ub = p[v] * abar + p[3 + v] * bbar;
        if (d_of_u(FK) || FK == FVM_FLUX_LINEAR) ub = ub + p[6 + v] * gbar;  // gamma enters q
        cr[r][v] = ub;
        if (FK == FVM_FLUX_D_NODAL) cr[r][3 + v] = lbar[v] * Dn[v];
      }
    }
  }
  __syncthreads();  // every thread has read the weights / D / coordinates that s_c overwrites
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
    if (lt < nt) {
#pragma unroll
      for (int v = 0; v < CW; ++v) s_c[CW * lt + v] = cr[r][v];
    }
  }
  __syncthreads();
  for (int l = tid; l < nn; l += BLOCK) {
    double acc = 0.0, accl = 0.0;
    const int k0 = a.inc_start[n0 + l], k1 = a.inc_start[n0 + l + 1];
    for (int k = k0; k < k1; ++k) {
      const int w = a.inc[k];
      acc += s_c[CW * (w >> 2) + (w & 3)];
      if (FK == FVM_FLUX_D_NODAL) accl += s_c[CW * (w >> 2) + 3 + (w & 3)];
    }
    const int pi = a.part_idx[n0 + l];
    if (pi >= 0) {
      a.partial[pi] = acc;
      a.partial2[pi] = accl;
    } else {
      const int g = a.tile_nodes[n0 + l];
      a.out[g] = acc + node_adj(a, g, s_u[l]);
      if (a.out2) a.out2[g] = accl;
    }
  }
}

template <int FK>
__global__ __launch_bounds__(256) void vjp_fix_kernel(KArgs a) {
  const int fi = blockIdx.x * blockDim.x + threadIdx.x;
  if (fi >= a.n_fix) return;
  const int g = a.fix_nodes[fi];
  double acc = 0.0, accl = 0.0;
  for (int p = a.fix_part_start[fi]; p < a.fix_part_start[fi + 1]; ++p) {
    acc += a.partial[p];
    accl += a.partial2[p];
  }
  for (int k = a.fix_bvert_start[fi]; k < a.fix_bvert_start[fi + 1]; ++k) {
    acc += a.bflux[2 * k];
    accl += a.bflux[2 * k + 1];
  }
  a.out[g] = acc + node_adj(a, g, a.u[g]);
  if (a.out2) a.out2[g] = accl;
}

__global__ __launch_bounds__(256) void dirichlet_kernel(int n, const int32_t* __restrict__ nodes,
                                                        const int32_t* __restrict__ fidx,
                                                        const double* __restrict__ X,
                                                        const double* __restrict__ Y, CondTable C,
                                                        double t, double* __restrict__ u) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = nodes[k];
  u[g] = eval_cond(C, fidx[g], X[g], Y[g], t, u[g], C.node_g ? C.node_g[g] : 0.0);
}

KArgs make_args(const fvm_ctx* c) {
  const fvm_mesh* m = c->m;
  KArgs a;
  a.n_tiles = m->n_tiles;
  a.n_fix = m->n_fix;
  a.T = m->T;
  a.B = m->B;
  a.tile_tri_start = m->tile_tri_start;
  a.tile_node_start = m->tile_node_start;
  a.tile_nodes = m->tile_nodes;
  a.loc = m->loc;
  a.inc_start = m->inc_start;
  a.inc = m->inc;
  a.part_idx = m->part_idx;
  a.fix_nodes = m->fix_nodes;
  a.fix_part_start = m->fix_part_start;
  a.fix_bedge_start = m->fix_bedge_start;
  a.fix_meta = c->fix_meta;
  a.fix_cv = c->fix_cv;
  a.fix_bedge = m->fix_bedge;
  a.fix_bvert_start = m->fix_bvert_start;
  a.fix_bvert = m->fix_bvert;
  a.props = m->props;
  a.tri = m->tri;
  a.x = m->x;
  a.y = m->y;
  a.cv = m->cv;
  a.bedges = m->bedges;
  a.badj = m->badj;
  a.code = c->kind;
  a.fidx = c->fidx;
  a.bneu = c->bneu;
  a.d0 = c->dpoly[0];
  a.d1 = c->dpoly[1];
  a.d2 = c->dpoly[2];
  a.d3 = c->dpoly[3];
  a.d_edge = c->d_edge;
  a.d_bedge = c->d_bedge;
  a.log_d = c->log_d;
  a.s0 = c->spoly[0];
  a.s1 = c->spoly[1];
  a.s2 = c->spoly[2];
  a.s3 = c->spoly[3];
  a.st = c->st;
  a.s_node = c->s_node;
  a.C = c->cond();
  a.u = a.v = nullptr;
  a.out = a.out2 = nullptr;
  a.partial = c->partial;
  a.partial2 = c->partial + (m->n_parts > 0 ? m->n_parts : 1);
  a.lpart = c->lpart;
  a.lcv = m->lcv;
  a.bflux = c->bflux;
  a.n_bitems = m->n_bedge_items;
  a.t = 0;
  a.tile_dt = a.dtv = nullptr;
  return a;
}

// Grid: the boundary workgroups followed by one workgroup per tile.
inline dim3 tile_grid(const KArgs& a, int block) {
  return dim3((unsigned)xcd_grid(a.n_tiles + (a.n_bitems + block - 1) / block));
}

template <int FK, int MODE>
void launch_tiles(const KArgs& a, int tile_tris, hipStream_t s) {
  if (tile_tris <= 256)
    hipLaunchKernelGGL((tile_kernel<FK, MODE, 256, 1>), tile_grid(a, 256), dim3(256), 0, s, a);
  else if (tile_tris <= 512)
    hipLaunchKernelGGL((tile_kernel<FK, MODE, 256, 2>), tile_grid(a, 256), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((tile_kernel<FK, MODE, 512, 2>), tile_grid(a, 512), dim3(512), 0, s, a);
}

template <int FK, int MODE>
void launch_fix(const KArgs& a, hipStream_t s) {
  if (a.n_fix > 0)
    hipLaunchKernelGGL((fix_kernel<FK, MODE>), dim3((unsigned)((a.n_fix + 255) / 256)), dim3(256), 0,
                       s, a);
}

template <int FK, int MODE>
int launch_fk(const KArgs& a, int tile_tris, hipStream_t s) {
  launch_tiles<FK, MODE>(a, tile_tris, s);
  launch_fix<FK, MODE>(a, s);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

template <int MODE>
int launch(const fvm_ctx* c, const KArgs& a, hipStream_t s) {
  switch (c->flux_kind) {
    case FVM_FLUX_D_POLY_U: return launch_fk<FVM_FLUX_D_POLY_U, MODE>(a, c->m->tile_tris, s);
    case FVM_FLUX_D_NODAL: return launch_fk<FVM_FLUX_D_NODAL, MODE>(a, c->m->tile_tris, s);
    case FVM_FLUX_LINEAR: return launch_fk<FVM_FLUX_LINEAR, MODE>(a, c->m->tile_tris, s);
    case FVM_FLUX_D_POWER: return launch_fk<FVM_FLUX_D_POWER, MODE>(a, c->m->tile_tris, s);
    default: return launch_fk<FVM_FLUX_D_EDGE, MODE>(a, c->m->tile_tris, s);
  }
}

template <int FK>
int launch_vjp_fk(const KArgs& a, int tile_tris, hipStream_t s) {
  if (tile_tris <= 256)
    hipLaunchKernelGGL((vjp_tile_kernel<FK, 256, 1>), tile_grid(a, 256), dim3(256), 0, s, a);
  else if (tile_tris <= 512)
    hipLaunchKernelGGL((vjp_tile_kernel<FK, 256, 2>), tile_grid(a, 256), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((vjp_tile_kernel<FK, 512, 2>), tile_grid(a, 512), dim3(512), 0, s, a);
  if (a.n_fix > 0)
    hipLaunchKernelGGL((vjp_fix_kernel<FK>), dim3((unsigned)((a.n_fix + 255) / 256)), dim3(256), 0, s,
                       a);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

template <class T>
int dev_copy(T** dst, const T* src, int64_t n) {
  *dst = nullptr;
  FVM_HIP(hipMalloc((void**)dst, sizeof(T) * (size_t)(n > 0 ? n : 1)));
  if (n > 0 && src) FVM_HIP(hipMemcpy(*dst, src, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
  return FVM_OK;
}

// out[k] = x[idx[k]]
__global__ __launch_bounds__(256) void gather_kernel(int64_t n, const int32_t* __restrict__ idx,
                                                     const double* __restrict__ x, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = x[idx[k]];
}

// FVM_FLUX_LINEAR: q = (a1 alpha + a2 beta + a3 gamma, a4 alpha + a5 beta + a6 gamma)
// is projected once onto each CV edge's unit normal and length, so the kernels form
// (q . n) l = c1 alpha + c2 beta + c3 gamma.  raw [6][3][T] sorted (row 6 e + k),
// out [3][3][T] (row 3 e + k).
__global__ __launch_bounds__(256) void project_edge_kernel(int64_t T, const double* __restrict__ props,
                                                           const double* __restrict__ raw, double* __restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= T) return;
  for (int e = 0; e < 3; ++e) {
    const double nx = props[(P_NX + 2 * e) * T + s], ny = props[(P_NX + 1 + 2 * e) * T + s];
    const double l = props[(P_L + e) * T + s];
    for (int k = 0; k < 3; ++k)
      out[(3 * e + k) * T + s] = (raw[(6 * e + k) * T + s] * nx + raw[(6 * e + 3 + k) * T + s] * ny) * l;
  }
}

// The boundary half-edges: raw [2B][6], out [2B][3] with the outward normal and the
// half-edge length of get_boundary_cv_components (control_volumes.jl:41-56).
__global__ __launch_bounds__(256) void project_bedge_kernel(KArgs a, const double* __restrict__ raw,
                                                            double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 2 * a.B) return;
  const BGeo G = bgeo(a, (int)(k >> 1));
  const double* A = raw + 6 * k;
  for (int c = 0; c < 3; ++c) out[3 * k + c] = (A[c] * G.nx + A[3 + c] * G.ny) * G.ell;
}

int project_linear(const fvm_mesh* m, const fvm_physics_desc* ph, fvm_ctx* c) {
  const int64_t T = m->T, B = m->B;
  std::vector<double> raw(18 * (size_t)T);
  for (int64_t s = 0; s < T; ++s)
    for (int j = 0; j < 18; ++j) raw[(size_t)j * T + s] = ph->d_edge[18 * (int64_t)m->h_perm[s] + j];
  double *d_raw = nullptr, *d_braw = nullptr;
  int rc = dev_copy(&d_raw, raw.data(), 18 * T);
  if (rc == FVM_OK) rc = dev_copy(&d_braw, ph->d_bedge, 12 * B);
  if (rc == FVM_OK) rc = dev_copy(&c->d_edge, (const double*)nullptr, 9 * T);
  if (rc == FVM_OK) rc = dev_copy(&c->d_bedge, (const double*)nullptr, 6 * B);
  if (rc == FVM_OK && T > 0)
    hipLaunchKernelGGL(project_edge_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, nullptr, T, m->props,
                       d_raw, c->d_edge);
  if (rc == FVM_OK && B > 0) {
    KArgs k{};
    k.B = (int32_t)B;
    k.x = m->x;
    k.y = m->y;
    k.bedges = m->bedges;
    hipLaunchKernelGGL(project_bedge_kernel, dim3((unsigned)((2 * B + 255) / 256)), dim3(256), 0, nullptr, k,
                       d_braw, c->d_bedge);
  }
  if (rc == FVM_OK && hipDeviceSynchronize() != hipSuccess) rc = FVM_ERR_HIP;
  if (rc == FVM_OK && hipGetLastError() != hipSuccess) rc = FVM_ERR_HIP;
  if (d_raw) (void)hipFree(d_raw);
  if (d_braw) (void)hipFree(d_braw);
  return rc;
}

void free_ctx(fvm_ctx* c) {
  if (!c) return;
  void* ptrs[] = {c->d_edge, c->d_bedge, c->log_d_own, c->s_node, c->kind, c->fidx,
                  c->bneu, c->fn, c->node_g, c->bedge_g, c->dir_nodes, c->partial,
                  c->lpart, c->bflux, c->fix_meta, c->fix_cv};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete c;
}

int build_ctx(const fvm_mesh* m, const fvm_physics_desc* ph, fvm_ctx* c) {
  const int64_t N = m->N, T = m->T, B = m->B;
  FVM_CHECK(ph->flux_kind >= FVM_FLUX_D_POLY_U && ph->flux_kind <= FVM_FLUX_D_POWER,
            "fvm_ctx_create: unknown flux_kind");
  FVM_CHECK(ph->flux_kind != FVM_FLUX_D_POWER || ph->dpoly[2] > 0,
            "fvm_ctx_create: FVM_FLUX_D_POWER needs a positive exponent dpoly[2]");
  FVM_CHECK(ph->node_kind, "fvm_ctx_create: node_kind is required");
  FVM_CHECK(ph->node_fidx, "fvm_ctx_create: node_fidx is required");
  FVM_CHECK(B == 0 || ph->bedge_neumann_fidx, "fvm_ctx_create: bedge_neumann_fidx is required");
  FVM_CHECK(ph->n_functions >= 0, "fvm_ctx_create: negative n_functions");
  FVM_CHECK(ph->n_functions == 0 || (ph->fn_c0 && ph->fn_cu && ph->fn_ct),
            "fvm_ctx_create: condition function coefficients missing");
  if (ph->flux_kind == FVM_FLUX_D_EDGE || ph->flux_kind == FVM_FLUX_LINEAR)
    FVM_CHECK(ph->d_edge && (B == 0 || ph->d_bedge), "fvm_ctx_create: d_edge / d_bedge required");
  if (ph->flux_kind == FVM_FLUX_D_NODAL) FVM_CHECK(ph->log_d, "fvm_ctx_create: log_d required");
  c->m = m;
  c->flux_kind = ph->flux_kind;
  for (int k = 0; k < 4; ++k) {
    c->dpoly[k] = ph->dpoly[k];
    c->spoly[k] = ph->spoly[k];
  }
  c->st = ph->st;
  c->n_fn = ph->n_functions;
  // node codes: 3 marks points that are not vertices (du = 0, source_contributions.jl:36-37)
  std::vector<int8_t> code(N);
  std::vector<int32_t> dir;
  for (int64_t g = 0; g < N; ++g) {
    const int k = ph->node_kind[g];
    FVM_CHECK(k >= 0 && k <= 2, "fvm_ctx_create: node_kind entries must be 0, 1 or 2");
    if (k != 0) {
      const int f = ph->node_fidx[g];
      FVM_CHECK(f >= 0 && f < ph->n_functions, "fvm_ctx_create: node_fidx out of range");
    }
    const bool vert = m->h_node_tri_start[g + 1] > m->h_node_tri_start[g];
    code[g] = vert ? (int8_t)k : (int8_t)3;
    if (k == FVM_NODE_DIRICHLET) dir.push_back((int32_t)g);
  }
  for (int64_t e = 0; e < B; ++e) {
    const int f = ph->bedge_neumann_fidx[e];
    FVM_CHECK(f == -1 || (f >= 0 && f < ph->n_functions), "fvm_ctx_create: bedge_neumann_fidx out of range");
  }
  int rc;
#define DC(dst, src, n) \
  if ((rc = dev_copy(&(dst), src, n)) != FVM_OK) return rc
  DC(c->kind, code.data(), N);
  DC(c->fidx, ph->node_fidx, N);
  DC(c->bneu, ph->bedge_neumann_fidx, B);
  {
    const int n = ph->n_functions;
    std::vector<double> fn(6 * (size_t)(n > 0 ? n : 1), 0.0);
    for (int f = 0; f < n; ++f) {
      fn[f] = ph->fn_c0[f];
      fn[n + f] = ph->fn_cu[f];
      fn[2 * n + f] = ph->fn_ct[f];
      fn[3 * n + f] = ph->fn_cx ? ph->fn_cx[f] : 0.0;
      fn[4 * n + f] = ph->fn_cy ? ph->fn_cy[f] : 0.0;
      fn[5 * n + f] = ph->fn_lam ? ph->fn_lam[f] : 0.0;
    }
    DC(c->fn, fn.data(), (int64_t)fn.size());
  }
  if (ph->node_g) DC(c->node_g, ph->node_g, N);
  if (ph->bedge_g) DC(c->bedge_g, ph->bedge_g, 2 * B);
  if (ph->s_node) DC(c->s_node, ph->s_node, N);
  if (ph->flux_kind == FVM_FLUX_D_EDGE) {
    std::vector<double> de(3 * (size_t)T);
    for (int64_t s = 0; s < T; ++s)
      for (int e = 0; e < 3; ++e) de[(size_t)e * T + s] = ph->d_edge[3 * (int64_t)m->h_perm[s] + e];
    DC(c->d_edge, de.data(), 3 * T);
    DC(c->d_bedge, ph->d_bedge, 2 * B);
  }
  if (ph->flux_kind == FVM_FLUX_LINEAR && (rc = project_linear(m, ph, c)) != FVM_OK) return rc;
  if (ph->log_d) {
    DC(c->log_d_own, ph->log_d, N);
    c->log_d = c->log_d_own;
  }
  c->n_dir = (int32_t)dir.size();
  DC(c->dir_nodes, dir.data(), (int64_t)dir.size());
  DC(c->partial, (const double*)nullptr, 2 * (int64_t)(m->n_parts > 0 ? m->n_parts : 1));
  {
    // exclusive entries carry their node's code: -1 - code
    std::vector<int32_t> lp(m->n_local);
    for (int64_t e = 0; e < m->n_local; ++e) {
      const int32_t p = m->h_part_idx[e];
      lp[e] = p >= 0 ? p : -1 - (int32_t)code[m->h_tile_nodes[e]];
    }
    DC(c->lpart, lp.data(), m->n_local);
  }
  {
    const int64_t nb = std::max<int64_t>(m->n_bedge_items, 2 * (int64_t)m->n_bvert_items);
    DC(c->bflux, (const double*)nullptr, nb);
  }
  {
    // per fix node: node, first slot, first boundary item, slots << 20 | items << 8 | code
    const int64_t nf = m->n_fix;
    std::vector<int32_t> meta(4 * (size_t)std::max<int64_t>(nf, 1), 0);
    std::vector<int32_t> hfix(std::max<int64_t>(nf, 1), 0);
    if (nf > 0) FVM_HIP(hipMemcpy(hfix.data(), m->fix_nodes, sizeof(int32_t) * nf, hipMemcpyDeviceToHost));
    for (int64_t f = 0; f < nf; ++f) {
      const int32_t np = m->h_fix_part_start[f + 1] - m->h_fix_part_start[f];
      const int32_t nb = m->h_fix_bedge_start[f + 1] - m->h_fix_bedge_start[f];
      FVM_CHECK(np < 4096 && nb < 4096, "fvm_ctx_create: a node shared by 4096 or more tiles");
      meta[f] = hfix[f];
      meta[nf + f] = m->h_fix_part_start[f];
      meta[2 * nf + f] = m->h_fix_bedge_start[f];
      meta[3 * nf + f] = (int32_t)(((uint32_t)np << 20) | ((uint32_t)nb << 8) | (uint32_t)(uint8_t)code[hfix[f]]);
    }
    DC(c->fix_meta, meta.data(), (int64_t)meta.size());
    DC(c->fix_cv, (const double*)nullptr, std::max<int64_t>(nf, 1));
    if (nf > 0) {
      hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, 0, nf, m->fix_nodes,
                         (const double*)m->cv, c->fix_cv);
      FVM_HIP(hipGetLastError());
      FVM_HIP(hipDeviceSynchronize());
    }
  }
#undef DC
  return FVM_OK;
}

}  // namespace

namespace fvm {

int euler_step(fvm_ctx* c, double* d_u, const double* tile_dt, const double* dtv, double t, hipStream_t s) {
  KArgs a = make_args(c);
  a.u = d_u;
  a.out = d_u;  // in place: a tile writes only its exclusive nodes, the fix-up the shared ones
  a.t = t;
  a.tile_dt = tile_dt;
  a.dtv = dtv;
  return launch<M_EULER>(c, a, s);
}

}  // namespace fvm

extern "C" {

int fvm_ctx_create(const fvm_mesh* m, const fvm_physics_desc* ph, fvm_ctx** out) {
  FVM_CHECK(m && ph && out, "fvm_ctx_create: null argument");
  *out = nullptr;
  fvm_ctx* c = new fvm_ctx();
  int rc = build_ctx(m, ph, c);
  if (rc != FVM_OK) {
    free_ctx(c);
    return rc;
  }
  *out = c;
  return FVM_OK;
}

int fvm_ctx_destroy(fvm_ctx* c) {
  free_ctx(c);
  return FVM_OK;
}

int fvm_ctx_set_log_d(fvm_ctx* c, const double* d_log_d) {
  FVM_CHECK(c && d_log_d, "fvm_ctx_set_log_d: null argument");
  FVM_CHECK(c->flux_kind == FVM_FLUX_D_NODAL, "fvm_ctx_set_log_d: context has no nodal field");
  c->log_d = d_log_d;
  return FVM_OK;
}

int fvm_rhs(fvm_ctx* c, const double* d_u, double* d_du, double t, fvm_stream s) {
  FVM_CHECK(c && d_u && d_du, "fvm_rhs: null argument");
  FVM_CHECK(d_u != d_du, "fvm_rhs: du must not alias u");
  KArgs a = make_args(c);
  a.u = d_u;
  a.out = d_du;
  a.t = t;
  return launch<M_RHS>(c, a, (hipStream_t)s);
}

int fvm_rhs_timed(fvm_ctx* c, const double* d_u, double* d_du, double t, fvm_stream s,
                  float* ms_tiles, float* ms_fixup) {
  FVM_CHECK(c && d_u && d_du && ms_tiles && ms_fixup, "fvm_rhs_timed: null argument");
  FVM_CHECK(d_u != d_du, "fvm_rhs_timed: du must not alias u");
  KArgs a = make_args(c);
  a.u = d_u;
  a.out = d_du;
  a.t = t;
  hipStream_t st = (hipStream_t)s;
  hipEvent_t ev[3];
  for (int k = 0; k < 3; ++k) FVM_HIP(hipEventCreate(&ev[k]));
  FVM_HIP(hipEventRecord(ev[0], st));
  int rc = FVM_OK;
  const int tt = c->m->tile_tris;
  switch (c->flux_kind) {
    case FVM_FLUX_D_POLY_U: launch_tiles<FVM_FLUX_D_POLY_U, M_RHS>(a, tt, st); break;
    case FVM_FLUX_D_NODAL: launch_tiles<FVM_FLUX_D_NODAL, M_RHS>(a, tt, st); break;
    case FVM_FLUX_LINEAR: launch_tiles<FVM_FLUX_LINEAR, M_RHS>(a, tt, st); break;
    case FVM_FLUX_D_POWER: launch_tiles<FVM_FLUX_D_POWER, M_RHS>(a, tt, st); break;
    default: launch_tiles<FVM_FLUX_D_EDGE, M_RHS>(a, tt, st);
  }
  FVM_HIP(hipEventRecord(ev[1], st));
  switch (c->flux_kind) {
    case FVM_FLUX_D_POLY_U: launch_fix<FVM_FLUX_D_POLY_U, M_RHS>(a, st); break;
    case FVM_FLUX_D_NODAL: launch_fix<FVM_FLUX_D_NODAL, M_RHS>(a, st); break;
    case FVM_FLUX_LINEAR: launch_fix<FVM_FLUX_LINEAR, M_RHS>(a, st); break;
    case FVM_FLUX_D_POWER: launch_fix<FVM_FLUX_D_POWER, M_RHS>(a, st); break;
    default: launch_fix<FVM_FLUX_D_EDGE, M_RHS>(a, st);
  }
  FVM_HIP(hipEventRecord(ev[2], st));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
  if (e == hipSuccess) e = hipEventElapsedTime(ms_tiles, ev[0], ev[1]);
  if (e == hipSuccess) e = hipEventElapsedTime(ms_fixup, ev[1], ev[2]);
  for (int k = 0; k < 3; ++k) (void)hipEventDestroy(ev[k]);
  if (e != hipSuccess) rc = hip_fail(e, "fvm_rhs_timed");
  return rc;
}

int fvm_jvp(fvm_ctx* c, const double* d_u, const double* d_v, double* d_jv, double t, fvm_stream s) {
  FVM_CHECK(c && d_u && d_v && d_jv, "fvm_jvp: null argument");
  FVM_CHECK(d_jv != d_u && d_jv != d_v, "fvm_jvp: output must not alias inputs");
  KArgs a = make_args(c);
  a.u = d_u;
  a.v = d_v;
  a.out = d_jv;
  a.t = t;
  return launch<M_JVP>(c, a, (hipStream_t)s);
}

int fvm_vjp(fvm_ctx* c, const double* d_u, const double* d_w, double* d_ubar, double* d_logd_bar,
            double t, fvm_stream s) {
  FVM_CHECK(c && d_u && d_w && d_ubar, "fvm_vjp: null argument");
  FVM_CHECK(d_ubar != d_u && d_ubar != d_w, "fvm_vjp: output must not alias inputs");
  FVM_CHECK(d_logd_bar == nullptr || c->flux_kind == FVM_FLUX_D_NODAL,
            "fvm_vjp: d_logd_bar needs a FVM_FLUX_D_NODAL context");
  KArgs a = make_args(c);
  a.u = d_u;
  a.v = d_w;
  a.out = d_ubar;
  a.out2 = d_logd_bar;
  a.t = t;
  a.n_bitems = c->m->n_bvert_items;
  hipStream_t st = (hipStream_t)s;
  switch (c->flux_kind) {
    case FVM_FLUX_D_POLY_U: return launch_vjp_fk<FVM_FLUX_D_POLY_U>(a, c->m->tile_tris, st);
    case FVM_FLUX_D_NODAL: return launch_vjp_fk<FVM_FLUX_D_NODAL>(a, c->m->tile_tris, st);
    case FVM_FLUX_LINEAR: return launch_vjp_fk<FVM_FLUX_LINEAR>(a, c->m->tile_tris, st);
    case FVM_FLUX_D_POWER: return launch_vjp_fk<FVM_FLUX_D_POWER>(a, c->m->tile_tris, st);
    default: return launch_vjp_fk<FVM_FLUX_D_EDGE>(a, c->m->tile_tris, st);
  }
}

int fvm_dirichlet_apply(fvm_ctx* c, double* d_u, double t, fvm_stream s) {
  FVM_CHECK(c && d_u, "fvm_dirichlet_apply: null argument");
  if (c->n_dir == 0) return FVM_OK;
  hipLaunchKernelGGL(dirichlet_kernel, dim3((unsigned)((c->n_dir + 255) / 256)), dim3(256), 0,
                     (hipStream_t)s, c->n_dir, c->dir_nodes, c->fidx, c->m->x, c->m->y, c->cond(), t,
                     d_u);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

}  // extern "C"
