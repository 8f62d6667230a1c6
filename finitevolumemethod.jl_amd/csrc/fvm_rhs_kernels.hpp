// Device code of fvm_eqs! and its derivatives: the triangle-tile kernel, the
// shared-node fix-up, the VJP kernels and the Dirichlet update (see fvm_rhs.hip for
// the scheme).  Compiled into libfvmhip.so for the traced closure forms and, for
// closures outside them (FVM_FLUX_JIT), at run time by hipRTC together with the
// generated JitPhysics specialisation (fvm_jit.hip); under hipRTC nothing is included.
#pragma once
#include "fvm_device.hpp"

namespace fvm_k {

using namespace fvm;

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
__host__ __device__ constexpr bool d_of_u(int fk) {
  return fk == FVM_FLUX_D_POLY_U || fk == FVM_FLUX_D_POWER || fk == FVM_FLUX_JIT;
}
// Closures compiled at run time (hipRTC): the generated code specialises
// JitPhysics<FVM_FLUX_JIT> with D(x, y, t, u), dD/du, S(x, y, t, u) and dS/du; the
// precompiled kernels are never instantiated with FK = FVM_FLUX_JIT.
template <int FK>
struct JitPhysics;
// D at the point (x, y) with solution value u there
template <int FK>
__device__ __forceinline__ double dfun(const KArgs& a, double u, double x, double y) {
  if constexpr (FK == FVM_FLUX_JIT) return JitPhysics<FK>::D(x, y, a.t, u);
  if (FK == FVM_FLUX_D_POWER) return a.d0 + a.d1 * pow(fabs(u), a.d2);
  return dpoly(a, u);
}
// dD/du; for the power law d1 d2 |u|^(d2 - 1) sign(u), taken as 0 at u = 0
template <int FK>
__device__ __forceinline__ double dfun_du(const KArgs& a, double u, double x, double y) {
  if constexpr (FK == FVM_FLUX_JIT) return JitPhysics<FK>::dD(x, y, a.t, u);
  if (FK == FVM_FLUX_D_POWER) {
    const double au = fabs(u);
    if (!(au > 0.0)) return 0.0;
    const double d = a.d1 * a.d2 * pow(au, a.d2 - 1.0);
    return u > 0.0 ? d : -d;
  }
  return dpoly_du(a, u);
}

// S(x, y, t, u) of the source term (source_contributions.jl:2-14)
template <int FK>
__device__ __forceinline__ double source(const KArgs& a, int g, double u) {
  if constexpr (FK == FVM_FLUX_JIT) return JitPhysics<FK>::S(a.x[g], a.y[g], a.t, u);
  double S = a.s0 + u * (a.s1 + u * (a.s2 + u * a.s3));
  if (a.s_node) S = a.s_node[g] + S;
  if (a.st != 0.0) S = S + a.st * a.t;
  return S;
}
template <int FK>
__device__ __forceinline__ double source_du(const KArgs& a, int g, double u) {
  if constexpr (FK == FVM_FLUX_JIT) return JitPhysics<FK>::dS(a.x[g], a.y[g], a.t, u);
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
      D = dfun<FK>(a, us, x, y);
      if (MODE == M_JVP) Dd = dfun_du<FK>(a, us, x, y) * (dal * x + dbe * y + dga);
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
      D = dfun<FK>(a, um, mx, my);
      if (MODE == M_JVP) Dd = dfun_du<FK>(a, um, mx, my) * (dal * mx + dbe * my + dga);
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
template <int FK, int MODE>
__device__ __forceinline__ double finish_pre(const KArgs& a, int g, double acc, double ug, double vg,
                                             int c, double V, double dt) {
  if (c == 0) {
    if (MODE == M_RHS) return acc / V + source<FK>(a, g, ug);
    if (MODE == M_EULER) return ug + dt * (acc / V + source<FK>(a, g, ug));
    return acc / V + source_du<FK>(a, g, ug) * vg;
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
  // space (fewer LDS bytes per workgroup, more workgroups per CU).  The fluxes are
  // four rows of R = TILE + 1 doubles, (Q3, Q1, Q2, Q3) of triangle lt in column lt,
  // so that a thread's four stores go to consecutive addresses across the wave (no
  // bank conflicts), and the incidence entry of (lt, slot), restaged as
  // w = slot R + lt, addresses its vertex's two terms directly: s_q[w] is the flux
  // added and s_q[w + R] the flux subtracted (i: Q3 - Q1, j: Q1 - Q2, k: Q2 - Q3).
  // Column TILE of rows 0 and 1 holds zeros for the padding entries of the rounds.
  constexpr int R = TILE + 1;
  constexpr int XY = RC ? 2 * MAXN : 0;
  static_assert(XY <= 4 * R, "the coordinates fit in the flux rows");
  static_assert(3 * R <= 65535, "restaged entries fit 16 bits");
  __shared__ double s_xyq[4 * R];
  double* const s_x = s_xyq;
  double* const s_y = s_xyq + MAXN;
  double* const s_q = s_xyq;
  __shared__ uint16_t s_inc[3 * TILE + 4];  // + 4: a round may read past a node's list
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

  // Every lane loads: lanes past the tile's nodes, entries or triangles read the last
  // valid one (their values are never used), so the loads need no branches, and each
  // takes a 32-bit offset from a base in scalar registers.
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
  double P[TPT][P_COUNT];
  double Ds[TPT][FK == FVM_FLUX_LINEAR ? 9 : 3];
  uint32_t L[TPT];
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
    L[r] = ld32(a.loc, (uint32_t)t0 + umin((uint32_t)lt, (uint32_t)nt - 1u));
    if (lt < nt) {
      const int64_t s = t0 + lt;
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
  double ug[NS], vg[NS], lg[NS], xg[NS], yg[NS];
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const uint32_t gi = (uint32_t)g[r];
    ug[r] = ld32(a.u, gi);
    vg[r] = MODE == M_JVP ? ld32(a.v, gi) : 0.0;
    lg[r] = FK == FVM_FLUX_D_NODAL ? ld32(a.log_d, gi) : 0.0;
    xg[r] = RC ? ld32(a.x, gi) : 0.0;
    yg[r] = RC ? ld32(a.y, gi) : 0.0;
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
      s_is[l] = is[r] - ib0;
    }
  }
  if (tid == 0) s_is[nn] = 3 * nt;
#pragma unroll
  for (int k = 0; k < 3 * TPT; ++k) {
    const int idx = tid + k * BLOCK;
    if (idx < 3 * nt) s_inc[idx] = (uint16_t)((iv[k] & 3) * R + (iv[k] >> 2));
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
      s_q[lt] = Q[r][2];
      s_q[R + lt] = Q[r][0];
      s_q[2 * R + lt] = Q[r][1];
      s_q[3 * R + lt] = Q[r][2];
    }
  }
  if (tid == 0) {
    s_q[TILE] = 0.0;
    s_q[R + TILE] = 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    if (l < nn) {
      double acc = 0.0;
      const int k1 = s_is[l + 1];
      // four incidences per round: their entries, then their flux pairs, then the
      // sums in list order (two LDS round trips per four terms); entries past the
      // list point at the two zeros, whose sum leaves acc unchanged bit for bit
      for (int k = s_is[l]; k < k1; k += 4) {
        int w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int e = s_inc[k + j];
          w[j] = k + j < k1 ? e : TILE;
        }
        double qa[4], qb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          qa[j] = s_q[w[j]];
          qb[j] = s_q[w[j] + R];
        }
        // update_du! (triangle_contributions.jl:10-15): i += Q3 - Q1, j += Q1 - Q2, k += Q2 - Q3
#pragma unroll
        for (int j = 0; j < 4; ++j) acc = acc + qa[j] - qb[j];
      }
      // the node's u (and v) are re-read from LDS and its volume from memory here rather
      // than held in registers across the barriers: fewer VGPRs, more waves per SIMD
      if (pi[r] >= 0)
        st32(a.partial, (uint32_t)pi[r], acc);
      else
        st32(a.out, (uint32_t)g[r],
             finish_pre<FK, MODE>(a, g[r], acc, s_u[l], MODE == M_JVP ? s_v[l] : 0.0, -1 - pi[r],
                                  ld32(a.lcv, (uint32_t)(n0 + l)), dt));
    }
  }
}

// Shared nodes: partial sums in tile order, then the boundary-edge fluxes (computed
// by the tile kernel's leading workgroups), then the node rule.  The first four slots are loaded without a loop-carried
// dependency (nearly every node has at most four).
template <int FK, int MODE>
__global__ __launch_bounds__(256) void fix_kernel(KArgs a) {
  const int fi = fix_wid() * blockDim.x + threadIdx.x;
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
  a.out[g] = finish_pre<FK, MODE>(a, g, acc, ug, vg, c, V, MODE == M_EULER ? a.dtv[g] : 0.0);
}

// ---------------------------------------------------------------------------
// VJP
// ---------------------------------------------------------------------------
// w / V at nodes whose du carries the fluxes, 0 elsewhere.
__device__ __forceinline__ double wt_of(const KArgs& a, int g) {
  return a.code[g] == 0 ? a.v[g] / a.cv[g] : 0.0;
}
// Adjoint of the node rule: w_g dS/du (free nodes) or w_g da/du (Dudt nodes).
template <int FK>
__device__ __forceinline__ double node_adj(const KArgs& a, int g, double ug) {
  const int c = a.code[g];
  if (c == 0) return a.v[g] * source_du<FK>(a, g, ug);
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
    D = dfun<FK>(a, us, x, y);
    Du = dfun_du<FK>(a, us, x, y);
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
  constexpr bool RC = rebuilds(FK);  // part of the triangle record rebuilt from coordinates
  // w~, D, x, y and u are dead once the per-vertex adjoints are formed: the adjoints
  // reuse their space (fewer LDS bytes per workgroup, more workgroups per CU; the
  // node rule reads u again from memory)
  constexpr int ND = FK == FVM_FLUX_D_NODAL ? MAXN : 0;
  constexpr int PRE = 2 * MAXN + ND + (RC ? 2 * MAXN : 0);
  // + CW zeros after the per-vertex adjoints: the target of the padding entries of
  // the incidence rounds
  __shared__ double s_big[PRE > CW * TILE + CW ? PRE : CW * TILE + CW];
  double* const s_w = s_big;
  double* const s_u = s_big + MAXN;
  double* const s_D = s_big + 2 * MAXN;
  double* const s_x = s_big + 2 * MAXN + ND;
  double* const s_y = s_x + MAXN;
  double* const s_c = s_big;
  __shared__ uint16_t s_inc[3 * TILE];
  __shared__ int s_is[MAXN + 1];
  constexpr int NS = MAXN / BLOCK;
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
  const int ib0 = 3 * t0;  // the tile's incidence block is [3 t0, 3 (t0 + nt))
  // every load of the tile issued before the first wait, branch-free (lanes past the
  // tile's nodes, entries or triangles read the last valid one and never use it), at
  // 32-bit offsets from bases in scalar registers
  int g[NS], is[NS];
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const uint32_t li = (uint32_t)n0 + umin((uint32_t)(tid + r * BLOCK), (uint32_t)nn - 1u);
    g[r] = ld32(a.tile_nodes, li);
    is[r] = ld32(a.inc_start, li);
  }
  uint16_t iv[3 * TPT];
#pragma unroll
  for (int k = 0; k < 3 * TPT; ++k)
    iv[k] = ld32(a.inc, (uint32_t)ib0 + umin((uint32_t)(tid + k * BLOCK), 3u * (uint32_t)nt - 1u));
  double P[TPT][P_COUNT];
  double Ds[TPT][FK == FVM_FLUX_LINEAR ? 9 : 3];
  uint32_t L[TPT];
#pragma unroll
  for (int r = 0; r < TPT; ++r) {
    const int lt = tid + r * BLOCK;
    L[r] = ld32(a.loc, (uint32_t)t0 + umin((uint32_t)lt, (uint32_t)nt - 1u));
    if (lt < nt) {
      const int64_t s = t0 + lt;
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
  double ug[NS], vg[NS], cvg[NS], lg[NS], xg[NS], yg[NS];
  int8_t cg[NS];
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const uint32_t gi = (uint32_t)g[r];
    ug[r] = ld32(a.u, gi);
    vg[r] = ld32(a.v, gi);
    cvg[r] = ld32(a.cv, gi);
    cg[r] = ld32(a.code, gi);
    lg[r] = FK == FVM_FLUX_D_NODAL ? ld32(a.log_d, gi) : 0.0;
    xg[r] = RC ? ld32(a.x, gi) : 0.0;
    yg[r] = RC ? ld32(a.y, gi) : 0.0;
  }
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    if (l < nn) {
      s_u[l] = ug[r];
      s_w[l] = cg[r] == 0 ? vg[r] / cvg[r] : 0.0;  // wt_of
      if (FK == FVM_FLUX_D_NODAL) s_D[l] = exp(lg[r]);
      if (RC) {
        s_x[l] = xg[r];
        s_y[l] = yg[r];
      }
      s_is[l] = is[r] - ib0;
    }
  }
  if (tid == 0) s_is[nn] = 3 * nt;
#pragma unroll
  for (int k = 0; k < 3 * TPT; ++k)
    if (tid + k * BLOCK < 3 * nt) s_inc[tid + k * BLOCK] = iv[k];
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
  if (tid < CW) s_c[CW * TILE + tid] = 0.0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NS; ++r) {
    const int l = tid + r * BLOCK;
    if (l >= nn) continue;
    double acc = 0.0, accl = 0.0;
    const int k1 = s_is[l + 1];
    // four incidences per round: their entries (reads clamped into the node's list),
    // then the sums in list order; entries past the list address the zeros, whose
    // sum leaves the accumulators unchanged bit for bit
    for (int k = s_is[l]; k < k1; k += 4) {
      int w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = s_inc[k + j < k1 ? k + j : k1 - 1];
        w[j] = k + j < k1 ? e : TILE << 2;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc += s_c[CW * (w[j] >> 2) + (w[j] & 3)];
        if (FK == FVM_FLUX_D_NODAL) accl += s_c[CW * (w[j] >> 2) + 3 + (w[j] & 3)];
      }
    }
    // the node's slot, id and u are read again here rather than held across the
    // barriers: fewer VGPRs, more waves per SIMD
    const int pi = ld32(a.part_idx, (uint32_t)(n0 + l));
    if (pi >= 0) {
      st32(a.partial, (uint32_t)pi, acc);
      st32(a.partial2, (uint32_t)pi, accl);
    } else {
      const uint32_t gl = (uint32_t)ld32(a.tile_nodes, (uint32_t)(n0 + l));
      st32(a.out, gl, acc + node_adj<FK>(a, (int)gl, ld32(a.u, gl)));
      if (a.out2) st32(a.out2, gl, accl);
    }
  }
}

template <int FK>
__global__ __launch_bounds__(256) void vjp_fix_kernel(KArgs a) {
  const int fi = fix_wid() * blockDim.x + threadIdx.x;
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
  a.out[g] = acc + node_adj<FK>(a, g, a.u[g]);
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

}  // namespace fvm_k
