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
#include "fvm_jit.hpp"
#include "fvm_rhs_kernels.hpp"

using namespace fvm;

using namespace fvm_k;

namespace {

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
    hipLaunchKernelGGL((fix_kernel<FK, MODE>), dim3((unsigned)fix_grid(a.n_fix)), dim3(256), 0,
                       s, a);
}

template <int FK, int MODE>
int launch_fk(const KArgs& a, int tile_tris, hipStream_t s) {
  launch_tiles<FK, MODE>(a, tile_tris, s);
  launch_fix<FK, MODE>(a, s);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

// FVM_FLUX_JIT: the same tile / fix-up pair from the run-time compiled module
int launch_jit(const fvm_ctx* c, const KArgs& a, int mode, hipStream_t s) {
  const JitKernels* j = c->jit;
  FVM_CHECK(mode == M_RHS || mode == M_JVP, "closures compiled at run time: RHS and JVP only");
  int rc = jit_launch(mode == M_RHS ? j->tile_rhs : j->tile_jvp, tile_grid(a, j->block).x, j->block, &a, s);
  if (rc == FVM_OK && a.n_fix > 0)
    rc = jit_launch(mode == M_RHS ? j->fix_rhs : j->fix_jvp, (unsigned)fix_grid(a.n_fix), 256, &a, s);
  return rc;
}

template <int MODE>
int launch(const fvm_ctx* c, const KArgs& a, hipStream_t s) {
  switch (c->flux_kind) {
    case FVM_FLUX_JIT: return launch_jit(c, a, MODE, s);
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
    hipLaunchKernelGGL((vjp_fix_kernel<FK>), dim3((unsigned)fix_grid(a.n_fix)), dim3(256), 0, s,
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
  FVM_CHECK(ph->flux_kind >= FVM_FLUX_D_POLY_U && ph->flux_kind <= FVM_FLUX_JIT,
            "fvm_ctx_create: unknown flux_kind");
  FVM_CHECK(ph->flux_kind != FVM_FLUX_JIT || ph->jit_source,
            "fvm_ctx_create: FVM_FLUX_JIT needs jit_source");
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
  if (ph->flux_kind == FVM_FLUX_JIT) {
    int jrc = jit_kernels(ph->jit_source, m->tile_tris, &c->jit);
    if (jrc != FVM_OK) return jrc;
  }
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
    case FVM_FLUX_JIT: rc = jit_launch(c->jit->tile_rhs, tile_grid(a, c->jit->block).x, c->jit->block, &a, st); break;
    case FVM_FLUX_D_POLY_U: launch_tiles<FVM_FLUX_D_POLY_U, M_RHS>(a, tt, st); break;
    case FVM_FLUX_D_NODAL: launch_tiles<FVM_FLUX_D_NODAL, M_RHS>(a, tt, st); break;
    case FVM_FLUX_LINEAR: launch_tiles<FVM_FLUX_LINEAR, M_RHS>(a, tt, st); break;
    case FVM_FLUX_D_POWER: launch_tiles<FVM_FLUX_D_POWER, M_RHS>(a, tt, st); break;
    default: launch_tiles<FVM_FLUX_D_EDGE, M_RHS>(a, tt, st);
  }
  FVM_HIP(hipEventRecord(ev[1], st));
  switch (c->flux_kind) {
    case FVM_FLUX_JIT:
      if (rc == FVM_OK && a.n_fix > 0)
        rc = jit_launch(c->jit->fix_rhs, (unsigned)fix_grid(a.n_fix), 256, &a, st);
      break;
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
    case FVM_FLUX_JIT: {
      const JitKernels* j = c->jit;
      int rc = jit_launch(j->vjp_tile, tile_grid(a, j->block).x, j->block, &a, st);
      if (rc == FVM_OK && a.n_fix > 0) rc = jit_launch(j->vjp_fix, (unsigned)fix_grid(a.n_fix), 256, &a, st);
      return rc;
    }
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
