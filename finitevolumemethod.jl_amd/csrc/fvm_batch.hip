// Ensembles: many FVMProblems on meshes of uneven size advanced together.
//
// The reference solves an ensemble by looping (EnsembleProblem over solve,
// solve.jl:222-256, one fvm_eqs! call per problem per stage).  Here the members'
// meshes are packed into one device mesh (disjoint node ranges; each member's
// triangles kept contiguous and in Morton order, so no tile mixes two members)
// and one context that carries every member's closures, so one launch of the
// tile and fix-up kernels evaluates the right-hand side of every member.  An
// explicit Euler step (u += dt_p du, then the Dirichlet callback of solve.jl:
// 133-165 at the member's new time) is fused into those two kernels (each tile
// updates the nodes only it touches, the fix-up the shared ones), and runs of
// steps are captured once in a HIP graph and replayed.  Members keep their own time step
// dt_p; closures that depend on t then require a common dt.
//
// An autotuner chooses the tile shape (triangles per tile, which fixes the
// workgroup size and triangles per thread of the tile kernel) by timing the
// packed RHS for each candidate.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fvm_internal.hpp"

using namespace fvm;

struct fvm_batch {
  int32_t P = 0;
  fvm_mesh* mesh = nullptr;
  fvm_ctx* ctx = nullptr;
  std::vector<int64_t> node_off;  // [P+1]
  std::vector<double> dt;         // [P]
  double* dtv = nullptr;          // [N] time step of each node's member
  double* tile_dt = nullptr;      // [n_tiles] time step of each tile's member
  std::vector<int64_t> tri_off;   // [P+1]
  double* du = nullptr;           // [N] scratch
  int8_t* code = nullptr;         // [N] node codes (0 free, 1 Dirichlet, 2 Dudt, 3 not a vertex)
  int32_t* fidx = nullptr;        // [N]
  double* fn = nullptr;           // condition table (6 * n_fn)
  double* node_g = nullptr;
  int32_t n_fn = 0;
  bool time_dependent = false;
  int32_t tile_tris = 0;
  int32_t graph_steps = 0;
  bool fused = true;               // Euler update fused into the RHS kernels
  hipStream_t cap = nullptr;
  hipGraphExec_t exec = nullptr;
  const double* exec_u = nullptr;  // the state buffer the graph was captured for
  std::vector<float> tune_ms;      // autotuner: RHS time per candidate (ms), 0 if not built
  std::vector<int32_t> tune_tiles;
};

namespace {

// u_g += dt_g du_g at free and Dudt nodes; Dirichlet nodes take a(x, y, t_new, u)
// (update_dirichlet_nodes!, dirichlet.jl:78-86), applied after the step as the
// reference's callback does.
__global__ __launch_bounds__(256) void euler_update(int64_t n, double* __restrict__ u,
                                                    const double* __restrict__ du,
                                                    const double* __restrict__ dtv,
                                                    const int8_t* __restrict__ code,
                                                    const int32_t* __restrict__ fidx, CondTable C,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ y, double t_new) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int c = code[g];
  if (c == 0 || c == 2) {
    u[g] = u[g] + dtv[g] * du[g];
  } else if (c == 1) {
    u[g] = eval_cond(C, fidx[g], x[g], y[g], t_new, u[g], C.node_g ? C.node_g[g] : 0.0);
  }
}

template <class T>
int dcopy(T** dst, const T* src, int64_t n) {
  *dst = nullptr;
  FVM_HIP(hipMalloc((void**)dst, sizeof(T) * (size_t)(n > 0 ? n : 1)));
  if (n > 0 && src) FVM_HIP(hipMemcpy(*dst, src, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
  return FVM_OK;
}

void free_batch(fvm_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->ctx) fvm_ctx_destroy(b->ctx);
  if (b->mesh) fvm_mesh_destroy(b->mesh);
  void* p[] = {b->dtv, b->tile_dt, b->du, b->code, b->fidx, b->fn, b->node_g};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete b;
}

// The members packed into one mesh description and one physics description.
struct Packed {
  std::vector<double> x, y;
  std::vector<int32_t> tri, bedges, group;
  std::vector<double> d_edge, d_bedge, log_d, s_node, node_g, bedge_g;
  std::vector<int8_t> node_kind;
  std::vector<int32_t> node_fidx, bneu;
  std::vector<double> c0, cu, ct, cx, cy, lam;
  bool any_s_node = false, any_node_g = false, any_bedge_g = false;
  fvm_mesh_desc md{};
  fvm_physics_desc pd{};
};

int pack(int32_t P, const fvm_mesh_desc* meshes, const fvm_physics_desc* ph, fvm_batch* b, Packed& k) {
  const int fk = ph[0].flux_kind;
  FVM_CHECK(fk != FVM_FLUX_LINEAR, "fvm_batch_create: linear flux forms (FVM_FLUX_LINEAR) are not batched");
  int64_t N = 0, T = 0, B = 0, F = 0;
  b->node_off.assign(P + 1, 0);
  for (int32_t p = 0; p < P; ++p) {
    const fvm_mesh_desc& m = meshes[p];
    const fvm_physics_desc& q = ph[p];
    FVM_CHECK(m.n_points > 0 && m.n_triangles > 0 && m.n_boundary_edges >= 0 && m.x && m.y &&
                  m.triangles && (m.n_boundary_edges == 0 || m.boundary_edges),
              "fvm_batch_create: invalid member mesh");
    FVM_CHECK(q.flux_kind == fk, "fvm_batch_create: every member needs the same flux_kind");
    FVM_CHECK(q.node_kind && q.node_fidx && (m.n_boundary_edges == 0 || q.bedge_neumann_fidx),
              "fvm_batch_create: member conditions missing");
    FVM_CHECK(q.n_functions >= 0 && (q.n_functions == 0 || (q.fn_c0 && q.fn_cu && q.fn_ct)),
              "fvm_batch_create: member condition coefficients missing");
    for (int c = 0; c < 4; ++c) {
      FVM_CHECK(q.spoly[c] == ph[0].spoly[c], "fvm_batch_create: members must share the source polynomial");
      if (fk == FVM_FLUX_D_POLY_U || fk == FVM_FLUX_D_POWER)
        FVM_CHECK(q.dpoly[c] == ph[0].dpoly[c], "fvm_batch_create: members must share D(u)");
    }
    FVM_CHECK(q.st == ph[0].st, "fvm_batch_create: members must share the source's t coefficient");
    if (fk == FVM_FLUX_D_EDGE)
      FVM_CHECK(q.d_edge && (m.n_boundary_edges == 0 || q.d_bedge), "fvm_batch_create: member d_edge missing");
    if (fk == FVM_FLUX_D_NODAL) FVM_CHECK(q.log_d, "fvm_batch_create: member log_d missing");
    N += m.n_points;
    b->tri_off.push_back(T);
    T += m.n_triangles;
    B += m.n_boundary_edges;
    F += q.n_functions;
    b->node_off[p + 1] = N;
  }
  b->tri_off.push_back(T);
  FVM_CHECK(N < INT32_MAX && T < ((int64_t)1 << 29) && F < INT32_MAX, "fvm_batch_create: ensemble too large");
  k.x.reserve(N);
  k.y.reserve(N);
  k.tri.reserve(3 * T);
  k.bedges.reserve(2 * B);
  k.group.reserve(T);
  k.node_kind.reserve(N);
  k.node_fidx.reserve(N);
  k.bneu.reserve(B);
  int64_t n0 = 0, f0 = 0;
  for (int32_t p = 0; p < P; ++p) {
    const fvm_mesh_desc& m = meshes[p];
    const fvm_physics_desc& q = ph[p];
    const int64_t n = m.n_points, t = m.n_triangles, be = m.n_boundary_edges, nf = q.n_functions;
    k.x.insert(k.x.end(), m.x, m.x + n);
    k.y.insert(k.y.end(), m.y, m.y + n);
    for (int64_t i = 0; i < 3 * t; ++i) {
      FVM_CHECK(m.triangles[i] >= 0 && m.triangles[i] < n, "fvm_batch_create: triangle vertex out of range");
      k.tri.push_back((int32_t)(m.triangles[i] + n0));
    }
    for (int64_t i = 0; i < 2 * be; ++i) {
      FVM_CHECK(m.boundary_edges[i] >= 0 && m.boundary_edges[i] < n,
                "fvm_batch_create: boundary edge vertex out of range");
      k.bedges.push_back((int32_t)(m.boundary_edges[i] + n0));
    }
    k.group.insert(k.group.end(), (size_t)t, p);
    for (int64_t i = 0; i < n; ++i) {
      k.node_kind.push_back(q.node_kind[i]);
      const int32_t f = q.node_fidx[i];
      FVM_CHECK(q.node_kind[i] == 0 || (f >= 0 && f < nf), "fvm_batch_create: member node_fidx out of range");
      k.node_fidx.push_back(q.node_kind[i] != 0 ? (int32_t)(f + f0) : f);
    }
    for (int64_t e = 0; e < be; ++e) {
      const int32_t f = q.bedge_neumann_fidx[e];
      FVM_CHECK(f == -1 || (f >= 0 && f < nf), "fvm_batch_create: member bedge_neumann_fidx out of range");
      k.bneu.push_back(f >= 0 ? (int32_t)(f + f0) : -1);
    }
    for (int64_t f = 0; f < nf; ++f) {
      k.c0.push_back(q.fn_c0[f]);
      k.cu.push_back(q.fn_cu[f]);
      k.ct.push_back(q.fn_ct[f]);
      k.cx.push_back(q.fn_cx ? q.fn_cx[f] : 0.0);
      k.cy.push_back(q.fn_cy ? q.fn_cy[f] : 0.0);
      k.lam.push_back(q.fn_lam ? q.fn_lam[f] : 0.0);
      if (q.fn_ct[f] != 0.0 || k.lam.back() != 0.0) b->time_dependent = true;
    }
    if (fk == FVM_FLUX_D_EDGE) {
      k.d_edge.insert(k.d_edge.end(), q.d_edge, q.d_edge + 3 * t);
      if (be) k.d_bedge.insert(k.d_bedge.end(), q.d_bedge, q.d_bedge + 2 * be);
    }
    if (fk == FVM_FLUX_D_NODAL) k.log_d.insert(k.log_d.end(), q.log_d, q.log_d + n);
    k.any_s_node |= q.s_node != nullptr;
    k.any_node_g |= q.node_g != nullptr;
    k.any_bedge_g |= q.bedge_g != nullptr;
    n0 += n;
    f0 += nf;
  }
  if (ph[0].st != 0.0) b->time_dependent = true;
  // optional per-node / per-edge arrays: zeros where a member gave none
  if (k.any_s_node || k.any_node_g || k.any_bedge_g) {
    if (k.any_s_node) k.s_node.assign(N, 0.0);
    if (k.any_node_g) k.node_g.assign(N, 0.0);
    if (k.any_bedge_g) k.bedge_g.assign(2 * B, 0.0);
    int64_t bo = 0;
    for (int32_t p = 0; p < P; ++p) {
      const int64_t no = b->node_off[p], n = meshes[p].n_points, be = meshes[p].n_boundary_edges;
      if (ph[p].s_node) std::copy(ph[p].s_node, ph[p].s_node + n, k.s_node.begin() + no);
      if (ph[p].node_g) std::copy(ph[p].node_g, ph[p].node_g + n, k.node_g.begin() + no);
      if (ph[p].bedge_g) std::copy(ph[p].bedge_g, ph[p].bedge_g + 2 * be, k.bedge_g.begin() + 2 * bo);
      bo += be;
    }
  }
  k.md.n_points = N;
  k.md.n_triangles = T;
  k.md.n_boundary_edges = B;
  k.md.x = k.x.data();
  k.md.y = k.y.data();
  k.md.triangles = k.tri.data();
  k.md.boundary_edges = k.bedges.data();
  k.pd.flux_kind = fk;
  for (int c = 0; c < 4; ++c) {
    k.pd.dpoly[c] = ph[0].dpoly[c];
    k.pd.spoly[c] = ph[0].spoly[c];
  }
  k.pd.st = ph[0].st;
  k.pd.d_edge = k.d_edge.empty() ? nullptr : k.d_edge.data();
  k.pd.d_bedge = k.d_bedge.empty() ? nullptr : k.d_bedge.data();
  k.pd.log_d = k.log_d.empty() ? nullptr : k.log_d.data();
  k.pd.s_node = k.any_s_node ? k.s_node.data() : nullptr;
  k.pd.node_kind = k.node_kind.data();
  k.pd.node_fidx = k.node_fidx.data();
  k.pd.bedge_neumann_fidx = k.bneu.data();
  k.pd.n_functions = (int32_t)F;
  k.pd.fn_c0 = k.c0.data();
  k.pd.fn_cu = k.cu.data();
  k.pd.fn_ct = k.ct.data();
  k.pd.fn_cx = k.cx.data();
  k.pd.fn_cy = k.cy.data();
  k.pd.fn_lam = k.lam.data();
  k.pd.node_g = k.any_node_g ? k.node_g.data() : nullptr;
  k.pd.bedge_g = k.any_bedge_g ? k.bedge_g.data() : nullptr;
  return FVM_OK;
}

// Mesh + context of the packed ensemble with a given tile size.
int build_packed(Packed& k, int tile_tris, fvm_mesh** mesh, fvm_ctx** ctx) {
  *mesh = nullptr;
  *ctx = nullptr;
  k.md.tile_triangles = tile_tris;
  k.md.tile_max_nodes = 0;
  int rc = mesh_create_grouped(&k.md, k.group.data(), mesh);
  if (rc) return rc;
  rc = fvm_ctx_create(*mesh, &k.pd, ctx);
  if (rc) {
    fvm_mesh_destroy(*mesh);
    *mesh = nullptr;
  }
  return rc;
}

// Mean time of one packed RHS evaluation (ms) over reps launches.
int time_rhs(fvm_ctx* c, const double* u, double* du, int reps, float* ms) {
  hipEvent_t e0, e1;
  FVM_HIP(hipEventCreate(&e0));
  FVM_HIP(hipEventCreate(&e1));
  int rc = FVM_OK;
  for (int r = 0; r < 2 && !rc; ++r) rc = fvm_rhs(c, u, du, 0.0, nullptr);
  if (!rc) {
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && !rc; ++r) rc = fvm_rhs(c, u, du, 0.0, nullptr);
    (void)hipEventRecord(e1, nullptr);
    if (!rc) {
      hipError_t e = hipEventSynchronize(e1);
      if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
      if (e != hipSuccess) rc = hip_fail(e, "fvm_batch_create: autotune timing");
      *ms /= reps;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int launch_step(fvm_batch* b, double* u, double t, hipStream_t s) {
  if (b->fused) return euler_step(b->ctx, u, b->tile_dt, b->dtv, t, s);
  int rc = fvm_rhs(b->ctx, u, b->du, t, s);
  if (rc) return rc;
  const int64_t N = b->mesh->N;
  CondTable C;
  C.n = b->n_fn;
  C.c0 = b->fn;
  C.cu = b->fn + b->n_fn;
  C.ct = b->fn + 2 * b->n_fn;
  C.cx = b->fn + 3 * b->n_fn;
  C.cy = b->fn + 4 * b->n_fn;
  C.lam = b->fn + 5 * b->n_fn;
  C.node_g = b->node_g;
  C.bedge_g = nullptr;
  const double t_new = t + (b->dt.empty() ? 0.0 : b->dt[0]);
  hipLaunchKernelGGL(euler_update, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, u, b->du, b->dtv,
                     b->code, b->fidx, C, b->mesh->x, b->mesh->y, t_new);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

int capture(fvm_batch* b, double* u) {
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  FVM_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  int rc = FVM_OK;
  for (int k = 0; k < b->graph_steps && !rc; ++k) rc = launch_step(b, u, 0.0, b->cap);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->cap, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "fvm_batch_euler: graph capture");
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->exec = nullptr;
    return hip_fail(e, "fvm_batch_euler: graph instantiation");
  }
  b->exec_u = u;
  return FVM_OK;
}

}  // namespace

extern "C" {

int fvm_batch_create(int32_t n_problems, const fvm_mesh_desc* meshes, const fvm_physics_desc* physics,
                     const double* dt, const fvm_batch_opts* opts, fvm_batch** out) {
  FVM_CHECK(out && meshes && physics && n_problems > 0, "fvm_batch_create: invalid argument");
  *out = nullptr;
  fvm_batch* b = new fvm_batch();
  b->P = n_problems;
  Packed k;
  int rc = pack(n_problems, meshes, physics, b, k);
  if (rc) {
    free_batch(b);
    return rc;
  }
  if (dt) {
    b->dt.assign(dt, dt + n_problems);
    for (double v : b->dt)
      if (!(v > 0.0 && std::isfinite(v))) {
        free_batch(b);
        set_error("fvm_batch_create: time steps must be positive and finite");
        return FVM_ERR_ARG;
      }
    for (double v : b->dt)
      if (v != b->dt[0] && b->time_dependent) {
        free_batch(b);
        set_error("fvm_batch_create: closures that depend on t need a common time step");
        return FVM_ERR_ARG;
      }
  }
  const int32_t tile_opt = opts ? opts->tile_triangles : 0;
  const int32_t autotune = opts ? opts->autotune : 0;
  b->graph_steps = opts ? std::max(0, opts->graph_steps) : 0;
  const int64_t N = k.md.n_points;
  std::vector<int32_t> cands;
  if (autotune && tile_opt == 0) cands = {256, 512, 1024};
  else cands = {tile_opt > 0 ? tile_opt : 512};
  double* tu = nullptr;
  double* tdu = nullptr;
  if (cands.size() > 1) {
    if ((rc = dcopy(&tu, (const double*)nullptr, N)) || (rc = dcopy(&tdu, (const double*)nullptr, N))) {
      if (tu) (void)hipFree(tu);
      free_batch(b);
      return rc;
    }
    std::vector<double> h(N);
    for (int64_t i = 0; i < N; ++i) h[i] = 0.5 + 0.25 * std::sin(0.001 * (double)i);
    (void)hipMemcpy(tu, h.data(), sizeof(double) * N, hipMemcpyHostToDevice);
  }
  float best = 0;
  for (int32_t tt : cands) {
    fvm_mesh* m = nullptr;
    fvm_ctx* c = nullptr;
    rc = build_packed(k, tt, &m, &c);
    if (rc) break;
    float ms = 0;
    if (cands.size() > 1) rc = time_rhs(c, tu, tdu, 10, &ms);
    if (rc) {
      fvm_ctx_destroy(c);
      fvm_mesh_destroy(m);
      break;
    }
    b->tune_tiles.push_back(tt);
    b->tune_ms.push_back(ms);
    if (!b->mesh || ms < best) {
      if (b->ctx) fvm_ctx_destroy(b->ctx);
      if (b->mesh) fvm_mesh_destroy(b->mesh);
      b->mesh = m;
      b->ctx = c;
      b->tile_tris = tt;
      best = ms;
    } else {
      fvm_ctx_destroy(c);
      fvm_mesh_destroy(m);
    }
  }
  if (tu) (void)hipFree(tu);
  if (tdu) (void)hipFree(tdu);
  if (rc) {
    free_batch(b);
    return rc;
  }
  // per-node step and the update kernel's condition data
  std::vector<double> dtv(N, 0.0);
  if (!b->dt.empty())
    for (int32_t p = 0; p < n_problems; ++p)
      std::fill(dtv.begin() + b->node_off[p], dtv.begin() + b->node_off[p + 1], b->dt[p]);
  // each tile lies inside one member (tiles are cut at member boundaries)
  std::vector<double> tdt(b->mesh->n_tiles, 0.0);
  if (!b->dt.empty()) {
    std::vector<int32_t> tts(b->mesh->n_tiles + 1);
    FVM_HIP(hipMemcpy(tts.data(), b->mesh->tile_tri_start, sizeof(int32_t) * tts.size(), hipMemcpyDeviceToHost));
    for (int32_t tl = 0; tl < b->mesh->n_tiles; ++tl) {
      const int64_t tin = b->mesh->h_perm[tts[tl]];
      const int64_t p = std::upper_bound(b->tri_off.begin(), b->tri_off.end(), tin) - b->tri_off.begin() - 1;
      tdt[tl] = b->dt[p];
    }
  }
  b->fused = !(opts && opts->unfused == 1);
  std::vector<int8_t> code(N);
  FVM_HIP(hipMemcpy(code.data(), b->ctx->kind, sizeof(int8_t) * N, hipMemcpyDeviceToHost));
  b->n_fn = k.pd.n_functions;
  std::vector<double> fn(6 * (size_t)std::max(b->n_fn, 1), 0.0);
  for (int f = 0; f < b->n_fn; ++f) {
    fn[f] = k.c0[f];
    fn[b->n_fn + f] = k.cu[f];
    fn[2 * b->n_fn + f] = k.ct[f];
    fn[3 * b->n_fn + f] = k.cx[f];
    fn[4 * b->n_fn + f] = k.cy[f];
    fn[5 * b->n_fn + f] = k.lam[f];
  }
  if ((rc = dcopy(&b->dtv, dtv.data(), N)) || (rc = dcopy(&b->du, (const double*)nullptr, N)) ||
      (rc = dcopy(&b->tile_dt, tdt.data(), (int64_t)tdt.size())) ||
      (rc = dcopy(&b->code, code.data(), N)) || (rc = dcopy(&b->fidx, k.node_fidx.data(), N)) ||
      (rc = dcopy(&b->fn, fn.data(), (int64_t)fn.size())) ||
      (k.any_node_g && (rc = dcopy(&b->node_g, k.node_g.data(), N)))) {
    free_batch(b);
    return rc;
  }
  hipError_t e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    free_batch(b);
    return hip_fail(e, "fvm_batch_create: stream");
  }
  *out = b;
  return FVM_OK;
}

int fvm_batch_destroy(fvm_batch* b) {
  free_batch(b);
  return FVM_OK;
}

int fvm_batch_get_info(const fvm_batch* b, fvm_batch_info* info, int64_t* node_offsets) {
  FVM_CHECK(b && info, "fvm_batch_get_info: null argument");
  info->n_problems = b->P;
  info->n_points = b->mesh->N;
  info->n_triangles = b->mesh->T;
  info->n_boundary_edges = b->mesh->B;
  info->n_tiles = b->mesh->n_tiles;
  info->n_fix_nodes = b->mesh->n_fix;
  info->tile_triangles = b->tile_tris;
  info->n_candidates = (int32_t)b->tune_tiles.size();
  for (int k = 0; k < 4; ++k) {
    info->candidate_tiles[k] = k < (int)b->tune_tiles.size() ? b->tune_tiles[k] : 0;
    info->candidate_ms[k] = k < (int)b->tune_ms.size() ? b->tune_ms[k] : 0.0f;
  }
  if (node_offsets) std::copy(b->node_off.begin(), b->node_off.end(), node_offsets);
  return FVM_OK;
}

int fvm_batch_mesh(const fvm_batch* b, const fvm_mesh** mesh, const fvm_ctx** ctx) {
  FVM_CHECK(b, "fvm_batch_mesh: null batch");
  if (mesh) *mesh = b->mesh;
  if (ctx) *ctx = b->ctx;
  return FVM_OK;
}

int fvm_batch_rhs(fvm_batch* b, const double* d_u, double* d_du, double t, fvm_stream s) {
  FVM_CHECK(b && d_u && d_du, "fvm_batch_rhs: null argument");
  return fvm_rhs(b->ctx, d_u, d_du, t, s);
}

int fvm_batch_euler(fvm_batch* b, double* d_u, int32_t n_steps, double t0, fvm_stream s) {
  FVM_CHECK(b && d_u, "fvm_batch_euler: null argument");
  FVM_CHECK(!b->dt.empty(), "fvm_batch_euler: the batch was created without time steps");
  FVM_CHECK(n_steps >= 0, "fvm_batch_euler: negative step count");
  hipStream_t st = (hipStream_t)s;
  int32_t done = 0;
  if (b->graph_steps > 0 && !b->time_dependent && n_steps >= b->graph_steps) {
    if (b->exec == nullptr || b->exec_u != d_u) {
      // capture only records the launches (on a stream of its own); the graph then
      // runs on s after the caller's earlier work
      int rc = capture(b, d_u);
      if (rc) return rc;
    }
    for (; done + b->graph_steps <= n_steps; done += b->graph_steps) FVM_HIP(hipGraphLaunch(b->exec, st));
  }
  const double dt = b->dt[0];
  for (; done < n_steps; ++done) {
    int rc = launch_step(b, d_u, t0 + done * dt, st);
    if (rc) return rc;
  }
  return FVM_OK;
}

}  // extern "C"
