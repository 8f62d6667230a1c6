// Internal definitions shared by the libfvmhip.so translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fvm_hip.h"
#include "fvm_ctrl.hpp"
#include "fvm_device.hpp"
#include "fvm_host.h"

// Build-time tuning (fvm_build_info reports the values in use):
//   FVM_GEO_LEVEL       how much of each triangle record the tile kernels rebuild from
//                       the vertex coordinates instead of streaming (1, 2 or 3 = all;
//                       4 = all, with the CV-edge normals unnormalised and lengths 1)
//   FVM_SOLVER_THREADS  workgroup size of the template solver kernels
//   FVM_SPMV_ROWS       CSR rows per lane group and step in the solver SpMV
// Revision of the RHS kernels' memory layout, part of fvm_build_info so that measured
// traffic is matched to the build that produced it (2: fix-up records in fix order;
// 3: tile kernels' fluxes share LDS with the node coordinates; 4: also in the VJP;
// 5: node data re-read after the barriers; 6: fluxes stored (Q3, Q1, Q2, Q3) per
// triangle and node sums taken four incidences per round, in the VJP too; 7: the
// four flux copies stored as rows of TILE + 1 doubles, so the stores are free of LDS
// bank conflicts; 8: the tile kernel's loads branch-free, at 32-bit offsets; 9: the
// VJP tile kernel's too, with its incidence block staged in LDS; 10: fix nodes in the
// order of their first tile, so that a tile's partial sums are contiguous, and the
// fix-up workgroups XCD-contiguous).
#define FVM_RHS_REV 10
#ifndef FVM_GEO_LEVEL
#define FVM_GEO_LEVEL 4
#endif
#ifndef FVM_SOLVER_THREADS
#define FVM_SOLVER_THREADS 512
#endif
#ifndef FVM_SPMV_ROWS
#define FVM_SPMV_ROWS 2
#endif

namespace fvm {

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// fvm_mesh_create with the triangles of each tri_group kept contiguous (NULL: one group)
int mesh_create_grouped(const fvm_mesh_desc* d, const int32_t* tri_group, fvm_mesh** out);
// One forward-Euler step of fvm_eqs! in place, fused into the RHS kernels:
// u += dt du (dt per tile, and per node for the shared nodes), then the Dirichlet
// callback at t + dt.  Exclusive nodes are read and written by one tile only, and
// shared nodes only by the fix-up after every tile has read them.
int euler_step(fvm_ctx* c, double* d_u, const double* tile_dt, const double* dtv, double t, hipStream_t s);

#define FVM_HIP(call)                                        \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) return ::fvm::hip_fail(e_, #call); \
  } while (0)

#define FVM_CHECK(cond, msg)          \
  do {                                \
    if (!(cond)) {                    \
      ::fvm::set_error(msg);          \
      return FVM_ERR_ARG;             \
    }                                 \
  } while (0)

inline int xcd_grid(int n) { return kXCD * ((n + kXCD - 1) / kXCD); }
// workgroups of 256 threads of the fix-up kernels over n_fix nodes (fix_wid)
inline int fix_grid(int n_fix) { return xcd_grid((n_fix + 255) / 256); }

// Runs launch(d_flag) on stream st with a cleared device flag, waits, and returns
// FVM_ERR_ARG ("<what>: a queried edge or triangle is not in the mesh") when a
// kernel set it (the pointwise query kernels).
// d_bad is the mesh's own query flag (fvm_mesh::qflag): queries on one mesh must not
// overlap on different streams or host threads.
template <class Launch>
int run_flagged(hipStream_t st, const char* what, int32_t* d_bad, Launch launch) {
  int32_t bad = 0;
  hipError_t e = hipMemcpyAsync(d_bad, &bad, sizeof(int32_t), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    launch(d_bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, what);
  if (bad) {
    set_error(std::string(what) + ": a queried edge or triangle is not in the mesh");
    return FVM_ERR_ARG;
  }
  return FVM_OK;
}

// Sorted index of the triangle whose counter-clockwise vertex order holds the
// directed edge a -> b, or -1: scans the triangles incident to a (node_tri entries
// are (sorted triangle << 2) | slot of a).
__device__ __forceinline__ int32_t find_directed_edge(const int32_t* __restrict__ node_tri_start,
                                                      const int32_t* __restrict__ node_tri,
                                                      const int32_t* __restrict__ tri, int64_t T,
                                                      int32_t a, int32_t b) {
  const int32_t k1 = node_tri_start[a + 1];
  for (int32_t k = node_tri_start[a]; k < k1; ++k) {
    const int32_t w = node_tri[k];
    const int32_t t = w >> 2, s = w & 3;
    if (tri[(int64_t)(s == 2 ? 0 : s + 1) * T + t] == b) return t;
  }
  return -1;
}


// fvm_ode_solve's Tsit5 loop for a small ODE template in one workgroup
// (fvm_linalg.hip); FVM_ERR_STATE when the template does not qualify.
int template_tsit5_small(const fvm_template* t, const fvm_ode_opts* o, double* const k[7], double* utmp,
                         double* u, double* unew, double* d_u, double t0, double dt, int64_t it0, int64_t ks0,
                         const double A[7][6], const double C[7], const double Bt[7], fvm_ode_stats* st,
                         hipStream_t s);

// Classical Gram-Schmidt applied twice to w against xs[0..k) (k <= 64) plus |w|^2,
// intermediates on the device, one host synchronisation: h_host[0..k), h_host[k..2k)
// the two passes' coefficients, h_host[2k] = |w|^2 (fvm_vec.hip).
int cgs2_norm(int64_t n, int k, const double* const* xs, double* w, double* h_host, hipStream_t s);

}  // namespace fvm

// Device-resident mesh: geometry SoA in tile (sorted) triangle order plus the
// tile / fix-up structures from fvmh_tiles_build.
struct fvm_mesh {
  int64_t N = 0, T = 0, B = 0;
  int32_t n_tiles = 0, n_fix = 0, n_parts = 0, tile_tris = 0, max_nodes = 0;
  int64_t n_local = 0;
  // device arrays
  double *x = nullptr, *y = nullptr, *cv = nullptr, *props = nullptr;
  int32_t* tri = nullptr;  // [3][T] sorted order
  int32_t *tile_tri_start = nullptr, *tile_node_start = nullptr, *tile_nodes = nullptr;
  uint32_t* loc = nullptr;
  int32_t* inc_start = nullptr;
  uint16_t* inc = nullptr;
  int32_t *part_idx = nullptr, *fix_nodes = nullptr, *fix_part_start = nullptr;
  int32_t *fix_bedge_start = nullptr, *fix_bedge = nullptr;
  int32_t *fix_bvert_start = nullptr, *fix_bvert = nullptr;
  int32_t *node_tri_start = nullptr, *node_tri = nullptr;
  int32_t *bedges = nullptr, *badj = nullptr;  // bedges [2][B]
  int32_t* perm = nullptr;                      // sorted -> input triangle
  uint8_t* is_vertex = nullptr;
  double* lcv = nullptr;            // [n_local] control volume of each local-node entry
  int32_t* qflag = nullptr;         // device flag of the pointwise queries (run_flagged)
  bool convex = false;              // one boundary loop that turns left (or goes straight) at
                                    // every vertex: a point outside a boundary edge is outside
  int32_t n_bedge_items = 0;        // entries of fix_bedge (boundary edge endpoints)
  int32_t n_bvert_items = 0;        // entries of fix_bvert (boundary triangle vertices)
  // host copies needed to build contexts
  std::vector<int32_t> h_perm;
  std::vector<int32_t> h_node_tri_start;
  std::vector<int32_t> h_part_idx;    // [n_local]
  std::vector<int32_t> h_tile_nodes;  // [n_local]
  std::vector<int32_t> h_fix_part_start;   // [n_fix + 1]
  std::vector<int32_t> h_fix_bedge_start;  // [n_fix + 1]
};

namespace fvm {
struct JitKernels;
}

struct fvm_ctx {
  const fvm_mesh* m = nullptr;
  int32_t flux_kind = 0;
  const fvm::JitKernels* jit = nullptr;  // FVM_FLUX_JIT: the run-time compiled kernels
  double dpoly[4] = {0, 0, 0, 0};
  double spoly[4] = {0, 0, 0, 0};
  double st = 0;
  double* d_edge = nullptr;   // [3][T] sorted order
  double* d_bedge = nullptr;  // [2][B]
  double* log_d_own = nullptr;
  const double* log_d = nullptr;
  double* s_node = nullptr;
  int8_t* kind = nullptr;     // 0 none, 1 Dirichlet, 2 Dudt, 3 not a vertex
  int32_t* fidx = nullptr;
  int32_t* bneu = nullptr;    // [B]
  double* fn = nullptr;       // 6 * n_functions: c0, cu, ct, cx, cy, lam
  double* node_g = nullptr;
  double* bedge_g = nullptr;
  int32_t n_fn = 0;
  int32_t* dir_nodes = nullptr;
  int32_t n_dir = 0;
  double* partial = nullptr;   // [n_parts] (x2 for the VJP)
  int32_t* lpart = nullptr;    // [n_local] partial slot, or -1 - node code for exclusive entries
  double* bflux = nullptr;     // boundary terms per fix_bedge item (RHS/JVP) or 2 per fix_bvert item (VJP)
  int32_t* fix_meta = nullptr; // [4][n_fix] per fix node: node, first slot, first boundary item,
                               // slots << 20 | boundary items << 8 | code
  double* fix_cv = nullptr;    // [n_fix] control volume of each fix node
  fvm::CondTable cond() const {
    fvm::CondTable C;
    C.n = n_fn;
    C.c0 = fn;
    C.cu = fn + n_fn;
    C.ct = fn + 2 * n_fn;
    C.cx = fn + 3 * n_fn;
    C.cy = fn + 4 * n_fn;
    C.lam = fn + 5 * n_fn;
    C.node_g = node_g;
    C.bedge_g = bedge_g;
    return C;
  }
};
