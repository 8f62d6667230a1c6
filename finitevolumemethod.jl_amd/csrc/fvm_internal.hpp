// Internal definitions shared by the libfvmhip.so translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fvm_hip.h"
#include "fvm_host.h"

// Build-time tuning (fvm_build_info reports the values in use):
//   FVM_GEO_LEVEL       how much of each triangle record the tile kernels rebuild from
//                       the vertex coordinates instead of streaming (1, 2 or 3 = all;
//                       4 = all, with the CV-edge normals unnormalised and lengths 1)
//   FVM_SOLVER_THREADS  workgroup size of the template solver kernels
//   FVM_SPMV_ROWS       CSR rows per lane group and step in the solver SpMV
// Revision of the RHS kernels' memory layout, part of fvm_build_info so that measured
// traffic is matched to the build that produced it (2: fix-up records in fix order;
// 3: tile kernels' fluxes share LDS with the node coordinates; 4: also in the VJP).
#define FVM_RHS_REV 5
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

// Number of XCDs of an MI355X; consecutive workgroups are dispatched round-robin
// over them, each XCD with its own L2.
constexpr int kXCD = 8;

// Maps workgroup id to a work index so that each XCD processes one contiguous
// range of work items (neighbouring tiles share nodes, hence L2 lines).  The
// grid must hold kXCD * ceil(n / kXCD) workgroups; ids >= n are idle.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  int per = nblocks / kXCD;
  return (bid % kXCD) * per + bid / kXCD;
}
inline int xcd_grid(int n) { return kXCD * ((n + kXCD - 1) / kXCD); }

// Per-triangle record in the device SoA: field f of sorted triangle s lives at
// props[f * T + s].  Fields follow TriangleProperties (geometry.jl:21-26).
enum : int {
  P_S1 = 0,   // s1..s9 at 0..8
  P_MX = 9,   // CV-edge midpoints m1x m1y m2x m2y m3x m3y at 9..14
  P_NX = 15,  // normals n1x n1y n2x n2y n3x n3y at 15..20
  P_L = 21,   // lengths l1 l2 l3 at 21..23
  P_COUNT = 24
};

// a / b correctly rounded, given y = RN(1 / b): Markstein's correction step
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, 2nd ed., 2018, sec. 4.9): q0 = RN(a y), the remainder
// r = a - b q0 exactly by one FMA, then RN(q0 + r y).  Equal to a / b for finite
// operands away from overflow and underflow (checked against IEEE division on 2e8
// random pairs and on every quotient of the geometry of 4M-node meshes).
__device__ __forceinline__ double div_rn(double a, double b, double y) {
  const double q0 = a * y;
  const double r = __builtin_fma(-b, q0, a);
  return __builtin_fma(r, y, q0);
}

// The triangle record (TriangleProperties, geometry.jl:99-169) of vertices p, q, r,
// with the geometry kernel's operations in the reference's order, so every value
// equals the stored one bit for bit up to LEVEL 3.  LEVEL >= 2: s1..s6; always
// s7..s9 and the CV-edge midpoints; LEVEL 3: normals and lengths; LEVEL 4: the
// quotients by Delta and by 3 from one reciprocal (div_rn, still bit-identical),
// and the CV-edge normals unnormalised with lengths 1 ((q . n) l = q . (e_y, -e_x)).
template <int LEVEL>
__device__ __forceinline__ void rebuild_record(double px, double py, double qx, double qy,
                                               double rx, double ry, double* P) {
  constexpr double kThird = 1.0 / 3;
  const double sx = px + qx + rx, sy = py + qy + ry;
  const double cx = LEVEL >= 4 ? div_rn(sx, 3.0, kThird) : sx / 3;
  const double cy = LEVEL >= 4 ? div_rn(sy, 3.0, kThird) : sy / 3;
  const double m1x = (px + qx) / 2, m1y = (py + qy) / 2;
  const double m2x = (qx + rx) / 2, m2y = (qy + ry) / 2;
  const double m3x = (rx + px) / 2, m3y = (ry + py) / 2;
  const double D = qx * ry - qy * rx - px * ry + rx * py + px * qy - qx * py;
  const double yD = LEVEL >= 4 ? 1.0 / D : 0.0;
  auto quot = [&](double a) { return LEVEL >= 4 ? div_rn(a, D, yD) : a / D; };
  if (LEVEL >= 2) {
    P[0] = quot(qy - ry);
    P[1] = quot(ry - py);
    P[2] = quot(py - qy);
    P[3] = quot(rx - qx);
    P[4] = quot(px - rx);
    P[5] = quot(qx - px);
  }
  P[6] = quot(qx * ry - rx * qy);
  P[7] = quot(rx * py - px * ry);
  P[8] = quot(px * qy - qx * py);
  P[9] = (m1x + cx) / 2;
  P[10] = (m1y + cy) / 2;
  P[11] = (m2x + cx) / 2;
  P[12] = (m2y + cy) / 2;
  P[13] = (m3x + cx) / 2;
  P[14] = (m3y + cy) / 2;
  if (LEVEL >= 4) {
    // (q . n) l with n = (e_y, -e_x) / l is q . (e_y, -e_x): keep the normal
    // unnormalised and the length 1, saving six divisions and three square roots
    const double e1x = cx - m1x, e1y = cy - m1y;
    const double e2x = cx - m2x, e2y = cy - m2y;
    const double e3x = cx - m3x, e3y = cy - m3y;
    P[15] = e1y;
    P[16] = -e1x;
    P[17] = e2y;
    P[18] = -e2x;
    P[19] = e3y;
    P[20] = -e3x;
    P[21] = 1.0;
    P[22] = 1.0;
    P[23] = 1.0;
  } else if (LEVEL >= 3) {
    const double e1x = cx - m1x, e1y = cy - m1y;
    const double e2x = cx - m2x, e2y = cy - m2y;
    const double e3x = cx - m3x, e3y = cy - m3y;
    const double l1 = sqrt(e1x * e1x + e1y * e1y);
    const double l2 = sqrt(e2x * e2x + e2y * e2y);
    const double l3 = sqrt(e3x * e3x + e3y * e3y);
    P[15] = e1y / l1;
    P[16] = -e1x / l1;
    P[17] = e2y / l2;
    P[18] = -e2x / l2;
    P[19] = e3y / l3;
    P[20] = -e3x / l3;
    P[21] = l1;
    P[22] = l2;
    P[23] = l3;
  }
}

// Condition functions a(x,y,t,u) = c0 + cu u + ct t + cx x + cy y + g, with g
// multiplied by exp(lam t) where lam (may be null) is nonzero.
struct CondTable {
  int32_t n;
  const double* c0;
  const double* cu;
  const double* ct;
  const double* cx;
  const double* cy;
  const double* lam;      // may be null
  const double* node_g;   // may be null
  const double* bedge_g;  // may be null
};

__device__ __forceinline__ double eval_cond(const CondTable& C, int f, double x, double y,
                                            double t, double u, double g) {
  double v = C.c0[f] + C.cu[f] * u;
  v = v + C.ct[f] * t;
  v = v + C.cx[f] * x;
  v = v + C.cy[f] * y;
  if (C.lam && C.lam[f] != 0.0) return v + g * exp(C.lam[f] * t);
  return v + g;
}

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

struct fvm_ctx {
  const fvm_mesh* m = nullptr;
  int32_t flux_kind = 0;
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
