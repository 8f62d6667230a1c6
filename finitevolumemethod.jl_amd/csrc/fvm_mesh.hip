// Device-resident FVMGeometry (reference src/geometry.jl:99-169).
//
// fvm_mesh_create builds the triangle tiles on the host (fvm_host.cpp), uploads
// them, and computes every TriangleProperties field and the control-volume
// areas on the GPU.  The arithmetic follows geometry.jl token for token, with
// FMA contraction off (the library is built with -ffp-contract=off): the
// shape-function determinant cancels badly on small triangles far from the
// origin, so any other evaluation order changes s1..s9 well above 1e-12.
#include <cstring>
#include <vector>

#include "fvm_internal.hpp"

using namespace fvm;

namespace {

template <class T>
int upload(T** dst, const T* src, int64_t n) {
  *dst = nullptr;
  size_t bytes = sizeof(T) * (size_t)(n > 0 ? n : 1);
  FVM_HIP(hipMalloc((void**)dst, bytes));
  if (n > 0 && src) FVM_HIP(hipMemcpy(*dst, src, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
  return FVM_OK;
}

// One thread per (sorted) triangle: TriangleProperties and the three sub-control-
// volume areas S1..S3 (geometry.jl:107-166).
__global__ __launch_bounds__(256) void geometry_kernel(int64_t T, const int32_t* __restrict__ tri,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ Y,
                                                       double* __restrict__ props,
                                                       double* __restrict__ sub) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= T) return;
  const int32_t i = tri[s], j = tri[T + s], k = tri[2 * T + s];
  const double px = X[i], py = Y[i], qx = X[j], qy = Y[j], rx = X[k], ry = Y[k];
  const double cx = (px + qx + rx) / 3, cy = (py + qy + ry) / 3;
  const double m1x = (px + qx) / 2, m1y = (py + qy) / 2;
  const double m2x = (qx + rx) / 2, m2y = (qy + ry) / 2;
  const double m3x = (rx + px) / 2, m3y = (ry + py) / 2;
  const double pcx = cx - px, pcy = cy - py;
  const double qcx = cx - qx, qcy = cy - qy;
  const double rcx = cx - rx, rcy = cy - ry;
  const double m13x = m1x - m3x, m13y = m1y - m3y;
  const double m21x = m2x - m1x, m21y = m2y - m1y;
  const double m32x = m3x - m2x, m32y = m3y - m2y;
  sub[s] = 0.5 * fabs(pcx * m13y - pcy * m13x);
  sub[T + s] = 0.5 * fabs(qcx * m21y - qcy * m21x);
  sub[2 * T + s] = 0.5 * fabs(rcx * m32y - rcy * m32x);
  const double D = qx * ry - qy * rx - px * ry + rx * py + px * qy - qx * py;
  double* P = props + s;
  P[0 * T] = (qy - ry) / D;
  P[1 * T] = (ry - py) / D;
  P[2 * T] = (py - qy) / D;
  P[3 * T] = (rx - qx) / D;
  P[4 * T] = (px - rx) / D;
  P[5 * T] = (qx - px) / D;
  P[6 * T] = (qx * ry - rx * qy) / D;
  P[7 * T] = (rx * py - px * ry) / D;
  P[8 * T] = (px * qy - qx * py) / D;
  P[9 * T] = (m1x + cx) / 2;
  P[10 * T] = (m1y + cy) / 2;
  P[11 * T] = (m2x + cx) / 2;
  P[12 * T] = (m2y + cy) / 2;
  P[13 * T] = (m3x + cx) / 2;
  P[14 * T] = (m3y + cy) / 2;
  const double e1x = cx - m1x, e1y = cy - m1y;
  const double e2x = cx - m2x, e2y = cy - m2y;
  const double e3x = cx - m3x, e3y = cy - m3y;
  const double l1 = sqrt(e1x * e1x + e1y * e1y);
  const double l2 = sqrt(e2x * e2x + e2y * e2y);
  const double l3 = sqrt(e3x * e3x + e3y * e3y);
  P[15 * T] = e1y / l1;
  P[16 * T] = -e1x / l1;
  P[17 * T] = e2y / l2;
  P[18 * T] = -e2x / l2;
  P[19 * T] = e3y / l3;
  P[20 * T] = -e3x / l3;
  P[21 * T] = l1;
  P[22 * T] = l2;
  P[23 * T] = l3;
}

// One thread per node: cv_volumes[i] = sum of its sub-volumes in ascending
// triangle order (deterministic replacement of geometry.jl:133-135's +=).
__global__ __launch_bounds__(256) void volume_kernel(int64_t N, int64_t T,
                                                     const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ inc,
                                                     const double* __restrict__ sub,
                                                     double* __restrict__ cv,
                                                     uint8_t* __restrict__ is_vertex) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  double v = 0.0;
  const int32_t k0 = start[g], k1 = start[g + 1];
  for (int32_t k = k0; k < k1; ++k) {
    const int32_t w = inc[k];
    v += sub[(int64_t)(w & 3) * T + (w >> 2)];
  }
  cv[g] = v;
  is_vertex[g] = k1 > k0;
}

// Control volume of every local-node entry, read contiguously by the tile kernels.
__global__ __launch_bounds__(256) void local_cv_kernel(int64_t n, const int32_t* __restrict__ nodes,
                                                       const double* __restrict__ cv,
                                                       double* __restrict__ lcv) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) lcv[e] = cv[nodes[e]];
}

// One thread per triangle: the record as the tile kernels rebuild it from the vertex
// coordinates (rebuild_record at FVM_GEO_LEVEL, the whole record from level 3 on).
__global__ __launch_bounds__(256) void rebuild_kernel(int64_t T, const int32_t* __restrict__ tri,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      double* __restrict__ P) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int32_t i = tri[t], j = tri[T + t], k = tri[2 * T + t];
  double R[P_COUNT];
  for (int f = 0; f < P_COUNT; ++f) R[f] = 0.0;
  rebuild_record<(FVM_GEO_LEVEL >= 3 ? FVM_GEO_LEVEL : 3)>(x[i], y[i], x[j], y[j], x[k], y[k], R);
#pragma unroll
  for (int f = 0; f < P_COUNT; ++f) P[f * T + t] = R[f];
}

void free_mesh(fvm_mesh* m) {
  if (!m) return;
  void* ptrs[] = {m->x, m->y, m->cv, m->props, m->tri, m->tile_tri_start, m->tile_node_start,
                  m->tile_nodes, m->loc, m->inc_start, m->inc, m->part_idx, m->fix_nodes,
                  m->fix_part_start, m->fix_bedge_start, m->fix_bedge, m->fix_bvert_start,
                  m->fix_bvert, m->lcv, m->node_tri_start,
                  m->node_tri, m->bedges, m->badj, m->perm, m->is_vertex, m->qflag};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete m;
}

// Whether the mesh covers a convex domain: its boundary edges (counter-clockwise, the
// domain on their left) form one closed loop, V - E + F = 1 (a disk: no holes), and
// the loop turns left or goes straight at every vertex.  Then a point strictly on the
// outer side of a boundary edge lies outside the domain.
bool convex_domain(const fvm_mesh_desc* d) {
  const int64_t N = d->n_points, T = d->n_triangles, B = d->n_boundary_edges;
  if (B < 3) return false;
  std::vector<int32_t> next(N, -1);
  for (int64_t e = 0; e < B; ++e) {
    const int32_t i = d->boundary_edges[2 * e], j = d->boundary_edges[2 * e + 1];
    if (next[i] >= 0) return false;  // a vertex where two boundary loops touch
    next[i] = j;
  }
  // one loop through every boundary edge
  const int32_t s = d->boundary_edges[0];
  int32_t v = s;
  int64_t steps = 0;
  do {
    const int32_t w = next[v];
    if (w < 0) return false;
    const int32_t z = next[w];
    if (z < 0) return false;
    const double ax = d->x[w] - d->x[v], ay = d->y[w] - d->y[v];
    const double bx = d->x[z] - d->x[w], by = d->y[z] - d->y[w];
    if (ax * by - ay * bx < 0.0) return false;  // a right turn: reflex vertex
    v = w;
    ++steps;
  } while (v != s && steps <= B);
  if (v != s || steps != B) return false;
  // Euler characteristic of the solid triangles' vertices, edges and faces: 1 for a disk
  std::vector<uint8_t> isv(N, 0);
  for (int64_t t = 0; t < 3 * T; ++t) isv[d->triangles[t]] = 1;
  int64_t nv = 0;
  for (int64_t i = 0; i < N; ++i) nv += isv[i];
  const int64_t E = (3 * T + B) / 2;  // every interior edge in two triangles, boundary edges in one
  return (3 * T + B) % 2 == 0 && nv - E + T == 1;
}

int build(const fvm_mesh_desc* d, const int32_t* tri_group, fvm_mesh* m) {
  // Tiles of up to 512 triangles run on 256-thread workgroups with room for 512
  // local nodes; tiles of up to 1024 triangles on 512-thread workgroups with 1024.
  const int tile_tris = d->tile_triangles > 0 ? d->tile_triangles : 512;
  FVM_CHECK(tile_tris <= 1024, "fvm_mesh_create: tile_triangles must be <= 1024");
  const int node_limit = tile_tris <= 512 ? 512 : 1024;
  const int max_nodes = d->tile_max_nodes > 0 ? d->tile_max_nodes : node_limit;
  FVM_CHECK(max_nodes <= node_limit,
            "fvm_mesh_create: tile_max_nodes must be <= 512 (<= 1024 for tiles above 512 triangles)");
  fvmh_tiles* tl = nullptr;
  if (fvmh_tiles_build_grouped(d->n_points, d->x, d->y, d->n_triangles, d->triangles, tri_group,
                               d->n_boundary_edges, d->boundary_edges, tile_tris, max_nodes, &tl) != 0) {
    set_error(std::string("fvm_mesh_create: ") + fvmh_last_error());
    return FVM_ERR_ARG;
  }
  struct Guard {
    fvmh_tiles* t;
    ~Guard() { fvmh_tiles_free(t); }
  } guard{tl};
  m->N = d->n_points;
  m->T = d->n_triangles;
  m->B = d->n_boundary_edges;
  m->n_tiles = tl->n_tiles;
  m->n_fix = tl->n_fix;
  m->n_parts = tl->n_parts;
  m->n_local = tl->n_local;
  m->tile_tris = tile_tris;
  m->max_nodes = max_nodes;
  const int64_t N = m->N, T = m->T, B = m->B;
  // the RHS tile kernel addresses these arrays with 32-bit byte offsets
  constexpr int64_t kOff = (int64_t)1 << 32;
  FVM_CHECK(8 * N < kOff && 8 * (int64_t)m->n_local < kOff && 6 * T < kOff,
            "fvm_mesh_create: mesh too large (the node and tile-node arrays must stay below 4 GiB: "
            "at most 536,870,911 points and tile-node slots)");
  int rc;
#define UP(dst, src, n) \
  if ((rc = upload(&(dst), src, n)) != FVM_OK) return rc
  UP(m->x, d->x, N);
  UP(m->y, d->y, N);
  UP(m->tri, tl->tri_sorted, 3 * T);
  UP(m->tile_tri_start, tl->tile_tri_start, m->n_tiles + 1);
  UP(m->tile_node_start, tl->tile_node_start, m->n_tiles + 1);
  UP(m->tile_nodes, tl->tile_nodes, m->n_local);
  UP(m->loc, tl->loc, T);
  UP(m->inc_start, tl->inc_start, m->n_local + 1);
  UP(m->inc, tl->inc, 3 * T);
  UP(m->part_idx, tl->part_idx, m->n_local);
  UP(m->fix_nodes, tl->fix_nodes, m->n_fix);
  UP(m->fix_part_start, tl->fix_part_start, m->n_fix + 1);
  UP(m->fix_bedge_start, tl->fix_bedge_start, m->n_fix + 1);
  UP(m->fix_bedge, tl->fix_bedge, tl->fix_bedge_start[m->n_fix]);
  UP(m->fix_bvert_start, tl->fix_bvert_start, m->n_fix + 1);
  UP(m->fix_bvert, tl->fix_bvert, tl->fix_bvert_start[m->n_fix]);
  UP(m->node_tri_start, tl->node_tri_start, N + 1);
  UP(m->node_tri, tl->node_tri, 3 * T);
  UP(m->badj, tl->badj_sorted, B);
  UP(m->perm, tl->perm, T);
  {
    std::vector<int32_t> be(2 * B);
    for (int64_t e = 0; e < B; ++e) {
      be[e] = d->boundary_edges[2 * e];
      be[B + e] = d->boundary_edges[2 * e + 1];
    }
    UP(m->bedges, be.data(), 2 * B);
  }
  UP(m->props, (const double*)nullptr, P_COUNT * T);
  UP(m->cv, (const double*)nullptr, N);
  UP(m->is_vertex, (const uint8_t*)nullptr, N);
  UP(m->lcv, (const double*)nullptr, m->n_local);
  UP(m->qflag, (const int32_t*)nullptr, 1);
#undef UP
  m->convex = convex_domain(d);
  m->n_bedge_items = tl->fix_bedge_start[m->n_fix];
  m->n_bvert_items = tl->fix_bvert_start[m->n_fix];
  m->h_part_idx.assign(tl->part_idx, tl->part_idx + m->n_local);
  m->h_tile_nodes.assign(tl->tile_nodes, tl->tile_nodes + m->n_local);
  m->h_perm.assign(tl->perm, tl->perm + T);
  m->h_node_tri_start.assign(tl->node_tri_start, tl->node_tri_start + N + 1);
  m->h_fix_part_start.assign(tl->fix_part_start, tl->fix_part_start + m->n_fix + 1);
  m->h_fix_bedge_start.assign(tl->fix_bedge_start, tl->fix_bedge_start + m->n_fix + 1);
  double* sub = nullptr;
  FVM_HIP(hipMalloc((void**)&sub, sizeof(double) * 3 * (size_t)(T > 0 ? T : 1)));
  if (T > 0)
    hipLaunchKernelGGL(geometry_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, 0, T, m->tri,
                       m->x, m->y, m->props, sub);
  if (N > 0)
    hipLaunchKernelGGL(volume_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, 0, N, T,
                       m->node_tri_start, m->node_tri, sub, m->cv, m->is_vertex);
  if (m->n_local > 0)
    hipLaunchKernelGGL(local_cv_kernel, dim3((unsigned)((m->n_local + 255) / 256)), dim3(256), 0, 0,
                       m->n_local, m->tile_nodes, m->cv, m->lcv);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(sub);
  if (e != hipSuccess) return hip_fail(e, "fvm_mesh_create: geometry kernels");
  return FVM_OK;
}

}  // namespace

namespace fvm {

int mesh_create_grouped(const fvm_mesh_desc* d, const int32_t* tri_group, fvm_mesh** out) {
  FVM_CHECK(d && out, "fvm_mesh_create: null argument");
  *out = nullptr;
  FVM_CHECK(d->n_points > 0 && d->n_points < INT32_MAX, "fvm_mesh_create: n_points out of range");
  FVM_CHECK(d->n_triangles > 0, "fvm_mesh_create: the mesh has no triangles");
  FVM_CHECK(d->n_boundary_edges >= 0, "fvm_mesh_create: negative boundary edge count");
  FVM_CHECK(d->x && d->y && d->triangles, "fvm_mesh_create: null mesh array");
  FVM_CHECK(d->n_boundary_edges == 0 || d->boundary_edges, "fvm_mesh_create: null boundary edges");
  fvm_mesh* m = new fvm_mesh();
  int rc = build(d, tri_group, m);
  if (rc != FVM_OK) {
    free_mesh(m);
    return rc;
  }
  *out = m;
  return FVM_OK;
}

}  // namespace fvm

extern "C" {

int fvm_mesh_create(const fvm_mesh_desc* d, fvm_mesh** out) {
  return fvm::mesh_create_grouped(d, nullptr, out);
}

int fvm_mesh_destroy(fvm_mesh* m) {
  free_mesh(m);
  return FVM_OK;
}

int fvm_mesh_get_info(const fvm_mesh* m, fvm_mesh_info* info) {
  FVM_CHECK(m && info, "fvm_mesh_get_info: null argument");
  info->n_points = m->N;
  info->n_triangles = m->T;
  info->n_boundary_edges = m->B;
  info->n_tiles = m->n_tiles;
  info->n_fix_nodes = m->n_fix;
  info->n_partials = m->n_parts;
  info->n_local_nodes = m->n_local;
  info->d_cv_volumes = m->cv;
  info->d_x = m->x;
  info->d_y = m->y;
  return FVM_OK;
}

int fvm_mesh_download_geometry(const fvm_mesh* m, double* cv_volumes, double* props) {
  FVM_CHECK(m, "fvm_mesh_download_geometry: null mesh");
  const int64_t N = m->N, T = m->T;
  if (cv_volumes) FVM_HIP(hipMemcpy(cv_volumes, m->cv, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (props) {
    std::vector<double> soa((size_t)P_COUNT * T);
    FVM_HIP(hipMemcpy(soa.data(), m->props, sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < T; ++s) {
      double* dst = props + (int64_t)P_COUNT * m->h_perm[s];
      for (int f = 0; f < P_COUNT; ++f) dst[f] = soa[(size_t)f * T + s];
    }
  }
  return FVM_OK;
}

int fvm_mesh_rebuilt_geometry(const fvm_mesh* m, double* props) {
  FVM_CHECK(m && props, "fvm_mesh_rebuilt_geometry: null argument");
  const int64_t T = m->T;
  double* d = nullptr;
  FVM_HIP(hipMalloc((void**)&d, sizeof(double) * (size_t)P_COUNT * T));
  hipLaunchKernelGGL(rebuild_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, 0, T, m->tri, m->x,
                     m->y, d);
  std::vector<double> soa((size_t)P_COUNT * T);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(soa.data(), d, sizeof(double) * soa.size(), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "fvm_mesh_rebuilt_geometry");
  for (int64_t s = 0; s < T; ++s) {
    double* dst = props + (int64_t)P_COUNT * m->h_perm[s];
    for (int f = 0; f < P_COUNT; ++f) dst[f] = soa[(size_t)f * T + s];
  }
  return FVM_OK;
}

}  // extern "C"
