// Host-side mesh generation and mesh preprocessing for the device library.
//
// Everything here runs on the CPU, once per mesh: it builds the inputs that the
// HIP kernels stream (triangle tiles, local node lists, incidence lists).  It is a
// plain C++17 library (libfvmhost.so) so that it can be exercised without a GPU.
//
// Index conventions: all node / triangle / edge indices are 0-based int32.  A
// reference node index k (1-based, Julia) is node k-1 here.
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// ---------------------------------------------------------------------------
// Mesh generation
// ---------------------------------------------------------------------------

// Sizes of the mesh produced by fvmh_rectangle for an nx x ny point grid.
int fvmh_rectangle_sizes(int nx, int ny, int single_boundary, int64_t* n_points,
                         int64_t* n_triangles, int64_t* n_boundary_edges,
                         int32_t* n_segments);

// Rectangle [a,b]x[c,d] with nx*ny points numbered x-fastest (node i + nx*j is
// (a + i*hx, c + j*hy)), two triangles per cell.
//   jitter == 0: every cell (BL,BR,TL,TR) is split along BR-TL into the
//     counter-clockwise triangles (BL,BR,TL) and (BR,TR,TL), the layout of
//     DelaunayTriangulation.triangulate_rectangle (test/equations.jl:42-56 pins
//     node nx+1 = (0,h) and the first triangle (1,2,nx+1)).
//   jitter > 0: interior points move by U(-jitter*hx, jitter*hx) x
//     U(-jitter*hy, jitter*hy) drawn from SplitMix64(seed) in node order (x then
//     y per node), and each cell takes the diagonal that makes it locally
//     Delaunay.  For jitter <= 0.25 every grid edge is then locally Delaunay as
//     well, so the result is the Delaunay triangulation of the points.
// Boundary edges run counter-clockwise (interior on the left):
//   bottom (left->right), right (bottom->top), top (right->left), left (top->bottom).
// bseg[e] is the boundary segment of edge e: 0 for all edges if single_boundary,
// else 0 bottom, 1 right, 2 top, 3 left (the order of the reference's
// BoundaryConditions for triangulate_rectangle(...; single_boundary=false)).
int fvmh_rectangle(double a, double b, double c, double d, int nx, int ny,
                   int single_boundary, double jitter, uint64_t seed, double* x,
                   double* y, int32_t* tri, int32_t* bedges, int32_t* bseg);

// Number of interior edges of the triangulation that are not locally Delaunay
// (the opposite vertex lies strictly inside the circumcircle by more than
// rel_tol relative to the determinant's scale).  0 means Delaunay.
int64_t fvmh_count_non_delaunay(int64_t n_points, const double* x, const double* y,
                                int64_t n_tri, const int32_t* tri, double rel_tol);

// For each boundary edge (i,j) the index of the triangle holding the directed
// edge i->j, i.e. (i, j, get_adjacent(tri, i, j)) up to rotation.  -1 if none.
int fvmh_boundary_adjacent(int64_t n_tri, const int32_t* tri, int64_t n_bedges,
                           const int32_t* bedges, int32_t* badj);

// ---------------------------------------------------------------------------
// Tiles: the layout streamed by the RHS / JVP / VJP kernels.
// ---------------------------------------------------------------------------
//
// Triangles are sorted along a Morton curve of their centroids and cut into
// tiles of at most tile_tris triangles and at most max_local_nodes distinct
// vertices.  Each tile owns a sorted list of its vertices ("local nodes"); a
// triangle stores its three local indices packed 10 bits each.  Within a tile
// every local node sums its incident triangles' fluxes in a fixed order (the
// local incidence list), so the scatter is deterministic without atomics.
// A node is "exclusive" when all of its triangles lie in one tile and it is not
// a vertex of a triangle touching the mesh boundary: that tile writes its du
// directly.  Every other node (tile-boundary nodes, vertices of boundary
// triangles, points that are not vertices) is a "fix" node: each tile touching
// it writes one partial sum into a dedicated slot and a second pass adds the
// slots in tile order plus the boundary-edge terms.
typedef struct fvmh_tiles {
  int64_t n_points, n_tri, n_bedges;
  int32_t n_tiles, tile_tris, max_local_nodes;
  int64_t n_local;          // total local-node entries over all tiles
  int32_t n_fix, n_parts;   // fix nodes, partial slots
  int32_t* perm;            // [n_tri] sorted position -> input triangle index
  int32_t* tri_sorted;      // [3*n_tri] SoA: a[n_tri], b[n_tri], c[n_tri] (sorted order)
  int32_t* tile_tri_start;  // [n_tiles+1]
  int32_t* tile_node_start; // [n_tiles+1]
  int32_t* tile_nodes;      // [n_local] global node ids, ascending within a tile
  uint32_t* loc;            // [n_tri] local ids a | b<<10 | c<<20 (sorted order)
  int32_t* inc_start;       // [n_local+1] absolute offsets into inc
  uint16_t* inc;            // [3*n_tri] (local triangle << 2) | vertex slot
  int32_t* part_idx;        // [n_local] partial slot of this entry, -1 if exclusive
  int32_t* fix_nodes;       // [n_fix] ascending global ids
  int32_t* fix_part_start;  // [n_fix+1] contiguous slot range per fix node
  int32_t* fix_bedge_start; // [n_fix+1]
  int32_t* fix_bedge;       // (boundary edge << 1) | endpoint (0 = i, 1 = j)
  int32_t* fix_bvert_start; // [n_fix+1]
  int32_t* fix_bvert;       // (boundary edge << 2) | slot of this node in the edge's triangle
  int32_t* node_tri_start;  // [n_points+1] global incidence list
  int32_t* node_tri;        // [3*n_tri] (sorted triangle << 2) | slot, ascending
  int32_t* badj_sorted;     // [n_bedges] adjacent triangle (sorted index)
  int32_t* tile_of_node_excl; // [n_points] tile that owns an exclusive node, -1 otherwise
} fvmh_tiles;

int fvmh_tiles_build(int64_t n_points, const double* x, const double* y, int64_t n_tri,
                     const int32_t* tri, int64_t n_bedges, const int32_t* bedges,
                     int tile_tris, int max_local_nodes, fvmh_tiles** out);
// fvmh_tiles_build with the triangles ordered by tri_group first (then Morton
// order within a group): an ensemble of disjoint meshes packed into one mesh keeps
// each member's triangles contiguous, so no tile interleaves two members.
int fvmh_tiles_build_grouped(int64_t n_points, const double* x, const double* y, int64_t n_tri,
                             const int32_t* tri, const int32_t* tri_group, int64_t n_bedges,
                             const int32_t* bedges, int tile_tris, int max_local_nodes,
                             fvmh_tiles** out);
void fvmh_tiles_free(fvmh_tiles* t);

// Sparsity of the Jacobian / template operators: row i holds i and every
// vertex sharing an edge with i (ascending).  Rows of points that are not
// vertices hold only the diagonal.  Mirrors jacobian_sparsity (solve.jl:56-77).
// Call with col == NULL to get nnz in *nnz_out; then with arrays sized nnz.
int fvmh_csr_pattern(int64_t n_points, int64_t n_tri, const int32_t* tri,
                     int64_t* row_ptr, int32_t* col, int64_t* nnz_out);

// Greedy distance-2 colouring of a symmetric CSR pattern (row i holds i): columns
// that share a row get different colours, so one JVP probe per colour recovers the
// matrix (sparse Jacobian extraction).  colour[n]; returns the number of colours, or
// -1 on an index out of range.
int fvmh_colour_d2(int64_t n, const int64_t* row_ptr, const int32_t* col, int32_t* colour);

// Last error message of this library (thread-local).
const char* fvmh_last_error(void);

#ifdef __cplusplus
}
#endif
