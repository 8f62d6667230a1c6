// Host-side mesh generation and preprocessing (see fvm_host.h).
#include "fvm_host.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const char* msg) {
  g_err = msg;
  return -1;
}

// SplitMix64 (Steele, Lea, Flood 2014): state advances by the golden gamma.
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
};

// Element i of Julia's LinRange(a, b, n): t = i/(n-1); fma(t, b, fma(-t, a, a)).
inline double linrange(double a, double b, int i, int n) {
  if (n == 1) return a;
  double t = (double)i / (double)(n - 1);
  return std::fma(t, b, std::fma(-t, a, a));
}

// > 0 when d lies inside the circumcircle of the counter-clockwise triangle abc.
inline double incircle(double ax, double ay, double bx, double by, double cx, double cy,
                       double dx, double dy, double* scale) {
  double adx = ax - dx, ady = ay - dy;
  double bdx = bx - dx, bdy = by - dy;
  double cdx = cx - dx, cdy = cy - dy;
  double alift = adx * adx + ady * ady;
  double blift = bdx * bdx + bdy * bdy;
  double clift = cdx * cdx + cdy * cdy;
  double det = alift * (bdx * cdy - cdx * bdy) + blift * (cdx * ady - adx * cdy) +
               clift * (adx * bdy - bdx * ady);
  if (scale)
    *scale = alift * (std::fabs(bdx * cdy) + std::fabs(cdx * bdy)) +
             blift * (std::fabs(cdx * ady) + std::fabs(adx * cdy)) +
             clift * (std::fabs(adx * bdy) + std::fabs(bdx * ady));
  return det;
}

inline uint32_t spread16(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

template <class T>
T* xalloc(int64_t n) {
  return (T*)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));
}

}  // namespace

extern "C" {

const char* fvmh_last_error(void) { return g_err.c_str(); }

int fvmh_rectangle_sizes(int nx, int ny, int single_boundary, int64_t* n_points,
                         int64_t* n_triangles, int64_t* n_boundary_edges,
                         int32_t* n_segments) {
  if (nx < 2 || ny < 2) return fail("fvmh_rectangle_sizes: need nx >= 2 and ny >= 2");
  if (n_points) *n_points = (int64_t)nx * ny;
  if (n_triangles) *n_triangles = 2 * (int64_t)(nx - 1) * (ny - 1);
  if (n_boundary_edges) *n_boundary_edges = 2 * (int64_t)(nx - 1) + 2 * (int64_t)(ny - 1);
  if (n_segments) *n_segments = single_boundary ? 1 : 4;
  return 0;
}

int fvmh_rectangle(double a, double b, double c, double d, int nx, int ny,
                   int single_boundary, double jitter, uint64_t seed, double* x,
                   double* y, int32_t* tri, int32_t* bedges, int32_t* bseg) {
  if (nx < 2 || ny < 2) return fail("fvmh_rectangle: need nx >= 2 and ny >= 2");
  if (!(b > a) || !(d > c)) return fail("fvmh_rectangle: need a < b and c < d");
  if (jitter < 0 || jitter > 0.25) return fail("fvmh_rectangle: jitter must lie in [0, 0.25]");
  if ((int64_t)nx * ny > INT32_MAX / 4) return fail("fvmh_rectangle: mesh too large for int32 indices");
  const double hx = (b - a) / (nx - 1), hy = (d - c) / (ny - 1);
  SplitMix64 rng{seed};
  for (int j = 0; j < ny; ++j) {
    double yj = linrange(c, d, j, ny);
    for (int i = 0; i < nx; ++i) {
      int64_t k = i + (int64_t)nx * j;
      double xi = linrange(a, b, i, nx), yv = yj;
      if (jitter > 0 && i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
        double ux = rng.uniform(), uy = rng.uniform();
        xi += (2.0 * ux - 1.0) * jitter * hx;
        yv += (2.0 * uy - 1.0) * jitter * hy;
      }
      x[k] = xi;
      y[k] = yv;
    }
  }
  int64_t t = 0;
  for (int j = 0; j < ny - 1; ++j) {
    for (int i = 0; i < nx - 1; ++i) {
      int32_t bl = i + nx * j, br = bl + 1, tl = bl + nx, tr = tl + 1;
      bool flip = false;
      if (jitter > 0) {
        double det = incircle(x[bl], y[bl], x[br], y[br], x[tl], y[tl], x[tr], y[tr], nullptr);
        flip = det > 0;
      }
      int32_t* T0 = tri + 3 * t;
      int32_t* T1 = tri + 3 * (t + 1);
      if (!flip) {
        T0[0] = bl; T0[1] = br; T0[2] = tl;
        T1[0] = br; T1[1] = tr; T1[2] = tl;
      } else {
        T0[0] = bl; T0[1] = br; T0[2] = tr;
        T1[0] = bl; T1[1] = tr; T1[2] = tl;
      }
      t += 2;
    }
  }
  int64_t e = 0;
  auto push = [&](int32_t u, int32_t v, int32_t s) {
    bedges[2 * e] = u;
    bedges[2 * e + 1] = v;
    bseg[e] = single_boundary ? 0 : s;
    ++e;
  };
  for (int i = 0; i < nx - 1; ++i) push(i, i + 1, 0);
  for (int j = 0; j < ny - 1; ++j) push(nx - 1 + nx * j, nx - 1 + nx * (j + 1), 1);
  for (int i = nx - 1; i > 0; --i) push(i + nx * (ny - 1), i - 1 + nx * (ny - 1), 2);
  for (int j = ny - 1; j > 0; --j) push(nx * j, nx * (j - 1), 3);
  return 0;
}

int64_t fvmh_count_non_delaunay(int64_t n_points, const double* x, const double* y,
                                int64_t n_tri, const int32_t* tri, double rel_tol) {
  // Directed edges keyed by the undirected pair; each interior edge appears twice.
  std::vector<std::pair<uint64_t, int64_t>> edges;
  edges.reserve(3 * n_tri);
  for (int64_t t = 0; t < n_tri; ++t)
    for (int s = 0; s < 3; ++s) {
      uint64_t u = (uint32_t)tri[3 * t + s], v = (uint32_t)tri[3 * t + (s + 1) % 3];
      uint64_t key = u < v ? (u << 32 | v) : (v << 32 | u);
      edges.push_back({key, 3 * t + s});
    }
  std::sort(edges.begin(), edges.end());
  int64_t bad = 0;
  for (size_t k = 0; k + 1 < edges.size(); ++k) {
    if (edges[k].first != edges[k + 1].first) continue;
    int64_t t1 = edges[k].second / 3, s1 = edges[k].second % 3;
    int64_t t2 = edges[k + 1].second / 3, s2 = edges[k + 1].second % 3;
    int32_t opp2 = tri[3 * t2 + (s2 + 2) % 3];
    (void)s1;
    const int32_t* T = tri + 3 * t1;
    double scale;
    double det = incircle(x[T[0]], y[T[0]], x[T[1]], y[T[1]], x[T[2]], y[T[2]], x[opp2],
                          y[opp2], &scale);
    if (det > rel_tol * scale) ++bad;
    ++k;
  }
  (void)n_points;
  return bad;
}

int fvmh_boundary_adjacent(int64_t n_tri, const int32_t* tri, int64_t n_bedges,
                           const int32_t* bedges, int32_t* badj) {
  std::unordered_map<uint64_t, int64_t> want;
  want.reserve(2 * n_bedges + 1);
  for (int64_t e = 0; e < n_bedges; ++e) {
    badj[e] = -1;
    want[((uint64_t)(uint32_t)bedges[2 * e] << 32) | (uint32_t)bedges[2 * e + 1]] = e;
  }
  for (int64_t t = 0; t < n_tri; ++t)
    for (int s = 0; s < 3; ++s) {
      uint64_t key = ((uint64_t)(uint32_t)tri[3 * t + s] << 32) | (uint32_t)tri[3 * t + (s + 1) % 3];
      auto it = want.find(key);
      if (it != want.end()) badj[it->second] = (int32_t)t;
    }
  for (int64_t e = 0; e < n_bedges; ++e)
    if (badj[e] < 0) return fail("fvmh_boundary_adjacent: a boundary edge has no triangle on its left");
  return 0;
}

void fvmh_tiles_free(fvmh_tiles* t) {
  if (!t) return;
  free(t->perm); free(t->tri_sorted); free(t->tile_tri_start); free(t->tile_node_start);
  free(t->tile_nodes); free(t->loc); free(t->inc_start); free(t->inc); free(t->part_idx);
  free(t->fix_nodes); free(t->fix_part_start); free(t->fix_bedge_start); free(t->fix_bedge);
  free(t->fix_bvert_start); free(t->fix_bvert);
  free(t->node_tri_start); free(t->node_tri); free(t->badj_sorted); free(t->tile_of_node_excl);
  free(t);
}

int fvmh_tiles_build(int64_t n_points, const double* x, const double* y, int64_t n_tri,
                     const int32_t* tri, int64_t n_bedges, const int32_t* bedges,
                     int tile_tris, int max_local_nodes, fvmh_tiles** out) {
  return fvmh_tiles_build_grouped(n_points, x, y, n_tri, tri, nullptr, n_bedges, bedges, tile_tris,
                                  max_local_nodes, out);
}

int fvmh_tiles_build_grouped(int64_t n_points, const double* x, const double* y, int64_t n_tri,
                             const int32_t* tri, const int32_t* tri_group, int64_t n_bedges,
                             const int32_t* bedges, int tile_tris, int max_local_nodes,
                             fvmh_tiles** out) {
  *out = nullptr;
  if (tile_tris < 1 || tile_tris > 4096) return fail("fvmh_tiles_build: tile_tris must lie in [1, 4096]");
  if (max_local_nodes < 3 || max_local_nodes > 1024)
    return fail("fvmh_tiles_build: max_local_nodes must lie in [3, 1024]");
  if (n_tri >= (int64_t)1 << 29) return fail("fvmh_tiles_build: too many triangles");
  for (int64_t t = 0; t < 3 * n_tri; ++t)
    if (tri[t] < 0 || tri[t] >= n_points) return fail("fvmh_tiles_build: triangle vertex out of range");
  for (int64_t e = 0; e < 2 * n_bedges; ++e)
    if (bedges[e] < 0 || bedges[e] >= n_points) return fail("fvmh_tiles_build: boundary edge vertex out of range");

  fvmh_tiles* T = (fvmh_tiles*)calloc(1, sizeof(fvmh_tiles));
  T->n_points = n_points; T->n_tri = n_tri; T->n_bedges = n_bedges;
  T->tile_tris = tile_tris; T->max_local_nodes = max_local_nodes;

  // 1. Morton order of centroids (16 bits per axis, ties by input index).
  double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
  for (int64_t i = 0; i < n_points; ++i) {
    xmin = std::min(xmin, x[i]); xmax = std::max(xmax, x[i]);
    ymin = std::min(ymin, y[i]); ymax = std::max(ymax, y[i]);
  }
  double sx = (xmax > xmin) ? 65535.0 / (xmax - xmin) : 0.0;
  double sy = (ymax > ymin) ? 65535.0 / (ymax - ymin) : 0.0;
  std::vector<uint64_t> keys(n_tri);
  for (int64_t t = 0; t < n_tri; ++t) {
    const int32_t* V = tri + 3 * t;
    double cx = (x[V[0]] + x[V[1]] + x[V[2]]) / 3.0, cy = (y[V[0]] + y[V[1]] + y[V[2]]) / 3.0;
    uint32_t qx = (uint32_t)std::min(65535.0, std::max(0.0, (cx - xmin) * sx));
    uint32_t qy = (uint32_t)std::min(65535.0, std::max(0.0, (cy - ymin) * sy));
    uint64_t m = spread16(qx) | (spread16(qy) << 1);
    keys[t] = (m << 32) | (uint64_t)t;
  }
  std::sort(keys.begin(), keys.end());
  if (tri_group)  // members stay contiguous, each in its own Morton order
    std::stable_sort(keys.begin(), keys.end(), [&](uint64_t a, uint64_t b) {
      return tri_group[a & 0xFFFFFFFFu] < tri_group[b & 0xFFFFFFFFu];
    });
  T->perm = xalloc<int32_t>(n_tri);
  std::vector<int32_t> invperm(n_tri);
  for (int64_t s = 0; s < n_tri; ++s) {
    T->perm[s] = (int32_t)(keys[s] & 0xFFFFFFFFu);
    invperm[T->perm[s]] = (int32_t)s;
  }
  keys.clear(); keys.shrink_to_fit();
  T->tri_sorted = xalloc<int32_t>(3 * n_tri);
  for (int64_t s = 0; s < n_tri; ++s)
    for (int v = 0; v < 3; ++v) T->tri_sorted[v * n_tri + s] = tri[3 * (int64_t)T->perm[s] + v];
  auto vert = [&](int64_t s, int v) { return T->tri_sorted[v * n_tri + s]; };

  // 2. Greedy tiles along the curve.
  std::vector<int32_t> stamp(n_points, -1), tile_start;
  std::vector<int32_t> node_tiles(n_points, 0);
  std::vector<int32_t> first_tile(n_points, INT32_MAX);  // the first (lowest) tile holding each node
  {
    int32_t tile = 0, count = 0, nodes = 0;
    tile_start.push_back(0);
    for (int64_t s = 0; s < n_tri; ++s) {
      int add = 0;
      for (int v = 0; v < 3; ++v) add += stamp[vert(s, v)] != tile;
      // a triangle's vertices are distinct, so add counts new nodes exactly
      const bool new_group = tri_group && s > 0 && tri_group[T->perm[s]] != tri_group[T->perm[s - 1]];
      if (count == tile_tris || nodes + add > max_local_nodes || (new_group && count > 0)) {
        ++tile;
        tile_start.push_back((int32_t)s);
        count = 0; nodes = 0;
        add = 3;
      }
      for (int v = 0; v < 3; ++v) {
        int32_t g = vert(s, v);
        if (stamp[g] != tile) {
          stamp[g] = tile;
          if (++node_tiles[g] == 1) first_tile[g] = tile;
        }
      }
      nodes += add;
      ++count;
    }
    tile_start.push_back((int32_t)n_tri);
    if (n_tri == 0) tile_start.assign(1, 0);
  }
  const int32_t n_tiles = (int32_t)tile_start.size() - 1;
  T->n_tiles = n_tiles;
  T->tile_tri_start = xalloc<int32_t>(n_tiles + 1);
  std::copy(tile_start.begin(), tile_start.end(), T->tile_tri_start);

  // Boundary edges: adjacent triangle (input index); its vertices and the edge's
  // endpoints all receive boundary-edge terms (the endpoints in the RHS, the
  // triangle's vertices in the adjoint), so they go through the fix-up pass.
  std::vector<int32_t> badj(n_bedges);
  if (n_bedges > 0 && fvmh_boundary_adjacent(n_tri, tri, n_bedges, bedges, badj.data()) != 0) {
    fvmh_tiles_free(T);
    return -1;
  }
  std::vector<char> on_bnd(n_points, 0);
  for (int64_t e = 0; e < 2 * n_bedges; ++e) on_bnd[bedges[e]] = 1;
  for (int64_t e = 0; e < n_bedges; ++e)
    for (int v = 0; v < 3; ++v) on_bnd[tri[3 * (int64_t)badj[e] + v]] = 1;
  // fix nodes: shared between tiles, on the boundary, or not a vertex; ordered by their
  // first tile (then by index), so that the fix-up, whose workgroups take contiguous
  // ranges of fix nodes per XCD as the tile kernel's take contiguous ranges of tiles,
  // reads partial sums the same XCD's tiles wrote
  std::vector<int32_t> fix_of(n_points, -1);
  std::vector<int32_t> fix_nodes;
  for (int64_t g = 0; g < n_points; ++g)
    if (node_tiles[g] != 1 || on_bnd[g]) fix_nodes.push_back((int32_t)g);
  std::stable_sort(fix_nodes.begin(), fix_nodes.end(),
                   [&](int32_t a, int32_t b) { return first_tile[a] < first_tile[b]; });
  for (size_t f = 0; f < fix_nodes.size(); ++f) fix_of[fix_nodes[f]] = (int32_t)f;
  T->n_fix = (int32_t)fix_nodes.size();
  T->fix_nodes = xalloc<int32_t>(T->n_fix);
  std::copy(fix_nodes.begin(), fix_nodes.end(), T->fix_nodes);
  T->fix_part_start = xalloc<int32_t>(T->n_fix + 1);
  T->fix_part_start[0] = 0;
  for (int32_t f = 0; f < T->n_fix; ++f)
    T->fix_part_start[f + 1] = T->fix_part_start[f] + node_tiles[fix_nodes[f]];
  T->n_parts = T->fix_part_start[T->n_fix];
  std::vector<int32_t> next_slot(T->n_fix);
  for (int32_t f = 0; f < T->n_fix; ++f) next_slot[f] = T->fix_part_start[f];

  // 3. Local node lists, packed local ids, local incidence lists.
  int64_t n_local = 0;
  for (int64_t g = 0; g < n_points; ++g) n_local += node_tiles[g];
  T->n_local = n_local;
  T->tile_node_start = xalloc<int32_t>(n_tiles + 1);
  T->tile_nodes = xalloc<int32_t>(n_local);
  T->part_idx = xalloc<int32_t>(n_local);
  T->loc = xalloc<uint32_t>(n_tri);
  T->inc_start = xalloc<int32_t>(n_local + 1);
  T->inc = xalloc<uint16_t>(3 * n_tri);
  T->tile_of_node_excl = xalloc<int32_t>(n_points);
  for (int64_t g = 0; g < n_points; ++g) T->tile_of_node_excl[g] = -1;
  std::vector<int32_t> lid(n_points, -1);
  std::vector<int32_t> nodes;
  std::vector<int32_t> cnt;
  int64_t pos = 0, ipos = 0;
  T->tile_node_start[0] = 0;
  for (int32_t tl = 0; tl < n_tiles; ++tl) {
    int32_t s0 = tile_start[tl], s1 = tile_start[tl + 1];
    nodes.clear();
    for (int32_t s = s0; s < s1; ++s)
      for (int v = 0; v < 3; ++v) nodes.push_back(vert(s, v));
    std::sort(nodes.begin(), nodes.end());
    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
    if ((int)nodes.size() > max_local_nodes) { fvmh_tiles_free(T); return fail("fvmh_tiles_build: internal tile overflow"); }
    for (size_t k = 0; k < nodes.size(); ++k) {
      int32_t g = nodes[k];
      lid[g] = (int32_t)k;
      T->tile_nodes[pos + k] = g;
      if (fix_of[g] >= 0) T->part_idx[pos + k] = next_slot[fix_of[g]]++;
      else { T->part_idx[pos + k] = -1; T->tile_of_node_excl[g] = tl; }
    }
    cnt.assign(nodes.size() + 1, 0);
    for (int32_t s = s0; s < s1; ++s) {
      uint32_t a = lid[vert(s, 0)], b = lid[vert(s, 1)], c = lid[vert(s, 2)];
      T->loc[s] = a | (b << 10) | (c << 20);
      ++cnt[a + 1]; ++cnt[b + 1]; ++cnt[c + 1];
    }
    for (size_t k = 0; k < nodes.size(); ++k) cnt[k + 1] += cnt[k];
    for (size_t k = 0; k < nodes.size(); ++k) T->inc_start[pos + k] = (int32_t)(ipos + cnt[k]);
    std::vector<int32_t> fillp(cnt.begin(), cnt.end() - 1);
    for (int32_t s = s0; s < s1; ++s)
      for (int v = 0; v < 3; ++v) {
        int32_t l = lid[vert(s, v)];
        T->inc[ipos + fillp[l]++] = (uint16_t)(((s - s0) << 2) | v);
      }
    ipos += 3 * (s1 - s0);
    pos += (int64_t)nodes.size();
    T->tile_node_start[tl + 1] = (int32_t)pos;
  }
  T->inc_start[n_local] = (int32_t)ipos;

  // 4. Global incidence list (ascending sorted triangle index).
  T->node_tri_start = xalloc<int32_t>(n_points + 1);
  T->node_tri = xalloc<int32_t>(3 * n_tri);
  {
    std::vector<int64_t> c2(n_points + 1, 0);
    for (int64_t s = 0; s < n_tri; ++s)
      for (int v = 0; v < 3; ++v) ++c2[vert(s, v) + 1];
    for (int64_t g = 0; g < n_points; ++g) c2[g + 1] += c2[g];
    for (int64_t g = 0; g <= n_points; ++g) T->node_tri_start[g] = (int32_t)c2[g];
    for (int64_t s = 0; s < n_tri; ++s)
      for (int v = 0; v < 3; ++v) T->node_tri[c2[vert(s, v)]++] = (int32_t)((s << 2) | v);
  }

  // 5. Boundary edges: adjacent triangle (sorted index) and per-fix-node edge lists.
  T->badj_sorted = xalloc<int32_t>(n_bedges);
  for (int64_t e = 0; e < n_bedges; ++e) T->badj_sorted[e] = invperm[badj[e]];
  T->fix_bedge_start = xalloc<int32_t>(T->n_fix + 1);
  T->fix_bedge = xalloc<int32_t>(2 * n_bedges);
  {
    std::vector<int32_t> c3(T->n_fix + 1, 0);
    for (int64_t e = 0; e < n_bedges; ++e)
      for (int ep = 0; ep < 2; ++ep) ++c3[fix_of[bedges[2 * e + ep]] + 1];
    for (int32_t f = 0; f < T->n_fix; ++f) c3[f + 1] += c3[f];
    std::copy(c3.begin(), c3.end(), T->fix_bedge_start);
    for (int64_t e = 0; e < n_bedges; ++e)
      for (int ep = 0; ep < 2; ++ep) T->fix_bedge[c3[fix_of[bedges[2 * e + ep]]]++] = (int32_t)((e << 1) | ep);
  }
  T->fix_bvert_start = xalloc<int32_t>(T->n_fix + 1);
  T->fix_bvert = xalloc<int32_t>(3 * n_bedges);
  {
    std::vector<int32_t> c4(T->n_fix + 1, 0);
    for (int64_t e = 0; e < n_bedges; ++e)
      for (int v = 0; v < 3; ++v) ++c4[fix_of[tri[3 * (int64_t)badj[e] + v]] + 1];
    for (int32_t f = 0; f < T->n_fix; ++f) c4[f + 1] += c4[f];
    std::copy(c4.begin(), c4.end(), T->fix_bvert_start);
    for (int64_t e = 0; e < n_bedges; ++e)
      for (int v = 0; v < 3; ++v)
        T->fix_bvert[c4[fix_of[tri[3 * (int64_t)badj[e] + v]]]++] = (int32_t)((e << 2) | v);
  }
  *out = T;
  return 0;
}

int fvmh_csr_pattern(int64_t n_points, int64_t n_tri, const int32_t* tri, int64_t* row_ptr,
                     int32_t* col, int64_t* nnz_out) {
  std::vector<int64_t> c(n_points + 1, 0);
  for (int64_t t = 0; t < n_tri; ++t)
    for (int v = 0; v < 3; ++v) ++c[tri[3 * t + v] + 1];
  for (int64_t g = 0; g < n_points; ++g) c[g + 1] += c[g];
  std::vector<int32_t> inc(3 * n_tri);
  {
    std::vector<int64_t> f(c.begin(), c.end() - 1);
    for (int64_t t = 0; t < n_tri; ++t)
      for (int v = 0; v < 3; ++v) inc[f[tri[3 * t + v]]++] = (int32_t)t;
  }
  int64_t nnz = 0;
  std::vector<int32_t> row;
  for (int64_t g = 0; g < n_points; ++g) {
    row.clear();
    row.push_back((int32_t)g);
    for (int64_t k = c[g]; k < c[g + 1]; ++k)
      for (int v = 0; v < 3; ++v) row.push_back(tri[3 * (int64_t)inc[k] + v]);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    if (row_ptr) row_ptr[g] = nnz;
    if (col) std::copy(row.begin(), row.end(), col + nnz);
    nnz += (int64_t)row.size();
  }
  if (row_ptr) row_ptr[n_points] = nnz;
  if (nnz_out) *nnz_out = nnz;
  return 0;
}

int fvmh_colour_d2(int64_t n, const int64_t* row_ptr, const int32_t* col, int32_t* colour) {
  // colours are taken in node order; a node's forbidden set is every colour within
  // distance two, marked in a stamp array (no clearing between nodes)
  std::fill(colour, colour + n, -1);
  std::vector<int64_t> stamp;
  int32_t n_colours = 0;
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t a = row_ptr[j]; a < row_ptr[j + 1]; ++a) {
      const int64_t i = col[a];
      if (i < 0 || i >= n) return fail("fvmh_colour_d2: column index out of range");
      for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; ++b) {
        const int32_t c = colour[col[b]];
        if (c >= 0) {
          if ((int64_t)stamp.size() <= c) stamp.resize(c + 1, -1);
          stamp[c] = j;
        }
      }
    }
    int32_t c = 0;
    while (c < (int32_t)stamp.size() && stamp[c] == j) ++c;
    colour[j] = c;
    if (c + 1 > n_colours) n_colours = c + 1;
  }
  return n_colours;
}

}  // extern "C"
