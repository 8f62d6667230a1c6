// Linear templates on the device: CSR assembly of the template operators, SpMV,
// and the steady solvers that replace the reference's sparse direct solve.
//
// Reference: src/specific_problems/abstract_templates.jl:73-325 (triangle and
// boundary-edge contributions, right-hand sides, Dirichlet rows, missing
// vertices), poissons_equation.jl:58-88, laplaces_equation.jl:51-78,
// mean_exit_time.jl:57-94, diffusion_equation.jl:69-101,
// linear_reaction_diffusion_equations.jl:76-125.  The reference fills a dense
// n x n matrix and converts it with sparse(); here each row is assembled by one
// thread into the CSR pattern of jacobian_sparsity (the node and its mesh
// neighbours, solve.jl:50-77), in a fixed order, so the result is deterministic.
//
// Steady solves.  A (rows scaled by 1/V_i, Dirichlet and missing-vertex rows =
// identity) is not symmetric.  M = -diag(V) A restricted to the free rows and
// columns (identity on the fixed rows) is the P1 stiffness matrix when D is
// constant per triangle and close to symmetric otherwise, so Jacobi-preconditioned
// CG on M d = -V (b - A x) converges; the stopping test is always the FP64 true
// residual |b - A x|_2 <= rtol |b|_2 on A itself.
//   FVM_SOLVE_CG        FP64 CG; a restart from the true residual if the
//                       recursive residual has drifted.
//   FVM_SOLVE_CG_MIXED  iterative refinement: the correction comes from CG with
//                       M, vectors and arithmetic in FP32 (dot products summed in
//                       FP64), stopped at a relative drop of inner_rtol; the
//                       residual and the solution are FP64.
//   FVM_SOLVE_CG_RELIABLE  one FP32 CG sequence with reliable updates: the FP32
//                       correction is folded into the FP64 x and the residual
//                       recomputed in FP64 whenever it has dropped by inner_rtol
//                       (default 0.1), keeping the search direction.
//   FVM_SOLVE_BICGSTAB  FP64 BiCGStab on M with Jacobi right preconditioning.
// Each reduction leaves one partial per workgroup, and every workgroup of the next
// kernel re-adds all of them in the same fixed order, so solver scalars stay on the
// device and results are bitwise reproducible; the host reads the scalars only
// every check_every iterations.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>
#include <vector>

#include "fvm_internal.hpp"

using namespace fvm;

namespace {

constexpr int kRT = FVM_SOLVER_THREADS;  // threads per workgroup of every solver kernel
constexpr int kMaxG = 1024;              // workgroups of the solver kernels (fixed per template)
constexpr int kLPR = 8;                  // lanes per CSR row in the SpMV kernels
constexpr int kRowsPerBlock = kRT / kLPR;
constexpr int kRPG = FVM_SPMV_ROWS;      // rows per lane group per step of the SpMV loops
constexpr int kEllMin = 4, kEllMax = 10; // unrolled slots of the sliced ELL form: the longest row's, in 4..10
constexpr int kSlice = 64;               // rows per slice of the sliced ELL form: one wave

// solver scalars (device doubles)
// (S_RZ, S_ALPHA, S_RHO and S_OMEGA take two slots each: value of an even / odd iteration)
enum {
  S_RZ = 0, S_ALPHA = 2, S_RHO = 4, S_OMEGA = 6,
  S_WR = 8, S_TOL2, S_ITER, S_DONE, S_RES2, S_B2, S_FLAG, S_COUNT = 16
};
// partial-sum slots (kMaxG doubles each)
enum { PT_RZ = 0, PT_WR, PT_PQ, PT_TT, PT_RES, PT_COUNT };

}  // namespace

struct fvm_template {
  const fvm_mesh* m = nullptr;
  int32_t kind = 0;
  int64_t n = 0, nnz = 0;
  int32_t* row_ptr = nullptr;
  int32_t* col = nullptr;
  double* val = nullptr;     // A
  double* valM = nullptr;    // M, FP64
  float* valM32 = nullptr;   // M, FP32
  double* b = nullptr;
  uint8_t* fixed = nullptr;  // identity rows of A
  double* vol = nullptr;     // V_i
  double* dinv = nullptr;    // 1 / M_ii
  double* wv = nullptr;      // 1 / V_i on free rows, 0 on fixed rows
  float* dinv32 = nullptr;
  float* wv32 = nullptr;
  double* work = nullptr;    // 10 n doubles
  double* part = nullptr;    // [PT_COUNT * kMaxG] reduction partials
  double* sc = nullptr;      // [S_COUNT] solver scalars
  int G = 1;
  // Sliced ELL copy of M for the Krylov products (ellW = 0: not built).  The rows
  // go in slices of kSlice (one wave each); inside a slice they are ordered by
  // their number of off-diagonals, longest first (sl_perm: the row, relative to the
  // slice, of each lane), and slot s of a slice holds the s-th off-diagonal (CSR
  // order) of every row that has one: lanes 0 .. cnt_s - 1, stored contiguously,
  // so no slot is padding.  sl_off[(ellW + 1) b + s], s = 0 .. ellW, is the first
  // entry of slot s of slice b; slices with rows longer than ellW list the ends of
  // slots ellW, ellW + 1, ... in sl_rem_off[sl_rem_ptr[b] .. sl_rem_ptr[b + 1]).
  // The diagonal (FP32: the row sum of the difference form) in diagM*, row order.
  int ellW = 0;
  int64_t ell_overflow = 0;    // entries in slots ellW and beyond
  bool use_ell = false;        // this solve's choice (fvm_solve_opts.spmv_format)
  int32_t* sl_off = nullptr;
  uint8_t* sl_perm = nullptr;
  int32_t* sl_rem_ptr = nullptr;
  int32_t* sl_rem_off = nullptr;
  int32_t* ell_col = nullptr;  // column of each entry, or (ell_off) its offset from the row
  int16_t* ell_off = nullptr;  // used instead of ell_col when every |column - row| < 2^15
  double* ellM = nullptr;
  float* ellM32 = nullptr;
  double* diagM = nullptr;
  float* diagM32 = nullptr;
  double* ellA = nullptr;      // the same entries of A itself (fvm_template_apply / _spmv)
  double* diagA = nullptr;
};

namespace {

// ----------------------------------------------------------------------------
// assembly
// ----------------------------------------------------------------------------
struct AsmArgs {
  int64_t N, T, B;
  int32_t kind;
  const int32_t* row_ptr;
  const int32_t* col;
  double* val;
  double* b;
  const int32_t* node_tri_start;
  const int32_t* node_tri;
  const int32_t* tri;       // [3][T] sorted
  const double* props;      // [24][T] sorted
  const double* cv;
  const double* x;
  const double* y;
  const uint8_t* is_vertex;
  const double* d_edge;     // [3][T] sorted
  const double* d_bedge;    // [2B] interleaved
  const int32_t* bedges;    // [2][B]
  const int32_t* badj;      // [B] sorted triangle
  const int32_t* nb_start;  // [N+1] boundary-edge endpoints per node
  const int32_t* nb;        // (edge << 1) | endpoint
  const int8_t* kindn;      // 0 none, 1 Dirichlet, 2 Dudt
  const int8_t* neu;        // [B] 1 for a Neumann edge
  const double* f_node;     // may be null
  const double* r_node;     // may be null
  const double* dir_val;    // may be null
  const double* dudt_val;   // may be null
  const double* neu_a;      // [2B] may be null
  uint8_t* fixed;
};

__device__ __forceinline__ void add_at(const AsmArgs& a, int k0, int k1, int c, double v) {
  for (int k = k0; k < k1; ++k)
    if (a.col[k] == c) {
      a.val[k] = a.val[k] + v;
      return;
    }
}

__device__ __forceinline__ void set_at(const AsmArgs& a, int k0, int k1, int c, double v) {
  for (int k = k0; k < k1; ++k)
    if (a.col[k] == c) {
      a.val[k] = v;
      return;
    }
}

// One thread per row r: the row's share of triangle_contributions!
// (abstract_templates.jl:73-99), of the boundary-edge contributions (:149-267, the
// j row divided by V_i as at :262; Neumann terms into b, :175-191), of
// linear_source_contributions! (linear_reaction_diffusion_equations.jl:115-125),
// then b and the identity rows: create_rhs_b / apply_steady_dirichlet_conditions!
// (:278-309), create_met_b! (mean_exit_time.jl:84-94), apply_dudt_conditions!
// (:127-135) and fix_missing_vertices! (:317-325).  Rows with a Dirichlet or Dudt
// condition receive no flux terms.
__global__ __launch_bounds__(256) void assemble_kernel(AsmArgs a) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.N) return;
  const int k0 = a.row_ptr[r], k1 = a.row_ptr[r + 1];
  for (int k = k0; k < k1; ++k) a.val[k] = 0.0;
  const int kd = a.kindn[r];
  const bool hascond = kd != 0;
  const bool vertex = a.is_vertex[r] != 0;
  double bneu = 0.0;
  if (vertex && !hascond) {
    const double Vr = a.cv[r];
    for (int q = a.node_tri_start[r]; q < a.node_tri_start[r + 1]; ++q) {
      const int w = a.node_tri[q];
      const int64_t s = w >> 2;
      const int v = w & 3;
      const int ijk[3] = {a.tri[s], a.tri[a.T + s], a.tri[2 * a.T + s]};
      double sc[6];
#pragma unroll
      for (int f = 0; f < 6; ++f) sc[f] = a.props[f * a.T + s];
      // CV edge e joins (ijk[e], ijk[e+1]): r is its first endpoint when e == v
      // and its second endpoint when e == v + 2 (mod 3).
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const bool first = e == v, second = e == (v + 2) % 3;
        if (!first && !second) continue;
        const double nx = a.props[(P_NX + 2 * e) * a.T + s], ny = a.props[(P_NX + 1 + 2 * e) * a.T + s];
        const double l = a.props[(P_L + e) * a.T + s];
        const double Dl = a.d_edge[e * a.T + s] * l;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double av = Dl * (sc[c] * nx + sc[3 + c] * ny);
          add_at(a, k0, k1, ijk[c], first ? av / Vr : -(av / Vr));
        }
      }
    }
    if (a.kind != 1) {
      for (int q = a.nb_start[r]; q < a.nb_start[r + 1]; ++q) {
        const int w = a.nb[q];
        const int e = w >> 1, ep = w & 1;
        const int i = a.bedges[e], j = a.bedges[a.B + e];
        const double px = a.x[i], py = a.y[i], qx = a.x[j], qy = a.y[j];
        const double lij = sqrt((qx - px) * (qx - px) + (qy - py) * (qy - py));
        const double nx = (qy - py) / lij, ny = -(qx - px) / lij;
        const double mijx = (px + qx) / 2, mijy = (py + qy) / 2;
        const double ell = sqrt((mijx - px) * (mijx - px) + (mijy - py) * (mijy - py));
        const double De = a.d_bedge[2 * e + ep];
        if (!a.neu[e]) {
          const int64_t s = a.badj[e];
          const int ijk[3] = {a.tri[s], a.tri[a.T + s], a.tri[2 * a.T + s]};
          const double Vi = a.cv[i];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double s1 = a.props[c * a.T + s], s2 = a.props[(3 + c) * a.T + s];
            add_at(a, k0, k1, ijk[c], De * ell * (s1 * nx + s2 * ny) / Vi);
          }
        } else if (a.neu_a) {
          bneu = bneu + De * a.neu_a[2 * e + ep] * ell / Vr;
        }
      }
    }
    if (a.r_node) add_at(a, k0, k1, (int)r, a.r_node[r]);
  }
  double bv = 0.0;
  bool fix = false;
  if (!vertex) {
    fix = true;
  } else if (a.kind == 0) {
    if (kd == 1) {
      bv = a.dir_val ? a.dir_val[r] : 0.0;
      fix = true;
    } else {
      bv = (a.f_node ? a.f_node[r] : 0.0) + bneu;
    }
  } else if (a.kind == 1) {
    if (kd == 1) fix = true;
    else bv = -1.0;
  } else {
    if (kd == 2) bv = a.dudt_val ? a.dudt_val[r] : 0.0;
    else if (kd == 0) bv = bneu;
  }
  if (fix) set_at(a, k0, k1, (int)r, 1.0);
  a.b[r] = bv;
  a.fixed[r] = fix;
}

// M = -V A on free rows and columns, identity on fixed rows; Jacobi inverse of M
// and the weights that turn a residual of M back into one of A.  The FP32 copy is
// kept in difference form: off-diagonal m_ij as they are, and in the diagonal slot
// the row sum s_i = sum_j m_ij (in FP64, then rounded), so that
// (M p)_i = sum_{j != i} m_ij (p_j - p_i) + s_i p_i.  The two forms are equal in
// exact arithmetic; in FP32 the direct form loses about eps |M| |p| to
// cancellation, which for a smooth p (a diffusion operator's row sums are near
// zero) is kappa(M) times the size of M p itself, while the difference form's
// rounding stays relative to the small differences.
__global__ __launch_bounds__(256) void scaled_kernel(int64_t N, const int32_t* __restrict__ row_ptr,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ valA,
                                                     const uint8_t* __restrict__ fixed,
                                                     const double* __restrict__ vol,
                                                     double* __restrict__ valM, float* __restrict__ valM32,
                                                     double* __restrict__ dinv, float* __restrict__ dinv32,
                                                     double* __restrict__ wv, float* __restrict__ wv32) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  double dM = 1.0, rsum = 0.0;
  const bool fr = fixed[r] != 0;
  for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
    const int c = col[k];
    double v;
    if (fr) v = c == r ? 1.0 : 0.0;
    else v = fixed[c] ? 0.0 : -vol[r] * valA[k];
    valM[k] = v;
    rsum += v;
    if (c == r) dM = v;
  }
  for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    valM32[k] = col[k] == r ? (float)rsum : (float)valM[k];
  const double di = dM != 0.0 ? 1.0 / dM : 1.0;
  dinv[r] = di;
  dinv32[r] = (float)di;
  const double w = fr || vol[r] == 0.0 ? 0.0 : 1.0 / vol[r];
  wv[r] = w;
  wv32[r] = (float)w;
}

// The rows of a solver kernel's workgroup: with a grid that is a multiple of the XCD
// count, each XCD takes contiguous ranges of rows (xcd_remap), the same in every
// solver kernel, so that a row's neighbours, gathered by nearby rows, and the vectors
// the previous kernel wrote for them tend to be in that XCD's own L2.
__device__ __forceinline__ int64_t solver_wid() {
  return gridDim.x % kXCD == 0 ? xcd_remap((int)blockIdx.x, (int)gridDim.x) : (int64_t)blockIdx.x;
}

// ----------------------------------------------------------------------------
// reductions
// ----------------------------------------------------------------------------
// A kernel leaves one partial sum per workgroup in part[]; the next kernel's
// workgroups each add the partials again, in the same fixed order, so every
// workgroup holds the same total without atomics or fences, and solver scalars
// never leave the device.  Scalars that later kernels need are stored by
// workgroup 0 in slots alternating with the iteration's parity.

// Sum over the workgroup (every thread gets the total): each wave folds its 64
// values by shuffles, then every thread adds the waves' sums in wave order, so the
// order is fixed and one barrier replaces a tree of them.
__device__ __forceinline__ double wg_sum(double v) {
  static_assert(kRT % 64 == 0, "whole waves");
  __shared__ double sh[kRT / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kRT / 64; ++w) t += sh[w];
  __syncthreads();  // sh is reused by the next call
  return t;
}

// Total of the previous kernel's partials part[0..G).
__device__ __forceinline__ double total_of(const double* __restrict__ part) {
  double s = 0.0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += kRT) s += part[i];
  return wg_sum(s);
}

// part[blockIdx.x] = this workgroup's sum of v
__device__ __forceinline__ void store_partial(double v, double* __restrict__ part) {
  const double t = wg_sum(v);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// The solver flag, read once per workgroup (a kernel may see it set by its own
// workgroup 0 only where every workgroup reaches the same decision anyway).
__device__ __forceinline__ bool wg_done(const double* sc) {
  __shared__ int f;
  if (threadIdx.x == 0) f = sc[S_DONE] != 0.0;
  __syncthreads();
  const bool d = f != 0;
  __syncthreads();
  return d;
}

// Sum over the kLPR lanes of a row group.
template <class V>
__device__ __forceinline__ V group_sum(V a) {
  a += __shfl_xor(a, 4, kLPR);
  a += __shfl_xor(a, 2, kLPR);
  a += __shfl_xor(a, 1, kLPR);
  return a;
}

// The FP32 operator is stored in difference form (scaled_kernel), the FP64 one directly.
template <class V>
constexpr bool kDiffForm = sizeof(V) == sizeof(float);

// Rows are taken kRPG at a time by each kLPR-lane group, with every load of the
// kRPG rows (row starts, then column / value pairs, then the gathers) issued
// before any is used: the SpMV is bound by memory latency, and this multiplies
// the loads in flight.  body(i, y_i, lane) runs for each row i < n with
// y_i = sum_k val[k] g(col[k]) valid in every lane of the group, or, for a matrix
// in difference form (DIFF), y_i = sum_{k: col[k] != i} val[k] (g(col[k]) - g(i))
// + val[diagonal] g(i).
template <bool DIFF, class V, class Gather, class Body>
__device__ __forceinline__ void for_rows(int64_t n, const int32_t* __restrict__ rp,
                                         const int32_t* __restrict__ col, const V* __restrict__ val,
                                         Gather g, Body body) {
  const int lane = threadIdx.x % kLPR;
  const int64_t step = (int64_t)gridDim.x * kRowsPerBlock * kRPG;
  for (int64_t i0 = ((int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / kLPR) * kRPG; i0 < n; i0 += step) {
    int rs[kRPG + 1];
#pragma unroll
    for (int j = 0; j <= kRPG; ++j) rs[j] = rp[i0 + j < n ? i0 + j : n];
    int c[kRPG];
    V v[kRPG];
#pragma unroll
    for (int j = 0; j < kRPG; ++j) {
      const int k = rs[j] + lane;
      c[j] = -1;
      v[j] = 0;
      if (k < rs[j + 1]) {
        c[j] = col[k];
        v[j] = val[k];
      }
    }
    V gi[kRPG];
    if constexpr (DIFF) {
#pragma unroll
      for (int j = 0; j < kRPG; ++j) gi[j] = i0 + j < n ? (V)g((int)(i0 + j)) : (V)0;
    }
    auto term = [&](int j, int cc) -> V {
      if constexpr (DIFF) return cc == (int)(i0 + j) ? gi[j] : (V)(g(cc) - gi[j]);
      else return g(cc);
    };
    V acc[kRPG];
#pragma unroll
    for (int j = 0; j < kRPG; ++j) {
      acc[j] = 0;
      if (c[j] >= 0) acc[j] = v[j] * term(j, c[j]);
    }
#pragma unroll
    for (int j = 0; j < kRPG; ++j)
      for (int k = rs[j] + lane + kLPR; k < rs[j + 1]; k += kLPR) acc[j] += val[k] * term(j, col[k]);
#pragma unroll
    for (int j = 0; j < kRPG; ++j) {
      const V y = group_sum(acc[j]);
      if (i0 + j < n) body(i0 + j, y, lane);
    }
  }
}

// ----------------------------------------------------------------------------
// CG on M (precision V for the matrix, vectors and arithmetic)
// ----------------------------------------------------------------------------
// start from d = 0 with residual r: z = D^-1 r, p = 0; partials of r.z and |r|_A^2
// A reliable update passes keep = the current direction instead (p == nullptr):
// the direction and the iteration count are kept, and the partials of r.keep
// (slot PT_TT) let the next iteration make the direction orthogonal to the new r.
template <class V>
__global__ __launch_bounds__(kRT) void cg_init(int64_t n, const V* __restrict__ r, const V* __restrict__ dinv,
                                               const V* __restrict__ wv, V* __restrict__ d, V* __restrict__ z,
                                               V* __restrict__ p, const V* __restrict__ keep, double* sc,
                                               double* part) {
  double rz = 0.0, wr = 0.0, rk = 0.0;
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT) {
    const V ri = r[i];
    const V zi = dinv[i] * ri;
    z[i] = zi;
    d[i] = 0;
    if (p) p[i] = 0;
    if (keep) rk += (double)ri * (double)keep[i];
    rz += (double)ri * (double)zi;
    const double w = (double)ri * (double)wv[i];
    wr += w * w;
  }
  store_partial(rz, part + PT_RZ * kMaxG);
  store_partial(wr, part + PT_WR * kMaxG);
  if (keep) store_partial(rk, part + PT_TT * kMaxG);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_DONE] = 0.0;
    sc[S_FLAG] = 0.0;
    if (p) sc[S_ITER] = 0.0;
  }
}

// Iteration k, part 1: stop test on |r|_A^2 (the tolerance is fixed at k = 0 as
// max(irt2 |r0|_A^2, tabs2)), beta = rz_k / rz_{k-1}; p = z + beta p_old with the
// neighbours' p formed on the fly from z and p_old; q = M p; partials of p.q.
// First iteration after a reliable update: the replaced r is not exactly
// orthogonal to p_old, as CG assumes, so p_old first loses its component along z
// (proj 1): p_old - g z with g = r.p_old / r.z, i.e. p = (1 - beta g) z + beta p_old;
// or the direction restarts, p = z (proj 2).
// The stop test and the direction coefficients shared by both forms of the
// direction kernel: false when the solve has stopped (the kernel returns),
// otherwise p = va z + vb p_old.
template <class V>
__device__ __forceinline__ bool cg_dir_coeffs(int k, int proj, double irt2, double tabs2, double* sc,
                                              double* part, V* va, V* vb) {
  if (wg_done(sc)) return false;  // launches after the stop leave S_ITER at the stopping iteration
  const double rz = total_of(part + PT_RZ * kMaxG);
  const double wr = total_of(part + PT_WR * kMaxG);
  const double tol2 = k == 0 ? fmax(irt2 * wr, tabs2) : sc[S_TOL2];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!(wr > tol2)) {  // converged (or NaN): every workgroup decides the same
    if (lead) {
      sc[S_DONE] = 1.0;
      sc[S_ITER] = (double)k;
      sc[S_WR] = wr;
    }
    return false;
  }
  const double beta = k == 0 || proj == 2 ? 0.0 : rz / sc[S_RZ + ((k - 1) & 1)];
  if (lead) {
    sc[S_RZ + (k & 1)] = rz;
    sc[S_WR] = wr;
    if (k == 0) sc[S_TOL2] = tol2;
  }
  *vb = (V)beta;
  *va = proj == 1 ? (V)(1.0 - beta * (total_of(part + PT_TT * kMaxG) / rz)) : (V)1;
  return true;
}

template <class V>
__global__ __launch_bounds__(kRT) void cg_dir_spmv(int64_t n, int k, int proj, double irt2, double tabs2,
                                                   const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col, const V* __restrict__ val,
                                                   const V* __restrict__ z, const V* __restrict__ pold,
                                                   V* __restrict__ pnew, V* __restrict__ q, double* sc,
                                                   double* part) {
  V va, vb;
  if (!cg_dir_coeffs<V>(k, proj, irt2, tabs2, sc, part, &va, &vb)) return;
  double pq = 0.0;
  auto dir = [&](int c) { return va * z[c] + vb * pold[c]; };
  for_rows<kDiffForm<V>>(n, rp, col, val, dir, [&](int64_t i, V qi, int lane) {
    if (lane == 0) {
      const V pi = dir((int)i);
      pnew[i] = pi;
      q[i] = qi;
      pq += (double)pi * (double)qi;
    }
  });
  store_partial(pq, part + PT_PQ * kMaxG);
}

// The column of an ELL entry: stored as is (int32) or as its offset from the row
// (int16, when the whole operator allows it: two bytes fewer per entry).
__device__ __forceinline__ int ell_column(int32_t c, int64_t) { return c; }
__device__ __forceinline__ int ell_column(int16_t d, int64_t i) { return (int)i + (int)d; }

// The index arrays of the sliced ELL form (fvm_template: sl_off, sl_perm,
// sl_rem_ptr, sl_rem_off).
struct SellIdx {
  const int32_t* off;
  const uint8_t* perm;
  const int32_t* rem_ptr;
  const int32_t* rem_off;
};

// One slice of the sliced ELL form, taken by one wave: lane l computes row
// i = 64 b + perm[64 b + l], its W first slots' column / value loads issued together
// (slot s's entries of the lanes that have one are contiguous: coalesced, and no
// entry is padding), then their gathers; rows longer than W continue slot by slot.
// body(i, g(i), y_i) with y_i the off-diagonal terms in CSR order (v (g(c) - g(i))
// in the difference form, v g(c) otherwise), then diag_i g(i).  The loads and
// gathers are not predicated, so that all of them are in flight together: a lane
// past its row's last slot reads the entry arrays' kSlice-entry tail or another
// lane's entry and uses column i and value 0 instead (an exact zero term, as in a
// padded ELL row).  The slot starts are wave-uniform: scalar loads.
template <bool DIFF, int W, class V, class CI, class Gather, class Body>
__device__ __forceinline__ void sell_slice(int64_t n, int64_t b, const SellIdx& S, const CI* __restrict__ ecol,
                                           const V* __restrict__ eval, const V* __restrict__ diag, Gather g,
                                           Body body) {
  using cint = const __attribute__((address_space(4))) int32_t;
  const int lane = threadIdx.x & (kSlice - 1);
  const int64_t r0 = b * kSlice;
  int o[W + 1];
#pragma unroll
  for (int s = 0; s <= W; ++s) o[s] = ((cint*)S.off)[(W + 1) * b + s];
  const int q0 = ((cint*)S.rem_ptr)[b], q1 = ((cint*)S.rem_ptr)[b + 1];
  if (lane >= n - r0) return;
  const int64_t i = r0 + S.perm[r0 + lane];
  int c[W];
  V v[W];
#pragma unroll
  for (int s = 0; s < W; ++s) {
    const int e = o[s] + lane;
    const bool on = e < o[s + 1];
    const int cs = ell_column(ecol[e], i);
    const V vs = eval[e];
    c[s] = on ? cs : (int)i;
    v[s] = on ? vs : (V)0;
  }
  const V gi = g((int)i);
  V acc = 0;
#pragma unroll
  for (int s = 0; s < W; ++s) acc += v[s] * (DIFF ? g(c[s]) - gi : g(c[s]));
  for (int q = q0, st = o[W]; q < q1; ++q) {  // slots W, W + 1, ... (rows longer than W)
    const int en = ((cint*)S.rem_off)[q], e = st + lane;
    if (e < en) {
      const int cc = ell_column(ecol[e], i);
      acc += eval[e] * (DIFF ? g(cc) - gi : g(cc));
    }
    st = en;
  }
  acc += diag[i] * gi;
  body(i, gi, acc);
}

// Every slice of the sliced ELL form, one wave per slice, grid-stride.
template <bool DIFF, int W, class V, class CI, class Gather, class Body>
__device__ __forceinline__ void sell_rows(int64_t n, const SellIdx& S, const CI* __restrict__ ecol,
                                          const V* __restrict__ eval, const V* __restrict__ diag, Gather g,
                                          Body body) {
  constexpr int kWaves = kRT / kSlice;
  const int64_t ns = (n + kSlice - 1) / kSlice;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kSlice);
  for (int64_t b = solver_wid() * kWaves + w; b < ns; b += (int64_t)gridDim.x * kWaves)
    sell_slice<DIFF, W>(n, b, S, ecol, eval, diag, g, body);
}

// The CG direction iteration on the sliced ELL copy of M.
template <class V, int W, class CI>
__global__ __launch_bounds__(kRT) void cg_dir_ell(int64_t n, int k, int proj, double irt2, double tabs2,
                                                  SellIdx S, const CI* __restrict__ ecol,
                                                  const V* __restrict__ eval, const V* __restrict__ diag,
                                                  const V* __restrict__ z, const V* __restrict__ pold,
                                                  V* __restrict__ pnew, V* __restrict__ q, double* sc,
                                                  double* part) {
  V va, vb;
  if (!cg_dir_coeffs<V>(k, proj, irt2, tabs2, sc, part, &va, &vb)) return;
  double pq = 0.0;
  auto dir = [&](int c) { return va * z[c] + vb * pold[c]; };
  sell_rows<kDiffForm<V>, W>(n, S, ecol, eval, diag, dir, [&](int64_t i, V pi, V qi) {
    pnew[i] = pi;
    q[i] = qi;
    pq += (double)pi * (double)qi;
  });
  store_partial(pq, part + PT_PQ * kMaxG);
}

// out[j] = src[j] >= 0 ? v[src[j]] : 0, in FP64 and (given o32) in FP32: the ELL values
__global__ __launch_bounds__(256) void gather_vals(int64_t m, const int32_t* __restrict__ src,
                                                   const double* __restrict__ v64, const float* __restrict__ v32,
                                                   double* __restrict__ o64, float* __restrict__ o32) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const int k = src[j];
    o64[j] = k >= 0 ? v64[k] : 0.0;
    if (o32) o32[j] = k >= 0 ? v32[k] : 0.0f;
  }
}

// Iteration k, part 2: alpha = rz_k / (p.q); d += alpha p (or, given x64, the
// FP64 solution itself: x64 += alpha p, exact products of the FP32 factors);
// r -= alpha q; z = D^-1 r; partials of r.z and |r|_A^2.
template <class V>
__global__ __launch_bounds__(kRT) void cg_update(int64_t n, int k, V* __restrict__ d, double* __restrict__ x64,
                                                 V* __restrict__ r,
                                                 const V* __restrict__ p, const V* __restrict__ q,
                                                 const V* __restrict__ dinv, const V* __restrict__ wv,
                                                 V* __restrict__ z, double* sc, double* part) {
  if (wg_done(sc)) return;
  const double pq = total_of(part + PT_PQ * kMaxG);
  if (!(pq > 0.0)) {  // M is not positive definite on this direction: stop
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc[S_FLAG] = 2.0;
      sc[S_DONE] = 1.0;
      sc[S_ITER] = (double)k;
    }
    return;
  }
  const V alpha = (V)(sc[S_RZ + (k & 1)] / pq);
  double rz = 0.0, wr = 0.0;
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT) {
    if (x64) x64[i] = x64[i] + (double)alpha * (double)p[i];
    else d[i] = d[i] + alpha * p[i];
    const V ri = r[i] - alpha * q[i];
    r[i] = ri;
    const V zi = dinv[i] * ri;
    z[i] = zi;
    rz += (double)ri * (double)zi;
    const double w = (double)ri * (double)wv[i];
    wr += w * w;
  }
  store_partial(rz, part + PT_RZ * kMaxG);
  store_partial(wr, part + PT_WR * kMaxG);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[S_ITER] = (double)(k + 1);
}

// ----------------------------------------------------------------------------
// BiCGStab on M, FP64, Jacobi right preconditioning
// ----------------------------------------------------------------------------
// d = 0, rhat = r, p = v = 0; partials of rhat.r and |r|_A^2
__global__ __launch_bounds__(kRT) void bi_init(int64_t n, const double* __restrict__ r,
                                               const double* __restrict__ wv, double* __restrict__ d,
                                               double* __restrict__ rhat, double* __restrict__ p,
                                               double* __restrict__ v, double* sc, double* part) {
  double rho = 0.0, wr = 0.0;
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT) {
    const double ri = r[i];
    rhat[i] = ri;
    d[i] = 0.0;
    p[i] = 0.0;
    v[i] = 0.0;
    rho += ri * ri;
    const double w = ri * wv[i];
    wr += w * w;
  }
  store_partial(rho, part + PT_RZ * kMaxG);
  store_partial(wr, part + PT_WR * kMaxG);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_DONE] = 0.0;
    sc[S_FLAG] = 0.0;
    sc[S_ITER] = 0.0;
  }
}

// stop test; beta = (rho_k / rho_{k-1}) (alpha_{k-1} / omega_{k-1});
// p = r + beta (p - omega_{k-1} v); phat = D^-1 p
__global__ __launch_bounds__(kRT) void bi_dir(int64_t n, int k, double tabs2, const double* __restrict__ r,
                                              double* __restrict__ p, const double* __restrict__ v,
                                              const double* __restrict__ dinv, double* __restrict__ phat,
                                              double* sc, const double* part) {
  if (wg_done(sc)) return;
  const double rho = total_of(part + PT_RZ * kMaxG);
  const double wr = total_of(part + PT_WR * kMaxG);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  double beta = 0.0, omega = 0.0;
  bool stop = !(wr > tabs2);
  if (!stop && k > 0) {
    const double rho0 = sc[S_RHO + ((k - 1) & 1)];
    const double alpha0 = sc[S_ALPHA + ((k - 1) & 1)];
    omega = sc[S_OMEGA + ((k - 1) & 1)];
    if (rho == 0.0 || omega == 0.0 || rho0 == 0.0) {
      stop = true;
      if (lead) sc[S_FLAG] = 2.0;
    } else {
      beta = (rho / rho0) * (alpha0 / omega);
    }
  }
  if (stop) {
    if (lead) {
      sc[S_DONE] = 1.0;
      sc[S_ITER] = (double)k;
      sc[S_WR] = wr;
    }
    return;
  }
  if (lead) {
    sc[S_RHO + (k & 1)] = rho;
    sc[S_WR] = wr;
  }
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT) {
    const double pi = r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    phat[i] = dinv[i] * pi;
  }
}

// y = M x with partials of a.y (slot A) and, for MODE 1, of y.y (slot B)
template <int MODE>
__global__ __launch_bounds__(kRT) void bi_spmv(int64_t n, const int32_t* __restrict__ rp,
                                               const int32_t* __restrict__ col, const double* __restrict__ val,
                                               const double* __restrict__ x, double* __restrict__ y,
                                               const double* __restrict__ a, const double* sc, double* part) {
  if (wg_done(sc)) return;
  double ay = 0.0, yy = 0.0;
  for_rows<false>(n, rp, col, val, [&](int c) { return x[c]; }, [&](int64_t i, double yi, int lane) {
    if (lane == 0) {
      y[i] = yi;
      ay += a[i] * yi;
      if (MODE == 1) yy += yi * yi;
    }
  });
  store_partial(ay, part + PT_PQ * kMaxG);
  if (MODE == 1) store_partial(yy, part + PT_TT * kMaxG);
}

// bi_spmv on the sliced ELL copy of M (FP64).
template <int MODE, int W, class CI>
__global__ __launch_bounds__(kRT) void bi_spmv_ell(int64_t n, SellIdx S, const CI* __restrict__ ecol,
                                                   const double* __restrict__ eval, const double* __restrict__ diag,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   const double* __restrict__ a, const double* sc, double* part) {
  if (wg_done(sc)) return;
  double ay = 0.0, yy = 0.0;
  sell_rows<false, W>(n, S, ecol, eval, diag, [&](int c) { return x[c]; },
                     [&](int64_t i, double, double yi) {
                       y[i] = yi;
                       ay += a[i] * yi;
                       if (MODE == 1) yy += yi * yi;
                     });
  store_partial(ay, part + PT_PQ * kMaxG);
  if (MODE == 1) store_partial(yy, part + PT_TT * kMaxG);
}

// alpha = rho_k / (rhat.v); s = r - alpha v (into r); shat = D^-1 s
__global__ __launch_bounds__(kRT) void bi_s(int64_t n, int k, double* __restrict__ r,
                                            const double* __restrict__ v, const double* __restrict__ dinv,
                                            double* __restrict__ shat, double* sc, const double* part) {
  if (wg_done(sc)) return;
  const double rv = total_of(part + PT_PQ * kMaxG);
  if (rv == 0.0 || rv != rv) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc[S_FLAG] = 2.0;
      sc[S_DONE] = 1.0;
      sc[S_ITER] = (double)k;
    }
    return;
  }
  const double alpha = sc[S_RHO + (k & 1)] / rv;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[S_ALPHA + (k & 1)] = alpha;
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT) {
    const double si = r[i] - alpha * v[i];
    r[i] = si;
    shat[i] = dinv[i] * si;
  }
}

// omega = (t.s) / (t.t); d += alpha phat + omega shat; r = s - omega t;
// partials of rhat.r and |r|_A^2
__global__ __launch_bounds__(kRT) void bi_update(int64_t n, int k, double* __restrict__ d, double* __restrict__ r,
                                                 const double* __restrict__ phat,
                                                 const double* __restrict__ shat, const double* __restrict__ t,
                                                 const double* __restrict__ rhat, const double* __restrict__ wv,
                                                 double* sc, double* part) {
  if (wg_done(sc)) return;
  const double ts = total_of(part + PT_PQ * kMaxG);
  const double tt = total_of(part + PT_TT * kMaxG);
  const double omega = tt != 0.0 ? ts / tt : 0.0;
  const double alpha = sc[S_ALPHA + (k & 1)];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_OMEGA + (k & 1)] = omega;
    sc[S_ITER] = (double)(k + 1);
  }
  double rho = 0.0, wr = 0.0;
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT) {
    d[i] = d[i] + alpha * phat[i] + omega * shat[i];
    const double ri = r[i] - omega * t[i];
    r[i] = ri;
    rho += rhat[i] * ri;
    const double w = ri * wv[i];
    wr += w * w;
  }
  store_partial(rho, part + PT_RZ * kMaxG);
  store_partial(wr, part + PT_WR * kMaxG);
}

// ----------------------------------------------------------------------------
// residuals and small vector kernels
// ----------------------------------------------------------------------------
// ra = b - A x (FP64); r = -V ra on free rows, 0 on fixed rows (precision V);
// partials of |ra|^2
template <class V>
__global__ __launch_bounds__(kRT) void true_residual(int64_t n, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ val, const double* __restrict__ b,
                                                     const double* __restrict__ x, const uint8_t* __restrict__ fixed,
                                                     const double* __restrict__ vol, V* __restrict__ r, double* part) {
  double acc = 0.0;
  for_rows<false>(n, rp, col, val, [&](int c) { return x[c]; }, [&](int64_t i, double ax, int lane) {
    if (lane == 0) {
      const double ra = b[i] - ax;
      if (r) r[i] = fixed[i] ? (V)0 : (V)(-vol[i] * ra);
      acc += ra * ra;
    }
  });
  store_partial(acc, part + PT_RES * kMaxG);
}

__global__ __launch_bounds__(kRT) void norm2_kernel(int64_t n, const double* __restrict__ x, double* part) {
  double acc = 0.0;
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT)
    acc += x[i] * x[i];
  store_partial(acc, part + PT_RES * kMaxG);
}

// sc[slot] = total of the G partials in slot PT_RES (one workgroup)
__global__ __launch_bounds__(kRT) void finish_sum(int G, const double* __restrict__ part, double* sc, int slot) {
  double s = 0.0;
  for (int i = threadIdx.x; i < G; i += kRT) s += part[PT_RES * kMaxG + i];
  const double t = wg_sum(s);
  if (threadIdx.x == 0) sc[slot] = t;
}

// x = 0 (unless keep) with x_i = b_i on fixed rows
__global__ __launch_bounds__(kRT) void init_x(int64_t n, const uint8_t* __restrict__ fixed,
                                              const double* __restrict__ b, double* __restrict__ x, int keep) {
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT) {
    if (fixed[i]) x[i] = b[i];
    else if (!keep) x[i] = 0.0;
  }
}

template <class V>
__global__ __launch_bounds__(kRT) void add_correction(int64_t n, const V* __restrict__ d, double* __restrict__ x) {
  for (int64_t i = solver_wid() * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT)
    x[i] = x[i] + (double)d[i];
}

// y = A x (+ b) in precision V, one row group per row
template <class V>
__global__ __launch_bounds__(kRT) void spmv_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col, const V* __restrict__ val,
                                                   const double* __restrict__ x, const double* __restrict__ b,
                                                   double* __restrict__ y) {
  for_rows<kDiffForm<V>>(n, rp, col, val, [&](int c) { return (V)x[c]; }, [&](int64_t i, V yi, int lane) {
    if (lane == 0) y[i] = b ? (double)yi + b[i] : (double)yi;
  });
}

// y = A x (+ b) on a sliced ELL copy (precision V, the difference form for FP32)
template <class V, int W, class CI>
__global__ __launch_bounds__(kRT) void spmv_ell_kernel(int64_t n, SellIdx S, const CI* __restrict__ ecol,
                                                       const V* __restrict__ eval, const V* __restrict__ diag,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       double* __restrict__ y) {
  sell_rows<kDiffForm<V>, W>(n, S, ecol, eval, diag, [&](int c) { return (V)x[c]; },
                            [&](int64_t i, V, V yi) { y[i] = b ? (double)yi + b[i] : (double)yi; });
}

// ----------------------------------------------------------------------------
// Tsit5 on a small ODE template in one workgroup
// ----------------------------------------------------------------------------
// The adaptive loop of fvm_ode_solve (fvm_ode.hip) for du/dt = A u + b when the
// state is small: one workgroup of 1024 threads runs every step (stages, products
// on the ELL copy of A, error norm, step control, saveat, first same as last) with
// barriers between the phases, so the integration is one launch instead of ~13
// launches and a host synchronisation per step.  Each element's arithmetic is the
// multi-kernel path's: the same row products (sell_slice's order), the same linear
// combinations (terms left to right), and the error norm's sum in the fixed order
// of its 512-workgroup grid, emulated here block by block; the step controller's
// power is fvm_ctrl.hpp's, the host loops' own: the same states bit for bit.
constexpr int kSmallThreads = 1024;
// One workgroup pays off below this many unknowns: a step costs about 16 us + 17 ns
// per unknown on one CU (2,500 unknowns: 60 us; 10,000: 190 us) against about 80 us
// for the launches of the multi-kernel step on the whole chip.
constexpr int64_t kSmallTsit5Max = 3600;
constexpr int kRedBlocksE = 512, kRedThreadsE = 256;  // the grid whose summation order is reproduced

struct SmallTsit5 {
  int64_t n;
  SellIdx S;            // the sliced ELL copy of A
  const int32_t* ecol;  // one of ecol / eoff
  const int16_t* eoff;
  const double* eval;
  const double* diag;
  const double* b;
  double* k[7];  // the stages (k[0] = f(u) on entry)
  double* utmp;
  double* u;     // the state on entry
  double* unew;
  double* u_out;
  double t0, tf, dt;
  double abstol, reltol;
  int64_t maxiters, it0;
  int32_t fsal;
  int64_t n_save, ks0;
  const double* save_t;  // device copy
  double* saves;
  double* stats;  // [0] t, [1] naccept, [2] nreject, [3] nf, [4] nsaved, [5] retcode
  double A[7][6], C[7], Bt[7];
};

// out = A x + b over the state, the workgroup's 16 waves taking the slices of the
// sliced ELL copy in turn.
template <int W, bool OFF>
__device__ __forceinline__ void small_apply(const SmallTsit5& a, const double* x, double* out) {
  const int64_t ns = (a.n + kSlice - 1) / kSlice;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kSlice);
  auto g = [&](int c) { return x[c]; };
  auto body = [&](int64_t i, double, double acc) { out[i] = acc + a.b[i]; };
  for (int64_t b = w; b < ns; b += kSmallThreads / kSlice) {
    if (OFF)
      sell_slice<false, W>(a.n, b, a.S, a.eoff, a.eval, a.diag, g, body);
    else
      sell_slice<false, W>(a.n, b, a.S, a.ecol, a.eval, a.diag, g, body);
  }
}

// The pairwise tree of a 256-thread reduction, sh[t] += sh[t + s] for s = 128, ..., 1,
// for each group of 256 threads of the workgroup at once: the two levels that pair
// waves go through LDS, the six inside a wave through shuffles (lane l adds lane
// l + s, the same pairs in the same order).  Thread 0 of each group returns its sum.
__device__ __forceinline__ double tree256(double* sh, double x) {
  const int tid = threadIdx.x, lt = tid % kRedThreadsE;
  sh[tid] = x;
  __syncthreads();
  if (lt < 128) sh[tid] = sh[tid] + sh[tid + 128];
  __syncthreads();
  double v = 0.0;
  if (lt < 64) {
    v = sh[tid] + sh[tid + 64];
    for (int s = 32; s > 0; s >>= 1) v += __shfl_down(v, s, 64);
  }
  __syncthreads();  // sh is free again
  return v;
}

template <int W, bool OFF>
__global__ __launch_bounds__(kSmallThreads) void tsit5_small_kernel(SmallTsit5 a) {
  __shared__ double sh[kSmallThreads];
  __shared__ double part[kRedBlocksE];
  const int tid = threadIdx.x;
  const int64_t n = a.n;
  double* k[7];
  for (int j = 0; j < 7; ++j) k[j] = a.k[j];
  double* utmp = a.utmp;
  double* u = a.u;
  double* unew = a.unew;
  double t = a.t0, dt = a.dt, qold = 1e-4;
  int64_t it = a.it0, ks = a.ks0, naccept = 0, nreject = 0, nf = 0;
  int retcode = FVM_ODE_SUCCESS;
  const int nblk = (int)((n + kRedThreadsE - 1) / kRedThreadsE);  // nonempty blocks of the emulated grid
  while (t < a.tf && it < a.maxiters) {
    ++it;
    dt = fmin(dt, a.tf - t);
    for (int j = 1; j < 7; ++j) {
      double c[7];
      c[0] = 1.0;
      for (int i = 0; i < j; ++i) c[i + 1] = dt * a.A[j][i];
      double* dst = j == 6 ? unew : utmp;
      for (int64_t e = tid; e < n; e += kSmallThreads) {
        double v = c[0] * u[e];
        for (int i = 0; i < j; ++i) v = v + c[i + 1] * k[i][e];
        dst[e] = v;
      }
      __syncthreads();
      small_apply<W, OFF>(a, dst, k[j]);
      ++nf;
      __syncthreads();
    }
    // the error norm: per element r = (sum_j dt bt_j k_j) / (abstol + reltol max(|u|, |unew|)),
    // summed as the 512 x 256 grid of the multi-kernel path sums it
    double cb[7];
    for (int j = 0; j < 7; ++j) cb[j] = dt * a.Bt[j];
    for (int b0 = 0; b0 < nblk; b0 += kSmallThreads / kRedThreadsE) {
      const int blk = b0 + tid / kRedThreadsE, lt = tid % kRedThreadsE;
      const int64_t e = (int64_t)blk * kRedThreadsE + lt;
      double acc = 0.0;
      if (blk < nblk && e < n) {
        double v = cb[0] * k[0][e];
        for (int j = 1; j < 7; ++j) v = v + cb[j] * k[j][e];
        const double sc = a.abstol + a.reltol * fmax(fabs(u[e]), fabs(unew[e]));
        const double r = v / sc;
        acc += r * r;
      }
      acc = tree256(sh, acc);
      if (lt == 0 && blk < nblk) part[blk] = acc;
    }
    for (int p = nblk + tid; p < kRedBlocksE; p += kSmallThreads) part[p] = 0.0;
    __syncthreads();
    double acc = 0.0;
    if (tid < kRedThreadsE)
      for (int i = tid; i < kRedBlocksE; i += kRedThreadsE) acc += part[i];
    acc = tree256(sh, acc);
    if (tid == 0) sh[0] = acc;
    __syncthreads();
    const double EEst = sqrt(sh[0] / (double)n);
    __syncthreads();  // every thread has read sh[0]
    if (!isfinite(EEst)) {
      retcode = FVM_ODE_UNSTABLE;
      break;
    }
    if (EEst <= 1.0) {
      const double tn = t + dt;
      while (ks < a.n_save && a.save_t[ks] <= tn + 1e-12 * fmax(1.0, fabs(tn))) {
        const double th = fmin(1.0, fmax(0.0, (a.save_t[ks] - t) / dt));
        const double t2 = th * th, t3 = t2 * th;
        const double h00 = 2 * t3 - 3 * t2 + 1;
        const double h10 = t3 - 2 * t2 + th;
        const double h01 = -2 * t3 + 3 * t2;
        const double h11 = t3 - t2;
        const double c0 = h00, c1 = h10 * dt, c2 = h01, c3 = h11 * dt;
        double* out = a.saves + ks * n;
        for (int64_t e = tid; e < n; e += kSmallThreads) {
          double v = c0 * u[e];
          v = v + c1 * k[0][e];
          v = v + c2 * unew[e];
          v = v + c3 * k[6][e];
          out[e] = v;
        }
        ++ks;
      }
      double* sw = u;
      u = unew;
      unew = sw;
      t = tn;
      ++naccept;
      if (a.fsal) {
        double* k0 = k[0];
        k[0] = k[6];
        k[6] = k0;
      } else {
        __syncthreads();
        small_apply<W, OFF>(a, u, k[0]);
        ++nf;
      }
      double q = fvm::ctrl_pow(fmax(EEst, 1e-16), 7.0 / 50.0) / fvm::ctrl_pow(qold, 2.0 / 25.0);
      q = fmin(5.0, fmax(0.1, q / 0.9));
      dt = dt / q;
      qold = fmax(EEst, 1e-4);
    } else {
      ++nreject;
      const double q = fmin(5.0, fvm::ctrl_pow(fmax(EEst, 1e-16), 7.0 / 50.0) / 0.9);
      dt = dt / q;
    }
    __syncthreads();
  }
  if (retcode == FVM_ODE_SUCCESS && !(t >= a.tf)) retcode = FVM_ODE_MAXITERS;
  for (int64_t e = tid; e < n; e += kSmallThreads) a.u_out[e] = u[e];
  if (tid == 0) {
    a.stats[0] = t;
    a.stats[1] = (double)naccept;
    a.stats[2] = (double)nreject;
    a.stats[3] = (double)nf;
    a.stats[4] = (double)ks;
    a.stats[5] = (double)retcode;
  }
}

// ----------------------------------------------------------------------------
// host helpers
// ----------------------------------------------------------------------------
template <class T>
int dalloc(T** dst, int64_t n) {
  *dst = nullptr;
  FVM_HIP(hipMalloc((void**)dst, sizeof(T) * (size_t)(n > 0 ? n : 1)));
  return FVM_OK;
}

template <class T>
int dupload(T** dst, const T* src, int64_t n) {
  int rc = dalloc(dst, n);
  if (rc) return rc;
  if (n > 0) FVM_HIP(hipMemcpy(*dst, src, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
  return FVM_OK;
}

void free_tmpl(fvm_template* t) {
  if (!t) return;
  void* p[] = {t->row_ptr, t->col, t->val, t->valM, t->valM32, t->b, t->fixed, t->vol, t->dinv,
               t->wv, t->dinv32, t->wv32, t->work, t->part, t->sc, t->sl_off, t->sl_perm, t->sl_rem_ptr,
               t->sl_rem_off, t->ell_col, t->ell_off, t->ellM, t->ellM32, t->diagM, t->diagM32, t->ellA,
               t->diagA};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete t;
}

int read_scalars(const fvm_template* t, double* h, hipStream_t s) {
  FVM_HIP(hipGetLastError());
  FVM_HIP(hipMemcpyAsync(h, t->sc, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s));
  FVM_HIP(hipStreamSynchronize(s));
  return FVM_OK;
}

inline dim3 grid(const fvm_template* t) { return dim3((unsigned)t->G); }

template <class V>
struct EllVals;
template <>
struct EllVals<double> {
  static const double* ell(const fvm_template* t) { return t->ellM; }
  static const double* diag(const fvm_template* t) { return t->diagM; }
};
template <>
struct EllVals<float> {
  static const float* ell(const fvm_template* t) { return t->ellM32; }
  static const float* diag(const fvm_template* t) { return t->diagM32; }
};

inline SellIdx sell_idx(const fvm_template* t) { return SellIdx{t->sl_off, t->sl_perm, t->sl_rem_ptr, t->sl_rem_off}; }

// f(std::integral_constant<int, W>) for the template's unrolled ELL slots
template <class F>
void by_width(int W, F&& f) {
  static_assert(kEllMin == 4 && kEllMax == 10, "one case per ELL width");
  switch (W) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 9: f(std::integral_constant<int, 9>{}); break;
    default: f(std::integral_constant<int, 10>{}); break;
  }
}

// f(width tag, column array) for the template's sliced ELL copy: its unrolled slots
// as a constant and its columns as int32 columns or int16 offsets
template <class P>
using ci_t = std::remove_cv_t<std::remove_pointer_t<P>>;
template <class F>
void by_ell(const fvm_template* t, F&& f) {
  by_width(t->ellW, [&](auto w) {
    if (t->ell_off)
      f(w, (const int16_t*)t->ell_off);
    else
      f(w, (const int32_t*)t->ell_col);
  });
}

// One CG direction step (p, q = M p, partials of p.q) on this solve's form of M;
// valM is the CSR value array of the precision V.
template <class V>
void launch_dir(const fvm_template* t, const V* valM, int k, int proj, double irt2, double tabs2, const V* z,
                const V* pold, V* pnew, V* q, hipStream_t s) {
  if (!t->use_ell) {
    hipLaunchKernelGGL((cg_dir_spmv<V>), grid(t), dim3(kRT), 0, s, t->n, k, proj, irt2, tabs2, t->row_ptr,
                       t->col, valM, z, pold, pnew, q, t->sc, t->part);
    return;
  }
  by_ell(t, [&](auto w, auto ec) {
    hipLaunchKernelGGL((cg_dir_ell<V, decltype(w)::value, ci_t<decltype(ec)>>), grid(t), dim3(kRT), 0, s, t->n, k,
                       proj, irt2, tabs2, sell_idx(t), ec, EllVals<V>::ell(t), EllVals<V>::diag(t), z, pold, pnew,
                       q, t->sc, t->part);
  });
}

// y = M x with the partials of a.y (and y.y for MODE 1), BiCGStab's products.
template <int MODE>
void launch_bi_spmv(const fvm_template* t, const double* x, double* y, const double* a, hipStream_t s) {
  if (!t->use_ell) {
    hipLaunchKernelGGL(bi_spmv<MODE>, grid(t), dim3(kRT), 0, s, t->n, t->row_ptr, t->col, t->valM, x, y, a, t->sc,
                       t->part);
    return;
  }
  by_ell(t, [&](auto w, auto ec) {
    hipLaunchKernelGGL((bi_spmv_ell<MODE, decltype(w)::value, ci_t<decltype(ec)>>), grid(t), dim3(kRT), 0, s, t->n,
                       sell_idx(t), ec, t->ellM, t->diagM, x, y, a, t->sc, t->part);
  });
}

// y = X x (+ b) for X = A (which 0), M in FP64 (1) or M in FP32 (2), on the ELL
// copy when there is one, else on CSR.
void spmv_any(const fvm_template* t, int which, const double* x, const double* b, double* y, hipStream_t s) {
  if (t->ellW > 0) {
    by_ell(t, [&](auto w, auto ec) {
      constexpr int W = decltype(w)::value;
      using CI = ci_t<decltype(ec)>;
      const SellIdx S = sell_idx(t);
      if (which == 0)
        hipLaunchKernelGGL((spmv_ell_kernel<double, W, CI>), grid(t), dim3(kRT), 0, s, t->n, S, ec, t->ellA,
                           t->diagA, x, b, y);
      else if (which == 1)
        hipLaunchKernelGGL((spmv_ell_kernel<double, W, CI>), grid(t), dim3(kRT), 0, s, t->n, S, ec, t->ellM,
                           t->diagM, x, b, y);
      else
        hipLaunchKernelGGL((spmv_ell_kernel<float, W, CI>), grid(t), dim3(kRT), 0, s, t->n, S, ec, t->ellM32,
                           t->diagM32, x, b, y);
    });
    return;
  }
  const void* v = which == 0 ? (const void*)t->val : which == 1 ? (const void*)t->valM : (const void*)t->valM32;
  if (which == 2)
    hipLaunchKernelGGL((spmv_kernel<float>), grid(t), dim3(kRT), 0, s, t->n, t->row_ptr, t->col, (const float*)v, x,
                       b, y);
  else
    hipLaunchKernelGGL((spmv_kernel<double>), grid(t), dim3(kRT), 0, s, t->n, t->row_ptr, t->col, (const double*)v,
                       x, b, y);
}

// The sliced ELL copy of M and A (values gathered from val / valM / valM32 after
// their assembly; layout at fvm_template).
int build_ell(fvm_template* t, const std::vector<int32_t>& rp, const std::vector<int32_t>& col) {
  const int64_t N = t->n, ns = (N + kSlice - 1) / kSlice;
  std::vector<int32_t> deg(N, 0), dsrc(N, -1);
  int dmax = 0;
  for (int64_t i = 0; i < N; ++i) {
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      if (col[k] == i)
        dsrc[i] = k;
      else
        ++deg[i];
    }
    dmax = std::max(dmax, deg[i]);
  }
  const int W = std::min(kEllMax, std::max(kEllMin, dmax));
  const int64_t n_off = (int64_t)rp[N] - std::count_if(dsrc.begin(), dsrc.end(), [](int32_t k) { return k >= 0; });
  FVM_CHECK(n_off < ((int64_t)1 << 31) - kSlice, "fvm_template_create: too many off-diagonal entries for the ELL copy");
  std::vector<int32_t> off((size_t)(W + 1) * ns), rem_ptr(ns + 1, 0), rem_off, ecol, esrc, erow;
  std::vector<uint8_t> perm(N);
  ecol.reserve(n_off);
  esrc.reserve(n_off);
  erow.reserve(n_off);
  bool fits16 = true;
  int64_t over = 0;
  int lanes[kSlice], cur[kSlice];
  for (int64_t b = 0; b < ns; ++b) {
    const int64_t r0 = b * kSlice;
    const int nr = (int)std::min<int64_t>(kSlice, N - r0);
    for (int l = 0; l < nr; ++l) lanes[l] = l;
    std::stable_sort(lanes, lanes + nr, [&](int x, int y) { return deg[r0 + x] > deg[r0 + y]; });
    for (int l = 0; l < nr; ++l) {
      perm[r0 + l] = (uint8_t)lanes[l];
      cur[l] = rp[r0 + lanes[l]];
    }
    const int smax = nr ? deg[r0 + lanes[0]] : 0;
    for (int sl = 0; sl < smax; ++sl) {
      if (sl <= W) off[(size_t)(W + 1) * b + sl] = (int32_t)ecol.size();
      for (int l = 0; l < nr && deg[r0 + lanes[l]] > sl; ++l) {
        const int64_t row = r0 + lanes[l];
        int k = cur[l];
        while (col[k] == row) ++k;  // the diagonal is not a slot
        cur[l] = k + 1;
        const int64_t d = (int64_t)col[k] - row;
        fits16 = fits16 && d >= -32768 && d <= 32767;
        ecol.push_back(col[k]);
        esrc.push_back(k);
        erow.push_back((int32_t)row);
        over += sl >= W;
      }
      if (sl >= W) rem_off.push_back((int32_t)ecol.size());
    }
    for (int sl = smax; sl <= W; ++sl) off[(size_t)(W + 1) * b + sl] = (int32_t)ecol.size();
    rem_ptr[b + 1] = (int32_t)rem_off.size();
  }
  const int64_t ne = (int64_t)ecol.size();
  // a kSlice-entry tail: sell_slice reads up to 63 entries past a slot's last one
  ecol.resize(ne + kSlice, 0);
  erow.resize(ne + kSlice, 0);
  esrc.resize(ne + kSlice, -1);
  int32_t *d_esrc = nullptr, *d_dsrc = nullptr;
  int rc = FVM_OK;
  do {
    if ((rc = dupload(&t->sl_off, off.data(), (int64_t)off.size())) ||
        (rc = dupload(&t->sl_perm, perm.data(), N)) || (rc = dupload(&t->sl_rem_ptr, rem_ptr.data(), ns + 1)) ||
        (rc = dupload(&t->sl_rem_off, rem_off.data(), (int64_t)rem_off.size())))
      break;
    if (fits16) {
      std::vector<int16_t> eoff(ne + kSlice);
      for (int64_t e = 0; e < ne + kSlice; ++e) eoff[e] = (int16_t)(ecol[e] - erow[e]);
      if ((rc = dupload(&t->ell_off, eoff.data(), ne + kSlice))) break;
    } else if ((rc = dupload(&t->ell_col, ecol.data(), ne + kSlice))) {
      break;
    }
    const int64_t np = ne + kSlice;
    if ((rc = dalloc(&t->ellM, np)) || (rc = dalloc(&t->ellM32, np)) || (rc = dalloc(&t->diagM, N)) ||
        (rc = dalloc(&t->diagM32, N)) || (rc = dalloc(&t->ellA, np)) || (rc = dalloc(&t->diagA, N)))
      break;
    if ((rc = dupload(&d_esrc, esrc.data(), np)) || (rc = dupload(&d_dsrc, dsrc.data(), N))) break;
    auto g = [](int64_t m) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, 4096))); };
    hipLaunchKernelGGL(gather_vals, g(np), dim3(256), 0, 0, np, d_esrc, t->valM, t->valM32, t->ellM, t->ellM32);
    hipLaunchKernelGGL(gather_vals, g(np), dim3(256), 0, 0, np, d_esrc, t->val, (const float*)nullptr, t->ellA,
                       (float*)nullptr);
    hipLaunchKernelGGL(gather_vals, g(N), dim3(256), 0, 0, N, d_dsrc, t->valM, t->valM32, t->diagM, t->diagM32);
    hipLaunchKernelGGL(gather_vals, g(N), dim3(256), 0, 0, N, d_dsrc, t->val, (const float*)nullptr, t->diagA,
                       (float*)nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = hip_fail(e, "fvm_template_create: ELL copy");
  } while (false);
  for (void* p : {(void*)d_esrc, (void*)d_dsrc})
    if (p) (void)hipFree(p);
  if (rc == FVM_OK) {
    t->ellW = W;
    t->ell_overflow = over;
  }
  return rc;
}

// |b - A x|^2 into sc[S_RES2] (and r = -V (b - A x) on free rows when r is given)
template <class V>
int true_res(fvm_template* t, const double* x, V* r, hipStream_t s) {
  hipLaunchKernelGGL((true_residual<V>), grid(t), dim3(kRT), 0, s, t->n, t->row_ptr, t->col, t->val, t->b, x,
                     t->fixed, t->vol, r, t->part);
  hipLaunchKernelGGL(finish_sum, dim3(1), dim3(kRT), 0, s, t->G, t->part, t->sc, (int)S_RES2);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

// Inner CG on M d = r from d = 0, until |r|_A^2 <= max(irt2 |r0|_A^2, tabs2) or
// maxit iterations; returns the iterations taken and the breakdown flag.
template <class V>
int inner_cg(fvm_template* t, const V* valM, const V* dinv, const V* wv, V* d, V* r, V* z, V* p0, V* p1, V* q,
             double irt2, double tabs2, int maxit, int check, hipStream_t s, int* iters, int* flag) {
  const int64_t n = t->n;
  hipLaunchKernelGGL((cg_init<V>), grid(t), dim3(kRT), 0, s, n, r, dinv, wv, d, z, p0, (const V*)nullptr, t->sc,
                     t->part);
  double h[S_COUNT];
  int k = 0;
  V* pp[2] = {p0, p1};
  for (;;) {
    const int batch = std::min(check, maxit - k);
    for (int j = 0; j < batch; ++j, ++k) {
      V* pold = pp[k & 1];
      V* pnew = pp[(k + 1) & 1];
      launch_dir<V>(t, valM, k, 0, irt2, tabs2, z, pold, pnew, q, s);
      hipLaunchKernelGGL((cg_update<V>), grid(t), dim3(kRT), 0, s, n, k, d, (double*)nullptr, r, pnew, q, dinv, wv,
                         z, t->sc, t->part);
    }
    int rc = read_scalars(t, h, s);
    if (rc) return rc;
    if (h[S_DONE] != 0.0 || k >= maxit) break;
  }
  *iters = (int)h[S_ITER];
  *flag = (int)h[S_FLAG];
  return FVM_OK;
}

// FP32 CG with reliable updates (Sleijpen & van der Vorst 1996): one Krylov
// sequence for the whole solve.  Each step goes straight into the FP64 x (an FP32
// correction vector would carry a rounding error of about eps32 kappa(M) of the
// residual for smooth solutions, which at 10^6 unknowns is the size of a whole
// segment's reduction); whenever the recursive residual estimate has dropped by
// delta since the last replacement (tested every `check` iterations) or reaches
// the target, r is replaced by -V (b - A x) computed in FP64, keeping the
// direction p and rz of the last iteration, so beta continues the sequence.  If
// the true residual is more than twice the recursive one (the recursion ran far
// below what FP32 can represent of the correction, e.g. when CG converges in a few
// iterations between two tests), beta = rz_new / rz_old would swamp z and the
// direction restarts instead.  On entry r holds -V (b - A x) and *res = |b - A x|_2.
int reliable_cg(fvm_template* t, double* x, double target, double delta, int maxit, int max_updates, int check,
                hipStream_t s, int* iters, int* updates, int* flag, double* res) {
  const int64_t n = t->n;
  float* w = (float*)t->work;
  float *d = w, *r = w + n, *z = w + 2 * n, *q = w + 5 * n;
  float* pp[2] = {w + 3 * n, w + 4 * n};
  const double tabs2 = 0.25 * target * target;
  hipLaunchKernelGGL((cg_init<float>), grid(t), dim3(kRT), 0, s, n, r, t->dinv32, t->wv32, d, z, pp[0],
                     (const float*)nullptr, t->sc, t->part);
  double h[S_COUNT];
  double wref = (*res) * (*res);  // |b - A x|^2 at the last replacement
  int k = 0, upd = 0, proj = 0;
  *flag = 0;
  for (;;) {
    const int batch = std::min(check, maxit - k);
    for (int j = 0; j < batch; ++j, ++k) {
      launch_dir<float>(t, t->valM32, k, j == 0 ? proj : 0, 0.0, tabs2, z, pp[k & 1], pp[(k + 1) & 1], q, s);
      hipLaunchKernelGGL((cg_update<float>), grid(t), dim3(kRT), 0, s, n, k, d, x, r, pp[(k + 1) & 1], q,
                         t->dinv32, t->wv32, z, t->sc, t->part);
    }
    int rc = read_scalars(t, h, s);
    if (rc) return rc;
    const bool stop = h[S_DONE] != 0.0;
    if (stop) k = (int)h[S_ITER];  // the iteration that stopped; p of the one before is in pp[k & 1]
    if (h[S_FLAG] != 0.0) *flag = (int)h[S_FLAG];
    proj = 0;
    if (!(stop || h[S_WR] < delta * delta * wref || k >= maxit)) continue;
    if ((rc = true_res<float>(t, x, r, s)) || (rc = read_scalars(t, h, s))) return rc;
    ++upd;
    *res = std::sqrt(h[S_RES2]);
    if (*res <= target || k >= maxit || upd >= max_updates || *flag || !(*res == *res)) break;
    hipLaunchKernelGGL((cg_init<float>), grid(t), dim3(kRT), 0, s, n, r, t->dinv32, t->wv32, d, z,
                       (float*)nullptr, (const float*)pp[k & 1], t->sc, t->part);
    wref = (*res) * (*res);
    proj = k == 0 ? 0 : *res > 2.0 * std::sqrt(h[S_WR]) ? 2 : 1;
  }
  *iters = k;
  *updates = upd;
  return FVM_OK;
}

int inner_bicgstab(fvm_template* t, double* d, double* r, double tabs2, int maxit, int check, hipStream_t s,
                   int* iters, int* flag) {
  const int64_t n = t->n;
  double* w = t->work;
  double *rhat = w + 3 * n, *p = w + 4 * n, *v = w + 5 * n, *phat = w + 6 * n, *shat = w + 7 * n,
         *tt = w + 8 * n;
  const dim3 g = grid(t), b(kRT);
  hipLaunchKernelGGL(bi_init, g, b, 0, s, n, r, t->wv, d, rhat, p, v, t->sc, t->part);
  double h[S_COUNT];
  int k = 0;
  for (;;) {
    const int batch = std::min(check, maxit - k);
    for (int j = 0; j < batch; ++j, ++k) {
      hipLaunchKernelGGL(bi_dir, g, b, 0, s, n, k, tabs2, r, p, v, t->dinv, phat, t->sc, t->part);
      launch_bi_spmv<0>(t, phat, v, rhat, s);
      hipLaunchKernelGGL(bi_s, g, b, 0, s, n, k, r, v, t->dinv, shat, t->sc, t->part);
      launch_bi_spmv<1>(t, shat, tt, r, s);
      hipLaunchKernelGGL(bi_update, g, b, 0, s, n, k, d, r, phat, shat, tt, rhat, t->wv, t->sc, t->part);
    }
    int rc = read_scalars(t, h, s);
    if (rc) return rc;
    if (h[S_DONE] != 0.0 || k >= maxit) break;
  }
  *iters = (int)h[S_ITER];
  *flag = (int)h[S_FLAG];
  return FVM_OK;
}

}  // namespace

extern "C" {

int fvm_template_create(const fvm_mesh* m, const fvm_template_desc* d, fvm_template** out) {
  FVM_CHECK(m && d && out, "fvm_template_create: null argument");
  *out = nullptr;
  FVM_CHECK(d->kind >= 0 && d->kind <= 2, "fvm_template_create: kind must be 0, 1 or 2");
  FVM_CHECK(d->d_edge, "fvm_template_create: d_edge required");
  FVM_CHECK(m->B == 0 || d->d_bedge, "fvm_template_create: d_bedge required");
  FVM_CHECK(d->node_kind, "fvm_template_create: node_kind required");
  FVM_CHECK(m->B == 0 || d->bedge_neumann_fidx, "fvm_template_create: bedge_neumann_fidx required");
  const int64_t N = m->N, T = m->T, B = m->B;
  for (int64_t g = 0; g < N; ++g)
    FVM_CHECK(d->node_kind[g] >= 0 && d->node_kind[g] <= 2, "fvm_template_create: node_kind must be 0, 1 or 2");
  // jacobian_sparsity's pattern from the input-order triangles
  std::vector<int32_t> hs(3 * T), tri_in(3 * T);
  FVM_HIP(hipMemcpy(hs.data(), m->tri, sizeof(int32_t) * 3 * T, hipMemcpyDeviceToHost));
  for (int64_t s = 0; s < T; ++s)
    for (int v = 0; v < 3; ++v) tri_in[3 * (int64_t)m->h_perm[s] + v] = hs[v * T + s];
  int64_t nnz = 0;
  FVM_CHECK(fvmh_csr_pattern(N, T, tri_in.data(), nullptr, nullptr, &nnz) == 0, "fvm_template_create: pattern");
  FVM_CHECK(nnz < INT32_MAX, "fvm_template_create: more than 2^31 nonzeros");
  std::vector<int64_t> rp64(N + 1);
  std::vector<int32_t> col(nnz);
  fvmh_csr_pattern(N, T, tri_in.data(), rp64.data(), col.data(), &nnz);
  std::vector<int32_t> rp(rp64.begin(), rp64.end());
  // boundary-edge endpoints per node
  std::vector<int32_t> hbe(2 * B);
  if (B) FVM_HIP(hipMemcpy(hbe.data(), m->bedges, sizeof(int32_t) * 2 * B, hipMemcpyDeviceToHost));
  std::vector<int32_t> nbs(N + 1, 0), nb(2 * B);
  for (int64_t e = 0; e < B; ++e)
    for (int ep = 0; ep < 2; ++ep) ++nbs[hbe[ep * B + e] + 1];
  for (int64_t i = 0; i < N; ++i) nbs[i + 1] += nbs[i];
  {
    std::vector<int32_t> f(nbs.begin(), nbs.end() - 1);
    for (int64_t e = 0; e < B; ++e)
      for (int ep = 0; ep < 2; ++ep) nb[f[hbe[ep * B + e]]++] = (int32_t)((e << 1) | ep);
  }
  std::vector<int8_t> neu(B);
  for (int64_t e = 0; e < B; ++e) neu[e] = d->bedge_neumann_fidx[e] >= 0;
  std::vector<double> de(3 * T);
  for (int64_t s = 0; s < T; ++s)
    for (int e = 0; e < 3; ++e) de[(size_t)e * T + s] = d->d_edge[3 * (int64_t)m->h_perm[s] + e];

  fvm_template* t = new fvm_template();
  t->m = m;
  t->kind = d->kind;
  t->n = N;
  t->nnz = nnz;
  t->G = (int)std::max<int64_t>(1, std::min<int64_t>((N + kRT - 1) / kRT, kMaxG));
  int rc = FVM_OK;
  int32_t *d_nbs = nullptr, *d_nb = nullptr;
  int8_t *d_kind = nullptr, *d_neu = nullptr;
  double *d_de = nullptr, *d_dbe = nullptr, *d_f = nullptr, *d_r = nullptr, *d_dir = nullptr,
         *d_dudt = nullptr, *d_na = nullptr;
  auto opt = [&](double** dst, const double* src, int64_t n) -> int {
    *dst = nullptr;
    return src ? dupload(dst, src, n) : FVM_OK;
  };
  do {
    if ((rc = dupload(&t->row_ptr, rp.data(), N + 1))) break;
    if ((rc = dupload(&t->col, col.data(), nnz))) break;
    if ((rc = dalloc(&t->val, nnz)) || (rc = dalloc(&t->valM, nnz)) || (rc = dalloc(&t->valM32, nnz))) break;
    if ((rc = dalloc(&t->b, N)) || (rc = dalloc(&t->fixed, N)) || (rc = dalloc(&t->vol, N))) break;
    if ((rc = dalloc(&t->dinv, N)) || (rc = dalloc(&t->wv, N)) || (rc = dalloc(&t->dinv32, N)) ||
        (rc = dalloc(&t->wv32, N)))
      break;
    if ((rc = dalloc(&t->work, 10 * N)) || (rc = dalloc(&t->part, PT_COUNT * kMaxG)) ||
        (rc = dalloc(&t->sc, S_COUNT)))
      break;
    if ((rc = dupload(&d_nbs, nbs.data(), N + 1)) || (rc = dupload(&d_nb, nb.data(), 2 * B))) break;
    if ((rc = dupload(&d_kind, d->node_kind, N)) || (rc = dupload(&d_neu, neu.data(), B))) break;
    if ((rc = dupload(&d_de, de.data(), 3 * T)) || (rc = opt(&d_dbe, d->d_bedge, 2 * B))) break;
    if ((rc = opt(&d_f, d->f_node, N)) || (rc = opt(&d_r, d->r_node, N)) || (rc = opt(&d_dir, d->dir_val, N)) ||
        (rc = opt(&d_dudt, d->dudt_val, N)) || (rc = opt(&d_na, d->neu_a, 2 * B)))
      break;
    hipError_t e = hipMemset(t->sc, 0, sizeof(double) * S_COUNT);
    if (e == hipSuccess) e = hipMemcpy(t->vol, m->cv, sizeof(double) * N, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
      rc = hip_fail(e, "fvm_template_create: initialise");
      break;
    }
    AsmArgs a;
    a.N = N;
    a.T = T;
    a.B = B;
    a.kind = d->kind;
    a.row_ptr = t->row_ptr;
    a.col = t->col;
    a.val = t->val;
    a.b = t->b;
    a.node_tri_start = m->node_tri_start;
    a.node_tri = m->node_tri;
    a.tri = m->tri;
    a.props = m->props;
    a.cv = m->cv;
    a.x = m->x;
    a.y = m->y;
    a.is_vertex = m->is_vertex;
    a.d_edge = d_de;
    a.d_bedge = d_dbe;
    a.bedges = m->bedges;
    a.badj = m->badj;
    a.nb_start = d_nbs;
    a.nb = d_nb;
    a.kindn = d_kind;
    a.neu = d_neu;
    a.f_node = d_f;
    a.r_node = d_r;
    a.dir_val = d_dir;
    a.dudt_val = d_dudt;
    a.neu_a = d_na;
    a.fixed = t->fixed;
    const dim3 g((unsigned)((N + 255) / 256));
    hipLaunchKernelGGL(assemble_kernel, g, dim3(256), 0, 0, a);
    hipLaunchKernelGGL(scaled_kernel, g, dim3(256), 0, 0, N, t->row_ptr, t->col, t->val, t->fixed, t->vol,
                       t->valM, t->valM32, t->dinv, t->dinv32, t->wv, t->wv32);
    hipLaunchKernelGGL(norm2_kernel, grid(t), dim3(kRT), 0, 0, N, t->b, t->part);
    hipLaunchKernelGGL(finish_sum, dim3(1), dim3(kRT), 0, 0, t->G, t->part, t->sc, (int)S_B2);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      rc = hip_fail(e, "fvm_template_create: assembly");
      break;
    }
    rc = build_ell(t, rp, col);
  } while (false);
  void* tmp[] = {d_nbs, d_nb, d_kind, d_neu, d_de, d_dbe, d_f, d_r, d_dir, d_dudt, d_na};
  for (void* p : tmp)
    if (p) (void)hipFree(p);
  if (rc != FVM_OK) {
    free_tmpl(t);
    return rc;
  }
  *out = t;
  return FVM_OK;
}

int fvm_template_create_csr(const fvm_mesh* m, int64_t n, const int32_t* row_ptr, const int32_t* col,
                            const double* val, const double* b, const uint8_t* fixed, fvm_template** out) {
  FVM_CHECK(m && row_ptr && col && val && b && fixed && out, "fvm_template_create_csr: null argument");
  *out = nullptr;
  FVM_CHECK(n == m->N && n > 0, "fvm_template_create_csr: n must be the mesh's number of points");
  FVM_CHECK(row_ptr[0] == 0, "fvm_template_create_csr: row_ptr[0] must be 0");
  for (int64_t i = 0; i < n; ++i)
    FVM_CHECK(row_ptr[i + 1] >= row_ptr[i], "fvm_template_create_csr: row_ptr must not decrease");
  const int64_t nnz = row_ptr[n];
  std::vector<int32_t> rp(row_ptr, row_ptr + n + 1), cl(col, col + nnz);
  std::vector<double> v(val, val + nnz);
  std::vector<uint8_t> fx(fixed, fixed + n);
  for (int64_t i = 0; i < n; ++i) {
    bool diag = false;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      FVM_CHECK(cl[k] >= 0 && cl[k] < n, "fvm_template_create_csr: column out of range");
      diag = diag || cl[k] == i;
      if (fx[i]) v[k] = cl[k] == i ? 1.0 : 0.0;  // a fixed row is x_i = b_i
    }
    FVM_CHECK(diag, "fvm_template_create_csr: every row must hold its diagonal");
    fx[i] = fx[i] != 0;
  }
  fvm_template* t = new fvm_template();
  t->m = m;
  t->kind = FVM_TEMPLATE_POISSON;  // a linear system A x = b
  t->n = n;
  t->nnz = nnz;
  t->G = (int)std::max<int64_t>(1, std::min<int64_t>((n + kRT - 1) / kRT, kMaxG));
  int rc = FVM_OK;
  do {
    if ((rc = dupload(&t->row_ptr, rp.data(), n + 1)) || (rc = dupload(&t->col, cl.data(), nnz)) ||
        (rc = dupload(&t->val, v.data(), nnz)) || (rc = dupload(&t->b, b, n)) ||
        (rc = dupload(&t->fixed, fx.data(), n)))
      break;
    if ((rc = dalloc(&t->valM, nnz)) || (rc = dalloc(&t->valM32, nnz)) || (rc = dalloc(&t->vol, n))) break;
    if ((rc = dalloc(&t->dinv, n)) || (rc = dalloc(&t->wv, n)) || (rc = dalloc(&t->dinv32, n)) ||
        (rc = dalloc(&t->wv32, n)))
      break;
    if ((rc = dalloc(&t->work, 10 * n)) || (rc = dalloc(&t->part, PT_COUNT * kMaxG)) ||
        (rc = dalloc(&t->sc, S_COUNT)))
      break;
    hipError_t e = hipMemset(t->sc, 0, sizeof(double) * S_COUNT);
    if (e == hipSuccess) e = hipMemcpy(t->vol, m->cv, sizeof(double) * n, hipMemcpyDeviceToDevice);
    const dim3 g((unsigned)((n + 255) / 256));
    if (e == hipSuccess) {
      hipLaunchKernelGGL(scaled_kernel, g, dim3(256), 0, 0, n, t->row_ptr, t->col, t->val, t->fixed, t->vol,
                         t->valM, t->valM32, t->dinv, t->dinv32, t->wv, t->wv32);
      hipLaunchKernelGGL(norm2_kernel, grid(t), dim3(kRT), 0, 0, n, t->b, t->part);
      hipLaunchKernelGGL(finish_sum, dim3(1), dim3(kRT), 0, 0, t->G, t->part, t->sc, (int)S_B2);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      rc = hip_fail(e, "fvm_template_create_csr");
      break;
    }
    rc = build_ell(t, rp, cl);
  } while (false);
  if (rc != FVM_OK) {
    free_tmpl(t);
    return rc;
  }
  *out = t;
  return FVM_OK;
}

int fvm_template_destroy(fvm_template* t) {
  free_tmpl(t);
  return FVM_OK;
}

int fvm_template_ell(const fvm_template* t, int32_t* width, int64_t* overflow, int32_t* index_bytes) {
  FVM_CHECK(t && width && overflow, "fvm_template_ell: null argument");
  *width = t->ellW;
  *overflow = t->ell_overflow;
  if (index_bytes) *index_bytes = t->ell_off ? 2 : 4;
  return FVM_OK;
}

int fvm_template_info(const fvm_template* t, int64_t* n, int64_t* nnz) {
  FVM_CHECK(t, "fvm_template_info: null template");
  if (n) *n = t->n;
  if (nnz) *nnz = t->nnz;
  return FVM_OK;
}

int fvm_template_download(const fvm_template* t, int32_t* row_ptr, int32_t* col, double* val, double* b) {
  FVM_CHECK(t, "fvm_template_download: null template");
  if (row_ptr) FVM_HIP(hipMemcpy(row_ptr, t->row_ptr, sizeof(int32_t) * (t->n + 1), hipMemcpyDeviceToHost));
  if (col) FVM_HIP(hipMemcpy(col, t->col, sizeof(int32_t) * t->nnz, hipMemcpyDeviceToHost));
  if (val) FVM_HIP(hipMemcpy(val, t->val, sizeof(double) * t->nnz, hipMemcpyDeviceToHost));
  if (b) FVM_HIP(hipMemcpy(b, t->b, sizeof(double) * t->n, hipMemcpyDeviceToHost));
  return FVM_OK;
}

int fvm_template_apply(const fvm_template* t, const double* d_u, double* d_du, fvm_stream s) {
  FVM_CHECK(t && d_u && d_du, "fvm_template_apply: null argument");
  FVM_CHECK(d_u != d_du, "fvm_template_apply: du must not alias u");
  spmv_any(t, 0, d_u, t->b, d_du, (hipStream_t)s);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

int fvm_template_spmv(const fvm_template* t, int32_t which, const double* d_x, double* d_y, fvm_stream s) {
  FVM_CHECK(t && d_x && d_y, "fvm_template_spmv: null argument");
  FVM_CHECK(d_x != d_y, "fvm_template_spmv: y must not alias x");
  FVM_CHECK(which >= 0 && which <= 2, "fvm_template_spmv: which must be 0, 1 or 2");
  spmv_any(t, which, d_x, nullptr, d_y, (hipStream_t)s);
  FVM_HIP(hipGetLastError());
  return FVM_OK;
}

int fvm_template_residual(fvm_template* t, const double* d_x, double* rel, fvm_stream s) {
  FVM_CHECK(t && d_x && rel, "fvm_template_residual: null argument");
  hipStream_t st = (hipStream_t)s;
  int rc = true_res<double>(t, d_x, nullptr, st);
  if (rc) return rc;
  double h[S_COUNT];
  if ((rc = read_scalars(t, h, st))) return rc;
  const double bn = std::sqrt(h[S_B2]);
  *rel = std::sqrt(h[S_RES2]) / (bn > 0 ? bn : 1.0);
  return FVM_OK;
}

int fvm_template_solve(fvm_template* t, const fvm_solve_opts* o, double* d_x, fvm_solve_stats* st,
                       fvm_stream s) {
  FVM_CHECK(t && o && d_x && st, "fvm_template_solve: null argument");
  FVM_CHECK(o->method >= 0 && o->method <= 3, "fvm_template_solve: unknown method");
  FVM_CHECK(o->rtol > 0 && o->maxiter > 0, "fvm_template_solve: rtol and maxiter must be positive");
  FVM_CHECK(t->kind != 2, "fvm_template_solve: a diffusion template is an ODE, not a linear system");
  FVM_CHECK(o->spmv_format >= FVM_SPMV_AUTO && o->spmv_format <= FVM_SPMV_ELL,
            "fvm_template_solve: unknown spmv_format");
  FVM_CHECK(o->spmv_format != FVM_SPMV_ELL || t->ellW > 0, "fvm_template_solve: no ELL copy of this operator");
  t->use_ell = t->ellW > 0 && o->spmv_format != FVM_SPMV_CSR;
  hipStream_t str = (hipStream_t)s;
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = t->n;
  const int check = o->check_every > 0 ? o->check_every : 32;
  const int max_outer = o->max_outer > 0 ? o->max_outer : 50;
  const double irt = o->method == 1 ? (o->inner_rtol > 0 ? o->inner_rtol : 1e-4) : 0.0;
  double* w = t->work;
  double *d = w, *r = w + n, *z = w + 2 * n, *p0 = w + 3 * n, *p1 = w + 4 * n, *q = w + 5 * n;
  float *df = (float*)w, *rf = df + n, *zf = df + 2 * n, *p0f = df + 3 * n, *p1f = df + 4 * n, *qf = df + 5 * n;
  hipLaunchKernelGGL(init_x, grid(t), dim3(kRT), 0, str, n, t->fixed, t->b, d_x, o->use_x0 ? 1 : 0);
  double h[S_COUNT];
  const bool fp32 = o->method == FVM_SOLVE_CG_MIXED || o->method == FVM_SOLVE_CG_RELIABLE;
  int rc = fp32 ? true_res<float>(t, d_x, rf, str) : true_res<double>(t, d_x, r, str);
  if (rc || (rc = read_scalars(t, h, str))) return rc;
  const double bn = std::sqrt(h[S_B2]) > 0 ? std::sqrt(h[S_B2]) : 1.0;
  const double target = o->rtol * bn;
  double res = std::sqrt(h[S_RES2]);
  // the recursive residual must reach half the target before the true residual is checked
  const double tabs2 = 0.25 * target * target;
  int total = 0, outer = 0, flag = 0;
  if (o->method == FVM_SOLVE_CG_RELIABLE && res > target) {
    const double delta = o->inner_rtol > 0 ? o->inner_rtol : 0.1;
    rc = reliable_cg(t, d_x, target, delta, o->maxiter, o->max_outer > 0 ? o->max_outer : 1000, check, str,
                     &total, &outer, &flag, &res);
    if (rc) return rc;
  }
  while (o->method != FVM_SOLVE_CG_RELIABLE && res > target && total < o->maxiter && outer < max_outer) {
    int it = 0;
    if (o->method == 1) {
      rc = inner_cg<float>(t, t->valM32, t->dinv32, t->wv32, df, rf, zf, p0f, p1f, qf, irt * irt, tabs2,
                           o->maxiter - total, check, str, &it, &flag);
      if (rc) return rc;
      hipLaunchKernelGGL((add_correction<float>), grid(t), dim3(kRT), 0, str, n, df, d_x);
      rc = true_res<float>(t, d_x, rf, str);
    } else {
      if (o->method == 0)
        rc = inner_cg<double>(t, t->valM, t->dinv, t->wv, d, r, z, p0, p1, q, 0.0, tabs2, o->maxiter - total,
                              check, str, &it, &flag);
      else
        rc = inner_bicgstab(t, d, r, tabs2, o->maxiter - total, check, str, &it, &flag);
      if (rc) return rc;
      hipLaunchKernelGGL((add_correction<double>), grid(t), dim3(kRT), 0, str, n, d, d_x);
      rc = true_res<double>(t, d_x, r, str);
    }
    if (rc || (rc = read_scalars(t, h, str))) return rc;
    total += it;
    ++outer;
    res = std::sqrt(h[S_RES2]);
    if (!(res == res) || it == 0) break;  // NaN, or no progress possible
  }
  st->iterations = total;
  st->outer = outer;
  st->flag = flag;
  st->residual = res / bn;
  st->converged = res <= target ? 1 : 0;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return FVM_OK;
}

}  // extern "C"

namespace fvm {

// The Tsit5 loop of fvm_ode_solve on the template's ELL copy of A in one launch
// (tsit5_small_kernel), from time t0 with step dt, it0 steps and ks0 saves done,
// k[0] = f(u) already computed; the final state goes to d_u.  Returns
// FVM_ERR_STATE when the template has no ELL copy or n exceeds kSmallTsit5Max, so
// that the caller keeps the multi-kernel loop.
int template_tsit5_small(const fvm_template* t, const fvm_ode_opts* o, double* const k[7], double* utmp,
                         double* u, double* unew, double* d_u, double t0, double dt, int64_t it0, int64_t ks0,
                         const double A[7][6], const double C[7], const double Bt[7], fvm_ode_stats* st,
                         hipStream_t s) {
  if (!t || t->ellW <= 0 || o->n != t->n || o->n > kSmallTsit5Max) return FVM_ERR_STATE;
  SmallTsit5 a;
  a.n = o->n;
  a.S = sell_idx(t);
  a.ecol = t->ell_col;
  a.eoff = t->ell_off;
  a.eval = t->ellA;
  a.diag = t->diagA;
  a.b = t->b;
  for (int j = 0; j < 7; ++j) a.k[j] = k[j];
  a.utmp = utmp;
  a.u = u;
  a.unew = unew;
  a.u_out = d_u;
  a.t0 = t0;
  a.tf = o->tf;
  a.dt = dt;
  a.abstol = o->abstol;
  a.reltol = o->reltol;
  a.maxiters = o->maxiters;
  a.it0 = it0;
  a.fsal = o->fsal;
  a.n_save = o->n_save;
  a.ks0 = ks0;
  a.saves = o->d_saves;
  for (int j = 0; j < 7; ++j) {
    for (int i = 0; i < 6; ++i) a.A[j][i] = A[j][i];
    a.C[j] = C[j];
    a.Bt[j] = Bt[j];
  }
  double* dev = nullptr;  // save times, then 6 result slots
  FVM_HIP(hipMalloc((void**)&dev, sizeof(double) * (size_t)(o->n_save + 6)));
  a.save_t = dev;
  a.stats = dev + o->n_save;
  int rc = FVM_OK;
  hipError_t e = hipSuccess;
  if (o->n_save > 0)
    e = hipMemcpyAsync(dev, o->save_t, sizeof(double) * (size_t)o->n_save, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    by_width(t->ellW, [&](auto w) {
      if (t->ell_off)
        hipLaunchKernelGGL((tsit5_small_kernel<decltype(w)::value, true>), dim3(1), dim3(kSmallThreads), 0, s, a);
      else
        hipLaunchKernelGGL((tsit5_small_kernel<decltype(w)::value, false>), dim3(1), dim3(kSmallThreads), 0, s, a);
    });
    e = hipGetLastError();
  }
  double r[6] = {0, 0, 0, 0, 0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(r, a.stats, sizeof r, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(dev);
  if (e != hipSuccess) return hip_fail(e, "fvm_ode_solve (one-workgroup Tsit5)");
  st->t = r[0];
  st->naccept += (int64_t)r[1];
  st->nreject += (int64_t)r[2];
  st->nf += (int64_t)r[3];
  st->nsaved = (int64_t)r[4];
  st->retcode = (int32_t)r[5];
  return rc;
}

}  // namespace fvm
