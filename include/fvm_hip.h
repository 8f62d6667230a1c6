/* C ABI of libfvmhip.so: the MI355X (gfx950) implementation of the vertex-centred
 * finite-volume hot path of FiniteVolumeMethod.jl v1.2.3.
 *
 * Conventions
 *  - Every function returns 0 (FVM_OK) or a negative status; fvm_last_error()
 *    returns the message of the calling thread's last failure.  No C++ exception
 *    crosses this boundary.
 *  - Pointers named d_* are device pointers (from fvm_alloc); all others are host
 *    pointers.  Kernels run asynchronously on the given stream (NULL = the null
 *    stream); only *_synchronize and the host copies (fvm_memcpy_h2d, fvm_memcpy_d2h) block.
 *  - Indices are 0-based int32: reference node k (1-based Julia) is node k-1.
 *  - A context (fvm_ctx) holds scratch buffers: calls on one context must not
 *    overlap on different streams.
 */
#ifndef FVM_HIP_H
#define FVM_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVM_OK 0
#define FVM_ERR_ARG -1
#define FVM_ERR_HIP -2
#define FVM_ERR_ALLOC -3
#define FVM_ERR_STATE -4

typedef void* fvm_stream; /* hipStream_t */
typedef void* fvm_event;  /* hipEvent_t */
typedef struct fvm_mesh fvm_mesh;
typedef struct fvm_ctx fvm_ctx;
typedef struct fvm_template fvm_template;
typedef struct fvm_batch fvm_batch;
typedef struct fvm_system fvm_system;

/* ---- runtime ------------------------------------------------------------- */
/* Error text of the calling thread's last failure (copied into buf, truncated). */
int fvm_last_error(char* buf, size_t n);
/* Build configuration of this library (target and tuning constants), as text.  No
 * reference counterpart: lets benchmarks match profiles to the build that made them. */
int fvm_build_info(char* buf, size_t n);
int fvm_device_count(int* n);
int fvm_device_synchronize(void);
/* Device memory (what Julia would wrap as a device vector with a finalizer). */
int fvm_alloc(void** d_ptr, size_t bytes);
int fvm_free(void* d_ptr);
int fvm_memcpy_h2d(void* d_dst, const void* src, size_t bytes, fvm_stream s);
int fvm_memcpy_d2h(void* dst, const void* d_src, size_t bytes, fvm_stream s); /* blocks */
int fvm_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, fvm_stream s);
int fvm_memset(void* d_dst, int value, size_t bytes, fvm_stream s);
int fvm_stream_create(fvm_stream* s);
int fvm_stream_destroy(fvm_stream s);
int fvm_stream_synchronize(fvm_stream s);
int fvm_event_create(fvm_event* e);
int fvm_event_destroy(fvm_event e);
int fvm_event_record(fvm_event e, fvm_stream s);
int fvm_event_elapsed_ms(float* ms, fvm_event start, fvm_event stop); /* waits for stop */

/* ---- geometry (FVMGeometry, reference src/geometry.jl:99-169) -------------- */
typedef struct fvm_mesh_desc {
  int64_t n_points;          /* DelaunayTriangulation.num_points (vertices or not) */
  int64_t n_triangles;       /* solid triangles */
  int64_t n_boundary_edges;
  const double* x;           /* [n_points] */
  const double* y;           /* [n_points] */
  const int32_t* triangles;  /* [3*n_triangles], counter-clockwise (each_solid_triangle) */
  const int32_t* boundary_edges; /* [2*n_boundary_edges], counter-clockwise, interior on the left
                                    (keys of get_boundary_edge_map) */
  int32_t tile_triangles;    /* triangles per tile, <= 1024; 0: default (512) */
  int32_t tile_max_nodes;    /* vertices per tile, <= 512 (<= 1024 above 512 triangles); 0: the limit */
} fvm_mesh_desc;

/* Replaces FVMGeometry(tri) (geometry.jl:99): uploads the mesh, builds the
 * triangle tiles, and computes the control-volume geometry on the device. */
int fvm_mesh_create(const fvm_mesh_desc* desc, fvm_mesh** out);
int fvm_mesh_destroy(fvm_mesh* m);

typedef struct fvm_mesh_info {
  int64_t n_points, n_triangles, n_boundary_edges;
  int32_t n_tiles, n_fix_nodes, n_partials;
  int64_t n_local_nodes;
  const double* d_cv_volumes;     /* [n_points] device */
  const double* d_x;              /* [n_points] device */
  const double* d_y;              /* [n_points] device */
} fvm_mesh_info;
int fvm_mesh_get_info(const fvm_mesh* m, fvm_mesh_info* info);

/* Copies the geometry back in the input triangle order: cv_volumes [n_points]
 * (FVMGeometry.cv_volumes) and props [24*n_triangles], per triangle
 * s1..s9, CV-edge midpoints (m1x m1y m2x m2y m3x m3y), unit normals (n1x n1y
 * n2x n2y n3x n3y), lengths (l1 l2 l3) -- TriangleProperties, geometry.jl:21-26. */
int fvm_mesh_download_geometry(const fvm_mesh* m, double* cv_volumes, double* props);
/* The record as the RHS tile kernels rebuild it from the vertex coordinates, in the
 * layout of fvm_mesh_download_geometry (waits for completion).  No reference
 * counterpart: lets tests pin the rebuilt values against the stored geometry (s1..s9
 * and the midpoints bit for bit; with the build's geometry level 4 the normals are
 * unnormalised, (e_y, -e_x) = n l, and the lengths 1). */
int fvm_mesh_rebuilt_geometry(const fvm_mesh* m, double* props);

/* ---- problem (FVMProblem + Conditions, problem.jl:118-163, conditions.jl:486-552) ---- */
enum {
  FVM_FLUX_D_POLY_U = 0, /* q = -D(u_s)(alpha, beta), D = d0 + u(d1 + u(d2 + u d3)),
                            u_s = alpha x_s + beta y_s + gamma (construct_flux_function) */
  FVM_FLUX_D_NODAL = 1,  /* q = -D(x_s)(alpha, beta), D the piecewise-linear interpolant of
                            exp(logD) (pl_interpolate of a nodal field) */
  FVM_FLUX_D_EDGE = 2,   /* q = -D_s (alpha, beta) with D_s given at every CV-edge midpoint and
                            boundary half-edge point (D(x, y) independent of t and u) */
  FVM_FLUX_LINEAR = 3,   /* q = (a1 alpha + a2 beta + a3 gamma, a4 alpha + a5 beta + a6 gamma), the
                            a's given at every CV-edge midpoint and boundary half-edge point
                            (independent of t and u): advection-diffusion, anisotropic
                            diffusion, any flux linear in u and its gradient */
  FVM_FLUX_D_POWER = 4,  /* q = -D(u_s)(alpha, beta), D = dpoly[0] + dpoly[1] |u_s|^dpoly[2],
                            dpoly[2] > 0 (porous-medium diffusivity) */
  FVM_FLUX_JIT = 5       /* q = -D(x_s, y_s, t, u_s)(alpha, beta) and the source S(x, y, t, u) of
                            any closure (problem.jl:118-163, 425-440), given as HIP source in
                            jit_source and compiled with hipRTC at fvm_ctx_create; spoly, st
                            and s_node are then unused */
};
enum { FVM_NODE_NONE = 0, FVM_NODE_DIRICHLET = 1, FVM_NODE_DUDT = 2 };

typedef struct fvm_physics_desc {
  int32_t flux_kind;
  double dpoly[4];              /* FVM_FLUX_D_POLY_U; FVM_FLUX_D_POWER: d0, c, p */
  const double* d_edge;         /* [3*n_triangles] input triangle order, CV edge (i,j),(j,k),(k,i);
                                   FVM_FLUX_LINEAR: [18*n_triangles], a1..a6 of each CV edge */
  const double* d_bedge;        /* [2*n_boundary_edges] at the half-edge points m_i, m_j;
                                   FVM_FLUX_LINEAR: [12*n_boundary_edges], a1..a6 of each */
  const double* log_d;          /* [n_points] FVM_FLUX_D_NODAL (replaceable by fvm_set_log_d) */
  double spoly[4];              /* S(x,y,t,u) = s_node[i] + s0 + u(s1 + u(s2 + u s3)) + st t */
  double st;
  const double* s_node;         /* [n_points] or NULL */
  const int8_t* node_kind;      /* [n_points] FVM_NODE_*; a node in both Dicts is Dirichlet */
  const int32_t* node_fidx;     /* [n_points] condition function of Dirichlet / Dudt nodes */
  const int32_t* bedge_neumann_fidx; /* [n_boundary_edges] function of a Neumann edge, -1 otherwise */
  int32_t n_functions;          /* condition functions a(x,y,t,u) = c0 + cu u + ct t + cx x + cy y + g,
                                   or + g exp(lam t) where fn_lam is given and nonzero */
  const double* fn_c0;
  const double* fn_cu;
  const double* fn_ct;
  const double* fn_cx;          /* may be NULL (0) */
  const double* fn_cy;          /* may be NULL (0) */
  const double* node_g;         /* [n_points] g at Dirichlet / Dudt nodes, or NULL (0) */
  const double* bedge_g;        /* [2*n_boundary_edges] g at Neumann half-edge points, or NULL */
  const double* fn_lam;         /* [n_functions] exp rate of the g term, or NULL (0) */
  const char* jit_source;       /* FVM_FLUX_JIT: HIP source defining, for the point (x, y), time t
                                   and solution value u,
                                     __device__ double fvm_jit_D(double x, double y, double t, double u);
                                     __device__ double fvm_jit_dD(double x, double y, double t, double u);
                                     __device__ double fvm_jit_S(double x, double y, double t, double u);
                                     __device__ double fvm_jit_dS(double x, double y, double t, double u);
                                   (dD, dS: derivatives in u); NULL otherwise */
} fvm_physics_desc;

int fvm_ctx_create(const fvm_mesh* m, const fvm_physics_desc* phys, fvm_ctx** out);
/* Compiles the FVM_FLUX_JIT kernels of jit_source for gfx950 (tiles of tile_triangles,
 * 0 = the default 512) without loading them: no GPU needed.  Returns FVM_ERR_ARG with
 * the compiler's diagnostics in fvm_last_error when the source does not compile;
 * code_bytes (may be NULL) receives the code object's size.  Julia compiles a closure
 * into fvm_eqs! on first call (problem.jl:425-440); this is that step, checked ahead. */
int fvm_jit_check(const char* jit_source, int32_t tile_triangles, size_t* code_bytes);
int fvm_ctx_destroy(fvm_ctx* c);
/* FVM_FLUX_D_NODAL: use the device field d_log_d [n_points] from now on. */
int fvm_ctx_set_log_d(fvm_ctx* c, const double* d_log_d);

/* fvm_eqs!(du, u, p, t) (main_equations.jl:28-81): triangle fluxes, boundary-edge
 * fluxes and source / condition terms, du = 0 at Dirichlet nodes and at points
 * that are not vertices. */
int fvm_rhs(fvm_ctx* c, const double* d_u, double* d_du, double t, fvm_stream s);

/* fvm_rhs with HIP events around its two kernels (triangle tiles, then the
 * shared-node fix-up); waits for completion and returns both durations.  No
 * reference counterpart: a measurement hook for benchmarks. */
int fvm_rhs_timed(fvm_ctx* c, const double* d_u, double* d_du, double t, fvm_stream s,
                  float* ms_tiles, float* ms_fixup);

/* update_dirichlet_nodes! (equations/dirichlet.jl:38-86): u_i = a(x_i, y_i, t, u_i)
 * at Dirichlet nodes (the Dirichlet callback of solve.jl:133-165). */
int fvm_dirichlet_apply(fvm_ctx* c, double* d_u, double t, fvm_stream s);

/* Jacobian-vector product J_u v of fvm_eqs! (what ForwardDiff through DiffCache
 * computes for implicit steppers, main_equations.jl:56). */
int fvm_jvp(fvm_ctx* c, const double* d_u, const double* d_v, double* d_jv, double t, fvm_stream s);

/* Vector-Jacobian products of fvm_eqs!: d_ubar = J_u^T w and, for
 * FVM_FLUX_D_NODAL, d_logd_bar = J_logD^T w (d_logd_bar may be NULL).  No
 * reference counterpart (the reference differentiates only forward, by
 * ForwardDiff); it is the adjoint of fvm_rhs. */
int fvm_vjp(fvm_ctx* c, const double* d_u, const double* d_w, double* d_ubar,
            double* d_logd_bar, double t, fvm_stream s);

/* ---- pointwise queries ------------------------------------------------------
 * compute_flux (problem.jl:458-487) and pl_interpolate (utils.jl:23-27), each for a
 * batch of queries (the reference answers one per call).  Both wait for completion
 * and return FVM_ERR_ARG when a queried edge or triangle is not in the mesh. */
/* out[q] = q(m, t, alpha, beta, gamma) . n across the edge (d_i[q], d_j[q]): m the
 * edge midpoint, n the clockwise normal of i -> j, the shape function from the
 * triangle on the right of i -> j (or on its left for a boundary edge with the
 * domain on the left).  d_dq [n]: D at the edge midpoints, required for
 * FVM_FLUX_D_EDGE (D(x, y) is a host closure); [6n] a1..a6 at the edge midpoints
 * for FVM_FLUX_LINEAR; NULL otherwise. */
int fvm_compute_flux(fvm_ctx* c, int64_t n, const int32_t* d_i, const int32_t* d_j, const double* d_u,
                     double t, const double* d_dq, double* d_out, fvm_stream s);
/* out[q * n_vars + v] = alpha_v x_q + beta_v y_q + gamma_v on the mesh triangle
 * (d_tri[3q], d_tri[3q+1], d_tri[3q+2]) given in any rotation of its counter-clockwise
 * order; d_u [n_vars * n_points] column-major (n_vars = 1 for an FVMProblem). */
int fvm_pl_interpolate(const fvm_mesh* m, int32_t n_vars, int64_t n, const int32_t* d_tri,
                       const double* d_x, const double* d_y, const double* d_u, double* d_out,
                       fvm_stream s);
/* d_tri_out[q] = the (input) index of a mesh triangle containing (d_x[q], d_y[q]) (on an
 * edge or vertex: one of the triangles sharing it), -1 outside the mesh.  Each query
 * walks from a triangle of its start vertex d_start[q] (NULL: vertex 0) across the
 * edges the point lies outside of; a walk trapped by the boundary of a non-convex
 * domain is finished by a scan of all triangles.  Replaces DelaunayTriangulation's
 * jump_and_march / find_triangle, which the reference's tutorials call before
 * pl_interpolate.  Blocks until done. */
int fvm_locate_points(const fvm_mesh* m, int64_t n, const double* d_x, const double* d_y, const int32_t* d_start,
                      int32_t* d_tri_out, fvm_stream s);

/* ---- linear templates (src/specific_problems/) ----------------------------
 * The operator A and vector b of PoissonsEquation / LaplacesEquation
 * (poissons_equation.jl:58-88, laplaces_equation.jl:51-78), MeanExitTimeProblem
 * (mean_exit_time.jl:57-94), DiffusionEquation and LinearReactionDiffusionEquation
 * (diffusion_equation.jl:69-101, linear_reaction_diffusion_equations.jl:76-125),
 * assembled on the device in the CSR pattern of jacobian_sparsity (solve.jl:50-77):
 * the contributions of abstract_templates.jl:73-325, including the division of the
 * j row of a non-Neumann boundary edge by V_i (:262).  Rows carry every point
 * (num_points); points that are not vertices get identity rows. */
enum {
  FVM_TEMPLATE_POISSON = 0,        /* A u = b, Dirichlet rows A_ii = 1, b_i = value */
  FVM_TEMPLATE_MEAN_EXIT_TIME = 1, /* triangle terms only, b = -1, Dirichlet rows A_ii = 1 */
  FVM_TEMPLATE_DIFFUSION = 2       /* du/dt = A u + b (Dudt rows: b_i = value, A row 0) */
};

typedef struct fvm_template_desc {
  int32_t kind;                       /* FVM_TEMPLATE_* */
  const double* d_edge;               /* [3*n_triangles] D(x, y, p) at the three CV-edge
                                         midpoints of each input triangle (edges (i,j), (j,k), (k,i)) */
  const double* d_bedge;              /* [2*n_boundary_edges] D at the half-edge points m_i, m_j */
  const int8_t* node_kind;            /* [n_points] 0 none, 1 Dirichlet, 2 Dudt */
  const int32_t* bedge_neumann_fidx;  /* [n_boundary_edges] >= 0 for a Neumann edge, else -1 */
  const double* f_node;               /* [n_points] Poisson source f(x_i, y_i, p), or NULL (0) */
  const double* r_node;               /* [n_points] linear reaction coefficient, or NULL (0) */
  const double* neu_a;                /* [2*n_boundary_edges] Neumann value at m_i, m_j, or NULL */
  const double* dir_val;              /* [n_points] Dirichlet value (Poisson), or NULL (0) */
  const double* dudt_val;             /* [n_points] Dudt value (diffusion), or NULL (0) */
} fvm_template_desc;

int fvm_template_create(const fvm_mesh* m, const fvm_template_desc* d, fvm_template** out);
/* A linear system A x = b handed in as host CSR (0-based, int32 columns, every row
 * holding its diagonal; rows with fixed[i] != 0 are taken as x_i = b_i), set up for
 * fvm_template_solve and the products exactly like an assembled PoissonsEquation (the
 * mesh gives the control-volume areas of the M = -diag(V) A scaling).  For the
 * SteadyFVMProblem of an fvm_eqs! that is affine in u: A = its Jacobian (from
 * fvm_jvp probes), b = -f(0), the Dirichlet nodes fixed -- the LinearProblem the
 * reference builds for the same equation with a template (poissons_equation.jl:58-88,
 * abstract_templates.jl:302-309), solved on the device instead of by Newton steps
 * (solve.jl:185-195, 209-220 hand SteadyFVMProblems to NonlinearSolve). */
int fvm_template_create_csr(const fvm_mesh* m, int64_t n, const int32_t* row_ptr, const int32_t* col,
                            const double* val, const double* b, const uint8_t* fixed, fvm_template** out);
int fvm_template_destroy(fvm_template* t);
int fvm_template_info(const fvm_template* t, int64_t* n, int64_t* nnz);
/* The sliced ELL copy of the operator used by the Krylov products and
 * fvm_template_apply / fvm_template_spmv (64-row slices, rows ordered by length inside
 * a slice, no padding): its unrolled slots, the longest row's off-diagonal count
 * clamped to 4..10 (0: no copy, CSR throughout), the number of off-diagonals past
 * those slots and the bytes of each entry's column index (2: offsets from the row,
 * when every one fits 16 bits; 4: columns).  No counterpart in the reference (its
 * operator is a SparseMatrixCSC handed to KLU). */
int fvm_template_ell(const fvm_template* t, int32_t* width, int64_t* overflow, int32_t* index_bytes);
/* the assembled A (CSR, 0-based, int32) and b; any pointer may be NULL */
int fvm_template_download(const fvm_template* t, int32_t* row_ptr, int32_t* col, double* val,
                          double* b);
/* du = A u + b: the MatrixOperator of the augmented [A b; 0 0] system
 * (diffusion_equation.jl:82-94) applied to (u, 1). */
int fvm_template_apply(const fvm_template* t, const double* d_u, double* d_du, fvm_stream s);
/* y = A x (which 0), y = M x in FP64 (1) or FP32 (2), where M = -diag(V) A on the free
 * rows and columns and the identity on fixed rows: the operator of the steady solvers. */
int fvm_template_spmv(const fvm_template* t, int32_t which, const double* d_x, double* d_y,
                      fvm_stream s);
/* |b - A x|_2 / |b|_2 (waits for completion) */
int fvm_template_residual(fvm_template* t, const double* d_x, double* rel, fvm_stream s);

enum {
  FVM_SOLVE_CG = 0,       /* FP64 Jacobi-PCG on M */
  FVM_SOLVE_CG_MIXED = 1, /* FP64 iterative refinement around FP32 Jacobi-PCG on M */
  FVM_SOLVE_BICGSTAB = 2, /* FP64 Jacobi-BiCGStab on M */
  FVM_SOLVE_CG_RELIABLE = 3 /* one FP32 Jacobi-PCG sequence with FP64 reliable updates of x and the residual */
};

/* Forms of the scaled operator M in the Krylov products: CSR with eight lanes per row, or an ELL
 * copy built at fvm_template_create (one thread per row, W off-diagonal slots
 * stored slot-major, W the fewest of 4..10 that hold every off-diagonal of 99 % of
 * the rows, the rest in a per-row overflow list).  AUTO takes ELL. */
enum { FVM_SPMV_AUTO = 0, FVM_SPMV_CSR = 1, FVM_SPMV_ELL = 2 };

typedef struct fvm_solve_opts {
  int32_t method;       /* FVM_SOLVE_* */
  int32_t maxiter;      /* total inner iterations */
  int32_t check_every;  /* iterations between host reads of the device scalars (0: 32) */
  int32_t max_outer;    /* refinement / restart steps (0: 50; FVM_SOLVE_CG_RELIABLE: updates, 0: 1000) */
  int32_t use_x0;       /* start from d_x (its fixed rows are reset to b) instead of 0 */
  int32_t spmv_format;  /* FVM_SPMV_*: the form of M in the CG and BiCGStab products (the
                           FP64 residuals b - A x always use CSR) */
  double rtol;          /* stop when |b - A x|_2 <= rtol |b|_2 */
  double inner_rtol;    /* FVM_SOLVE_CG_MIXED: relative drop of each FP32 solve (0: 1e-4);
                           FVM_SOLVE_CG_RELIABLE: residual drop between updates (0: 0.1) */
} fvm_solve_opts;

typedef struct fvm_solve_stats {
  int32_t converged;
  int32_t iterations;   /* inner iterations over all refinement steps */
  int32_t outer;        /* refinement / restart steps */
  int32_t flag;         /* 2: breakdown (non-positive curvature or a zero BiCGStab scalar) */
  double residual;      /* final |b - A x|_2 / |b|_2 */
  double seconds;       /* wall time of the call, including its host synchronisations */
} fvm_solve_stats;

/* LinearProblem(A, b) of a steady template solved on the device (replaces the
 * sparse direct solve behind solve(::AbstractFVMTemplate), abstract_templates.jl:58-60).
 * d_x [n_points] receives the solution.  A template's work vectors and solver
 * scalars serve one solve at a time: calls on the same template must not overlap
 * (different templates are independent). */
int fvm_template_solve(fvm_template* t, const fvm_solve_opts* o, double* d_x, fvm_solve_stats* st,
                       fvm_stream s);

/* ---- ensembles ------------------------------------------------------------
 * Many FVMProblems on meshes of uneven size advanced together (what the reference
 * does by looping solve over an EnsembleProblem, solve.jl:222-256): the members are
 * packed into one device mesh and context, so one launch of the RHS kernels covers
 * every member.  Members share flux_kind (and D(u) for FVM_FLUX_D_POLY_U) and the
 * source polynomial; their meshes, D fields, conditions and s_node are their own. */
typedef struct fvm_batch_opts {
  int32_t tile_triangles; /* 0: 512, or the autotuner's choice */
  int32_t autotune;       /* nonzero (with tile_triangles 0): time 256-, 512- and 1024-triangle
                             tiles on the packed RHS and keep the fastest */
  int32_t graph_steps;    /* > 0: fvm_batch_euler replays a HIP graph of this many steps
                             (closures independent of t only) */
  int32_t unfused;        /* 1: Euler steps as RHS + a separate update kernel (for comparison);
                             0: the update fused into the RHS kernels */
} fvm_batch_opts;

typedef struct fvm_batch_info {
  int32_t n_problems;
  int32_t tile_triangles;          /* the tile size in use */
  int64_t n_points, n_triangles, n_boundary_edges;  /* of the packed mesh */
  int32_t n_tiles, n_fix_nodes;
  int32_t n_candidates;            /* autotuner candidates timed */
  int32_t candidate_tiles[4];
  float candidate_ms[4];           /* packed RHS time of each candidate */
} fvm_batch_info;

/* dt [n_problems]: each member's explicit time step (NULL: fvm_batch_euler unavailable).
 * Closures that depend on t need one common dt. */
int fvm_batch_create(int32_t n_problems, const fvm_mesh_desc* meshes,
                     const fvm_physics_desc* physics, const double* dt,
                     const fvm_batch_opts* opts, fvm_batch** out);
int fvm_batch_destroy(fvm_batch* b);
/* node_offsets [n_problems + 1] (may be NULL): member p owns packed nodes
 * [node_offsets[p], node_offsets[p+1]), in its own node order. */
int fvm_batch_get_info(const fvm_batch* b, fvm_batch_info* info, int64_t* node_offsets);
/* the packed mesh and context (owned by the batch): fvm_rhs / fvm_jvp / fvm_vjp on them
 * evaluate every member at once */
int fvm_batch_mesh(const fvm_batch* b, const fvm_mesh** mesh, const fvm_ctx** ctx);
/* fvm_eqs! of every member at time t, one launch of the tile and fix-up kernels */
int fvm_batch_rhs(fvm_batch* b, const double* d_u, double* d_du, double t, fvm_stream s);
/* n_steps explicit Euler steps of every member from t0: u += dt_p fvm_eqs!(u), then the
 * Dirichlet callback (solve.jl:133-165) at the member's new time. */
int fvm_batch_euler(fvm_batch* b, double* d_u, int32_t n_steps, double t0, fvm_stream s);

/* ---- systems (FVMSystem, problem.jl:196-413) -------------------------------
 * n_vars coupled variables on one mesh.  The state is the reference's matrix
 * u[var, i] stored column-major: u[var + n_vars * i].  Device forms of the closures:
 *   flux       q_v = -sum_w K[v][w] (alpha_w, beta_w)             (eval_flux_function for
 *              a system, individual_flux_contributions.jl:36-50; K constant)
 *   source     S_v = s_node[v][i] + sum over terms k of variable v of
 *              coef_k prod_w u_w^exp_k[w] + st[v] t                 (source_contributions.jl:17-30)
 *   condition  a = c0 + sum_w cu[w] u_w + ct t + cx x + cy y + g exp(lam t), g per location
 * Condition functions are numbered in the merged order of merge_problem_conditions
 * (problem.jl:368-400): variable v's function f is number f + (functions of vars < v). */
typedef struct fvm_system_desc {
  int32_t n_vars;                    /* 1..4 */
  const double* flux_k;              /* [n_vars * n_vars] row-major K[v][w] */
  int32_t n_terms;                   /* source monomials, at most 24 */
  const int32_t* term_var;           /* [n_terms] the variable whose source holds the term */
  const int32_t* term_exp;           /* [n_terms * n_vars] exponents 0..3 */
  const double* term_coef;           /* [n_terms] */
  const double* s_node;              /* [n_vars * n_points] var-major, or NULL */
  const double* st;                  /* [n_vars] or NULL */
  const int8_t* node_kind;           /* [n_vars * n_points] FVM_NODE_* per variable */
  const int32_t* node_fidx;          /* [n_vars * n_points] merged function index */
  const int32_t* bedge_neumann_fidx; /* [n_vars * n_boundary_edges], -1: not Neumann for the variable */
  int32_t n_functions;
  const double* fn_c0;               /* [n_functions] */
  const double* fn_cu;               /* [n_functions * n_vars] */
  const double* fn_ct;               /* [n_functions] or NULL (0) */
  const double* fn_cx;               /* [n_functions] or NULL (0) */
  const double* fn_cy;               /* [n_functions] or NULL (0) */
  const double* fn_lam;              /* [n_functions] or NULL (0: g enters without a time factor) */
  const double* node_g;              /* [n_vars * n_points] or NULL (0) */
  const double* bedge_g;             /* [n_vars * 2 * n_boundary_edges], half-edge points m_i, m_j,
                                        or NULL (0) */
} fvm_system_desc;

/* A system holds scratch buffers (partials, boundary terms), like a context: calls on
 * one system must not overlap on different streams. */
int fvm_system_create(const fvm_mesh* m, const fvm_system_desc* d, fvm_system** out);
int fvm_system_destroy(fvm_system* s);
/* fvm_eqs!(du, u, p, t) for an FVMSystem (main_equations.jl:28-81 with the system
 * methods in src/equations/): du[var, i] for every variable, 0 at Dirichlet
 * entries and at points that are not vertices. */
int fvm_system_rhs(fvm_system* s, const double* d_u, double* d_du, double t, fvm_stream st);
/* update_dirichlet_nodes! for a system (dirichlet.jl:46-53): variable by variable,
 * u[var, i] = a(x_i, y_i, t, u[:, i]) with the variables already updated. */
int fvm_system_dirichlet_apply(fvm_system* s, double* d_u, double t, fvm_stream st);
/* Jacobian-vector product of fvm_system_rhs along d_v (what ForwardDiff through
 * DiffCache computes for implicit steppers on a system, main_equations.jl:56): the
 * flux of v, the source monomials' derivatives and cu . v for Dudt / Neumann
 * conditions; 0 at Dirichlet entries and points that are not vertices. */
int fvm_system_jvp(fvm_system* s, const double* d_u, const double* d_v, double* d_jv, double t,
                   fvm_stream st);
/* compute_flux for a system (problem.jl:458-487, is_system branch): out[q * n_vars + v]
 * is variable v's q . n across the edge (d_i[q], d_j[q]); waits for completion. */
int fvm_system_compute_flux(fvm_system* s, int64_t n, const int32_t* d_i, const int32_t* d_j,
                            const double* d_u, double t, double* d_out, fvm_stream st);

/* ---- device vector operations (used by the integrators and solvers) ------
 * The reference broadcasts over Vector{Float64} inside OrdinaryDiffEq; device-
 * resident state needs these as kernels.  Reductions use a fixed grid and order,
 * so they are bitwise reproducible; they block until their result is on the host
 * and share one scratch area per process, so calls from several host threads are
 * serialised (each is still correct). */
/* d_out = sum_k coefs[k] * d_xs[k], 1 <= nterms <= 8 (d_out may alias an input). */
int fvm_vec_lincomb(int64_t n, double* d_out, int nterms, const double* coefs,
                    const double* const* d_xs, fvm_stream s);
/* d_out = d_x * d_y elementwise (d_out may alias an input). */
int fvm_vec_mul(int64_t n, double* d_out, const double* d_x, const double* d_y, fvm_stream s);
int fvm_vec_dot(int64_t n, const double* d_x, const double* d_y, double* result, fvm_stream s);
/* results[j] = d_xs[j] . d_y for j < k <= 64, one pass over d_y (waits for completion;
 * the Gram-Schmidt step of the Krylov solvers). */
int fvm_vec_mdot(int64_t n, int32_t k, const double* const* d_xs, const double* d_y, double* results,
                 fvm_stream s);
/* fvm_vec_mdot without allocating: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The k dot products (1 <= k <= 64, more is an
 * error) go to the caller's device array d_results, equal bit for bit to
 * fvm_vec_mdot's; the partial sums live in a static per-device array of the library
 * (64 x 512 doubles = 256 KiB per GPU, held until the process ends, every partial
 * rewritten by each call).  Asynchronous on s; calls that overlap on one GPU race
 * on that array. */
int fvm_vec_mdot_noalloc(int64_t n, int32_t k, const double* const* d_xs, const double* d_y, double* d_results,
                         fvm_stream s);
/* d_y += sum_j coefs[j] * d_xs[j], k <= 64 (no d_xs[j] may alias d_y). */
int fvm_vec_maxpy(int64_t n, double* d_y, int32_t k, const double* coefs, const double* const* d_xs,
                  fvm_stream s);
/* sqrt(mean((err_i / (abstol + reltol * max(|u0_i|, |u1_i|)))^2)): the
 * error norm of OrdinaryDiffEq's adaptive step control. */
int fvm_vec_err_norm(int64_t n, const double* d_err, const double* d_u0, const double* d_u1,
                     double abstol, double reltol, double* result, fvm_stream s);
/* fvm_vec_err_norm of err = sum_k coefs[k] * d_xs[k] (1 <= nterms <= 8) without
 * storing err: the Runge-Kutta error estimate (the b - bhat weights times the
 * stages) and its norm in one pass; equal bit for bit to fvm_vec_lincomb followed
 * by fvm_vec_err_norm. */
int fvm_vec_lincomb_err_norm(int64_t n, int nterms, const double* coefs, const double* const* d_xs,
                             const double* d_u0, const double* d_u1, double abstol, double reltol,
                             double* result, fvm_stream s);

/* ---- TR-BDF2 for small affine problems (solve.jl:167-183 with TRBDF2) -------
 * f(u) = J u + c with J constant (linear fluxes, sources affine in u, conditions
 * constant in t): the whole TR-BDF2 integration (the tutorials' stiff integrator,
 * there with a KLU factorisation of J) in one workgroup: Newton with GMRES on
 * W = I - d h J per stage, the W^-1-filtered error estimate, the step controller,
 * the Dirichlet update after every accepted step (constant values) and saveat.
 * J comes in ELL form (width columns per row, slot-major [width][n], padding slots
 * with the row's own column and a zero value, further entries in an overflow CSR).
 * Small n only (<= 65536): larger problems use the host-driven integrator. */
typedef struct fvm_affine_op {
  int64_t n;
  int32_t width;
  const int32_t* d_col;      /* [width * n] */
  const double* d_val;       /* [width * n] */
  const double* d_diag;      /* [n] J_ii */
  const int32_t* d_ovf_rp;   /* [n + 1] */
  const int32_t* d_ovf_col;  /* [ovf_rp[n]] (may be NULL when empty) */
  const double* d_ovf_val;
  const double* d_c;         /* [n] f(0) */
  int32_t n_dirichlet;
  const int32_t* d_dir_nodes; /* [n_dirichlet] */
  const double* d_dir_values; /* [n_dirichlet] */
} fvm_affine_op;

typedef struct fvm_trbdf2_opts {
  double t0, tf, dt;          /* dt: the first step (> 0) */
  double abstol, reltol;
  int64_t maxiters;
  double newton_kappa;        /* Newton stops when theta / (1 - theta) |dz| <= kappa */
  int32_t max_newton;
  int32_t krylov_restart;     /* 1..30 */
  double krylov_rtol;
  int32_t smooth_est;         /* filter the error estimate through W^-1 */
  int64_t n_save;
  const double* save_t;       /* host [n_save], ascending, > t0 */
  double* d_saves;            /* device [n_save * n] */
} fvm_trbdf2_opts;

typedef struct fvm_trbdf2_stats {
  int64_t nf, njvp, naccept, nreject, nnewton, nkrylov, nnewton_fail, nsaved;
  int32_t retcode;            /* 0 Success, 1 MaxIters, 2 Unstable, 3 ConvergenceFailure */
  int32_t reserved;
  double t;
} fvm_trbdf2_stats;

/* d_u [n]: the initial state in, the state at st->t out.  Blocks until done. */
int fvm_trbdf2_affine(const fvm_affine_op* op, const fvm_trbdf2_opts* o, double* d_u, fvm_trbdf2_stats* st,
                      fvm_stream s);

/* ---- explicit time integration (solve(::FVMProblem), solve.jl:133-220) -----
 * The whole adaptive or fixed-step integration of du/dt = f(u, t) in one call, f one
 * of the library's operators, with the Dirichlet update (dirichlet.jl:78-86) after
 * every accepted step, as the reference's DiscreteCallback(true, ...) of
 * solve.jl:133-165, and the states at the saveat times written to device memory.
 * Replaces handing fvm_eqs! to OrdinaryDiffEq (solve.jl:167-220) for Tsit5 and the
 * fixed-step explicit methods: a step costs its kernel launches and, for Tsit5, one
 * host synchronisation. */
typedef enum {
  FVM_ODE_TSIT5 = 0,  /* Tsitouras 5(4), PI controller (beta1 7/50, beta2 2/25, safety 0.9,
                         step ratio in [1/10, 5]), Hermite dense output for saveat */
  FVM_ODE_EULER = 1,  /* forward Euler, fixed dt */
  FVM_ODE_SSPRK3 = 2  /* Shu-Osher SSPRK3, fixed dt */
} fvm_ode_method;

typedef enum {
  FVM_OP_PROBLEM = 0,  /* op: fvm_ctx*     f = fvm_rhs, update = fvm_dirichlet_apply */
  FVM_OP_SYSTEM = 1,   /* op: fvm_system*  f = fvm_system_rhs, update = fvm_system_dirichlet_apply */
  FVM_OP_TEMPLATE = 2  /* op: fvm_template* f = fvm_template_apply (A u + b), no update */
} fvm_op_kind;

typedef enum { FVM_ODE_SUCCESS = 0, FVM_ODE_MAXITERS = 1, FVM_ODE_UNSTABLE = 2 } fvm_ode_retcode;

typedef struct fvm_ode_opts {
  int32_t method;         /* FVM_ODE_* */
  int32_t op_kind;        /* FVM_OP_*: what op points to */
  void* op;
  int32_t dirichlet;      /* apply the operator's Dirichlet update after every accepted step */
  int32_t fsal;           /* Tsit5: nonzero when the Dirichlet update cannot change u after the
                             first accepted step (none, or values constant in t and u): f(u)
                             after an accepted step is then the step's last stage, not a new
                             evaluation (first same as last) */
  int64_t n;              /* state length */
  double t0, tf;
  double dt;              /* the fixed step; Tsit5: the first step (<= 0: automatic) */
  double abstol, reltol;  /* Tsit5 error control */
  int64_t maxiters;       /* Tsit5: accepted + rejected steps */
  int64_t n_save;         /* number of saveat times (0: none) */
  const double* save_t;   /* host [n_save], ascending */
  double* d_saves;        /* device [n_save * n]: the state at save_t[k] at d_saves + k n */
} fvm_ode_opts;

typedef struct fvm_ode_stats {
  int64_t nf;             /* evaluations of f */
  int64_t naccept, nreject;
  int64_t nsaved;         /* save slots written */
  int32_t retcode;        /* fvm_ode_retcode */
  int32_t reserved;
  double t;               /* time reached */
} fvm_ode_stats;

/* d_u [n]: the initial state in, the state at st->t out.  Blocks until done. */
int fvm_ode_solve(const fvm_ode_opts* o, double* d_u, fvm_ode_stats* st, fvm_stream s);

/* One Arnoldi step of restarted GMRES for A = a I + b J(d_u, t), J the Jacobian of an
 * operator (fvm_op_kind: fvm_jvp, fvm_system_jvp, or the template matrix itself):
 * w = A d_V[j], then classical Gram-Schmidt applied twice against d_V[0..j] on the
 * device, the coefficients never leaving it until one final read.  h [2 (j+1) + 1]
 * receives the first pass's coefficients, the second's, and |w|^2; the caller
 * normalises w into d_V[j+1].  The Newton-Krylov stage solves of TR-BDF2
 * (W = I - d h J) and of SteadyFVMProblem (J), which the reference does with KLU
 * or a Krylov method from LinearSolve (solve.jl:167-220).  Blocks until done. */
int fvm_arnoldi_step(int32_t op_kind, void* op, int64_t n, const double* d_u, double t, double a, double b,
                     int32_t j, double* const* d_V, double* d_w, double* h, fvm_stream s);
/* fvm_arnoldi_step for the right-preconditioned operator A diag(d_scale): w =
 * (a I + b J)(d_scale * d_V[j]), the product's input staged in d_tmp (n doubles, not
 * d_w and not a basis vector).  With d_scale the inverse diagonal of A this is the
 * Jacobi-preconditioned GMRES of TR-BDF2's and SteadyFVMProblem's Newton steps,
 * where the reference factorises the Jacobian with KLU (solve.jl:50-77, 167-183).
 * Blocks until done. */
int fvm_arnoldi_step_scaled(int32_t op_kind, void* op, int64_t n, const double* d_u, double t, double a,
                            double b, const double* d_scale, double* d_tmp, int32_t j, double* const* d_V,
                            double* d_w, double* h, fvm_stream s);

#ifdef __cplusplus
}
#endif
#endif
