/* CPU restatement of FiniteVolumeMethod.jl v1.2.3's vertex-centred finite-volume
 * hot path: FVMGeometry, fvm_eqs! (serial and threaded), the Dirichlet update and
 * the linear-template assembly.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
 * (libfvmhip.so) never links, loads or calls it.
 *
 * Each function cites the reference file:line it restates.  Arithmetic follows the
 * reference token for token (same operation order, no FMA contraction: build with
 * -ffp-contract=off) so that geometry matches the reference bit for bit.
 *
 * Indices are 0-based (reference node k is node k-1 here).
 */
#ifndef FVM_ORACLE_H
#define FVM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-triangle record, 24 doubles in input-triangle order:
 *   [0..8]   s1..s9            (TriangleProperties.shape_function_coefficients)
 *   [9..14]  CV-edge midpoints m1x m1y m2x m2y m3x m3y  (edges (i,j),(j,k),(k,i))
 *   [15..20] unit normals      n1x n1y n2x n2y n3x n3y
 *   [21..23] lengths           l1 l2 l3                  (geometry.jl:21-26) */
#define ORC_PROPS 24

/* FVMGeometry(tri), geometry.jl:99-169.  cv has n_points entries, zero for points
 * that are no triangle's vertex. */
void orc_geometry(int64_t n_points, const double* x, const double* y, int64_t n_tri,
                  const int32_t* tri, double* cv, double* props);

typedef struct orc_mesh {
  int64_t n_points, n_tri, n_bedges;
  const double *x, *y;
  const int32_t* tri;      /* 3*n_tri, counter-clockwise */
  const double* props;     /* ORC_PROPS*n_tri */
  const double* cv;        /* n_points */
  const int32_t* bedges;   /* 2*n_bedges, counter-clockwise boundary edges (i,j) */
  const int32_t* badj;     /* n_bedges: triangle holding the directed edge i->j */
  const uint8_t* is_vertex;/* n_points: DelaunayTriangulation.has_vertex */
} orc_mesh;

/* Flattened Conditions (conditions.jl:310-324, 486-552):
 *   node_kind: 0 none, 1 Dirichlet, 2 Dudt (a node in both Dicts is Dirichlet,
 *   source_contributions.jl:7-11)
 *   node_fidx: function index of the Dirichlet / Dudt condition
 *   bedge_neumann_fidx: function index if the oriented edge is in neumann_edges, else -1 */
typedef struct orc_conditions {
  const int8_t* node_kind;
  const int32_t* node_fidx;
  const int32_t* bedge_neumann_fidx;
} orc_conditions;

/* The user's closures.  flux is q(x,y,t,alpha,beta,gamma,p) (problem.jl:113-116);
 * when the problem was built from a diffusion function it is the closure of
 * construct_flux_function (problem.jl:425-440).  tri is the triangle whose
 * shape function is in use (a closure may use it to locate (x,y)).
 * source is S(x,y,t,u,p); cond is the fidx-th condition function a(x,y,t,u,p). */
typedef void (*orc_flux_fn)(double x, double y, double t, double a, double b, double g,
                            int64_t tri, void* ctx, double* qx, double* qy);
typedef double (*orc_source_fn)(double x, double y, double t, double u, int64_t node, void* ctx);
typedef double (*orc_cond_fn)(int32_t fidx, double x, double y, double t, double u, void* ctx);
typedef struct orc_physics {
  orc_flux_fn flux;
  orc_source_fn source;
  orc_cond_fn cond;
  void* ctx;
} orc_physics;

/* serial_fvm_eqs!, main_equations.jl:38-44: zero -> triangles (triangle_contributions.jl:28-43)
 * -> boundary edges (boundary_edge_contributions.jl:41-94) -> sources
 * (source_contributions.jl:2-14, 33-47, 71-76). */
void orc_fvm_eqs_serial(const orc_mesh* m, const orc_conditions* c, const orc_physics* ph,
                        const double* u, double t, double* du);

/* parallel_fvm_eqs!, main_equations.jl:47-81 with get_multithreading_parameters
 * (solve.jl:1-27): nthreads private du columns (chunks of the triangle and
 * boundary-edge lists), a serial column-sum, then the source pass.  scratch holds
 * n_points*nthreads doubles.  Uses OpenMP with nthreads threads. */
void orc_fvm_eqs_parallel(const orc_mesh* m, const orc_conditions* c, const orc_physics* ph,
                          const double* u, double t, double* du, int nthreads, double* scratch);

/* Per-node magnitude scale of fvm_eqs! (sum |Q| / V_i + |S_i|), for parity bounds. */
void orc_fvm_scale(const orc_mesh* m, const orc_conditions* c, const orc_physics* ph,
                   const double* u, double t, double* sc);

/* update_dirichlet_nodes!, dirichlet.jl:38-43: u[i] = a(x_i, y_i, t, u_i). */
void orc_update_dirichlet(const orc_mesh* m, const orc_conditions* c, const orc_physics* ph,
                          double t, double* u);

/* ---- FVMSystem (problem.jl:196-413) ----
 * nv variables on one mesh; u and du are [n_points][nv] (du[var, i] of the reference,
 * column-major, at var + nv*i).  Per-variable conditions: node_kind / node_fidx are
 * [nv][n_points], bedge_neumann_fidx is [nv][n_bedges]; function indices are already
 * mapped into the merged function list (map_fidx, problem.jl:322).
 * flux(x, y, t, alpha[nv], beta[nv], gamma[nv]) returns every variable's flux
 * (eval_flux_function over the tuple of flux functions, problem.jl:354-357);
 * source(var, x, y, t, u[nv]) and cond(fidx, x, y, t, u[nv]) take the vector of
 * every variable's value. */
typedef void (*orc_sys_flux_fn)(double x, double y, double t, const double* a, const double* b,
                                const double* g, void* ctx, double* qx, double* qy);
typedef double (*orc_sys_source_fn)(int32_t var, double x, double y, double t, const double* u,
                                    void* ctx);
typedef double (*orc_sys_cond_fn)(int32_t fidx, double x, double y, double t, const double* u,
                                  void* ctx);
typedef struct orc_system {
  int32_t nv;
  const int8_t* node_kind;
  const int32_t* node_fidx;
  const int32_t* bedge_neumann_fidx;
  orc_sys_flux_fn flux;
  orc_sys_source_fn source;
  orc_sys_cond_fn cond;
  void* ctx;
} orc_system;

/* serial_fvm_eqs! for an FVMSystem: triangles (triangle_contributions.jl:18-35 and the
 * system get_flux, individual_flux_contributions.jl:36-50), boundary edges
 * (boundary_edge_contributions.jl:15-72, 76-86), sources
 * (source_contributions.jl:17-30, 49-68).  Requires nv <= 8. */
void orc_fvm_eqs_system(const orc_mesh* m, const orc_system* s, const double* u, double t,
                        double* du);
/* Per (node, variable) magnitude scale sum |Q| / V + |S| of orc_fvm_eqs_system. */
void orc_fvm_scale_system(const orc_mesh* m, const orc_system* s, const double* u, double t,
                          double* sc);
/* serial_update_dirichlet_nodes! for a system (dirichlet.jl:46-53): variable by
 * variable, u[var, i] = a(x_i, y_i, t, u[:, i]). */
void orc_update_dirichlet_system(const orc_mesh* m, const orc_system* s, double t, double* u);

/* ---- built-in closures (C versions of the closures used by the workloads) ---- */
enum {
  ORC_FLUX_POLY_U = 0, /* q = -(d0 + d1 u + d2 u^2 + d3 u^3) (alpha, beta), u = alpha x + beta y + gamma */
  ORC_FLUX_NODAL = 1,  /* q = -D(x,y) (alpha, beta), D the P1 interpolant of exp(logD_v) on tri */
  ORC_FLUX_DXY = 2     /* q = -dscale * exp(0.5 sin(4 pi x) cos(4 pi y)) (alpha, beta) */
};
typedef struct orc_builtin {
  int flux_kind;
  double dpoly[4];
  double dscale;
  const double* logD;       /* n_points, ORC_FLUX_NODAL */
  const double* props;      /* the mesh's props, ORC_FLUX_NODAL */
  const int32_t* tri;       /* the mesh's triangles, ORC_FLUX_NODAL */
  double spoly[4];          /* S = s0 + s1 u + s2 u^2 + s3 u^3 (+ s_node[node] if given) */
  const double* s_node;
  int32_t n_fn;             /* condition functions a = c0 + cu u + ct t (+ cx x + cy y) */
  const double* fn_c0;
  const double* fn_cu;
  const double* fn_ct;
  const double* fn_cx;
  const double* fn_cy;
} orc_builtin;
void orc_builtin_flux(double x, double y, double t, double a, double b, double g, int64_t tri,
                      void* ctx, double* qx, double* qy);
double orc_builtin_source(double x, double y, double t, double u, int64_t node, void* ctx);
double orc_builtin_cond(int32_t fidx, double x, double y, double t, double u, void* ctx);

/* Complex-step version of orc_fvm_eqs_serial with the built-in closures: u and
 * logD are complex (re, im interleaved), du likewise.  Im(du(u + i h v))/h is the
 * Jacobian-vector product to machine precision (Squire & Trapp 1998). */
void orc_fvm_eqs_serial_cs(const orc_mesh* m, const orc_conditions* c, const orc_builtin* b,
                           const double* u_c, const double* logD_c, double t, double* du_c);

/* ---- template assembly (specific_problems/abstract_templates.jl) ----
 * Writes the COO triplets the reference adds into its dense matrix, in the
 * reference's order, and the vector b.  Duplicates are summed by the caller
 * (sparse(A) does the same).  The template diffusion function is D(x,y,p).
 * kind: 0 Poisson / Laplace (poissons_equation.jl:58-88, laplaces_equation.jl:51-78),
 * 1 MeanExitTime (mean_exit_time.jl:57-94), 2 DiffusionEquation /
 * LinearReactionDiffusionEquation (diffusion_equation.jl:69-101,
 * linear_reaction_diffusion_equations.jl:76-125; A and b of du/dt = A u + b,
 * without the augmented row/column).
 * f_node: Poisson source values f(x_i, y_i) per node (ignored otherwise).
 * neu_a: Neumann function values at the two half-edge points of each boundary
 * edge (2*n_bedges, ignored for non-Neumann edges).  dir_val / dudt_val: condition
 * values per node for Dirichlet / Dudt nodes.  r_node: linear reaction
 * coefficient added to A_ii of rows without a condition (NULL: none).
 * Identity rows (A_ii = 1.0 assignments in the reference) are emitted once.
 * Returns the number of triplets written (call with rows == NULL to count). */
typedef double (*orc_dxy_fn)(double x, double y, void* ctx);
int64_t orc_template_assemble(const orc_mesh* m, const orc_conditions* c, int kind,
                              orc_dxy_fn D, void* Dctx, const double* f_node,
                              const double* neu_a, const double* dir_val, const double* dudt_val,
                              const double* r_node, int64_t* rows, int64_t* cols, double* vals, double* bvec);
double orc_dxy_exp_sincos(double x, double y, void* ctx); /* *(double*)ctx * exp(0.5 sin 4pi x cos 4pi y) */
double orc_dxy_const(double x, double y, void* ctx);      /* *(double*)ctx */

#ifdef __cplusplus
}
#endif
#endif
