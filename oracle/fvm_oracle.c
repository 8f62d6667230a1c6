/* CPU restatement of FiniteVolumeMethod.jl's vertex-centred FVM hot path.
 * TEST INFRASTRUCTURE ONLY (the checker); see fvm_oracle.h. */
#include "fvm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* FVMGeometry(tri), geometry.jl:99-169.                                      */
/* ------------------------------------------------------------------------ */
void orc_geometry(int64_t n_points, const double* x, const double* y, int64_t n_tri,
                  const int32_t* tri, double* cv, double* props) {
  for (int64_t i = 0; i < n_points; ++i) cv[i] = 0.0;
  for (int64_t T = 0; T < n_tri; ++T) {
    const int32_t i = tri[3 * T], j = tri[3 * T + 1], k = tri[3 * T + 2];
    const double px = x[i], py = y[i], qx = x[j], qy = y[j], rx = x[k], ry = y[k];
    /* DelaunayTriangulation statistics: centroid (p+q+r)/3, edge midpoints of
     * (p,q), (q,r), (r,p) (pinned approximately by test/geometry.jl:45-48). */
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
    const double S1 = 1.0 / 2 * fabs(pcx * m13y - pcy * m13x);
    const double S2 = 1.0 / 2 * fabs(qcx * m21y - qcy * m21x);
    const double S3 = 1.0 / 2 * fabs(rcx * m32y - rcy * m32x);
    cv[i] += S1;
    cv[j] += S2;
    cv[k] += S3;
    const double D = qx * ry - qy * rx - px * ry + rx * py + px * qy - qx * py;
    double* P = props + ORC_PROPS * T;
    P[0] = (qy - ry) / D;
    P[1] = (ry - py) / D;
    P[2] = (py - qy) / D;
    P[3] = (rx - qx) / D;
    P[4] = (px - rx) / D;
    P[5] = (qx - px) / D;
    P[6] = (qx * ry - rx * qy) / D;
    P[7] = (rx * py - px * ry) / D;
    P[8] = (px * qy - qx * py) / D;
    P[9] = (m1x + cx) / 2;
    P[10] = (m1y + cy) / 2;
    P[11] = (m2x + cx) / 2;
    P[12] = (m2y + cy) / 2;
    P[13] = (m3x + cx) / 2;
    P[14] = (m3y + cy) / 2;
    const double e1x = cx - m1x, e1y = cy - m1y;
    const double e2x = cx - m2x, e2y = cy - m2y;
    const double e3x = cx - m3x, e3y = cy - m3y;
    /* LinearAlgebra.norm of a 2-tuple: sqrt(a*a + b*b) */
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

/* ------------------------------------------------------------------------ */
/* Built-in closures.                                                         */
/* ------------------------------------------------------------------------ */
static const double ORC_PI = 3.14159265358979323846;

/* construct_flux_function's closure (problem.jl:428-434) for D given by the kind. */
void orc_builtin_flux(double x, double y, double t, double a, double b, double g, int64_t tri,
                      void* ctx, double* qx, double* qy) {
  const orc_builtin* B = (const orc_builtin*)ctx;
  (void)t;
  double Dval;
  if (B->flux_kind == ORC_FLUX_POLY_U) {
    double u = a * x + b * y + g;
    Dval = B->dpoly[0] + u * (B->dpoly[1] + u * (B->dpoly[2] + u * B->dpoly[3]));
  } else if (B->flux_kind == ORC_FLUX_NODAL) {
    /* pl_interpolate (geometry.jl:192-196) of the nodal field exp(logD) on tri */
    const double* P = B->props + ORC_PROPS * tri;
    const int32_t i = B->tri[3 * tri], j = B->tri[3 * tri + 1], k = B->tri[3 * tri + 2];
    double Di = exp(B->logD[i]), Dj = exp(B->logD[j]), Dk = exp(B->logD[k]);
    double da = P[0] * Di + P[1] * Dj + P[2] * Dk;
    double db = P[3] * Di + P[4] * Dj + P[5] * Dk;
    double dg = P[6] * Di + P[7] * Dj + P[8] * Dk;
    Dval = da * x + db * y + dg;
  } else {
    Dval = B->dscale * exp(0.5 * sin(4 * ORC_PI * x) * cos(4 * ORC_PI * y));
  }
  *qx = -Dval * a;
  *qy = -Dval * b;
}

double orc_builtin_source(double x, double y, double t, double u, int64_t node, void* ctx) {
  const orc_builtin* B = (const orc_builtin*)ctx;
  (void)x; (void)y; (void)t;
  double S = B->spoly[0] + u * (B->spoly[1] + u * (B->spoly[2] + u * B->spoly[3]));
  if (B->s_node) S = B->s_node[node] + S;
  return S;
}

double orc_builtin_cond(int32_t fidx, double x, double y, double t, double u, void* ctx) {
  const orc_builtin* B = (const orc_builtin*)ctx;
  double v = B->fn_c0[fidx] + B->fn_cu[fidx] * u;
  v = v + B->fn_ct[fidx] * t;
  if (B->fn_cx) v = v + B->fn_cx[fidx] * x;
  if (B->fn_cy) v = v + B->fn_cy[fidx] * y;
  return v;
}

/* ------------------------------------------------------------------------ */
/* Real instantiation (user closures as callbacks).                          */
/* ------------------------------------------------------------------------ */
#define REAL double
#define SFX _real
#define PHYS_DECL const orc_physics* ph
#define FLUX(x, y, t, a, b, g, T, qx, qy) ph->flux(x, y, t, a, b, g, T, ph->ctx, qx, qy)
#define SRC(x, y, t, u, n) ph->source(x, y, t, u, n, ph->ctx)
#define COND(f, x, y, t, u) ph->cond(f, x, y, t, u, ph->ctx)
#include "orc_rhs_body.inc"
#undef REAL
#undef SFX
#undef PHYS_DECL
#undef FLUX
#undef SRC
#undef COND

void orc_fvm_eqs_serial(const orc_mesh* m, const orc_conditions* c, const orc_physics* ph,
                        const double* u, double t, double* du) {
  for (int64_t i = 0; i < m->n_points; ++i) du[i] = 0.0;
  orc_tri_range_real(m, ph, u, t, 0, m->n_tri, du);
  orc_bedge_range_real(m, c, ph, u, t, 0, m->n_bedges, du);
  orc_source_range_real(m, c, ph, u, t, 0, m->n_points, du);
}

/* ChunkSplitters.index_chunks(1:len; n): n contiguous chunks, the first len % n one longer. */
static void chunk(int64_t len, int n, int k, int64_t* a, int64_t* b) {
  int64_t base = len / n, rem = len % n;
  *a = k * base + (k < rem ? k : rem);
  *b = *a + base + (k < rem ? 1 : 0);
}

void orc_fvm_eqs_parallel(const orc_mesh* m, const orc_conditions* c, const orc_physics* ph,
                          const double* u, double t, double* du, int nthreads, double* scratch) {
  const int64_t N = m->n_points;
  for (int64_t i = 0; i < N; ++i) du[i] = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int k = 0; k < nthreads; ++k) {
    double* col = scratch + (int64_t)k * N;
    memset(col, 0, sizeof(double) * (size_t)N);
    int64_t a, b;
    chunk(m->n_tri, nthreads, k, &a, &b);
    orc_tri_range_real(m, ph, u, t, a, b, col);
  }
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int k = 0; k < nthreads; ++k) {
    int64_t a, b;
    chunk(m->n_bedges, nthreads, k, &a, &b);
    orc_bedge_range_real(m, c, ph, u, t, a, b, scratch + (int64_t)k * N);
  }
  /* combine_duplicated_du!: serial column sum (main_equations.jl:70-81) */
  for (int k = 0; k < nthreads; ++k) {
    const double* col = scratch + (int64_t)k * N;
    for (int64_t i = 0; i < N; ++i) du[i] += col[i];
  }
  /* get_parallel_source_contributions!: disjoint writes */
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < N; ++i) orc_source_range_real(m, c, ph, u, t, i, i + 1, du);
}


/* Per-node scale of the parity criterion: sum over the node's CV edges of |Q|,
 * divided by V_i, plus |S_i| (0 at Dirichlet / non-vertex nodes, |a| at Dudt
 * nodes).  Same loops as orc_fvm_eqs_serial with every term taken by magnitude. */
void orc_fvm_scale(const orc_mesh* m, const orc_conditions* c, const orc_physics* ph,
                   const double* u, double t, double* sc) {
  for (int64_t i = 0; i < m->n_points; ++i) sc[i] = 0.0;
  for (int64_t T = 0; T < m->n_tri; ++T) {
    const int32_t i = m->tri[3 * T], j = m->tri[3 * T + 1], k = m->tri[3 * T + 2];
    const double* P = m->props + ORC_PROPS * T;
    double alpha = P[0] * u[i] + P[1] * u[j] + P[2] * u[k];
    double beta = P[3] * u[i] + P[4] * u[j] + P[5] * u[k];
    double gamma = P[6] * u[i] + P[7] * u[j] + P[8] * u[k];
    double Q[3];
    for (int e = 0; e < 3; ++e) {
      double qx, qy;
      ph->flux(P[9 + 2 * e], P[10 + 2 * e], t, alpha, beta, gamma, T, ph->ctx, &qx, &qy);
      Q[e] = fabs((qx * P[15 + 2 * e] + qy * P[16 + 2 * e]) * P[21 + e]);
    }
    sc[i] += Q[2] + Q[0];
    sc[j] += Q[0] + Q[1];
    sc[k] += Q[1] + Q[2];
  }
  double* tmp = (double*)calloc((size_t)m->n_points, sizeof(double));
  orc_bedge_range_real(m, c, ph, u, t, 0, m->n_bedges, tmp);
  for (int64_t i = 0; i < m->n_points; ++i) sc[i] += fabs(tmp[i]);
  free(tmp);
  for (int64_t i = 0; i < m->n_points; ++i) {
    if (!m->is_vertex[i]) { sc[i] = 0.0; continue; }
    const int kind = c->node_kind[i];
    if (kind == 0)
      sc[i] = sc[i] / m->cv[i] + fabs(ph->source(m->x[i], m->y[i], t, u[i], i, ph->ctx));
    else if (kind == 1)
      sc[i] = 0.0;
    else
      sc[i] = fabs(ph->cond(c->node_fidx[i], m->x[i], m->y[i], t, u[i], ph->ctx));
  }
}

void orc_update_dirichlet(const orc_mesh* m, const orc_conditions* c, const orc_physics* ph,
                          double t, double* u) {
  for (int64_t i = 0; i < m->n_points; ++i)
    if (c->node_kind[i] == 1) u[i] = ph->cond(c->node_fidx[i], m->x[i], m->y[i], t, u[i], ph->ctx);
}

/* ------------------------------------------------------------------------ */
/* Complex-step instantiation (built-in closures only).                      */
/* ------------------------------------------------------------------------ */
typedef double complex cplx;
typedef struct {
  const orc_builtin* b;
  const cplx* logD;
} orc_cs_phys;

static void cs_flux(const orc_cs_phys* P, double x, double y, double t, cplx a, cplx b, cplx g,
                    int64_t tri, cplx* qx, cplx* qy) {
  const orc_builtin* B = P->b;
  (void)t;
  cplx Dval;
  if (B->flux_kind == ORC_FLUX_POLY_U) {
    cplx u = a * x + b * y + g;
    Dval = B->dpoly[0] + u * (B->dpoly[1] + u * (B->dpoly[2] + u * B->dpoly[3]));
  } else if (B->flux_kind == ORC_FLUX_NODAL) {
    const double* Pr = B->props + ORC_PROPS * tri;
    const int32_t i = B->tri[3 * tri], j = B->tri[3 * tri + 1], k = B->tri[3 * tri + 2];
    cplx Di = cexp(P->logD[i]), Dj = cexp(P->logD[j]), Dk = cexp(P->logD[k]);
    cplx da = Pr[0] * Di + Pr[1] * Dj + Pr[2] * Dk;
    cplx db = Pr[3] * Di + Pr[4] * Dj + Pr[5] * Dk;
    cplx dg = Pr[6] * Di + Pr[7] * Dj + Pr[8] * Dk;
    Dval = da * x + db * y + dg;
  } else {
    Dval = B->dscale * exp(0.5 * sin(4 * ORC_PI * x) * cos(4 * ORC_PI * y));
  }
  *qx = -Dval * a;
  *qy = -Dval * b;
}
static cplx cs_source(const orc_cs_phys* P, double x, double y, double t, cplx u, int64_t node) {
  const orc_builtin* B = P->b;
  (void)x; (void)y; (void)t;
  cplx S = B->spoly[0] + u * (B->spoly[1] + u * (B->spoly[2] + u * B->spoly[3]));
  if (B->s_node) S = B->s_node[node] + S;
  return S;
}
static cplx cs_cond(const orc_cs_phys* P, int32_t f, double x, double y, double t, cplx u) {
  const orc_builtin* B = P->b;
  cplx v = B->fn_c0[f] + B->fn_cu[f] * u;
  v = v + B->fn_ct[f] * t;
  if (B->fn_cx) v = v + B->fn_cx[f] * x;
  if (B->fn_cy) v = v + B->fn_cy[f] * y;
  return v;
}

#define REAL cplx
#define SFX _cs
#define PHYS_DECL const orc_cs_phys* ph
#define FLUX(x, y, t, a, b, g, T, qx, qy) cs_flux(ph, x, y, t, a, b, g, T, qx, qy)
#define SRC(x, y, t, u, n) cs_source(ph, x, y, t, u, n)
#define COND(f, x, y, t, u) cs_cond(ph, f, x, y, t, u)
#include "orc_rhs_body.inc"
#undef REAL
#undef SFX
#undef PHYS_DECL
#undef FLUX
#undef SRC
#undef COND

void orc_fvm_eqs_serial_cs(const orc_mesh* m, const orc_conditions* c, const orc_builtin* b,
                           const double* u_c, const double* logD_c, double t, double* du_c) {
  const cplx* u = (const cplx*)u_c;
  cplx* du = (cplx*)du_c;
  orc_cs_phys P = {b, (const cplx*)logD_c};
  for (int64_t i = 0; i < m->n_points; ++i) du[i] = 0.0;
  orc_tri_range_cs(m, &P, u, t, 0, m->n_tri, du);
  orc_bedge_range_cs(m, c, &P, u, t, 0, m->n_bedges, du);
  orc_source_range_cs(m, c, &P, u, t, 0, m->n_points, du);
}

/* ------------------------------------------------------------------------ */
/* Template assembly (specific_problems/abstract_templates.jl).               */
/* ------------------------------------------------------------------------ */
double orc_dxy_exp_sincos(double x, double y, void* ctx) {
  return *(const double*)ctx * exp(0.5 * sin(4 * ORC_PI * x) * cos(4 * ORC_PI * y));
}
double orc_dxy_const(double x, double y, void* ctx) {
  (void)x; (void)y;
  return *(const double*)ctx;
}

typedef struct {
  int64_t n;
  int64_t *r, *c;
  double* v;
} coo;
static void push(coo* A, int64_t r, int64_t c, double v) {
  if (A->r) {
    A->r[A->n] = r;
    A->c[A->n] = c;
    A->v[A->n] = v;
  }
  A->n++;
}

static int has_cond(const orc_conditions* c, int64_t i) { return c->node_kind[i] != 0; }

/* boundary CV components (control_volumes.jl:41-56) */
static void bcv(const orc_mesh* m, int64_t e, double* nx, double* ny, double* mix, double* miy,
                double* mjx, double* mjy, double* ell) {
  const int32_t i = m->bedges[2 * e], j = m->bedges[2 * e + 1];
  double px = m->x[i], py = m->y[i], qx = m->x[j], qy = m->y[j];
  double lij = sqrt((qx - px) * (qx - px) + (qy - py) * (qy - py));
  *nx = (qy - py) / lij;
  *ny = -(qx - px) / lij;
  double mijx = (px + qx) / 2, mijy = (py + qy) / 2;
  *mix = (px + mijx) / 2;
  *miy = (py + mijy) / 2;
  *mjx = (qx + mijx) / 2;
  *mjy = (qy + mijy) / 2;
  *ell = sqrt((mijx - px) * (mijx - px) + (mijy - py) * (mijy - py));
}

int64_t orc_template_assemble(const orc_mesh* m, const orc_conditions* c, int kind,
                              orc_dxy_fn Dfn, void* Dctx, const double* f_node,
                              const double* neu_a, const double* dir_val, const double* dudt_val,
                              const double* r_node, int64_t* rows, int64_t* cols, double* vals, double* bvec) {
  coo A = {0, rows, cols, vals};
  const int64_t N = m->n_points;
  if (bvec) for (int64_t i = 0; i < N; ++i) bvec[i] = 0.0;
  /* create_rhs_b (abstract_templates.jl:278-288) for Poisson */
  if (kind == 0 && bvec)
    for (int64_t i = 0; i < N; ++i)
      if (m->is_vertex[i] && c->node_kind[i] != 1) bvec[i] = f_node[i];
  /* triangle_contributions! (abstract_templates.jl:73-99) */
  for (int64_t T = 0; T < m->n_tri; ++T) {
    const int32_t ijk[3] = {m->tri[3 * T], m->tri[3 * T + 1], m->tri[3 * T + 2]};
    const double* P = m->props + ORC_PROPS * T;
    for (int e = 0; e < 3; ++e) {
      const int32_t e1 = ijk[e], e2 = ijk[(e + 1) % 3];
      double x = P[9 + 2 * e], y = P[10 + 2 * e];
      double nx = P[15 + 2 * e], ny = P[16 + 2 * e];
      double ell = P[21 + e];
      double D = Dfn(x, y, Dctx);
      double Dl = D * ell;
      double a123[3] = {Dl * (P[0] * nx + P[3] * ny), Dl * (P[1] * nx + P[4] * ny),
                        Dl * (P[2] * nx + P[5] * ny)};
      int h1 = has_cond(c, e1), h2 = has_cond(c, e2);
      for (int v = 0; v < 3; ++v) {
        if (!h1) push(&A, e1, ijk[v], a123[v] / m->cv[e1]);
        if (!h2) push(&A, e2, ijk[v], -(a123[v] / m->cv[e2]));
      }
    }
  }
  if (kind == 0 || kind == 2) {
    /* non_neumann_boundary_edge_contributions! (abstract_templates.jl:237-267);
     * the j row is divided by V_i as at :262 */
    for (int64_t e = 0; e < m->n_bedges; ++e) {
      if (c->bedge_neumann_fidx[e] >= 0) continue;
      const int32_t i = m->bedges[2 * e], j = m->bedges[2 * e + 1];
      double nx, ny, mix, miy, mjx, mjy, ell;
      bcv(m, e, &nx, &ny, &mix, &miy, &mjx, &mjy, &ell);
      const int64_t T = m->badj[e];
      const int32_t ijk[3] = {m->tri[3 * T], m->tri[3 * T + 1], m->tri[3 * T + 2]};
      const double* P = m->props + ORC_PROPS * T;
      double Di = Dfn(mix, miy, Dctx), Dj = Dfn(mjx, mjy, Dctx);
      int hi = has_cond(c, i), hj = has_cond(c, j);
      for (int v = 0; v < 3; ++v) {
        double ai = Di * ell * (P[v] * nx + P[3 + v] * ny);
        double aj = Dj * ell * (P[v] * nx + P[3 + v] * ny);
        if (!hi) push(&A, i, ijk[v], ai / m->cv[i]);
        if (!hj) push(&A, j, ijk[v], aj / m->cv[i]);
      }
    }
    /* neumann_boundary_edge_contributions! (abstract_templates.jl:175-191) */
    for (int64_t e = 0; e < m->n_bedges; ++e) {
      if (c->bedge_neumann_fidx[e] < 0) continue;
      const int32_t i = m->bedges[2 * e], j = m->bedges[2 * e + 1];
      double nx, ny, mix, miy, mjx, mjy, ell;
      bcv(m, e, &nx, &ny, &mix, &miy, &mjx, &mjy, &ell);
      double Di = Dfn(mix, miy, Dctx), Dj = Dfn(mjx, mjy, Dctx);
      double ai = neu_a[2 * e], aj = neu_a[2 * e + 1];
      if (bvec) {
        if (!has_cond(c, i)) bvec[i] += Di * ai * ell / m->cv[i];
        if (!has_cond(c, j)) bvec[j] += Dj * aj * ell / m->cv[j];
      }
    }
  }
  if (kind == 0) {
    /* apply_steady_dirichlet_conditions! (abstract_templates.jl:302-309) */
    for (int64_t i = 0; i < N; ++i)
      if (c->node_kind[i] == 1) {
        if (bvec) bvec[i] = dir_val[i];
        push(&A, i, i, 1.0);
      }
  } else if (kind == 1) {
    /* create_met_b! (mean_exit_time.jl:273-283) */
    for (int64_t i = 0; i < N; ++i) {
      if (!m->is_vertex[i]) continue;
      if (c->node_kind[i] != 1) {
        if (bvec) bvec[i] = -1.0;
      } else {
        push(&A, i, i, 1.0);
      }
    }
  } else {
    /* linear_source_contributions! (linear_reaction_diffusion_equations.jl:115-125) */
    if (r_node)
      for (int64_t i = 0; i < N; ++i)
        if (m->is_vertex[i] && !has_cond(c, i)) push(&A, i, i, r_node[i]);
    /* apply_dudt_conditions! (abstract_templates.jl:127-135) */
    if (bvec)
      for (int64_t i = 0; i < N; ++i)
        if (c->node_kind[i] == 2) bvec[i] = dudt_val[i];
  }
  /* fix_missing_vertices! (abstract_templates.jl:317-325); A_ii = 1.0 is an
   * assignment, so a Poisson Dirichlet row that is also a missing vertex keeps 1 */
  for (int64_t i = 0; i < N; ++i)
    if (!m->is_vertex[i]) {
      if (!(kind == 0 && c->node_kind[i] == 1)) push(&A, i, i, 1.0);
      if (bvec) bvec[i] = 0.0;
    }
  return A.n;
}

/* ------------------------------------------------------------------------ */
/* FVMSystem: serial fvm_eqs! over nv variables.                              */
/* ------------------------------------------------------------------------ */
#define ORC_MAXV 8

/* get_shape_function_coefficients for a system (shape_functions.jl:12-19) */
static void sys_shape(const orc_mesh* m, int nv, const double* u, int64_t T, double* al, double* be,
                      double* ga) {
  const int32_t i = m->tri[3 * T], j = m->tri[3 * T + 1], k = m->tri[3 * T + 2];
  const double* P = m->props + ORC_PROPS * T;
  for (int v = 0; v < nv; ++v) {
    al[v] = P[0] * u[v + nv * i] + P[1] * u[v + nv * j] + P[2] * u[v + nv * k];
    be[v] = P[3] * u[v + nv * i] + P[4] * u[v + nv * j] + P[5] * u[v + nv * k];
    ga[v] = P[6] * u[v + nv * i] + P[7] * u[v + nv * j] + P[8] * u[v + nv * k];
  }
}

static void sys_body(const orc_mesh* m, const orc_system* s, const double* u, double t, double* du,
                     int scale) {
  const int nv = s->nv;
  const int64_t N = m->n_points;
  for (int64_t i = 0; i < nv * N; ++i) du[i] = 0.0;
  double al[ORC_MAXV], be[ORC_MAXV], ga[ORC_MAXV], qx[ORC_MAXV], qy[ORC_MAXV];
  double Q[3][ORC_MAXV];
  /* triangle_contributions.jl:10-43 (system update_du! at :18-25) */
  for (int64_t T = 0; T < m->n_tri; ++T) {
    const int32_t i = m->tri[3 * T], j = m->tri[3 * T + 1], k = m->tri[3 * T + 2];
    const double* P = m->props + ORC_PROPS * T;
    sys_shape(m, nv, u, T, al, be, ga);
    for (int e = 0; e < 3; ++e) {
      const double x = P[9 + 2 * e], y = P[10 + 2 * e], nx = P[15 + 2 * e], ny = P[16 + 2 * e];
      const double ell = P[21 + e];
      s->flux(x, y, t, al, be, ga, s->ctx, qx, qy);
      for (int v = 0; v < nv; ++v) Q[e][v] = (qx[v] * nx + qy[v] * ny) * ell;
    }
    for (int v = 0; v < nv; ++v) {
      if (scale) {
        const double a1 = fabs(Q[0][v]), a2 = fabs(Q[1][v]), a3 = fabs(Q[2][v]);
        du[v + nv * i] += a3 + a1;
        du[v + nv * j] += a1 + a2;
        du[v + nv * k] += a2 + a3;
      } else {
        du[v + nv * i] = du[v + nv * i] + Q[2][v] - Q[0][v];
        du[v + nv * j] = du[v + nv * j] + Q[0][v] - Q[1][v];
        du[v + nv * k] = du[v + nv * k] + Q[1][v] - Q[2][v];
      }
    }
  }
  /* boundary_edge_contributions.jl:15-72, 76-86 */
  for (int64_t e = 0; e < m->n_bedges; ++e) {
    const int32_t i = m->bedges[2 * e], j = m->bedges[2 * e + 1];
    double nx, ny, mix, miy, mjx, mjy, ell;
    bcv(m, e, &nx, &ny, &mix, &miy, &mjx, &mjy, &ell);
    sys_shape(m, nv, u, m->badj[e], al, be, ga);
    double ui[ORC_MAXV], uj[ORC_MAXV], q1[ORC_MAXV], q2[ORC_MAXV];
    for (int v = 0; v < nv; ++v) {
      ui[v] = al[v] * mix + be[v] * miy + ga[v];
      uj[v] = al[v] * mjx + be[v] * mjy + ga[v];
    }
    double qxi[ORC_MAXV], qyi[ORC_MAXV], qxj[ORC_MAXV], qyj[ORC_MAXV];
    s->flux(mix, miy, t, al, be, ga, s->ctx, qxi, qyi);
    s->flux(mjx, mjy, t, al, be, ga, s->ctx, qxj, qyj);
    for (int v = 0; v < nv; ++v) {
      const int32_t f = s->bedge_neumann_fidx[v * m->n_bedges + e];
      if (f < 0) {
        q1[v] = (qxi[v] * nx + qyi[v] * ny) * ell;
        q2[v] = (qxj[v] * nx + qyj[v] * ny) * ell;
      } else {
        q1[v] = s->cond(f, mix, miy, t, ui, s->ctx) * ell;
        q2[v] = s->cond(f, mjx, mjy, t, uj, s->ctx) * ell;
      }
      if (scale) {
        du[v + nv * i] += fabs(q1[v]);
        du[v + nv * j] += fabs(q2[v]);
      } else {
        du[v + nv * i] = du[v + nv * i] - q1[v];
        du[v + nv * j] = du[v + nv * j] - q2[v];
      }
    }
  }
  /* source_contributions.jl:17-30, 49-68 */
  for (int64_t i = 0; i < N; ++i) {
    if (!m->is_vertex[i]) {
      for (int v = 0; v < nv; ++v) du[v + nv * i] = 0.0;
      continue;
    }
    const double* ui = u + nv * i;
    for (int v = 0; v < nv; ++v) {
      const int kd = s->node_kind[v * N + i];
      double S;
      if (kd == 0) S = s->source(v, m->x[i], m->y[i], t, ui, s->ctx);
      else if (kd == 1) S = 0.0;
      else S = s->cond(s->node_fidx[v * N + i], m->x[i], m->y[i], t, ui, s->ctx);
      if (scale) du[v + nv * i] = kd == 0 ? du[v + nv * i] / m->cv[i] + fabs(S) : fabs(S);
      else du[v + nv * i] = kd == 0 ? du[v + nv * i] / m->cv[i] + S : S;
    }
  }
}

void orc_fvm_eqs_system(const orc_mesh* m, const orc_system* s, const double* u, double t,
                        double* du) {
  sys_body(m, s, u, t, du, 0);
}

void orc_fvm_scale_system(const orc_mesh* m, const orc_system* s, const double* u, double t,
                          double* sc) {
  sys_body(m, s, u, t, sc, 1);
}

void orc_update_dirichlet_system(const orc_mesh* m, const orc_system* s, double t, double* u) {
  const int nv = s->nv;
  const int64_t N = m->n_points;
  for (int v = 0; v < nv; ++v)
    for (int64_t i = 0; i < N; ++i)
      if (s->node_kind[v * N + i] == 1)
        u[v + nv * i] = s->cond(s->node_fidx[v * N + i], m->x[i], m->y[i], t, u + nv * i, s->ctx);
}
