// Explicit time integration on the device in one call: the adaptive Tsitouras
// 5(4) loop with its PI step controller and the fixed-step Euler and SSPRK3
// loops, over one of the library's operators (a problem context, a system, or a
// linear template), with the Dirichlet update after every accepted step and the
// states at the requested save times written to device memory.
//
// Reference: src/solve.jl:133-220 hands fvm_eqs! to OrdinaryDiffEq with a
// DiscreteCallback(true, update_dirichlet_nodes!) after every step
// (dirichlet.jl:78-86) and saveat.  Here a step costs its kernel launches and one
// host synchronisation (the error norm of the adaptive loop), instead of a
// host-language call per stage: what matters on small meshes, where every kernel
// runs in a few microseconds.  The arithmetic is the package's host-language
// integrator's, in the same order (the same fused linear combinations, step
// control and Hermite dense output computed in double on the host), so the two
// give the same states bit for bit.
#include <algorithm>
#include <cmath>

#include "fvm_internal.hpp"

using namespace fvm;

namespace {

// Tsitouras (2011): nodes, stage coefficients (row s: a_s1 .. a_s,s; row 6 is the
// solution, first-same-as-last) and the error weights b - bhat.
constexpr double kC[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0};
constexpr double kA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774},
};
constexpr double kBt[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                           -0.1447110071732629,     0.5823571654525552,     -0.45808210592918697,
                           0.015151515151515152};

struct Op {
  int kind;
  void* h;
  bool dirichlet;
  int64_t* nf;

  int rhs(const double* u, double* du, double t, fvm_stream s) const {
    ++*nf;
    switch (kind) {
      case FVM_OP_PROBLEM: return fvm_rhs((fvm_ctx*)h, u, du, t, s);
      case FVM_OP_SYSTEM: return fvm_system_rhs((fvm_system*)h, u, du, t, s);
      default: return fvm_template_apply((const fvm_template*)h, u, du, s);
    }
  }
  int jvp(const double* u, const double* v, double* jv, double t, fvm_stream s) const {
    switch (kind) {
      case FVM_OP_PROBLEM: return fvm_jvp((fvm_ctx*)h, u, v, jv, t, s);
      case FVM_OP_SYSTEM: return fvm_system_jvp((fvm_system*)h, u, v, jv, t, s);
      default: return fvm_template_spmv((const fvm_template*)h, 0, v, jv, s);
    }
  }
  int apply_dirichlet(double* u, double t, fvm_stream s) const {
    if (!dirichlet) return FVM_OK;
    switch (kind) {
      case FVM_OP_PROBLEM: return fvm_dirichlet_apply((fvm_ctx*)h, u, t, s);
      case FVM_OP_SYSTEM: return fvm_system_dirichlet_apply((fvm_system*)h, u, t, s);
      default: return FVM_OK;
    }
  }
};

// Device work vectors of one solve, released on every exit path.
struct Work {
  double* base = nullptr;
  ~Work() {
    if (base) (void)hipFree(base);
  }
};

int lc(int64_t n, double* out, int k, const double* c, const double* const* x, fvm_stream s) {
  return fvm_vec_lincomb(n, out, k, c, x, s);
}

int norm(int64_t n, const double* e, const double* a, const double* b, double abstol, double reltol,
         double* r, fvm_stream s) {
  return fvm_vec_err_norm(n, e, a, b, abstol, reltol, r, s);
}

// Hairer, Norsett & Wanner (1993), II.4: the starting step, (0.01 / m)^(1/5).
int initial_dt(const Op& op, int64_t n, const double* u0, const double* f0, double t0, double tf, double abstol,
               double reltol, double* tmp, double* f1, fvm_stream s, double* out) {
  double d0, d1, d2;
  int rc;
  if ((rc = norm(n, u0, u0, u0, abstol, reltol, &d0, s)) || (rc = norm(n, f0, u0, u0, abstol, reltol, &d1, s)))
    return rc;
  double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
  dt0 = std::min(dt0, tf - t0);
  {
    const double c[2] = {1.0, dt0};
    const double* x[2] = {u0, f0};
    if ((rc = lc(n, tmp, 2, c, x, s)) || (rc = op.rhs(tmp, f1, t0 + dt0, s))) return rc;
  }
  {
    const double c[2] = {1.0 / dt0, -1.0 / dt0};
    const double* x[2] = {f1, f0};
    if ((rc = lc(n, tmp, 2, c, x, s)) || (rc = norm(n, tmp, u0, u0, abstol, reltol, &d2, s))) return rc;
  }
  const double m = std::max(d1, d2);
  const double dt1 = m <= 1e-15 ? std::max(1e-6, dt0 * 1e-3) : std::pow(0.01 / m, 1.0 / 5.0);
  *out = std::min(std::min(100 * dt0, dt1), tf - t0);
  return FVM_OK;
}

int save(const fvm_ode_opts* o, int64_t k, const double* v, fvm_stream s) {
  FVM_HIP(hipMemcpyAsync(o->d_saves + k * o->n, v, sizeof(double) * (size_t)o->n, hipMemcpyDeviceToDevice,
                         (hipStream_t)s));
  return FVM_OK;
}

int tsit5(const fvm_ode_opts* o, const Op& op, double* d_u, fvm_ode_stats* st, fvm_stream s) {
  const int64_t n = o->n;
  Work w;
  FVM_HIP(hipMalloc((void**)&w.base, sizeof(double) * (size_t)(n > 0 ? n : 1) * 11));
  double* k[7];
  for (int j = 0; j < 7; ++j) k[j] = w.base + j * n;
  double *utmp = w.base + 7 * n, *unew = w.base + 8 * n, *interp = w.base + 9 * n, *u = w.base + 10 * n;
  FVM_HIP(hipMemcpyAsync(u, d_u, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, (hipStream_t)s));
  double t = o->t0;
  const double tf = o->tf;
  int64_t ks = 0;
  int rc = op.rhs(u, k[0], t, s);
  if (rc) return rc;
  double dt = o->dt;
  if (!(dt > 0)) {
    if ((rc = initial_dt(op, n, u, k[0], t, tf, o->abstol, o->reltol, utmp, unew, s, &dt))) return rc;
  }
  double qold = 1e-4;
  while (ks < o->n_save && o->save_t[ks] <= t + 1e-14 * std::max(1.0, std::fabs(t)))
    if ((rc = save(o, ks++, u, s))) return rc;
  int64_t it = 0;
  st->retcode = FVM_ODE_SUCCESS;
  if (op.kind == FVM_OP_TEMPLATE && !op.dirichlet) {
    // small templates: the whole loop in one workgroup (same arithmetic per element)
    rc = fvm::template_tsit5_small((const fvm_template*)op.h, o, k, utmp, u, unew, d_u, t, dt, it, ks, kA, kC, kBt,
                                   st, (hipStream_t)s);
    if (rc != FVM_ERR_STATE) return rc;
    rc = FVM_OK;
  }
  while (t < tf && it < o->maxiters) {
    ++it;
    dt = std::min(dt, tf - t);
    for (int j = 1; j < 7; ++j) {
      double c[7];
      const double* x[7];
      c[0] = 1.0;
      x[0] = u;
      for (int i = 0; i < j; ++i) {
        c[i + 1] = dt * kA[j][i];
        x[i + 1] = k[i];
      }
      double* dst = j == 6 ? unew : utmp;
      if ((rc = lc(n, dst, j + 1, c, x, s)) || (rc = op.rhs(dst, k[j], t + kC[j] * dt, s))) return rc;
    }
    double EEst;
    {
      double c[7];
      for (int j = 0; j < 7; ++j) c[j] = dt * kBt[j];
      if ((rc = fvm_vec_lincomb_err_norm(n, 7, c, k, u, unew, o->abstol, o->reltol, &EEst, s))) return rc;
    }
    if (!std::isfinite(EEst)) {
      st->retcode = FVM_ODE_UNSTABLE;
      break;
    }
    if (EEst <= 1.0) {
      const double tn = t + dt;
      while (ks < o->n_save && o->save_t[ks] <= tn + 1e-12 * std::max(1.0, std::fabs(tn))) {
        const double th = std::min(1.0, std::max(0.0, (o->save_t[ks] - t) / dt));
        // cubic Hermite through (u, f) at both ends of the step (explicit products:
        // a compiler may turn pow(th, 2) into th * th, which libm's pow need not match)
        const double t2 = th * th, t3 = t2 * th;
        const double h00 = 2 * t3 - 3 * t2 + 1;
        const double h10 = t3 - 2 * t2 + th;
        const double h01 = -2 * t3 + 3 * t2;
        const double h11 = t3 - t2;
        const double c[4] = {h00, h10 * dt, h01, h11 * dt};
        const double* x[4] = {u, k[0], unew, k[6]};
        if ((rc = lc(n, interp, 4, c, x, s)) || (rc = save(o, ks++, interp, s))) return rc;
      }
      std::swap(u, unew);
      t = tn;
      ++st->naccept;
      if ((rc = op.apply_dirichlet(u, t, s))) return rc;
      // first same as last: the last stage is f(u_new, t + dt), and u is u_new unless
      // the Dirichlet update changed it, which a constant update can do only once
      if (o->fsal && (!op.dirichlet || st->naccept >= 2)) std::swap(k[0], k[6]);
      else if ((rc = op.rhs(u, k[0], t, s))) return rc;
      double q = fvm::ctrl_pow(std::max(EEst, 1e-16), 7.0 / 50.0) / fvm::ctrl_pow(qold, 2.0 / 25.0);
      q = std::min(5.0, std::max(0.1, q / 0.9));
      dt = dt / q;
      qold = std::max(EEst, 1e-4);
    } else {
      ++st->nreject;
      const double q = std::min(5.0, fvm::ctrl_pow(std::max(EEst, 1e-16), 7.0 / 50.0) / 0.9);
      dt = dt / q;
    }
  }
  if (st->retcode == FVM_ODE_SUCCESS && !(t >= tf)) st->retcode = FVM_ODE_MAXITERS;
  st->t = t;
  st->nsaved = ks;
  FVM_HIP(hipMemcpyAsync(d_u, u, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, (hipStream_t)s));
  FVM_HIP(hipStreamSynchronize((hipStream_t)s));
  return FVM_OK;
}

int fixed_step(const fvm_ode_opts* o, const Op& op, double* d_u, fvm_ode_stats* st, fvm_stream s) {
  const int64_t n = o->n;
  Work w;
  FVM_HIP(hipMalloc((void**)&w.base, sizeof(double) * (size_t)(n > 0 ? n : 1) * 4));
  double *k1 = w.base, *u1 = w.base + n, *u2 = w.base + 2 * n, *u = w.base + 3 * n;
  FVM_HIP(hipMemcpyAsync(u, d_u, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, (hipStream_t)s));
  const double h = o->dt;
  const int64_t nsteps = (int64_t)std::ceil((o->tf - o->t0) / h - 1e-9);
  double t = o->t0;
  int64_t ks = 0;
  int rc;
  while (ks < o->n_save && o->save_t[ks] <= t + 1e-14)
    if ((rc = save(o, ks++, u, s))) return rc;
  for (int64_t step = 0; step < nsteps; ++step) {
    const double hh = std::min(h, o->tf - t);
    if ((rc = op.rhs(u, k1, t, s))) return rc;
    {
      const double c[2] = {1.0, hh};
      const double* x[2] = {u, k1};
      if ((rc = lc(n, u1, 2, c, x, s))) return rc;
    }
    if (o->method == FVM_ODE_SSPRK3) {
      if ((rc = op.rhs(u1, k1, t + hh, s))) return rc;
      {
        const double c[3] = {0.75, 0.25, 0.25 * hh};
        const double* x[3] = {u, u1, k1};
        if ((rc = lc(n, u2, 3, c, x, s)) || (rc = op.rhs(u2, k1, t + 0.5 * hh, s))) return rc;
      }
      const double c[3] = {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * hh};
      const double* x[3] = {u, u2, k1};
      if ((rc = lc(n, u1, 3, c, x, s))) return rc;
    }
    const double tn = t + hh;
    while (ks < o->n_save && o->save_t[ks] <= tn + 1e-12) {
      const double th = (o->save_t[ks] - t) / hh;
      const double c[2] = {1 - th, th};
      const double* x[2] = {u, u1};
      if ((rc = lc(n, u2, 2, c, x, s)) || (rc = save(o, ks++, u2, s))) return rc;
    }
    std::swap(u, u1);
    t = tn;
    ++st->naccept;
    if ((rc = op.apply_dirichlet(u, t, s))) return rc;
  }
  st->retcode = FVM_ODE_SUCCESS;
  st->t = t;
  st->nsaved = ks;
  FVM_HIP(hipMemcpyAsync(d_u, u, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, (hipStream_t)s));
  FVM_HIP(hipStreamSynchronize((hipStream_t)s));
  return FVM_OK;
}

}  // namespace

extern "C" {

int fvm_arnoldi_step(int32_t op_kind, void* op, int64_t n, const double* d_u, double t, double a, double b,
                     int32_t j, double* const* d_V, double* d_w, double* h, fvm_stream s) {
  FVM_CHECK(op && d_u && d_V && d_w && h, "fvm_arnoldi_step: null argument");
  FVM_CHECK(op_kind >= FVM_OP_PROBLEM && op_kind <= FVM_OP_TEMPLATE, "fvm_arnoldi_step: unknown operator kind");
  FVM_CHECK(j >= 0 && j < 64 && n > 0, "fvm_arnoldi_step: 0 <= j < 64 and n > 0 required");
  for (int i = 0; i <= j; ++i) FVM_CHECK(d_V[i] && d_V[i] != d_w, "fvm_arnoldi_step: null or aliased basis vector");
  int64_t nf = 0;
  Op o{op_kind, op, false, &nf};
  int rc = o.jvp(d_u, d_V[j], d_w, t, s);
  if (rc) return rc;
  if (!(a == 0.0 && b == 1.0)) {  // A = J takes the product as it is
    const double c[2] = {a, b};
    const double* x[2] = {d_V[j], d_w};
    if ((rc = lc(n, d_w, 2, c, x, s))) return rc;
  }
  return fvm::cgs2_norm(n, j + 1, d_V, d_w, h, (hipStream_t)s);
}

int fvm_arnoldi_step_scaled(int32_t op_kind, void* op, int64_t n, const double* d_u, double t, double a,
                            double b, const double* d_scale, double* d_tmp, int32_t j, double* const* d_V,
                            double* d_w, double* h, fvm_stream s) {
  FVM_CHECK(op && d_u && d_scale && d_tmp && d_V && d_w && h, "fvm_arnoldi_step_scaled: null argument");
  FVM_CHECK(op_kind >= FVM_OP_PROBLEM && op_kind <= FVM_OP_TEMPLATE,
            "fvm_arnoldi_step_scaled: unknown operator kind");
  FVM_CHECK(j >= 0 && j < 64 && n > 0, "fvm_arnoldi_step_scaled: 0 <= j < 64 and n > 0 required");
  for (int i = 0; i <= j; ++i)
    FVM_CHECK(d_V[i] && d_V[i] != d_w && d_V[i] != d_tmp, "fvm_arnoldi_step_scaled: null or aliased basis vector");
  FVM_CHECK(d_tmp != d_w, "fvm_arnoldi_step_scaled: d_tmp and d_w must differ");
  // z = scale * V[j] (the right preconditioner), then w = a z + b J z
  int rc = fvm_vec_mul(n, d_tmp, d_scale, d_V[j], s);
  if (rc) return rc;
  int64_t nf = 0;
  Op o{op_kind, op, false, &nf};
  if ((rc = o.jvp(d_u, d_tmp, d_w, t, s))) return rc;
  if (!(a == 0.0 && b == 1.0)) {
    const double c[2] = {a, b};
    const double* x[2] = {d_tmp, d_w};
    if ((rc = lc(n, d_w, 2, c, x, s))) return rc;
  }
  return fvm::cgs2_norm(n, j + 1, d_V, d_w, h, (hipStream_t)s);
}

int fvm_ode_solve(const fvm_ode_opts* o, double* d_u, fvm_ode_stats* st, fvm_stream s) {
  FVM_CHECK(o && d_u && st && o->op, "fvm_ode_solve: null argument");
  FVM_CHECK(o->op_kind >= FVM_OP_PROBLEM && o->op_kind <= FVM_OP_TEMPLATE, "fvm_ode_solve: unknown operator kind");
  FVM_CHECK(o->method >= FVM_ODE_TSIT5 && o->method <= FVM_ODE_SSPRK3, "fvm_ode_solve: unknown method");
  FVM_CHECK(o->n > 0, "fvm_ode_solve: the state must not be empty");
  FVM_CHECK(o->tf >= o->t0, "fvm_ode_solve: tf must not precede t0");
  FVM_CHECK(o->n_save >= 0 && (o->n_save == 0 || (o->save_t && o->d_saves)),
            "fvm_ode_solve: n_save save times need save_t and d_saves");
  for (int64_t j = 1; j < o->n_save; ++j) FVM_CHECK(o->save_t[j] >= o->save_t[j - 1], "fvm_ode_solve: save_t must ascend");
  if (o->method == FVM_ODE_TSIT5) {
    FVM_CHECK(o->abstol > 0 && o->reltol >= 0, "fvm_ode_solve: abstol must be positive, reltol non-negative");
    FVM_CHECK(o->maxiters > 0, "fvm_ode_solve: maxiters must be positive");
  } else {
    FVM_CHECK(o->dt > 0, "fvm_ode_solve: fixed-step methods need dt > 0");
  }
  *st = fvm_ode_stats{};
  Op op{o->op_kind, const_cast<void*>(o->op), o->dirichlet != 0, &st->nf};
  return o->method == FVM_ODE_TSIT5 ? tsit5(o, op, d_u, st, s) : fixed_step(o, op, d_u, st, s);
}

}  // extern "C"
