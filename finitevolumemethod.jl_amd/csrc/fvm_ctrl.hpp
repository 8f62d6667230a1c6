// x^p for the integrators' step-size controllers, the same on the host and the
// device: a power built from +, -, *, /, frexp, ldexp and floor only (all exact or
// correctly rounded in IEEE double), so the host-driven loops and the device loops
// that run whole integrations (and the package's own host loop, which restates it)
// choose the same step sizes bit for bit.  The math libraries' pow differ between
// host and device in the last place, and step sizes feed back into every later step.
// Accuracy ~1e-15 relative (a controller needs far less).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include <cmath>
#endif

namespace fvm {

// log(x), x > 0 finite: x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m - 1) / (m + 1),
// log m = 2 (s + s^3 / 3 + ... + s^25 / 25) (|s| <= 0.172: the next term is below 1e-21)
__host__ __device__ inline double ctrl_log(double x) {
  int e = 0;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 25.0;
  for (int k = 23; k >= 1; k -= 2) p = p * s2 + 1.0 / (double)k;
  return 2.0 * s * p + (double)e * 0.69314718055994531;
}

// exp(y), |y| < 700: y = k ln 2 + r with |r| <= ln 2 / 2, exp r by its Taylor
// polynomial to r^20 / 20! (Horner), then times 2^k
__host__ __device__ inline double ctrl_exp(double y) {
  const double k = floor(y / 0.69314718055994531 + 0.5);
  const double r = y - k * 0.69314718055994531;
  double p = 1.0;
  for (int i = 20; i >= 1; --i) p = 1.0 + p * r / (double)i;
  return ldexp(p, (int)k);
}

__host__ __device__ inline double ctrl_pow(double x, double p) { return ctrl_exp(p * ctrl_log(x)); }

}  // namespace fvm
