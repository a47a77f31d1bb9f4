// Short-range gamma of the SCC Coulomb interaction, device side (shared by sparse.hip and
// batch.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace dftbhip {

// ------------------------------------------------------------------------------------------
// SCC electrostatics: gamma = 1/R (or Ewald) - short-range exponential gamma (+U on-site)
// ------------------------------------------------------------------------------------------
__device__ inline double gsub(double r, double t1, double t2) {
  return exp(-t1 * r) * ((0.5 * t2 * t2 * t2 * t2 * t1 / ((t1 * t1 - t2 * t2) * (t1 * t1 - t2 * t2))) -
                         (t2 * t2 * t2 * t2 * t2 * t2 - 3.0 * t2 * t2 * t2 * t2 * t1 * t1) /
                             (r * (t1 * t1 - t2 * t2) * (t1 * t1 - t2 * t2) * (t1 * t1 - t2 * t2)));
}

__device__ inline double gsubp(double r, double t1, double t2) {
  const double d2 = t1 * t1 - t2 * t2;
  const double a = 0.5 * t2 * t2 * t2 * t2 * t1 / (d2 * d2);
  const double b = (t2 * t2 * t2 * t2 * t2 * t2 - 3.0 * t2 * t2 * t2 * t2 * t1 * t1) / (d2 * d2 * d2);
  return -t1 * exp(-t1 * r) * (a - b / r) + exp(-t1 * r) * b / (r * r);
}

// expGamma (shortgammafuncs.F90:97-148)
__device__ inline double expgamma(double r, double ua, double ub) {
  const double ta = 3.2 * ua, tb = 3.2 * ub;
  if (r < 1.0e-5) {
    if (fabs(ua - ub) < 1.0e-10) return -0.5 * (ua + ub);
    return -0.5 * ((ta * tb) / (ta + tb) + (ta * tb) * (ta * tb) / ((ta + tb) * (ta + tb) * (ta + tb)));
  }
  if (fabs(ua - ub) < 1.0e-10) {
    const double tm = 0.5 * (ta + tb);
    return exp(-tm * r) * (1.0 / r + 0.6875 * tm + 0.1875 * r * tm * tm +
                           0.02083333333333333333 * r * r * tm * tm * tm);
  }
  return gsub(r, ta, tb) + gsub(r, tb, ta);
}

// expGammaPrime (shortgammafuncs.F90:152-196)
__device__ inline double expgamma_prime(double r, double ua, double ub) {
  if (r < 1.0e-5) return 0.0;
  if (fabs(ua - ub) < 1.0e-10) {
    const double tm = 3.2 * 0.5 * (ua + ub);
    return -tm * exp(-tm * r) *
               (1.0 / r + 0.6875 * tm + 0.1875 * r * tm * tm +
                0.02083333333333333333 * r * r * tm * tm * tm) +
           exp(-tm * r) * (-1.0 / (r * r) + 0.1875 * tm * tm +
                           2.0 * 0.02083333333333333333 * r * tm * tm * tm);
  }
  const double ta = 3.2 * ua, tb = 3.2 * ub;
  return gsubp(r, ta, tb) + gsubp(r, tb, ta);
}


}  // namespace dftbhip
