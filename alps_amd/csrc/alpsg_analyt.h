// The pre-coded f0 that eval_fit continues for species with ACmethod 0
// (ALPS_analyt.f90:77-79), restated from distribution/distribution_analyt.f90:30-96
// as shipped: species 1 and 2 are Maxwellians with beta = 1 and masses 1 and
// 1/1836,
//   f0 = pi^(-3/2) / (m beta)^(3/2) exp(-(ppar^2/(beta m) + pperp^2/(beta m))),
// at real pperp and complex ppar.  Like the reference's file, this is the
// place a user edits to continue another closed-form f0; the library is then
// rebuilt.  distribution_analyt_defined() names the species the function
// covers: alpsg_create refuses ACmethod 0 for any other species, where the
// reference would read an undefined value.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "alpsg_device.h"

namespace {

__host__ __device__ inline bool distribution_analyt_defined(int is) { return is == 1 || is == 2; }

// is: the 1-based species index of the run
__device__ inline cd distribution_analyt(int is, double pperp, cd ppar) {
  const double pi = 3.141592653589793238462643383279502884;
  double beta = 1.0, ms = 1.0;
  if (is == 2) ms = 1.0 / 1836.0;
  const double bm = beta * ms;
  const double norm = pow(pi, -1.5) / pow(bm, 1.5);
  const cd p2 = ppar * ppar;
  const cd arg = mk(-(p2.re / bm + (pperp * pperp) / bm), -(p2.im / bm));
  const cd e = cexpd(arg);
  return mk(norm * e.re, norm * e.im);
}

}  // namespace
