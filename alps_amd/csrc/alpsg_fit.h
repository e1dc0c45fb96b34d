// Device evaluation of the analytic continuation of f0 (eval_fit,
// ALPS_analyt.f90:32-363), shared by the resonant kernels of the
// non-relativistic and relativistic paths.
#pragma once

#include "alpsg_analyt.h"
#include "alpsg_device.h"
#include "alpsg_lean.h"

namespace {

// eval_fit (ALPS_analyt.f90:32-363) at (iperp, complex p).
__device__ __noinline__ cd eval_fit(const SpecDev &S, int iperp, cd p) {
  if (S.acmethod == 0)  // the pre-coded f0 (ALPS_analyt.f90:77-79)
    return distribution_analyt(S.anum, S.pperp[iperp], p);
  if (S.acmethod == 2) {
    // fit_function_poly, Chebyshev kind 1 (ALPS_analyt.f90:307-358)
    cd t = mk((p.re - S.poly_norm1) / S.poly_norm2, p.im / S.poly_norm2);
    if (cabsd(t) > 1.0) return mk(0.0, 0.0);
    const double *c = S.polyc + (size_t)iperp * (S.poly_order + 1);
    cd bm2 = mk(1.0, 0.0), bm1 = t, f = c[0] * bm2;
    f = f + c[1] * bm1;
    for (int n = 2; n <= S.poly_order; ++n) {
      cd b = (2.0 * t) * bm1 - bm2;
      f = f + c[n] * b;
      bm2 = bm1;
      bm1 = b;
    }
    if (S.poly_logfit) {
      const double lm = S.poly_log_max;
      if (f.re < -lm || f.im < -lm || f.re > lm || f.im > lm) return mk(0.0, 0.0);
      const double ln10 = 2.302585092994045684017991454684364208;
      return cexpd(mk(f.re * ln10, f.im * ln10));
    }
    return f;
  }
  cd f = mk(0.0, 0.0);
  for (int fi = 0; fi < S.nfits; ++fi) {
    const double *c = S.fitc + ((size_t)iperp * S.nfits + fi) * FIT_STRIDE;
    const cd dq = mk(p.re - c[2], p.im);
    const cd sq = dq * dq;
    switch (S.fit_type[fi]) {
      case 1:  // Maxwell: c0 = u1*exp(-pc*pperp^2), c1 = u2, c2 = u3
        f = f + c[0] * cexpd(mk(-c[1] * sq.re, -c[1] * sq.im));
        break;
      case 2: {  // kappa: c3 = u4, c4 = pc*u5*pperp^2
        const cd kp = mk(1.0 + c[1] * sq.re + c[4], c[1] * sq.im);
        f = f + c[0] * cpowr(kp, c[3]);
      } break;
      case 3: {  // Juettner in (pperp, ppar): c3 = pperp^2, c4 = vA^2/m^2 (ALPS_analyt.f90:238-242)
        const cd sq2 = mk(1.0 + (c[3] + sq.re) * c[4], sq.im * c[4]);
        const cd r = csqrtd(sq2);
        f = f + c[0] * cexpd(mk(-c[1] * r.re, -c[1] * r.im));
      } break;
      case 6: {  // bi-Moyal: c3 = u4*pc*pperp^2
        const cd arg = mk(c[3] + c[1] * sq.re, c[1] * sq.im);
        const cd e = cexpd(arg);
        f = f + c[0] * cexpd(mk(0.5 * (arg.re - e.re), 0.5 * (arg.im - e.im)));
      } break;
      default:
        break;
    }
  }
  return f;
}

// eval_fit of a species whose every fit is in the Maxwell form (case 1 above,
// the same operations), inlined: the relativistic species, whose fits over
// Gamma the host stores in that form, take it in k_rel_res without a call
__device__ __forceinline__ cd eval_fit_maxwell(const SpecDev &S, int iperp, cd p) {
  cd f = mk(0.0, 0.0);
#pragma unroll 1
  for (int fi = 0; fi < S.nfits; ++fi) {
    const double *c = S.fitc + ((size_t)iperp * S.nfits + fi) * FIT_STRIDE;
    const cd dq = mk(p.re - c[2], p.im);
    const cd sq = dq * dq;
    f = f + c[0] * cexpd(mk(-c[1] * sq.re, -c[1] * sq.im));
  }
  return f;
}

// eval_fit of a species whose every fit is a kappa fit (case 2 above), inlined
// into the Landau pass: the same kappa part, raised to u4 by the register-lean
// complex power of alpsg_lean.h (the library functions of cpowr, inlined,
// hold that kernel to two wavefronts per SIMD, a call of eval_fit to one).
// The host gives a species this form only when every |u4| <= ALPSG_LEAN_MAXC.
// The fits are summed one at a time (unroll 1), so that their temporaries are
// not live together.
__device__ __forceinline__ cd eval_fit_kappa(const SpecDev &S, int iperp, cd p) {
  cd f = mk(0.0, 0.0);
#pragma unroll 1
  for (int fi = 0; fi < S.nfits; ++fi) {
    const double *c = S.fitc + ((size_t)iperp * S.nfits + fi) * FIT_STRIDE;
    const cd dq = mk(p.re - c[2], p.im);
    const cd sq = dq * dq;
    const cd kp = mk(1.0 + c[1] * sq.re + c[4], c[1] * sq.im);
    double re, im;
    alpsg_lean::cpow_(kp.re, kp.im, c[3], re, im);
    f = f + c[0] * mk(re, im);
  }
  return f;
}

}  // namespace
