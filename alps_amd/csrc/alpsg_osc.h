// secant_osc (ALPS_fns.f90:1919-2101) as a state machine: want() gives the
// next frequency whose disp() the solve needs, feed() takes that value.  The
// batched solvers (alpsg_secant_batch, the server) serve the requests of many
// machines with one device launch per step; the same code also runs on the
// device (k_osc_round), where the machines of a batch step between the disp
// launches of a captured graph.
//
// Complex arithmetic is written out (products without fused multiply-adds,
// quotients by Smith's algorithm, moduli from one quotient and a square root)
// with operations that round the same way on the host and on the device, so
// a solve takes the same steps wherever its state machine runs.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/alpsg.h"

namespace alpsg_osc {

struct zc {
  double re, im;
};

__host__ __device__ inline zc zmk(double r, double i) { return zc{r, i}; }
__host__ __device__ inline zc operator+(zc a, zc b) {
#pragma clang fp contract(off)
  return zc{a.re + b.re, a.im + b.im};
}
__host__ __device__ inline zc operator-(zc a, zc b) {
#pragma clang fp contract(off)
  return zc{a.re - b.re, a.im - b.im};
}
__host__ __device__ inline zc operator*(zc a, zc b) {
#pragma clang fp contract(off)
  return zc{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline zc operator*(double s, zc a) {
#pragma clang fp contract(off)
  return zc{s * a.re, s * a.im};
}
__host__ __device__ inline zc operator*(zc a, double s) {
#pragma clang fp contract(off)
  return zc{a.re * s, a.im * s};
}
// a / b by Smith's algorithm
__host__ __device__ inline zc operator/(zc a, zc b) {
#pragma clang fp contract(off)
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return zc{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return zc{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__host__ __device__ inline zc operator/(zc a, double s) { return zc{a.re / s, a.im / s}; }
// |a| = x sqrt(1 + (y/x)^2), x = max(|re|, |im|), y = min: correctly rounded
// division and square root, no fused operations
__host__ __device__ inline double zabs(zc a) {
#pragma clang fp contract(off)
  double x = fabs(a.re), y = fabs(a.im);
  if (x < y) {
    const double t = x;
    x = y;
    y = t;
  }
  if (x == 0.0 || x > 1.7976931348623157e308) return x + y;
  const double r = y / x;
  return x * sqrt(1.0 + r * r);
}

struct OscMachine {
  // phases: 0 = need D(om) initial, 1 = need D(prevom) initial, 2 = need D(om) iteration,
  //         3 = need D(prevom+1e-8), 4 = need D(om*(1+delta)), 5 = need D(om*(1-delta))
  zc om, prevom, prev2om, prev3om, prev4om, prevD, D, minom, minD, jump, dplus;
  double damping;
  int iter, osc, phase, done;
  int numiter;
  int done_round;  // the round (disp launch) that finished this solve
  double D_threshold, D_prec;

  __host__ __device__ static zc delta() { return zc{1.e-6, 1.e-8}; }

  __host__ __device__ void start(zc w, int numiter_, double D_threshold_, double D_prec_) {
    om = w;
    prevom = prev2om = prev3om = prev4om = prevD = D = minom = minD = jump = dplus = zc{0.0, 0.0};
    phase = 0;
    done = 0;
    iter = 0;
    osc = 0;
    damping = 1.0;
    done_round = -1;
    numiter = numiter_;
    D_threshold = D_threshold_;
    D_prec = D_prec_;
  }

  __host__ __device__ zc want() const {
#pragma clang fp contract(off)
    const zc one{1.0, 0.0};
    switch (phase) {
      case 0: return om;
      case 1: return prevom;
      case 2: return om;
      case 3: return prevom;
      case 4: return om * (one + delta());
      default: return om * (one - delta());
    }
  }

  __host__ __device__ zc result() const { return iter >= numiter ? minom : om; }

  __host__ __device__ void feed(zc Dv) {
#pragma clang fp contract(off)
    switch (phase) {
      case 0:
        D = Dv;
        minom = om;
        minD = Dv;
        prevom = om * (1.0 - D_prec);
        prev2om = prev3om = prev4om = om;
        phase = 1;
        break;
      case 1:
        prevD = Dv;
        if (zabs(prevD) < zabs(minD)) {
          minom = prevom;
          minD = prevD;
        }
        iter = 0;
        start_iter();
        break;
      case 2:
        D = Dv;
        if (zabs(D - prevD) < 1.e-80) {
          prevom = prevom + zc{1.e-8, 0.0};
          phase = 3;
          break;
        }
        decide();
        if (phase == 6) {
          finish();
          start_iter();
        }
        break;
      case 3:
        prevD = Dv;
        decide();
        if (phase == 6) {
          finish();
          start_iter();
        }
        break;
      case 4:
        dplus = Dv;
        phase = 5;
        break;
      case 5: {
        const double lambda = (double)0.1f;
        const zc Dprime = (dplus - Dv) / ((2.0 * om) * delta());
        jump = D / (Dprime + lambda * D);
        finish();
        start_iter();
      } break;
    }
  }

 private:
  __host__ __device__ void start_iter() {
    if (iter <= numiter - 1) {
      iter = iter + 1;
      phase = 2;
    } else {
      done = 1;
    }
  }
  // the convergence test and the rest of the iteration body once D (and the
  // optional prevD refresh) are known; may request the Newton evaluations
  __host__ __device__ void decide() {
#pragma clang fp contract(off)
    const double osc_threshold = (double)1.e-3f;
    if (zabs(D) < D_threshold) {
      jump = zc{0.0, 0.0};
      done = 1;
      return;
    }
    if (iter > 4) {
      const zc cand[4] = {prevom, prev2om, prev3om, prev4om};
      bool any = false;
      for (int k = 0; k < 4; ++k) {
        const zc a = cand[k];
        any = any || ((fabs(om.re - a.re) < (fabs(om.re) * osc_threshold)) &&
                      (fabs(om.im - a.im) < (fabs(om.im) * osc_threshold)));
      }
      if (any) {
        osc = osc + 1;
        damping = fmin(0.5, damping * 0.75);
      }
    }
    if (osc > 1) {
      phase = 4;
      return;
    }
    jump = ((damping * D) * (om - prevom)) / (D - prevD);
    phase = 6;
  }
  __host__ __device__ void finish() {
#pragma clang fp contract(off)
    const double cap = (double)0.1f;
    if (zabs(jump) > cap * zabs(om)) jump = (cap * zabs(om)) * (jump / zabs(jump));
    if (zabs(D) > zabs(prevD)) jump = 0.5 * jump;
    prev4om = prev3om;
    prev3om = prev2om;
    prev2om = prevom;
    prevom = om;
    prevD = D;
    if (zabs(D) < zabs(minD)) {
      minom = om;
      minD = D;
    }
    om = om - jump;
    phase = 2;
  }
};

}  // namespace alpsg_osc
