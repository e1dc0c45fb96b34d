// secant_osc (ALPS_fns.f90:1919-2101) as a state machine: want() gives the
// next frequency whose disp() the solve needs, feed() takes that value.  The
// batched solvers (alpsg_secant_batch, the server) serve the requests of many
// machines with one device launch per step.
#pragma once

#include <cmath>
#include <complex>

#include "../../include/alpsg.h"

namespace alpsg_host {

typedef std::complex<double> cplx;

struct OscMachine {
  // phases: 0 = need D(om) initial, 1 = need D(prevom) initial, 2 = need D(om) iteration,
  //         3 = need D(prevom+1e-8), 4 = need D(om*(1+delta)), 5 = need D(om*(1-delta))
  cplx om, prevom, prev2om, prev3om, prev4om, prevD, D, minom, minD, jump, dplus;
  double damping = 1.0;
  int iter = 0, osc = 0, phase = 0, done = 0;
  int numiter = 50;
  double D_threshold = 0.0, D_prec = 0.0;

  static cplx delta() { return cplx(1.e-6, 1.e-8); }

  void start(cplx w, const alpsg_root_params &rp) {
    om = w;
    phase = 0;
    done = 0;
    iter = 0;
    osc = 0;
    damping = 1.0;
    numiter = rp.numiter;
    D_threshold = rp.D_threshold;
    D_prec = rp.D_prec;
  }

  cplx want() const {
    switch (phase) {
      case 0: return om;
      case 1: return prevom;
      case 2: return om;
      case 3: return prevom;
      case 4: return om * (1.0 + delta());
      default: return om * (1.0 - delta());
    }
  }

  cplx result() const { return iter >= numiter ? minom : om; }

  void feed(cplx Dv) {
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
        if (std::abs(prevD) < std::abs(minD)) {
          minom = prevom;
          minD = prevD;
        }
        iter = 0;
        start_iter();
        break;
      case 2:
        D = Dv;
        if (std::abs(D - prevD) < 1.e-80) {
          prevom = prevom + 1.e-8;
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
        cplx Dprime = (dplus - Dv) / (2.0 * om * delta());
        jump = D / (Dprime + lambda * D);
        finish();
        start_iter();
      } break;
    }
  }

 private:
  void start_iter() {
    if (iter <= numiter - 1) {
      iter = iter + 1;
      phase = 2;
    } else {
      done = 1;
    }
  }
  // the convergence test and the rest of the iteration body once D (and the
  // optional prevD refresh) are known; may request the Newton evaluations
  void decide() {
    const double osc_threshold = (double)1.e-3f;
    if (std::abs(D) < D_threshold) {
      jump = 0.0;
      done = 1;
      return;
    }
    if (iter > 4) {
      auto close = [&](cplx a) {
        return (std::fabs(om.real() - a.real()) < (std::fabs(om.real()) * osc_threshold)) &&
               (std::fabs(om.imag() - a.imag()) < (std::fabs(om.imag()) * osc_threshold));
      };
      if (close(prevom) || close(prev2om) || close(prev3om) || close(prev4om)) {
        osc = osc + 1;
        damping = std::min(0.5, damping * 0.75);
      }
    }
    if (osc > 1) {
      phase = 4;
      return;
    }
    jump = damping * D * (om - prevom) / (D - prevD);
    phase = 6;
  }
  void finish() {
    const double cap = (double)0.1f;
    if (std::abs(jump) > cap * std::abs(om)) jump = (cap * std::abs(om)) * (jump / std::abs(jump));
    if (std::abs(D) > std::abs(prevD)) jump = 0.5 * jump;
    prev4om = prev3om;
    prev3om = prev2om;
    prev2om = prevom;
    prevom = om;
    prevD = D;
    if (std::abs(D) < std::abs(minD)) {
      minom = om;
      minD = D;
    }
    om = om - jump;
    phase = 2;
  }
};

}  // namespace alpsg_host
