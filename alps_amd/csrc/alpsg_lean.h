// Complex power z^c (z complex, c real) for the kappa continuation
// (1 + u2 (p - u3)^2 + pc u5 pperp^2)^u4 of eval_fit (ALPS_analyt.f90:231-235),
// written so that an inlined evaluation stays small in registers: each
// elementary function is a reduction to a short interval and a truncated
// Taylor (or atanh) series with constant coefficients.  The methods:
//
//  exp(u):   u = k ln2 + r, |r| <= ln2/2 (Cody-Waite, ln2 in two parts, the
//            first with 29 significant bits so k ln2_hi is exact), then the
//            Taylor series of e^r to r^13/13! (remainder below 5e-18), scaled
//            by 2^k;
//  log(q):   q = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//            s = (m - 1)/(m + 1), |s| <= 0.1716, series to s^21/21
//            (remainder below 1e-18);
//  atan2:    t = min(|x|,|y|)/max(|x|,|y|) in [0, 1], atan t = atan(c) +
//            atan((t - c)/(1 + t c)) with c = k/8 the nearest eighth (table of
//            atan(k/8)), |argument| <= 1/16, series to ^15/15; octant and
//            quadrant from the signs and the order of |x|, |y|;
//  sin, cos: a = k pi/2 + r, |r| <= pi/4 (pi/2 in three parts, the first two
//            with 33 significant bits, exact for |k| < 2^20), Taylor series to
//            r^17/17! and r^18/18!, quadrant from k mod 4.
//
// Domain: the callers keep |c| <= ALPSG_LEAN_MAXC (so |c arg z| <= 3.2e4, far
// inside the exact range of the sin/cos reduction); z = 0 gives 0 as cpowr
// does.  The constants were computed for this file to 60 digits (ln 2, pi/2,
// atan(k/8)); the series coefficients are 1/n! and 1/(2k+1).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define ALPSG_HD __host__ __device__
#else
#define ALPSG_HD
#endif

#define ALPSG_LEAN_MAXC 1.0e4

namespace alpsg_lean {

// ln 2 = LN2_HI + LN2_LO (LN2_HI: 29 significant bits)
constexpr double LN2_HI = 0x1.62e42ffp-1;
constexpr double LN2_LO = -4.2009150726810846e-11;
constexpr double LOG2E = 1.4426950408889634;
// pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (PIO2_1, PIO2_2: 33 significant bits)
constexpr double PIO2_1 = 0x1.921fb544p+0;
constexpr double PIO2_2 = 0x1.0b4611a6p-34;
constexpr double PIO2_3 = 2.0222662487959506e-21;
constexpr double TWO_OVER_PI = 0.6366197723675814;
// pi and pi/2 as a double plus the rest
constexpr double PI_HI = 3.141592653589793, PI_LO = 1.2246467991473532e-16;
constexpr double PIO2_HI = 1.5707963267948966, PIO2_LO = 6.123233995736766e-17;

// e^u
ALPSG_HD inline double exp_(double u) {
  double k = rint(u * LOG2E);
  k = fmin(fmax(k, -2200.0), 2200.0);  // beyond this e^u is 0 or inf anyway
  const double r = fma(-k, LN2_LO, fma(-k, LN2_HI, u));
  double p = 1.0 / 6227020800.0;        // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);     // 1/12!
  p = fma(p, r, 1.0 / 39916800.0);      // 1/11!
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

// log m for q = 2^E m, m in [sqrt(1/2), sqrt(2)) (q > 0 and normal); E is
// returned, so that log q = E ln 2 + log m
ALPSG_HD inline double log_m(double q, int &E) {
  int e;
  double m = frexp(q, &e);  // m in [0.5, 1)
  if (m < 0.70710678118654752) {
    m = 2.0 * m;
    e -= 1;
  }
  E = e;
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  return fma(2.0 * s * s2, p, 2.0 * s);  // 2 s (1 + s2 p)
}

// atan2(y, x) for (x, y) != (0, 0)
ALPSG_HD inline double atan2_(double y, double x) {
  const double a = fabs(y), b = fabs(x);
  const double mx = fmax(a, b), mn = fmin(a, b);
  const double t = mn / mx;
  const double kf = rint(8.0 * t);
  const double c = 0.125 * kf;
  const double u = (t - c) / fma(t, c, 1.0);
  const double u2 = u * u;
  double p = 1.0 / 15.0;
  p = fma(p, u2, -1.0 / 13.0);
  p = fma(p, u2, 1.0 / 11.0);
  p = fma(p, u2, -1.0 / 9.0);
  p = fma(p, u2, 1.0 / 7.0);
  p = fma(p, u2, -1.0 / 5.0);
  p = fma(p, u2, 1.0 / 3.0);
  const double at = fma(-u * u2, p, u);  // atan u = u - u^3/3 + ...
  const int k = (int)kf;
  double ac = 0.0;  // atan(k/8)
  ac = (k == 1) ? 0.12435499454676144 : ac;
  ac = (k == 2) ? 0.24497866312686414 : ac;
  ac = (k == 3) ? 0.35877067027057225 : ac;
  ac = (k == 4) ? 0.4636476090008061 : ac;
  ac = (k == 5) ? 0.5585993153435624 : ac;
  ac = (k == 6) ? 0.6435011087932844 : ac;
  ac = (k == 7) ? 0.7188299996216245 : ac;
  ac = (k == 8) ? 0.7853981633974483 : ac;
  double th = ac + at;
  if (a > b) th = (PIO2_HI - th) + PIO2_LO;
  if (x < 0.0) th = (PI_HI - th) + PI_LO;
  return copysign(th, y);
}

// sin and cos of a, |a| < 2^20 pi/2
ALPSG_HD inline void sincos_(double a, double &s, double &c) {
  const double k = rint(a * TWO_OVER_PI);
  double r = fma(-k, PIO2_1, a);
  r = fma(-k, PIO2_2, r);
  r = fma(-k, PIO2_3, r);
  const double r2 = r * r;
  double ps = 1.0 / 355687428096000.0;  // 1/17!
  ps = fma(ps, r2, -1.0 / 1307674368000.0);
  ps = fma(ps, r2, 1.0 / 6227020800.0);
  ps = fma(ps, r2, -1.0 / 39916800.0);
  ps = fma(ps, r2, 1.0 / 362880.0);
  ps = fma(ps, r2, -1.0 / 5040.0);
  ps = fma(ps, r2, 1.0 / 120.0);
  ps = fma(ps, r2, -1.0 / 6.0);
  const double sn = fma(r * r2, ps, r);
  double pc = 1.0 / 6402373705728000.0;  // 1/18!
  pc = fma(pc, r2, -1.0 / 20922789888000.0);
  pc = fma(pc, r2, 1.0 / 87178291200.0);
  pc = fma(pc, r2, -1.0 / 479001600.0);
  pc = fma(pc, r2, 1.0 / 3628800.0);
  pc = fma(pc, r2, -1.0 / 40320.0);
  pc = fma(pc, r2, 1.0 / 720.0);
  pc = fma(pc, r2, -1.0 / 24.0);
  pc = fma(pc, r2, 0.5);
  const double cs = fma(-r2, pc, 1.0);
  const int q = ((int)k) & 3;
  const double ss = (q & 1) ? cs : sn, cc = (q & 1) ? sn : cs;
  s = (q & 2) ? -ss : ss;
  c = ((q + 1) & 2) ? -cc : cc;
}

// z^c = exp(c log|z|) (cos(c arg z), sin(c arg z)); 0 for z = 0
ALPSG_HD inline void cpow_(double x, double y, double c, double &re, double &im) {
  if (x == 0.0 && y == 0.0) {
    re = im = 0.0;
    return;
  }
  int e;
  (void)frexp(fmax(fabs(x), fabs(y)), &e);
  const double xs = ldexp(x, -e), ys = ldexp(y, -e);  // exact; max(|xs|, |ys|) in [0.5, 1)
  int E;
  const double lm = log_m(fma(xs, xs, ys * ys), E);
  // log|z| = (2 e + E) ln2 / 2 + log(m) / 2
  const double n = 2.0 * e + E;
  const double lr = 0.5 * (fma(n, LN2_HI, fma(n, LN2_LO, lm)));
  const double mag = exp_(c * lr);
  double sn, cs;
  sincos_(c * atan2_(y, x), sn, cs);
  re = mag * cs;
  im = mag * sn;
}

}  // namespace alpsg_lean
