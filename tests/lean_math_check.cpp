// Host driver of the lean complex power (alps_amd/csrc/alpsg_lean.h) for
// tests/test_lean_math.py: reads "x y c" per line and prints z^c, exp(c),
// (E, log m) of x (x > 0), atan2(y, x), sin(c) and cos(c).
#include <cstdio>

#include "../alps_amd/csrc/alpsg_lean.h"

int main() {
  double x, y, c;
  while (std::scanf("%lf %lf %lf", &x, &y, &c) == 3) {
    double re, im, s, co;
    alpsg_lean::cpow_(x, y, c, re, im);
    int E = 0;
    const double lm = alpsg_lean::log_m(x > 0.0 ? x : 1.0, E);
    const double at = (x == 0.0 && y == 0.0) ? 0.0 : alpsg_lean::atan2_(y, x);
    alpsg_lean::sincos_(c, s, co);
    std::printf("%.17g %.17g %.17g %d %.17g %.17g %.17g %.17g\n", re, im, alpsg_lean::exp_(c), E,
                lm, at, s, co);
  }
  return 0;
}
