// The update kernels form the Metropolis proposal offset a (2u - 1), u = w 2^-32,
// as fma(w, a 2^-31, -a); the oracle forms it in two steps.  Both are the
// correctly rounded value of the same real number, so they must agree bit
// for bit: checked here for every step the tests use, pi, and random steps in
// (0, pi], over edge words and 2^22 random words each.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "u1_math.h"

static double two_step(uint32_t w, double a) {
  const double u = (double)w * 0x1p-32;
  return a * (2.0 * u - 1.0);
}
static double one_fma(uint32_t w, double a) { return std::fma((double)w, a * 0x1p-31, -a); }

int main() {
  std::vector<double> steps = {0.5, 0.25, 1.0, 0.1, M_PI, 3.141592653589793};
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> us(1e-6, M_PI);
  for (int i = 0; i < 64; ++i) steps.push_back(us(rng));
  const uint32_t edge[] = {0u, 1u, 2u, 0x7fffffffu, 0x80000000u, 0x80000001u, 0xfffffffeu, 0xffffffffu};
  long bad = 0, n = 0;
  for (double a : steps) {
    auto check = [&](uint32_t w) {
      const double x = two_step(w, a), y = one_fma(w, a);
      ++n;
      if (std::memcmp(&x, &y, sizeof x) != 0) {
        if (bad < 5) std::printf("mismatch a=%.17g w=%u: %.17g vs %.17g\n", a, w, x, y);
        ++bad;
      }
    };
    for (uint32_t w : edge) check(w);
    for (int i = 0; i < (1 << 22); ++i) check((uint32_t)rng());
  }
  std::printf("%ld of %ld differ\n", bad, n);
  if (bad) return 1;
  std::printf("proposal ok\n");

  // wrap3_d (binary32 overrelaxation) against the stepwise wrap_d on
  // [-3 pi - 1e-6, 3 pi + 1e-6]: range [-pi, pi), within 1 ulp of binary64
  long wbad = 0, wn = 0;
  auto wcheck = [&](double x) {
    const double a = u1m::wrap_d(x), b = u1m::wrap3_d(x);
    ++wn;
    const bool in = b >= -u1m::K_PI && b < u1m::K_PI;
    const double ulp = std::nextafter(std::fabs(a), INFINITY) - std::fabs(a);
    if (!in || std::fabs(a - b) > ulp) {
      if (wbad < 5) std::printf("wrap mismatch x=%.17g: %.17g vs %.17g\n", x, a, b);
      ++wbad;
    }
  };
  const double lim = 3.0 * u1m::K_PI + 1e-6;
  std::uniform_real_distribution<double> ux(-lim, lim);
  for (int i = 0; i < (1 << 22); ++i) wcheck(ux(rng));
  for (int m = -3; m <= 3; ++m)  // odd multiples of pi and their neighbours
    for (double base : {m * u1m::K_PI, m * u1m::K_PI + 1e-6, m * u1m::K_PI - 1e-6}) {
      double x = base;
      for (int s = 0; s < 64; ++s) x = std::nextafter(x, INFINITY);
      double y = base;
      for (int s = 0; s < 128; ++s) {
        wcheck(y);
        y = std::nextafter(y, INFINITY);
      }
      (void)x;
    }
  std::printf("wrap: %ld of %ld out of range or beyond 1 ulp\n", wbad, wn);
  if (wbad) return 1;
  std::printf("wrap ok\n");
  return 0;
}
