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

  // wrap_near_d (the binary32 update kernels' wrap, in binary64) against the
  // stepwise wrap_d on [-3 pi - 1, 3 pi + 1] (the overrelaxation angle
  // reaches 3 pi, theta + delta 3 pi between two wraps): the same angle within
  // 4 binary64 ulps of pi, in [-pi, pi] up to that; and, rounded to binary32
  // as the kernels store it, within one binary32 ulp of pi.
  long wbad = 0, wn = 0;
  const double ulp_pi_d = std::nextafter(u1m::K_PI, INFINITY) - u1m::K_PI;
  const double ulp_pi_f = (double)std::nextafter((float)u1m::K_PI, INFINITY) - (double)(float)u1m::K_PI;
  auto ang = [](double d) { return d - u1m::K_TWO_PI * std::rint(d / u1m::K_TWO_PI); };
  auto wcheck = [&](double x) {
    const double a = u1m::wrap_d(x), b = u1m::wrap_near_d(x);
    const double bf = (double)(float)u1m::wrap_near_d((double)(float)x), af = u1m::wrap_d((double)(float)x);
    ++wn;
    if (std::fabs(b) > u1m::K_PI + 4 * ulp_pi_d || std::fabs(ang(b - a)) > 4 * ulp_pi_d ||
        std::fabs(ang(bf - af)) > ulp_pi_f) {
      if (wbad < 5) std::printf("wrap mismatch x=%.17g: %.17g vs %.17g (binary32 %.9g)\n", x, a, b, bf);
      ++wbad;
    }
  };
  const double lim = 3.0 * u1m::K_PI + 1.0;
  std::uniform_real_distribution<double> ux(-lim, lim);
  for (int i = 0; i < (1 << 22); ++i) wcheck(ux(rng));
  for (int m = -3; m <= 3; m += 2)  // odd multiples of pi: 256 binary64 neighbours each side
    for (double dir : {-INFINITY, INFINITY}) {
      double y = m * u1m::K_PI;
      for (int s = 0; s < 256; ++s) {
        wcheck(y);
        y = std::nextafter(y, dir);
      }
    }
  std::printf("wrap: %ld of %ld out of range or off the wrap_d angle\n", wbad, wn);
  if (wbad) return 1;
  std::printf("wrap ok\n");
  return 0;
}
