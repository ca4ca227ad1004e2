// CPU check of the Gamma-method (include/u1/actime.h) on AR(1) series
// x_{i+1} = rho x_i + sqrt(1 - rho^2) e_i, whose exact integrated
// autocorrelation time is (1 + rho) / (2 (1 - rho)).  Prints one line per rho.
#include <cstdio>
#include <random>

#include "actime.h"

int main() {
  const double rhos[] = {0.0, 0.5, 0.8, 0.9};
  for (double rho : rhos) {
    std::mt19937_64 gen(12345);
    std::normal_distribution<double> n01(0.0, 1.0);
    std::vector<double> x(200000);
    double v = 0.0;
    for (auto& e : x) {
      v = rho * v + std::sqrt(1.0 - rho * rho) * n01(gen);
      e = v + 3.0;
    }
    const U1::GammaResult r = U1::gamma_method(x);
    const double exact = (1.0 + rho) / (2.0 * (1.0 - rho));
    // error of the mean of N unit-variance samples: sqrt(2 tau_int / N)
    std::printf("%g %.6f %.6f %.6f %.6e %.6e %d %d\n", rho, exact, r.tauint, r.dtauint, r.dvalue,
                std::sqrt(2.0 * exact / x.size()), r.wopt, r.window_ok ? 1 : 0);
  }
  return 0;
}
