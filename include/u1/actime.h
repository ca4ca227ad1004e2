// actime.h -- drop-in for the reference's include/actime.h (actime.cpp:19-145):
// integrated autocorrelation time and error of a Monte Carlo series by the
// Gamma method with automatic windowing (U. Wolff, "Monte Carlo errors with
// less errors", Comput. Phys. Commun. 156 (2004) 143, hep-lat/0306017).
// Header-only host code.
#ifndef U1_DROPIN_ACTIME_H
#define U1_DROPIN_ACTIME_H
#include <cmath>
#include <iostream>
#include <vector>

namespace U1 {

struct GammaResult {
  double value;     // mean
  double dvalue;    // error including autocorrelations
  double ddvalue;   // error of the error
  double tauint;    // integrated autocorrelation time (1/2 for uncorrelated data)
  double dtauint;   // its error
  int wopt;         // summation window
  bool window_ok;   // false if the windowing criterion never fired
};

// S is Wolff's tau/tau_int guess (1.5 by default).
inline GammaResult gamma_method(const std::vector<double>& a, double S = 1.5) {
  GammaResult r{0, 0, 0, 0, 0, 0, false};
  const int N = (int)a.size();
  if (N < 2) return r;
  double mean = 0.0;
  for (double v : a) mean += v;
  mean /= N;
  r.value = mean;
  int tmax = N / 2;
  std::vector<double> G;  // Gamma(t)
  double sum_rho = 0.0;   // sum_{t>=1} Gamma(t)/Gamma(0)
  int W = 0;
  bool found = false;
  for (int t = 0; t <= tmax; ++t) {
    double g = 0.0;
    for (int i = 0; i + t < N; ++i) g += (a[i] - mean) * (a[i + t] - mean);
    g /= (double)(N - t);
    G.push_back(g);
    if (t == 0 || found) continue;
    if (G[0] <= 0.0) break;  // constant series
    sum_rho += g / G[0];
    // tau(W) = S / ln((2 tau_int + 1)/(2 tau_int - 1)), 2 tau_int = 1 + 2 sum_rho
    const double tau = sum_rho <= 0.0 ? 1e-300 : S / std::log((sum_rho + 1.0) / sum_rho);
    const double gW = std::exp(-(double)t / tau) - tau / std::sqrt((double)t * N);
    if (gW < 0.0) {
      W = t;
      found = true;
      tmax = std::min(tmax, 2 * t);
    }
  }
  if (!found) W = (int)G.size() - 1;
  r.window_ok = found;
  r.wopt = W;
  if (G[0] <= 0.0) {
    r.tauint = 0.5;
    return r;
  }
  double C = G[0];
  for (int t = 1; t <= W; ++t) C += 2.0 * G[t];
  for (double& g : G) g += C / N;  // bias correction of Gamma
  C = G[0];
  for (int t = 1; t <= W; ++t) C += 2.0 * G[t];
  r.dvalue = std::sqrt(C / N);
  r.ddvalue = r.dvalue * std::sqrt((W + 0.5) / N);
  r.tauint = C / (2.0 * G[0]);
  r.dtauint = r.tauint * 2.0 * std::sqrt((W - r.tauint + 0.5) / N);
  return r;
}

// The reference's entry point: analyse gamma once it holds at least miniter
// values, print the results and set nsweep to the number of sweeps between
// effectively independent measurements, ceil(2 tau_int).
inline void calculateCorTime(int miniter, int iter, std::vector<double>& gamma, int& nsweep) {
  (void)iter;
  if ((int)gamma.size() < miniter) return;
  const GammaResult r = gamma_method(gamma);
  if (!r.window_ok) std::cout << "WARNING: windowing condition failed up to W = " << r.wopt << std::endl;
  std::cout << "value: " << r.value << "\ndvalue: " << r.dvalue << "\nddvalue: " << r.ddvalue
            << "\ntauint: " << r.tauint << "\ndtauint: " << r.dtauint << "\nWopt: " << r.wopt << std::endl;
  nsweep = (int)std::ceil(2.0 * r.tauint);
}
inline void calculateCorTime1(int miniter, int iter, std::vector<double>& gamma, int& nsweep) {
  calculateCorTime(miniter, iter, gamma, nsweep);
}

}  // namespace U1
#endif
