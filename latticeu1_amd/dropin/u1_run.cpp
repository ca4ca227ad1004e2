// u1_run -- the reference's driver loop (u1.cpp:37-176) on the drop-in layer.
//
// Start, SetupLatticeParameters, a zeroed (cold) or hot-started lattice and a
// CudaRNG seeded 1234, the plaquette and Polyakov loop of the start, then
// maxIter sweeps of UpdateLattice(metrop, ovrn).  Every printiter sweeps the
// plaquette and Polyakov loop go to plaq_<name>.dat and poly_<name>.dat
// (name = GetLatticeName(), precision 12, the reference's line formats:
//   iter \t Re P_ss \t Im P_ss \t Re P_st \t Im P_st
//   iter \t Re L \t Im L \t |L|: |L|
// ), and from iteration 1000 on, every printiter sweeps, Poly2 (plain and
// multihit) and MultiLevel(10, 1, 20, 1, 2, 5) write their Pot_, Pot_mhit_
// and Pot_mlevel_ files.  Finalize(0) saves the tune cache.
//
// usage: u1_run [Nx Ny Nz Nt beta aniso metrop ovrn maxIter printiter hotstart [fp32]]
// defaults: the reference's 24 24 24 12 1.0 1.0 1 3 1000 100 0.  With "fp32"
// the angles are binary32 (FP64 reductions); the files are the same.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "array.h"
#include "complex.h"
#include "multilevel.h"
#include "parameters.h"
#include "plaquette.h"
#include "polyakov.h"
#include "random.h"
#include "timer.h"
#include "update.h"

using namespace U1;

namespace {

std::ofstream open_file(const std::string& name) {
  std::ofstream f(name, std::ios::out);
  if (!f.is_open()) {
    std::cout << "Cannot create file: " << name << std::endl;
    std::exit(1);
  }
  std::cout << "Creating file: " << name << std::endl;
  f.precision(12);
  return f;
}

template <typename Real>
int run(int maxIter, int printiter, bool hotstart) {
  Array<Real>* lattice = new Array<Real>(Device, Volume() * Dirs());  // zero-filled: cold start
  CudaRNG* rng = new CudaRNG(1234, HalfVolume());
  if (hotstart) HotStart(lattice, rng);

  const std::string filename = GetLatticeName() + ".dat";
  std::ofstream fileout = open_file("plaq_" + filename);
  std::ofstream fileout1 = open_file("poly_" + filename);

  complexd* plaqv = new complexd[2];
  auto measure = [&]() {
    Plaquette(lattice, plaqv, false);
    const complexd L = Polyakov(lattice);
    std::cout << PARAMS::iter << '\t' << plaqv[0] << '\t' << plaqv[1] << std::endl;
    std::cout << "L: " << L.real() << '\t' << L.imag() << "\t|L|: " << L.abs() << std::endl;
    fileout << PARAMS::iter << '\t' << plaqv[0] << '\t' << plaqv[1] << std::endl;
    fileout1 << PARAMS::iter << '\t' << L.real() << '\t' << L.imag() << "\t|L|: " << L.abs() << std::endl;
  };
  PARAMS::iter = 0;
  std::cout << "Iter: " << PARAMS::iter << " \t";
  measure();
  for (PARAMS::iter = 1; PARAMS::iter <= maxIter; ++PARAMS::iter) {
    UpdateLattice(lattice, rng, PARAMS::metrop, PARAMS::ovrn);
    if (PARAMS::iter % printiter == 0) measure();
    if (PARAMS::iter >= 1000 && PARAMS::iter % printiter == 0) {
      Timer p0, p1, p2;
      p0.start();
      delete Poly2(lattice, false);
      p0.stop();
      std::cout << "p0: " << p0.getElapsedTime() << " s" << std::endl;
      p1.start();
      delete Poly2(lattice, true);
      p1.stop();
      std::cout << "p1: " << p1.getElapsedTime() << " s" << std::endl;
      p2.start();
      delete MultiLevel(lattice, rng, 10, 1, 20, 1, 2, 5);
      p2.stop();
      std::cout << "p2: " << p2.getElapsedTime() << " s" << std::endl;
    }
  }
  fileout.close();
  fileout1.close();
  delete lattice;
  delete rng;
  delete[] plaqv;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  int Nx = 24, Ny = 24, Nz = 24, Nt = 12, metrop = 1, ovrn = 3, maxIter = 1000, printiter = 100, hot = 0;
  double beta = 1.0, aniso = 1.0;
  if (argc > 1 && argc < 12) {
    std::cout << "usage: u1_run [Nx Ny Nz Nt beta aniso metrop ovrn maxIter printiter hotstart [fp32]]" << std::endl;
    return 1;
  }
  if (argc >= 12) {
    Nx = std::atoi(argv[1]);
    Ny = std::atoi(argv[2]);
    Nz = std::atoi(argv[3]);
    Nt = std::atoi(argv[4]);
    beta = std::atof(argv[5]);
    aniso = std::atof(argv[6]);
    metrop = std::atoi(argv[7]);
    ovrn = std::atoi(argv[8]);
    maxIter = std::atoi(argv[9]);
    printiter = std::atoi(argv[10]);
    hot = std::atoi(argv[11]);
  }
  const bool fp32 = argc >= 13 && std::string(argv[12]) == "fp32";
  if (printiter < 1) {
    std::cout << "printiter must be >= 1" << std::endl;
    return 1;
  }
  Timer t0;
  t0.start();
  Start(0, VERBOSE, TUNE_YES);
  SetupLatticeParameters(Nx, Ny, Nz, Nt, 4, beta, aniso, metrop, ovrn);
  const int rc = fp32 ? run<float>(maxIter, printiter, hot != 0) : run<double>(maxIter, printiter, hot != 0);
  t0.stop();
  std::cout << "Time: " << t0.getElapsedTime() << " s" << std::endl;
  Finalize(0);  // saves the tuned launch shapes
  return rc;
}
