// A small driver written against the drop-in headers exactly as the
// reference's u1.cpp uses them (Start, SetupLatticeParameters, Array, CudaRNG,
// HotStart, UpdateLattice, Plaquette, Polyakov, Poly2, MultiLevel, Finalize),
// plus LatticeConvert, PlaquetteFields and the FP32 overloads.  Compiled with
// g++ against include/u1 and libu1.so only.  Prints "KEY value ..." lines and
// writes the final FP64 lattice to lattice.bin for the test to check.
#include <cstdio>
#include <fstream>
#include <iostream>

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

int main(int argc, char** argv) {
  const int nsweeps = argc > 1 ? std::atoi(argv[1]) : 20;
  Start(0, SUMMARIZE, TUNE_YES);
  SetupLatticeParameters(8, 8, 8, 8, 4, 1.0, 1.0, 1, 3);
  setProposal(PROPOSAL_DELTA, 4, 0.5);
  Array<double>* lattice = new Array<double>(Device, Volume() * Dirs());
  CudaRNG* rng = new CudaRNG(1234, HalfVolume());
  HotStart(lattice, rng);
  complexd* plaqv = new complexd[2];
  for (PARAMS::iter = 1; PARAMS::iter <= nsweeps; ++PARAMS::iter) UpdateLattice(lattice, rng, PARAMS::metrop, PARAMS::ovrn);
  Plaquette(lattice, plaqv, false);
  std::cout.precision(17);
  std::cout << "PLAQ " << plaqv[0].real() << " " << plaqv[0].imag() << " " << plaqv[1].real() << " " << plaqv[1].imag()
            << std::endl;
  complexd L = Polyakov(lattice);
  std::cout << "POLY " << L.real() << " " << L.imag() << std::endl;
  PARAMS::iter = nsweeps;
  Array<complexd>* p2 = Poly2(lattice, false);
  std::cout << "POLY2 " << p2->at(0).real() << " " << p2->at(0).imag() << std::endl;
  delete p2;
  Array<complexd>* p2m = Poly2(lattice, true);
  std::cout << "POLY2MH " << p2m->at(0).real() << " " << p2m->at(0).imag() << std::endl;
  delete p2m;
  Array<complexd>* ml = MultiLevel(lattice, rng, 2, 1, 2, 1, 1, 3);
  std::cout << "ML " << ml->at(0).real() << " " << ml->at(0).imag() << std::endl;
  delete ml;
  Array<double>* no = LatticeConvert(lattice, true);
  Array<double>* eo = LatticeConvert(no, false);
  Array<double> h1(Host), h2(Host);
  h1.Copy(lattice);
  h2.Copy(eo);
  bool same = true;
  for (size_t i = 0; i < h1.Size(); ++i) same = same && (h1[i] == h2[i]);
  std::cout << "CONVERT_ROUNDTRIP " << (same ? 1 : 0) << std::endl;
  PlaqFieldArg pf;
  PlaquetteFields(lattice, &pf, true, true);
  std::cout << "FIELD0 " << pf.plaq[0].real() << " " << pf.plaq[0].imag() << std::endl;
  double hp[2];
  plaquette(h1.getPtr(), hp);
  std::cout << "HOSTPLAQ " << hp[0] << " " << hp[1] << std::endl;
  Array<float>* lat32 = new Array<float>(Device, Volume() * Dirs());
  CudaRNG rng32(1234, HalfVolume());
  HotStart(lat32, &rng32);
  for (int s = 0; s < 5; ++s) UpdateLattice(lat32, &rng32, 1, 3);
  Plaquette(lat32, plaqv, false);
  std::cout << "PLAQ32 " << plaqv[0].real() << " " << plaqv[1].real() << std::endl;
  std::ofstream out("lattice.bin", std::ios::binary);
  out.write((const char*)h1.getPtr(), h1.Size() * sizeof(double));
  out.close();
  // host-array passes (update.cu:51-98): Metropolis, overrelaxation, twice,
  // on a host copy of the final lattice
  PARAMS::accept_ratio = 0.0;
  for (int s = 0; s < 2; ++s) {
    metropolis(h1.getPtr());
    overrelaxation(h1.getPtr());
  }
  std::cout << "HOSTACCEPT " << PARAMS::accept_ratio << std::endl;
  std::ofstream out2("host_passes.bin", std::ios::binary);
  out2.write((const char*)h1.getPtr(), h1.Size() * sizeof(double));
  out2.close();
  delete lat32;
  delete no;
  delete eo;
  delete lattice;
  delete rng;
  delete[] plaqv;
  Finalize(0);
  return 0;
}
