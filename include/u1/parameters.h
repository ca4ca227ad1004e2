// parameters.h -- drop-in for the reference's include/parameters.h: global
// geometry set by SetupLatticeParameters (parameters.cu:146-208) and its host
// accessors (parameters.h:82-146).  There is no __constant__ copy: every call
// into the MI355X library passes the geometry explicitly, built from PARAMS.
#ifndef U1_DROPIN_PARAMETERS_H
#define U1_DROPIN_PARAMETERS_H
#include <cstddef>
#include <sstream>
#include <string>

#include "enum.h"
#include "runtime.h"

namespace U1 {

void SetupLatticeParameters(int Nx, int Ny, int Nz, int Nt, int dirs, double beta, double aniso, int imetrop,
                            int ovrn);

// Metropolis proposal used by UpdateLattice and MultiLevel (default: the
// reference's independent angle, one hit).  setProposal(PROPOSAL_DELTA, 4, 0.5)
// gives theta' = theta + U[-0.5, 0.5) with 4 hits.
void setProposal(ProposalKind kind, int nhit, double step);
// Multihit flavour of Poly2(multihit=true) and MultiLevel (default analytic,
// lattice_functions.h:92-105); the stochastic flavour averages nhit hits.
void setMultiHit(MultiHitKind kind, int nhit = 100, double step = 0.5);

TuneMode getTuning();
Verbosity getVerbosity();
void setTuning(TuneMode kerneltunein);
void setVerbosity(Verbosity verbosein);

std::string GetLatticeName();
std::string GetLatticeNameI();

// Device properties kept by Start (the reference stores a cudaDeviceProp).
struct DeviceProp {
  char name[256];
  int multiProcessorCount;
};

namespace PARAMS {
extern double Beta;
extern double Aniso;
extern int DIRS;
extern int TDir;
extern int volume;
extern int half_volume;
extern int spatial_volume;
extern int Grid[4];
extern int Offset[4];
extern int iter;
extern double accept_ratio;
extern int ovrn;
extern int metrop;
extern DeviceProp deviceProp;
extern int proposal;
extern int nhit;
extern double step;
extern int mhit_kind;
extern int mhit_nhit;
extern double mhit_step;
}  // namespace PARAMS

inline int Volume() { return PARAMS::volume; }
inline int HalfVolume() { return PARAMS::half_volume; }
inline int SpatialVolume() { return PARAMS::spatial_volume; }
inline int Grid(int dim) { return PARAMS::Grid[dim]; }
inline int Offset(int dim) { return PARAMS::Offset[dim]; }
inline double Beta() { return PARAMS::Beta; }
inline int Dirs() { return PARAMS::DIRS; }
inline int TDir() { return PARAMS::TDir; }
inline double Aniso() { return PARAMS::Aniso; }

// Geometry of the current PARAMS in the C ABI's form.
u1_geom CurrentGeom();
u1_metro CurrentMetro();

template <class T>
inline std::string ToString(const T number) {
  std::stringstream ss;
  ss << number;
  return ss.str();
}

}  // namespace U1
#endif
