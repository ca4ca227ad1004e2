// update.h -- drop-in for the reference's include/update.h:10-12.
#ifndef U1_DROPIN_UPDATE_H
#define U1_DROPIN_UPDATE_H
#include "array.h"
#include "random.h"

namespace U1 {

void HotStart(Array<double>* dev_lat, CudaRNG* rng_state);
void UpdateLattice(Array<double>* dev_lat, CudaRNG* rng_state, int metrop, int ovrn);

// FP32-bulk lattices (binary32 angles, binary64 reductions)
void HotStart(Array<float>* dev_lat, CudaRNG* rng_state);
void UpdateLattice(Array<float>* dev_lat, CudaRNG* rng_state, int metrop, int ovrn);

// Host-array versions of the reference's OpenMP routines (update.cu:51-98),
// computed on the GPU: the host lattice is copied in and out.
void metropolis(double* lat);
void overrelaxation(double* lat);

}  // namespace U1
#endif
