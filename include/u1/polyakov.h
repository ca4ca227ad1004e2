// polyakov.h -- drop-in for the reference's include/polyakov.h:10-16.
#ifndef U1_DROPIN_POLYAKOV_H
#define U1_DROPIN_POLYAKOV_H
#include "array.h"
#include "complex.h"

namespace U1 {

void polyakov(double* lat, double* poly);
complexd Polyakov(Array<double>* dev_lat, bool print = false);
complexd Polyakov(Array<float>* dev_lat, bool print = false);
// <P(0) P*(R)>, R = 1..Nx/2, as a host Array owned by the caller; writes
// Pot_[mhit_]<name>.dat as the reference (polyakov.cu:618-654).
Array<complexd>* Poly2(Array<double>* dev_lat, bool multihit);
Array<complexd>* Poly2(Array<float>* dev_lat, bool multihit);

}  // namespace U1
#endif
