// plaquette.h -- drop-in for the reference's include/plaquette.h:10-34.
#ifndef U1_DROPIN_PLAQUETTE_H
#define U1_DROPIN_PLAQUETTE_H
#include "alloc.h"
#include "array.h"
#include "complex.h"

namespace U1 {

// plaq[0] = <1 - e^{iP}> over space-space plaquettes, plaq[1] over space-time
// (plaquette.cu:156-177).  Writes into the caller's complexd[2].
complexd* Plaquette(Array<double>* dev_lat, complexd* plaq, bool print = false);
complexd* Plaquette(Array<float>* dev_lat, complexd* plaq, bool print = false);
// Host lattice: plaq[0] = <cos P>, plaq[1] = <sin P> (plaquette.cu:32-57).
void plaquette(double* lat, double* plaq);

class PlaqFieldArg {
 public:
  size_t size;
  complexd* plaqfield;  // device, 6 * size components
  complexd* plaq;       // host, the 6 means
  PlaqFieldArg() : size(0), plaqfield(0), plaq(0) {}
  ~PlaqFieldArg() {
    if (plaqfield) dev_free(plaqfield);
    if (plaq) host_free(plaq);
    size = 0;
  }
};

void PlaquetteFields(Array<double>* lat, PlaqFieldArg* plaqfield, bool spacetime, bool evenoddOrder);

}  // namespace U1
#endif
