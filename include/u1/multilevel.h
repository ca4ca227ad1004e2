// multilevel.h -- drop-in for the reference's include/multilevel.h:9-25.
#ifndef U1_DROPIN_MULTILEVEL_H
#define U1_DROPIN_MULTILEVEL_H
#include "array.h"
#include "complex.h"
#include "random.h"

namespace U1 {

Array<complexd>* MultiLevel(Array<double>* dev_lat, CudaRNG* rng_state, int n4, int k4, int n2, int k2, int metrop,
                            int ovrn);
Array<complexd>* MultiLevel(Array<float>* dev_lat, CudaRNG* rng_state, int n4, int k4, int n2, int k2, int metrop,
                            int ovrn);

class MultiLevelRes {
 public:
  Array<complexd>* poly;     // host, R = 1..Nx/2
  Array<complexd>* ppSpace;  // device, SV * (Nx/2) * 3
  MultiLevelRes() : poly(0), ppSpace(0) {}
  ~MultiLevelRes() {
    if (poly) delete poly;
    if (ppSpace) delete ppSpace;
  }
};

MultiLevelRes* MultiLevelField(Array<double>* dev_lat, CudaRNG* rng_state, int n4, int k4, int n2, int k2, int metrop,
                               int ovrn);

}  // namespace U1
#endif
