// checkpoint.h -- configuration save / resume (no counterpart in the
// reference, which persists only its tune cache).  The file holds the angles in
// lexicographic order plus the generator's seed and sweep counter, so a
// counter-based chain resumes exactly where it stopped.
#ifndef U1_DROPIN_CHECKPOINT_H
#define U1_DROPIN_CHECKPOINT_H
#include <string>

#include "array.h"
#include "random.h"

namespace U1 {

void SaveConfiguration(Array<double>* dev_lat, CudaRNG* rng_state, const std::string& path);
void LoadConfiguration(Array<double>* dev_lat, CudaRNG* rng_state, const std::string& path);
void SaveConfiguration(Array<float>* dev_lat, CudaRNG* rng_state, const std::string& path);
void LoadConfiguration(Array<float>* dev_lat, CudaRNG* rng_state, const std::string& path);

}  // namespace U1
#endif
