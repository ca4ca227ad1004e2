// random.h -- drop-in for CudaRNG (include/random.h:31-58).  The reference keeps
// one cuRAND MRG32k3a state per half-lattice site; here randomness is
// counter-based Philox4x32-10, so the "state" is just the key (seed) and the
// counters of the chain: a draw is a pure function of (seed, link, sweep,
// stream, hit), nothing is read or written per site, and Backup/Restore save
// and restore the counters.  Unlike the reference (random.cu:34 hard-codes
// 1234), the seed argument is honoured.
#ifndef U1_DROPIN_RANDOM_H
#define U1_DROPIN_RANDOM_H
#include <cstddef>
#include <cstdint>

namespace U1 {

struct PhiloxState {
  uint64_t seed;
  uint64_t sweep;     // UpdateLattice calls so far (Philox counter word 1)
  uint64_t mcalls;    // measurement calls that draw random numbers (multihit, multilevel)
};
typedef PhiloxState cuRNGState;

class CudaRNG {
 public:
  CudaRNG() : st{0, 0, 0}, bk{0, 0, 0}, size(0) {}
  CudaRNG(int seed, size_t size_) : st{(uint64_t)(uint32_t)seed, 0, 0}, bk{0, 0, 0}, size(size_) {}
  cuRNGState* getPtr() { return &st; }
  size_t Size() const { return size; }
  void Backup() { bk = st; }
  void Restore() { st = bk; }
  ~CudaRNG() {}

 private:
  PhiloxState st, bk;
  size_t size;
};

}  // namespace U1
#endif
