// enum.h -- drop-in for the reference's include/enum.h:13-31 (Verbosity,
// TuneMode, ReadMode) on the MI355X build.  GetMemcpyKind returns the copy kind
// of the C ABI (U1_H2H ... U1_D2D) instead of a CUDA enum.
#ifndef U1_DROPIN_ENUM_H
#define U1_DROPIN_ENUM_H
#include <string>

#include "../latticeu1_amd.h"

namespace U1 {

typedef enum Verbosity_s { SILENT, SUMMARIZE, VERBOSE, DEBUG_VERBOSE } Verbosity;
typedef enum TuneMode_S { TUNE_NO, TUNE_YES } TuneMode;
typedef enum ReadMode_S { Host, Device } ReadMode;

// Metropolis proposal of UpdateLattice / MultiLevel (not in the reference,
// which always draws an independent angle with one hit, update.cu:181-182).
typedef enum ProposalKind_s { PROPOSAL_INDEPENDENT = U1_PROPOSAL_INDEPENDENT, PROPOSAL_DELTA = U1_PROPOSAL_DELTA } ProposalKind;

// Multihit flavour used by Poly2(multihit=true) and MultiLevel.
typedef enum MultiHitKind_s { MULTIHIT_ANALYTIC = U1_POLY_MULTIHIT, MULTIHIT_STOCHASTIC = U1_POLY_MULTIHIT_STOCHASTIC } MultiHitKind;

inline int GetMemcpyKind(const ReadMode in, const ReadMode out) {
  if (in == Host) return out == Host ? U1_H2H : U1_H2D;
  return out == Host ? U1_D2H : U1_D2D;
}

inline std::string GetMemcpyKindString(int kind) {
  switch (kind) {
    case U1_H2H: return "HostToHost";
    case U1_H2D: return "HostToDevice";
    case U1_D2H: return "DeviceToHost";
    default: return "DeviceToDevice";
  }
}

}  // namespace U1
#endif
