// runtime.h -- drop-in for Start / Finalize and the fatal-error macros of the
// reference's include/cuda_error_check.h:45-60.  Failures of the C ABI print
// the library's error text and end in Finalize(1), as the reference's
// cudaSafeCall / errorU1 do (cuda_error_check.cpp:80-101).
#ifndef U1_DROPIN_RUNTIME_H
#define U1_DROPIN_RUNTIME_H
#include <cstdio>
#include <cstdlib>

#include "enum.h"

namespace U1 {

void Start(int gpuid, Verbosity verb = VERBOSE, TuneMode tune = TUNE_YES);
[[noreturn]] void Finalize_(const char* func, const char* file, int line, int error);
#define Finalize(_error_) U1::Finalize_(__func__, __FILE__, __LINE__, _error_)

// Check a C ABI status; on failure print the error and Finalize(1).
void u1Check_(const char* func, const char* file, int line, int status);
#define u1SafeCall(call) U1::u1Check_(__func__, __FILE__, __LINE__, (call))
void devSync();

#define printfU1(abc, ...) printf(abc, ##__VA_ARGS__);
#define errorU1(abc, ...)                                                  \
  do {                                                                     \
    printf("Error in %s: %d in %s()\n\t", __FILE__, __LINE__, __func__);   \
    printf(abc, ##__VA_ARGS__);                                            \
    Finalize(1);                                                           \
  } while (0)

}  // namespace U1
#endif
