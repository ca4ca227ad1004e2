// alloc.h -- drop-in for the allocation helpers of the reference's
// include/alloc.h:60-67.  dev_malloc zero-fills (alloc.cpp:193).
#ifndef U1_DROPIN_ALLOC_H
#define U1_DROPIN_ALLOC_H
#include <cstddef>

namespace U1 {
void* dev_malloc(size_t size);
void dev_free(void* ptr);
void* safe_malloc(size_t size);  // zero-filled host memory
void host_free(void* ptr);
void printPeakMemUsage();
}  // namespace U1
#endif
