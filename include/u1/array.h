// array.h -- drop-in for the reference's Array<Real> (include/array.h:19-182):
// an owning buffer on the host or the device, zero-filled on allocation, with
// Copy / Clear / Backup / Restore and element access (host arrays only).
#ifndef U1_DROPIN_ARRAY_H
#define U1_DROPIN_ARRAY_H
#include <cstring>
#include <iostream>

#include "alloc.h"
#include "enum.h"
#include "parameters.h"
#include "runtime.h"

namespace U1 {

template <class Real>
class Array {
 public:
  Array() : ptr(0), ptr_backup(0), location(Host), size(0) {}
  explicit Array(ReadMode location_) : ptr(0), ptr_backup(0), location(location_), size(0) {}
  Array(ReadMode location_, size_t size_) : ptr(0), ptr_backup(0), location(location_), size(size_) {
    ptr = Allocate(size);
  }
  ~Array() { Release(); }
  Array(const Array&) = delete;
  Array& operator=(const Array&) = delete;

  Real* getPtr() { return ptr; }
  const Real* getPtr() const { return ptr; }
  size_t Size() const { return size; }
  ReadMode Location() const { return location; }

  Real& operator()(const int i) { return ptr[i]; }
  Real operator()(const int i) const { return ptr[i]; }
  Real& operator[](const int i) { return ptr[i]; }
  Real operator[](const int i) const { return ptr[i]; }
  Real& at(const int i) { return ptr[i]; }
  Real at(const int i) const { return ptr[i]; }

  // array.h:60-82: (re)allocate to the source size, then copy across spaces.
  void Copy(Array* in) {
    if (ptr && size != in->Size()) {
      Free(ptr);
      ptr = 0;
    }
    if (!ptr) {
      size = in->Size();
      ptr = Allocate(size);
    }
    u1SafeCall(u1_memcpy(ptr, in->getPtr(), in->Size() * sizeof(Real), GetMemcpyKind(in->Location(), location)));
  }
  void Clear() {
    if (location == Host)
      std::memset(ptr, 0, size * sizeof(Real));
    else
      u1SafeCall(u1_zero(ptr, (size * sizeof(Real) + 3) / 4 * 4));
  }
  void Backup() {
    if (!ptr_backup) ptr_backup = Allocate(size);
    u1SafeCall(u1_memcpy(ptr_backup, ptr, size * sizeof(Real), GetMemcpyKind(location, location)));
  }
  void Restore() {
    if (!ptr_backup) return;
    u1SafeCall(u1_memcpy(ptr, ptr_backup, size * sizeof(Real), GetMemcpyKind(location, location)));
    Free(ptr_backup);
    ptr_backup = 0;
  }

  friend std::ostream& operator<<(std::ostream& out, Array<Real>* M) {
    for (size_t i = 0; i < M->size; i++) out << i << '\t' << M->ptr[i] << std::endl;
    return out;
  }
  friend std::ostream& operator<<(std::ostream& out, Array<Real>& M) { return out << &M; }

 private:
  Real* ptr;
  Real* ptr_backup;
  ReadMode location;
  size_t size;

  Real* Allocate(size_t n) {
    return (Real*)(location == Host ? safe_malloc(n * sizeof(Real)) : dev_malloc(n * sizeof(Real)));
  }
  void Free(Real* p) {
    if (location == Host)
      host_free(p);
    else
      dev_free(p);
  }
  void Release() {
    if (ptr) Free(ptr);
    if (ptr_backup) Free(ptr_backup);
    ptr = ptr_backup = 0;
  }
};

// EO <-> lexicographic copy of a device lattice (convertOrder.cu:136-145);
// returns a new device Array owned by the caller.
Array<double>* LatticeConvert(Array<double>* lat, bool eo_to_no);
Array<float>* LatticeConvert(Array<float>* lat, bool eo_to_no);

}  // namespace U1
#endif
