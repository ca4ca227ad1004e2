// u1_selftest.hip -- device evaluation of the shared elementary functions and
// of Philox4x32-10, so tests can check bit identity with the host oracle.
#include "u1_kernels.h"

namespace u1d {

namespace {

__global__ __launch_bounds__(256) void k_math(int fn, const double* __restrict__ x, const double* __restrict__ y,
                                              double* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (fn) {
    case 0: {
      double s, c;
      u1m::sincos_d(x[i], &s, &c);
      out[2 * i] = s;
      out[2 * i + 1] = c;
      break;
    }
    case 1: out[i] = u1m::exp_d(x[i]); break;
    case 2: out[i] = u1m::atan2_d(x[i], y[i]); break;
    case 3: {
      float s, c;
      u1m::sincos_f((float)x[i], &s, &c);
      out[2 * i] = s;
      out[2 * i + 1] = c;
      break;
    }
    case 4: out[i] = u1m::atan2_f((float)x[i], (float)y[i]); break;
    case 5: {
      double a, b;
      u1m::bessel_i0i1_d(x[i], &a, &b);
      out[2 * i] = a;
      out[2 * i + 1] = b;
      break;
    }
    case 6: out[i] = u1m::wrap_d(x[i]); break;
    case 7: {
      float s, c;
      UMath<float>::sincos((float)x[i], &s, &c);
      out[2 * i] = s;
      out[2 * i + 1] = c;
      break;
    }
    default: out[i] = UMath<float>::atan2((float)x[i], (float)y[i]); break;
  }
}

__global__ __launch_bounds__(256) void k_philox(const uint32_t* __restrict__ ctr, const uint32_t* __restrict__ key,
                                                uint32_t* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u1r::U32x4 c;
  for (int j = 0; j < 4; ++j) c.v[j] = ctr[4 * i + j];
  const u1r::U32x4 o = u1r::philox4x32_10(c, key[2 * i], key[2 * i + 1]);
  for (int j = 0; j < 4; ++j) out[4 * i + j] = o.v[j];
}

}  // namespace

hipError_t launch_selftest_math(int fn, const double* x, const double* y, double* out, int64_t n, hipStream_t st) {
  k_math<<<(int)((n + 255) / 256), 256, 0, st>>>(fn, x, y, out, n);
  return hipGetLastError();
}

hipError_t launch_selftest_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int64_t n,
                                  hipStream_t st) {
  k_philox<<<(int)((n + 255) / 256), 256, 0, st>>>(ctr, key, out, n);
  return hipGetLastError();
}

}  // namespace u1d
