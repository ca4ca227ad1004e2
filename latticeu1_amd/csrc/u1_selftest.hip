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
    case 8: out[i] = UMath<float>::atan2((float)x[i], (float)y[i]); break;
    case 10: out[i] = UMath<double>::exp_in_loop(x[i]); break;
    default: out[i] = (float)u1m::wrap_near_d((double)(float)x[i]); break;
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

// Streaming references for the roofline: the same bytes as one (parity, mu)
// phase with an ideal access pattern.  Pattern 0 reads 9n floats (as float4),
// pattern 1 copies 4.5n floats (float4), pattern 2 reads eight arrays of n
// floats and writes one (the phase's 8:1 mix, fully coalesced, no stencil).
__global__ __launch_bounds__(256) void k_stream_read(const float4* __restrict__ a, int64_t n4, float* __restrict__ sink) {
  float s = 0.0f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 v = a[i];
    s += (v.x + v.y) + (v.z + v.w);
  }
  if (s == -1.0f) sink[0] = s;  // keeps the loads; the data are zeros, so never taken
}
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, int64_t n4) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
__global__ __launch_bounds__(256) void k_stream_r8w1(const float* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[i + k * n];
  b[i] = s;
}

hipError_t launch_stream(int pattern, float* a, float* b, int64_t n, hipStream_t st) {
  const int grid = 4096;  // grid-stride: 16 workgroups per CU
  if (pattern == 0) {
    k_stream_read<<<grid, 256, 0, st>>>((const float4*)a, 9 * n / 4, b);
  } else if (pattern == 1) {
    k_stream_copy<<<grid, 256, 0, st>>>((const float4*)a, (float4*)b, 9 * n / 8);
  } else if (pattern == 2) {
    k_stream_r8w1<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a, b, n);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_selftest_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int64_t n,
                                  hipStream_t st) {
  k_philox<<<(int)((n + 255) / 256), 256, 0, st>>>(ctr, key, out, n);
  return hipGetLastError();
}

}  // namespace u1d
