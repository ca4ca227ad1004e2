// u1_multilevel.hip -- tensor accumulation of the two-level multilevel
// Polyakov-loop correlator (reference multilevel.cu:331-553).
//
// Layout of one T2 / T4 layer (multilevel.cu:350): element (x, R, dir) at
//     x + SV*(R-1) + SV*Rmax*dir,  layer stride sz = SV*Rmax*3,
// complex FP64, interleaved (re, im).  All accumulation is FP64.
// The masked sublattice updates reuse the update phase kernels (ml = 4 or 2),
// and the temporal-link field reuses k_temporal_links.
#include "u1_kernels.h"

namespace u1d {

namespace {

struct C64 {
  double re, im;
};
__device__ __forceinline__ C64 cmul(C64 a, C64 b) { return {a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im}; }
__device__ __forceinline__ C64 ldc(const double* p, int64_t i) { return {p[2 * i], p[2 * i + 1]}; }

// multilevel.cu:331-360: l2[x,R,dir,layer] += prod_{t in layer} P(x,t) *
// prod_{t in layer} conj P(x + R e_dir, t); one thread per (x, R, dir).
__global__ __launch_bounds__(256) void k_ml_l2(const double* __restrict__ tl, double* __restrict__ l2, RepDesc r,
                                               int rmax) {
  const int64_t sz = (int64_t)r.SV * rmax * 3;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= sz) return;
  const int32_t id = (int32_t)(idx % r.SV);
  const int32_t R = (int32_t)((idx / r.SV) % rmax) + 1;
  const int32_t dir = (int32_t)(idx / ((int64_t)r.SV * rmax));
  int32_t x[3] = {id % r.N[0], (id / r.N[0]) % r.N[1], id / (r.N[0] * r.N[1])};
  int32_t y[3] = {x[0], x[1], x[2]};
  y[dir] = (y[dir] + R) % r.N[dir];
  const int64_t s0 = x[0] + (int64_t)r.N[0] * (x[1] + (int64_t)r.N[1] * x[2]);
  const int64_t s1 = y[0] + (int64_t)r.N[0] * (y[1] + (int64_t)r.N[1] * y[2]);
  int layer = 0;
  for (int t = 0; t < r.N[3]; t += 2) {
    C64 p0 = {1.0, 0.0}, p1 = {1.0, 0.0};
    for (int tt = t; tt < t + 2; ++tt) {
      p0 = cmul(p0, ldc(tl, s0 + (int64_t)tt * r.SV));
      const C64 q = ldc(tl, s1 + (int64_t)tt * r.SV);
      p1 = cmul(p1, C64{q.re, -q.im});
    }
    const C64 pl = cmul(p0, p1);
    const int64_t pos = idx + sz * layer;
    l2[2 * pos] = pl.re + l2[2 * pos];
    l2[2 * pos + 1] = pl.im + l2[2 * pos + 1];
    ++layer;
  }
}

// multilevel.cu:439-456: l4[.,j] += (l2[.,2j] * l2norm) (l2[.,2j+1] * l2norm).
__global__ __launch_bounds__(256) void k_ml_l4(const double* __restrict__ l2, double* __restrict__ l4, RepDesc r,
                                               int rmax, double l2norm) {
  const int64_t sz = (int64_t)r.SV * rmax * 3;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= sz) return;
  const int nl2 = r.N[3] / 2;
  int j4 = 0;
  for (int a = 0; a < nl2; a += 2) {
    C64 pl = {1.0, 0.0};
    for (int layer = a; layer < a + 2; ++layer) {
      const C64 v = ldc(l2, idx + sz * layer);
      pl = cmul(pl, C64{v.re * l2norm, v.im * l2norm});
    }
    const int64_t pos = idx + sz * j4;
    l4[2 * pos] = pl.re + l4[2 * pos];
    l4[2 * pos + 1] = pl.im + l4[2 * pos + 1];
    ++j4;
  }
}

// multilevel.cu:532-553: C(R) = sum_x sum_dir prod_j (l4[.,j] * l4norm);
// optional per-site field ppSpace; partial rows of 2*rmax per workgroup.
__global__ __launch_bounds__(256) void k_ml_final(const double* __restrict__ l4, double* __restrict__ ppspace,
                                                  RepDesc r, int rmax, double l4norm,
                                                  double* __restrict__ partials) {
  const int32_t id = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  const int64_t sz = (int64_t)r.SV * rmax * 3;
  const int nl4 = r.N[3] / 4;
  for (int R = 0; R < rmax; ++R) {
    double v[2] = {0.0, 0.0};
    if (id < r.SV) {
      for (int dir = 0; dir < 3; ++dir) {
        C64 pl = {1.0, 0.0};
        const int64_t base = id + (int64_t)r.SV * R + (int64_t)r.SV * rmax * dir;
        for (int j = 0; j < nl4; ++j) {
          const C64 q = ldc(l4, base + sz * j);
          pl = cmul(pl, C64{q.re * l4norm, q.im * l4norm});
        }
        v[0] += pl.re;
        v[1] += pl.im;
        if (ppspace) {
          ppspace[2 * base] = pl.re;
          ppspace[2 * base + 1] = pl.im;
        }
      }
    }
    double o[2];
    block_sum<2>(v, o);
    if (threadIdx.x == 0) {
      partials[((int64_t)blockIdx.x * rmax + R) * 2] = o[0];
      partials[((int64_t)blockIdx.x * rmax + R) * 2 + 1] = o[1];
    }
  }
}

}  // namespace

hipError_t launch_ml_l2(const double* tl, double* l2, const RepDesc& r, int rmax, int block, hipStream_t st) {
  const int64_t sz = (int64_t)r.SV * rmax * 3;
  k_ml_l2<<<(int)((sz + block - 1) / block), block, 0, st>>>(tl, l2, r, rmax);
  return hipGetLastError();
}

hipError_t launch_ml_l4(const double* l2, double* l4, const RepDesc& r, int rmax, double l2norm, int block,
                        hipStream_t st) {
  const int64_t sz = (int64_t)r.SV * rmax * 3;
  k_ml_l4<<<(int)((sz + block - 1) / block), block, 0, st>>>(l2, l4, r, rmax, l2norm);
  return hipGetLastError();
}

hipError_t launch_ml_final(const double* l4, double* ppspace, const RepDesc& r, int rmax, double l4norm,
                           double* partials, int nblk, int block, hipStream_t st) {
  k_ml_final<<<nblk, block, 0, st>>>(l4, ppspace, r, rmax, l4norm, partials);
  return hipGetLastError();
}

}  // namespace u1d
