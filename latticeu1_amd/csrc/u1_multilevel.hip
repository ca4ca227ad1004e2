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

// The two-slice products of the temporal-link field, once per level-2 step:
// p2[l SV + x] = prod_{t = 2l, 2l+1} P(x, t) and q2[l SV + x] = the same
// product of the conjugates, each from 1 in the reference's operation order
// (multilevel.cu:342-347), so k_ml_l2 needs two loads per layer instead of
// four and no longer recomputes them for each of its 3 Rmax (R, dir).
__global__ __launch_bounds__(256) void k_ml_pair2(const double* __restrict__ tl, double* __restrict__ p2,
                                                  double* __restrict__ q2, RepDesc r) {
  // 32-bit indices: the host checks SV * Rmax * 3 * Nt/2 < 2^31
  const uint32_t n = (uint32_t)r.SV * (uint32_t)(r.N[3] / 2);
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const uint32_t l = idx / (uint32_t)r.SV, x = idx - l * (uint32_t)r.SV;
  C64 p0 = {1.0, 0.0}, p1 = {1.0, 0.0};
  for (uint32_t tt = 0; tt < 2; ++tt) {
    const C64 q = ldc(tl, x + (2 * l + tt) * (uint32_t)r.SV);
    p0 = cmul(p0, q);
    p1 = cmul(p1, C64{q.re, -q.im});
  }
  p2[2 * idx] = p0.re;
  p2[2 * idx + 1] = p0.im;
  q2[2 * idx] = p1.re;
  q2[2 * idx + 1] = p1.im;
}

// multilevel.cu:331-360: l2[x,R,dir,layer] += p2(x, layer) * q2(x + R e_dir,
// layer); one thread per (x, R, dir).
__global__ __launch_bounds__(256) void k_ml_l2(const double* __restrict__ p2, const double* __restrict__ q2,
                                               double* __restrict__ l2, RepDesc r, int rmax) {
  // 32-bit indices: the host checks SV * Rmax * 3 * Nt/2 < 2^31
  const uint32_t SV = (uint32_t)r.SV, sz = SV * (uint32_t)rmax * 3u;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= sz) return;
  const uint32_t q = idx / SV, id = idx - q * SV;
  const uint32_t dir = q / (uint32_t)rmax, R = q - dir * (uint32_t)rmax + 1u;
  const uint32_t n0 = (uint32_t)r.N[0], n1 = (uint32_t)r.N[1];
  const uint32_t x1x2 = id / n0, x0 = id - x1x2 * n0, x2 = x1x2 / n1, x1 = x1x2 - x2 * n1;
  uint32_t y[3] = {x0, x1, x2};
  y[dir] = (y[dir] + R) % (uint32_t)r.N[dir];
  const uint32_t s0 = id, s1 = y[0] + n0 * (y[1] + n1 * y[2]);
  const uint32_t nl2 = (uint32_t)r.N[3] / 2u;
  for (uint32_t layer = 0; layer < nl2; ++layer) {
    const C64 pl = cmul(ldc(p2, s0 + layer * SV), ldc(q2, s1 + layer * SV));
    const uint32_t pos = idx + sz * layer;
    l2[2 * pos] = pl.re + l2[2 * pos];
    l2[2 * pos + 1] = pl.im + l2[2 * pos + 1];
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

hipError_t launch_ml_l2(const double* tl, double* pq, double* l2, const RepDesc& r, int rmax, int block,
                        hipStream_t st) {
  const int64_t n = (int64_t)r.SV * (r.N[3] / 2);
  double* p2 = pq;
  double* q2 = pq + 2 * n;
  k_ml_pair2<<<(int)((n + block - 1) / block), block, 0, st>>>(tl, p2, q2, r);
  const int64_t sz = (int64_t)r.SV * rmax * 3;
  k_ml_l2<<<(int)((sz + block - 1) / block), block, 0, st>>>(p2, q2, l2, r, rmax);
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
