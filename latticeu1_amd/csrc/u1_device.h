// u1_device.h -- device-side building blocks of the gfx950 path.
//
// Lattice layout in HBM (same even/odd structure-of-arrays order as the
// reference, index.h:124-132, array.h): one contiguous plane per direction mu,
// each plane split into the even and the odd half,
//     angle(x, mu) = lat[mu*V + parity(x)*V/2 + (lex(x) >> 1)],
//     lex(x) = x0 + Nx*(x1 + Ny*(x2 + Nz*x3)).
// A wavefront of 64 consecutive eo-ids walks consecutive addresses of every
// plane it touches, so each of the 19 angle loads of a link update is a fully
// coalesced 256 B (FP32) or 512 B (FP64) request.  Neighbour eo-ids are
// separable: eo(x + a e_d + b e_e) = eo(x) + off_d(a) + off_e(b) for d != e,
// so eight per-site offsets (forward/backward per direction) give all twelve
// neighbours of a staple with adds only (no division per neighbour).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "u1_math.h"
#include "u1_philox.h"

namespace u1d {

// Per-lattice (per-replica) descriptor, read by every workgroup through the
// scalar cache.  Batched launches index a table of these.
struct RepDesc {
  int32_t N[4];        // Nx, Ny, Nz, Nt
  int32_t nxh;         // Nx / 2
  int32_t HV, V;       // half volume, volume
  int32_t SV;          // spatial volume
  int32_t sy, sz, st;  // eo strides of +y, +z, +t: Nx/2, Nx*Ny/2, Nx*Ny*Nz/2
  int32_t blk_begin;   // first workgroup of this replica in a batched launch
  int32_t nblk;        // workgroups of this replica
  uint32_t c3;         // Philox counter word 3 (replica id)
  uint32_t div_nxh_m, div_ny_m, div_nz_m;  // magic multipliers for id/nxh, /Ny, /Nz
  uint32_t div_nxh_s, div_ny_s, div_nz_s;  // and their shifts
  int32_t bpp;                             // update launches: workgroups per (z, t) plane
  uint32_t div_bpp_m, div_bpp_s;           // magic division by bpp
  float inv_aniso_f;                       // binary32 1 / aniso (binary32 space-space staple weight)
  double beta;
  double aniso;
  uint64_t seed;
  uint64_t lat_off;  // element offset of this replica's angles in the packed buffer
  // Update launches of a single lattice: 1 if the (z, t) planes are taken in
  // 2 x 4 tiles of tile_bz x tile_bt planes (set_phase_tiles), one tile per XCD
  uint32_t tile;
  uint32_t tile_bz, tile_bt;          // Nz / 2, Nt / 4
  uint32_t div_tile_m, div_tile_s;    // magic division by tile_bz * tile_bt
  uint32_t div_bz_m, div_bz_s;        // magic division by tile_bz
  // binary32 update constants, rounded once on the host: aniso, and
  // beta log2(e) for the exp2 form of the Metropolis weight
  float aniso_f, bl2e_f;
};

// Round-up magic division (Granlund & Montgomery 1994): for 1 <= d < 2^31 and
// n < 2^31, n / d = (umulhi(n, m) + n) >> s with s = ceil(log2 d),
// m = floor(2^32 (2^s - d) / d) + 1.
struct MagicDiv {
  uint32_t m, s;
};
inline MagicDiv make_magic(uint32_t d) {
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  const uint64_t m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
  return {(uint32_t)m, s};
}
__host__ __device__ __forceinline__ uint32_t mdiv(uint32_t n, uint32_t m, uint32_t s) {
  return ((uint32_t)(((uint64_t)n * m) >> 32) + n) >> s;  // v_mul_hi_u32 on the device
}

// Site coordinates and the eight neighbour offsets of the even/odd site
// (id, parity) -- the strength-reduced form of indexEO / indexEO_neg
// (index.h:114-153).
struct Site {
  int32_t x[4];
  int32_t fwd[4], bwd[4];
};

__host__ __device__ __forceinline__ void site_of(const RepDesc& r, int32_t id, int32_t parity, Site& s) {
  const uint32_t row = mdiv((uint32_t)id, r.div_nxh_m, r.div_nxh_s);
  const int32_t xh = id - (int32_t)row * r.nxh;
  const uint32_t zt = mdiv(row, r.div_ny_m, r.div_ny_s);
  const int32_t y = (int32_t)row - (int32_t)zt * r.N[1];
  const uint32_t t = mdiv(zt, r.div_nz_m, r.div_nz_s);
  const int32_t z = (int32_t)zt - (int32_t)t * r.N[2];
  const int32_t xodd = (y + z + (int32_t)t + parity) & 1;
  s.x[0] = 2 * xh + xodd;
  s.x[1] = y;
  s.x[2] = z;
  s.x[3] = (int32_t)t;
  // x direction: the eo-id moves only when crossing a pair boundary
  s.fwd[0] = xodd ? ((xh == r.nxh - 1) ? -(r.nxh - 1) : 1) : 0;
  s.bwd[0] = xodd ? 0 : ((xh == 0) ? (r.nxh - 1) : -1);
  s.fwd[1] = (y == r.N[1] - 1) ? -(r.N[1] - 1) * r.sy : r.sy;
  s.bwd[1] = (y == 0) ? (r.N[1] - 1) * r.sy : -r.sy;
  s.fwd[2] = (z == r.N[2] - 1) ? -(r.N[2] - 1) * r.sz : r.sz;
  s.bwd[2] = (z == 0) ? (r.N[2] - 1) * r.sz : -r.sz;
  s.fwd[3] = ((int32_t)t == r.N[3] - 1) ? -(r.N[3] - 1) * r.st : r.st;
  s.bwd[3] = (t == 0) ? (r.N[3] - 1) * r.st : -r.st;
}

// Loads and stores at a uniform base plus a 32-bit element index: the byte
// offset stays 32-bit, so the access compiles to global_load/store with an
// SGPR base and one VGPR offset.
template <typename T>
__device__ __forceinline__ T ld(const T* base, uint32_t idx) {
  return *(const T*)((const char*)base + idx * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(T* base, uint32_t idx, T v) {
  *(T*)((char*)base + idx * (uint32_t)sizeof(T)) = v;
}
// The same at device scope (relaxed atomics: global_load / global_store with
// the sc1 bit): the store is written through to the device's point of
// coherence, and the load is served from there, not from a line that this
// CU's or this XCD's cache filled earlier.  Used by the fused parity pass for
// the angles one workgroup writes and another reads within the launch.
template <typename T>
__device__ __forceinline__ T ldc(const T* base, uint32_t idx) {
  return __hip_atomic_load((T*)((const char*)base + idx * (uint32_t)sizeof(T)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void stc(T* base, uint32_t idx, T v) {
  __hip_atomic_store((T*)((char*)base + idx * (uint32_t)sizeof(T)), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Update-launch decomposition: workgroup b (relative to the replica's first)
// covers chunk (b % bpp) of plane b / bpp in launch order, which is the
// lexicographic plane z + Nz t, or with tiling (set_phase_tiles) the planes of
// one 2 x 4 tile after another; thread tid of a workgroup of bdim threads takes
// even/odd site id = (z + Nz t) * P + chunk * bdim + tid, P = Nx/2 * Ny.  z, t
// (and so the z/t neighbour offsets) are uniform across the workgroup.  valid is false for the tail threads of a plane,
// which are given the plane's last site.
struct PhaseSite {
  bool valid;
  uint32_t id, t;
  int32_t z;
  int32_t fwd[4], bwd[4];
};

// Site i_raw (in-plane index; past the end: the plane's last site, invalid)
// of plane (z, t).
__host__ __device__ __forceinline__ void phase_site_zt(const RepDesc& r, int32_t z, uint32_t t, uint32_t i_raw,
                                                       int32_t parity, PhaseSite& s) {
  const uint32_t P = (uint32_t)r.nxh * (uint32_t)r.N[1];
  s.valid = i_raw < P;
  // a thread past the end of the plane takes the plane's last site: its loads
  // stay in bounds (and hit the cache lines of that site); it stores nothing
  const uint32_t i = s.valid ? i_raw : P - 1;
  s.t = t;
  s.z = z;
  const uint32_t y = mdiv(i, r.div_nxh_m, r.div_nxh_s);
  const int32_t xh = (int32_t)(i - y * (uint32_t)r.nxh);
  s.id = ((uint32_t)z + (uint32_t)r.N[2] * t) * P + i;  // lexicographic plane z + Nz t
  const int32_t xodd = ((int32_t)y + z + (int32_t)t + parity) & 1;
  s.fwd[0] = xodd ? ((xh == r.nxh - 1) ? -(r.nxh - 1) : 1) : 0;
  s.bwd[0] = xodd ? 0 : ((xh == 0) ? (r.nxh - 1) : -1);
  s.fwd[1] = ((int32_t)y == r.N[1] - 1) ? -(r.N[1] - 1) * r.sy : r.sy;
  s.bwd[1] = (y == 0) ? (r.N[1] - 1) * r.sy : -r.sy;
  s.fwd[2] = (z == r.N[2] - 1) ? -(r.N[2] - 1) * r.sz : r.sz;
  s.bwd[2] = (z == 0) ? (r.N[2] - 1) * r.sz : -r.sz;
  s.fwd[3] = ((int32_t)t == r.N[3] - 1) ? -(r.N[3] - 1) * r.st : r.st;
  s.bwd[3] = (t == 0) ? (r.N[3] - 1) * r.st : -r.st;
}

__host__ __device__ __forceinline__ void phase_site(const RepDesc& r, uint32_t b, uint32_t tid, uint32_t bdim,
                                                    int32_t parity, PhaseSite& s) {
  const uint32_t plane = mdiv(b, r.div_bpp_m, r.div_bpp_s);  // in launch order
  const uint32_t chunk = b - plane * (uint32_t)r.bpp;
  uint32_t t;
  int32_t z;
  if (r.tile) {
    // tile k = 2 tb + zb covers z in [zb Bz, zb Bz + Bz), t in [tb Bt, tb Bt + Bt),
    // its planes in t-major order
    const uint32_t area = r.tile_bz * r.tile_bt;
    const uint32_t k = mdiv(plane, r.div_tile_m, r.div_tile_s);
    const uint32_t w = plane - k * area;
    const uint32_t wt = mdiv(w, r.div_bz_m, r.div_bz_s);
    t = (k >> 1) * r.tile_bt + wt;
    z = (int32_t)((k & 1u) * r.tile_bz + (w - wt * r.tile_bz));
  } else {
    t = mdiv(plane, r.div_nz_m, r.div_nz_s);
    z = (int32_t)(plane - t * (uint32_t)r.N[2]);
  }
  phase_site_zt(r, z, t, chunk * bdim + tid, parity, s);
}

// XCD-aware workgroup order.  The dispatcher hands workgroup b of a launch to
// XCD b % 8 (round robin), and each XCD has its own L2.  Returning logical
// index L such that XCD x gets the contiguous range
// [x*q + min(x, rem), ...) of logical indices (q = n / 8, rem = n % 8) keeps
// neighbouring (z, t) planes -- which share halo data -- in one L2.  It is a
// bijection of [0, n).
__host__ __device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
  const uint32_t x = b & 7u, q = n >> 3, rem = n & 7u;
  return x * q + (x < rem ? x : rem) + (b >> 3);
}

// Host helper: workgroup layout of update launches for a block size.
inline void set_phase_blocks(RepDesc& r, int block) {
  const int P = r.nxh * r.N[1];
  r.bpp = (P + block - 1) / block;
  const MagicDiv m = make_magic((uint32_t)r.bpp);
  r.div_bpp_m = m.m;
  r.div_bpp_s = m.s;
  r.nblk = r.bpp * r.N[2] * r.N[3];
}

// Host helper: take the (z, t) planes of a single lattice in 2 x 4 tiles
// (Nz / 2 x Nt / 4 planes), in the order of the tiles' XCDs under xcd_remap:
// each XCD then holds a tile, whose halo to the other XCDs is smaller than
// that of an Nt / 8-slice slab.  Needs Nt % 4 == 0 (Nz is even); otherwise,
// and for batched launches, planes stay in lexicographic order.
inline void set_phase_tiles(RepDesc& r, bool on) {
  r.tile = on && r.N[2] >= 2 && r.N[3] >= 4 && r.N[3] % 4 == 0 ? 1u : 0u;
  if (!r.tile) return;
  r.tile_bz = (uint32_t)r.N[2] / 2;
  r.tile_bt = (uint32_t)r.N[3] / 4;
  const MagicDiv a = make_magic(r.tile_bz * r.tile_bt), b = make_magic(r.tile_bz);
  r.div_tile_m = a.m;
  r.div_tile_s = a.s;
  r.div_bz_m = b.m;
  r.div_bz_s = b.s;
}

// ------------------------------------------------------- precision traits
template <typename Real>
struct Prec;

template <>
struct Prec<double> {
  static __device__ __forceinline__ void sincos(double x, double* s, double* c) { u1m::sincos_d(x, s, c); }
  static __device__ __forceinline__ double exp(double x) { return u1m::exp_d(x); }
  static __device__ __forceinline__ double atan2(double y, double x) { return u1m::atan2_d(y, x); }
  static __device__ __forceinline__ double u01(uint32_t w) { return u1r::u01_d(w); }
};

template <>
struct Prec<float> {
  static __device__ __forceinline__ void sincos(float x, float* s, float* c) { u1m::sincos_f(x, s, c); }
  // Hardware exp2 (v_exp_f32); used only for the Metropolis weight, whose
  // FP32 parity bar is statistical.
  static __device__ __forceinline__ float exp(float x) { return __expf(x); }
  static __device__ __forceinline__ float atan2(float y, float x) { return u1m::atan2_f(y, x); }
  static __device__ __forceinline__ float u01(uint32_t w) { return u1r::u01_f(w); }
};

// Update-kernel transcendentals: FP64 shares u1_math.h with the oracle; FP32
// uses the hardware units (input of v_sin/v_cos in revolutions).
template <typename Real>
struct UMath;
template <>
struct UMath<double> {
  // sincos_d's bits: sincos_d_main, then its small-argument case only in a
  // wavefront where some lane has |x| < 2^-27 (one vote instead of a test
  // and a 64-bit select per call)
  template <bool PIN = false>
  static __device__ __forceinline__ void sincos_voted(double x, double* s, double* c) {
    u1m::sincos_d_main<PIN>(x, s, c);
    if (!__all(__builtin_fabs(x) >= 0x1p-27))
      if (__builtin_fabs(x) < 0x1p-27) *s = x;
  }
  static __device__ __forceinline__ void sincos(double x, double* s, double* c) { sincos_voted<false>(x, s, c); }
  static __device__ __forceinline__ double exp(double x) { return u1m::exp_d(x); }
  // the same functions for calls inside a loop (coefficients pinned to
  // scalar registers, u1m::kc; identical results)
  static __device__ __forceinline__ void sincos_in_loop(double x, double* s, double* c) {
    sincos_voted<true>(x, s, c);
  }
  // exp_d's bits for every x: exp_d_inrange, then exp_d's three special
  // cases only in a wavefront where some lane's argument is outside
  // |x| < 708 (one vote per hit instead of three tests; the Metropolis
  // weights stay far inside at any usual coupling)
  static __device__ __forceinline__ double exp_in_loop(double x) {
    double y = u1m::exp_d_inrange<true>(x);
    if (!__all(__builtin_fabs(x) < 708.0)) {
      if (x != x) y = x;
      else if (x > 709.782712893384) y = __builtin_huge_val();
      else if (x < -745.1332191019412) y = 0.0;
    }
    return y;
  }
  static __device__ __forceinline__ double atan2(double y, double x) { return u1m::atan2_d(y, x); }
};
template <>
struct UMath<float> {
  static __device__ __forceinline__ void sincos(float x, float* s, float* c) {
    const float rev = x * 0x1.45f306p-3f;  // 1 / (2 pi)
    *s = __builtin_amdgcn_sinf(rev);
    *c = __builtin_amdgcn_cosf(rev);
  }
  static __device__ __forceinline__ float exp(float x) { return __expf(x); }
  // atan2: octant reduction to z = min/max in [0, 1] (division by v_rcp_f32),
  // then a minimax polynomial atan z = z + z^3 A(z^2) on all of [0, 1]
  // (relative error 1.8e-8, gen_math_constants.py: remez_atan01), so no
  // second reduction by tan(pi/8) is needed
  static __device__ __forceinline__ float atan2(float y, float x) {
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const bool swap = ay > ax;
    const float num = swap ? ax : ay, den = swap ? ay : ax;
    const float z = den > 0.0f ? num * __builtin_amdgcn_rcpf(den) : 0.0f;
    const float w = z * z;
    const float ca[u1m::N_ATAN01_MM32] = {U1M_ATAN01_MM32_COEFFS};
    float p = ca[0];
#pragma unroll
    for (int i = 1; i < u1m::N_ATAN01_MM32; ++i) p = __builtin_fmaf(p, w, ca[i]);
    float a = __builtin_fmaf(z * w, p, z);
    if (swap) a = (u1m::KF_PIO2_HI - a) + u1m::KF_PIO2_LO;
    if (__builtin_signbit(x)) a = u1m::KF_PI - a;
    return __builtin_signbit(y) ? -a : a;
  }
};

// ------------------------------------------------------- wave64 reductions
// DPP move of a binary64 value (two 32-bit halves) with a fixed dpp_ctrl:
// 0xB1 = quad_perm [1,0,3,2], 0x4E = quad_perm [2,3,0,1],
// 0x141 = row_half_mirror, 0x140 = row_mirror.  All lanes must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the 64 lanes of a wavefront, the same value in every lane.  Four
// DPP steps give every lane the sum of its 16-lane row (each step adds two
// equal-shaped partial sums, so partner lanes hold bit-identical values);
// the four row sums are then added in row order.  Fixed order: results are
// reproducible run to run.  Must be called with all 64 lanes active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return ((readlane_f64(v, 0) + readlane_f64(v, 16)) + readlane_f64(v, 32)) + readlane_f64(v, 48);
}

// Wavefront lane mask of a predicate (wave-uniform, in scalar registers), and
// the number of lanes of mask m for which p holds: the population count of
// the compare mask on the scalar unit.
__device__ __forceinline__ uint64_t wave_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ uint32_t count_wave(bool p, uint64_t m) {
  return (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(p) & m);
}

// Block sum of NV doubles per thread into out[0..NV) (valid in thread 0).
// Partial sums of the waves go through LDS and are added in wave order.
template <int NV, int MAXW = 16>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (&out)[NV]) {
  __shared__ double red[MAXW][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[w][k] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = 0; i < nw; ++i) s += red[i][k];
      out[k] = s;
    }
  }
  __syncthreads();
}

}  // namespace u1d
