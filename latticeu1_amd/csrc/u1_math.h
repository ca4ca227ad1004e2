// u1_math.h -- elementary functions shared bit-for-bit by host and gfx950 code.
//
// Why this exists: Metropolis and overrelaxation dynamics are chaotic, so a
// one-ulp difference between two sin() implementations grows until accept /
// reject decisions differ.  Sweep-by-sweep FP64 parity between the GPU path and
// the host FP64 restatement therefore needs the *same* arithmetic on both sides.
// Everything below uses only IEEE-754 correctly rounded operations (+ - * /,
// fused multiply-add, round-to-nearest-integer, exact power-of-two scaling),
// so any conforming compiler produces identical bits when floating-point
// contraction is disabled (-ffp-contract=off on both g++ and hipcc) and FMAs are
// written explicitly.
//
// Methods (all published, none taken from a vendor math library):
//  * sincos: Cody & Waite (1980) additive range reduction by pi/2 split into
//    three parts, then minimax polynomials (Remez exchange) of sin/cos on
//    |r| <= pi/4 (sin through r^13, cos through r^14), by Horner's rule.
//  * exp:    Cody-Waite reduction x = k ln2 + r with ln2 split in two parts,
//    a minimax polynomial of e^r on |r| <= ln2/2 through r^12, exact 2^k
//    scaling.
//  * atan2:  octant reduction to z = min/max in [0,1], then the identity
//    atan z = pi/4 + atan((z-1)/(z+1)) for z > tan(pi/8) (the argument
//    formed as (min - max)/(min + max)), then a minimax polynomial of atan on
//    |t| <= tan(pi/8) through t^23.
//  The minimax tables' own relative error is at most 0.11 units of 2^-53
//  (re-measured by the tests); the rest is the rounding of the evaluation.
//  * Bessel I0, I1: their everywhere-convergent power series
//    I0(x) = sum (x^2/4)^k/(k!)^2,  I1(x) = (x/2) sum (x^2/4)^k/(k!(k+1)!),
//    summed until the next term no longer changes the sum.
// Constants come from gen_math_constants.py (100-digit decimal derivation).
//
// Valid range of sincos: |x| < 2^20 (the angles on this path are sums of at
// most Nt link angles, each in [-pi, pi)).
#pragma once
#include <stdint.h>
#include "u1_math_constants.h"

#if defined(__HIPCC__)
#define U1_HD __host__ __device__ inline
#else
#define U1_HD inline
#endif

namespace u1m {

// A polynomial coefficient as an operand of its FMA.  With PIN on the GPU an
// empty asm statement ties each use to a scalar register pair, so that the
// compiler neither hoists the constant out of a surrounding loop into a
// vector register nor, in turn, copies it into the accumulator before a
// two-address FMA (one more vector instruction per polynomial step); the
// scalar moves that rebuild it run on the scalar unit.  For calls inside a
// loop (the Metropolis hits); elsewhere the extra scalar registers would
// only spill.  The value is unchanged, so are the results.
template <bool PIN>
U1_HD double kc(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (PIN) asm volatile("" : "+s"(c));
#endif
  return c;
}

U1_HD double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
U1_HD float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
U1_HD double rintd(double x) { return __builtin_rint(x); }
U1_HD float rintf_(float x) { return __builtin_rintf(x); }
U1_HD double absd(double x) { return __builtin_fabs(x); }
U1_HD float absf_(float x) { return __builtin_fabsf(x); }
U1_HD bool signbitd(double x) { return __builtin_signbit(x); }
U1_HD bool signbitf_(float x) { return __builtin_signbit(x); }

// ---------------------------------------------------------------- FP64 sincos
// sincos_d_main: sincos_d without its case |x| < 2^-27, where sincos_d
// returns sin x = x itself.  PIN: see kc (the same results either way).
template <bool PIN = false>
U1_HD void sincos_d_main(double x, double* sp, double* cp) {
  const double k = rintd(x * K_TWO_OVER_PI);
  double r = fmad(-k, K_PIO2_1, x);
  r = fmad(-k, K_PIO2_2, r);
  r = fmad(-k, K_PIO2_3, r);
  const int q = (int)k & 3;
  const double z = r * r;
  const double cs_[N_SIN64] = {U1M_SIN64_COEFFS};
  const double cc_[N_COS64] = {U1M_COS64_COEFFS};
  double ps = cs_[0];
#pragma unroll
  for (int i = 1; i < N_SIN64; ++i) ps = fmad(ps, z, kc<PIN>(cs_[i]));
  double pc = cc_[0];
#pragma unroll
  for (int i = 1; i < N_COS64; ++i) pc = fmad(pc, z, kc<PIN>(cc_[i]));
  const double sn = fmad(r * z, ps, r);  // r + r^3 P(r^2)
  const double cn = fmad(z, pc, 1.0);    // 1 + r^2 Q(r^2)
  const double a = (q & 1) ? cn : sn;
  const double b = (q & 1) ? sn : cn;
  // quadrant signs as a flip of the sign bit (exact, zeros included): one
  // integer op on the high word instead of a negation and a 64-bit select
  const double sa = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, a) ^ ((uint64_t)(q & 2) << 62));
  const double sb = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, b) ^ ((uint64_t)((q + 1) & 2) << 62));
  *sp = sa;
  *cp = sb;
}

template <bool PIN = false>
U1_HD void sincos_d(double x, double* sp, double* cp) {
  double sa;
  sincos_d_main<PIN>(x, &sa, cp);
  // |x| < 2^-27: sin x = x to binary64 precision, which also keeps sin(-0) =
  // -0 (the reduction turns -0 into +0); cos needs no such case, since there
  // k = 0, r = x, and 1 + r^2 Q(r^2) rounds to exactly 1
  *sp = absd(x) < 0x1p-27 ? x : sa;
}

U1_HD double sin_d(double x) { double s, c; sincos_d(x, &s, &c); return s; }
U1_HD double cos_d(double x) { double s, c; sincos_d(x, &s, &c); return c; }

// ------------------------------------------------------------------- FP64 exp
// exp_d for -745.1332191019412 <= x <= 709.782712893384 (no tests): e^r
// scaled by 2^k with ldexp, which rounds once where the result is subnormal
// and overflows to infinity exactly where the product would.
template <bool PIN = false>
U1_HD double exp_d_inrange(double x) {
  const double k = rintd(x * K_LOG2E);
  double r = fmad(-k, K_LN2_HI, x);
  r = fmad(-k, K_LN2_LO, r);
  const double ce[N_EXP64] = {U1M_EXP64_COEFFS};
  double p = ce[0];
#pragma unroll
  for (int i = 1; i < N_EXP64; ++i) p = fmad(p, r, kc<PIN>(ce[i]));
  const double t = fmad(r * r, p, r);  // r + r^2 E(r) = e^r - 1
  return __builtin_ldexp(1.0 + t, (int)k);
}

template <bool PIN = false>
U1_HD double exp_d(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_huge_val();
  if (x < -745.1332191019412) return 0.0;
  return exp_d_inrange<PIN>(x);
}

// ----------------------------------------------------------------- FP64 atan2
// atan(num / den) for 0 <= num <= den, den > 0, num finite: above tan(pi/8)
// through atan z = pi/4 + atan((z - 1) / (z + 1)) with the reduced argument
// formed directly as (num - den) / (num + den) -- one division either way.
U1_HD double atan_ratio_d(double num, double den) {
  const bool red = num > K_TAN_PIO8 * den;
  const double t = (red ? num - den : num) / (red ? num + den : den);
  const double t2 = t * t;
  const double ca[N_ATAN64] = {U1M_ATAN64_COEFFS};
  double p = ca[0];
#pragma unroll
  for (int i = 1; i < N_ATAN64; ++i) p = fmad(p, t2, ca[i]);
  const double at = fmad(t * t2, p, t);  // t + t^3 A(t^2)
  return red ? (at + K_PIO4_LO) + K_PIO4 : at;
}

U1_HD double atan2_d(double y, double x) {
  if (x != x || y != y) return x + y;
  const double ax = absd(x), ay = absd(y);
  const bool swap = ay > ax;
  const double num = swap ? ax : ay;
  const double den = swap ? ay : ax;
  double a;
  if (den == 0.0) {
    a = 0.0;
  } else if (num == __builtin_huge_val()) {
    a = K_PIO4;  // both infinite
  } else {
    a = atan_ratio_d(num, den);
  }
  if (swap) a = (K_PIO2_HI - a) + K_PIO2_LO;
  if (signbitd(x)) a = (K_PI - a) + K_PI_LO;
  // the sign of y, as a flip of the sign bit (exact)
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, a) ^
                                        (__builtin_bit_cast(uint64_t, y) & 0x8000000000000000ull));
}

// ------------------------------------------------------- wrap into [-pi, pi)
// Exact for inputs in [-5pi, 5pi): subtracts or adds the binary64 2*pi at most
// twice.  The update rules only produce arguments in (-3pi, 3pi].
U1_HD double wrap_d(double x) {
  if (x >= K_PI) x -= K_TWO_PI;
  if (x >= K_PI) x -= K_TWO_PI;
  if (x < -K_PI) x += K_TWO_PI;
  if (x < -K_PI) x += K_TWO_PI;
  return x;
}

// Wrap of the binary32 update kernels, done in binary64: x - 2pi rint(x / 2pi)
// (one rint, no correction steps).  The result lies in [-pi, pi] up to
// binary64 rounding and is the angle of wrap_d(x) within a few binary64 ulps
// (checked on the CPU); the kernels then round it once to binary32.
U1_HD double wrap_near_d(double x) {
  const double k = rintd(x * (K_TWO_OVER_PI * 0.25));
  return fmad(-k, K_TWO_PI, x);
}

// ------------------------------------------------------- Bessel I0, I1 (FP64)
// Returns I1(x)/I0(x) for x >= 0 via the two power series; used by the analytic
// multihit average <e^{i theta}> = conj(S) I1(beta|S|) / (I0(beta|S|) |S|).
U1_HD void bessel_i0i1_d(double x, double* i0, double* i1) {
  const double q = 0.25 * x * x;
  double t0 = 1.0, s0 = 1.0;  // terms of I0
  double t1 = 1.0, s1 = 1.0;  // terms of 2 I1 / x
  for (int k = 1; k < 400; ++k) {
    t0 = t0 * q / (double)(k * k);
    t1 = t1 * q / (double)(k * (k + 1));
    const double n0 = s0 + t0;
    const double n1 = s1 + t1;
    if (n0 == s0 && n1 == s1) break;
    s0 = n0;
    s1 = n1;
  }
  *i0 = s0;
  *i1 = 0.5 * x * s1;
}

// ---------------------------------------------------------------- FP32 sincos
// Same construction in binary32 (Taylor through r^9 / r^10); used by the
// FP32-bulk path, whose parity bar is a tolerance, not bit identity.
U1_HD void sincos_f(float x, float* sp, float* cp) {
  const float k = rintf_(x * KF_TWO_OVER_PI);
  float r = fmaf_(-k, KF_PIO2_1, x);
  r = fmaf_(-k, KF_PIO2_2, r);
  r = fmaf_(-k, KF_PIO2_3, r);
  const int q = (int)k & 3;
  const float z = r * r;
  const float cs_[N_SIN32] = {U1M_SIN32_COEFFS};
  const float cc_[N_COS32] = {U1M_COS32_COEFFS};
  float ps = cs_[0];
#pragma unroll
  for (int i = 1; i < N_SIN32; ++i) ps = fmaf_(ps, z, cs_[i]);
  float pc = cc_[0];
#pragma unroll
  for (int i = 1; i < N_COS32; ++i) pc = fmaf_(pc, z, cc_[i]);
  const float sn = fmaf_(r * z, ps, r);
  const float cn = fmaf_(z, pc, 1.0f);
  const float a = (q & 1) ? cn : sn;
  const float b = (q & 1) ? sn : cn;
  *sp = (q & 2) ? -a : a;
  *cp = ((q + 1) & 2) ? -b : b;
}

U1_HD float atan01_f(float z) {
  const bool red = z > KF_TAN_PIO8;
  const float t = red ? (z - 1.0f) / (z + 1.0f) : z;
  const float t2 = t * t;
  const float ca[N_ATAN32] = {U1M_ATAN32_COEFFS};
  float p = ca[0];
#pragma unroll
  for (int i = 1; i < N_ATAN32; ++i) p = fmaf_(p, t2, ca[i]);
  const float at = fmaf_(t * t2, p, t);
  return red ? at + KF_PIO4 : at;
}

U1_HD float atan2_f(float y, float x) {
  if (x != x || y != y) return x + y;
  const float ax = absf_(x), ay = absf_(y);
  const bool swap = ay > ax;
  const float num = swap ? ax : ay;
  const float den = swap ? ay : ax;
  float a;
  if (den == 0.0f) a = 0.0f;
  else if (num == __builtin_huge_valf()) a = KF_PIO4;
  else a = atan01_f(num / den);
  if (swap) a = (KF_PIO2_HI - a) + KF_PIO2_LO;
  if (signbitf_(x)) a = KF_PI - a;
  return signbitf_(y) ? -a : a;
}

}  // namespace u1m
