// u1_update.hip -- checkerboard Metropolis and overrelaxation phases, hot start.
//
// One launch per (pass, parity, mu) phase, as the reference (update.cu:177-190,
// 227-231, 301-316, 379-394), or one k_pair launch for the phases (parity, 0)
// and (parity, 1) of a single lattice (see k_pair).  Within a phase no link
// reads a link of the same phase: the staple of (x, mu) reads theta_nu
// (nu != mu) and theta_mu at the other parity (staple.h:85-92), so the launch
// is race free.
//
// Work decomposition (instruction-issue bound, so indexing is kept scalar):
// a workgroup covers a contiguous chunk of one (z, t) plane of the half
// lattice (P = Nx/2 * Ny even/odd sites per plane), so z, t, their boundary
// tests and the +-z, +-t neighbour offsets are uniform (SGPR) values; a thread
// derives only (x/2, y) from its index in the plane.  All angle loads use a
// uniform 64-bit plane base plus a 32-bit byte offset, which maps to
// global_load with an SGPR base.
//
// FP64 parity contract with the host oracle (oracle/u1_oracle.cpp): the
// staple, the Metropolis weight and the OR reflection below perform the same
// IEEE operations in the same order as the oracle, on the same Philox words,
// with the same u1_math.h functions.  The library is compiled with
// -ffp-contract=off so that no FMA is formed implicitly.
//
// FP32-bulk variant: angles, staples, proposals, the Metropolis weight and the
// OR angle in binary32 with the hardware transcendentals (v_sin_f32,
// v_cos_f32, v_exp_f32, v_rcp_f32) and explicit FMAs; the periodic wrap into
// [-pi, pi] in binary64 (wrap_near_d), rounded once to binary32.  It reads the
// same Philox words as the FP64 path.  Its parity bar is statistical.
#include "u1_chain.h"
#include "u1_kernels.h"

namespace u1d {

namespace {

template <typename Real>
struct Staple {
  Real re, im;
};

// Plaquette pieces of a link (x, MU) for the directions nu < MU: the sums of
// e^{i up_nu} (u) and e^{i lo_nu} (l).  With the link's new angle theta,
// theta + up_nu = -P_{nu,MU}(x) and theta + lo_nu = P_{nu,MU}(x - nu), so
// every plaquette of the planes (nu, MU), nu < MU, is one piece of exactly one
// odd-site MU link.
template <typename Real>
struct PlaqPieces {
  Real ur, ui, lr, li;
};

// First direction nu != mu of the space-time (st) or space-space class.
__host__ __device__ constexpr int first_nu(int mu, bool st) {
  for (int nu = 0; nu < 4; ++nu)
    if (nu != mu && ((mu == 3 || nu == 3) == st)) return nu;
  return -1;
}

// The 18 angles of the staple of link (x, MU) (staple.h:85-92): for the j-th
// direction nu != MU (j = nu for nu < MU, else nu - 1),
//   a = theta_nu(x+mu), b = theta_mu(x+nu), c = theta_nu(x),
//   d = theta_nu(x+mu-nu), e = theta_mu(x-nu), f = theta_nu(x-nu).
// fwd/bwd[d] are the eo-id offsets to x +- e_d.
template <typename Real>
struct StapleIn {
  Real a[3], b[3], c[3], d[3], e[3], f[3];
};

// Loads the staple angles.  The c and d angles of the directions nu in the
// bit mask DEFM (theta_nu at the link's own parity) are left out: for a fused
// pair (DEFM = 1) to phase_load_nu0_lds, for a fused pass to phase_load_dep.
template <typename Real, int MU, int DEFM = 0>
__device__ __forceinline__ void staple_load(const Real* __restrict__ lat, const RepDesc& r, uint32_t same,
                                            uint32_t opp, const int32_t (&fwd)[4], const int32_t (&bwd)[4],
                                            StapleIn<Real>& in) {
  const Real* pm = lat + (size_t)MU * r.V;
#pragma unroll
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == MU) continue;
    const int j = nu < MU ? nu : nu - 1;
    const Real* pn = lat + (size_t)nu * r.V;
    in.a[j] = ld(pn, opp + fwd[MU]);
    in.b[j] = ld(pm, opp + fwd[nu]);
    if (!((DEFM >> nu) & 1)) {
      in.c[j] = ld(pn, same);
      in.d[j] = ld(pn, same + fwd[MU] + bwd[nu]);
    }
    in.e[j] = ld(pm, opp + bwd[nu]);
    in.f[j] = ld(pn, opp + bwd[nu]);
  }
}
// Staple of link (x, MU) (staple.h:79-102) from its angles, with the operation
// order of the oracle.  With MEAS the plaquette pieces are collected on the
// side (the staple is unchanged).
template <typename Real, int MU, bool MEAS = false>
__device__ __forceinline__ Staple<Real> staple_sum(const StapleIn<Real>& in, Real xi, Real inv_xi, bool aniso,
                                                   PlaqPieces<Real>* pp = nullptr) {
  Real ssr = 0, ssi = 0, str = 0, sti = 0;
#pragma unroll
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == MU) continue;
    const int j = nu < MU ? nu : nu - 1;
    const Real up = (in.a[j] - in.b[j]) - in.c[j];
    const Real lo = ((-in.d[j]) - in.e[j]) + in.f[j];
    Real su, cu, sl, cl;
    UMath<Real>::sincos(up, &su, &cu);
    UMath<Real>::sincos(lo, &sl, &cl);
    const Real ar = cu + cl, ai = su + sl;
    if (MEAS && nu < MU) {
      pp->ur += cu;
      pp->ui += su;
      pp->lr += cl;
      pp->li += sl;
    }
    // binary32: the first term of each class initialises its sum (the oracle's
    // 0 + x is kept in binary64, where the kernels match it bit for bit)
    const bool st_class = MU == 3 || nu == 3;
    const bool first = sizeof(Real) == sizeof(float) && nu == first_nu(MU, st_class);
    if (st_class) {
      str = first ? ar : str + ar;
      sti = first ? ai : sti + ai;
    } else {
      ssr = first ? ar : ssr + ar;
      ssi = first ? ai : ssi + ai;
    }
  }
  if constexpr (sizeof(Real) == sizeof(float)) {
    // binary32: the weighted form with no test of xi; at xi == 1 both weights
    // are exactly 1 and the fused sum equals the plain sum ss + st
    if (MU == 3) return {str * xi, sti * xi};  // no space-space staples
    return {__builtin_fmaf(ssr, inv_xi, str * xi), __builtin_fmaf(ssi, inv_xi, sti * xi)};
  } else {
    if (!aniso) return {ssr + str, ssi + sti};  // xi == 1: identical to ss/1 + st*1
    return {ssr / xi + str * xi, ssi / xi + sti * xi};
  }
}

template <typename Real>
struct MetroState {
  Real th, re_old;
};

// One Metropolis hit with a fixed staple (lattice_functions.h:68-80 in the
// Re(S e^{i theta}) form), binary64: the oracle's operations in its order.
// wu: proposal word, wb: accept-test word.  For theta + delta one wrap step
// each way suffices because |delta| <= step <= pi.
template <typename Real>
__device__ __forceinline__ bool metro_hit(MetroState<Real>& m, Staple<Real> S, Real beta, int proposal, double step,
                                          uint32_t wu, uint32_t wb) {
  // a * (2u - 1) with u = wu 2^-32 as one FMA: 2u - 1 is exact in binary64,
  // so fma(wu, a 2^-31, -a) is the same correctly rounded product as the
  // oracle's two-step form (a 2^-31 is exact).
  double tnd;
  if (proposal == 0) {
    tnd = __builtin_fma((double)wu, u1m::K_PI * 0x1p-31, -u1m::K_PI);
  } else {
    tnd = (double)m.th + __builtin_fma((double)wu, step * 0x1p-31, -step);
    if (tnd >= u1m::K_PI) tnd -= u1m::K_TWO_PI;
    if (tnd < -u1m::K_PI) tnd += u1m::K_TWO_PI;
  }
  const Real tn = (Real)tnd;
  Real sn, cn;
  UMath<Real>::sincos_in_loop(tn, &sn, &cn);
  const Real re_new = S.re * cn - S.im * sn;
  const Real dS = UMath<Real>::exp_in_loop(beta * (re_new - m.re_old));
  const Real b = Prec<Real>::u01(wb);
  const bool ok = dS > b;
  if (ok) {
    m.th = tn;
    m.re_old = re_new;
  }
  return ok;
}

// The same hit in binary32, from the same two words.  The proposal is
// a (2 wu 2^-32 - 1) as one FMA, with a = pi (independent: pscale = pi 2^-31,
// poff = pi) or a = step (theta + delta).  theta + delta is not wrapped here:
// the caller wraps theta (x - 2 pi rint(x / 2 pi), any range) after every
// second pair of hits and after the last (|theta| <= pi + 4 step <= 5 pi in
// between, well inside the hardware sin/cos range).  The accept test
// exp(beta dRe) > wb 2^-32 is evaluated as 2^(beta log2(e) dRe + 32) > wb
// (bl2e = beta log2(e)): one v_exp_f32 and the word converted without scaling.
// thm is 1 for theta + delta and 0 for the independent proposal, so both
// forms are one more FMA (theta * thm + offset, rounded once).
__device__ __forceinline__ bool metro_hit_f32(MetroState<float>& m, Staple<float> S, float bl2e, float thm,
                                              float pscale, float poff, uint32_t wu, uint32_t wb) {
  const float tn = __builtin_fmaf(m.th, thm, __builtin_fmaf((float)wu, pscale, -poff));
  float sn, cn;
  UMath<float>::sincos(tn, &sn, &cn);
  const float re_new = __builtin_fmaf(S.re, cn, -(S.im * sn));
  const float x = __builtin_fmaf(bl2e, re_new - m.re_old, 32.0f);
  const bool ok = __builtin_amdgcn_exp2f(x) > (float)wb;
  if (ok) {
    m.th = tn;
    m.re_old = re_new;
  }
  return ok;
}

// SPT even/odd sites per thread: sites chunk*SPT*bdim + k*bdim + tid, k < SPT,
// of one (z, t) plane.  All angle loads of all SPT sites are issued before the
// transcendental work, so each wave keeps SPT x 19 loads in flight.  Threads
// past the end of a plane work on the plane's last site (in bounds) and do not
// store.
//
// MEAS (MU >= 1, parity 1 of the last pass of a sweep): the logical
// workgroup L also writes row L of a.meas with its sums over the plaquettes of
// the planes (nu, MU), nu < MU, after the update -- [sum (1 - cos P), sum
// -sin P] in the space-space (MU < 3) or space-time (MU = 3) slot.  The three
// phases (1, 1), (1, 2), (1, 3) come after every link of those planes has
// reached its final value, and together cover all six planes once.
//
// A thread's share of a phase is the sites tid0 + k blockDim.x (k < SPT) of
// chunk b in phase_site numbering (bdim sites per chunk), from phase_sites.
// phase_load issues all of their angle loads; phase_compute forms the staples, updates and
// stores the links, adds the plaquette sums of its valid sites to v (MEAS) and
// the accepted hits of its wavefront to acc_wave; the caller reduces both.
template <typename Real, int SPT>
struct PhaseIn {
  PhaseSite ps[SPT];
  uint32_t same[SPT];
  StapleIn<Real> in[SPT];
  Real th[SPT];
};

// The thread's sites; returns true if the workgroup's plane is frozen
// (multilevel.cu:41,141 -- spatial links on slices t % ml == 0 are frozen; t
// is uniform over the workgroup, so the whole workgroup skips the phase).
template <int MU, int SPT>
__device__ __forceinline__ bool phase_sites(const PhaseArgs& a, const RepDesc& r, uint32_t b, uint32_t tid0,
                                            uint32_t bdim, PhaseSite (&ps)[SPT]) {
#pragma unroll
  for (int k = 0; k < SPT; ++k) phase_site(r, b, tid0 + k * blockDim.x, bdim, a.parity, ps[k]);
  return MU != 3 && a.ml != 0 && (ps[0].t & (uint32_t)(a.ml - 1)) == 0;
}

// The angle loads of the sites in q.ps (but the c, d angles of DEFM).
template <typename Real, int MU, int SPT, int DEFM = 0>
__device__ __forceinline__ void phase_load(const Real* __restrict__ lat_base, const PhaseArgs& a, const RepDesc& r,
                                           PhaseIn<Real, SPT>& q) {
  const Real* lat = lat_base + r.lat_off;
#pragma unroll
  for (int k = 0; k < SPT; ++k) {
    q.same[k] = (uint32_t)(a.parity * r.HV) + q.ps[k].id;
    const uint32_t opp = (uint32_t)((1 - a.parity) * r.HV) + q.ps[k].id;
    staple_load<Real, MU, DEFM>(lat, r, q.same[k], opp, q.ps[k].fwd, q.ps[k].bwd, q.in[k]);
    q.th[k] = ld(lat + (size_t)MU * r.V, q.same[k]);
  }
}

// The c, d angles of the directions in DEFM that phase_load left out --
// theta_nu(x) and theta_nu(x + mu - nu) at the link's own parity, written
// earlier in the same launch by another workgroup -- as device-scope loads.
template <typename Real, int MU, int SPT, int DEFM>
__device__ __forceinline__ void phase_load_dep(const Real* __restrict__ lat_base, const RepDesc& r,
                                               PhaseIn<Real, SPT>& q) {
  const Real* lat = lat_base + r.lat_off;
#pragma unroll
  for (int k = 0; k < SPT; ++k)
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) {
      if (nu == MU || !((DEFM >> nu) & 1)) continue;
      const int j = nu < MU ? nu : nu - 1;
      const Real* pn = lat + (size_t)nu * r.V;
      q.in[k].c[j] = ldc(pn, q.same[k]);
      q.in[k].d[j] = ldc(pn, q.same[k] + (uint32_t)(q.ps[k].fwd[MU] + q.ps[k].bwd[nu]));
    }
}

// The two angles phase_load<..., DEFM = 1> left out -- theta_0 at the
// link's own parity at x and x + e1 - e0, both in the site's (z, t) plane --
// from lds_th, where phase 0 of this workgroup left the plane's new theta_0
// by in-plane index (id - pbase).
template <typename Real, int SPT>
__device__ __forceinline__ void phase_load_nu0_lds(const Real* lds_th, uint32_t pbase, PhaseIn<Real, SPT>& q) {
#pragma unroll
  for (int k = 0; k < SPT; ++k) {
    const uint32_t i = q.ps[k].id - pbase;
    q.in[k].c[0] = lds_th[i];
    q.in[k].d[0] = lds_th[i + (uint32_t)(q.ps[k].fwd[1] + q.ps[k].bwd[0])];
  }
}

// Phase (parity, 1) of a fused pair reads seven angles that phase (parity, 0)
// of the same site has already loaded, at the same addresses, and that
// neither phase writes: the link's own theta_1(x) (c of nu = 1 in phase 0),
// a = theta_0(x + e1) and b = theta_1(x + e0) of nu = 0 (b and a of nu = 1 in
// phase 0), and c = theta_nu(x), f = theta_nu(x - e_nu) of nu = 2, 3 (the
// same slots j = 1, 2 in both phases).  pair_shared keeps them from phase 0's
// staple angles, pair_take hands them to phase 1.
template <typename Real>
struct PairShared {
  Real th, a0, b0, c2, f2, c3, f3;
};
template <typename Real>
__device__ __forceinline__ PairShared<Real> pair_shared(const StapleIn<Real>& in0) {
  return {in0.c[0], in0.b[0], in0.a[0], in0.c[1], in0.f[1], in0.c[2], in0.f[2]};
}
template <typename Real>
__device__ __forceinline__ void pair_take(const PairShared<Real>& s, StapleIn<Real>& in1, Real& th1) {
  th1 = s.th;
  in1.a[0] = s.a0;
  in1.b[0] = s.b0;
  in1.c[1] = s.c2;
  in1.f[1] = s.f2;
  in1.c[2] = s.c3;
  in1.f[2] = s.f3;
}

// The phase-(parity, 1) angles of the sites in q.ps that neither the LDS
// exchange (c, d of nu = 0) nor pair_take provides: e, f of nu = 0 and a, b,
// d, e of nu = 2, 3 -- ten loads instead of seventeen (the addresses of
// staple_load for MU = 1).
template <typename Real, int SPT>
__device__ __forceinline__ void pair1_load(const Real* __restrict__ lat_base, const PhaseArgs& a, const RepDesc& r,
                                           PhaseIn<Real, SPT>& q) {
  const Real* lat = lat_base + r.lat_off;
  const Real* p0 = lat;
  const Real* p1 = lat + (size_t)r.V;
#pragma unroll
  for (int k = 0; k < SPT; ++k) {
    const PhaseSite& s = q.ps[k];
    q.same[k] = (uint32_t)(a.parity * r.HV) + s.id;
    const uint32_t opp = (uint32_t)((1 - a.parity) * r.HV) + s.id;
    q.in[k].e[0] = ld(p1, opp + s.bwd[0]);
    q.in[k].f[0] = ld(p0, opp + s.bwd[0]);
#pragma unroll
    for (int nu = 2; nu < 4; ++nu) {
      const Real* pn = lat + (size_t)nu * r.V;
      q.in[k].a[nu - 1] = ld(pn, opp + s.fwd[1]);
      q.in[k].b[nu - 1] = ld(p1, opp + s.fwd[nu]);
      q.in[k].d[nu - 1] = ld(pn, q.same[k] + s.fwd[1] + s.bwd[nu]);
      q.in[k].e[nu - 1] = ld(p1, opp + s.bwd[nu]);
    }
  }
}

// lds_th (nullable): the new angles are also written to lds_th[id - pbase],
// the site's index within its (z, t) plane (fused pairs, phase 0).  COH: the
// new angles are stored at device scope (fused passes).
template <typename Real, int MU, int KIND, int SPT, bool MEAS, bool COH = false>
__device__ __forceinline__ void phase_compute(Real* __restrict__ lat_base, const PhaseArgs& a, const RepDesc& r,
                                              PhaseIn<Real, SPT>& q, double (&v)[2], uint32_t& acc_wave,
                                              Real* lds_th = nullptr, uint32_t pbase = 0) {
  Real* pmu = lat_base + r.lat_off + (size_t)MU * r.V;
  const bool aniso = r.aniso != 1.0;
  const Real xi = sizeof(Real) == sizeof(float) ? (Real)r.aniso_f : (Real)r.aniso;
  const PhaseSite(&ps)[SPT] = q.ps;
  const uint32_t(&same)[SPT] = q.same;
  Staple<Real> S[SPT];
  Real th[SPT];
  PlaqPieces<Real> pp[SPT];
#pragma unroll
  for (int k = 0; k < SPT; ++k) {
    pp[k] = {0, 0, 0, 0};
    S[k] = staple_sum<Real, MU, MEAS>(q.in[k], xi, (Real)r.inv_aniso_f, aniso, &pp[k]);
    th[k] = q.th[k];
  }
  // accepted hits of this wavefront (valid sites only): a wave-uniform count
  // from the accept masks, kept on the scalar unit
  if (KIND == KIND_METROPOLIS) {
    MetroState<Real> m[SPT];
    uint32_t l[SPT];
    uint64_t vmask[SPT];  // lanes with a valid site
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      m[k].th = th[k];
      Real s0, c0;
      UMath<Real>::sincos(th[k], &s0, &c0);
      if constexpr (sizeof(Real) == sizeof(float))
        m[k].re_old = __builtin_fmaf(S[k].re, c0, -(S[k].im * s0));
      else
        m[k].re_old = S[k].re * c0 - S[k].im * s0;
      l[k] = same[k] + (uint32_t)MU * (uint32_t)r.V;
      vmask[k] = wave_mask(ps[k].valid);
    }
    const uint32_t sweep = a.sweep_dev ? *a.sweep_dev * a.sweep_mul + a.sweep : a.sweep;
    const Real beta = (Real)r.beta;
    // binary32 hit constants (metro_hit_f32), rounded on the host
    const float bl2e = r.bl2e_f;
    const float pscale = a.pscale_f, poff = a.poff_f, thm = a.thm_f;
    // hits in blocks of up to four (two Philox draws of four words each)
    for (int h4 = 0; h4 < a.nhit; h4 += 4) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int h = h4 + 2 * j;
        if (h >= a.nhit) break;
        const uint32_t c2 = u1r::ctr2(a.stream, a.pass, (uint32_t)(h >> 1));
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
          const u1r::U32x4 w = u1r::draw(r.seed, l[k], sweep, c2, r.c3);
          if constexpr (sizeof(Real) == sizeof(float)) {
            const bool ok0 = metro_hit_f32(m[k], S[k], bl2e, thm, pscale, poff, w.v[0], w.v[1]);
            acc_wave += count_wave(ok0, vmask[k]);
            if (h + 1 < a.nhit) {
              const bool ok1 = metro_hit_f32(m[k], S[k], bl2e, thm, pscale, poff, w.v[2], w.v[3]);
              acc_wave += count_wave(ok1, vmask[k]);
            }
          } else {
            const bool ok0 = metro_hit<Real>(m[k], S[k], beta, a.proposal, a.step, w.v[0], w.v[1]);
            acc_wave += count_wave(ok0, vmask[k]);
            if (h + 1 < a.nhit) {
              const bool ok1 = metro_hit<Real>(m[k], S[k], beta, a.proposal, a.step, w.v[2], w.v[3]);
              acc_wave += count_wave(ok1, vmask[k]);
            }
          }
        }
      }
      // binary32: back into [-pi, pi] after each block, i.e. after every
      // second pair of hits and after the last (|theta| <= pi + 4 step in
      // between)
      if constexpr (sizeof(Real) == sizeof(float)) {
#pragma unroll
        for (int k = 0; k < SPT; ++k) m[k].th = (float)u1m::wrap_near_d((double)m[k].th);
      }
    }
#pragma unroll
    for (int k = 0; k < SPT; ++k) th[k] = m[k].th;
  } else {
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      // lattice_functions.h:83-89: theta' = -theta - 2 arg(S), wrapped
      const Real t2 = UMath<Real>::atan2(S[k].im, S[k].re);
      if constexpr (sizeof(Real) == sizeof(float))  // -theta - 2 t2 and the wrap in binary64
        th[k] = (float)u1m::wrap_near_d(__builtin_fma(-2.0, (double)t2, -(double)th[k]));
      else
        th[k] = u1m::wrap_d(-th[k] - 2.0 * t2);
    }
  }
#pragma unroll
  for (int k = 0; k < SPT; ++k)
    if (ps[k].valid) {
      if (COH)
        stc(pmu, same[k], th[k]);
      else
        st(pmu, same[k], th[k]);
      if (lds_th) lds_th[ps[k].id - pbase] = th[k];
    }
  if (MEAS) {
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      Real sn, cn;
      UMath<Real>::sincos(th[k], &sn, &cn);
      // sum cos(theta + up) + cos(theta + lo); sum sin(theta + up) - sin(theta + lo)
      const Real re = cn * (pp[k].ur + pp[k].lr) - sn * (pp[k].ui + pp[k].li);
      const Real im = sn * (pp[k].ur - pp[k].lr) + cn * (pp[k].ui - pp[k].li);
      if (ps[k].valid) {
        v[0] += (double)(2 * MU) - (double)re;
        v[1] += (double)im;
      }
    }
  }
}

// Workgroup epilogue: the plaquette row L of a measuring phase (block_sum:
// every thread of the workgroup takes part) and the wavefront's accepts.
template <int MU, int KIND, bool MEAS>
__device__ __forceinline__ void phase_finish(const PhaseArgs& a, uint32_t L, double (&v)[2], uint32_t acc_wave) {
  if (MEAS) {
    double o[2];
    block_sum<2>(v, o);
    if (threadIdx.x == 0) {
      const int c = MU == 3 ? 2 : 0;
      double* row = a.meas + 4 * (size_t)L;
      row[c] = o[0];
      row[c + 1] = o[1];
      row[2 - c] = 0.0;
      row[3 - c] = 0.0;
    }
  }
  if (KIND == KIND_METROPOLIS && a.accepts && (threadIdx.x & 63) == 0 && acc_wave)
    atomicAdd(a.accepts, (unsigned long long)acc_wave);
}

template <typename Real, int MU, int KIND, bool BATCHED, int SPT, bool MEAS = false>
__global__ __launch_bounds__(512) void k_phase(Real* __restrict__ lat_base, PhaseArgs a) {
  const uint32_t L = a.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  const RepDesc& r = BATCHED ? a.reps[a.blockmap[L]] : a.single;
  const uint32_t b = L - (uint32_t)r.blk_begin;
  double v[2] = {0.0, 0.0};
  uint32_t acc_wave = 0;
  PhaseIn<Real, SPT> q;
  // (phase_sites and the early exit written out here: through the helper,
  // hipcc turns the site arithmetic into divergent branches)
#pragma unroll
  for (int k = 0; k < SPT; ++k) phase_site(r, b, threadIdx.x + k * blockDim.x, SPT * blockDim.x, a.parity, q.ps[k]);
  if (MU != 3 && a.ml != 0 && (q.ps[0].t & (uint32_t)(a.ml - 1)) == 0) return;
  phase_load<Real, MU, SPT>(lat_base, a, r, q);
  phase_compute<Real, MU, KIND, SPT, MEAS>(lat_base, a, r, q, v, acc_wave);
  phase_finish<MU, KIND, MEAS>(a, L, v, acc_wave);
}

// The phases (parity, 0) and (parity, 1) of one single lattice in one launch.
// Workgroup L owns the whole (z, t) plane L (set_phase_blocks(r, P): one chunk
// per plane) and runs phase (parity, 0) over it, SPT * blockDim.x sites per
// step, then phase (parity, 1), with a workgroup barrier in between.  This is
// race free and gives the results of the two separate launches bit for bit:
//  * the staple of (x, 1) reads the phase-0 links theta_0 at this parity only
//    at x and x + e1 - e0, both in the same (z, t) plane, so in this
//    workgroup: phase 0 also leaves them in LDS (P angles of the plane, by
//    in-plane index), and phase 1 reads those two from there after the
//    barrier instead of reloading them;
//  * phase (parity, 0) reads theta_1 at this parity only at x and x + e0 - e1,
//    also in the same plane, so before this workgroup's phase 1 rewrites them
//    (its loads have returned before the barrier);
//  * nothing else either phase reads is written by the launch.
// The pair reads the lattice once (4V angles) and writes V/2 + V/2: 5V angles
// against 9V for two launches.  MEAS applies to phase (parity, 1).  Dynamic
// LDS: P angles (pair_lds_bytes).  Of phase 1's nineteen angles two come
// from LDS; in the two forms below that keep a thread's sites over both
// phases, seven more come from phase 0's loads of the same site
// (pair_shared) and phase 1 loads ten:
//  * one step covers the plane (P <= SPT * blockDim.x, SPT <= 3): every load
//    of phase 1 is issued before phase 0 computes, so phase 1 waits on the
//    barrier and two LDS reads only;
//  * two steps (SPT = 1): the seven shared angles of both steps are kept in
//    registers across the barrier.
// Otherwise (more steps) phase 1 loads its seventeen angles after the
// barrier.
// FORM picks the form at compile time for a single lattice, whose plane
// size the host knows (PAIR_ONE_STEP, PAIR_TWO_STEPS, PAIR_LOOP): each
// kernel then holds the code of its own form only (in one kernel holding
// all three, the compiler's schedule of the measuring two-step form took
// 42 us where its own kernel takes 30 us at 48^4); PAIR_ANY (batched
// replicas, planes of several sizes) decides per workgroup, without the
// measuring two-step form.
enum { PAIR_ONE_STEP = 0, PAIR_TWO_STEPS = 1, PAIR_LOOP = 2, PAIR_ANY = 3 };
template <typename Real, int KIND, int SPT, bool MEAS, bool BATCHED = false, int FORM = PAIR_ANY>
__global__ __launch_bounds__(pair_max_block(sizeof(Real), SPT)) void k_pair(Real* __restrict__ lat_base, PhaseArgs a) {
  extern __shared__ __align__(8) unsigned char pair_lds[];
  Real* th0 = reinterpret_cast<Real*>(pair_lds);
  const uint32_t L = a.xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  const RepDesc& r = BATCHED ? a.reps[a.blockmap[L]] : a.single;
  const uint32_t b = L - (uint32_t)r.blk_begin;  // the replica's plane (one chunk per plane)
  const uint32_t P = (uint32_t)r.nxh * (uint32_t)r.N[1];
  const uint32_t step = SPT * blockDim.x;
  double v0[2] = {0.0, 0.0}, v[2] = {0.0, 0.0};
  uint32_t acc0 = 0, acc1 = 0;
  constexpr bool kOne = SPT <= 3 && (FORM == PAIR_ONE_STEP || FORM == PAIR_ANY);
  constexpr bool kTwo = SPT == 1 && (FORM == PAIR_TWO_STEPS || (FORM == PAIR_ANY && !MEAS));
  constexpr bool kLoop = FORM == PAIR_LOOP || FORM == PAIR_ANY;
  if (kOne && (FORM != PAIR_ANY || P <= step)) {
    PhaseIn<Real, SPT> q0, q1;
    // a frozen plane (multilevel mask, uniform over the workgroup) is frozen
    // for both phases: the whole workgroup leaves before the barrier
#pragma unroll
    for (int k = 0; k < SPT; ++k) phase_site(r, b, threadIdx.x + k * blockDim.x, step, a.parity, q0.ps[k]);
    if (a.ml != 0 && (q0.ps[0].t & (uint32_t)(a.ml - 1)) == 0) return;
    const uint32_t pbase = ((uint32_t)q0.ps[0].z + (uint32_t)r.N[2] * q0.ps[0].t) * P;  // the plane's first id
#pragma unroll
    for (int k = 0; k < SPT; ++k) q1.ps[k] = q0.ps[k];  // the same sites
    phase_load<Real, 0, SPT>(lat_base, a, r, q0);
    pair1_load<Real, SPT>(lat_base, a, r, q1);
    phase_compute<Real, 0, KIND, SPT, false>(lat_base, a, r, q0, v0, acc0, th0, pbase);
#pragma unroll
    for (int k = 0; k < SPT; ++k) pair_take(pair_shared(q0.in[k]), q1.in[k], q1.th[k]);
    __syncthreads();
    phase_load_nu0_lds<Real, SPT>(th0, pbase, q1);
    phase_compute<Real, 1, KIND, SPT, MEAS>(lat_base, a, r, q1, v, acc1);
  } else if (kTwo && (FORM != PAIR_ANY || P <= 2 * step)) {
    // two steps per phase: the angles phase 1 shares with phase 0 stay in
    // registers across the barrier (pair_shared), one set per step
    PhaseIn<Real, SPT> q;
    PairShared<Real> sh[2][SPT];
    uint32_t pbase = 0;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (phase_sites<0, SPT>(a, r, b, s * step + threadIdx.x, step, q.ps)) return;  // frozen plane
      pbase = ((uint32_t)q.ps[0].z + (uint32_t)r.N[2] * q.ps[0].t) * P;
      phase_load<Real, 0, SPT>(lat_base, a, r, q);
      phase_compute<Real, 0, KIND, SPT, false>(lat_base, a, r, q, v0, acc0, th0, pbase);
#pragma unroll
      for (int k = 0; k < SPT; ++k) sh[s][k] = pair_shared(q.in[k]);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      phase_sites<1, SPT>(a, r, b, s * step + threadIdx.x, step, q.ps);
      pair1_load<Real, SPT>(lat_base, a, r, q);
#pragma unroll
      for (int k = 0; k < SPT; ++k) pair_take(sh[s][k], q.in[k], q.th[k]);
      phase_load_nu0_lds<Real, SPT>(th0, pbase, q);
      phase_compute<Real, 1, KIND, SPT, MEAS>(lat_base, a, r, q, v, acc1);
    }
  } else if (kLoop) {
    uint32_t pbase = 0;
    for (uint32_t base = 0; base < P; base += step) {
      PhaseIn<Real, SPT> q;
      if (phase_sites<0, SPT>(a, r, b, base + threadIdx.x, step, q.ps)) return;  // frozen plane
      pbase = ((uint32_t)q.ps[0].z + (uint32_t)r.N[2] * q.ps[0].t) * P;
      phase_load<Real, 0, SPT>(lat_base, a, r, q);
      phase_compute<Real, 0, KIND, SPT, false>(lat_base, a, r, q, v0, acc0, th0, pbase);
    }
    __syncthreads();
    for (uint32_t base = 0; base < P; base += step) {
      PhaseIn<Real, SPT> q;
      phase_sites<1, SPT>(a, r, b, base + threadIdx.x, step, q.ps);
      phase_load<Real, 1, SPT, 1>(lat_base, a, r, q);
      phase_load_nu0_lds<Real, SPT>(th0, pbase, q);
      phase_compute<Real, 1, KIND, SPT, MEAS>(lat_base, a, r, q, v, acc1);
    }
  }
  phase_finish<1, KIND, MEAS>(a, L, v, acc0 + acc1);
}

// Reads a completion counter at device scope (sees other XCDs' increments);
// the value is wave-uniform.
__device__ __forceinline__ uint32_t chain_poll(const uint32_t* p) {
  return __builtin_amdgcn_readfirstlane(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Waits (one wavefront) until the completion counter *p reaches target, with
// a sleep between polls; after kChainSpin polls it records a timeout in *err
// and returns.
__device__ __forceinline__ void chain_wait(const uint32_t* p, uint32_t target, uint32_t* err) {
  for (uint32_t n = 0; chain_poll(p) < target; ++n) {
    if (n == kChainSpin) {
      if ((threadIdx.x & 63) == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      return;
    }
    __builtin_amdgcn_s_sleep(8);
  }
}

// One phase-MU workgroup of a fused pass (k_chain): chunk `chunk` of plane
// (z, t).  Each wavefront on its own:
//  1. (MU > 0) reads the counters of the phase-(MU - 1) planes of chain_deps
//     once, before its other loads, so that this read overlaps them;
//  2. issues the loads that no workgroup of this launch writes: the angles at
//     the other parity, the link's own angle, and theta_nu (nu > MU) at this
//     parity, which phase nu only overwrites after this wavefront is counted;
//  3. (MU > 0) waits until those counters are complete, if the first read did
//     not find them so;
//  4. loads theta_nu (nu < MU) at this parity at device scope;
//  5. updates its links and stores them at device scope (plus the
//     measurement row, the accept count);
//  6. waits for its stores to complete (s_waitcnt vmcnt(0)) and adds 1 to
//     the (MU, plane) counter at device scope.
// A (mu, plane) counter is complete when every wavefront of the plane's bpp
// workgroups has added to it (tgt: the launch's value, bpp * blockDim.x / 64
// above the previous launch's).  Device-scope stores and loads make the
// exchange coherent across the XCDs' L2 caches without the whole-cache
// write-back and invalidation of an agent-scope fence.
template <typename Real, int MU, int KIND, int SPT, bool MEAS>
__device__ __forceinline__ void chain_phase(Real* __restrict__ lat_base, const PhaseArgs& a, const ChainArgs& c,
                                            int32_t z, uint32_t t, uint32_t chunk) {
  const RepDesc& r = a.single;
  const uint32_t npl = (uint32_t)r.N[2] * (uint32_t)r.N[3];
  const uint32_t plane = (uint32_t)z + (uint32_t)r.N[2] * t;
  constexpr int DEFM = (1 << MU) - 1;
  PhaseIn<Real, SPT> q;
#pragma unroll
  for (int k = 0; k < SPT; ++k)
    phase_site_zt(r, z, t, chunk * (SPT * blockDim.x) + threadIdx.x + k * blockDim.x, a.parity, q.ps[k]);
  const uint32_t* dep[2] = {nullptr, nullptr};
  uint32_t seen[2] = {0, 0};
  int ndep = 0;
  if (MU > 0) {
    int dz[2], dt[2];
    ndep = chain_deps(MU, z, (int)t, r.N[2], r.N[3], dz, dt);
    for (int i = 0; i < ndep; ++i) {
      dep[i] = c.cnt + (size_t)(MU - 1) * npl + (uint32_t)dz[i] + (uint32_t)r.N[2] * (uint32_t)dt[i];
      seen[i] = chain_poll(dep[i]);
    }
  }
  phase_load<Real, MU, SPT, DEFM>(lat_base, a, r, q);
  if (MU > 0) {
    for (int i = 0; i < ndep; ++i)
      if (seen[i] < c.tgt[MU - 1]) chain_wait(dep[i], c.tgt[MU - 1], c.err);
    phase_load_dep<Real, MU, SPT, DEFM>(lat_base, r, q);
  }
  double v[2] = {0.0, 0.0};
  uint32_t acc = 0;
  phase_compute<Real, MU, KIND, SPT, MEAS, true>(lat_base, a, r, q, v, acc);
  phase_finish<MU, KIND, MEAS>(a, (uint32_t)((MU > 0 ? MU - 1 : 0) * npl + plane) * (uint32_t)r.bpp + chunk, v, acc);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wavefront's stores have completed
  if ((threadIdx.x & 63) == 0)
    __hip_atomic_fetch_add(c.cnt + (size_t)MU * npl + plane, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The phases (parity, 0..3) of one pass of a single lattice in one launch:
// workgroup b runs entry b >> 3 of the schedule list (u1_chain.h) on the
// tile of XCD b & 7.  Results are those of four separate launches, bit for
// bit: each link update does the same arithmetic on the same angles.  MEAS
// makes phases 1-3 measuring.
template <typename Real, int KIND, int SPT, bool MEAS>
__global__ __launch_bounds__(512) void k_chain(Real* __restrict__ lat_base, PhaseArgs a, ChainArgs c) {
  const uint32_t x = blockIdx.x & 7u;
  const uint32_t e = c.sched[blockIdx.x >> 3];
  const uint32_t w = e >> 10, chunk = (e >> 2) & 0xFFu;
  const uint32_t wt = mdiv(w, c.bz_m, c.bz_s);
  const uint32_t t = (x >> c.tz_log2) * c.bt + wt;
  const int32_t z = (int32_t)((x & ((1u << c.tz_log2) - 1u)) * c.bz + (w - wt * c.bz));
  switch (e & 3u) {
    case 0: chain_phase<Real, 0, KIND, SPT, false>(lat_base, a, c, z, t, chunk); break;
    case 1: chain_phase<Real, 1, KIND, SPT, MEAS>(lat_base, a, c, z, t, chunk); break;
    case 2: chain_phase<Real, 2, KIND, SPT, MEAS>(lat_base, a, c, z, t, chunk); break;
    default: chain_phase<Real, 3, KIND, SPT, MEAS>(lat_base, a, c, z, t, chunk); break;
  }
}

template <typename Real>
__global__ __launch_bounds__(256) void k_hotstart(Real* __restrict__ lat, RepDesc r) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= 4 * (int64_t)r.V) return;
  const u1r::U32x4 w = u1r::draw(r.seed, (uint32_t)l, 0u, u1r::ctr2(u1r::STREAM_HOTSTART, 0, 0), r.c3);
  lat[l] = (Real)(u1m::K_PI * (2.0 * u1r::u01_d(w.v[0]) - 1.0));
}

template <typename Real, int KIND, bool BATCHED, int SPT>
hipError_t launch_mu(int mu, Real* lat, const PhaseArgs& a, int nblocks, int block, hipStream_t s) {
  if (a.meas != nullptr) {
    // measuring phases: mu >= 1, no multilevel mask
    if (mu < 1 || a.ml != 0) return hipErrorInvalidValue;
    switch (mu) {
      case 1: k_phase<Real, 1, KIND, BATCHED, SPT, true><<<nblocks, block, 0, s>>>(lat, a); break;
      case 2: k_phase<Real, 2, KIND, BATCHED, SPT, true><<<nblocks, block, 0, s>>>(lat, a); break;
      default: k_phase<Real, 3, KIND, BATCHED, SPT, true><<<nblocks, block, 0, s>>>(lat, a); break;
    }
    return hipGetLastError();
  }
  switch (mu) {
    case 0: k_phase<Real, 0, KIND, BATCHED, SPT><<<nblocks, block, 0, s>>>(lat, a); break;
    case 1: k_phase<Real, 1, KIND, BATCHED, SPT><<<nblocks, block, 0, s>>>(lat, a); break;
    case 2: k_phase<Real, 2, KIND, BATCHED, SPT><<<nblocks, block, 0, s>>>(lat, a); break;
    default: k_phase<Real, 3, KIND, BATCHED, SPT><<<nblocks, block, 0, s>>>(lat, a); break;
  }
  return hipGetLastError();
}

template <typename Real, int KIND, bool BATCHED>
hipError_t launch_spt(int spt, int mu, Real* lat, const PhaseArgs& a, int nblocks, int block, hipStream_t s) {
  if (spt == 2) return launch_mu<Real, KIND, BATCHED, 2>(mu, lat, a, nblocks, block, s);
  if (spt == 4) {
    // four sites per thread only for binary32 (binary64 would exceed the register budget)
    if (sizeof(Real) != sizeof(float)) return hipErrorInvalidValue;
    return launch_mu<Real, KIND, BATCHED, (sizeof(Real) == sizeof(float) ? 4 : 1)>(mu, lat, a, nblocks, block, s);
  }
  return launch_mu<Real, KIND, BATCHED, 1>(mu, lat, a, nblocks, block, s);
}

template <typename Real>
hipError_t launch_prec(int kind, int mu, Real* lat, const PhaseArgs& a, int nblocks, int block, int spt,
                       hipStream_t s) {
  const bool batched = a.reps != nullptr;
  if (kind == KIND_METROPOLIS)
    return batched ? launch_spt<Real, KIND_METROPOLIS, true>(spt, mu, lat, a, nblocks, block, s)
                   : launch_spt<Real, KIND_METROPOLIS, false>(spt, mu, lat, a, nblocks, block, s);
  return batched ? launch_spt<Real, KIND_OVERRELAX, true>(spt, mu, lat, a, nblocks, block, s)
                 : launch_spt<Real, KIND_OVERRELAX, false>(spt, mu, lat, a, nblocks, block, s);
}

// binary32 proposal constants: a (2 wu 2^-32 - 1) with a = pi (independent
// angle) or step (theta + delta); a 2^-31 is exact
PhaseArgs with_f32_constants(const PhaseArgs& a) {
  PhaseArgs b = a;
  const float amp = a.proposal == 0 ? u1m::KF_PI : (float)a.step;
  b.pscale_f = amp * 0x1p-31f;
  b.poff_f = amp;
  b.thm_f = a.proposal == 0 ? 0.0f : 1.0f;
  return b;
}

template <typename Real, int KIND, int SPT, bool MEAS>
void launch_pair_form(int form, Real* lat, const PhaseArgs& a, int nblocks, int block, size_t lds, hipStream_t s) {
  if constexpr (SPT <= 3)
    if (form == PAIR_ONE_STEP) {
      k_pair<Real, KIND, SPT, MEAS, false, PAIR_ONE_STEP><<<nblocks, block, lds, s>>>(lat, a);
      return;
    }
  if constexpr (SPT == 1)
    if (form == PAIR_TWO_STEPS) {
      k_pair<Real, KIND, SPT, MEAS, false, PAIR_TWO_STEPS><<<nblocks, block, lds, s>>>(lat, a);
      return;
    }
  k_pair<Real, KIND, SPT, MEAS, false, PAIR_LOOP><<<nblocks, block, lds, s>>>(lat, a);
}

template <typename Real, int KIND, int SPT>
hipError_t launch_pair_meas(Real* lat, const PhaseArgs& a, int nblocks, int block, size_t lds, hipStream_t s) {
  if (a.reps != nullptr) {  // batched replicas: no measuring pairs; the form per workgroup
    k_pair<Real, KIND, SPT, false, true><<<nblocks, block, lds, s>>>(lat, a);
    return hipGetLastError();
  }
  // a single lattice: its plane size fixes the form (the kernel's own test)
  const int64_t P = (int64_t)a.single.nxh * a.single.N[1], step = (int64_t)SPT * block;
  const bool meas = a.meas != nullptr;
  const int form = SPT <= 3 && P <= step ? PAIR_ONE_STEP
                   : SPT == 1 && P <= 2 * step ? PAIR_TWO_STEPS
                                                         : PAIR_LOOP;
  if (meas)
    launch_pair_form<Real, KIND, SPT, true>(form, lat, a, nblocks, block, lds, s);
  else
    launch_pair_form<Real, KIND, SPT, false>(form, lat, a, nblocks, block, lds, s);
  return hipGetLastError();
}

template <typename Real, int KIND>
hipError_t launch_pair_spt(int spt, Real* lat, const PhaseArgs& a, int nblocks, int block, size_t lds,
                           hipStream_t s) {
  constexpr bool F = sizeof(Real) == sizeof(float);
  if (spt == 2) return launch_pair_meas<Real, KIND, 2>(lat, a, nblocks, block, lds, s);
  if (spt == 3) return launch_pair_meas<Real, KIND, (F ? 3 : 1)>(lat, a, nblocks, block, lds, s);
  if (spt == 4) return launch_pair_meas<Real, KIND, (F ? 4 : 1)>(lat, a, nblocks, block, lds, s);
  return launch_pair_meas<Real, KIND, 1>(lat, a, nblocks, block, lds, s);
}

template <typename Real>
hipError_t launch_pair_prec(int kind, Real* lat, const PhaseArgs& a, int nblocks, int block, int spt, size_t lds,
                            hipStream_t s) {
  if (kind == KIND_METROPOLIS) return launch_pair_spt<Real, KIND_METROPOLIS>(spt, lat, a, nblocks, block, lds, s);
  return launch_pair_spt<Real, KIND_OVERRELAX>(spt, lat, a, nblocks, block, lds, s);
}

template <typename Real, int KIND, int SPT>
hipError_t launch_chain_meas(Real* lat, const PhaseArgs& a, const ChainArgs& c, int nblocks, int block,
                             hipStream_t s) {
  if (a.meas != nullptr)
    k_chain<Real, KIND, SPT, true><<<nblocks, block, 0, s>>>(lat, a, c);
  else
    k_chain<Real, KIND, SPT, false><<<nblocks, block, 0, s>>>(lat, a, c);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_chain_prec(int kind, Real* lat, const PhaseArgs& a, const ChainArgs& c, int nblocks, int block,
                             int spt, hipStream_t s) {
  if (kind == KIND_METROPOLIS)
    return spt == 2 ? launch_chain_meas<Real, KIND_METROPOLIS, 2>(lat, a, c, nblocks, block, s)
                    : launch_chain_meas<Real, KIND_METROPOLIS, 1>(lat, a, c, nblocks, block, s);
  return spt == 2 ? launch_chain_meas<Real, KIND_OVERRELAX, 2>(lat, a, c, nblocks, block, s)
                  : launch_chain_meas<Real, KIND_OVERRELAX, 1>(lat, a, c, nblocks, block, s);
}

}  // namespace

hipError_t launch_phase(int prec, int kind, int mu, void* lat_base, const PhaseArgs& a, int nblocks, int block,
                        int spt, hipStream_t s) {
  if (prec == PREC_F64) return launch_prec<double>(kind, mu, (double*)lat_base, a, nblocks, block, spt, s);
  return launch_prec<float>(kind, mu, (float*)lat_base, with_f32_constants(a), nblocks, block, spt, s);
}

hipError_t launch_pair(int prec, int kind, void* lat_base, const PhaseArgs& a, int nblocks, int block, int spt,
                       int plane_sites, hipStream_t s) {
  // whole wavefronts within the kernel's launch bound; the plane's angles in LDS
  if (!(spt == 1 || spt == 2 || ((spt == 3 || spt == 4) && prec != PREC_F64))) return hipErrorInvalidValue;
  const int elem = prec == PREC_F64 ? 8 : 4;
  const int maxb = pair_max_block(elem, spt);
  if (block < 64 || block > maxb || block % 64 != 0 || plane_sites < 1 || plane_sites > pair_max_plane(elem) ||
      (a.meas != nullptr && (a.ml != 0 || a.reps != nullptr)))
    return hipErrorInvalidValue;
  // single lattice: one chunk per plane, one workgroup per plane, plane_sites
  // its plane (batched: the replica table and block map carry the same layout
  // per replica, plane_sites the largest plane)
  if (a.reps == nullptr && (a.single.bpp != 1 || nblocks != a.single.N[2] * a.single.N[3] ||
                            plane_sites != a.single.nxh * a.single.N[1]))
    return hipErrorInvalidValue;
  const size_t lds = (size_t)plane_sites * (size_t)elem;
  if (prec == PREC_F64) return launch_pair_prec<double>(kind, (double*)lat_base, a, nblocks, block, spt, lds, s);
  return launch_pair_prec<float>(kind, (float*)lat_base, with_f32_constants(a), nblocks, block, spt, lds, s);
}

hipError_t launch_chain(int prec, int kind, void* lat_base, const PhaseArgs& a, const ChainArgs& c, int nblocks,
                        int block, int spt, hipStream_t s) {
  const RepDesc& r = a.single;
  const int64_t P = (int64_t)r.nxh * r.N[1], chunk = (int64_t)block * spt;
  // a single unmasked lattice in a tz x (8 / tz) tiling, its planes cut into
  // bpp chunks of block * spt sites, whole wavefronts within the launch bound
  const int tz = c.tz_log2 <= 3 ? 1 << c.tz_log2 : 0;
  if (!(spt == 1 || spt == 2) || block < 64 || block > 512 || block % 64 != 0 || a.reps != nullptr || a.ml != 0 ||
      !chain_possible(r.N[2], r.N[3], r.bpp, tz) || c.bz * (uint32_t)tz != (uint32_t)r.N[2] ||
      c.bt * (uint32_t)(8 / tz) != (uint32_t)r.N[3] || (int64_t)r.bpp * chunk < P ||
      (int64_t)(r.bpp - 1) * chunk >= P || nblocks <= 0 || nblocks % 8 != 0 || c.sched == nullptr ||
      c.cnt == nullptr || c.err == nullptr)
    return hipErrorInvalidValue;
  if (prec == PREC_F64) return launch_chain_prec<double>(kind, (double*)lat_base, a, c, nblocks, block, spt, s);
  return launch_chain_prec<float>(kind, (float*)lat_base, with_f32_constants(a), c, nblocks, block, spt, s);
}

hipError_t launch_hotstart(int prec, void* lat, const RepDesc& r, int block, hipStream_t s) {
  const int64_t n = 4 * (int64_t)r.V;
  const int nb = (int)((n + block - 1) / block);
  if (prec == PREC_F64)
    k_hotstart<double><<<nb, block, 0, s>>>((double*)lat + r.lat_off, r);
  else
    k_hotstart<float><<<nb, block, 0, s>>>((float*)lat + r.lat_off, r);
  return hipGetLastError();
}

}  // namespace u1d
