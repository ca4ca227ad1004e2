// u1_update.hip -- checkerboard Metropolis and overrelaxation phases, hot start.
//
// One workgroup thread per even/odd site of the active parity; one launch per
// (pass, parity, mu) phase, as the reference (update.cu:177-190, 227-231,
// 301-316, 379-394).  Within a phase no link reads a link of the same phase:
// the staple of (x, mu) reads theta_nu (nu != mu) and theta_mu at the other
// parity (staple.h:85-92), so the launch is race free.
//
// FP64 parity contract with the host oracle (oracle/u1_oracle.cpp): the
// staple, the Metropolis weight and the OR reflection below perform the same
// IEEE operations in the same order as the oracle (see the comments there), on
// the same Philox words, with the same u1_math.h functions.  The library is
// compiled with -ffp-contract=off so that no FMA is formed implicitly.
//
// FP32-bulk variant: angles, staples, the Metropolis weight and the OR angle
// in binary32; proposals and the wrap into [-pi, pi) in binary64 before the
// single rounding to binary32.
#include "u1_kernels.h"

namespace u1d {

namespace {

template <typename Real>
struct Staple {
  Real re, im;
};

// Staple of link (x, MU) at even/odd site id of the given parity
// (staple.h:79-102), with the operation order of the oracle.
template <typename Real, int MU>
__device__ __forceinline__ Staple<Real> staple_at(const Real* __restrict__ lat, const RepDesc& r, const Site& s,
                                                  int32_t id, int32_t parity) {
  const int32_t same = parity * r.HV + id;
  const int32_t opp = (1 - parity) * r.HV + id;
  Real ssr = 0, ssi = 0, str = 0, sti = 0;
#pragma unroll
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == MU) continue;
    const Real* pm = lat + (int64_t)MU * r.V;
    const Real* pn = lat + (int64_t)nu * r.V;
    const Real a = pn[opp + s.fwd[MU]];                  // theta_nu(x+mu)
    const Real b = pm[opp + s.fwd[nu]];                  // theta_mu(x+nu)
    const Real c = pn[same];                             // theta_nu(x)
    const Real d = pn[same + s.fwd[MU] + s.bwd[nu]];     // theta_nu(x+mu-nu)
    const Real e = pm[opp + s.bwd[nu]];                  // theta_mu(x-nu)
    const Real f = pn[opp + s.bwd[nu]];                  // theta_nu(x-nu)
    const Real up = (a - b) - c;
    const Real lo = ((-d) - e) + f;
    Real su, cu, sl, cl;
    Prec<Real>::sincos(up, &su, &cu);
    Prec<Real>::sincos(lo, &sl, &cl);
    const Real ar = cu + cl, ai = su + sl;
    if (MU == 3 || nu == 3) {
      str += ar;
      sti += ai;
    } else {
      ssr += ar;
      ssi += ai;
    }
  }
  const Real xi = (Real)r.aniso;
  return {ssr / xi + str * xi, ssi / xi + sti * xi};
}

// Proposal in FP64 (both precisions): independent U[-pi,pi) or wrapped theta+delta.
__device__ __forceinline__ double propose(int proposal, double theta, double step, uint32_t w) {
  const double u = u1r::u01_d(w);
  return proposal == 0 ? u1m::K_PI * (2.0 * u - 1.0) : u1m::wrap_d(theta + step * (2.0 * u - 1.0));
}

template <typename Real>
struct MetroState {
  Real th, s0, c0, re_old;
  int acc;
};

// One Metropolis hit with a fixed staple (lattice_functions.h:68-80 in the
// Re(S e^{i theta}) form).  wu: proposal word, wb: accept-test word.
template <typename Real>
__device__ __forceinline__ void metro_hit(MetroState<Real>& m, Staple<Real> S, Real beta, int proposal, double step,
                                          uint32_t wu, uint32_t wb) {
  const Real tn = (Real)propose(proposal, (double)m.th, step, wu);
  Real sn, cn;
  Prec<Real>::sincos(tn, &sn, &cn);
  const Real re_new = S.re * cn - S.im * sn;
  const Real dS = Prec<Real>::exp(beta * (re_new - m.re_old));
  const Real b = Prec<Real>::u01(wb);
  if (dS > b) {
    m.th = tn;
    m.re_old = re_new;
    m.s0 = sn;
    m.c0 = cn;
    ++m.acc;
  }
}

template <typename Real, int MU, int KIND, bool BATCHED>
__global__ __launch_bounds__(256) void k_phase(Real* __restrict__ lat_base, PhaseArgs a) {
  const RepDesc& r = BATCHED ? a.reps[a.blockmap[blockIdx.x]] : a.single;
  const int32_t id = ((int32_t)blockIdx.x - r.blk_begin) * (int32_t)blockDim.x + (int32_t)threadIdx.x;
  bool active = id < r.HV;
  Real* lat = lat_base + r.lat_off;
  int acc = 0;
  if (active) {
    Site s;
    site_of(r, id, a.parity, s);
    // multilevel.cu:41,141 -- spatial links on slices t % ml == 0 are frozen
    if (MU != 3 && a.ml != 0 && (s.x[3] & (a.ml - 1)) == 0) active = false;
    if (active) {
      const Staple<Real> S = staple_at<Real, MU>(lat, r, s, id, a.parity);
      const int32_t l = id + a.parity * r.HV + MU * r.V;
      Real th = lat[l];
      if (KIND == KIND_METROPOLIS) {
        MetroState<Real> m;
        m.th = th;
        Prec<Real>::sincos(th, &m.s0, &m.c0);
        m.re_old = S.re * m.c0 - S.im * m.s0;
        m.acc = 0;
        const uint32_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
        const Real beta = (Real)r.beta;
        for (int h = 0; h < a.nhit; h += 2) {
          const u1r::U32x4 w = u1r::draw(r.seed, (uint32_t)l, sweep, u1r::ctr2(a.stream, a.pass, (uint32_t)(h >> 1)), r.c3);
          metro_hit<Real>(m, S, beta, a.proposal, a.step, w.v[0], w.v[1]);
          if (h + 1 < a.nhit) metro_hit<Real>(m, S, beta, a.proposal, a.step, w.v[2], w.v[3]);
        }
        th = m.th;
        acc = m.acc;
      } else {
        // lattice_functions.h:83-89: theta' = -theta - 2 arg(S), wrapped in FP64
        const Real t2 = Prec<Real>::atan2(S.im, S.re);
        th = (Real)u1m::wrap_d(-(double)th - 2.0 * (double)t2);
      }
      lat[l] = th;
    }
  }
  if (KIND == KIND_METROPOLIS && a.accepts) {
    const uint64_t tot = wave_sum_u64((uint64_t)acc);
    if ((threadIdx.x & 63) == 0 && tot) atomicAdd(a.accepts, (unsigned long long)tot);
  }
}

template <typename Real>
__global__ __launch_bounds__(256) void k_hotstart(Real* __restrict__ lat, RepDesc r) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= 4 * (int64_t)r.V) return;
  const u1r::U32x4 w = u1r::draw(r.seed, (uint32_t)l, 0u, u1r::ctr2(u1r::STREAM_HOTSTART, 0, 0), r.c3);
  lat[l] = (Real)(u1m::K_PI * (2.0 * u1r::u01_d(w.v[0]) - 1.0));
}

template <typename Real, int KIND, bool BATCHED>
hipError_t launch_mu(int mu, Real* lat, const PhaseArgs& a, int nblocks, int block, hipStream_t st) {
  switch (mu) {
    case 0: k_phase<Real, 0, KIND, BATCHED><<<nblocks, block, 0, st>>>(lat, a); break;
    case 1: k_phase<Real, 1, KIND, BATCHED><<<nblocks, block, 0, st>>>(lat, a); break;
    case 2: k_phase<Real, 2, KIND, BATCHED><<<nblocks, block, 0, st>>>(lat, a); break;
    default: k_phase<Real, 3, KIND, BATCHED><<<nblocks, block, 0, st>>>(lat, a); break;
  }
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_prec(int kind, int mu, Real* lat, const PhaseArgs& a, int nblocks, int block, hipStream_t st) {
  const bool batched = a.reps != nullptr;
  if (kind == KIND_METROPOLIS)
    return batched ? launch_mu<Real, KIND_METROPOLIS, true>(mu, lat, a, nblocks, block, st)
                   : launch_mu<Real, KIND_METROPOLIS, false>(mu, lat, a, nblocks, block, st);
  return batched ? launch_mu<Real, KIND_OVERRELAX, true>(mu, lat, a, nblocks, block, st)
                 : launch_mu<Real, KIND_OVERRELAX, false>(mu, lat, a, nblocks, block, st);
}

}  // namespace

hipError_t launch_phase(int prec, int kind, int mu, void* lat_base, const PhaseArgs& a, int nblocks, int block,
                        hipStream_t st) {
  if (prec == PREC_F64) return launch_prec<double>(kind, mu, (double*)lat_base, a, nblocks, block, st);
  return launch_prec<float>(kind, mu, (float*)lat_base, a, nblocks, block, st);
}

hipError_t launch_hotstart(int prec, void* lat, const RepDesc& r, int block, hipStream_t st) {
  const int64_t n = 4 * (int64_t)r.V;
  const int nb = (int)((n + block - 1) / block);
  if (prec == PREC_F64)
    k_hotstart<double><<<nb, block, 0, st>>>((double*)lat + r.lat_off, r);
  else
    k_hotstart<float><<<nb, block, 0, st>>>((float*)lat + r.lat_off, r);
  return hipGetLastError();
}

}  // namespace u1d
