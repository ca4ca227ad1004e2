// u1_measure.hip -- plaquette, Polyakov loop, Polyakov-loop volume (plain,
// analytic and stochastic multihit), correlator PP, layout conversions and
// plaquette fields.
//
// Reductions are deterministic: every workgroup writes its partial sums (DPP
// row sums within each wave64, then waves in order through LDS) to a partials
// buffer, and k_fold_rows / k_finish fold the rows in order -- replacing the reference's
// shared-memory fold plus one FP64 atomicAdd per block (reduce_block_1d.h:41-64,
// cudaAtomic.h:147-177), whose summation order is nondeterministic.  All sums
// are FP64 for both lattice precisions.
#include "u1_kernels.h"

namespace u1d {

namespace {

struct C64 {
  double re, im;
};
// complex.h:205-217 ordering
__device__ __forceinline__ C64 cmul(C64 a, C64 b) { return {a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im}; }

template <typename Real>
__device__ __forceinline__ C64 exp_i(Real a) {
  Real s, c;
  Prec<Real>::sincos(a, &s, &c);
  return {(double)c, (double)s};
}

// Temporal-link staple (staple.h:79-102 with mu = TDir): all six staples are
// space-time, weight xi.  Same operation order as the update kernels.
template <typename Real>
__device__ __forceinline__ void staple_t(const Real* __restrict__ lat, const RepDesc& r, const Site& s, int32_t id,
                                         int32_t parity, Real& Sre, Real& Sim) {
  const int32_t same = parity * r.HV + id;
  const int32_t opp = (1 - parity) * r.HV + id;
  const Real* pm = lat + 3 * (int64_t)r.V;
  Real ssr = 0, ssi = 0, str = 0, sti = 0;
#pragma unroll
  for (int nu = 0; nu < 3; ++nu) {
    const Real* pn = lat + (int64_t)nu * r.V;
    const Real a = pn[opp + s.fwd[3]];
    const Real b = pm[opp + s.fwd[nu]];
    const Real c = pn[same];
    const Real d = pn[same + s.fwd[3] + s.bwd[nu]];
    const Real e = pm[opp + s.bwd[nu]];
    const Real f = pn[opp + s.bwd[nu]];
    const Real up = (a - b) - c;
    const Real lo = ((-d) - e) + f;
    Real su, cu, sl, cl;
    Prec<Real>::sincos(up, &su, &cu);
    Prec<Real>::sincos(lo, &sl, &cl);
    str += cu + cl;
    sti += su + sl;
  }
  const Real xi = (Real)r.aniso;
  Sre = ssr / xi + str * xi;
  Sim = ssi / xi + sti * xi;
}

// Temporal link value at even/odd site (id, parity): e^{i theta_t}, the analytic
// multihit mean (lattice_functions.h:92-105) or the stochastic nhit mean.
template <typename Real>
__device__ C64 temporal_value(const Real* __restrict__ lat, const RepDesc& r, int32_t id, int32_t parity, int mode,
                              int proposal, int nhit, double step, uint32_t stream, uint32_t c1, uint32_t c3) {
  const int32_t l = id + parity * r.HV + 3 * r.V;
  if (mode == 0) return exp_i<Real>(lat[l]);
  Site s;
  site_of(r, id, parity, s);
  Real Sre, Sim;
  staple_t<Real>(lat, r, s, id, parity, Sre, Sim);
  if (mode == 1) {
    const double sr = (double)Sre, si = (double)Sim;
    const double alpha = __builtin_sqrt(sr * sr + si * si);
    if (alpha == 0.0) return {0.0, 0.0};
    double i0, i1;
    u1m::bessel_i0i1_d(r.beta * alpha, &i0, &i1);
    const double temp = i1 / (i0 * alpha);
    return {temp * sr, -temp * si};
  }
  // stochastic: nhit Metropolis hits with the staple fixed, mean of e^{i theta}
  Real th = lat[l], s0, c0;
  Prec<Real>::sincos(th, &s0, &c0);
  Real re_old = Sre * c0 - Sim * s0;
  const Real beta = (Real)r.beta;
  double ac = 0.0, as = 0.0;
  auto hit = [&](uint32_t wu, uint32_t wb) {
    const double u = u1r::u01_d(wu);
    const Real tn = (Real)(proposal == 0 ? u1m::K_PI * (2.0 * u - 1.0) : u1m::wrap_d((double)th + step * (2.0 * u - 1.0)));
    Real sn, cn;
    Prec<Real>::sincos(tn, &sn, &cn);
    const Real re_new = Sre * cn - Sim * sn;
    const Real dS = Prec<Real>::exp(beta * (re_new - re_old));
    if (dS > Prec<Real>::u01(wb)) {
      th = tn;
      re_old = re_new;
      s0 = sn;
      c0 = cn;
    }
    ac += (double)c0;
    as += (double)s0;
  };
  // hits 2j and 2j + 1 read words (0, 1) and (2, 3) of one Philox call
  for (int h = 0; h < nhit; h += 2) {
    const u1r::U32x4 w = u1r::draw(r.seed, (uint32_t)l, c1, u1r::ctr2(stream, 0, (uint32_t)(h >> 1)), c3);
    hit(w.v[0], w.v[1]);
    if (h + 1 < nhit) hit(w.v[2], w.v[3]);
  }
  return {ac / (double)nhit, as / (double)nhit};
}

// plaquette.cu:93-118: per site, sum over mu < nu of 1 - e^{iP}, split into
// space-space and space-time; P = ((t_mu(x) + t_nu(x+mu)) - t_mu(x+nu)) - t_nu(x).
//
// Decomposition as the update phases (one chunk of one (z, t) plane per
// workgroup, uniform z/t, SGPR-based loads, XCD-aware order), over both
// parities: the first nblk/2 workgroups of a replica take the even sites, the
// rest the odd sites.  SPT sites per thread.  A site reads its four links and
// the twelve forward neighbours it needs (16 loads; every link is read by four
// sites, mostly from L2).  The three plaquettes of each class are added in
// Real, then once per site into the FP64 accumulators.
template <typename Real, bool BATCHED, int SPT>
__global__ __launch_bounds__(256) void k_plaquette(const Real* __restrict__ lat_base, RepDesc single,
                                                   const RepDesc* reps, const uint16_t* blockmap,
                                                   double* __restrict__ partials) {
  const uint32_t L = xcd_remap(blockIdx.x, gridDim.x);
  const RepDesc& r = BATCHED ? reps[blockmap[L]] : single;
  const Real* lat = lat_base + r.lat_off;
  uint32_t b = L - (uint32_t)r.blk_begin;
  const uint32_t half = (uint32_t)r.nblk >> 1;
  const int32_t parity = b >= half ? 1 : 0;
  if (parity) b -= half;
  PhaseSite ps[SPT];
#pragma unroll
  for (int k = 0; k < SPT; ++k) phase_site(r, b, threadIdx.x + k * blockDim.x, SPT * blockDim.x, parity, ps[k]);
  Real tm[SPT][4], tf[SPT][4][4];  // tf[k][nu][mu] = theta_nu(x + mu), mu < nu or mu > nu
#pragma unroll
  for (int k = 0; k < SPT; ++k) {
    const uint32_t same = (uint32_t)(parity * r.HV) + ps[k].id;
    const uint32_t opp = (uint32_t)((1 - parity) * r.HV) + ps[k].id;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) tm[k][mu] = ld(lat + (size_t)mu * r.V, same);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int nu = 0; nu < 4; ++nu)
        if (nu != mu) tf[k][nu][mu] = ld(lat + (size_t)nu * r.V, opp + ps[k].fwd[mu]);
  }
  double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < SPT; ++k) {
    Real re[2] = {0, 0}, im[2] = {0, 0};
#pragma unroll
    for (int mu = 0; mu < 3; ++mu)
#pragma unroll
      for (int nu = mu + 1; nu < 4; ++nu) {
        const Real p = ((tm[k][mu] + tf[k][nu][mu]) - tf[k][mu][nu]) - tm[k][nu];
        Real sp, cp;
        Prec<Real>::sincos(p, &sp, &cp);
        const int c = nu == 3 ? 1 : 0;
        re[c] += (Real)1 - cp;
        im[c] -= sp;
      }
    if (ps[k].valid) {
      v[0] += (double)re[0];
      v[1] += (double)im[0];
      v[2] += (double)re[1];
      v[3] += (double)im[1];
    }
  }
  double out[4];
  block_sum<4>(v, out);
  // row L (the logical workgroup): a batched fold reads each replica's rows
  // [blk_begin, blk_begin + nblk)
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) partials[4 * (int64_t)L + k] = out[k];
}

// polyakov.cu:60-77: sum of the temporal angles along t (in FP64), then e^{i.}.
template <typename Real>
__global__ __launch_bounds__(256) void k_polyakov(const Real* __restrict__ lat, RepDesc r,
                                                  double* __restrict__ partials) {
  const int32_t idx = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  double v[2] = {0.0, 0.0};
  if (idx < r.SV) {
    const int32_t half = r.SV / 2;
    const int32_t parity = idx >= half ? 1 : 0;
    const int32_t id = idx - parity * half;  // eo-id of (x, t = 0)
    const Real* pt = lat + 3 * (int64_t)r.V;
    double tmp = 0.0;
    for (int t = 0; t < r.N[3]; ++t) tmp += (double)pt[((parity + t) & 1) * r.HV + id + t * half];
    double s, c;
    u1m::sincos_d(tmp, &s, &c);
    v[0] = c;
    v[1] = s;
  }
  double out[2];
  block_sum<2>(v, out);
  if (threadIdx.x == 0) {
    partials[2 * (int64_t)blockIdx.x] = out[0];
    partials[2 * (int64_t)blockIdx.x + 1] = out[1];
  }
}

// Fold nblk partial rows of nv values in workgroup order; one workgroup per value.
__global__ __launch_bounds__(256) void k_finish(const double* __restrict__ partials, int nblk, int nv, double scale,
                                                double* __restrict__ out) {
  const int k = blockIdx.x;
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) v[0] += partials[(int64_t)i * nv + k];
  double o[1];
  block_sum<1>(v, o);
  if (threadIdx.x == 0) out[k] = o[0] * scale;
}

// First level of a long fold: workgroup (g, k) adds rows [g R, g R + R) of
// column k into row g of l1 (same row layout, nv values per row).
__global__ __launch_bounds__(256) void k_fold_rows(const double* __restrict__ partials, int nrows, int nv, int R,
                                                   double* __restrict__ l1) {
  const int g = blockIdx.x, k = blockIdx.y;
  const int r0 = g * R, r1 = min(nrows, r0 + R);
  double v[1] = {0.0};
  for (int i = r0 + (int)threadIdx.x; i < r1; i += blockDim.x) v[0] += partials[(int64_t)i * nv + k];
  double o[1];
  block_sum<1>(v, o);
  if (threadIdx.x == 0) l1[(int64_t)g * nv + k] = o[0];
}

// polyakov.cu:415-440: P(x) = prod_t T(x, t), stored at the lexicographic
// spatial index (indexIdS).
template <typename Real>
__global__ __launch_bounds__(256) void k_poly_volume(const Real* __restrict__ lat, RepDesc r, int mode, int proposal,
                                                     int nhit, double step, uint32_t stream, uint32_t c1, uint32_t c3,
                                                     double* __restrict__ P) {
  const int32_t idx = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (idx >= r.SV) return;
  const int32_t half = r.SV / 2;
  const int32_t parity = idx >= half ? 1 : 0;
  const int32_t id = idx - parity * half;
  C64 res = {1.0, 0.0};
  for (int t = 0; t < r.N[3]; ++t)
    res = cmul(res, temporal_value<Real>(lat, r, id + t * half, (parity + t) & 1, mode, proposal, nhit, step, stream,
                                         c1, c3));
  Site s;
  site_of(r, id, parity, s);
  const int32_t j = s.x[0] + r.N[0] * (s.x[1] + r.N[1] * s.x[2]);
  P[2 * (int64_t)j] = res.re;
  P[2 * (int64_t)j + 1] = res.im;
}

// multilevel.cu:239-258: per-site temporal link at the lexicographic 4D index.
template <typename Real>
__global__ __launch_bounds__(256) void k_temporal_links(const Real* __restrict__ lat, RepDesc r, int mode,
                                                        int proposal, int nhit, double step, uint32_t stream,
                                                        uint32_t c1, uint32_t c3, double* __restrict__ out,
                                                        const uint32_t* c1_dev) {
  const int32_t idx = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (idx >= r.V) return;
  if (c1_dev) c1 = *c1_dev;
  const int32_t parity = idx >= r.HV ? 1 : 0;
  const int32_t id = idx - parity * r.HV;
  const C64 v = temporal_value<Real>(lat, r, id, parity, mode, proposal, nhit, step, stream, c1, c3);
  // lex = 2*id + (x0 & 1); x0 parity = (y + z + t + parity) & 1
  Site s;
  site_of(r, id, parity, s);
  const int64_t j = s.x[0] + (int64_t)r.N[0] * (s.x[1] + (int64_t)r.N[1] * (s.x[2] + (int64_t)r.N[2] * s.x[3]));
  out[2 * j] = v.re;
  out[2 * j + 1] = v.im;
}

// Product over t of a per-site temporal-link field (lexicographic 4D index),
// in t order as polyakov.cu:430-437: P(x) = prod_t T(x, t).
__global__ __launch_bounds__(256) void k_tprod(const double* __restrict__ tl, RepDesc r, double* __restrict__ P) {
  const int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= r.SV) return;
  C64 res = {1.0, 0.0};
  for (int t = 0; t < r.N[3]; ++t) {
    const int64_t k = j + (int64_t)t * r.SV;
    res = cmul(res, C64{tl[2 * k], tl[2 * k + 1]});
  }
  P[2 * (int64_t)j] = res.re;
  P[2 * (int64_t)j + 1] = res.im;
}

// polyakov.cu:514-535: sum over x and the three spatial axes of
// P(x) conj(P(x + R e_dir)), R = 1..rmax; one partial row of 2*rmax per workgroup.
__global__ __launch_bounds__(256) void k_pp(const double* __restrict__ P, RepDesc r, int rmax,
                                            double* __restrict__ partials) {
  const int32_t id = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  const bool act = id < r.SV;
  int32_t x[3] = {0, 0, 0};
  C64 p0 = {0.0, 0.0};
  if (act) {
    x[0] = id % r.N[0];
    x[1] = (id / r.N[0]) % r.N[1];
    x[2] = id / (r.N[0] * r.N[1]);
    p0 = {P[2 * (int64_t)id], P[2 * (int64_t)id + 1]};
  }
  for (int R = 1; R <= rmax; ++R) {
    double v[2] = {0.0, 0.0};
    if (act) {
#pragma unroll
      for (int dir = 0; dir < 3; ++dir) {
        int32_t y[3] = {x[0], x[1], x[2]};
        y[dir] = (y[dir] + R) % r.N[dir];
        const int64_t j = y[0] + (int64_t)r.N[0] * (y[1] + (int64_t)r.N[1] * y[2]);
        const C64 q = {P[2 * j], -P[2 * j + 1]};
        const C64 t = cmul(p0, q);
        v[0] += t.re;
        v[1] += t.im;
      }
    }
    double o[2];
    block_sum<2>(v, o);
    if (threadIdx.x == 0) {
      partials[((int64_t)blockIdx.x * rmax + (R - 1)) * 2] = o[0];
      partials[((int64_t)blockIdx.x * rmax + (R - 1)) * 2 + 1] = o[1];
    }
  }
}

// convertOrder.cu:32-63: even/odd <-> lexicographic, all four planes.
template <typename Real>
__global__ __launch_bounds__(256) void k_convert_layout(const Real* __restrict__ in, Real* __restrict__ out,
                                                        RepDesc r, int eo_to_no) {
  const int32_t idx = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (idx >= r.V) return;
  if (eo_to_no) {
    const int32_t parity = idx >= r.HV ? 1 : 0;
    const int32_t id = idx - parity * r.HV;
    Site s;
    site_of(r, id, parity, s);
    const int32_t n = s.x[0] + r.N[0] * (s.x[1] + r.N[1] * (s.x[2] + r.N[2] * s.x[3]));
#pragma unroll
    for (int d = 0; d < 4; ++d) out[n + (int64_t)d * r.V] = in[idx + (int64_t)d * r.V];
  } else {
    const int32_t x0 = idx % r.N[0];
    const int32_t rest = idx / r.N[0];
    const int32_t x1 = rest % r.N[1];
    const int32_t x2 = (rest / r.N[1]) % r.N[2];
    const int32_t x3 = rest / (r.N[1] * r.N[2]);
    const int32_t parity = (x0 + x1 + x2 + x3) & 1;
    const int32_t e = (idx >> 1) + parity * r.HV;
#pragma unroll
    for (int d = 0; d < 4; ++d) out[e + (int64_t)d * r.V] = in[idx + (int64_t)d * r.V];
  }
}

template <typename A, typename B>
__global__ __launch_bounds__(256) void k_convert_precision(const A* __restrict__ in, B* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (B)in[i];
}

__global__ __launch_bounds__(256) void k_fill_zero(uint32_t* __restrict__ p, int64_t nwords) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) p[i] = 0u;
}

__global__ void k_increment(uint32_t* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = c[0] + 1u;
}

// SixPlaquette order (lattice_functions.h:38-45): (0,3) (1,3) (2,3) (1,2) (2,0) (0,1).
__device__ __forceinline__ void six_pair(int d, int& mu, int& nu) {
  const int m[6] = {0, 1, 2, 1, 2, 0};
  const int n[6] = {3, 3, 3, 2, 0, 1};
  mu = m[d];
  nu = n[d];
}

// Plaquette 1 - e^{iP} at the site with lexicographic coordinates x (even/odd input).
template <typename Real>
__device__ __forceinline__ C64 plaq_eo(const Real* __restrict__ lat, const RepDesc& r, const int32_t* x, int mu,
                                       int nu) {
  const int32_t lexc = x[0] + r.N[0] * (x[1] + r.N[1] * (x[2] + r.N[2] * x[3]));
  const int32_t par = (x[0] + x[1] + x[2] + x[3]) & 1;
  auto eo = [&](int d) {
    int32_t y[4] = {x[0], x[1], x[2], x[3]};
    y[d] = (y[d] + 1) % r.N[d];
    const int32_t ly = y[0] + r.N[0] * (y[1] + r.N[1] * (y[2] + r.N[2] * y[3]));
    return (ly >> 1) + (1 - par) * r.HV;
  };
  const int32_t here = (lexc >> 1) + par * r.HV;
  Real p = lat[(int64_t)mu * r.V + here];
  p += lat[(int64_t)nu * r.V + eo(mu)];
  p -= lat[(int64_t)mu * r.V + eo(nu)];
  p -= lat[(int64_t)nu * r.V + here];
  const C64 e = exp_i<Real>(p);
  return {1.0 - e.re, -e.im};
}

// Same with a lexicographic input lattice (lattice_functions.h:40-46).
template <typename Real>
__device__ __forceinline__ C64 plaq_no(const Real* __restrict__ lat, const RepDesc& r, const int32_t* x, int mu,
                                       int nu) {
  const int32_t here = x[0] + r.N[0] * (x[1] + r.N[1] * (x[2] + r.N[2] * x[3]));
  auto nb = [&](int d) {
    int32_t y[4] = {x[0], x[1], x[2], x[3]};
    y[d] = (y[d] + 1) % r.N[d];
    return y[0] + r.N[0] * (y[1] + r.N[1] * (y[2] + r.N[2] * y[3]));
  };
  Real p = lat[(int64_t)mu * r.V + here];
  p += lat[(int64_t)nu * r.V + nb(mu)];
  p -= lat[(int64_t)mu * r.V + nb(nu)];
  p -= lat[(int64_t)nu * r.V + here];
  const C64 e = exp_i<Real>(p);
  return {1.0 - e.re, -e.im};
}

// plaquette_comps.cu:34-107: six plaquette components per site (spacetime) or
// per spatial site averaged over t; field written at the lexicographic index,
// component-major; partial sums of the six means per workgroup.
template <typename Real>
__global__ __launch_bounds__(256) void k_plaquette_fields(const Real* __restrict__ lat, RepDesc r, int spacetime,
                                                          int eo_order, double* __restrict__ field,
                                                          double* __restrict__ partials) {
  const int32_t size = spacetime ? r.V : r.SV;
  const int32_t idx = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  double v[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = 0.0;
  if (idx < size) {
    int32_t x[4];
    int32_t pos;
    if (eo_order) {
      const int32_t half = size / 2;
      const int32_t parity = idx >= half ? 1 : 0;
      const int32_t id = idx - parity * half;
      Site s;
      site_of(r, id, parity, s);
      for (int k = 0; k < 4; ++k) x[k] = s.x[k];
    } else {
      x[0] = idx % r.N[0];
      x[1] = (idx / r.N[0]) % r.N[1];
      x[2] = (idx / (r.N[0] * r.N[1])) % r.N[2];
      x[3] = spacetime ? idx / (r.N[0] * r.N[1] * r.N[2]) : 0;
    }
    const int nt = spacetime ? 1 : r.N[3];
    for (int t = 0; t < nt; ++t) {
      if (!spacetime) x[3] = t;
      for (int d = 0; d < 6; ++d) {
        int mu, nu;
        six_pair(d, mu, nu);
        const C64 p = eo_order ? plaq_eo<Real>(lat, r, x, mu, nu) : plaq_no<Real>(lat, r, x, mu, nu);
        v[2 * d] += p.re;
        v[2 * d + 1] += p.im;
      }
    }
    if (!spacetime)
      for (int k = 0; k < 12; ++k) v[k] /= (double)r.N[3];
    pos = spacetime ? x[0] + r.N[0] * (x[1] + r.N[1] * (x[2] + r.N[2] * x[3]))
                    : x[0] + r.N[0] * (x[1] + r.N[1] * x[2]);
    for (int d = 0; d < 6; ++d) {
      field[2 * ((int64_t)pos + (int64_t)d * size)] = v[2 * d];
      field[2 * ((int64_t)pos + (int64_t)d * size) + 1] = v[2 * d + 1];
    }
  }
  double o[12];
  block_sum<12>(v, o);
  if (threadIdx.x == 0)
    for (int k = 0; k < 12; ++k) partials[12 * (int64_t)blockIdx.x + k] = o[k];
}

// Per-replica fold of batched plaquette partials, in workgroup order.
__global__ __launch_bounds__(256) void k_batched_finish(const double* __restrict__ partials,
                                                        const RepDesc* __restrict__ reps, double* __restrict__ out) {
  const RepDesc& r = reps[blockIdx.x];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < r.nblk; i += blockDim.x)
    for (int k = 0; k < 4; ++k) v[k] += partials[4 * (int64_t)(r.blk_begin + i) + k];
  double o[4];
  block_sum<4>(v, o);
  if (threadIdx.x == 0) {
    const double n = 1.0 / (3.0 * (double)r.V);
    for (int k = 0; k < 4; ++k) out[4 * blockIdx.x + k] = o[k] * n;
  }
}

// Per-replica fold of the plaquette rows written by the measuring update
// phases of a batched sweep (rows of all segments in order), stored at the
// series row of the current sweep.
__global__ __launch_bounds__(256) void k_ensemble_plaquette_fold(const double* __restrict__ rows, int nseg,
                                                                 int seg_rows, const RepDesc* __restrict__ reps,
                                                                 int nrep, const uint32_t* sweep_dev,
                                                                 const uint32_t* base_dev, int cap,
                                                                 double* __restrict__ series) {
  const RepDesc& r = reps[blockIdx.x];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int sgi = 0; sgi < nseg; ++sgi)
    for (int i = threadIdx.x; i < r.nblk; i += blockDim.x) {
      const double* row = rows + 4 * ((int64_t)sgi * seg_rows + r.blk_begin + i);
      for (int k = 0; k < 4; ++k) v[k] += row[k];
    }
  double o[4];
  block_sum<4>(v, o);
  if (threadIdx.x == 0) {
    const uint32_t idx = *sweep_dev - *base_dev;
    if (idx < (uint32_t)cap) {
      const double n = 1.0 / (3.0 * (double)r.V);
      for (int k = 0; k < 4; ++k) series[4 * ((int64_t)idx * nrep + blockIdx.x) + k] = o[k] * n;
    }
  }
}

inline int nblocks(int64_t n, int block) { return (int)((n + block - 1) / block); }

}  // namespace

namespace {
template <typename Real, bool BATCHED>
hipError_t plaquette_spt(int spt, const Real* lat, const RepDesc& r, const RepDesc* reps, const uint16_t* map,
                         double* partials, int nblk, int block, hipStream_t st) {
  if (spt == 4 && sizeof(Real) == sizeof(float))
    k_plaquette<Real, BATCHED, 4><<<nblk, block, 0, st>>>(lat, r, reps, map, partials);
  else if (spt == 2)
    k_plaquette<Real, BATCHED, 2><<<nblk, block, 0, st>>>(lat, r, reps, map, partials);
  else if (spt == 1)
    k_plaquette<Real, BATCHED, 1><<<nblk, block, 0, st>>>(lat, r, reps, map, partials);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
}  // namespace

hipError_t launch_plaquette(int prec, const void* lat, const RepDesc& r, double* partials, int nblk, int block,
                            int spt, hipStream_t st) {
  if (prec == PREC_F64)
    return plaquette_spt<double, false>(spt, (const double*)lat, r, nullptr, nullptr, partials, nblk, block, st);
  return plaquette_spt<float, false>(spt, (const float*)lat, r, nullptr, nullptr, partials, nblk, block, st);
}

hipError_t launch_plaquette_batched(int prec, const void* lat_base, const RepDesc* reps, const uint16_t* blockmap,
                                    double* partials, int nblk, int block, int spt, hipStream_t st) {
  RepDesc dummy{};
  if (prec == PREC_F64)
    return plaquette_spt<double, true>(spt, (const double*)lat_base, dummy, reps, blockmap, partials, nblk, block,
                                       st);
  return plaquette_spt<float, true>(spt, (const float*)lat_base, dummy, reps, blockmap, partials, nblk, block, st);
}

hipError_t launch_ensemble_plaquette_fold(const double* rows, int nseg, int seg_rows, const RepDesc* reps, int nrep,
                                          const uint32_t* sweep_dev, const uint32_t* base_dev, int cap,
                                          double* series, hipStream_t st) {
  k_ensemble_plaquette_fold<<<nrep, 256, 0, st>>>(rows, nseg, seg_rows, reps, nrep, sweep_dev, base_dev, cap, series);
  return hipGetLastError();
}

hipError_t launch_batched_finish(const double* partials, const RepDesc* reps, int nrep, double* out, hipStream_t st) {
  k_batched_finish<<<nrep, 256, 0, st>>>(partials, reps, out);
  return hipGetLastError();
}

hipError_t launch_polyakov(int prec, const void* lat, const RepDesc& r, double* partials, int nblk, int block,
                           hipStream_t st) {
  if (prec == PREC_F64)
    k_polyakov<double><<<nblk, block, 0, st>>>((const double*)lat, r, partials);
  else
    k_polyakov<float><<<nblk, block, 0, st>>>((const float*)lat, r, partials);
  return hipGetLastError();
}

hipError_t launch_finish_sums(const double* partials, int nblk, int nv, double scale, double* out, double* scratch,
                              hipStream_t st) {
  if (nblk <= kFoldRows) {
    k_finish<<<nv, 256, 0, st>>>(partials, nblk, nv, scale, out);
    return hipGetLastError();
  }
  const int G = fold_groups(nblk);
  k_fold_rows<<<dim3(G, nv), 256, 0, st>>>(partials, nblk, nv, kFoldRows, scratch);
  k_finish<<<nv, 256, 0, st>>>(scratch, G, nv, scale, out);
  return hipGetLastError();
}

hipError_t launch_poly_volume(int prec, const void* lat, const RepDesc& r, int mode, int proposal, int nhit,
                              double step, uint32_t stream, uint32_t c1, uint32_t c3, double* P, int block,
                              hipStream_t st) {
  const int nb = nblocks(r.SV, block);
  if (prec == PREC_F64)
    k_poly_volume<double><<<nb, block, 0, st>>>((const double*)lat, r, mode, proposal, nhit, step, stream, c1, c3, P);
  else
    k_poly_volume<float><<<nb, block, 0, st>>>((const float*)lat, r, mode, proposal, nhit, step, stream, c1, c3, P);
  return hipGetLastError();
}

hipError_t launch_tprod(const double* tl, const RepDesc& r, double* P, hipStream_t st) {
  k_tprod<<<nblocks(r.SV, 256), 256, 0, st>>>(tl, r, P);
  return hipGetLastError();
}

hipError_t launch_temporal_links(int prec, const void* lat, const RepDesc& r, int mode, int proposal, int nhit,
                                 double step, uint32_t stream, uint32_t c1, uint32_t c3, double* out, int block,
                                 hipStream_t st, const uint32_t* c1_dev) {
  const int nb = nblocks(r.V, block);
  if (prec == PREC_F64)
    k_temporal_links<double><<<nb, block, 0, st>>>((const double*)lat, r, mode, proposal, nhit, step, stream, c1, c3,
                                                   out, c1_dev);
  else
    k_temporal_links<float><<<nb, block, 0, st>>>((const float*)lat, r, mode, proposal, nhit, step, stream, c1, c3,
                                                  out, c1_dev);
  return hipGetLastError();
}

hipError_t launch_pp(const double* P, const RepDesc& r, int rmax, double* partials, int nblk, int block,
                     hipStream_t st) {
  k_pp<<<nblk, block, 0, st>>>(P, r, rmax, partials);
  return hipGetLastError();
}

hipError_t launch_convert_layout(int prec, const void* in, void* out, const RepDesc& r, int eo_to_no, int block,
                                 hipStream_t st) {
  const int nb = nblocks(r.V, block);
  if (prec == PREC_F64)
    k_convert_layout<double><<<nb, block, 0, st>>>((const double*)in, (double*)out, r, eo_to_no);
  else
    k_convert_layout<float><<<nb, block, 0, st>>>((const float*)in, (float*)out, r, eo_to_no);
  return hipGetLastError();
}

hipError_t launch_convert_precision(const void* in, int in_prec, void* out, int out_prec, int64_t n, int block,
                                    hipStream_t st) {
  const int nb = nblocks(n, block);
  if (in_prec == PREC_F64 && out_prec == PREC_F32)
    k_convert_precision<double, float><<<nb, block, 0, st>>>((const double*)in, (float*)out, n);
  else if (in_prec == PREC_F32 && out_prec == PREC_F64)
    k_convert_precision<float, double><<<nb, block, 0, st>>>((const float*)in, (double*)out, n);
  else if (in_prec == PREC_F64)
    k_convert_precision<double, double><<<nb, block, 0, st>>>((const double*)in, (double*)out, n);
  else
    k_convert_precision<float, float><<<nb, block, 0, st>>>((const float*)in, (float*)out, n);
  return hipGetLastError();
}

hipError_t launch_plaquette_fields(int prec, const void* lat, const RepDesc& r, int spacetime, int eo_order,
                                   double* field, double* partials, int nblk, int block, hipStream_t st) {
  if (prec == PREC_F64)
    k_plaquette_fields<double><<<nblk, block, 0, st>>>((const double*)lat, r, spacetime, eo_order, field, partials);
  else
    k_plaquette_fields<float><<<nblk, block, 0, st>>>((const float*)lat, r, spacetime, eo_order, field, partials);
  return hipGetLastError();
}

hipError_t launch_fill_zero(void* p, int64_t bytes, hipStream_t st) {
  const int64_t nw = bytes / 4;
  if (nw > 0) k_fill_zero<<<nblocks(nw, 256), 256, 0, st>>>((uint32_t*)p, nw);
  return hipGetLastError();
}

hipError_t launch_increment(uint32_t* counter, hipStream_t st) {
  k_increment<<<1, 64, 0, st>>>(counter);
  return hipGetLastError();
}

}  // namespace u1d
