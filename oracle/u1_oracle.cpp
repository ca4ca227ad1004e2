// u1_oracle.cpp -- FP64 host oracle for the LatticeU1 hot path.
//
// TEST INFRASTRUCTURE ONLY (see u1_oracle.h).  A plain, unoptimised
// restatement of the reference's algorithms in C++17 + OpenMP.  Every function
// cites the reference file:line it follows (paths relative to the reference
// tree: include/ and src/).  Randomness: Philox4x32-10 with the counter layout
// of u1_philox.h (replacing cuRAND MRG32k3a, random.h:31-58).  Elementary
// functions: u1_math.h, shared with the GPU path so that FP64 dynamics agree
// bit for bit.  Build with -ffp-contract=off (see oracle/Makefile).
//
// Deliberate, documented deltas from the reference (all shared with the GPU):
//  * angles live in [-pi, pi) and are wrapped in FP64 (reference: [0, 2pi)
//    via fmod, lattice_functions.h:87; hot start U(0,2pi], update.cu:38);
//  * the Metropolis weight ratio uses Re(S e^{i theta}) = Sre cos(theta) -
//    Sim sin(theta), the identity behind the reference's
//    |S| cos(theta + atan2(Sim, Sre)) (lattice_functions.h:68-80);
//  * the proposal is parametrised: independent U[-pi,pi) with 1 hit (the
//    reference, update.cu:181-182) or theta + step*U[-1,1) with nhit hits
//    reusing the fixed staple;
//  * analytic multihit returns 0 for a vanishing staple (the reference
//    evaluates 0/0 there, lattice_functions.h:101).
#include "u1_oracle.h"

#include <omp.h>

#include <cstring>
#include <vector>

#include "../latticeu1_amd/csrc/u1_math.h"
#include "../latticeu1_amd/csrc/u1_philox.h"

#ifdef ORC_USE_LIBM
// Variant used only to pin the shared elementary functions: same algorithms,
// glibc's sin/cos/exp/atan2 instead of u1_math.h.
#include <cmath>
static inline void SINCOS(double x, double* s, double* c) { *s = std::sin(x); *c = std::cos(x); }
static inline double EXP(double x) { return std::exp(x); }
static inline double ATAN2(double y, double x) { return std::atan2(y, x); }
#else
static inline void SINCOS(double x, double* s, double* c) { u1m::sincos_d(x, s, c); }
static inline double EXP(double x) { return u1m::exp_d(x); }
static inline double ATAN2(double y, double x) { return u1m::atan2_d(y, x); }
#endif

namespace {

constexpr int TDIR = 3;  // 4D: time direction (parameters.cu:151)

// Geometry restated from parameters.cu:146-186 and index.h:14-153.
struct Geo {
  int N[4];
  int V, HV, SV;
  double beta, aniso;
  explicit Geo(const orc_geom* g) {
    for (int i = 0; i < 4; ++i) N[i] = g->dims[i];
    V = N[0] * N[1] * N[2] * N[3];
    HV = V / 2;
    SV = N[0] * N[1] * N[2];
    beta = g->beta;
    aniso = g->aniso;
  }
  // index.h:114-122 (indexEO)
  void eo_coords(int id, int parity, int x[4]) const {
    int za = id / (N[0] / 2);
    int zb = za / N[1];
    x[1] = za - zb * N[1];
    x[3] = zb / N[2];
    x[2] = zb - x[3] * N[2];
    int xodd = (x[1] + x[2] + x[3] + parity) & 1;
    x[0] = (2 * id + xodd) - za * N[0];
  }
  // index.h:25-30 (indexId(x)), lexicographic site index
  int lex(const int x[4]) const { return x[0] + N[0] * (x[1] + N[1] * (x[2] + N[2] * x[3])); }
  int lex3(const int x[4]) const { return x[0] + N[0] * (x[1] + N[1] * x[2]); }
  static int parity_of(const int x[4]) { return (x[0] + x[1] + x[2] + x[3]) & 1; }
  // index.h:124-132 (indexId(x, dir)): EO position of link (x, dir)
  int link(const int x[4], int dir) const { return (lex(x) >> 1) + parity_of(x) * HV + dir * V; }
  // index.h:134-153 (indexEO_neg): EO site position (eo_id + parity*HV) of x + lmu*mu + lnu*nu
  int shift(const int x0[4], int mu, int lmu, int nu = 0, int lnu = 0) const {
    int x[4] = {x0[0], x0[1], x0[2], x0[3]};
    x[mu] = (x[mu] + lmu + N[mu]) % N[mu];
    x[nu] = (x[nu] + lnu + N[nu]) % N[nu];
    return (lex(x) >> 1) + parity_of(x) * HV;
  }
};

struct Cplx {
  double re, im;
};
// complex.h:205-217: (a.x b.x - a.y b.y, a.y b.x + a.x b.y)
inline Cplx cmul(Cplx a, Cplx b) { return {a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im}; }
inline Cplx conj(Cplx a) { return {a.re, -a.im}; }
inline Cplx exp_ir(double a) {  // index.h:14-16
  Cplx z;
  SINCOS(a, &z.im, &z.re);
  return z;
}

// staple.h:79-102 (Staple): weighted sum of the six staples of link (x, mu).
// Order of operations is part of the parity contract with the GPU kernels:
//   up = (theta_nu(x+mu) - theta_mu(x+nu)) - theta_nu(x)
//   lo = ((-theta_nu(x+mu-nu)) - theta_mu(x-nu)) + theta_nu(x-nu)
//   S  = SS / xi + ST * xi, SS and ST accumulated in nu = 0..3 order.
Cplx staple(const Geo& g, const double* lat, int id, int parity, int mu) {
  int x[4];
  g.eo_coords(id, parity, x);
  double ssr = 0., ssi = 0., str = 0., sti = 0.;
  const int idmu1 = g.shift(x, mu, 1);
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    double up = lat[idmu1 + g.V * nu];
    up -= lat[g.shift(x, nu, 1) + g.V * mu];
    up -= lat[id + parity * g.HV + nu * g.V];
    const int idnm = g.shift(x, nu, -1);
    double lo = -lat[g.shift(x, mu, 1, nu, -1) + g.V * nu];
    lo -= lat[idnm + g.V * mu];
    lo += lat[idnm + g.V * nu];
    double su, cu, sl, cl;
    SINCOS(up, &su, &cu);
    SINCOS(lo, &sl, &cl);
    const double ar = cu + cl, ai = su + sl;
    if (mu == TDIR || nu == TDIR) {
      str += ar;
      sti += ai;
    } else {
      ssr += ar;
      ssi += ai;
    }
  }
  return {ssr / g.aniso + str * g.aniso, ssi / g.aniso + sti * g.aniso};
}

// Metropolis on one link with a fixed staple (lattice_functions.h:68-80,
// update.cu:177-190 generalised to nhit hits).  Returns accepted hits.
// Hit h uses Philox call h/2 and words 2*(h%2), 2*(h%2)+1.
inline int metropolis_link(double& theta, Cplx S, double beta, const orc_metro& m, uint64_t seed,
                           uint32_t c0, uint32_t c1, uint32_t stream, uint32_t pass, uint32_t c3,
                           double* acc_cos, double* acc_sin) {
  double s0, c0v;
  SINCOS(theta, &s0, &c0v);
  double re_old = S.re * c0v - S.im * s0;
  int acc = 0;
  u1r::U32x4 w{};
  for (int h = 0; h < m.nhit; ++h) {
    if ((h & 1) == 0) w = u1r::draw(seed, c0, c1, u1r::ctr2(stream, pass, (uint32_t)(h >> 1)), c3);
    const double u = u1r::u01_d(w.v[2 * (h & 1)]);
    const double b = u1r::u01_d(w.v[2 * (h & 1) + 1]);
    double tn;
    if (m.proposal == 0)
      tn = u1m::K_PI * (2.0 * u - 1.0);
    else
      tn = u1m::wrap_d(theta + m.step * (2.0 * u - 1.0));
    double sn, cn;
    SINCOS(tn, &sn, &cn);
    const double re_new = S.re * cn - S.im * sn;
    const double dS = EXP(beta * (re_new - re_old));
    if (dS > b) {
      theta = tn;
      re_old = re_new;
      s0 = sn;
      c0v = cn;
      ++acc;
    }
    if (acc_cos) {
      *acc_cos += c0v;
      *acc_sin += s0;
    }
  }
  return acc;
}

// Overrelaxation on one link (lattice_functions.h:83-89): reflection of theta
// about -arg(S), which keeps Re(S e^{i theta}) fixed.
inline double overrelax_link(double theta, Cplx S) {
  const double t2 = ATAN2(S.im, S.re);
  return u1m::wrap_d(-theta - 2.0 * t2);
}

// Multilevel mask (multilevel.cu:41,141): spatial links on time slices
// t % ml == 0 are frozen; temporal links always update.
inline bool updates(const Geo& g, int id, int parity, int mu, int ml) {
  if (ml == 0 || mu == TDIR) return true;
  int x[4];
  g.eo_coords(id, parity, x);
  return (x[TDIR] % ml) != 0;
}

void metropolis_phase(const Geo& g, double* lat, int parity, int mu, const orc_metro& m, uint64_t seed,
                      uint32_t stream, uint32_t c1, uint32_t pass, uint32_t c3, int ml, uint64_t* accepts) {
  uint64_t acc = 0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
  for (int id = 0; id < g.HV; ++id) {
    if (!updates(g, id, parity, mu, ml)) continue;
    const int l = id + parity * g.HV + mu * g.V;
    const Cplx S = staple(g, lat, id, parity, mu);
    double th = lat[l];
    acc += metropolis_link(th, S, g.beta, m, seed, (uint32_t)l, c1, stream, pass, c3, nullptr, nullptr);
    lat[l] = th;
  }
  if (accepts) *accepts += acc;
}

void overrelax_phase(const Geo& g, double* lat, int parity, int mu, int ml) {
#pragma omp parallel for schedule(static)
  for (int id = 0; id < g.HV; ++id) {
    if (!updates(g, id, parity, mu, ml)) continue;
    const int l = id + parity * g.HV + mu * g.V;
    lat[l] = overrelax_link(lat[l], staple(g, lat, id, parity, mu));
  }
}

// update.cu:449-456 with the Metropolis (update.cu:301-316) and OR
// (update.cu:379-394) loop orders: pass, parity, mu.
void update(const Geo& g, double* lat, int metrop, int ovrn, const orc_metro& m, uint64_t seed,
            uint32_t stream, uint32_t c1, uint32_t c3, int ml, uint64_t* accepts) {
  for (int p = 0; p < metrop; ++p)
    for (int parity = 0; parity < 2; ++parity)
      for (int mu = 0; mu < 4; ++mu)
        metropolis_phase(g, lat, parity, mu, m, seed, stream, c1, (uint32_t)p, c3, ml, accepts);
  for (int o = 0; o < ovrn; ++o)
    for (int parity = 0; parity < 2; ++parity)
      for (int mu = 0; mu < 4; ++mu) overrelax_phase(g, lat, parity, mu, ml);
}

// lattice_functions.h:92-105 (MultiHit): analytic infinite-hit mean of e^{i theta}.
Cplx multihit_analytic(const Geo& g, const double* lat, int id, int parity) {
  const Cplx S = staple(g, lat, id, parity, TDIR);
  const double alpha = __builtin_sqrt(S.re * S.re + S.im * S.im);
  if (alpha == 0.0) return {0.0, 0.0};
  double i0, i1;
  u1m::bessel_i0i1_d(g.beta * alpha, &i0, &i1);
  const double temp = i1 / (i0 * alpha);
  return {temp * S.re, -temp * S.im};
}

// Stochastic multihit: mean of e^{i theta} over nhit Metropolis hits with the
// staple held fixed (the finite-hit estimator whose limit is multihit_analytic).
Cplx multihit_stochastic(const Geo& g, const double* lat, int id, int parity, const orc_metro& mh,
                         uint64_t seed, uint32_t stream, uint32_t c1, uint32_t c3) {
  const Cplx S = staple(g, lat, id, parity, TDIR);
  const int l = id + parity * g.HV + TDIR * g.V;
  double th = lat[l];
  double ac = 0.0, as = 0.0;
  metropolis_link(th, S, g.beta, mh, seed, (uint32_t)l, c1, stream, 0u, c3, &ac, &as);
  return {ac / (double)mh.nhit, as / (double)mh.nhit};
}

Cplx temporal_link(const Geo& g, const double* lat, int id, int parity, int mode, const orc_metro* mh,
                   uint64_t seed, uint32_t stream, uint32_t c1, uint32_t c3) {
  if (mode == 1) return multihit_analytic(g, lat, id, parity);
  if (mode == 2) return multihit_stochastic(g, lat, id, parity, *mh, seed, stream, c1, c3);
  return exp_ir(lat[id + parity * g.HV + TDIR * g.V]);
}

// PlaquetteF (lattice_functions.h:29-36): 1 - e^{iP},
// P = ((theta_mu(x) + theta_nu(x+mu)) - theta_mu(x+nu)) - theta_nu(x).
inline Cplx plaq_term(const Geo& g, const double* lat, const int x[4], int id, int parity, int mu, int nu) {
  double p = lat[id + parity * g.HV + mu * g.V];
  p += lat[g.shift(x, mu, 1) + g.V * nu];
  p -= lat[g.shift(x, nu, 1) + g.V * mu];
  p -= lat[id + parity * g.HV + nu * g.V];
  const Cplx e = exp_ir(p);
  return {1.0 - e.re, -e.im};
}

// Lexicographic-layout variant (lattice_functions.h:40-46, indexNO_neg).
inline Cplx plaq_term_no(const Geo& g, const double* lat, int idx, int mu, int nu) {
  int x[4];
  x[3] = idx / (g.N[0] * g.N[1] * g.N[2]);
  x[2] = (idx / (g.N[0] * g.N[1])) % g.N[2];
  x[1] = (idx / g.N[0]) % g.N[1];
  x[0] = idx % g.N[0];
  auto sh = [&](int d) {
    int y[4] = {x[0], x[1], x[2], x[3]};
    y[d] = (y[d] + 1) % g.N[d];
    return g.lex(y);
  };
  double p = lat[idx + mu * g.V];
  p += lat[sh(mu) + g.V * nu];
  p -= lat[sh(nu) + g.V * mu];
  p -= lat[idx + nu * g.V];
  const Cplx e = exp_ir(p);
  return {1.0 - e.re, -e.im};
}

// SixPlaquette order (lattice_functions.h:38-45).
constexpr int SIX_MU[6] = {0, 1, 2, 1, 2, 0};
constexpr int SIX_NU[6] = {3, 3, 3, 2, 0, 1};

}  // namespace

extern "C" {

void orc_sincos(const double* x, double* s, double* c, int64_t n) {
  for (int64_t i = 0; i < n; ++i) SINCOS(x[i], &s[i], &c[i]);
}
void orc_exp(const double* x, double* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = EXP(x[i]);
}
void orc_atan2(const double* y, const double* x, double* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = ATAN2(y[i], x[i]);
}
void orc_bessel_i0i1(const double* x, double* i0, double* i1, int64_t n) {
  for (int64_t i = 0; i < n; ++i) u1m::bessel_i0i1_d(x[i], &i0[i], &i1[i]);
}
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    u1r::U32x4 c;
    for (int j = 0; j < 4; ++j) c.v[j] = ctr[4 * i + j];
    const u1r::U32x4 o = u1r::philox4x32_10(c, key[2 * i], key[2 * i + 1]);
    for (int j = 0; j < 4; ++j) out[4 * i + j] = o.v[j];
  }
}
int orc_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }

void orc_eo_coords(const orc_geom* gg, int32_t id, int32_t parity, int32_t x[4]) {
  Geo g(gg);
  int xx[4];
  g.eo_coords(id, parity, xx);
  for (int i = 0; i < 4; ++i) x[i] = xx[i];
}

int32_t orc_eo_neighbor(const orc_geom* gg, int32_t id, int32_t parity, int32_t mu, int32_t lmu, int32_t nu,
                        int32_t lnu) {
  Geo g(gg);
  int x[4];
  g.eo_coords(id, parity, x);
  return g.shift(x, mu, lmu, nu, lnu);
}

void orc_staple(const orc_geom* gg, const double* lat, int32_t id, int32_t parity, int32_t mu, double out[2]) {
  Geo g(gg);
  const Cplx S = staple(g, lat, id, parity, mu);
  out[0] = S.re;
  out[1] = S.im;
}

// Hot start (update.cu:30-41): each link gets an independent uniform angle,
// here U[-pi, pi) from Philox stream HOTSTART with c0 = link index.
void orc_hotstart(const orc_geom* gg, double* lat, uint64_t seed, uint32_t replica) {
  Geo g(gg);
  const int n = 4 * g.V;
#pragma omp parallel for schedule(static)
  for (int l = 0; l < n; ++l) {
    const u1r::U32x4 w = u1r::draw(seed, (uint32_t)l, 0u, u1r::ctr2(u1r::STREAM_HOTSTART, 0, 0), replica);
    lat[l] = u1m::K_PI * (2.0 * u1r::u01_d(w.v[0]) - 1.0);
  }
}

void orc_update(const orc_geom* gg, double* lat, int32_t metrop, int32_t ovrn, const orc_metro* m,
                uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t c3, int32_t mask_ml,
                uint64_t* accepts) {
  Geo g(gg);
  update(g, lat, metrop, ovrn, *m, seed, stream, sweep, c3, mask_ml, accepts);
}

void orc_metropolis_phase(const orc_geom* gg, double* lat, int32_t parity, int32_t mu, const orc_metro* m,
                          uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t pass, uint32_t c3,
                          uint64_t* accepts) {
  Geo g(gg);
  metropolis_phase(g, lat, parity, mu, *m, seed, stream, sweep, pass, c3, 0, accepts);
}

void orc_overrelax_phase(const orc_geom* gg, double* lat, int32_t parity, int32_t mu) {
  Geo g(gg);
  overrelax_phase(g, lat, parity, mu, 0);
}

// plaquette.cu:93-118 (kernel_plaquette) with the norms of plaquette.cu:156-165.
void orc_plaquette(const orc_geom* gg, const double* lat, double out[4]) {
  Geo g(gg);
  double ssr = 0, ssi = 0, str = 0, sti = 0;
  for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(static) reduction(+ : ssr, ssi, str, sti)
    for (int id = 0; id < g.HV; ++id) {
      int x[4];
      g.eo_coords(id, parity, x);
      for (int mu = 0; mu < 3; ++mu)
        for (int nu = mu + 1; nu < 4; ++nu) {
          const Cplx t = plaq_term(g, lat, x, id, parity, mu, nu);
          if (mu == TDIR || nu == TDIR) {
            str += t.re;
            sti += t.im;
          } else {
            ssr += t.re;
            ssi += t.im;
          }
        }
    }
  }
  const double nss = 1.0 / (double(g.V) * 3.0), nst = 1.0 / (double(g.V) * 3.0);
  out[0] = ssr * nss;
  out[1] = ssi * nss;
  out[2] = str * nst;
  out[3] = sti * nst;
}

// plaquette.cu:32-57 (host plaquette): <cos P>, <sin P> over 6V plaquettes.
void orc_plaquette_host(const orc_geom* gg, const double* lat, double out[2]) {
  Geo g(gg);
  double pc = 0, ps = 0;
  for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(static) reduction(+ : pc, ps)
    for (int id = 0; id < g.HV; ++id) {
      int x[4];
      g.eo_coords(id, parity, x);
      for (int mu = 0; mu < 3; ++mu)
        for (int nu = mu + 1; nu < 4; ++nu) {
          const Cplx t = plaq_term(g, lat, x, id, parity, mu, nu);
          pc += 1.0 - t.re;
          ps += -t.im;
        }
    }
  }
  const double norm = 1.0 / (double(g.V) * 6.0);
  out[0] = pc * norm;
  out[1] = ps * norm;
}

// polyakov.cu:60-77 (kernel_polyakov) / polyakov.cu:33-49: sum the temporal
// angles along t, then one e^{i.}; mean over spatial sites.
void orc_polyakov(const orc_geom* gg, const double* lat, double out[2]) {
  Geo g(gg);
  double pr = 0, pi = 0;
  for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(static) reduction(+ : pr, pi)
    for (int id = 0; id < g.SV / 2; ++id) {
      int x[4];
      g.eo_coords(id, parity, x);
      double tmp = 0.;
      for (x[TDIR] = 0; x[TDIR] < g.N[TDIR]; ++x[TDIR]) tmp += lat[g.link(x, TDIR)];
      const Cplx e = exp_ir(tmp);
      pr += e.re;
      pi += e.im;
    }
  }
  out[0] = pr / double(g.SV);
  out[1] = pi / double(g.SV);
}

// polyakov.cu:415-440 (kernel_polyakov_volume<multihit>).
void orc_poly_volume(const orc_geom* gg, const double* lat, int32_t mode, const orc_metro* mh, uint64_t seed,
                     uint32_t stream, uint32_t c1, uint32_t c3, double* P) {
  Geo g(gg);
#pragma omp parallel for schedule(static)
  for (int idx = 0; idx < g.SV; ++idx) {
    int parity = 0, id = idx;
    if (id >= g.SV / 2) {
      parity = 1;
      id -= g.SV / 2;
    }
    int x[4];
    g.eo_coords(id, parity, x);
    Cplx res = {1.0, 0.0};
    for (x[TDIR] = 0; x[TDIR] < g.N[TDIR]; ++x[TDIR]) {
      const int pos = g.lex(x) >> 1, odd = Geo::parity_of(x);
      res = cmul(res, temporal_link(g, lat, pos, odd, mode, mh, seed, stream, c1, c3));
    }
    const int s = g.lex3(x);
    P[2 * s] = res.re;
    P[2 * s + 1] = res.im;
  }
}

// polyakov.cu:514-535 (kernel_PP) and the norm of PP (polyakov.cu:560).
void orc_pp(const orc_geom* gg, const double* P, int32_t rmax, double* out) {
  Geo g(gg);
  for (int r = 1; r <= rmax; ++r) {
    double sr = 0, si = 0;
#pragma omp parallel for schedule(static) reduction(+ : sr, si)
    for (int id = 0; id < g.SV; ++id) {
      int x[3] = {id % g.N[0], (id / g.N[0]) % g.N[1], (id / (g.N[0] * g.N[1])) % g.N[2]};
      const Cplx p0 = {P[2 * id], P[2 * id + 1]};
      for (int dir = 0; dir < 3; ++dir) {
        int y[3] = {x[0], x[1], x[2]};
        y[dir] = (y[dir] + r) % g.N[dir];
        const int j = y[0] + g.N[0] * (y[1] + g.N[1] * y[2]);
        const Cplx t = cmul(p0, conj({P[2 * j], P[2 * j + 1]}));
        sr += t.re;
        si += t.im;
      }
    }
    const double norm = 1.0 / (double(g.SV) * 3.0);
    out[2 * (r - 1)] = sr * norm;
    out[2 * (r - 1) + 1] = si * norm;
  }
}

// multilevel.cu:239-258 (kernel_l2_multilevel_0): per-site temporal link,
// stored at the lexicographic 4D index.
void orc_temporal_links(const orc_geom* gg, const double* lat, int32_t mode, const orc_metro* mh, uint64_t seed,
                        uint32_t stream, uint32_t c1, uint32_t c3, double* out) {
  Geo g(gg);
#pragma omp parallel for schedule(static)
  for (int idx = 0; idx < g.V; ++idx) {
    int parity = 0, id = idx;
    if (id >= g.HV) {
      parity = 1;
      id -= g.HV;
    }
    int x[4];
    g.eo_coords(id, parity, x);
    const Cplx v = temporal_link(g, lat, id, parity, mode, mh, seed, stream, c1, c3);
    const int s = g.lex(x);
    out[2 * s] = v.re;
    out[2 * s + 1] = v.im;
  }
}

// multilevel.cu:639-732 (MultiLevel) with its kernels: masked updates
// (:34-51, :134-144), T2 accumulation (:331-360), T2 -> T4 (:439-456),
// final correlator (:532-553) and norms (:696-701, :581).
void orc_multilevel(const orc_geom* gg, const double* lat_in, int32_t n4, int32_t k4, int32_t n2, int32_t k2,
                    int32_t metrop, int32_t ovrn, const orc_metro* m, int32_t mh_mode, const orc_metro* mh,
                    uint64_t seed, uint32_t call_id, double* out, double* ppspace) {
  Geo g(gg);
  const int R = g.N[0] / 2, SV = g.SV, Nt = g.N[TDIR];
  const int nl2 = Nt / 2, nl4 = Nt / 4;
  const size_t sz = (size_t)SV * R * 3;  // one layer of T2/T4
  std::vector<double> lat(lat_in, lat_in + 4 * (size_t)g.V);
  std::vector<Cplx> l2(sz * nl2), l4(sz * nl4, Cplx{0, 0}), tl(g.V);
  const double l2norm = 1.0 / double(n2), l4norm = 1.0 / double(n4);
  std::vector<double> tmp(2 * (size_t)g.V);
  for (int i = 0; i < n4; ++i) {
    for (int j = 0; j < k4; ++j)
      update(g, lat.data(), metrop, ovrn, *m, seed, u1r::STREAM_ML_L4, (uint32_t)(i * k4 + j), call_id, 4, nullptr);
    std::fill(l2.begin(), l2.end(), Cplx{0, 0});
    for (int k = 0; k < n2; ++k) {
      for (int l = 0; l < k2; ++l)
        update(g, lat.data(), metrop, ovrn, *m, seed, u1r::STREAM_ML_L2, (uint32_t)((i * n2 + k) * k2 + l),
               call_id, 2, nullptr);
      orc_temporal_links(gg, lat.data(), mh_mode, mh, seed, u1r::STREAM_ML_MULTIHIT, (uint32_t)(i * n2 + k),
                         call_id, tmp.data());
      for (int s = 0; s < g.V; ++s) tl[s] = {tmp[2 * s], tmp[2 * s + 1]};
      // kernel_l2_multilevel_1
#pragma omp parallel for schedule(static)
      for (int id = 0; id < SV; ++id) {
        int x[4] = {id % g.N[0], (id / g.N[0]) % g.N[1], (id / (g.N[0] * g.N[1])) % g.N[2], 0};
        for (int r = 1; r <= R; ++r)
          for (int dir = 0; dir < 3; ++dir) {
            int layer = 0;
            for (int t = 0; t < Nt; t += 2) {
              Cplx p0 = {1, 0}, p1 = {1, 0};
              for (x[3] = t; x[3] < t + 2; ++x[3]) {
                p0 = cmul(p0, tl[g.lex(x)]);
                int y[4] = {x[0], x[1], x[2], x[3]};
                y[dir] = (y[dir] + r) % g.N[dir];
                p1 = cmul(p1, conj(tl[g.lex(y)]));
              }
              const Cplx pl = cmul(p0, p1);
              const size_t pos = id + (size_t)SV * (r - 1) + (size_t)SV * R * dir + sz * layer;
              l2[pos].re = pl.re + l2[pos].re;
              l2[pos].im = pl.im + l2[pos].im;
              ++layer;
            }
          }
      }
    }
    // kernel_l2avg_l4_multilevel
#pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < (int64_t)sz; ++id) {
      int j4 = 0;
      for (int a = 0; a < nl2; a += 2) {
        Cplx pl = {1, 0};
        for (int layer = a; layer < a + 2; ++layer) {
          const Cplx v = l2[id + sz * layer];
          pl = cmul(pl, Cplx{v.re * l2norm, v.im * l2norm});
        }
        l4[id + sz * j4].re = pl.re + l4[id + sz * j4].re;
        l4[id + sz * j4].im = pl.im + l4[id + sz * j4].im;
        ++j4;
      }
    }
  }
  // kernel_l4avg_Final_multilevel
  for (int r = 0; r < R; ++r) {
    double sr = 0, si = 0;
#pragma omp parallel for schedule(static) reduction(+ : sr, si)
    for (int id = 0; id < SV; ++id) {
      for (int dir = 0; dir < 3; ++dir) {
        Cplx pl = {1, 0};
        for (int j = 0; j < nl4; ++j) {
          const Cplx v = l4[id + (size_t)SV * r + (size_t)SV * R * dir + sz * j];
          pl = cmul(pl, Cplx{v.re * l4norm, v.im * l4norm});
        }
        sr += pl.re;
        si += pl.im;
        if (ppspace) {
          const size_t p = id + (size_t)SV * r + (size_t)SV * R * dir;
          ppspace[2 * p] = pl.re;
          ppspace[2 * p + 1] = pl.im;
        }
      }
    }
    const double norm = 1.0 / (double(SV) * 3.0);
    out[2 * r] = sr * norm;
    out[2 * r + 1] = si * norm;
  }
}

// convertOrder.cu:32-63 (kernel_convert_EO_NO<bool>).
void orc_convert(const orc_geom* gg, const double* in, double* out, int32_t eo_to_no) {
  Geo g(gg);
#pragma omp parallel for schedule(static)
  for (int idx = 0; idx < g.V; ++idx) {
    if (eo_to_no) {
      int parity = 0, id = idx;
      if (id >= g.HV) {
        parity = 1;
        id -= g.HV;
      }
      int x[4];
      g.eo_coords(id, parity, x);
      const int n = g.lex(x);
      for (int d = 0; d < 4; ++d) out[n + d * g.V] = in[id + parity * g.HV + d * g.V];
    } else {
      int x[4] = {idx % g.N[0], (idx / g.N[0]) % g.N[1], (idx / (g.N[0] * g.N[1])) % g.N[2],
                  idx / (g.N[0] * g.N[1] * g.N[2])};
      const int e = g.lex(x) >> 1, parity = Geo::parity_of(x);
      for (int d = 0; d < 4; ++d) out[e + parity * g.HV + d * g.V] = in[idx + d * g.V];
    }
  }
}

// plaquette_comps.cu:34-107 (kernel_plaquette_comps<spacetime, evenoddOrder>)
// and PlaqFields::Run normalisation (plaquette_comps.cu:148-163).
void orc_plaquette_fields(const orc_geom* gg, const double* lat, int32_t spacetime, int32_t eo_order,
                          double* field, double* means) {
  Geo g(gg);
  const int size = spacetime ? g.V : g.SV;
  std::vector<double> acc(12 * (size_t)size);
#pragma omp parallel for schedule(static)
  for (int idx = 0; idx < size; ++idx) {
    Cplx pq[6];
    for (int d = 0; d < 6; ++d) pq[d] = {0, 0};
    int pos;
    if (eo_order) {
      int parity = 0, id = idx;
      const int half = spacetime ? g.HV : g.SV / 2;
      if (id >= half) {
        parity = 1;
        id -= half;
      }
      int x[4];
      g.eo_coords(id, parity, x);
      if (spacetime) {
        for (int d = 0; d < 6; ++d) {
          const Cplx t = plaq_term(g, lat, x, id, parity, SIX_MU[d], SIX_NU[d]);
          pq[d].re += t.re;
          pq[d].im += t.im;
        }
        pos = g.lex(x);
      } else {
        for (x[TDIR] = 0; x[TDIR] < g.N[TDIR]; ++x[TDIR]) {
          const int p2 = g.lex(x) >> 1, odd = Geo::parity_of(x);
          for (int d = 0; d < 6; ++d) {
            const Cplx t = plaq_term(g, lat, x, p2, odd, SIX_MU[d], SIX_NU[d]);
            pq[d].re += t.re;
            pq[d].im += t.im;
          }
        }
        for (int d = 0; d < 6; ++d) {
          pq[d].re /= double(g.N[TDIR]);
          pq[d].im /= double(g.N[TDIR]);
        }
        pos = g.lex3(x);
      }
    } else {
      if (spacetime) {
        for (int d = 0; d < 6; ++d) {
          const Cplx t = plaq_term_no(g, lat, idx, SIX_MU[d], SIX_NU[d]);
          pq[d].re += t.re;
          pq[d].im += t.im;
        }
      } else {
        for (int t = 0; t < g.N[TDIR]; ++t)
          for (int d = 0; d < 6; ++d) {
            const Cplx v = plaq_term_no(g, lat, idx + t * g.SV, SIX_MU[d], SIX_NU[d]);
            pq[d].re += v.re;
            pq[d].im += v.im;
          }
        for (int d = 0; d < 6; ++d) {
          pq[d].re /= double(g.N[TDIR]);
          pq[d].im /= double(g.N[TDIR]);
        }
      }
      pos = idx;
    }
    for (int d = 0; d < 6; ++d) {
      field[2 * ((size_t)pos + (size_t)d * size)] = pq[d].re;
      field[2 * ((size_t)pos + (size_t)d * size) + 1] = pq[d].im;
      acc[12 * (size_t)idx + 2 * d] = pq[d].re;
      acc[12 * (size_t)idx + 2 * d + 1] = pq[d].im;
    }
  }
  for (int k = 0; k < 12; ++k) {
    double s = 0;
    for (int i = 0; i < size; ++i) s += acc[12 * (size_t)i + k];
    means[k] = s / double(size);
  }
}

}  // extern "C"
