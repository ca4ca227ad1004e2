/* u1_oracle.h -- C interface of the FP64 host oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * gfx950 path and the timed CPU baseline; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product library never links
 * or calls it.
 *
 * It is a plain C++17/OpenMP restatement of the reference's FP64 algorithms
 * (file:line citations are on each function in u1_oracle.cpp), driven by the
 * same Philox4x32-10 streams and the same elementary functions (u1_math.h) as
 * the GPU path, so that the FP64 GPU path must reproduce it sweep by sweep.
 *
 * Pinning: every function here is checked in tests/ against a closed form, a
 * known-answer vector (Philox), an independent numpy restatement or a library
 * routine (glibc, scipy).  Parity unpinned: the trajectory of the reference
 * itself, which draws from cuRAND MRG32k3a and CUDA's cyl_bessel_i0/i1 (closed
 * code, no golden vectors in the reference); agreement with it is statistical.
 *
 * Lattice layout (reference array.h / index.h:124-132): angles theta_mu(x) in
 * even/odd structure-of-arrays order  lat[eo_id + parity*V/2 + mu*V],
 * eo_id = lex(x) >> 1,  lex = x0 + Nx*(x1 + Ny*(x2 + Nz*x3)).  4D only.
 */
#ifndef U1_ORACLE_H
#define U1_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t dims[4]; /* Nx, Ny, Nz, Nt (all even) */
  double beta;
  double aniso; /* xi: space-space staples weighted 1/xi, space-time xi */
} orc_geom;

typedef struct {
  int32_t proposal; /* 0: independent theta' ~ U[-pi,pi) (reference), 1: theta' = theta + delta */
  int32_t nhit;     /* Metropolis hits per link per pass (reference: 1) */
  double step;      /* delta ~ step * U[-1, 1) */
} orc_metro;

/* ---- elementary functions and RNG (exported so tests can pin them) ---- */
void orc_sincos(const double* x, double* s, double* c, int64_t n);
void orc_exp(const double* x, double* y, int64_t n);
void orc_atan2(const double* y, const double* x, double* out, int64_t n);
void orc_bessel_i0i1(const double* x, double* i0, double* i1, int64_t n);
void orc_philox(const uint32_t* ctr /*4n*/, const uint32_t* key /*2n*/, uint32_t* out /*4n*/, int64_t n);
int orc_threads(void);
void orc_set_threads(int n);

/* ---- geometry ---- */
void orc_eo_coords(const orc_geom* g, int32_t id, int32_t parity, int32_t x[4]);
int32_t orc_eo_neighbor(const orc_geom* g, int32_t id, int32_t parity, int32_t mu, int32_t lmu, int32_t nu, int32_t lnu);

/* ---- per-link pieces ---- */
void orc_staple(const orc_geom* g, const double* lat, int32_t id, int32_t parity, int32_t mu, double out[2]);

/* ---- updates ---- */
void orc_hotstart(const orc_geom* g, double* lat, uint64_t seed, uint32_t replica);
/* One UpdateLattice(metrop, ovrn) call (reference update.cu:449-456).
 * stream/c3: Philox counter words (see u1_philox.h); mask_ml = 0 (plain), 2 or 4
 * (multilevel: spatial links on time slices t % mask_ml == 0 stay frozen).
 * accepts (nullable) receives the number of accepted Metropolis hits. */
void orc_update(const orc_geom* g, double* lat, int32_t metrop, int32_t ovrn, const orc_metro* m,
                uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t c3, int32_t mask_ml,
                uint64_t* accepts);
/* Only the Metropolis part, a single (parity, mu) phase -- used by closed-form tests. */
void orc_metropolis_phase(const orc_geom* g, double* lat, int32_t parity, int32_t mu, const orc_metro* m,
                          uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t pass, uint32_t c3,
                          uint64_t* accepts);
void orc_overrelax_phase(const orc_geom* g, double* lat, int32_t parity, int32_t mu);

/* ---- measurements ---- */
/* out = {SS.re, SS.im, ST.re, ST.im} of sum(1 - e^{iP}) normalised as the
 * reference GPU Plaquette (plaquette.cu:93-118,156-177). */
void orc_plaquette(const orc_geom* g, const double* lat, double out[4]);
/* out = {<cos P>, <sin P>} over all 6V plaquettes (reference host plaquette(), plaquette.cu:32-57). */
void orc_plaquette_host(const orc_geom* g, const double* lat, double out[2]);
/* out = {Re L, Im L} (polyakov.cu:33-49 / 60-77). */
void orc_polyakov(const orc_geom* g, const double* lat, double out[2]);
/* Per spatial site P(x) (complex, 2*SV doubles, lexicographic spatial index).
 * mode 0: prod e^{i theta_t}; 1: analytic multihit; 2: stochastic multihit
 * (mh->nhit hits with fixed staples, step mh->step). polyakov.cu:415-440. */
void orc_poly_volume(const orc_geom* g, const double* lat, int32_t mode, const orc_metro* mh,
                     uint64_t seed, uint32_t stream, uint32_t c1, uint32_t c3, double* P);
/* C(R), R = 1..rmax (2*rmax doubles), normalised by 1/(3 SV) (polyakov.cu:514-613). */
void orc_pp(const orc_geom* g, const double* P, int32_t rmax, double* out);
/* Per-site temporal link field (2*V doubles, lexicographic 4D) of multilevel.cu:239-258. */
void orc_temporal_links(const orc_geom* g, const double* lat, int32_t mode, const orc_metro* mh,
                        uint64_t seed, uint32_t stream, uint32_t c1, uint32_t c3, double* out);
/* Two-level multilevel Polyakov correlator (multilevel.cu:639-732); lat is not modified.
 * out: 2*rmax doubles (rmax = Nx/2); ppspace (nullable): 2*SV*rmax*3 doubles. */
void orc_multilevel(const orc_geom* g, const double* lat, int32_t n4, int32_t k4, int32_t n2, int32_t k2,
                    int32_t metrop, int32_t ovrn, const orc_metro* m, int32_t mh_mode, const orc_metro* mh,
                    uint64_t seed, uint32_t call_id, double* out, double* ppspace);
/* EO <-> lexicographic layout (convertOrder.cu:32-63). */
void orc_convert(const orc_geom* g, const double* in, double* out, int32_t eo_to_no);
/* Six-component plaquette fields (plaquette_comps.cu:34-107): field has 2*6*size
 * doubles (size = V if spacetime else SV, component-major), means 12 doubles
 * (normalised by 1/size as PlaqFields::Run). */
void orc_plaquette_fields(const orc_geom* g, const double* lat, int32_t spacetime, int32_t eo_order,
                          double* field, double* means);

#ifdef __cplusplus
}
#endif
#endif
