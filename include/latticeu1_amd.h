/* latticeu1_amd.h -- C ABI of the MI355X (gfx950) LatticeU1 hot path.
 *
 * Plain pointers, sizes and status codes; no HIP, torch or C++ types.  Every
 * function returns U1_OK (0) or an error code; u1_last_error() describes the
 * last failure of the calling thread.  Device pointers are allocated with
 * u1_malloc (zero filled, as the reference's dev_malloc, alloc.cpp:193) and all
 * work runs in order on one library stream.  Functions that hand results to
 * host memory synchronise; updates are asynchronous until the next
 * synchronising call (u1_sync, a measurement, a copy to the host).
 *
 * Lattice layout (reference index.h:124-132): 4V angles theta_mu(x) in
 * even/odd structure-of-arrays order lat[eo + parity*V/2 + mu*V],
 * eo = lex(x) >> 1, lex = x0 + Nx*(x1 + Ny*(x2 + Nz*x3)); angles in [-pi, pi).
 * prec selects binary64 (U1_F64, "plain") or binary32 (U1_F32, "FP32 bulk")
 * angles; all reductions are binary64 either way.  4D lattices only.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   u1_init / u1_finalize        Start / Finalize            include/cuda_error_check.h:45-48
 *   u1_malloc/u1_free/u1_memcpy  dev_malloc, Array::Copy     include/alloc.h:60-67, include/array.h:60-82
 *   u1_hotstart                  HotStart                    include/update.h:10
 *   u1_update                    UpdateLattice               include/update.h:12
 *   u1_update_plaquette          UpdateLattice + Plaquette   include/update.h:12, include/plaquette.h:12
 *   u1_plaquette                 Plaquette                   include/plaquette.h:12
 *   u1_polyakov                  Polyakov                    include/polyakov.h:13
 *   u1_poly_volume, u1_pp,       Poly2 (Polyakov_Vol + PP)   include/polyakov.h:16, src/polyakov.cu:618-654
 *   u1_poly2
 *   u1_multilevel                MultiLevel / MultiLevelField include/multilevel.h:9,25
 *   u1_convert_layout            LatticeConvert              include/array.h:186
 *   u1_plaquette_fields          PlaquetteFields             include/plaquette.h:34
 *   u1_tune_*                    setTuning, tune cache       include/parameters.h:33-37, src/tune.cpp:160-295
 *   u1_ensemble_*                (no counterpart: the reference holds one
 *                                 lattice per process in __constant__ memory,
 *                                 src/parameters.cu:78-88)
 */
#ifndef LATTICEU1_AMD_H
#define LATTICEU1_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  U1_OK = 0,
  U1_ERR_ARG = 1,         /* invalid argument (odd extent, null pointer, ...) */
  U1_ERR_HIP = 2,         /* HIP runtime error */
  U1_ERR_STATE = 3,       /* library not initialised */
  U1_ERR_UNSUPPORTED = 4, /* valid in the reference but not built here */
  U1_ERR_IO = 5           /* file I/O (tune cache) */
};

enum { U1_F64 = 0, U1_F32 = 1 };
enum { U1_PROPOSAL_INDEPENDENT = 0, U1_PROPOSAL_DELTA = 1 };
enum { U1_POLY_PLAIN = 0, U1_POLY_MULTIHIT = 1, U1_POLY_MULTIHIT_STOCHASTIC = 2 };
enum { U1_H2H = 0, U1_H2D = 1, U1_D2H = 2, U1_D2D = 3 };
/* Philox stream tags (counter word 2, bits 24-31) */
enum {
  U1_STREAM_UPDATE = 0,
  U1_STREAM_HOTSTART = 1,
  U1_STREAM_MULTIHIT = 2,
  U1_STREAM_ML_L4 = 3,
  U1_STREAM_ML_L2 = 4,
  U1_STREAM_ML_MULTIHIT = 5
};

/* Geometry and couplings (reference SetupLatticeParameters, parameters.cu:146-208;
 * dims = Grid[0..3] = Nx, Ny, Nz, Nt, all even). */
typedef struct {
  int32_t dims[4];
  double beta;
  double aniso; /* xi: space-space staples weighted 1/xi, space-time xi (staple.h:75-76) */
} u1_geom;

/* Metropolis settings.  Reference behaviour: proposal INDEPENDENT, nhit 1
 * (update.cu:181-182).  The 48^4 benchmark runs: DELTA, nhit 4, step 0.5. */
typedef struct {
  int32_t proposal;
  int32_t nhit;
  double step; /* delta = step * U[-1, 1) */
} u1_metro;

/* ---------------------------------------------------------------- runtime */
int u1_init(int device);
int u1_finalize(void);
int u1_sync(void);
const char* u1_last_error(void);
int u1_device_name(char* buf, size_t len);

/* ---------------------------------------------------------------- memory */
int u1_malloc(void** ptr, size_t bytes);
int u1_free(void* ptr);
int u1_memcpy(void* dst, const void* src, size_t bytes, int kind);
int u1_zero(void* ptr, size_t bytes);
/* n elements, FP64 <-> FP32 (or same-precision copy) on the device */
int u1_convert_precision(const void* src, int src_prec, void* dst, int dst_prec, int64_t n);

/* ---------------------------------------------------------------- updates */
/* Hot start: theta = pi*(2u - 1), u from Philox stream HOTSTART, c0 = link index. */
int u1_hotstart(const u1_geom* g, int prec, void* lat, uint64_t seed, uint32_t replica);
/* One UpdateLattice(metrop, ovrn): metrop Metropolis passes then ovrn
 * overrelaxation passes, each over parity 0,1 and mu 0..3.  Philox counter
 * (link, sweep, stream|pass|hitpair, c3).  mask_ml 0, 2 or 4: multilevel
 * sublattice mask (spatial links on slices t % mask_ml == 0 frozen).
 * accepts (nullable, synchronises): accepted Metropolis hits are added. */
int u1_update(const u1_geom* g, int prec, void* lat, int32_t metrop, int32_t ovrn, const u1_metro* m,
              uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t c3, int32_t mask_ml, uint64_t* accepts);

/* u1_update (mask_ml 0, no accept count) fused with the plaquette of the
 * updated lattice, returned in u1_plaquette's form.  The plaquettes are
 * summed inside the last pass's (parity 1, mu = 1, 2, 3) launches from the
 * staple pieces and the new angles, saving a separate pass over the lattice.
 * Counterpart of the reference's UpdateLattice followed by Plaquette
 * (include/update.h:12, include/plaquette.h:12).  Binary64: equal to
 * u1_plaquette up to summation order; binary32: within ~1e-6 relative (the
 * pieces use the update kernels' hardware sin/cos). */
int u1_update_plaquette(const u1_geom* g, int prec, void* lat, int32_t metrop, int32_t ovrn, const u1_metro* m,
                        uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t c3, double out[4]);

/* One checkerboard phase: a single launch of kernel_metropolis (kind 0, with
 * Metropolis pass index `pass`) or kernel_overrelaxation (kind 1) on parity
 * `parity`, direction mu (update.cu:177-190, 227-231). */
int u1_phase(const u1_geom* g, int prec, void* lat, int32_t kind, int32_t parity, int32_t mu, const u1_metro* m,
             uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t pass, uint32_t c3, int32_t mask_ml);

/* ---------------------------------------------------------------- measurements */
/* out[4] = {SS.re, SS.im, ST.re, ST.im}: sum of 1 - e^{iP} over space-space and
 * space-time plaquettes, normalised by 1/(3V) each (plaquette.cu:156-177).
 * Mean plaquette <P> = 1 - (out[0] + out[2]) / 2. */
int u1_plaquette(const u1_geom* g, int prec, const void* lat, double out[4]);
/* u1_plaquette that allocates nothing, for callers that cannot allocate
 * device memory (code under hipGraph stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * result, in u1_plaquette's form, goes to out_dev (device memory, 4 doubles),
 * asynchronously on the library stream; nothing synchronises.  The reduction
 * rows (4 doubles per workgroup of the plaquette pass, plus 4 per 512 rows
 * for the fold) live in a fixed 8 MiB device scratch array of the library
 * (one per GPU, held until the process ends): every lattice up to 96^4 fits
 * in both precisions; a lattice that needs more is rejected with U1_ERR_ARG
 * before any launch.  The scratch keeps the last call's contents; every row
 * is written before it is read, so nothing needs clearing.  Calls that
 * overlap on one GPU share the scratch and race: issue them in order on one
 * stream (the library stream does). */
int u1_plaquette_noalloc(const u1_geom* g, int prec, const void* lat, double* out_dev);
/* out[2] = Polyakov loop (1/SV) sum_x e^{i sum_t theta_t(x,t)}. */
int u1_polyakov(const u1_geom* g, int prec, const void* lat, double out[2]);
/* Per spatial site P(x) into P_dev (2*SV doubles, lexicographic spatial index).
 * mode U1_POLY_*; mh (stochastic mode only): proposal / nhit / step; Philox
 * counter (link, c1, stream|0|hitpair, c3). */
int u1_poly_volume(const u1_geom* g, int prec, const void* lat, int32_t mode, const u1_metro* mh, uint64_t seed,
                   uint32_t stream, uint32_t c1, uint32_t c3, void* P_dev);
/* C(R) = (1/(3 SV)) sum_x sum_dir P(x) conj P(x + R e_dir), R = 1..rmax, into out[2*rmax]. */
int u1_pp(const u1_geom* g, const void* P_dev, int32_t rmax, double* out);
/* Poly2: u1_poly_volume then u1_pp with rmax = Nx/2; out[2*(Nx/2)]. */
int u1_poly2(const u1_geom* g, int prec, const void* lat, int32_t mode, const u1_metro* mh, uint64_t seed,
             uint32_t c1, uint32_t c3, double* out);
/* Per-site temporal-link field (2*V doubles, lexicographic 4D index), multilevel.cu:239-258. */
int u1_temporal_links(const u1_geom* g, int prec, const void* lat, int32_t mode, const u1_metro* mh,
                      uint64_t seed, uint32_t stream, uint32_t c1, uint32_t c3, void* out_dev);
/* Two-level multilevel correlator (multilevel.cu:639-732) on a copy of lat.
 * Sublattice updates: metrop/ovrn passes with m, Philox streams ML_L4 / ML_L2,
 * c1 = step index, c3 = call_id.  Temporal links: mh_mode (U1_POLY_MULTIHIT as
 * the reference, or U1_POLY_MULTIHIT_STOCHASTIC with mh).  out[2*(Nx/2)];
 * ppspace_dev (nullable): 2*SV*(Nx/2)*3 doubles (MultiLevelField).  Nt % 4 == 0. */
int u1_multilevel(const u1_geom* g, int prec, const void* lat, int32_t n4, int32_t k4, int32_t n2, int32_t k2,
                  int32_t metrop, int32_t ovrn, const u1_metro* m, int32_t mh_mode, const u1_metro* mh,
                  uint64_t seed, uint32_t call_id, double* out, void* ppspace_dev);
/* Layout permutation EO <-> lexicographic (convertOrder.cu:32-63); out distinct from in. */
int u1_convert_layout(const u1_geom* g, int prec, const void* in, void* out, int32_t eo_to_no);
/* Six plaquette components per site (plaquette_comps.cu:34-107): field_dev gets
 * 2*6*size doubles (size = V if spacetime else SV), means[12] the normalised means. */
int u1_plaquette_fields(const u1_geom* g, int prec, const void* lat, int32_t spacetime, int32_t eo_order,
                        void* field_dev, double means[12]);

/* ---------------------------------------------------------------- checkpoint */
/* Configuration file (no counterpart in the reference, which saves no
 * configurations): a 64-byte little-endian header -- magic "U1AMDCFG",
 * version 1, dims[4], beta, xi, seed, sweep -- then 4V binary64 angles in the
 * lexicographic order of LatticeConvert(eo_to_no) ([mu][lex]).  With the
 * counter-based generator, (configuration, seed, sweep) resumes a chain
 * exactly.  Load checks the file's geometry against g; seed/sweep (nullable)
 * receive the stored counters. */
int u1_config_save(const char* path, const u1_geom* g, int prec, const void* lat, uint64_t seed, uint64_t sweep);
int u1_config_load(const char* path, const u1_geom* g, int prec, void* lat, uint64_t* seed, uint64_t* sweep);

/* ---------------------------------------------------------------- batched replicas */
typedef struct u1_ensemble u1_ensemble;
/* nrep independent lattices (own geometry, beta, xi, seed; hot[r] != 0 for a hot
 * start, else cold) packed in one buffer of precision prec.  Replica r uses
 * Philox c3 = r.  Every sweep (metrop Metropolis + ovrn OR passes) of all
 * replicas is one launch per (pass, parity, mu), captured once in a hipGraph. */
int u1_ensemble_create(u1_ensemble** out, int32_t nrep, const u1_geom* geoms, const uint64_t* seeds,
                       const int32_t* hot, int prec, int32_t metrop, int32_t ovrn, const u1_metro* m);
int u1_ensemble_sweep(u1_ensemble* e, int32_t nsweeps, int32_t use_graph);
/* nsweeps graph sweeps of all replicas, each with the plaquette of every
 * replica summed inside its last pass (as u1_update_plaquette); series
 * (host, nsweeps * nrep * 4 doubles) receives them sweep-major in
 * u1_plaquette's form.  No counterpart in the reference, which holds one
 * lattice per process and measures with Plaquette (include/plaquette.h:12). */
int u1_ensemble_sweep_plaquette(u1_ensemble* e, int32_t nsweeps, double* series);
/* Timing helper (no counterpart in the reference): u1_ensemble_sweep_plaquette
 * over nsweeps (>= 1; ovrn >= 1) with elapsed_ms the device time from one HIP
 * event pair on the library stream around the graph replays (series copies
 * of all but the last chunk included).  kernel_ms (nullable, 20 doubles) as in
 * u1_bench_updates, for the ensemble's launches over all replicas: classes
 * 0 Metropolis, 1 overrelaxation, 2 measuring overrelaxation phase, 3 fused
 * Metropolis pair, 4 fused overrelaxation pair (5-8 unused), timed on a scratch
 * copy of the packed lattice after the timed region. */
int u1_ensemble_bench(u1_ensemble* e, int32_t nsweeps, double* series, double* elapsed_ms, double* kernel_ms);
/* out[4*nrep]: per-replica plaquette as u1_plaquette. */
int u1_ensemble_plaquette(u1_ensemble* e, double* out);
/* Replica r angles as FP64 host array (4V, EO layout); write sets them. */
int u1_ensemble_read(u1_ensemble* e, int32_t r, double* host);
int u1_ensemble_write(u1_ensemble* e, int32_t r, const double* host);
int u1_ensemble_info(u1_ensemble* e, uint64_t* sweeps_done, int64_t* total_links);
int u1_ensemble_destroy(u1_ensemble* e);

/* ---------------------------------------------------------------- autotuner */
/* Tuning picks the workgroup size per (kernel, geometry, precision) by timing
 * candidates on a scratch copy (results are independent of the choice). */
int u1_tune_enable(int32_t on);
/* Fused update launches (no reference counterpart): on a single lattice the
 * phases (parity, 0) and (parity, 1) of a pass can run as one launch, one
 * workgroup per (z, t) plane with a workgroup barrier between the phases; the
 * results are bit-identical to two launches.  mode 0: never; 1 (default):
 * where the tuner measures the fused launch faster (needs tuning on); 2:
 * always.  The tune cache keeps the decision as "MetropolisPair" /
 * "OverRelaxationPair" entries (block.x 0 = two launches). */
int u1_set_pair_fusion(int32_t mode);
/* Fused parity passes (no reference counterpart): on a single lattice without
 * multilevel mask whose (z, t) planes tile 2 x 4 (Nz even, Nt a multiple of
 * 4), the four phases (parity, 0..3) of a pass can run as one launch; the
 * workgroups of phase mu wait on device-side completion counters of the
 * phase-(mu - 1) planes they depend on instead of a launch boundary.  The
 * results are bit-identical to the separate launches.  mode 0: never; 1
 * (default): where the tuner measures the fused pass faster (needs tuning
 * on); 2: always.  Not used during stream capture.  The tune cache keeps the
 * decision as "MetropolisChain" / "OverRelaxationChain" entries (block.x 0 =
 * unfused).  A dependency wait that times out (bounded, about 0.1 s) is
 * reported as U1_ERR_HIP by the next call that synchronises the library
 * stream (u1_sync, or any call returning results to the host). */
int u1_set_pass_fusion(int32_t mode);
/* Multilevel level-2 steps (u1_multilevel): replayed from one captured
 * hipGraph (on = 1, default) or enqueued eagerly (0).  The same kernels read
 * the same device-side step counter either way, so the results are
 * identical.  No counterpart in the reference, which launches eagerly
 * (multilevel.cu:680-700).  Callable before u1_init. */
int u1_set_graph_capture(int32_t on);
/* Verbosity (parameters.h:33-37): 0 SILENT, 1 SUMMARIZE (the tuner reports
 * its choice on stderr), 2 VERBOSE and 3 DEBUG_VERBOSE (every candidate's
 * time).  Callable before u1_init. */
int u1_set_verbosity(int32_t level);
int u1_tune_load(const char* path);
int u1_tune_save(const char* path);
/* Number of launch-shape entries held (tuned in this process or loaded). */
int u1_tune_entries(int32_t* n);

/* ---------------------------------------------------------------- timing helper */
/* Run nsweeps of u1_update (sweep indices sweep0, sweep0+1, ...; a plaquette
 * measurement into plaq_series[4*i] after every sweep if measure_plaq: 1 fused
 * into the sweep as u1_update_plaquette, 2 a separate u1_plaquette pass) and
 * return the elapsed device time in ms, from HIP events on the library stream.
 * The timed region holds no other events.
 * kernel_ms (nullable, 20 doubles): kernel timing after the timed region, on
 * a scratch copy of the lattice (the chain is not advanced): per launch class
 * c of the sweep -- 0 Metropolis phase, 1 overrelaxation phase, 2 measuring
 * overrelaxation phase, 3 fused Metropolis pair, 4 fused overrelaxation pair,
 * 5 measuring fused overrelaxation pair (u1_set_pair_fusion), 6 fused
 * Metropolis pass, 7 fused overrelaxation pass, 8 measuring fused
 * overrelaxation pass (u1_set_pass_fusion) -- 32 back-to-back launches in the
 * sweep's shapes between one event pair;
 * kernel_ms[c] is the mean launch duration (0 if the sweep has none) and
 * kernel_ms[9 + c] the launches per sweep; kernel_ms[18] is the timed
 * region's ms per sweep and kernel_ms[19] the launches timed per class. */
int u1_bench_updates(const u1_geom* g, int prec, void* lat, int32_t metrop, int32_t ovrn, const u1_metro* m,
                     uint64_t seed, uint32_t sweep0, int32_t nsweeps, int32_t measure_plaq, double* plaq_series,
                     double* elapsed_ms, double* kernel_ms);

/* ---------------------------------------------------------------- self test */
/* Device evaluation of the shared elementary functions (no reference
 * counterpart; lets tests check bit identity with the host oracle).
 * fn 0: sincos_d(x) -> out[2i], out[2i+1] = sin, cos;  1: exp_d(x);
 *    2: atan2_d(x, y) -> atan2(x[i], y[i]);  3: sincos_f((float)x);
 *    4: atan2_f((float)x, (float)y);  5: Bessel I0, I1 (out 2 per input);
 *    6: wrap_d(x);  7: hardware sincos of the FP32 update kernels (2 outputs);
 *    8: fast atan2 of the FP32 update kernels;  9: the wrap of the FP32
 *    update kernels, (float)wrap_near_d((float)x);  10: the exp of the
 *    binary64 Metropolis hits (exp_d's bits by a per-wavefront fast path).
 *    x, y, out are host arrays. */
int u1_selftest_math(int32_t fn, const double* x, const double* y, double* out, int64_t n);
/* Philox4x32-10 on the device: ctr[4n], key[2n] -> out[4n] (host arrays). */
int u1_selftest_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int64_t n);
/* Streaming reference for the roofline (no counterpart in the reference):
 * moves the bytes of one (parity, mu) phase, 36 n (4.5 floats per each of n
 * updated links), with an ideal access pattern and returns the rate in GB/s
 * averaged over reps launches.  pattern 0: read 9n floats; 1: copy 4.5n
 * floats; 2: read eight n-float arrays, write one (the phase's 8:1 mix).
 * n % 8 == 0, n <= 2^30. */
int u1_selftest_stream(int32_t pattern, int64_t n, int32_t reps, double* gbs);

#ifdef __cplusplus
}
#endif
#endif
