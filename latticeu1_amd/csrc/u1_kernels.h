// u1_kernels.h -- host-side launchers of the gfx950 kernels (internal to the
// library; the public boundary is include/latticeu1_amd.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "u1_device.h"

namespace u1d {

// One checkerboard (parity, mu) phase of Metropolis or overrelaxation over one
// lattice, or over every replica of a batch in a single launch.
struct PhaseArgs {
  RepDesc single;               // used when reps == nullptr
  const RepDesc* reps;          // batched: replica table (device)
  const uint16_t* blockmap;     // batched: workgroup -> replica (device)
  int32_t parity;
  int32_t ml;                   // multilevel mask: 0 none, 2 or 4
  int32_t proposal;             // 0 independent, 1 theta + delta
  int32_t nhit;
  double step;
  // binary32 Metropolis proposal a (2 wu 2^-32 - 1) + thm theta, a = pi or
  // step: pscale = a 2^-31, poff = a, thm = 0 or 1 (set by launch_phase)
  float pscale_f, poff_f, thm_f;
  uint32_t stream, pass, sweep;
  uint32_t xcd;                 // 1: XCD-aware workgroup order (xcd_remap)
  const uint32_t* sweep_dev;    // if set, sweep index = *sweep_dev * sweep_mul + sweep (graph replays)
  uint32_t sweep_mul;
  unsigned long long* accepts;  // if set, accepted hits are added here
  double* meas;                 // if set (mu >= 1): plaquette rows, 4 doubles per (logical) workgroup
};

enum { KIND_METROPOLIS = 0, KIND_OVERRELAX = 1 };
enum { PREC_F64 = 0, PREC_F32 = 1 };

// spt: even/odd sites per thread (1 or 2); the replica descriptors' plane
// decomposition must have been made with set_phase_blocks(r, block * spt).
hipError_t launch_phase(int prec, int kind, int mu, void* lat_base, const PhaseArgs& a, int nblocks, int block,
                        int spt, hipStream_t st);

// The phases (parity, 0) and (parity, 1) in one launch (k_pair): one
// workgroup per (z, t) plane.  A single lattice: a.single carries
// set_phase_blocks(r, Nx/2 * Ny), nblocks = Nz * Nt and plane_sites =
// Nx/2 * Ny.  Batched replicas (a.reps, a.blockmap): every replica descriptor
// carries its own set_phase_blocks(r, Nx/2 * Ny) with blk_begin its first
// workgroup, the block map covers nblocks, plane_sites is the largest plane;
// no measurement.  spt 1, 2 (both precisions), 3 or 4 (binary32); block a
// multiple of 64 up to pair_max_block; plane_sites up to pair_max_plane (the
// plane's new theta_0 is exchanged through LDS).  a.meas makes phase
// (parity, 1) measuring.
constexpr int pair_max_block(int elem_bytes, int spt) { return elem_bytes == 8 || spt >= 3 ? 512 : 1024; }
constexpr int pair_max_plane(int elem_bytes) { return 65536 / elem_bytes; }
hipError_t launch_pair(int prec, int kind, void* lat_base, const PhaseArgs& a, int nblocks, int block, int spt,
                       int plane_sites, hipStream_t st);

// The four phases (parity, 0..3) of one pass in one launch (k_chain, schedule
// and dependencies in u1_chain.h).  A single lattice without multilevel mask:
// a.single carries set_phase_blocks(r, block * spt); sched is the list of
// chain_schedule(Nz, Nt, bpp, tz) in device memory, nblocks = 8 * its length,
// and (tz_log2, bz, bt, bz_m, bz_s) its tiling: 2^tz_log2 tiles of bz planes
// along z, 8 / 2^tz_log2 of bt along t, magic division by bz.  cnt holds 4 Nz Nt completion counters (row mu,
// plane z + Nz t); every wavefront adds 1 to its (mu, plane) counter when
// done, and phase mu waits until the counters of row mu - 1 it depends on
// reach tgt[mu - 1].  A wait that times out sets *err (host-visible memory).
// a.meas makes phases 1-3 measuring: row (mu - 1) Nz Nt bpp + (z + Nz t) bpp
// + chunk.  spt 1 or 2, block a multiple of 64 up to 512.
struct ChainArgs {
  const uint32_t* sched;
  uint32_t* cnt;
  uint32_t* err;
  uint32_t tgt[4];
  uint32_t tz_log2, bz, bt, bz_m, bz_s;
};
hipError_t launch_chain(int prec, int kind, void* lat_base, const PhaseArgs& a, const ChainArgs& c, int nblocks,
                        int block, int spt, hipStream_t st);

hipError_t launch_hotstart(int prec, void* lat, const RepDesc& r, int block, hipStream_t st);

// measurement launchers (u1_measure.hip)
// partial sums per workgroup; finish_sums folds them in workgroup order.
// Plaquette: r carries the plane decomposition of set_phase_blocks(r, block *
// spt) with r.nblk doubled (even sites, then odd sites); nblk = r.nblk.
hipError_t launch_plaquette(int prec, const void* lat, const RepDesc& r, double* partials /*4*nblk*/, int nblk,
                            int block, int spt, hipStream_t st);
hipError_t launch_polyakov(int prec, const void* lat, const RepDesc& r, double* partials /*2*nblk*/, int nblk,
                           int block, hipStream_t st);
// Fold of nblk rows of nv partial sums, in row order, times scale.  Beyond
// kFoldRows rows the fold takes two levels; scratch then holds
// fold_groups(nblk) * nv doubles.
constexpr int kFoldRows = 512;
inline int fold_groups(int nblk) { return (nblk + kFoldRows - 1) / kFoldRows; }
inline size_t fold_scratch(int nblk, int nv) { return (size_t)nv * (size_t)fold_groups(nblk); }
hipError_t launch_finish_sums(const double* partials, int nblk, int nv, double scale, double* out, double* scratch,
                              hipStream_t st);
hipError_t launch_poly_volume(int prec, const void* lat, const RepDesc& r, int mode, int proposal, int nhit,
                              double step, uint32_t stream, uint32_t c1, uint32_t c3, double* P, int block,
                              hipStream_t st);
// c1_dev (nullable): Philox counter word c1 read from device memory (graph replays)
hipError_t launch_temporal_links(int prec, const void* lat, const RepDesc& r, int mode, int proposal, int nhit,
                                 double step, uint32_t stream, uint32_t c1, uint32_t c3, double* out, int block,
                                 hipStream_t st, const uint32_t* c1_dev = nullptr);
// P(x) = prod_t T(x, t) from a temporal-link field (2*V doubles, lexicographic 4D)
hipError_t launch_tprod(const double* tl, const RepDesc& r, double* P, hipStream_t st);
hipError_t launch_pp(const double* P, const RepDesc& r, int rmax, double* partials /*2*rmax*nblk*/, int nblk,
                     int block, hipStream_t st);
hipError_t launch_convert_layout(int prec, const void* in, void* out, const RepDesc& r, int eo_to_no, int block,
                                 hipStream_t st);
hipError_t launch_convert_precision(const void* in, int in_prec, void* out, int out_prec, int64_t n, int block,
                                    hipStream_t st);
hipError_t launch_plaquette_fields(int prec, const void* lat, const RepDesc& r, int spacetime, int eo_order,
                                   double* field, double* partials /*12*nblk*/, int nblk, int block,
                                   hipStream_t st);
hipError_t launch_fill_zero(void* p, int64_t bytes, hipStream_t st);

// multilevel (u1_multilevel.hip)
// T2 accumulation of one level-2 step; pq: scratch of 4 SV (Nt/2) doubles for
// the two-slice products.
hipError_t launch_ml_l2(const double* tl, double* pq, double* l2, const RepDesc& r, int rmax, int block,
                        hipStream_t st);
hipError_t launch_ml_l4(const double* l2, double* l4, const RepDesc& r, int rmax, double l2norm, int block,
                        hipStream_t st);
hipError_t launch_ml_final(const double* l4, double* ppspace, const RepDesc& r, int rmax, double l4norm,
                           double* partials /*2*rmax*nblk*/, int nblk, int block, hipStream_t st);

// batched replica plaquettes: one partial row per workgroup, tagged by replica.
hipError_t launch_plaquette_batched(int prec, const void* lat_base, const RepDesc* reps, const uint16_t* blockmap,
                                    double* partials /*4*nblk*/, int nblk, int block, int spt, hipStream_t st);
hipError_t launch_batched_finish(const double* partials, const RepDesc* reps, int nrep, double* out /*4*nrep*/,
                                 hipStream_t st);
hipError_t launch_increment(uint32_t* counter, hipStream_t st);
// Per-replica fold of the measuring phases' plaquette rows (nseg segments of
// seg_rows rows; replica r owns rows [blk_begin, blk_begin + nblk) of each)
// into series[(*sweep_dev - *base_dev) * nrep + r][4], skipped at or past cap.
hipError_t launch_ensemble_plaquette_fold(const double* rows, int nseg, int seg_rows, const RepDesc* reps, int nrep,
                                          const uint32_t* sweep_dev, const uint32_t* base_dev, int cap,
                                          double* series, hipStream_t st);

// self test (u1_selftest.hip)
hipError_t launch_selftest_math(int fn, const double* x, const double* y, double* out, int64_t n, hipStream_t st);
hipError_t launch_selftest_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int64_t n,
                                  hipStream_t st);
// streaming reference over 36 n bytes (a: 9n floats, b: n floats; n % 8 == 0)
hipError_t launch_stream(int pattern, float* a, float* b, int64_t n, hipStream_t st);

}  // namespace u1d
