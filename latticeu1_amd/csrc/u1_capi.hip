// u1_capi.hip -- implementation of the C ABI (include/latticeu1_amd.h):
// runtime context, argument checks, launch geometry, the autotuner, the
// multilevel driver and the batched-replica ensemble with its hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/latticeu1_amd.h"
#include "u1_chain.h"
#include "u1_kernels.h"

using u1d::PhaseArgs;
using u1d::RepDesc;

namespace {

thread_local std::string g_err;

// Scratch of u1_plaquette_noalloc: the reduction rows of the plaquette pass
// and the fold's second-level rows, in a file-scope device array (each GPU the
// library runs on holds its own copy until the process ends).  Host code never
// passes the array's name to a kernel: it uses the device address that
// hipGetSymbolAddress returns for the current device (kept per device in
// g_scratch_dev, set in u1_init).
constexpr size_t kNoAllocScratchDoubles = 1u << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];
constexpr int kMaxDevices = 64;
constexpr int kChainMaxMargin = 1024;  // fused-pass schedule margin accepted from a tune cache
// A fused pass must take under this fraction of the unfused pass's time:
// overrelaxation passes save bandwidth and ramps; Metropolis passes, bound by
// their arithmetic, only ramps, a gain the short tuning runs overstate (a
// binary32 Metropolis pass chosen at 5 % ahead ran no faster in full sweeps).
constexpr float kChainMinGainOR = 0.97f;
constexpr float kChainMinGainMetropolis = 0.90f;
double* g_scratch_dev[kMaxDevices] = {};

// Launch shape of an update phase kernel: workgroup size, even/odd sites per
// thread, XCD-aware workgroup order.
struct Cfg {
  int block, spt, xcd, tile = 0;  // tile: single-lattice planes in 2 x 4 tiles (set_phase_tiles)
  int margin = 0;                 // fused passes (tile: tiles along z): schedule margin (chain_schedule)
};

struct Ctx {
  bool init = false;
  int device = 0;
  hipStream_t st = nullptr;
  double* partials = nullptr;  // reduction scratch
  size_t partials_cap = 0;     // in doubles
  double* work = nullptr;      // Polyakov-volume and temporal-link scratch of u1_poly_volume / u1_poly2
  size_t work_cap = 0;         // in doubles
  unsigned long long* counter = nullptr;
  double* dres = nullptr;  // small device result buffer (64 doubles)
  bool tune = true;
  // fused (parity, 0) + (parity, 1) update launches on single lattices:
  // 0 never, 1 where the tuner measures them faster, 2 always
  int pair = 1;
  int verbosity = 0;                      // 0 silent, 1 summarize, 2 verbose (tuner output on stderr)
  // multilevel level-2 steps: replayed from one captured hipGraph (true) or
  // enqueued eagerly (false); the same kernels and counters either way
  bool graphs = true;
  // fused parity passes (k_chain) on single lattices: 0 never, 1 where the
  // tuner measures them faster, 2 always (where the geometry allows)
  int chain = 1;
  uint32_t* chain_cnt = nullptr;  // completion counters, 4 Nz Nt
  size_t chain_cnt_cap = 0;       // in counters
  uint32_t* chain_err = nullptr;  // wait-timeout flag (pinned host memory, written by the kernels)
  int chain_nz = 0, chain_nt = 0, chain_per = 0;  // layout of the counters' last use (per: wavefronts per plane)
  uint32_t chain_inc[4] = {};                     // launches that counted into each row since the last clear
  std::map<std::string, uint32_t*> chain_sched;   // "Nz Nt bpp" -> schedule list (device)
  std::map<std::string, int> chain_sched_len;
  std::map<std::string, Cfg> tunecache;  // key -> launch shape
  int tune_loaded = 0;
};
Ctx G;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return fail(U1_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define REQ_INIT() \
  if (!G.init) return fail(U1_ERR_STATE, "u1_init has not been called")
// Pointers handed to kernels or written on the host are never null.
#define REQ_PTR(p) \
  if (!(p)) return fail(U1_ERR_ARG, "null pointer argument: " #p)
#define RET(expr)             \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != U1_OK) return rc_; \
  } while (0)

// Temporary device buffer, freed on every return path (after the library
// stream has drained, so no queued kernel still uses it).
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) {
      (void)hipStreamSynchronize(G.st);
      (void)hipFree(p);
    }
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 4); }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

// HIP events owned by one call, destroyed on every return path.
struct EventSet {
  std::vector<hipEvent_t> ev;
  EventSet() = default;
  EventSet(const EventSet&) = delete;
  EventSet& operator=(const EventSet&) = delete;
  ~EventSet() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  hipError_t create(hipEvent_t* out) {
    const hipError_t e = hipEventCreate(out);
    if (e == hipSuccess) ev.push_back(*out);
    return e;
  }
};

// parameters.cu:163-168 rejects odd extents; the kernels here also need
// 4 * V < 2^31 (32-bit link index and Philox counter word).
int check_geom(const u1_geom* g) {
  if (!g) return fail(U1_ERR_ARG, "null geometry");
  int64_t v = 1;
  for (int i = 0; i < 4; ++i) {
    if (g->dims[i] < 2 || (g->dims[i] % 2) != 0)
      return fail(U1_ERR_ARG, "Number of points should be an even number >= 2 in every direction");
    v *= g->dims[i];
  }
  if (4 * v >= (int64_t)1 << 31) return fail(U1_ERR_ARG, "lattice too large: 4*V must be < 2^31");
  if (!(g->aniso > 0.0)) return fail(U1_ERR_ARG, "anisotropy must be > 0");
  return U1_OK;
}

int check_prec(int prec) {
  if (prec != U1_F64 && prec != U1_F32) return fail(U1_ERR_ARG, "precision must be U1_F64 or U1_F32");
  return U1_OK;
}

int check_metro(const u1_metro* m) {
  if (!m) return fail(U1_ERR_ARG, "null Metropolis settings");
  if (m->proposal != U1_PROPOSAL_INDEPENDENT && m->proposal != U1_PROPOSAL_DELTA)
    return fail(U1_ERR_ARG, "unknown proposal kind");
  if (m->nhit < 1 || m->nhit > 8190) return fail(U1_ERR_ARG, "nhit must be in [1, 8190]");
  if (m->proposal == U1_PROPOSAL_DELTA && !(m->step >= 0.0 && m->step <= 3.141592653589793))
    return fail(U1_ERR_ARG, "the delta proposal needs 0 <= step <= pi");
  return U1_OK;
}

RepDesc make_rep(const u1_geom* g, uint64_t seed, uint32_t c3) {
  RepDesc r{};
  for (int i = 0; i < 4; ++i) r.N[i] = g->dims[i];
  r.nxh = r.N[0] / 2;
  r.V = r.N[0] * r.N[1] * r.N[2] * r.N[3];
  r.HV = r.V / 2;
  r.SV = r.N[0] * r.N[1] * r.N[2];
  r.sy = r.nxh;
  r.sz = r.nxh * r.N[1];
  r.st = r.nxh * r.N[1] * r.N[2];
  r.blk_begin = 0;
  r.nblk = 0;
  r.c3 = c3;
  const u1d::MagicDiv a = u1d::make_magic((uint32_t)r.nxh), b = u1d::make_magic((uint32_t)r.N[1]),
                     c = u1d::make_magic((uint32_t)r.N[2]);
  r.div_nxh_m = a.m;
  r.div_nxh_s = a.s;
  r.div_ny_m = b.m;
  r.div_ny_s = b.s;
  r.div_nz_m = c.m;
  r.div_nz_s = c.s;
  r.beta = g->beta;
  r.aniso = g->aniso;
  r.inv_aniso_f = (float)(1.0 / g->aniso);
  r.aniso_f = (float)g->aniso;
  r.bl2e_f = (float)(g->beta * 1.4426950408889634);  // beta log2(e)
  r.seed = seed;
  r.lat_off = 0;
  return r;
}

void set_blocks(RepDesc& r, int block) { u1d::set_phase_blocks(r, block); }
// Single-lattice update launches: plane decomposition of a launch shape,
// planes in 2 x 4 tiles (one per XCD) if the shape asks for it.
void set_update_blocks(RepDesc& r, const Cfg& c) {
  u1d::set_phase_blocks(r, c.block * c.spt);
  u1d::set_phase_tiles(r, c.tile != 0);
}

size_t elem_size(int prec) { return prec == U1_F64 ? sizeof(double) : sizeof(float); }

int ensure_partials(size_t n) {
  if (n <= G.partials_cap) return U1_OK;
  HIPCHK(hipStreamSynchronize(G.st));  // no queued kernel still reads the old buffer
  if (G.partials) HIPCHK(hipFree(G.partials));
  G.partials = nullptr;
  size_t cap = std::max<size_t>(n, 1 << 16);
  HIPCHK(hipMalloc((void**)&G.partials, cap * sizeof(double)));
  G.partials_cap = cap;
  return U1_OK;
}

// Kept between calls and grown on demand (the reference allocates these
// per call, polyakov.cu:443-508, 618-654).
int ensure_work(size_t n) {
  if (n <= G.work_cap) return U1_OK;
  HIPCHK(hipStreamSynchronize(G.st));
  if (G.work) HIPCHK(hipFree(G.work));
  G.work = nullptr;
  G.work_cap = 0;
  HIPCHK(hipMalloc((void**)&G.work, n * sizeof(double)));
  G.work_cap = n;
  return U1_OK;
}

// Room for nblk rows of nv partial sums followed by the fold's scratch rows.
int ensure_fold(int nblk, int nv) { return ensure_partials((size_t)nv * nblk + u1d::fold_scratch(nblk, nv)); }
double* fold_scratch_ptr(int nblk, int nv) { return G.partials + (size_t)nv * nblk; }

// Launch shape of the plaquette: the (workgroup, sites per thread) pair whose
// chunks tile a (z, t) plane of P = Nx/2 * Ny sites with the fewest idle
// slots, then the most sites per workgroup (fewer partial rows).
void plaquette_shape(const RepDesc& r, int prec, int* block, int* spt) {
  const int P = r.nxh * r.N[1];
  long best_slots = -1;
  int bb = 64, bs = 1;
  for (int s : {1, 2, 4}) {
    if (s == 4 && prec == U1_F64) continue;
    for (int b : {64, 128, 192, 256}) {
      const long chunk = (long)b * s;
      const long slots = (P + chunk - 1) / chunk * chunk;
      if (best_slots < 0 || slots < best_slots || (slots == best_slots && chunk > (long)bb * bs)) {
        best_slots = slots;
        bb = b;
        bs = s;
      }
    }
  }
  *block = bb;
  *spt = bs;
}

// ------------------------------------------------------------------ autotuner
// Own tuner (replacing the QUDA-derived tune.cpp:306-416): launch shape per
// (kernel class, geometry, precision) -- workgroup size and even/odd sites per
// thread -- chosen by timing every candidate with HIP events on a scratch copy
// of the lattice, so tuning never perturbs the chain (the reference backs up
// and restores lattice and RNG instead, tune.h preTune/postTune).  Results do
// not depend on the launch shape.
const Cfg kShapes[] = {{64, 1, 0},  {128, 1, 0}, {192, 1, 0}, {256, 1, 0}, {384, 1, 0},
                       {512, 1, 0}, {64, 2, 0},  {128, 2, 0}, {192, 2, 0}, {256, 2, 0},
                       {384, 2, 0}, {64, 4, 0},  {128, 4, 0}, {192, 4, 0}, {256, 4, 0}};
// every shape is tried with and without the XCD-aware workgroup order and,
// for a single lattice, with lexicographic or 2 x 4 tiled planes: the tiles
// cut the halo between XCDs while the lattice lives in the MALL, and have
// measured slower on a lattice streamed from HBM (96^4)
std::vector<Cfg> candidates(bool tiles) {
  std::vector<Cfg> v;
  for (int t = 0; t < (tiles ? 2 : 1); ++t)
    for (int x = 0; x < 2; ++x)
      for (Cfg c : kShapes) {
        c.xcd = x;
        c.tile = t;
        v.push_back(c);
      }
  return v;
}
const Cfg kDefaultCfg = {256, 1, 1, 0};

std::string vol_str(const u1_geom* g) {
  std::ostringstream s;
  s << g->dims[0] << "x" << g->dims[1] << "x" << g->dims[2] << "x" << g->dims[3];
  return s.str();
}
std::string tune_key(const char* name, const u1_geom* g, int prec) {
  return vol_str(g) + "\t" + name + (prec == U1_F64 ? "_f64" : "_f32");
}

int time_phase(int prec, int kind, void* lat, const PhaseArgs& a, int nb, Cfg c, float* ms) {
  EventSet es;
  hipEvent_t e0, e1;
  HIPCHK(es.create(&e0));
  HIPCHK(es.create(&e1));
  HIPCHK(u1d::launch_phase(prec, kind, 0, lat, a, nb, c.block, c.spt, G.st));  // warm-up
  HIPCHK(hipEventRecord(e0, G.st));
  for (int rep = 0; rep < 2; ++rep)
    for (int mu = 0; mu < 4; ++mu) HIPCHK(u1d::launch_phase(prec, kind, mu, lat, a, nb, c.block, c.spt, G.st));
  HIPCHK(hipEventRecord(e1, G.st));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipEventElapsedTime(ms, e0, e1));
  return U1_OK;
}

// Launch shape for the update phase kernel class `kind` on this geometry.
int phase_cfg(int prec, int kind, const u1_geom* g, const void* lat, const PhaseArgs& proto, Cfg* cfg) {
  const char* name = kind == u1d::KIND_METROPOLIS ? "Metropolis" : "OverRelaxation";
  const std::string key = tune_key(name, g, prec);
  auto it = G.tunecache.find(key);
  if (it != G.tunecache.end()) {
    *cfg = it->second;
    return U1_OK;
  }
  if (!G.tune) {
    *cfg = kDefaultCfg;
    return U1_OK;
  }
  const int64_t bytes = 4 * (int64_t)proto.single.V * (int64_t)elem_size(prec);
  DevBuf scratch;
  HIPCHK(scratch.alloc(bytes));
  HIPCHK(hipMemcpyAsync(scratch.p, lat, bytes, hipMemcpyDeviceToDevice, G.st));
  // Three rounds over all candidates, best time per candidate: a round-robin
  // order keeps clock ramp-up and neighbour noise from favouring one shape.
  std::vector<Cfg> cands;
  for (const Cfg c : candidates(true))
    if (!(c.spt == 4 && prec != U1_F32)) cands.push_back(c);
  std::vector<float> tbest(cands.size(), 1e30f);
  for (int round = 0; round < 3; ++round)
    for (size_t i = 0; i < cands.size(); ++i) {
      const Cfg c = cands[i];
      PhaseArgs a = proto;
      a.accepts = nullptr;
      a.meas = nullptr;
      a.xcd = (uint32_t)c.xcd;
      set_update_blocks(a.single, c);
      float ms = 0;
      RET(time_phase(prec, kind, scratch.p, a, a.single.nblk, c, &ms));
      tbest[i] = std::min(tbest[i], ms);
    }
  size_t ib = 0;
  for (size_t i = 1; i < cands.size(); ++i)
    if (tbest[i] < tbest[ib]) ib = i;
  if (G.verbosity >= 2)
    for (size_t i = 0; i < cands.size(); ++i)
      std::fprintf(stderr, "tune %s: block %d, %d site(s)/thread, %s order%s: %.2f us per launch%s\n", key.c_str(),
                   cands[i].block, cands[i].spt, cands[i].xcd ? "XCD" : "plain", cands[i].tile ? ", tiled planes" : "",
                   1e3 * tbest[i] / 8.0, i == ib ? "  <- chosen" : "");
  else if (G.verbosity >= 1)
    std::fprintf(stderr, "tune %s: block %d, %d site(s)/thread, %s order%s: %.2f us per launch\n", key.c_str(),
                 cands[ib].block, cands[ib].spt, cands[ib].xcd ? "XCD" : "plain", cands[ib].tile ? ", tiled planes" : "",
                 1e3 * tbest[ib] / 8.0);
  G.tunecache[key] = cands[ib];
  *cfg = cands[ib];
  return U1_OK;
}

// Fused pair shapes of a geometry: whole (z, t) planes per workgroup, SPT
// sites per thread per step, the workgroup size for one step over the plane
// and for two, rounded up to whole wavefronts and capped at the kernel's
// bound (larger planes take more steps).  With one site per thread, also
// 1.25 and 1.5 times the two-step size: a full first step and a shorter
// second one.
std::vector<Cfg> pair_candidates(int prec, const RepDesc& r) {
  const int P = r.nxh * r.N[1];
  std::vector<Cfg> v;
  auto add = [&v](int b, int spt) {
    for (const Cfg& c : v)
      if (c.block == b && c.spt == spt) return;
    for (int t = 0; t < 2; ++t)
      for (int x = 0; x < 2; ++x) v.push_back(Cfg{b, spt, x, t});
  };
  for (int spt : {1, 2, 3, 4}) {
    if (spt >= 3 && prec != U1_F32) continue;
    const int maxb = u1d::pair_max_block((int)elem_size(prec), spt);
    for (int steps = 1; steps <= 2; ++steps)
      add(std::min(maxb, ((P + spt * steps - 1) / (spt * steps) + 63) / 64 * 64), spt);
    if (spt == 1) {
      const int b2 = (P + 1) / 2;
      for (int q : {5, 6})  // b2 * q / 4
        if (b2 * q / 4 < maxb) add((b2 * q / 4 + 63) / 64 * 64, spt);
    }
  }
  return v;
}
bool pair_shape_ok(int prec, const Cfg& c) {
  if (!(c.spt == 1 || c.spt == 2 || ((c.spt == 3 || c.spt == 4) && prec == U1_F32))) return false;
  const int maxb = u1d::pair_max_block((int)elem_size(prec), c.spt);
  return c.block >= 64 && c.block <= maxb && c.block % 64 == 0;
}
// Plane decomposition of the fused pair: one chunk (workgroup) per plane.
void set_pair_blocks(RepDesc& r, const Cfg& c) {
  u1d::set_phase_blocks(r, r.nxh * r.N[1]);
  u1d::set_phase_tiles(r, c.tile != 0);
}

int time_pair(int prec, int kind, void* lat, const PhaseArgs& a, const Cfg& c, float* ms) {
  EventSet es;
  hipEvent_t e0, e1;
  HIPCHK(es.create(&e0));
  HIPCHK(es.create(&e1));
  const int plane = a.single.nxh * a.single.N[1];
  HIPCHK(u1d::launch_pair(prec, kind, lat, a, a.single.nblk, c.block, c.spt, plane, G.st));  // warm-up
  HIPCHK(hipEventRecord(e0, G.st));
  for (int rep = 0; rep < 4; ++rep)
    HIPCHK(u1d::launch_pair(prec, kind, lat, a, a.single.nblk, c.block, c.spt, plane, G.st));
  HIPCHK(hipEventRecord(e1, G.st));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipEventElapsedTime(ms, e0, e1));
  return U1_OK;
}

// Shape of the fused (parity, 0) + (parity, 1) launch for kernel class `kind`
// on this geometry, or block = 0 for two separate launches.  With tuning, the
// best pair shape is timed against the two phases in the single-phase shape
// `single` (tuned), on a scratch copy; the decision is cached with the shapes.
int pair_cfg(int prec, int kind, const u1_geom* g, const void* lat, const PhaseArgs& proto, const Cfg& single,
             Cfg* cfg) {
  *cfg = Cfg{0, 1, 0, 0};
  if (G.pair == 0) return U1_OK;
  if (proto.single.nxh * proto.single.N[1] > u1d::pair_max_plane((int)elem_size(prec)))
    return U1_OK;  // the plane's angles do not fit the pair kernel's LDS: two launches
  const char* name = kind == u1d::KIND_METROPOLIS ? "MetropolisPair" : "OverRelaxationPair";
  const std::string key = tune_key(name, g, prec);
  const std::vector<Cfg> cands = pair_candidates(prec, proto.single);
  // forced without a tuned pair shape: XCD order, lexicographic planes, 2
  // (binary32) or 1 (binary64) sites per thread
  auto forced_default = [&]() {
    for (const Cfg& c : cands)
      if (c.spt == (prec == U1_F32 ? 2 : 1) && c.xcd == 1 && c.tile == 0) *cfg = c;
  };
  auto it = G.tunecache.find(key);
  if (it != G.tunecache.end()) {  // block 0: the tuner measured two launches faster
    if (it->second.block != 0) *cfg = it->second;
    else if (G.pair == 2) forced_default();
    return U1_OK;
  }
  if (!G.tune) {
    if (G.pair == 2) forced_default();
    return U1_OK;
  }
  const int64_t bytes = 4 * (int64_t)proto.single.V * (int64_t)elem_size(prec);
  DevBuf scratch;
  HIPCHK(scratch.alloc(bytes));
  HIPCHK(hipMemcpyAsync(scratch.p, lat, bytes, hipMemcpyDeviceToDevice, G.st));
  std::vector<float> tbest(cands.size(), 1e30f);
  float tsep = 1e30f;  // phases (parity, 0) and (parity, 1) as two launches each
  for (int round = 0; round < 3; ++round) {
    for (size_t i = 0; i < cands.size(); ++i) {
      PhaseArgs a = proto;
      a.accepts = nullptr;
      a.meas = nullptr;
      a.xcd = (uint32_t)cands[i].xcd;
      set_pair_blocks(a.single, cands[i]);
      float ms = 0;
      RET(time_pair(prec, kind, scratch.p, a, cands[i], &ms));
      tbest[i] = std::min(tbest[i], ms);
    }
    PhaseArgs a = proto;
    a.accepts = nullptr;
    a.meas = nullptr;
    a.xcd = (uint32_t)single.xcd;
    set_update_blocks(a.single, single);
    EventSet es;
    hipEvent_t e0, e1;
    HIPCHK(es.create(&e0));
    HIPCHK(es.create(&e1));
    HIPCHK(hipEventRecord(e0, G.st));
    for (int rep = 0; rep < 4; ++rep)
      for (int mu = 0; mu < 2; ++mu)
        HIPCHK(u1d::launch_phase(prec, kind, mu, scratch.p, a, a.single.nblk, single.block, single.spt, G.st));
    HIPCHK(hipEventRecord(e1, G.st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    tsep = std::min(tsep, ms);
  }
  size_t ib = 0;
  for (size_t i = 1; i < cands.size(); ++i)
    if (tbest[i] < tbest[ib]) ib = i;
  const bool use = G.pair == 2 || tbest[ib] < tsep;
  if (G.verbosity >= 2)
    for (size_t i = 0; i < cands.size(); ++i)
      std::fprintf(stderr, "tune %s: block %d, %d site(s)/thread, %s order%s: %.2f us per pair launch%s\n",
                   key.c_str(), cands[i].block, cands[i].spt, cands[i].xcd ? "XCD" : "plain",
                   cands[i].tile ? ", tiled planes" : "", 1e3 * tbest[i] / 4.0, i == ib ? "  <- best" : "");
  if (G.verbosity >= 1)
    std::fprintf(stderr, "tune %s: pair %.2f us (block %d, %d site(s)/thread) vs two launches %.2f us: %s\n",
                 key.c_str(), 1e3 * tbest[ib] / 4.0, cands[ib].block, cands[ib].spt, 1e3 * tsep / 4.0,
                 use ? "fused" : "separate");
  const Cfg chosen = use ? cands[ib] : Cfg{0, 1, 0, 0};
  G.tunecache[key] = tbest[ib] < tsep ? cands[ib] : Cfg{0, 1, 0, 0};
  *cfg = chosen;
  return U1_OK;
}

// ------------------------------------------------------------ fused passes
// Reports, and clears, a wait timeout of a fused pass (k_chain) since the
// last check; called after the library stream has been synchronised.
int chain_check() {
  if (G.chain_err && __atomic_load_n(G.chain_err, __ATOMIC_ACQUIRE) != 0) {
    __atomic_store_n(G.chain_err, 0u, __ATOMIC_RELEASE);
    return fail(U1_ERR_HIP, "fused parity pass: a dependency wait timed out; the updated lattice is not reliable");
  }
  return U1_OK;
}

// Shapes of a fused pass: Cfg{block, spt, 1, tz, margin}, planes cut into
// chunks of block * spt sites, the (z, t) planes in tz x (8 / tz) tiles (one
// per XCD), dependencies at least margin + 1 schedule steps back.
void set_chain_blocks(RepDesc& r, const Cfg& c) { u1d::set_phase_blocks(r, c.block * c.spt); }
bool chain_shape_ok(const RepDesc& r0, const Cfg& c) {
  if (!(c.spt == 1 || c.spt == 2) || c.block < 64 || c.block > 512 || c.block % 64 != 0 || c.margin < 0 ||
      c.margin > kChainMaxMargin)
    return false;
  RepDesc r = r0;
  set_chain_blocks(r, c);
  return u1d::chain_possible(r.N[2], r.N[3], r.bpp, c.tile);
}
// Workgroup shapes (first tuning stage, default tiling and margin), then the
// tilings and margins of the best shape (second stage).
std::vector<Cfg> chain_shapes(const RepDesc& r) {
  std::vector<Cfg> v;
  for (int spt : {1, 2})
    for (int b : {64, 128, 192, 256, 384, 512}) {
      Cfg c{b, spt, 1, 2, 64};
      if (!chain_shape_ok(r, c)) c.tile = 1;
      if (chain_shape_ok(r, c)) v.push_back(c);
    }
  return v;
}
std::vector<Cfg> chain_layouts(const RepDesc& r, const Cfg& shape) {
  std::vector<Cfg> v;
  for (int tz : {1, 2, 4, 8})
    for (int margin : {16, 32, 64, 128, 256}) {
      Cfg c = shape;
      c.tile = tz;
      c.margin = margin;
      if (chain_shape_ok(r, c)) v.push_back(c);
    }
  return v;
}

// Schedule list, completion counters and counter targets of a fused pass over
// r (which carries set_chain_blocks) in workgroups of `block` threads;
// *nblocks: its workgroups.  Every wavefront adds 1 to its (mu, plane)
// counter, so a launch adds per = bpp * block / 64 to each.  The counters only
// grow: a launch waits for the value the previous launches with the same
// layout have brought them to plus per, and they are cleared (by a fill
// kernel on the library stream) when the layout changes or before they could
// overflow.
int chain_prepare(const RepDesc& r, const Cfg& cc, u1d::ChainArgs* c, int* nblocks) {
  const int nz = r.N[2], nt = r.N[3], bpp = r.bpp, per = r.bpp * (cc.block / 64), tz = cc.tile;
  const std::string key = std::to_string(nz) + " " + std::to_string(nt) + " " + std::to_string(bpp) + " " +
                          std::to_string(tz) + " " + std::to_string(cc.margin);
  auto it = G.chain_sched.find(key);
  if (it == G.chain_sched.end()) {
    const u1d::ChainSchedule s = u1d::chain_schedule(nz, nt, bpp, tz, cc.margin);
    uint32_t* d = nullptr;
    HIPCHK(hipMalloc((void**)&d, s.list.size() * sizeof(uint32_t)));
    const hipError_t e = hipMemcpy(d, s.list.data(), s.list.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return fail(U1_ERR_HIP, std::string("chain schedule upload: ") + hipGetErrorString(e));
    }
    it = G.chain_sched.emplace(key, d).first;
    G.chain_sched_len[key] = (int)s.list.size();
  }
  const size_t need = 4 * (size_t)nz * (size_t)nt;
  bool clear = nz != G.chain_nz || nt != G.chain_nt || per != G.chain_per;
  for (int mu = 0; mu < 4; ++mu) clear = clear || (uint64_t)(G.chain_inc[mu] + 2) * (uint64_t)per >= (1ull << 31);
  if (need > G.chain_cnt_cap) {
    HIPCHK(hipStreamSynchronize(G.st));  // no queued launch still uses the old counters
    if (G.chain_cnt) HIPCHK(hipFree(G.chain_cnt));
    G.chain_cnt = nullptr;
    G.chain_cnt_cap = 0;
    HIPCHK(hipMalloc((void**)&G.chain_cnt, need * sizeof(uint32_t)));
    G.chain_cnt_cap = need;
    clear = true;
  }
  if (!G.chain_err) {
    HIPCHK(hipHostMalloc((void**)&G.chain_err, sizeof(uint32_t), hipHostMallocMapped));
    *G.chain_err = 0;
  }
  if (clear) {
    HIPCHK(u1d::launch_fill_zero(G.chain_cnt, (int64_t)(need * sizeof(uint32_t)), G.st));
    G.chain_nz = nz;
    G.chain_nt = nt;
    G.chain_per = per;
    for (int mu = 0; mu < 4; ++mu) G.chain_inc[mu] = 0;
  }
  void* err_dev = nullptr;
  HIPCHK(hipHostGetDevicePointer(&err_dev, G.chain_err, 0));
  c->sched = it->second;
  c->cnt = G.chain_cnt;
  c->err = (uint32_t*)err_dev;
  for (int mu = 0; mu < 4; ++mu) c->tgt[mu] = (G.chain_inc[mu] + 1) * (uint32_t)per;
  c->tz_log2 = tz == 8 ? 3 : tz == 4 ? 2 : tz == 2 ? 1 : 0;
  c->bz = (uint32_t)(nz / tz);
  c->bt = (uint32_t)(nt / (8 / tz));
  const u1d::MagicDiv md = u1d::make_magic(c->bz);
  c->bz_m = md.m;
  c->bz_s = md.s;
  *nblocks = 8 * G.chain_sched_len[key];
  return U1_OK;
}

// One fused pass (phases (a.parity, 0..3)) in shape cc; a.meas as for
// launch_chain.
int chain_launch(int prec, int kind, void* lat, const PhaseArgs& a0, const Cfg& cc) {
  PhaseArgs a = a0;
  a.xcd = 0;
  set_chain_blocks(a.single, cc);
  u1d::ChainArgs c{};
  int nb = 0;
  RET(chain_prepare(a.single, cc, &c, &nb));
  HIPCHK(u1d::launch_chain(prec, kind, lat, a, c, nb, cc.block, cc.spt, G.st));
  for (int mu = 0; mu < 4; ++mu) ++G.chain_inc[mu];
  return U1_OK;
}

// The four phases of one parity as the unfused path launches them: the
// fused (parity, 0) + (parity, 1) pair if pc says so, then the phases.
int unfused_parity_pass(int prec, int kind, void* lat, const PhaseArgs& a0, const Cfg& c, const Cfg& pc) {
  PhaseArgs a = a0;
  for (int mu = 0; mu < 4; ++mu) {
    if (mu == 0 && pc.block != 0) {
      a.single = a0.single;
      a.xcd = (uint32_t)pc.xcd;
      set_pair_blocks(a.single, pc);
      HIPCHK(u1d::launch_pair(prec, kind, lat, a, a.single.nblk, pc.block, pc.spt, a.single.nxh * a.single.N[1], G.st));
      ++mu;
    } else {
      a.single = a0.single;
      a.xcd = (uint32_t)c.xcd;
      set_update_blocks(a.single, c);
      HIPCHK(u1d::launch_phase(prec, kind, mu, lat, a, a.single.nblk, c.block, c.spt, G.st));
    }
  }
  return U1_OK;
}

// Shape of the fused parity pass for kernel class `kind` on this geometry, or
// block = 0 for the unfused launches.  With tuning, every fused shape is timed
// against the unfused pass (tuned phase and pair shapes c, pc) on a scratch
// copy: four passes each (parities 0, 1, 0, 1); the decision is cached
// (MetropolisChain / OverRelaxationChain, block.x 0 = unfused).  Only for an
// unmasked single lattice whose (z, t) planes tile 2 x 4 (Nz even, Nt % 4 ==
// 0), and never during stream capture (the counter targets are per launch).
int chain_cfg(int prec, int kind, const u1_geom* g, const void* lat, const PhaseArgs& proto, const Cfg& c,
              const Cfg& pc, Cfg* cfg) {
  *cfg = Cfg{0, 1, 0, 0};
  if (G.chain == 0 || proto.ml != 0) return U1_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(G.st, &cs));
  if (cs != hipStreamCaptureStatusNone) return U1_OK;
  const std::vector<Cfg> shapes = chain_shapes(proto.single);
  if (shapes.empty()) return U1_OK;
  const char* name = kind == u1d::KIND_METROPOLIS ? "MetropolisChain" : "OverRelaxationChain";
  const std::string key = tune_key(name, g, prec);
  auto forced_default = [&]() {
    *cfg = shapes[0];
    for (const Cfg& k : shapes)
      if (k.block == 256 && k.spt == 1) *cfg = k;
  };
  auto it = G.tunecache.find(key);
  if (it != G.tunecache.end()) {  // block 0: the tuner measured the unfused launches faster
    if (it->second.block != 0 && chain_shape_ok(proto.single, it->second)) *cfg = it->second;
    else if (G.chain == 2) forced_default();
    return U1_OK;
  }
  if (!G.tune) {
    if (G.chain == 2) forced_default();
    return U1_OK;
  }
  const int64_t bytes = 4 * (int64_t)proto.single.V * (int64_t)elem_size(prec);
  DevBuf scratch;
  HIPCHK(scratch.alloc(bytes));
  HIPCHK(hipMemcpyAsync(scratch.p, lat, bytes, hipMemcpyDeviceToDevice, G.st));
  EventSet es;
  hipEvent_t e0, e1;
  HIPCHK(es.create(&e0));
  HIPCHK(es.create(&e1));
  PhaseArgs a = proto;
  a.accepts = nullptr;
  a.meas = nullptr;
  auto timed = [&](bool fused, const Cfg& cc, float* ms) -> int {
    auto one = [&](int i) -> int {
      a.parity = i & 1;
      if (fused) return chain_launch(prec, kind, scratch.p, a, cc);
      return unfused_parity_pass(prec, kind, scratch.p, a, c, pc);
    };
    RET(one(0));  // warm-up
    HIPCHK(hipEventRecord(e0, G.st));
    for (int i = 0; i < 4; ++i) RET(one(i));
    HIPCHK(hipEventRecord(e1, G.st));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    return chain_check();
  };
  // three rounds over the candidates, best time per candidate; the unfused
  // pass timed in every round too
  float tsep = 1e30f;
  auto search = [&](const std::vector<Cfg>& cands, std::vector<float>& tbest) -> int {
    tbest.assign(cands.size(), 1e30f);
    for (int round = 0; round < 3; ++round) {
      for (size_t i = 0; i < cands.size(); ++i) {
        float ms = 0;
        RET(timed(true, cands[i], &ms));
        tbest[i] = std::min(tbest[i], ms);
      }
      float ms = 0;
      RET(timed(false, Cfg{0, 1, 0, 0}, &ms));
      tsep = std::min(tsep, ms);
    }
    return U1_OK;
  };
  auto report = [&](const std::vector<Cfg>& cands, const std::vector<float>& tbest, size_t ib) {
    if (G.verbosity >= 2)
      for (size_t i = 0; i < cands.size(); ++i)
        std::fprintf(stderr, "tune %s: block %d, %d site(s)/thread, %d z tiles, margin %d: %.2f us per fused pass%s\n",
                     key.c_str(), cands[i].block, cands[i].spt, cands[i].tile, cands[i].margin,
                     1e3 * tbest[i] / 4.0, i == ib ? "  <- best" : "");
  };
  auto argmin = [](const std::vector<float>& t) {
    size_t ib = 0;
    for (size_t i = 1; i < t.size(); ++i)
      if (t[i] < t[ib]) ib = i;
    return ib;
  };
  std::vector<float> t1, t2;
  RET(search(shapes, t1));
  const size_t i1 = argmin(t1);
  report(shapes, t1, i1);
  const std::vector<Cfg> layouts = chain_layouts(proto.single, shapes[i1]);
  RET(search(layouts, t2));
  const size_t i2 = argmin(t2);
  report(layouts, t2, i2);
  const Cfg best = layouts[i2];
  // the search's best time is the minimum over many candidates and rounds,
  // biased low against the unfused pass's few: both are timed afresh, the
  // same number of times, alternating, before they are compared
  float tb = 1e30f;
  tsep = 1e30f;
  for (int round = 0; round < 5; ++round) {
    float ms = 0;
    RET(timed(true, best, &ms));
    tb = std::min(tb, ms);
    RET(timed(false, Cfg{0, 1, 0, 0}, &ms));
    tsep = std::min(tsep, ms);
  }
  // chosen only on a clear win: the two paths' timings move by a few per cent
  // from one tuning run to the next
  const float gain = kind == u1d::KIND_METROPOLIS ? kChainMinGainMetropolis : kChainMinGainOR;
  const bool faster = tb < gain * tsep;
  const bool use = G.chain == 2 || faster;
  if (G.verbosity >= 1)
    std::fprintf(stderr,
                 "tune %s: fused pass %.2f us (block %d, %d site(s)/thread, %d z tiles, margin %d) vs unfused %.2f "
                 "us: %s\n",
                 key.c_str(), 1e3 * tb / 4.0, best.block, best.spt, best.tile, best.margin, 1e3 * tsep / 4.0,
                 use ? "fused" : "unfused");
  G.tunecache[key] = faster ? best : Cfg{0, 1, 0, 0};
  *cfg = use ? best : Cfg{0, 1, 0, 0};
  return U1_OK;
}

// Settles (tunes, if enabled) every launch shape a sweep of this geometry
// uses, and the fused passes' schedules and counters, so that a later sweep
// allocates and synchronises nothing (graph capture, timed regions).
int settle_update_cfgs(int prec, const u1_geom* g, const void* lat, const PhaseArgs& proto, int metrop, int ovrn) {
  Cfg c, pc, cc;
  for (int kind = 0; kind < 2; ++kind) {
    if ((kind == u1d::KIND_METROPOLIS ? metrop : ovrn) <= 0) continue;
    RET(phase_cfg(prec, kind, g, lat, proto, &c));
    RET(pair_cfg(prec, kind, g, lat, proto, c, &pc));
    RET(chain_cfg(prec, kind, g, lat, proto, c, pc, &cc));
    if (cc.block != 0) {
      RepDesc r = proto.single;
      set_chain_blocks(r, cc);
      u1d::ChainArgs ca{};
      int nb = 0;
      RET(chain_prepare(r, cc, &ca, &nb));
    }
  }
  return U1_OK;
}

// One UpdateLattice on a single lattice (update.cu:449-456 loop order).
// meas_dst (device, 4 doubles, nullable): the plaquette of the updated
// lattice, in u1_plaquette's normalisation, from the measuring variants of
// phases (parity 1, mu = 1, 2, 3) of the sweep's last pass.
// sweep_dev (nullable): the Philox sweep index is *sweep_dev * sweep_mul +
// sweep, read by the kernels (for graph capture; phase_cfg must have run).
int run_update(const u1_geom* g, int prec, void* lat, int metrop, int ovrn, const u1_metro* m, uint64_t seed,
               uint32_t stream, uint32_t sweep, uint32_t c3, int ml, unsigned long long* acc_dev,
               double* meas_dst = nullptr, const uint32_t* sweep_dev = nullptr,
               uint32_t sweep_mul = 0) {
  // the Metropolis pass index takes 12 bits of Philox counter word 2
  if (metrop < 0 || ovrn < 0 || metrop > 4095) return fail(U1_ERR_ARG, "bad pass counts");
  PhaseArgs a{};
  a.single = make_rep(g, seed, c3);
  a.reps = nullptr;
  a.blockmap = nullptr;
  a.ml = ml;
  a.proposal = m->proposal;
  a.nhit = m->nhit;
  a.step = m->step;
  a.stream = stream;
  a.sweep = sweep;
  a.sweep_dev = sweep_dev;
  a.sweep_mul = sweep_mul;
  a.accepts = acc_dev;
  Cfg cm = kDefaultCfg, co = kDefaultCfg, pm{0, 1, 0, 0}, po{0, 1, 0, 0}, chm{0, 1, 0, 0}, cho{0, 1, 0, 0};
  if (metrop > 0) RET(phase_cfg(prec, u1d::KIND_METROPOLIS, g, lat, a, &cm));
  if (ovrn > 0) RET(phase_cfg(prec, u1d::KIND_OVERRELAX, g, lat, a, &co));
  if (metrop > 0) RET(pair_cfg(prec, u1d::KIND_METROPOLIS, g, lat, a, cm, &pm));
  if (ovrn > 0) RET(pair_cfg(prec, u1d::KIND_OVERRELAX, g, lat, a, co, &po));
  if (metrop > 0) RET(chain_cfg(prec, u1d::KIND_METROPOLIS, g, lat, a, cm, pm, &chm));
  if (ovrn > 0) RET(chain_cfg(prec, u1d::KIND_OVERRELAX, g, lat, a, co, po, &cho));
  RepDesc rm = a.single, ro = a.single, rpm = a.single, rpo = a.single;
  set_update_blocks(rm, cm);
  set_update_blocks(ro, co);
  set_pair_blocks(rpm, pm);
  set_pair_blocks(rpo, po);
  if (meas_dst && (ml != 0 || metrop + ovrn == 0)) return fail(U1_ERR_ARG, "no measuring pass");
  // measuring phases: parity 1, mu >= 1 of the last pass; rows of phase 1 (in
  // a fused pair: one per plane), then of phases 2 and 3 (a fused pass: its
  // own rows for phases 1-3, launch_chain)
  const bool meas_or = ovrn > 0;
  const Cfg& meas_chain = meas_or ? cho : chm;
  const int meas_nblk = meas_or ? ro.nblk : rm.nblk;
  const int meas_nblk1 = (meas_or ? po.block : pm.block) != 0 ? (meas_or ? rpo.nblk : rpm.nblk) : meas_nblk;
  int meas_rows_total = meas_nblk1 + 2 * meas_nblk;
  if (meas_chain.block != 0) {
    RepDesc rc = a.single;
    set_chain_blocks(rc, meas_chain);
    meas_rows_total = 3 * rc.N[2] * rc.N[3] * rc.bpp;
  }
  if (meas_dst) RET(ensure_fold(meas_rows_total, 4));
  auto meas_rows = [&](bool last_pass, int parity, int mu) -> double* {
    if (!meas_dst || !last_pass || parity != 1 || mu < 1) return nullptr;
    return G.partials + 4 * (mu == 1 ? 0 : (size_t)meas_nblk1 + (size_t)meas_nblk * (size_t)(mu - 2));
  };
  // one pass of kind `kind`: per parity, one fused pass (chc), or (parity, 0)
  // + (parity, 1) fused or not, then (parity, 2) and (parity, 3)
  auto pass = [&](int kind, uint32_t p, bool last, const Cfg& c, const RepDesc& r, const Cfg& pc,
                  const RepDesc& rp, const Cfg& chc) -> int {
    a.pass = p;
    for (int parity = 0; parity < 2; ++parity) {
      if (chc.block != 0) {
        a.parity = parity;
        a.meas = meas_rows(last, parity, 1);  // rows from 0: phases 1-3 of the fused pass
        RET(chain_launch(prec, kind, lat, a, chc));
        continue;
      }
      for (int mu = 0; mu < 4; ++mu) {
        const bool pair = mu == 0 && pc.block != 0;
        a.parity = parity;
        a.meas = meas_rows(last, parity, pair ? 1 : mu);
        if (pair) {
          a.single = rp;
          a.xcd = (uint32_t)pc.xcd;
          HIPCHK(u1d::launch_pair(prec, kind, lat, a, rp.nblk, pc.block, pc.spt, rp.nxh * rp.N[1], G.st));
          ++mu;
        } else {
          a.single = r;
          a.xcd = (uint32_t)c.xcd;
          HIPCHK(u1d::launch_phase(prec, kind, mu, lat, a, r.nblk, c.block, c.spt, G.st));
        }
      }
    }
    return U1_OK;
  };
  for (int p = 0; p < metrop; ++p)
    RET(pass(u1d::KIND_METROPOLIS, (uint32_t)p, ovrn == 0 && p == metrop - 1, cm, rm, pm, rpm, chm));
  for (int o = 0; o < ovrn; ++o) RET(pass(u1d::KIND_OVERRELAX, (uint32_t)o, o == ovrn - 1, co, ro, po, rpo, cho));
  if (meas_dst)
    HIPCHK(u1d::launch_finish_sums(G.partials, meas_rows_total, 4, 1.0 / (3.0 * (double)a.single.V), meas_dst,
                                   fold_scratch_ptr(meas_rows_total, 4), G.st));
  return U1_OK;
}

// Plaquette into a device result slot dst[4] (no synchronisation).
int plaquette_async(const u1_geom* g, int prec, const void* lat, double* dst) {
  RepDesc r = make_rep(g, 0, 0);
  int block, spt;
  plaquette_shape(r, prec, &block, &spt);
  set_blocks(r, block * spt);
  r.nblk *= 2;  // even sites, then odd sites
  const int nblk = r.nblk;
  RET(ensure_fold(nblk, 4));
  HIPCHK(u1d::launch_plaquette(prec, lat, r, G.partials, nblk, block, spt, G.st));
  HIPCHK(u1d::launch_finish_sums(G.partials, nblk, 4, 1.0 / (3.0 * (double)r.V), dst, fold_scratch_ptr(nblk, 4),
                                 G.st));
  return U1_OK;
}

// Polyakov-loop volume P (polyakov.cu:415-508) into P (2 SV doubles, device);
// the multihit modes first write the temporal-link field into tl (2 V
// doubles), then take the product along t.
int poly_volume_async(int prec, const void* lat, const RepDesc& r, int mode, const u1_metro* mh, uint32_t stream,
                      uint32_t c1, uint32_t c3, double* P, double* tl) {
  if (mode == U1_POLY_PLAIN) {
    HIPCHK(u1d::launch_poly_volume(prec, lat, r, mode, 0, 0, 0.0, stream, c1, c3, P, 256, G.st));
    return U1_OK;
  }
  HIPCHK(u1d::launch_temporal_links(prec, lat, r, mode, mh ? mh->proposal : 0, mh ? mh->nhit : 0,
                                    mh ? mh->step : 0.0, stream, c1, c3, tl, 256, G.st));
  HIPCHK(u1d::launch_tprod(tl, r, P, G.st));
  return U1_OK;
}

int copy_back(double* host, const double* dev, size_t n) {
  HIPCHK(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, G.st));
  HIPCHK(hipStreamSynchronize(G.st));
  return chain_check();
}

}  // namespace

struct u1_ensemble {
  int32_t nrep = 0;
  int prec = U1_F32;
  int32_t metrop = 1, ovrn = 3;
  u1_metro m{};
  std::vector<u1_geom> geoms;
  std::vector<RepDesc> upd_host, meas_host;  // per-replica descriptors (update / measurement launches)
  RepDesc* upd_dev = nullptr;
  RepDesc* meas_dev = nullptr;
  uint16_t* upd_map = nullptr;
  uint16_t* meas_map = nullptr;
  int upd_blocks = 0, meas_blocks = 0;
  void* lat = nullptr;
  int64_t total_links = 0;
  uint32_t* sweep_dev = nullptr;
  uint64_t sweeps_done = 0;
  double* res_dev = nullptr;
  double* partials = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // sweep with the plaquette summed in its last pass (u1_ensemble_sweep_plaquette)
  double* mrows = nullptr;      // 3 segments x upd_blocks rows x 4 doubles
  int mrows_blocks = 0;         // upd_blocks that mrows was sized for
  double* series_dev = nullptr;  // kEnsSeriesCap sweeps x nrep x 4 doubles
  uint32_t* base_dev = nullptr;  // sweep index of series row 0
  hipGraph_t mgraph = nullptr;
  hipGraphExec_t mexec = nullptr;
  Cfg cfg = {256, 1, 1, 0};  // launch shape of the update phases (tuned on first use)
  bool tuned = false;
  // fused (parity, 0) + (parity, 1) launches (block 0: none): per-replica
  // plane tables with one workgroup per (z, t) plane
  Cfg pcfg = {0, 1, 0, 0};
  RepDesc* pair_dev = nullptr;
  uint16_t* pair_map = nullptr;
  int pair_blocks = 0;
  int pair_plane = 0;  // largest plane (Nx/2 * Ny sites) of the replicas: the pair kernel's LDS
};

namespace {

const int kEnsSeriesCap = 256;  // sweeps per device-side plaquette series chunk
const int kEnsBlock = 64;  // measurement launches: workgroup size and sites per thread
const int kEnsSpt = 1;     // (small plane chunks suit the small replicas of a batch)

// (Re)build the update-launch tables for launch shape c: per-replica plane
// decomposition and first workgroup, and the workgroup -> replica map.
int ensemble_build_update(u1_ensemble* e, Cfg c) {
  std::vector<uint16_t> umap;
  int nb = 0;
  for (int i = 0; i < e->nrep; ++i) {
    RepDesc& u = e->upd_host[i];
    set_blocks(u, c.block * c.spt);
    u.blk_begin = nb;
    nb += u.nblk;
    for (int b = 0; b < u.nblk; ++b) umap.push_back((uint16_t)i);
  }
  if (e->upd_map) HIPCHK(hipFree(e->upd_map));
  e->upd_map = nullptr;
  HIPCHK(hipMalloc((void**)&e->upd_map, umap.size() * sizeof(uint16_t)));
  HIPCHK(hipMemcpyAsync(e->upd_map, umap.data(), umap.size() * sizeof(uint16_t), hipMemcpyHostToDevice, G.st));
  HIPCHK(hipMemcpyAsync(e->upd_dev, e->upd_host.data(), e->nrep * sizeof(RepDesc), hipMemcpyHostToDevice, G.st));
  HIPCHK(hipStreamSynchronize(G.st));
  e->upd_blocks = nb;
  e->cfg = c;
  return U1_OK;
}

int ensemble_tune(u1_ensemble* e);
int ensemble_build_pair(u1_ensemble* e, Cfg c);
// Settles the ensemble's launch shapes before either graph exists (they fix
// the launch layout): tuned, or, with tuning off and fused pairs forced
// (u1_set_pair_fusion(2)), a default pair shape.
int ensemble_settle(u1_ensemble* e) {
  if (e->tuned || e->exec || e->mexec) return U1_OK;
  if (G.tune) return ensemble_tune(e);
  if (G.pair == 2 && e->pcfg.block == 0) return ensemble_build_pair(e, Cfg{256, 1, 1, 0});
  return U1_OK;
}

// Per-replica tables of the fused pair launches for shape c (block 0: none).
int ensemble_build_pair(u1_ensemble* e, Cfg c) {
  e->pcfg = Cfg{0, 1, 0, 0};  // set to c only once its tables are in place
  if (c.block == 0) return U1_OK;
  int plane = 0;
  for (const RepDesc& r : e->upd_host) plane = std::max(plane, r.nxh * r.N[1]);
  if (plane > u1d::pair_max_plane((int)elem_size(e->prec))) return U1_OK;  // planes too large for the LDS exchange
  std::vector<RepDesc> reps(e->upd_host);
  std::vector<uint16_t> pmap;
  int nb = 0;
  for (int i = 0; i < e->nrep; ++i) {
    RepDesc& p = reps[i];
    u1d::set_phase_blocks(p, p.nxh * p.N[1]);  // one chunk per plane
    p.tile = 0;
    p.blk_begin = nb;
    nb += p.nblk;
    for (int b = 0; b < p.nblk; ++b) pmap.push_back((uint16_t)i);
  }
  if (!e->pair_dev) HIPCHK(hipMalloc((void**)&e->pair_dev, e->nrep * sizeof(RepDesc)));
  if (e->pair_map) HIPCHK(hipFree(e->pair_map));
  e->pair_map = nullptr;
  HIPCHK(hipMalloc((void**)&e->pair_map, pmap.size() * sizeof(uint16_t)));
  HIPCHK(hipMemcpyAsync(e->pair_map, pmap.data(), pmap.size() * sizeof(uint16_t), hipMemcpyHostToDevice, G.st));
  HIPCHK(hipMemcpyAsync(e->pair_dev, reps.data(), e->nrep * sizeof(RepDesc), hipMemcpyHostToDevice, G.st));
  HIPCHK(hipStreamSynchronize(G.st));
  e->pair_blocks = nb;
  e->pair_plane = plane;
  e->pcfg = c;
  return U1_OK;
}

// mrows (nullable): plaquette rows of the measuring phases (parity 1, mu >= 1
// of the last pass), segment mu - 1 at row offset (mu - 1) * upd_blocks.
// With fused pairs (e->pcfg), the phases (parity, 0) + (parity, 1) of every
// pass run as one launch, except the measuring parity 1 of the last pass.
int ensemble_enqueue_phases(u1_ensemble* e, void* lat, double* mrows = nullptr) {
  auto meas_rows = [&](bool last_pass, int parity, int mu) -> double* {
    if (!mrows || !last_pass || parity != 1 || mu < 1) return nullptr;
    return mrows + 4 * (size_t)e->upd_blocks * (size_t)(mu - 1);
  };
  PhaseArgs a{};
  a.reps = e->upd_dev;
  a.blockmap = e->upd_map;
  a.ml = 0;
  a.proposal = e->m.proposal;
  a.nhit = e->m.nhit;
  a.step = e->m.step;
  a.stream = u1r::STREAM_UPDATE;
  a.sweep = 0;
  a.sweep_dev = e->sweep_dev;
  a.sweep_mul = 1;
  a.accepts = nullptr;
  a.xcd = (uint32_t)e->cfg.xcd;
  const bool fused = e->pcfg.block != 0;
  auto pass = [&](int kind, uint32_t p, bool last) -> int {
    a.pass = p;
    for (int parity = 0; parity < 2; ++parity)
      for (int mu = 0; mu < 4; ++mu) {
        a.parity = parity;
        if (mu == 0 && fused && !meas_rows(last, parity, 1)) {
          PhaseArgs b = a;
          b.reps = e->pair_dev;
          b.blockmap = e->pair_map;
          b.meas = nullptr;
          b.xcd = (uint32_t)e->pcfg.xcd;
          HIPCHK(u1d::launch_pair(e->prec, kind, lat, b, e->pair_blocks, e->pcfg.block, e->pcfg.spt, e->pair_plane,
                                  G.st));
          ++mu;
          continue;
        }
        a.meas = meas_rows(last, parity, mu);
        HIPCHK(u1d::launch_phase(e->prec, kind, mu, lat, a, e->upd_blocks, e->cfg.block, e->cfg.spt, G.st));
      }
    return U1_OK;
  };
  for (int p = 0; p < e->metrop; ++p)
    RET(pass(u1d::KIND_METROPOLIS, (uint32_t)p, e->ovrn == 0 && p == e->metrop - 1));
  for (int o = 0; o < e->ovrn; ++o) RET(pass(u1d::KIND_OVERRELAX, (uint32_t)o, o == e->ovrn - 1));
  return U1_OK;
}

int ensemble_enqueue_sweep(u1_ensemble* e) {
  RET(ensemble_enqueue_phases(e, e->lat));
  HIPCHK(u1d::launch_increment(e->sweep_dev, G.st));
  return U1_OK;
}

// Time every launch-shape candidate over one full sweep of all replicas on a
// scratch copy of the packed lattice; keep the fastest.
int ensemble_tune(u1_ensemble* e) {
  const size_t bytes = (size_t)e->total_links * elem_size(e->prec);
  DevBuf scratch;
  HIPCHK(scratch.alloc(bytes));
  HIPCHK(hipMemcpyAsync(scratch.p, e->lat, bytes, hipMemcpyDeviceToDevice, G.st));
  EventSet es;
  hipEvent_t e0, e1;
  HIPCHK(es.create(&e0));
  HIPCHK(es.create(&e1));
  float best = 1e30f;
  Cfg bestc = e->cfg;
  for (const Cfg c : candidates(false)) {  // batched launches keep lexicographic planes
    if (c.spt == 4 && e->prec != U1_F32) continue;
    RET(ensemble_build_update(e, c));
    RET(ensemble_enqueue_phases(e, scratch.p));  // warm-up
    HIPCHK(hipEventRecord(e0, G.st));
    RET(ensemble_enqueue_phases(e, scratch.p));
    HIPCHK(hipEventRecord(e1, G.st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      bestc = c;
    }
  }
  RET(ensemble_build_update(e, bestc));
  // fused pairs: a few whole-plane shapes over full sweeps, against none
  if (G.pair != 0) {
    const Cfg none{0, 1, 0, 0};
    float tnone = best, tp = 1e30f;
    Cfg bestp = none;
    for (const Cfg pc : {Cfg{64, 1, 1, 0}, Cfg{128, 1, 1, 0}, Cfg{256, 1, 1, 0}, Cfg{512, 1, 1, 0}, Cfg{128, 2, 1, 0},
                         Cfg{256, 2, 1, 0}, Cfg{128, 1, 0, 0}, Cfg{256, 1, 0, 0}}) {
      RET(ensemble_build_pair(e, pc));
      RET(ensemble_enqueue_phases(e, scratch.p));  // warm-up
      HIPCHK(hipEventRecord(e0, G.st));
      RET(ensemble_enqueue_phases(e, scratch.p));
      HIPCHK(hipEventRecord(e1, G.st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < tp) {
        tp = ms;
        bestp = pc;
      }
    }
    const bool use = G.pair == 2 || tp < tnone;
    if (G.verbosity >= 1)
      std::fprintf(stderr, "tune ensemble (%d replicas): fused pairs %.1f us per sweep (block %d, %d site(s)/thread) "
                   "vs %.1f us without: %s\n", e->nrep, 1e3 * tp, bestp.block, bestp.spt, 1e3 * tnone,
                   use ? "fused" : "separate");
    RET(ensemble_build_pair(e, use ? bestp : none));
  }
  e->tuned = true;
  return U1_OK;
}

void ensemble_release(u1_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->mexec) (void)hipGraphExecDestroy(e->mexec);
  if (e->mgraph) (void)hipGraphDestroy(e->mgraph);
  if (e->mrows) (void)hipFree(e->mrows);
  if (e->series_dev) (void)hipFree(e->series_dev);
  if (e->base_dev) (void)hipFree(e->base_dev);
  if (e->upd_dev) (void)hipFree(e->upd_dev);
  if (e->meas_dev) (void)hipFree(e->meas_dev);
  if (e->upd_map) (void)hipFree(e->upd_map);
  if (e->pair_dev) (void)hipFree(e->pair_dev);
  if (e->pair_map) (void)hipFree(e->pair_map);
  if (e->meas_map) (void)hipFree(e->meas_map);
  if (e->lat) (void)hipFree(e->lat);
  if (e->sweep_dev) (void)hipFree(e->sweep_dev);
  if (e->res_dev) (void)hipFree(e->res_dev);
  if (e->partials) (void)hipFree(e->partials);
}

// The measuring sweep graph (once per ensemble): phases with the plaquette
// summed in the last pass, the per-replica fold into the series row of the
// sweep, sweep index + 1.
int ensemble_build_mgraph(u1_ensemble* e) {
  if (e->mexec) return U1_OK;
  // (kept across a failed capture and reused by the next attempt, unless a
  // later tuning has given the update launches more workgroups)
  if (e->mrows && e->mrows_blocks < e->upd_blocks) {
    HIPCHK(hipFree(e->mrows));
    e->mrows = nullptr;
  }
  if (!e->mrows) {
    HIPCHK(hipMalloc((void**)&e->mrows, 3 * 4 * (size_t)e->upd_blocks * sizeof(double)));
    e->mrows_blocks = e->upd_blocks;
  }
  if (!e->series_dev)
    HIPCHK(hipMalloc((void**)&e->series_dev, (size_t)kEnsSeriesCap * e->nrep * 4 * sizeof(double)));
  if (!e->base_dev) HIPCHK(hipMalloc((void**)&e->base_dev, sizeof(uint32_t)));
  HIPCHK(hipStreamBeginCapture(G.st, hipStreamCaptureModeThreadLocal));
  int rc = ensemble_enqueue_phases(e, e->lat, e->mrows);
  if (rc == U1_OK) {
    hipError_t f = u1d::launch_ensemble_plaquette_fold(e->mrows, 3, e->upd_blocks, e->upd_dev, e->nrep, e->sweep_dev,
                                                       e->base_dev, kEnsSeriesCap, e->series_dev, G.st);
    if (f == hipSuccess) f = u1d::launch_increment(e->sweep_dev, G.st);
    if (f != hipSuccess) rc = fail(U1_ERR_HIP, std::string("ensemble capture: ") + hipGetErrorString(f));
  }
  hipGraph_t gr = nullptr;
  const hipError_t ce = hipStreamEndCapture(G.st, &gr);
  if (rc != U1_OK) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  HIPCHK(ce);
  e->mgraph = gr;
  HIPCHK(hipGraphInstantiate(&e->mexec, e->mgraph, nullptr, nullptr, 0));
  return U1_OK;
}

// nsweeps replays of the measuring graph, the series copied to the host in
// chunks of kEnsSeriesCap sweeps.  done_ev (optional) is recorded on the
// library stream after the last sweep's work, before the last chunk's copy.
int ensemble_measured_sweeps(u1_ensemble* e, int32_t nsweeps, double* series, hipEvent_t done_ev = nullptr) {
  for (int32_t done = 0; done < nsweeps;) {
    const int32_t n = std::min<int32_t>(kEnsSeriesCap, nsweeps - done);
    HIPCHK(hipMemcpyAsync(e->base_dev, e->sweep_dev, sizeof(uint32_t), hipMemcpyDeviceToDevice, G.st));
    for (int32_t s = 0; s < n; ++s) HIPCHK(hipGraphLaunch(e->mexec, G.st));
    if (done_ev && done + n == nsweeps) HIPCHK(hipEventRecord(done_ev, G.st));
    HIPCHK(hipMemcpyAsync(series + 4 * (size_t)e->nrep * done, e->series_dev, 4 * (size_t)e->nrep * n * sizeof(double),
                          hipMemcpyDeviceToHost, G.st));
    HIPCHK(hipStreamSynchronize(G.st));
    done += n;
    e->sweeps_done += (uint64_t)n;
  }
  return U1_OK;
}

// Launch classes of a sweep, for kernel timing.
enum LaunchClass {
  LC_METRO = 0,
  LC_OR,
  LC_OR_MEAS,
  LC_METRO_PAIR,
  LC_OR_PAIR,
  LC_OR_PAIR_MEAS,
  LC_METRO_CHAIN,   // fused parity passes (single lattices only)
  LC_OR_CHAIN,
  LC_OR_CHAIN_MEAS,
  LC_COUNT
};
constexpr int kTimingLaunches = 32;  // back-to-back launches per class

// Kernel timing of the ensemble's launch classes (shapes settled, ovrn >= 1):
// on a scratch copy of the packed lattice, kTimingLaunches back-to-back
// launches of each class over all replicas, in the sweep's launch tables, with
// (parity, mu) cycling over the phases the class covers in a pass.  ms[c]:
// mean launch duration; per_sweep[c]: launches per measuring sweep (the last
// overrelaxation pass measures in its parity-1 launches of mu >= 1 and, with
// fused pairs, leaves its parity-1 pair unfused).
int time_ensemble_classes(u1_ensemble* e, double* ms, double* per_sweep) {
  const size_t bytes = (size_t)e->total_links * elem_size(e->prec);
  DevBuf scratch, rows;
  HIPCHK(scratch.alloc(bytes));
  HIPCHK(rows.alloc(4 * (size_t)e->upd_blocks * sizeof(double)));
  HIPCHK(hipMemcpyAsync(scratch.p, e->lat, bytes, hipMemcpyDeviceToDevice, G.st));
  EventSet es;
  hipEvent_t e0, e1;
  HIPCHK(es.create(&e0));
  HIPCHK(es.create(&e1));
  PhaseArgs a{};
  a.reps = e->upd_dev;
  a.blockmap = e->upd_map;
  a.proposal = e->m.proposal;
  a.nhit = e->m.nhit;
  a.step = e->m.step;
  a.stream = u1r::STREAM_UPDATE;
  a.sweep = (uint32_t)e->sweeps_done;
  a.xcd = (uint32_t)e->cfg.xcd;
  const bool fused = e->pcfg.block != 0;
  for (int c = 0; c < LC_COUNT; ++c) ms[c] = per_sweep[c] = 0.0;
  for (int c = 0; c < LC_COUNT; ++c) {
    if (c == LC_OR_PAIR_MEAS || c >= LC_METRO_CHAIN) continue;  // the measuring parity-1 pair stays unfused
    const bool pair = c == LC_METRO_PAIR || c == LC_OR_PAIR;
    const int kind = (c == LC_METRO || c == LC_METRO_PAIR) ? u1d::KIND_METROPOLIS : u1d::KIND_OVERRELAX;
    const int passes = kind == u1d::KIND_METROPOLIS ? e->metrop : e->ovrn;
    const bool cmeas = c == LC_OR_MEAS, measured_kind = kind == u1d::KIND_OVERRELAX;
    if (passes == 0 || (pair && !fused)) continue;
    std::vector<std::pair<int, int>> ph;  // (parity, mu); a pair is listed by mu 0
    for (int parity = 0; parity < 2; ++parity)
      for (int mu = 0; mu < 4; ++mu) {
        if (cmeas ? !(parity == 1 && mu >= 1) : (pair ? mu != 0 : (fused && mu < 2))) continue;
        ph.push_back({parity, mu});
      }
    if (cmeas)
      per_sweep[c] = 3;
    else if (pair)
      per_sweep[c] = 2.0 * passes - (measured_kind ? 1 : 0);
    else
      per_sweep[c] = (fused ? 4.0 : 8.0) * passes - (measured_kind ? (fused ? 1 : 3) : 0);
    PhaseArgs b = a;
    if (pair) {
      b.reps = e->pair_dev;
      b.blockmap = e->pair_map;
      b.xcd = (uint32_t)e->pcfg.xcd;
    }
    b.meas = cmeas ? rows.as<double>() : nullptr;
    auto launch = [&](int i) -> int {
      b.parity = ph[(size_t)i % ph.size()].first;
      b.pass = (uint32_t)(i / (int)ph.size());
      if (pair)
        HIPCHK(u1d::launch_pair(e->prec, kind, scratch.p, b, e->pair_blocks, e->pcfg.block, e->pcfg.spt,
                                e->pair_plane, G.st));
      else
        HIPCHK(u1d::launch_phase(e->prec, kind, ph[(size_t)i % ph.size()].second, scratch.p, b, e->upd_blocks,
                                 e->cfg.block, e->cfg.spt, G.st));
      return U1_OK;
    };
    RET(launch(0));  // warm-up
    HIPCHK(hipEventRecord(e0, G.st));
    for (int i = 0; i < kTimingLaunches; ++i) RET(launch(i));
    HIPCHK(hipEventRecord(e1, G.st));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    ms[c] = t / kTimingLaunches;
  }
  return U1_OK;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- runtime
// Start (cuda_error_check.cpp:24-45): select the device, create the stream.
int u1_init(int device) {
  if (G.init) return U1_OK;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (n == 0) return fail(U1_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(U1_ERR_ARG, "GPU id out of range");
  HIPCHK(hipSetDevice(device));
  // non-blocking: no implicit synchronisation with the legacy default
  // stream, so default-stream work of other code in the process (another
  // thread, a profiler, a framework) neither waits on the library's launches
  // nor lands inside one of its stream captures
  HIPCHK(hipStreamCreateWithFlags(&G.st, hipStreamNonBlocking));
  HIPCHK(hipMalloc((void**)&G.counter, sizeof(unsigned long long)));
  HIPCHK(hipMalloc((void**)&G.dres, 64 * sizeof(double)));
  if (device < kMaxDevices && !g_scratch_dev[device]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = (double*)p;
  }
  G.device = device;
  G.init = true;
  return U1_OK;
}

// Finalize (cuda_error_check.cpp:47-72) without the exit(): sync and release.
int u1_finalize(void) {
  if (!G.init) return U1_OK;
  HIPCHK(hipStreamSynchronize(G.st));
  if (G.partials) HIPCHK(hipFree(G.partials));
  if (G.work) HIPCHK(hipFree(G.work));
  G.work = nullptr;
  G.work_cap = 0;
  if (G.counter) HIPCHK(hipFree(G.counter));
  if (G.dres) HIPCHK(hipFree(G.dres));
  for (auto& kv : G.chain_sched) HIPCHK(hipFree(kv.second));
  G.chain_sched.clear();
  G.chain_sched_len.clear();
  if (G.chain_cnt) HIPCHK(hipFree(G.chain_cnt));
  if (G.chain_err) HIPCHK(hipHostFree(G.chain_err));
  G.chain_cnt = nullptr;
  G.chain_cnt_cap = 0;
  G.chain_err = nullptr;
  G.chain_nz = G.chain_nt = G.chain_per = 0;
  HIPCHK(hipStreamDestroy(G.st));
  G.partials = nullptr;
  G.partials_cap = 0;
  G.counter = nullptr;
  G.dres = nullptr;
  G.st = nullptr;
  G.init = false;
  return U1_OK;
}

int u1_sync(void) {
  REQ_INIT();
  HIPCHK(hipStreamSynchronize(G.st));
  return chain_check();
}

const char* u1_last_error(void) { return g_err.c_str(); }

int u1_device_name(char* buf, size_t len) {
  REQ_INIT();
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, G.device));
  std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
  return U1_OK;
}

// ---------------------------------------------------------------- memory
int u1_malloc(void** ptr, size_t bytes) {
  REQ_INIT();
  if (!ptr) return fail(U1_ERR_ARG, "null pointer");
  *ptr = nullptr;
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 4));
  HIPCHK(u1d::launch_fill_zero(*ptr, (int64_t)((bytes + 3) / 4 * 4), G.st));
  HIPCHK(hipStreamSynchronize(G.st));
  return U1_OK;
}

int u1_free(void* ptr) {
  REQ_INIT();
  if (!ptr) return U1_OK;
  HIPCHK(hipStreamSynchronize(G.st));
  HIPCHK(hipFree(ptr));
  return U1_OK;
}

int u1_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  REQ_INIT();
  if ((!dst || !src) && bytes) return fail(U1_ERR_ARG, "null pointer");
  static const hipMemcpyKind kinds[4] = {hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                                         hipMemcpyDeviceToDevice};
  if (kind < 0 || kind > 3) return fail(U1_ERR_ARG, "bad copy kind");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, kinds[kind], G.st));
  HIPCHK(hipStreamSynchronize(G.st));
  return U1_OK;
}

int u1_zero(void* ptr, size_t bytes) {
  REQ_INIT();
  if (bytes) REQ_PTR(ptr);
  if (bytes % 4) return fail(U1_ERR_ARG, "u1_zero needs a multiple of 4 bytes");
  HIPCHK(u1d::launch_fill_zero(ptr, (int64_t)bytes, G.st));
  return U1_OK;
}

int u1_convert_precision(const void* src, int src_prec, void* dst, int dst_prec, int64_t n) {
  REQ_INIT();
  if (n < 0) return fail(U1_ERR_ARG, "negative element count");
  if (n) REQ_PTR(src);
  if (n) REQ_PTR(dst);
  RET(check_prec(src_prec));
  RET(check_prec(dst_prec));
  HIPCHK(u1d::launch_convert_precision(src, src_prec, dst, dst_prec, n, 256, G.st));
  return U1_OK;
}

// ---------------------------------------------------------------- updates
int u1_hotstart(const u1_geom* g, int prec, void* lat, uint64_t seed, uint32_t replica) {
  REQ_INIT();
  REQ_PTR(lat);
  RET(check_geom(g));
  RET(check_prec(prec));
  const RepDesc r = make_rep(g, seed, replica);
  HIPCHK(u1d::launch_hotstart(prec, lat, r, 256, G.st));
  return U1_OK;
}

int u1_update(const u1_geom* g, int prec, void* lat, int32_t metrop, int32_t ovrn, const u1_metro* m,
              uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t c3, int32_t mask_ml, uint64_t* accepts) {
  REQ_INIT();
  RET(check_geom(g));
  RET(check_prec(prec));
  RET(check_metro(m));
  if (!lat) return fail(U1_ERR_ARG, "null lattice");
  if (metrop < 0 || ovrn < 0 || metrop > 4095) return fail(U1_ERR_ARG, "bad pass counts");
  if (mask_ml != 0 && mask_ml != 2 && mask_ml != 4) return fail(U1_ERR_ARG, "mask_ml must be 0, 2 or 4");
  if (accepts) HIPCHK(u1d::launch_fill_zero(G.counter, sizeof(unsigned long long), G.st));
  RET(run_update(g, prec, lat, metrop, ovrn, m, seed, stream, sweep, c3, mask_ml, accepts ? G.counter : nullptr));
  if (accepts) {
    unsigned long long v = 0;
    HIPCHK(hipMemcpyAsync(&v, G.counter, sizeof(v), hipMemcpyDeviceToHost, G.st));
    HIPCHK(hipStreamSynchronize(G.st));
    RET(chain_check());
    *accepts += v;
  }
  return U1_OK;
}

int u1_update_plaquette(const u1_geom* g, int prec, void* lat, int32_t metrop, int32_t ovrn, const u1_metro* m,
                        uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t c3, double out[4]) {
  REQ_INIT();
  RET(check_geom(g));
  RET(check_prec(prec));
  RET(check_metro(m));
  if (!lat || !out) return fail(U1_ERR_ARG, "null argument");
  if (metrop < 0 || ovrn < 0 || metrop > 4095 || metrop + ovrn == 0) return fail(U1_ERR_ARG, "bad pass counts");
  RET(run_update(g, prec, lat, metrop, ovrn, m, seed, stream, sweep, c3, 0, nullptr, G.dres));
  return copy_back(out, G.dres, 4);
}

int u1_phase(const u1_geom* g, int prec, void* lat, int32_t kind, int32_t parity, int32_t mu, const u1_metro* m,
             uint64_t seed, uint32_t stream, uint32_t sweep, uint32_t pass, uint32_t c3, int32_t mask_ml) {
  REQ_INIT();
  RET(check_geom(g));
  RET(check_prec(prec));
  RET(check_metro(m));
  if (!lat) return fail(U1_ERR_ARG, "null lattice");
  if (kind != 0 && kind != 1) return fail(U1_ERR_ARG, "kind must be 0 (Metropolis) or 1 (overrelaxation)");
  if (parity < 0 || parity > 1 || mu < 0 || mu > 3) return fail(U1_ERR_ARG, "bad parity or direction");
  if (mask_ml != 0 && mask_ml != 2 && mask_ml != 4) return fail(U1_ERR_ARG, "mask_ml must be 0, 2 or 4");
  PhaseArgs a{};
  a.single = make_rep(g, seed, c3);
  a.ml = mask_ml;
  a.proposal = m->proposal;
  a.nhit = m->nhit;
  a.step = m->step;
  a.stream = stream;
  a.sweep = sweep;
  a.pass = pass;
  a.parity = parity;
  Cfg c = kDefaultCfg;
  RET(phase_cfg(prec, kind, g, lat, a, &c));
  set_update_blocks(a.single, c);
  a.xcd = (uint32_t)c.xcd;
  HIPCHK(u1d::launch_phase(prec, kind, mu, lat, a, a.single.nblk, c.block, c.spt, G.st));
  return U1_OK;
}

// ---------------------------------------------------------------- measurements
int u1_plaquette(const u1_geom* g, int prec, const void* lat, double out[4]) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(out);
  RET(check_geom(g));
  RET(check_prec(prec));
  RET(plaquette_async(g, prec, lat, G.dres));
  return copy_back(out, G.dres, 4);
}

int u1_plaquette_noalloc(const u1_geom* g, int prec, const void* lat, double* out_dev) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(out_dev);
  RET(check_geom(g));
  RET(check_prec(prec));
  double* rows = G.device < kMaxDevices ? g_scratch_dev[G.device] : nullptr;
  if (!rows) return fail(U1_ERR_STATE, "no-alloc scratch is not available on this device");
  RepDesc r = make_rep(g, 0, 0);
  int block, spt;
  plaquette_shape(r, prec, &block, &spt);
  set_blocks(r, block * spt);
  r.nblk *= 2;  // even sites, then odd sites
  const int nblk = r.nblk;
  const size_t need = 4 * (size_t)nblk + u1d::fold_scratch(nblk, 4);
  if (need > kNoAllocScratchDoubles)
    return fail(U1_ERR_ARG, "lattice too large for u1_plaquette_noalloc: needs " + std::to_string(need) +
                                " scratch doubles, the maximum is " + std::to_string(kNoAllocScratchDoubles));
  // every row and every fold row is written before it is read: nothing to clear
  HIPCHK(u1d::launch_plaquette(prec, lat, r, rows, nblk, block, spt, G.st));
  HIPCHK(u1d::launch_finish_sums(rows, nblk, 4, 1.0 / (3.0 * (double)r.V), out_dev, rows + 4 * (size_t)nblk, G.st));
  return U1_OK;
}

int u1_polyakov(const u1_geom* g, int prec, const void* lat, double out[2]) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(out);
  RET(check_geom(g));
  RET(check_prec(prec));
  const RepDesc r = make_rep(g, 0, 0);
  const int block = 256, nblk = (r.SV + block - 1) / block;
  RET(ensure_fold(nblk, 2));
  HIPCHK(u1d::launch_polyakov(prec, lat, r, G.partials, nblk, block, G.st));
  HIPCHK(u1d::launch_finish_sums(G.partials, nblk, 2, 1.0 / (double)r.SV, G.dres, fold_scratch_ptr(nblk, 2), G.st));
  return copy_back(out, G.dres, 2);
}

int u1_poly_volume(const u1_geom* g, int prec, const void* lat, int32_t mode, const u1_metro* mh, uint64_t seed,
                   uint32_t stream, uint32_t c1, uint32_t c3, void* P_dev) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(P_dev);
  RET(check_geom(g));
  RET(check_prec(prec));
  if (mode < 0 || mode > 2) return fail(U1_ERR_ARG, "bad Polyakov mode");
  if (mode == U1_POLY_MULTIHIT_STOCHASTIC) RET(check_metro(mh));
  const RepDesc r = make_rep(g, seed, c3);
  if (mode != U1_POLY_PLAIN) RET(ensure_work(2 * (size_t)r.V));
  return poly_volume_async(prec, lat, r, mode, mh, stream, c1, c3, (double*)P_dev, G.work);
}

int u1_temporal_links(const u1_geom* g, int prec, const void* lat, int32_t mode, const u1_metro* mh,
                      uint64_t seed, uint32_t stream, uint32_t c1, uint32_t c3, void* out_dev) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(out_dev);
  RET(check_geom(g));
  RET(check_prec(prec));
  if (mode < 0 || mode > 2) return fail(U1_ERR_ARG, "bad temporal-link mode");
  if (mode == U1_POLY_MULTIHIT_STOCHASTIC) RET(check_metro(mh));
  const RepDesc r = make_rep(g, seed, c3);
  HIPCHK(u1d::launch_temporal_links(prec, lat, r, mode, mh ? mh->proposal : 0, mh ? mh->nhit : 0,
                                    mh ? mh->step : 0.0, stream, c1, c3, (double*)out_dev, 256, G.st));
  return U1_OK;
}

int u1_pp(const u1_geom* g, const void* P_dev, int32_t rmax, double* out) {
  REQ_INIT();
  REQ_PTR(P_dev);
  REQ_PTR(out);
  RET(check_geom(g));
  if (rmax < 1 || rmax > 32) return fail(U1_ERR_ARG, "rmax must be in [1, 32]");
  const RepDesc r = make_rep(g, 0, 0);
  const int block = 256, nblk = (r.SV + block - 1) / block;
  RET(ensure_fold(nblk, 2 * rmax));
  HIPCHK(u1d::launch_pp((const double*)P_dev, r, rmax, G.partials, nblk, block, G.st));
  HIPCHK(u1d::launch_finish_sums(G.partials, nblk, 2 * rmax, 1.0 / (3.0 * (double)r.SV), G.dres,
                                 fold_scratch_ptr(nblk, 2 * rmax), G.st));
  return copy_back(out, G.dres, 2 * (size_t)rmax);
}

int u1_poly2(const u1_geom* g, int prec, const void* lat, int32_t mode, const u1_metro* mh, uint64_t seed,
             uint32_t c1, uint32_t c3, double* out) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(out);
  RET(check_geom(g));
  RET(check_prec(prec));
  if (mode < 0 || mode > 2) return fail(U1_ERR_ARG, "bad Polyakov mode");
  if (mode == U1_POLY_MULTIHIT_STOCHASTIC) RET(check_metro(mh));
  const RepDesc r = make_rep(g, seed, c3);
  // work: P (2 SV doubles), then the temporal links (2 V doubles, multihit)
  RET(ensure_work(2 * (size_t)r.SV + (mode != U1_POLY_PLAIN ? 2 * (size_t)r.V : 0)));
  double* P = G.work;
  RET(poly_volume_async(prec, lat, r, mode, mh, U1_STREAM_MULTIHIT, c1, c3, P, G.work + 2 * (size_t)r.SV));
  return u1_pp(g, P, g->dims[0] / 2, out);
}

int u1_multilevel(const u1_geom* g, int prec, const void* lat, int32_t n4, int32_t k4, int32_t n2, int32_t k2,
                  int32_t metrop, int32_t ovrn, const u1_metro* m, int32_t mh_mode, const u1_metro* mh,
                  uint64_t seed, uint32_t call_id, double* out, void* ppspace_dev) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(out);
  RET(check_geom(g));
  RET(check_prec(prec));
  RET(check_metro(m));
  // multilevel.cu:641-644
  if (g->dims[3] % 4 != 0) return fail(U1_ERR_ARG, "Cannot apply MultiLevel algorithm: Nt is not a multiple of 4");
  if (n4 < 1 || n2 < 1 || k4 < 0 || k2 < 0) return fail(U1_ERR_ARG, "bad multilevel level counts");
  if (mh_mode != U1_POLY_PLAIN && mh_mode != U1_POLY_MULTIHIT && mh_mode != U1_POLY_MULTIHIT_STOCHASTIC)
    return fail(U1_ERR_ARG, "bad multihit mode");
  if (mh_mode == U1_POLY_MULTIHIT_STOCHASTIC) RET(check_metro(mh));
  const RepDesc r = make_rep(g, seed, call_id);
  const int rmax = r.N[0] / 2;
  if (rmax > 32) return fail(U1_ERR_ARG, "Nx/2 must be <= 32");
  const int64_t sz = (int64_t)r.SV * rmax * 3;
  const int64_t nl2 = r.N[3] / 2, nl4 = r.N[3] / 4;
  if (sz * nl2 >= ((int64_t)1 << 31)) return fail(U1_ERR_ARG, "lattice too large for the multilevel T2 tensor");
  const size_t latb = 4 * (size_t)r.V * elem_size(prec);
  void *wl = nullptr, *tl = nullptr, *pq = nullptr, *l2 = nullptr, *l4 = nullptr;
  uint32_t* ctr = nullptr;  // level-2 step counter i * n2 + k, advanced by the graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int rc = U1_OK;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(G.st);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (wl) (void)hipFree(wl);
    if (tl) (void)hipFree(tl);
    if (pq) (void)hipFree(pq);
    if (l2) (void)hipFree(l2);
    if (l4) (void)hipFree(l4);
    if (ctr) (void)hipFree(ctr);
  };
#define MLCHK(expr)                                                                          \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      cleanup();                                                                             \
      return fail(U1_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    }                                                                                        \
  } while (0)
  MLCHK(hipMalloc(&wl, latb));
  MLCHK(hipMalloc(&tl, 2 * (size_t)r.V * sizeof(double)));
  MLCHK(hipMalloc(&pq, 4 * (size_t)r.SV * (size_t)nl2 * sizeof(double)));
  MLCHK(hipMalloc(&l2, 2 * (size_t)sz * nl2 * sizeof(double)));
  MLCHK(hipMalloc(&l4, 2 * (size_t)sz * nl4 * sizeof(double)));
  MLCHK(hipMemcpyAsync(wl, lat, latb, hipMemcpyDeviceToDevice, G.st));
  MLCHK(u1d::launch_fill_zero(l4, 2 * sz * nl4 * (int64_t)sizeof(double), G.st));
  MLCHK(hipMalloc((void**)&ctr, sizeof(uint32_t)));
  MLCHK(u1d::launch_fill_zero(ctr, sizeof(uint32_t), G.st));
  const double l2norm = 1.0 / (double)n2, l4norm = 1.0 / (double)n4;
  const int mp = mh ? mh->proposal : 0, mn = mh ? mh->nhit : 0;
  const double ms = mh ? mh->step : 0.0;
  {  // launch shapes are settled before capture (the tuner allocates and synchronises)
    PhaseArgs a{};
    a.single = r;
    a.nhit = m->nhit;
    a.proposal = m->proposal;
    a.step = m->step;
    rc = settle_update_cfgs(prec, g, wl, a, metrop, ovrn);
    if (rc != U1_OK) {
      cleanup();
      return rc;
    }
  }
  // One level-2 step -- k2 masked sweeps (Philox sweep index ctr * k2 + l),
  // the temporal-link field (c1 = ctr), the T2 accumulation and ctr += 1 --
  // captured once and replayed n4 * n2 times (multilevel.cu:680-700 order;
  // sweep and c1 indices equal to ((i n2 + k) k2 + l) and (i n2 + k)), or
  // enqueued eagerly with the same device counter when graphs are off.
  auto level2_step = [&]() -> int {
    for (int l = 0; l < k2; ++l)
      RET(run_update(g, prec, wl, metrop, ovrn, m, seed, u1r::STREAM_ML_L2, (uint32_t)l, call_id, 2, nullptr,
                     nullptr, ctr, (uint32_t)k2));
    hipError_t e = u1d::launch_temporal_links(prec, wl, r, mh_mode, mp, mn, ms, u1r::STREAM_ML_MULTIHIT, 0, call_id,
                                              (double*)tl, 256, G.st, ctr);
    if (e == hipSuccess) e = u1d::launch_ml_l2((const double*)tl, (double*)pq, (double*)l2, r, rmax, 256, G.st);
    if (e == hipSuccess) e = u1d::launch_increment(ctr, G.st);
    if (e != hipSuccess) return fail(U1_ERR_HIP, std::string("multilevel level-2 step: ") + hipGetErrorString(e));
    return U1_OK;
  };
  if (G.graphs) {
    MLCHK(hipStreamBeginCapture(G.st, hipStreamCaptureModeThreadLocal));
    rc = level2_step();
    const hipError_t ce = hipStreamEndCapture(G.st, &graph);
    if (rc != U1_OK) {
      cleanup();
      return rc;
    }
    MLCHK(ce);
    MLCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  for (int i = 0; i < n4 && rc == U1_OK; ++i) {
    for (int j = 0; j < k4 && rc == U1_OK; ++j)
      rc = run_update(g, prec, wl, metrop, ovrn, m, seed, u1r::STREAM_ML_L4, (uint32_t)(i * k4 + j), call_id, 4,
                      nullptr);
    if (rc != U1_OK) break;
    MLCHK(u1d::launch_fill_zero(l2, 2 * sz * nl2 * (int64_t)sizeof(double), G.st));
    for (int k = 0; k < n2 && rc == U1_OK; ++k) {
      if (exec)
        MLCHK(hipGraphLaunch(exec, G.st));
      else
        rc = level2_step();
    }
    if (rc != U1_OK) break;
    MLCHK(u1d::launch_ml_l4((const double*)l2, (double*)l4, r, rmax, l2norm, 256, G.st));
  }
  if (rc == U1_OK) {
    const int block = 256, nblk = (r.SV + block - 1) / block;
    rc = ensure_fold(nblk, 2 * rmax);
    if (rc == U1_OK) {
      MLCHK(u1d::launch_ml_final((const double*)l4, (double*)ppspace_dev, r, rmax, l4norm, G.partials, nblk, block,
                                 G.st));
      MLCHK(u1d::launch_finish_sums(G.partials, nblk, 2 * rmax, 1.0 / (3.0 * (double)r.SV), G.dres,
                                    fold_scratch_ptr(nblk, 2 * rmax), G.st));
      rc = copy_back(out, G.dres, 2 * (size_t)rmax);
    }
  }
#undef MLCHK
  cleanup();
  return rc;
}

int u1_convert_layout(const u1_geom* g, int prec, const void* in, void* out, int32_t eo_to_no) {
  REQ_INIT();
  REQ_PTR(in);
  REQ_PTR(out);
  RET(check_geom(g));
  RET(check_prec(prec));
  if (in == out) return fail(U1_ERR_ARG, "in and out must differ");
  const RepDesc r = make_rep(g, 0, 0);
  HIPCHK(u1d::launch_convert_layout(prec, in, out, r, eo_to_no, 256, G.st));
  return U1_OK;
}

int u1_plaquette_fields(const u1_geom* g, int prec, const void* lat, int32_t spacetime, int32_t eo_order,
                        void* field_dev, double means[12]) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(field_dev);
  REQ_PTR(means);
  RET(check_geom(g));
  RET(check_prec(prec));
  const RepDesc r = make_rep(g, 0, 0);
  const int size = spacetime ? r.V : r.SV;
  const int block = 256, nblk = (size + block - 1) / block;
  RET(ensure_fold(nblk, 12));
  HIPCHK(u1d::launch_plaquette_fields(prec, lat, r, spacetime, eo_order, (double*)field_dev, G.partials, nblk, block,
                                      G.st));
  HIPCHK(u1d::launch_finish_sums(G.partials, nblk, 12, 1.0 / (double)size, G.dres, fold_scratch_ptr(nblk, 12),
                                 G.st));
  return copy_back(means, G.dres, 12);
}

// ---------------------------------------------------------------- checkpoint
namespace {
struct CfgHeader {
  char magic[8];
  uint32_t version;
  int32_t dims[4];
  uint32_t reserved;
  double beta, aniso;
  uint64_t seed, sweep;
};
static_assert(sizeof(CfgHeader) == 64, "config header must be 64 bytes");
}  // namespace

int u1_config_save(const char* path, const u1_geom* g, int prec, const void* lat, uint64_t seed, uint64_t sweep) {
  REQ_INIT();
  RET(check_geom(g));
  RET(check_prec(prec));
  if (!path || !lat) return fail(U1_ERR_ARG, "null argument");
  const RepDesc r = make_rep(g, 0, 0);
  const size_t n = 4 * (size_t)r.V;
  std::vector<double> host(n);
  {
    DevBuf no, d64;
    HIPCHK(no.alloc(n * elem_size(prec)));
    HIPCHK(d64.alloc(n * sizeof(double)));
    HIPCHK(u1d::launch_convert_layout(prec, lat, no.p, r, 1, 256, G.st));
    HIPCHK(u1d::launch_convert_precision(no.p, prec, d64.p, U1_F64, (int64_t)n, 256, G.st));
    HIPCHK(hipMemcpyAsync(host.data(), d64.p, n * sizeof(double), hipMemcpyDeviceToHost, G.st));
    HIPCHK(hipStreamSynchronize(G.st));
  }
  CfgHeader h{};
  std::memcpy(h.magic, "U1AMDCFG", 8);
  h.version = 1;
  for (int i = 0; i < 4; ++i) h.dims[i] = g->dims[i];
  h.beta = g->beta;
  h.aniso = g->aniso;
  h.seed = seed;
  h.sweep = sweep;
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(U1_ERR_IO, std::string("cannot write ") + path);
  f.write((const char*)&h, sizeof(h));
  f.write((const char*)host.data(), n * sizeof(double));
  return f ? U1_OK : fail(U1_ERR_IO, std::string("write failed: ") + path);
}

int u1_config_load(const char* path, const u1_geom* g, int prec, void* lat, uint64_t* seed, uint64_t* sweep) {
  if (!path || !g) return fail(U1_ERR_ARG, "null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(U1_ERR_IO, std::string("cannot open ") + path);
  CfgHeader h{};
  f.read((char*)&h, sizeof(h));
  if (!f || std::memcmp(h.magic, "U1AMDCFG", 8) != 0 || h.version != 1)
    return fail(U1_ERR_IO, std::string("not a configuration file: ") + path);
  for (int i = 0; i < 4; ++i)
    if (h.dims[i] != g->dims[i]) return fail(U1_ERR_ARG, "configuration file geometry differs from the lattice");
  REQ_INIT();
  RET(check_geom(g));
  RET(check_prec(prec));
  if (!lat) return fail(U1_ERR_ARG, "null lattice");
  const RepDesc r = make_rep(g, 0, 0);
  const size_t n = 4 * (size_t)r.V;
  std::vector<double> host(n);
  f.read((char*)host.data(), n * sizeof(double));
  if (!f) return fail(U1_ERR_IO, std::string("truncated configuration file: ") + path);
  {
    DevBuf no, d64;
    HIPCHK(no.alloc(n * elem_size(prec)));
    HIPCHK(d64.alloc(n * sizeof(double)));
    HIPCHK(hipMemcpyAsync(d64.p, host.data(), n * sizeof(double), hipMemcpyHostToDevice, G.st));
    HIPCHK(u1d::launch_convert_precision(d64.p, U1_F64, no.p, prec, (int64_t)n, 256, G.st));
    HIPCHK(u1d::launch_convert_layout(prec, no.p, lat, r, 0, 256, G.st));
    HIPCHK(hipStreamSynchronize(G.st));
  }
  if (seed) *seed = h.seed;
  if (sweep) *sweep = h.sweep;
  return U1_OK;
}

// ---------------------------------------------------------------- ensemble
int u1_ensemble_create(u1_ensemble** out, int32_t nrep, const u1_geom* geoms, const uint64_t* seeds,
                       const int32_t* hot, int prec, int32_t metrop, int32_t ovrn, const u1_metro* m) {
  REQ_INIT();
  if (!out || !geoms || !seeds || nrep < 1 || nrep > 65535) return fail(U1_ERR_ARG, "bad ensemble arguments");
  if (metrop < 0 || ovrn < 0 || metrop > 4095 || ovrn > 4095) return fail(U1_ERR_ARG, "bad pass counts");
  RET(check_prec(prec));
  RET(check_metro(m));
  *out = nullptr;
  u1_ensemble* e = new u1_ensemble();
  e->nrep = nrep;
  e->prec = prec;
  e->metrop = metrop;
  e->ovrn = ovrn;
  e->m = *m;
  std::vector<uint16_t> mmap;
  int64_t off = 0;
  for (int i = 0; i < nrep; ++i) {
    int rc = check_geom(&geoms[i]);
    if (rc != U1_OK) {
      delete e;
      return rc;
    }
    e->geoms.push_back(geoms[i]);
    RepDesc r = make_rep(&geoms[i], seeds[i], (uint32_t)i);
    r.lat_off = (uint64_t)off;
    off += 4 * (int64_t)r.V;
    RepDesc u = r, q = r;
    set_blocks(q, kEnsBlock * kEnsSpt);
    q.nblk *= 2;  // even sites, then odd sites
    q.blk_begin = e->meas_blocks;
    e->meas_blocks += q.nblk;
    for (int b = 0; b < q.nblk; ++b) mmap.push_back((uint16_t)i);
    e->upd_host.push_back(u);
    e->meas_host.push_back(q);
  }
  e->total_links = off;
  auto bail = [&](const std::string& msg) {
    ensemble_release(e);
    delete e;
    return fail(U1_ERR_HIP, msg);
  };
#define ECHK(expr)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return bail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
  ECHK(hipMalloc(&e->lat, (size_t)off * elem_size(prec)));
  ECHK(u1d::launch_fill_zero(e->lat, off * (int64_t)elem_size(prec), G.st));
  ECHK(hipMalloc((void**)&e->upd_dev, nrep * sizeof(RepDesc)));
  ECHK(hipMalloc((void**)&e->meas_dev, nrep * sizeof(RepDesc)));
  ECHK(hipMalloc((void**)&e->meas_map, mmap.size() * sizeof(uint16_t)));
  ECHK(hipMalloc((void**)&e->sweep_dev, sizeof(uint32_t)));
  ECHK(hipMalloc((void**)&e->res_dev, 4 * (size_t)nrep * sizeof(double)));
  ECHK(hipMalloc((void**)&e->partials, 4 * (size_t)e->meas_blocks * sizeof(double)));
  ECHK(hipMemcpyAsync(e->meas_dev, e->meas_host.data(), nrep * sizeof(RepDesc), hipMemcpyHostToDevice, G.st));
  ECHK(hipMemcpyAsync(e->meas_map, mmap.data(), mmap.size() * sizeof(uint16_t), hipMemcpyHostToDevice, G.st));
  ECHK(u1d::launch_fill_zero(e->sweep_dev, sizeof(uint32_t), G.st));
  if (int rc = ensemble_build_update(e, e->cfg); rc != U1_OK) {
    ensemble_release(e);
    delete e;
    return rc;
  }
  for (int i = 0; i < nrep; ++i)
    if (hot && hot[i]) ECHK(u1d::launch_hotstart(prec, e->lat, e->upd_host[i], 256, G.st));
  ECHK(hipStreamSynchronize(G.st));
#undef ECHK
  *out = e;
  return U1_OK;
}

int u1_ensemble_sweep(u1_ensemble* e, int32_t nsweeps, int32_t use_graph) {
  REQ_INIT();
  if (!e) return fail(U1_ERR_ARG, "null ensemble");
  if (nsweeps < 0) return fail(U1_ERR_ARG, "nsweeps must be >= 0");
  // tuning changes the launch layout, so it only runs before either graph exists
  RET(ensemble_settle(e));
  if (use_graph && !e->exec) {
    HIPCHK(hipStreamBeginCapture(G.st, hipStreamCaptureModeThreadLocal));
    int rc = ensemble_enqueue_sweep(e);
    hipGraph_t gr = nullptr;
    hipError_t ce = hipStreamEndCapture(G.st, &gr);
    if (rc != U1_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    HIPCHK(ce);
    e->graph = gr;
    HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  }
  for (int s = 0; s < nsweeps; ++s) {
    if (use_graph)
      HIPCHK(hipGraphLaunch(e->exec, G.st));
    else
      RET(ensemble_enqueue_sweep(e));
  }
  e->sweeps_done += (uint64_t)nsweeps;
  return U1_OK;
}

int u1_ensemble_sweep_plaquette(u1_ensemble* e, int32_t nsweeps, double* series) {
  REQ_INIT();
  if (!e || !series) return fail(U1_ERR_ARG, "null argument");
  if (nsweeps < 0) return fail(U1_ERR_ARG, "nsweeps must be >= 0");
  if (e->metrop + e->ovrn == 0) return fail(U1_ERR_ARG, "no measuring pass");
  if (nsweeps == 0) return U1_OK;
  RET(ensemble_settle(e));
  RET(ensemble_build_mgraph(e));
  return ensemble_measured_sweeps(e, nsweeps, series);
}

int u1_ensemble_bench(u1_ensemble* e, int32_t nsweeps, double* series, double* elapsed_ms, double* kernel_ms) {
  REQ_INIT();
  if (!e || !series || !elapsed_ms) return fail(U1_ERR_ARG, "null argument");
  if (nsweeps < 1) return fail(U1_ERR_ARG, "nsweeps must be >= 1");
  if (e->ovrn < 1) return fail(U1_ERR_ARG, "the ensemble timing needs an overrelaxation pass (ovrn >= 1)");
  RET(ensemble_settle(e));
  RET(ensemble_build_mgraph(e));
  EventSet es;
  hipEvent_t e0, e1;
  HIPCHK(es.create(&e0));
  HIPCHK(es.create(&e1));
  HIPCHK(hipStreamSynchronize(G.st));
  HIPCHK(hipEventRecord(e0, G.st));
  RET(ensemble_measured_sweeps(e, nsweeps, series, e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *elapsed_ms = ms;
  if (kernel_ms) {
    RET(time_ensemble_classes(e, kernel_ms, kernel_ms + LC_COUNT));
    kernel_ms[2 * LC_COUNT] = (double)ms / nsweeps;
    kernel_ms[2 * LC_COUNT + 1] = kTimingLaunches;
  }
  return U1_OK;
}

int u1_ensemble_plaquette(u1_ensemble* e, double* out) {
  REQ_INIT();
  if (!e || !out) return fail(U1_ERR_ARG, "null argument");
  HIPCHK(u1d::launch_plaquette_batched(e->prec, e->lat, e->meas_dev, e->meas_map, e->partials, e->meas_blocks,
                                       kEnsBlock, kEnsSpt, G.st));
  HIPCHK(u1d::launch_batched_finish(e->partials, e->meas_dev, e->nrep, e->res_dev, G.st));
  return copy_back(out, e->res_dev, 4 * (size_t)e->nrep);
}

int u1_ensemble_read(u1_ensemble* e, int32_t r, double* host) {
  REQ_INIT();
  if (!e || !host || r < 0 || r >= e->nrep) return fail(U1_ERR_ARG, "bad replica");
  const RepDesc& d = e->upd_host[r];
  const int64_t n = 4 * (int64_t)d.V;
  if (e->prec == U1_F64) {
    HIPCHK(hipMemcpyAsync(host, (double*)e->lat + d.lat_off, n * sizeof(double), hipMemcpyDeviceToHost, G.st));
  } else {
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), (float*)e->lat + d.lat_off, n * sizeof(float), hipMemcpyDeviceToHost, G.st));
    HIPCHK(hipStreamSynchronize(G.st));
    for (int64_t i = 0; i < n; ++i) host[i] = (double)tmp[i];
  }
  HIPCHK(hipStreamSynchronize(G.st));
  return U1_OK;
}

int u1_ensemble_write(u1_ensemble* e, int32_t r, const double* host) {
  REQ_INIT();
  if (!e || !host || r < 0 || r >= e->nrep) return fail(U1_ERR_ARG, "bad replica");
  const RepDesc& d = e->upd_host[r];
  const int64_t n = 4 * (int64_t)d.V;
  if (e->prec == U1_F64) {
    HIPCHK(hipMemcpyAsync((double*)e->lat + d.lat_off, host, n * sizeof(double), hipMemcpyHostToDevice, G.st));
    HIPCHK(hipStreamSynchronize(G.st));
  } else {
    std::vector<float> tmp(n);
    for (int64_t i = 0; i < n; ++i) tmp[i] = (float)host[i];
    HIPCHK(hipMemcpyAsync((float*)e->lat + d.lat_off, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, G.st));
    HIPCHK(hipStreamSynchronize(G.st));
  }
  return U1_OK;
}

int u1_ensemble_info(u1_ensemble* e, uint64_t* sweeps_done, int64_t* total_links) {
  if (!e) return fail(U1_ERR_ARG, "null ensemble");
  if (sweeps_done) *sweeps_done = e->sweeps_done;
  if (total_links) *total_links = e->total_links;
  return U1_OK;
}

int u1_ensemble_destroy(u1_ensemble* e) {
  if (!e) return U1_OK;
  if (G.init) (void)hipStreamSynchronize(G.st);
  ensemble_release(e);
  delete e;
  return U1_OK;
}

// ---------------------------------------------------------------- autotuner
int u1_tune_enable(int32_t on) {
  G.tune = on != 0;
  return U1_OK;
}

int u1_set_graph_capture(int32_t on) {
  if (on != 0 && on != 1) return fail(U1_ERR_ARG, "graph capture switch must be 0 or 1");
  G.graphs = on != 0;
  return U1_OK;
}

int u1_set_pair_fusion(int32_t mode) {
  if (mode < 0 || mode > 2) return fail(U1_ERR_ARG, "pair fusion mode must be 0, 1 or 2");
  G.pair = mode;
  return U1_OK;
}

int u1_set_pass_fusion(int32_t mode) {
  if (mode < 0 || mode > 2) return fail(U1_ERR_ARG, "pass fusion mode must be 0, 1 or 2");
  G.chain = mode;
  return U1_OK;
}

int u1_set_verbosity(int32_t level) {
  if (level < 0 || level > 3) return fail(U1_ERR_ARG, "verbosity must be in [0, 3]");
  G.verbosity = level;
  return U1_OK;
}

// tunecache.csv format of the reference (tune.cpp:109-122, 238-295):
//   tunecache <version> <build hash>
//   volume name aux block.x block.y block.z grid.x grid.y grid.z shared_bytes comment
static const char* kTuneVersion = "latticeu1_amd-1";
static const char* kTuneHash = "gpu_arch=gfx950";

int u1_tune_load(const char* path) {
  if (!path) return fail(U1_ERR_ARG, "null path");
  std::ifstream f(path);
  if (!f) return fail(U1_ERR_IO, std::string("cannot open ") + path);
  std::string line, tok, ver, hash;
  if (!std::getline(f, line)) return fail(U1_ERR_IO, "empty tune cache");
  {
    std::istringstream ls(line);
    ls >> tok >> ver >> hash;
    if (tok != "tunecache") return fail(U1_ERR_IO, std::string("Bad format in ") + path);
    if (ver != kTuneVersion || hash != kTuneHash)
      return fail(U1_ERR_IO, std::string("Cache file ") + path + " does not match this build");
  }
  std::getline(f, line);  // column description
  std::vector<std::pair<std::string, Cfg>> parsed;  // applied only if every line is valid
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    std::istringstream ls(line);
    std::string vol, name, aux;
    int bx, by, bz, gx, gy, gz, sh;
    if (!(std::getline(ls, vol, '\t') && std::getline(ls, name, '\t') && std::getline(ls, aux, '\t')))
      return fail(U1_ERR_IO, std::string("Bad format in ") + path);
    if (!(ls >> bx >> by >> bz >> gx >> gy >> gz >> sh)) return fail(U1_ERR_IO, std::string("Bad format in ") + path);
    // only the shapes the tuner itself tries are accepted: the kernels assume
    // whole wavefronts, at most 512 threads and 4 sites per thread (binary32)
    const bool f32 = name.size() > 4 && name.compare(name.size() - 4, 4, "_f32") == 0;
    bool known = false;
    if (name.find("Chain_") != std::string::npos) {
      // fused pass: block 0 = unfused, block.z = tiles along z, grid.x = schedule margin
      if (!((bx == 0 && by == 1) || ((by == 1 || by == 2) && bx >= 64 && bx <= 512 && bx % 64 == 0 &&
                                     (bz == 1 || bz == 2 || bz == 4 || bz == 8) && gx >= 0 && gx <= kChainMaxMargin)))
        return fail(U1_ERR_IO, std::string("Bad launch shape in ") + path + ": " + line);
      parsed.push_back({vol + "\t" + name, Cfg{bx, by, 1, bz, bx == 0 ? 0 : gx}});
      continue;
    }
    if (name.find("Pair_") != std::string::npos)  // fused pair: block 0 = two launches
      known = (bx == 0 && by == 1) || pair_shape_ok(f32 ? U1_F32 : U1_F64, Cfg{bx, by, 0, 0});
    else
      for (const Cfg& c : kShapes) known = known || (c.block == bx && c.spt == by);
    if (!known || (by == 4 && !f32) || bz < 1 || bz > 4)
      return fail(U1_ERR_IO, std::string("Bad launch shape in ") + path + ": " + line);
    parsed.push_back({vol + "\t" + name, Cfg{bx, by, (bz - 1) & 1, (bz - 1) >> 1}});
  }
  for (const auto& kv : parsed) G.tunecache[kv.first] = kv.second;
  G.tune_loaded = (int)parsed.size();
  return U1_OK;
}

int u1_tune_save(const char* path) {
  if (!path) return fail(U1_ERR_ARG, "null path");
  std::ofstream f(path);
  if (!f) return fail(U1_ERR_IO, std::string("cannot write ") + path);
  f << "tunecache\t" << kTuneVersion << "\t" << kTuneHash << "\n";
  f << "volume\tname\taux\tblock.x\tblock.y\tblock.z\tgrid.x\tgrid.y\tgrid.z\tshared_bytes\tcomment\n";
  for (const auto& kv : G.tunecache) {
    const std::string& key = kv.first;
    const size_t tab = key.find('\t');
    const std::string vol = key.substr(0, tab), name = key.substr(tab + 1);
    const bool chain = name.find("Chain_") != std::string::npos;
    f << vol << "\t" << name << "\tthreads=halfvolume\t" << kv.second.block << "\t" << kv.second.spt << "\t"
      << (chain ? kv.second.tile : 1 + kv.second.xcd + 2 * kv.second.tile) << "\t" << (chain ? kv.second.margin : 0)
      << "\t1\t1\t0\t# block.x = workgroup size, block.y = sites per thread, block.z - 1 = XCD-aware order (1) + "
         "tiled planes (2); Pair: fused (parity, 0) + (parity, 1) launch, block.x 0 = two launches; Chain: fused "
         "parity pass (phases 0-3) in one launch, block.x 0 = unfused, block.z = tiles along z, grid.x = schedule "
         "margin\n";
  }
  return f ? U1_OK : fail(U1_ERR_IO, "write failed");
}

int u1_tune_entries(int32_t* n) {
  if (!n) return fail(U1_ERR_ARG, "null pointer");
  *n = (int32_t)G.tunecache.size();
  return U1_OK;
}

// ---------------------------------------------------------------- timing helper

// Kernel timing of the launch classes of a sweep (1 <= metrop + ovrn, shapes
// already settled): on a scratch copy of the lattice, kTimingLaunches
// back-to-back launches of each class in the sweep's shapes (parity and mu
// cycling over the phases the class covers in a pass), bracketed by one event
// pair.  ms[c]: mean launch duration; per_sweep[c]: launches per sweep.
int time_launch_classes(const u1_geom* g, int prec, const void* lat, int metrop, int ovrn, const u1_metro* m,
                        uint64_t seed, uint32_t sweep, bool meas, double* ms, double* per_sweep) {
  PhaseArgs a{};
  a.single = make_rep(g, seed, 0);
  a.proposal = m->proposal;
  a.nhit = m->nhit;
  a.step = m->step;
  a.stream = U1_STREAM_UPDATE;
  a.sweep = sweep;
  Cfg cfg[2] = {kDefaultCfg, kDefaultCfg}, pcfg[2] = {{0, 1, 0, 0}, {0, 1, 0, 0}},
      ccfg[2] = {{0, 1, 0, 0}, {0, 1, 0, 0}};
  const int passes[2] = {metrop, ovrn};
  for (int kind = 0; kind < 2; ++kind)
    if (passes[kind] > 0) {
      RET(phase_cfg(prec, kind, g, lat, a, &cfg[kind]));
      RET(pair_cfg(prec, kind, g, lat, a, cfg[kind], &pcfg[kind]));
      RET(chain_cfg(prec, kind, g, lat, a, cfg[kind], pcfg[kind], &ccfg[kind]));
    }
  const int64_t bytes = 4 * (int64_t)a.single.V * (int64_t)elem_size(prec);
  DevBuf scratch;
  HIPCHK(scratch.alloc(bytes));
  HIPCHK(hipMemcpyAsync(scratch.p, lat, bytes, hipMemcpyDeviceToDevice, G.st));
  EventSet es;
  hipEvent_t e0, e1;
  HIPCHK(es.create(&e0));
  HIPCHK(es.create(&e1));
  for (int c = 0; c < LC_COUNT; ++c) ms[c] = per_sweep[c] = 0.0;
  // with the plaquette, the last overrelaxation pass measures in its parity-1
  // launches of mu >= 1 (the parity-1 pair, when fused)
  const bool meas_or = meas && ovrn > 0;
  for (int c = 0; c < LC_COUNT; ++c) {
    const bool chain = c >= LC_METRO_CHAIN;
    const bool pair = c >= LC_METRO_PAIR && !chain;
    const int kind = (c == LC_METRO || c == LC_METRO_PAIR || c == LC_METRO_CHAIN) ? u1d::KIND_METROPOLIS
                                                                                  : u1d::KIND_OVERRELAX;
    const bool cmeas = c == LC_OR_MEAS || c == LC_OR_PAIR_MEAS || c == LC_OR_CHAIN_MEAS;
    const Cfg& pc = pcfg[kind];
    const bool fused = pc.block != 0;
    const bool chained = ccfg[kind].block != 0;
    if (passes[kind] == 0 || (cmeas && !meas_or) || chain != chained) continue;
    if (chain) {
      // one launch per parity and pass; the parity-1 launch of the last pass measures
      per_sweep[c] = cmeas ? 1.0 : 2.0 * passes[kind] - (meas_or && kind == u1d::KIND_OVERRELAX ? 1 : 0);
      if (cmeas) {
        RepDesc rc = a.single;
        set_chain_blocks(rc, ccfg[kind]);
        RET(ensure_fold(3 * rc.N[2] * rc.N[3] * rc.bpp, 4));
      }
      PhaseArgs b = a;
      auto launch = [&](int i) -> int {
        b.parity = cmeas ? 1 : i & 1;
        b.pass = (uint32_t)(i >> 1);
        b.meas = cmeas ? G.partials : nullptr;
        return chain_launch(prec, kind, scratch.p, b, ccfg[kind]);
      };
      RET(launch(0));  // warm-up
      HIPCHK(hipEventRecord(e0, G.st));
      for (int i = 0; i < kTimingLaunches; ++i) RET(launch(i));
      HIPCHK(hipEventRecord(e1, G.st));
      HIPCHK(hipEventSynchronize(e1));
      RET(chain_check());
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, e0, e1));
      ms[c] = t / kTimingLaunches;
      continue;
    }
    if (pair && !fused) continue;
    // the (parity, mu) launches of one pass in this class (a pair is listed by mu 0)
    std::vector<std::pair<int, int>> ph;
    for (int parity = 0; parity < 2; ++parity)
      for (int mu = 0; mu < 4; ++mu) {
        if (pair ? mu != 0 : (fused && mu < 2)) continue;
        const bool is_meas = parity == 1 && (pair || mu >= 1);
        if (!cmeas || is_meas) ph.push_back({parity, mu});
      }
    if (cmeas)
      per_sweep[c] = (double)ph.size();
    else {
      const int n_meas = meas_or && kind == u1d::KIND_OVERRELAX ? (pair ? 1 : (fused ? 2 : 3)) : 0;
      per_sweep[c] = (double)ph.size() * passes[kind] - n_meas;
    }
    PhaseArgs b = a;
    b.xcd = (uint32_t)(pair ? pc.xcd : cfg[kind].xcd);
    if (pair)
      set_pair_blocks(b.single, pc);
    else
      set_update_blocks(b.single, cfg[kind]);
    if (cmeas) RET(ensure_fold(b.single.nblk, 4));
    auto launch = [&](int i) -> int {
      b.parity = ph[(size_t)i % ph.size()].first;
      b.pass = (uint32_t)(i / (int)ph.size());
      b.meas = cmeas ? G.partials : nullptr;
      if (pair)
        HIPCHK(u1d::launch_pair(prec, kind, scratch.p, b, b.single.nblk, pc.block, pc.spt, b.single.nxh * b.single.N[1],
                                G.st));
      else
        HIPCHK(u1d::launch_phase(prec, kind, ph[(size_t)i % ph.size()].second, scratch.p, b, b.single.nblk,
                                 cfg[kind].block, cfg[kind].spt, G.st));
      return U1_OK;
    };
    RET(launch(0));  // warm-up
    HIPCHK(hipEventRecord(e0, G.st));
    for (int i = 0; i < kTimingLaunches; ++i) RET(launch(i));
    HIPCHK(hipEventRecord(e1, G.st));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    ms[c] = t / kTimingLaunches;
  }
  return U1_OK;
}

int u1_bench_updates(const u1_geom* g, int prec, void* lat, int32_t metrop, int32_t ovrn, const u1_metro* m,
                     uint64_t seed, uint32_t sweep0, int32_t nsweeps, int32_t measure_plaq, double* plaq_series,
                     double* elapsed_ms, double* kernel_ms) {
  REQ_INIT();
  REQ_PTR(lat);
  REQ_PTR(elapsed_ms);
  RET(check_geom(g));
  RET(check_prec(prec));
  RET(check_metro(m));
  if (nsweeps < 1) return fail(U1_ERR_ARG, "nsweeps must be >= 1");
  if (measure_plaq < 0 || measure_plaq > 2) return fail(U1_ERR_ARG, "measure_plaq must be 0, 1 or 2");
  if (measure_plaq == 1 && metrop + ovrn == 0) return fail(U1_ERR_ARG, "no measuring pass");
  DevBuf series_buf;
  if (measure_plaq) HIPCHK(series_buf.alloc(4 * (size_t)nsweeps * sizeof(double)));
  double* series = series_buf.as<double>();
  {  // tune outside the timed region
    PhaseArgs a{};
    a.single = make_rep(g, seed, 0);
    a.nhit = m->nhit;
    a.proposal = m->proposal;
    a.step = m->step;
    RET(settle_update_cfgs(prec, g, lat, a, metrop, ovrn));
  }
  // Timed region: one event pair around the whole region, nothing inside.
  EventSet es;
  hipEvent_t e0, e1;
  HIPCHK(es.create(&e0));
  HIPCHK(es.create(&e1));
  HIPCHK(hipStreamSynchronize(G.st));
  HIPCHK(hipEventRecord(e0, G.st));
  for (int s = 0; s < nsweeps; ++s) {
    RET(run_update(g, prec, lat, metrop, ovrn, m, seed, U1_STREAM_UPDATE, sweep0 + (uint32_t)s, 0, 0, nullptr,
                   measure_plaq == 1 ? series + 4 * s : nullptr));
    if (measure_plaq == 2) RET(plaquette_async(g, prec, lat, series + 4 * s));
  }
  HIPCHK(hipEventRecord(e1, G.st));
  HIPCHK(hipEventSynchronize(e1));
  RET(chain_check());
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *elapsed_ms = ms;
  if (kernel_ms) {
    RET(time_launch_classes(g, prec, lat, metrop, ovrn, m, seed, sweep0 + (uint32_t)nsweeps, measure_plaq == 1,
                            kernel_ms, kernel_ms + LC_COUNT));
    kernel_ms[2 * LC_COUNT] = (double)ms / nsweeps;
    kernel_ms[2 * LC_COUNT + 1] = kTimingLaunches;
  }
  if (measure_plaq && plaq_series)
    RET(copy_back(plaq_series, series, 4 * (size_t)nsweeps));
  return U1_OK;
}

// ---------------------------------------------------------------- self test
int u1_selftest_math(int32_t fn, const double* x, const double* y, double* out, int64_t n) {
  REQ_INIT();
  if (fn < 0 || fn > 10 || !x || !out || n < 0) return fail(U1_ERR_ARG, "bad self-test arguments");
  if (n == 0) return U1_OK;
  const int nout = (fn == 0 || fn == 3 || fn == 5 || fn == 7) ? 2 : 1;
  DevBuf dx, dy, dout;
  HIPCHK(dx.alloc(n * sizeof(double)));
  HIPCHK(dy.alloc(n * sizeof(double)));
  HIPCHK(dout.alloc(nout * n * sizeof(double)));
  HIPCHK(hipMemcpyAsync(dx.p, x, n * sizeof(double), hipMemcpyHostToDevice, G.st));
  HIPCHK(hipMemcpyAsync(dy.p, y ? y : x, n * sizeof(double), hipMemcpyHostToDevice, G.st));
  HIPCHK(u1d::launch_selftest_math(fn, dx.as<double>(), dy.as<double>(), dout.as<double>(), n, G.st));
  HIPCHK(hipMemcpyAsync(out, dout.p, nout * n * sizeof(double), hipMemcpyDeviceToHost, G.st));
  HIPCHK(hipStreamSynchronize(G.st));
  return U1_OK;
}

int u1_selftest_stream(int32_t pattern, int64_t n, int32_t reps, double* gbs) {
  REQ_INIT();
  if (pattern < 0 || pattern > 2 || n < 8 || n % 8 != 0 || n > ((int64_t)1 << 30) || reps < 1 || !gbs)
    return fail(U1_ERR_ARG, "bad stream arguments");
  float *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc((void**)&a, 9 * (size_t)n * sizeof(float)));
  int rc = U1_OK;
  if (hipMalloc((void**)&b, 9 * (size_t)n * sizeof(float)) != hipSuccess) {
    (void)hipFree(a);
    return fail(U1_ERR_HIP, "hipMalloc failed");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto run = [&]() -> hipError_t {
    hipError_t e;
    if ((e = u1d::launch_fill_zero(a, 36 * n, G.st)) != hipSuccess) return e;
    if ((e = u1d::launch_stream(pattern, a, b, n, G.st)) != hipSuccess) return e;  // warm-up
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) return e;
    if ((e = hipEventRecord(e0, G.st)) != hipSuccess) return e;
    for (int i = 0; i < reps; ++i)
      if ((e = u1d::launch_stream(pattern, a, b, n, G.st)) != hipSuccess) return e;
    if ((e = hipEventRecord(e1, G.st)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return e;
    *gbs = 36.0 * (double)n * reps / (ms * 1e-3) / 1e9;
    return hipSuccess;
  };
  const hipError_t e = run();
  if (e != hipSuccess) rc = fail(U1_ERR_HIP, std::string("stream probe: ") + hipGetErrorString(e));
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(G.st);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

int u1_selftest_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int64_t n) {
  REQ_INIT();
  if (!ctr || !key || !out || n < 0) return fail(U1_ERR_ARG, "bad self-test arguments");
  if (n == 0) return U1_OK;
  DevBuf dc, dk, dout;
  HIPCHK(dc.alloc(4 * n * sizeof(uint32_t)));
  HIPCHK(dk.alloc(2 * n * sizeof(uint32_t)));
  HIPCHK(dout.alloc(4 * n * sizeof(uint32_t)));
  HIPCHK(hipMemcpyAsync(dc.p, ctr, 4 * n * sizeof(uint32_t), hipMemcpyHostToDevice, G.st));
  HIPCHK(hipMemcpyAsync(dk.p, key, 2 * n * sizeof(uint32_t), hipMemcpyHostToDevice, G.st));
  HIPCHK(u1d::launch_selftest_philox(dc.as<uint32_t>(), dk.as<uint32_t>(), dout.as<uint32_t>(), n, G.st));
  HIPCHK(hipMemcpyAsync(out, dout.p, 4 * n * sizeof(uint32_t), hipMemcpyDeviceToHost, G.st));
  HIPCHK(hipStreamSynchronize(G.st));
  return U1_OK;
}

}  // extern "C"
