// u1_dropin.cpp -- the reference's C++ host API (update.h, plaquette.h,
// polyakov.h, multilevel.h, array.h, parameters.h, cuda_error_check.h) on top
// of the MI355X C ABI (include/latticeu1_amd.h).  Plain C++17, compiled with
// g++: no HIP or CUDA header is included; all lattice work runs as gfx950
// kernels inside libu1amd.so.
//
// Semantics kept from the reference: global geometry in PARAMS, synchronous
// calls (every Run() ended in a device sync, e.g. update.cu:309), caller-owned
// results (Poly2 / MultiLevel return host Arrays the caller deletes), the data
// files Pot_*.dat / Pot_mhit_*.dat / Pot_mlevel_*.dat with precision 12, and
// fatal errors via Finalize(1).
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "../../include/u1/array.h"
#include "../../include/u1/checkpoint.h"
#include "../../include/u1/multilevel.h"
#include "../../include/u1/parameters.h"
#include "../../include/u1/plaquette.h"
#include "../../include/u1/polyakov.h"
#include "../../include/u1/random.h"
#include "../../include/u1/runtime.h"
#include "../../include/u1/timer.h"
#include "../../include/u1/update.h"

namespace U1 {

namespace PARAMS {
double Beta = 1.0;
double Aniso = 1.0;
int DIRS = 4;
int TDir = 3;
int volume = 0;
int half_volume = 0;
int spatial_volume = 0;
int Grid[4] = {1, 1, 1, 1};
int Offset[4] = {0, 0, 0, 0};
int iter = 0;
double accept_ratio = 0.;
int ovrn = 3;
int metrop = 1;
DeviceProp deviceProp = {{0}, 0};
int proposal = U1_PROPOSAL_INDEPENDENT;
int nhit = 1;
double step = 0.5;
int mhit_kind = U1_POLY_MULTIHIT;
int mhit_nhit = 100;
double mhit_step = 0.5;
}  // namespace PARAMS

namespace {
TuneMode kerneltune = TUNE_YES;
Verbosity verbose = VERBOSE;
Timer total_time;
std::string resource_path;
size_t dev_bytes = 0, dev_peak = 0;
// Random streams of the host-array entry points and of Poly2's stochastic
// multihit, which take no CudaRNG argument in the reference API.
PhiloxState host_rng = {1234, 0, 0};
uint64_t poly2_calls = 0;

std::string cache_file() { return resource_path + "/tunecache.csv"; }
}  // namespace

// ------------------------------------------------------------------ runtime
void u1Check_(const char* func, const char* file, int line, int status) {
  if (status != U1_OK) {
    std::fprintf(stderr, "U1 error in file %s:%i in %s: %s.\n", file, line, func, u1_last_error());
    Finalize_(func, file, line, 1);
  }
}

void devSync() { u1SafeCall(u1_sync()); }

// cuda_error_check.cpp:24-45
void Start(int gpuid, Verbosity verb, TuneMode tune) {
  total_time.start();
  u1SafeCall(u1_init(gpuid));
  char name[256];
  u1SafeCall(u1_device_name(name, sizeof(name)));
  std::snprintf(PARAMS::deviceProp.name, sizeof(PARAMS::deviceProp.name), "%s", name);
  if (verb >= VERBOSE) std::cout << "Using gpu id: " << gpuid << " (" << name << ")" << std::endl;
  setTuning(tune);
  setVerbosity(verb);
  // tune.cpp:160-236: cache in $U1_RESOURCE_PATH or the working directory
  const char* path = std::getenv("U1_RESOURCE_PATH");
  struct stat pstat;
  if (path && stat(path, &pstat) == 0 && S_ISDIR(pstat.st_mode)) {
    resource_path = path;
  } else {
    char cwd[4096];
    resource_path = getcwd(cwd, sizeof(cwd)) ? cwd : ".";
    if (verb >= SUMMARIZE) printfU1("warning: using default tune cache directory: %s\n", resource_path.c_str());
  }
  u1_tune_enable(tune == TUNE_YES ? 1 : 0);
  if (tune == TUNE_YES) {
    std::ifstream f(cache_file());
    if (f.good()) {
      f.close();
      if (u1_tune_load(cache_file().c_str()) != U1_OK) errorU1("%s\n", u1_last_error());
      if (verb >= SUMMARIZE) printfU1("Loaded cached launch parameters from %s\n", cache_file().c_str());
    } else if (verb >= SUMMARIZE) {
      printfU1("warning: Cache file not found.  All kernels will be re-tuned.\n");
    }
  }
}

// cuda_error_check.cpp:47-72
void Finalize_(const char* func, const char* file, int line, int error) {
  (void)func;
  (void)file;
  (void)line;
  u1_sync();
  if (getTuning() == TUNE_YES && !resource_path.empty()) u1_tune_save(cache_file().c_str());
  printPeakMemUsage();
  total_time.stop();
  if (getVerbosity() >= SUMMARIZE) std::cout << "Total Time: " << total_time.getElapsedTime() << " s" << std::endl;
  u1_finalize();
  std::fflush(stdout);
  std::exit(error ? error : EXIT_SUCCESS);
}

// ------------------------------------------------------------------ memory
void* dev_malloc(size_t size) {
  void* p = nullptr;
  u1SafeCall(u1_malloc(&p, size));
  dev_bytes += size;
  if (dev_bytes > dev_peak) dev_peak = dev_bytes;
  return p;
}
void dev_free(void* ptr) { u1SafeCall(u1_free(ptr)); }
void* safe_malloc(size_t size) {
  void* p = std::calloc(size ? size : 1, 1);
  if (!p) errorU1("host allocation of %zu bytes failed\n", size);
  return p;
}
void host_free(void* ptr) { std::free(ptr); }
void printPeakMemUsage() {
  if (getVerbosity() >= SUMMARIZE)
    printfU1("Device memory requested in total: %.3f MiB\n", (double)dev_peak / 1048576.0);
}

// ------------------------------------------------------------------ parameters
TuneMode getTuning() { return kerneltune; }
Verbosity getVerbosity() { return verbose; }
void setTuning(TuneMode t) {
  kerneltune = t;
  u1_tune_enable(t == TUNE_YES ? 1 : 0);
}
void setVerbosity(Verbosity v) {
  verbose = v;
  (void)u1_set_verbosity((int32_t)v);  // tuner reporting follows the same level
}

void setProposal(ProposalKind kind, int nhit, double step) {
  PARAMS::proposal = kind;
  PARAMS::nhit = nhit;
  PARAMS::step = step;
}
void setMultiHit(MultiHitKind kind, int nhit, double step) {
  PARAMS::mhit_kind = kind;
  PARAMS::mhit_nhit = nhit;
  PARAMS::mhit_step = step;
}

// parameters.cu:44-60
std::string GetLatticeName() {
  std::string name = "";
  for (int i = 0; i < Dirs(); i++) name += ToString(PARAMS::Grid[i]) + "_";
  name += "beta_" + ToString(PARAMS::Beta);
  name += "_xi_" + ToString(PARAMS::Aniso);
  name += "_mtr_" + ToString(PARAMS::metrop);
  name += "_ovr_" + ToString(PARAMS::ovrn);
  return name;
}
std::string GetLatticeNameI() { return GetLatticeName() + "_iter_" + ToString(PARAMS::iter); }

// parameters.cu:146-208
void SetupLatticeParameters(int Nx, int Ny, int Nz, int Nt, int dirs, double beta, double aniso, int imetrop,
                            int ovrn) {
  if (dirs != 4) {
    std::cout << "Error: this build supports 4D lattices (dirs = 4) only..." << std::endl;
    Finalize(1);
  }
  PARAMS::DIRS = dirs;
  PARAMS::TDir = dirs - 1;
  PARAMS::Grid[0] = Nx;
  PARAMS::Grid[1] = Ny;
  PARAMS::Grid[2] = Nz;
  PARAMS::Grid[3] = Nt;
  for (int i = 0; i < dirs; ++i)
    if ((Grid(i) % 2) != 0) {
      std::cout << "Error: Number of points should be an even number..." << std::endl;
      Finalize(1);
    }
  PARAMS::Offset[0] = 1;
  for (int i = 1; i < dirs; ++i) PARAMS::Offset[i] = PARAMS::Grid[i - 1] * PARAMS::Offset[i - 1];
  PARAMS::volume = Nx * Ny * Nz * Nt;
  PARAMS::half_volume = PARAMS::volume / 2;
  PARAMS::spatial_volume = Nx * Ny * Nz;
  PARAMS::Beta = beta;
  PARAMS::Aniso = aniso;
  PARAMS::metrop = imetrop;
  PARAMS::ovrn = ovrn;
  if (getVerbosity() >= VERBOSE) {
    std::cout << "==========================================" << std::endl;
    std::cout << "Lattice properties:" << std::endl;
    std::cout << "    Volume: " << Nx << "x" << Ny << "x" << Nz << "x" << Nt << std::endl;
    std::cout << "    Number of directions: " << Dirs() << std::endl;
    std::cout << "    Time indice: " << TDir() << std::endl;
    std::cout << "    Beta: " << Beta() << std::endl;
    std::cout << "    Anisotropy: " << Aniso() << std::endl;
    std::cout << "    Metropolis updates: " << PARAMS::metrop << std::endl;
    std::cout << "    Overrelaxation updates: " << PARAMS::ovrn << std::endl;
    std::cout << "==========================================" << std::endl;
  }
}

u1_geom CurrentGeom() {
  u1_geom g;
  for (int i = 0; i < 4; ++i) g.dims[i] = PARAMS::Grid[i];
  g.beta = PARAMS::Beta;
  g.aniso = PARAMS::Aniso;
  return g;
}
u1_metro CurrentMetro() {
  u1_metro m;
  m.proposal = PARAMS::proposal;
  m.nhit = PARAMS::nhit;
  m.step = PARAMS::step;
  return m;
}

namespace {
u1_metro MultiHitMetro() {
  u1_metro m;
  m.proposal = U1_PROPOSAL_DELTA;
  m.nhit = PARAMS::mhit_nhit;
  m.step = PARAMS::mhit_step;
  return m;
}
template <class Real>
int prec_of();
template <>
int prec_of<double>() { return U1_F64; }
template <>
int prec_of<float>() { return U1_F32; }

std::ofstream open_out(const std::string& filename) {
  std::ofstream f(filename.c_str());
  if (!f.is_open()) {
    std::cout << "Error opening file: " << filename << std::endl;
    std::exit(1);
  }
  f.precision(12);
  return f;
}
}  // namespace

// ------------------------------------------------------------------ updates
template <class Real>
static void hot_start(Array<Real>* lat, CudaRNG* rng) {
  const u1_geom g = CurrentGeom();
  u1SafeCall(u1_hotstart(&g, prec_of<Real>(), lat->getPtr(), rng->getPtr()->seed, 0));
  devSync();
}
void HotStart(Array<double>* lat, CudaRNG* rng) { hot_start(lat, rng); }
void HotStart(Array<float>* lat, CudaRNG* rng) { hot_start(lat, rng); }

// update.cu:449-456: metrop Metropolis passes, then ovrn overrelaxation passes.
template <class Real>
static void update_lattice(Array<Real>* lat, CudaRNG* rng, int metrop, int ovrn) {
  const u1_geom g = CurrentGeom();
  const u1_metro m = CurrentMetro();
  PhiloxState* st = rng->getPtr();
  u1SafeCall(u1_update(&g, prec_of<Real>(), lat->getPtr(), metrop, ovrn, &m, st->seed, U1_STREAM_UPDATE,
                       (uint32_t)st->sweep, 0, 0, nullptr));
  st->sweep++;
  devSync();
}
void UpdateLattice(Array<double>* lat, CudaRNG* rng, int metrop, int ovrn) { update_lattice(lat, rng, metrop, ovrn); }
void UpdateLattice(Array<float>* lat, CudaRNG* rng, int metrop, int ovrn) { update_lattice(lat, rng, metrop, ovrn); }

// update.cu:51-98 took host arrays; here the host array makes a round trip
// through the device, where one pass runs.
static void host_pass(double* lat, int metrop, int ovrn) {
  const u1_geom g = CurrentGeom();
  const u1_metro m = CurrentMetro();
  const size_t bytes = 4 * (size_t)Volume() * sizeof(double);
  void* d = dev_malloc(bytes);
  u1SafeCall(u1_memcpy(d, lat, bytes, U1_H2D));
  uint64_t acc = 0;
  u1SafeCall(u1_update(&g, U1_F64, d, metrop, ovrn, &m, host_rng.seed, U1_STREAM_UPDATE, (uint32_t)host_rng.sweep, 0,
                       0, metrop ? &acc : nullptr));
  host_rng.sweep++;
  PARAMS::accept_ratio += (double)acc;
  u1SafeCall(u1_memcpy(lat, d, bytes, U1_D2H));
  dev_free(d);
}
void metropolis(double* lat) { host_pass(lat, 1, 0); }
void overrelaxation(double* lat) { host_pass(lat, 0, 1); }

// ------------------------------------------------------------------ measurements
template <class Real>
static complexd* plaquette_dev(Array<Real>* lat, complexd* plaq, bool print) {
  const u1_geom g = CurrentGeom();
  double out[4];
  u1SafeCall(u1_plaquette(&g, prec_of<Real>(), lat->getPtr(), out));
  plaq[0] = complexd(out[0], out[1]);
  plaq[1] = complexd(out[2], out[3]);
  // plaquette.cu:224-225
  if (print) {
    std::cout << "0plaqSS: " << plaq[0] << "\tplaqST: " << plaq[1] << "\tMean: " << (plaq[0] + plaq[1]) * 0.5
              << std::endl;
    std::cout << "1plaqSS: " << plaq[0] / Aniso() << "\tplaqST: " << plaq[1] * Aniso()
              << "\tMean: " << (plaq[0] / Aniso() + plaq[1] * Aniso()) * 0.5
              << "\tAction: " << 0.5 * Beta() * (plaq[0] / Aniso() + plaq[1] * Aniso()) << std::endl;
  }
  return plaq;
}
complexd* Plaquette(Array<double>* lat, complexd* plaq, bool print) { return plaquette_dev(lat, plaq, print); }
complexd* Plaquette(Array<float>* lat, complexd* plaq, bool print) { return plaquette_dev(lat, plaq, print); }

void plaquette(double* lat, double* plaq) {
  const u1_geom g = CurrentGeom();
  const size_t bytes = 4 * (size_t)Volume() * sizeof(double);
  void* d = dev_malloc(bytes);
  u1SafeCall(u1_memcpy(d, lat, bytes, U1_H2D));
  double out[4];
  u1SafeCall(u1_plaquette(&g, U1_F64, d, out));
  dev_free(d);
  // <cos P>, <sin P> over the 6 planes (3 space-space + 3 space-time)
  plaq[0] = 1.0 - 0.5 * (out[0] + out[2]);
  plaq[1] = -0.5 * (out[1] + out[3]);
}

template <class Real>
static complexd polyakov_dev(Array<Real>* lat, bool print) {
  const u1_geom g = CurrentGeom();
  double out[2];
  u1SafeCall(u1_polyakov(&g, prec_of<Real>(), lat->getPtr(), out));
  complexd poly(out[0], out[1]);
  if (print) std::cout << "L: " << poly.real() << '\t' << poly.imag() << "\t|L|: " << poly.abs() << std::endl;
  return poly;
}
complexd Polyakov(Array<double>* lat, bool print) { return polyakov_dev(lat, print); }
complexd Polyakov(Array<float>* lat, bool print) { return polyakov_dev(lat, print); }

void polyakov(double* lat, double* poly) {
  const u1_geom g = CurrentGeom();
  const size_t bytes = 4 * (size_t)Volume() * sizeof(double);
  void* d = dev_malloc(bytes);
  u1SafeCall(u1_memcpy(d, lat, bytes, U1_H2D));
  u1SafeCall(u1_polyakov(&g, U1_F64, d, poly));
  dev_free(d);
}

// polyakov.cu:618-654
template <class Real>
static Array<complexd>* poly2_dev(Array<Real>* lat, bool multihit) {
  const u1_geom g = CurrentGeom();
  const int radius = Grid(0) / 2;
  const int mode = multihit ? PARAMS::mhit_kind : U1_POLY_PLAIN;
  const u1_metro mh = MultiHitMetro();
  Array<complexd>* res = new Array<complexd>(Host, radius);
  u1SafeCall(u1_poly2(&g, prec_of<Real>(), lat->getPtr(), mode, &mh, host_rng.seed, (uint32_t)poly2_calls++, 0,
                      (double*)res->getPtr()));
  std::ofstream fileout = open_out((multihit ? "Pot_mhit_" : "Pot_") + GetLatticeNameI() + ".dat");
  for (int r = 0; r < radius; ++r) {
    if (getVerbosity() >= VERBOSE) std::cout << r + 1 << '\t' << res->at(r).real() << '\t' << res->at(r).imag() << std::endl;
    fileout << r + 1 << '\t' << res->at(r).real() << '\t' << res->at(r).imag() << std::endl;
  }
  return res;
}
Array<complexd>* Poly2(Array<double>* lat, bool multihit) { return poly2_dev(lat, multihit); }
Array<complexd>* Poly2(Array<float>* lat, bool multihit) { return poly2_dev(lat, multihit); }

// multilevel.cu:639-837
template <class Real>
static MultiLevelRes* multilevel_dev(Array<Real>* lat, CudaRNG* rng, int n4, int k4, int n2, int k2, int metrop,
                                     int ovrn, bool field) {
  if (Grid(TDir()) % 4 != 0) {
    std::cout << "Error: Cannot Apply MultiLevel Algorithm...\n Nt is not multiple of 4...\n Exiting..." << std::endl;
    std::exit(1);
  }
  const u1_geom g = CurrentGeom();
  const u1_metro m = CurrentMetro();
  const u1_metro mh = MultiHitMetro();
  const int radius = Grid(0) / 2;
  MultiLevelRes* res = new MultiLevelRes;
  res->poly = new Array<complexd>(Host, radius);
  if (field) res->ppSpace = new Array<complexd>(Device, (size_t)SpatialVolume() * radius * (Dirs() - 1));
  PhiloxState* st = rng->getPtr();
  u1SafeCall(u1_multilevel(&g, prec_of<Real>(), lat->getPtr(), n4, k4, n2, k2, metrop, ovrn, &m, PARAMS::mhit_kind,
                           &mh, st->seed, (uint32_t)st->mcalls++, (double*)res->poly->getPtr(),
                           field ? res->ppSpace->getPtr() : nullptr));
  std::string filename = "Pot_mlevel_" + GetLatticeNameI();
  filename += "_" + ToString(n4) + "_" + ToString(k4);
  filename += "_" + ToString(n2) + "_" + ToString(k2);
  filename += "_" + ToString(metrop) + "_" + ToString(ovrn);
  filename += ".dat";
  std::ofstream fileout = open_out(filename);
  for (int r = 0; r < radius; ++r) {
    if (getVerbosity() >= VERBOSE) std::cout << r + 1 << '\t' << res->poly->at(r) << std::endl;
    fileout << r + 1 << '\t' << res->poly->at(r) << std::endl;
  }
  return res;
}
Array<complexd>* MultiLevel(Array<double>* lat, CudaRNG* rng, int n4, int k4, int n2, int k2, int metrop, int ovrn) {
  MultiLevelRes* r = multilevel_dev(lat, rng, n4, k4, n2, k2, metrop, ovrn, false);
  Array<complexd>* out = r->poly;
  r->poly = 0;
  delete r;
  return out;
}
Array<complexd>* MultiLevel(Array<float>* lat, CudaRNG* rng, int n4, int k4, int n2, int k2, int metrop, int ovrn) {
  MultiLevelRes* r = multilevel_dev(lat, rng, n4, k4, n2, k2, metrop, ovrn, false);
  Array<complexd>* out = r->poly;
  r->poly = 0;
  delete r;
  return out;
}
MultiLevelRes* MultiLevelField(Array<double>* lat, CudaRNG* rng, int n4, int k4, int n2, int k2, int metrop,
                               int ovrn) {
  return multilevel_dev(lat, rng, n4, k4, n2, k2, metrop, ovrn, true);
}

// convertOrder.cu:136-145
template <class Real>
static Array<Real>* convert_dev(Array<Real>* lat, bool eo_to_no) {
  const u1_geom g = CurrentGeom();
  Array<Real>* out = new Array<Real>(Device, lat->Size());
  u1SafeCall(u1_convert_layout(&g, prec_of<Real>(), lat->getPtr(), out->getPtr(), eo_to_no ? 1 : 0));
  devSync();
  return out;
}
Array<double>* LatticeConvert(Array<double>* lat, bool eo_to_no) { return convert_dev(lat, eo_to_no); }
Array<float>* LatticeConvert(Array<float>* lat, bool eo_to_no) { return convert_dev(lat, eo_to_no); }

// ------------------------------------------------------------------ checkpoint
template <class Real>
static void save_cfg(Array<Real>* lat, CudaRNG* rng, const std::string& path) {
  const u1_geom g = CurrentGeom();
  u1SafeCall(u1_config_save(path.c_str(), &g, prec_of<Real>(), lat->getPtr(), rng->getPtr()->seed,
                            rng->getPtr()->sweep));
}
template <class Real>
static void load_cfg(Array<Real>* lat, CudaRNG* rng, const std::string& path) {
  const u1_geom g = CurrentGeom();
  uint64_t seed = 0, sweep = 0;
  u1SafeCall(u1_config_load(path.c_str(), &g, prec_of<Real>(), lat->getPtr(), &seed, &sweep));
  rng->getPtr()->seed = seed;
  rng->getPtr()->sweep = sweep;
}
void SaveConfiguration(Array<double>* lat, CudaRNG* rng, const std::string& path) { save_cfg(lat, rng, path); }
void LoadConfiguration(Array<double>* lat, CudaRNG* rng, const std::string& path) { load_cfg(lat, rng, path); }
void SaveConfiguration(Array<float>* lat, CudaRNG* rng, const std::string& path) { save_cfg(lat, rng, path); }
void LoadConfiguration(Array<float>* lat, CudaRNG* rng, const std::string& path) { load_cfg(lat, rng, path); }

// plaquette_comps.cu:148-227
void PlaquetteFields(Array<double>* lat, PlaqFieldArg* pf, bool spacetime, bool evenoddOrder) {
  if (Dirs() < 4) {
    std::cout << "Only implemented for the 4D case...." << std::endl;
    Finalize(1);
  }
  const u1_geom g = CurrentGeom();
  const size_t size = spacetime ? (size_t)Volume() : (size_t)SpatialVolume();
  if (pf->plaqfield) dev_free(pf->plaqfield);
  if (pf->plaq) host_free(pf->plaq);
  pf->plaqfield = (complexd*)dev_malloc(6 * size * sizeof(complexd));
  pf->plaq = (complexd*)safe_malloc(6 * sizeof(complexd));
  pf->size = size;
  double means[12];
  u1SafeCall(u1_plaquette_fields(&g, U1_F64, lat->getPtr(), spacetime ? 1 : 0, evenoddOrder ? 1 : 0, pf->plaqfield,
                                 means));
  complexd mean = 0.;
  for (int i = 0; i < 6; i++) {
    pf->plaq[i] = complexd(means[2 * i], means[2 * i + 1]);
    mean += pf->plaq[i];
    if (getVerbosity() >= VERBOSE) std::cout << "plaq(" << i << "): " << pf->plaq[i] << std::endl;
  }
  mean /= 6.0;
  if (getVerbosity() >= VERBOSE) std::cout << "Mean plaquette: " << mean << std::endl;
}

}  // namespace U1
