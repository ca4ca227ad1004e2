// CPU check of the fused parity pass (u1_chain.h, k_chain):
//
// 1. Schedule: for several (Nz, Nt, workgroups per plane, tiles along z), the eight XCD
//    lists cover every (phase, plane, chunk) exactly once, each list is sorted
//    by schedule time, and every dependency of chain_deps has a strictly
//    smaller time -- so every workgroup a workgroup waits on has a smaller
//    workgroup index (the forward-progress argument).
// 2. Dependencies: an emulation of the launch with adversarial interleavings.
//    Each workgroup task reads the angles that no task of the pass writes
//    (other parity; its own link; theta_nu for nu > mu) when it starts, waits
//    until the chain_deps tasks of phase mu - 1 are complete, reads theta_nu
//    (nu < mu), and writes its links later; tasks advance in random order.
//    The final lattice must equal four phases applied one after another, for
//    every interleaving tried.  The update is an integer mix of the link and
//    its 18 staple angles (the read set of staple.h:85-92), so any stale or
//    premature read changes the result.  A planted mistake (one dependency
//    dropped) must be detected.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "u1_chain.h"

using u1d::chain_deps;

namespace {

int failures = 0;
#define CHECK(cond, ...)          \
  do {                            \
    if (!(cond)) {                \
      std::printf(__VA_ARGS__);   \
      std::printf("\n");          \
      ++failures;                 \
    }                             \
  } while (0)

// ------------------------------------------------------------ 1. schedule
void check_schedule(int Nz, int Nt, int bpp, int tz, int margin) {
  if (!u1d::chain_possible(Nz, Nt, bpp, tz)) return;
  const u1d::ChainSchedule s = u1d::chain_schedule(Nz, Nt, bpp, tz, margin);
  const int A = s.bz * s.bt;
  CHECK(s.list.size() == (size_t)4 * A * bpp, "schedule %dx%d bpp %d: %zu entries", Nz, Nt, bpp, s.list.size());
  auto tm = [&](uint32_t e) { return (int)(e >> 10) + s.off[e & 3u]; };
  std::vector<int> seen((size_t)4 * Nz * Nt * bpp, 0);
  std::vector<int> when((size_t)4 * Nz * Nt, -1);  // schedule time of (mu, plane)
  for (size_t i = 0; i < s.list.size(); ++i) {
    const uint32_t e = s.list[i];
    if (i > 0) CHECK(tm(s.list[i - 1]) <= tm(e), "schedule %dx%d: list not sorted at %zu", Nz, Nt, i);
    const int mu = (int)(e & 3u), chunk = (int)((e >> 2) & 0xFFu), w = (int)(e >> 10);
    CHECK(w < A && chunk < bpp, "schedule %dx%d: bad entry %u", Nz, Nt, e);
    for (int x = 0; x < 8; ++x) {
      int z, t;
      u1d::chain_tile_plane(x, w, tz, s.bz, s.bt, &z, &t);
      const size_t pl = (size_t)z + (size_t)Nz * t;
      ++seen[((size_t)mu * Nz * Nt + pl) * bpp + chunk];
      when[(size_t)mu * Nz * Nt + pl] = tm(e);
    }
  }
  for (size_t k = 0; k < seen.size(); ++k) CHECK(seen[k] == 1, "schedule %dx%d: task %zu seen %d times", Nz, Nt, k, seen[k]);
  for (int mu = 1; mu < 4; ++mu)
    for (int t = 0; t < Nt; ++t)
      for (int z = 0; z < Nz; ++z) {
        int dz[2], dt[2];
        const int n = chain_deps(mu, z, t, Nz, Nt, dz, dt);
        for (int i = 0; i < n; ++i)
          CHECK(when[(size_t)(mu - 1) * Nz * Nt + dz[i] + (size_t)Nz * dt[i]] + margin <
                    when[(size_t)mu * Nz * Nt + z + (size_t)Nz * t],
                "schedule %dx%d: dependency of (%d, %d, %d) not %d steps earlier", Nz, Nt, mu, z, t, margin + 1);
      }
}

// ------------------------------------------------------------ 2. emulation
struct Lat {
  int N[4];
  int V;
  std::vector<uint64_t> th[4];
  int idx(const int* x) const { return x[0] + N[0] * (x[1] + N[1] * (x[2] + N[2] * x[3])); }
  int shift(int site, int d, int s) const {
    int x[4];
    int r = site;
    for (int i = 0; i < 4; ++i) {
      x[i] = r % N[i];
      r /= N[i];
    }
    x[d] = (x[d] + s + N[d]) % N[d];
    return idx(x);
  }
};

uint64_t mix(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  h *= 0xBF58476D1CE4E5B9ull;
  return h ^ (h >> 31);
}

// The 18 staple angles of (x, mu) in a fixed order, then the link itself.
// dep_mask: which nu have c, d read in the dependent step (nu < mu); stage 0
// reads everything else, stage 1 the rest.
struct Reads {
  uint64_t v[19];
};
void read_inputs(const Lat& L, int x, int mu, int stage, Reads& r) {
  int k = 0;
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    const bool dep = nu < mu;
    const int xm = L.shift(x, mu, 1), xn = L.shift(x, nu, 1), xbn = L.shift(x, nu, -1);
    const int xmbn = L.shift(xm, nu, -1);
    if (stage == 0) {
      r.v[k + 0] = L.th[nu][xm];
      r.v[k + 1] = L.th[mu][xn];
      r.v[k + 4] = L.th[mu][xbn];
      r.v[k + 5] = L.th[nu][xbn];
      if (!dep) {
        r.v[k + 2] = L.th[nu][x];
        r.v[k + 3] = L.th[nu][xmbn];
      }
    } else if (dep) {
      r.v[k + 2] = L.th[nu][x];
      r.v[k + 3] = L.th[nu][xmbn];
    }
    k += 6;
  }
  if (stage == 0) r.v[18] = L.th[mu][x];
}
uint64_t update_value(int mu, const Reads& r) {
  uint64_t h = 0x1234567ull + (uint64_t)mu;
  for (int i = 0; i < 19; ++i) h = mix(h, r.v[i]);
  return h;
}

struct Task {
  int mu, z, t;
  std::vector<int> sites;  // sites of this chunk (parity p)
  std::vector<Reads> rd;
  int state = 0;  // 0 not started, 1 independent reads done, 2 all reads done, 3 written
};

// Returns true if the emulated launch matches the sequential phases.
bool emulate(const int* N, int parity, int bpp, uint64_t seed, bool drop_dep) {
  Lat L;
  for (int i = 0; i < 4; ++i) L.N[i] = N[i];
  L.V = N[0] * N[1] * N[2] * N[3];
  std::mt19937_64 rng(seed);
  for (int mu = 0; mu < 4; ++mu) {
    L.th[mu].resize(L.V);
    for (auto& v : L.th[mu]) v = rng();
  }
  // sequential reference
  Lat R = L;
  for (int mu = 0; mu < 4; ++mu) {
    std::vector<uint64_t> out(R.V);
    for (int x = 0; x < R.V; ++x) {
      int c[4], r = x;
      for (int i = 0; i < 4; ++i) { c[i] = r % N[i]; r /= N[i]; }
      if (((c[0] + c[1] + c[2] + c[3]) & 1) != parity) continue;
      Reads rd;
      read_inputs(R, x, mu, 0, rd);
      read_inputs(R, x, mu, 1, rd);
      out[x] = update_value(mu, rd);
    }
    for (int x = 0; x < R.V; ++x) {
      int c[4], r = x;
      for (int i = 0; i < 4; ++i) { c[i] = r % N[i]; r /= N[i]; }
      if (((c[0] + c[1] + c[2] + c[3]) & 1) == parity) R.th[mu][x] = out[x];
    }
  }
  // tasks: (mu, plane, chunk)
  const int Nz = N[2], Nt = N[3];
  std::vector<Task> tasks;
  std::vector<int> done((size_t)4 * Nz * Nt, 0);
  for (int mu = 0; mu < 4; ++mu)
    for (int t = 0; t < Nt; ++t)
      for (int z = 0; z < Nz; ++z) {
        std::vector<int> ps;
        for (int y = 0; y < N[1]; ++y)
          for (int x0 = 0; x0 < N[0]; ++x0)
            if (((x0 + y + z + t) & 1) == parity) {
              const int c[4] = {x0, y, z, t};
              ps.push_back(L.idx(c));
            }
        const int P = (int)ps.size(), S = (P + bpp - 1) / bpp;
        for (int k = 0; k < bpp; ++k) {
          Task tk;
          tk.mu = mu;
          tk.z = z;
          tk.t = t;
          for (int i = k * S; i < std::min(P, (k + 1) * S); ++i) tk.sites.push_back(ps[i]);
          tk.rd.resize(tk.sites.size());
          tasks.push_back(tk);
        }
      }
  auto deps_done = [&](const Task& tk) {
    int dz[2], dt[2];
    int n = chain_deps(tk.mu, tk.z, tk.t, Nz, Nt, dz, dt);
    if (drop_dep && tk.mu == 3) n = 1;  // planted mistake: (2, Q - e_z + e_t) not waited on
    for (int i = 0; i < n; ++i)
      if (done[(size_t)(tk.mu - 1) * Nz * Nt + dz[i] + (size_t)Nz * dt[i]] < bpp) return false;
    return true;
  };
  size_t finished = 0;
  while (finished < tasks.size()) {
    std::vector<size_t> ready;
    for (size_t i = 0; i < tasks.size(); ++i) {
      const Task& tk = tasks[i];
      if (tk.state == 0 || tk.state == 2 || (tk.state == 1 && deps_done(tk))) ready.push_back(i);
    }
    if (ready.empty()) {
      std::printf("emulation deadlock\n");
      return false;
    }
    Task& tk = tasks[ready[rng() % ready.size()]];
    if (tk.state == 0) {
      for (size_t i = 0; i < tk.sites.size(); ++i) read_inputs(L, tk.sites[i], tk.mu, 0, tk.rd[i]);
      tk.state = 1;
    } else if (tk.state == 1) {
      for (size_t i = 0; i < tk.sites.size(); ++i) read_inputs(L, tk.sites[i], tk.mu, 1, tk.rd[i]);
      tk.state = 2;
    } else {
      for (size_t i = 0; i < tk.sites.size(); ++i) L.th[tk.mu][tk.sites[i]] = update_value(tk.mu, tk.rd[i]);
      tk.state = 3;
      ++done[(size_t)tk.mu * Nz * Nt + tk.z + (size_t)Nz * tk.t];
      ++finished;
    }
  }
  for (int mu = 0; mu < 4; ++mu)
    if (L.th[mu] != R.th[mu]) return false;
  return true;
}

}  // namespace

int main() {
  const int sched_geoms[][3] = {{2, 4, 1},  {4, 4, 3},  {8, 8, 2},   {6, 12, 4}, {12, 8, 5},  {24, 12, 1},
                                {48, 48, 5}, {96, 96, 18}, {2, 8, 255}, {16, 32, 9}, {8, 2, 2},  {64, 16, 3}};
  int nsched = 0;
  for (const auto& g : sched_geoms)
    for (int tz : {1, 2, 4, 8})
      if (u1d::chain_possible(g[0], g[1], g[2], tz))
        for (int margin : {0, 3, 16}) {
          check_schedule(g[0], g[1], g[2], tz, margin);
          ++nsched;
        }
  CHECK(nsched >= 20, "only %d schedules checked", nsched);
  std::printf("schedule %s\n", failures == 0 ? "ok" : "FAILED");
  const int f0 = failures;

  const int emu_geoms[][4] = {{4, 4, 2, 4}, {4, 4, 4, 4}, {2, 4, 6, 4}, {4, 2, 4, 8}, {6, 4, 2, 8}};
  int runs = 0;
  for (const auto& N : emu_geoms)
    for (int parity = 0; parity < 2; ++parity)
      for (int bpp : {1, 3})
        for (uint64_t seed = 1; seed <= 12; ++seed) {
          ++runs;
          CHECK(emulate(N, parity, bpp, seed * 7919 + parity, false), "emulation mismatch: %dx%dx%dx%d parity %d bpp %d seed %llu",
                N[0], N[1], N[2], N[3], parity, bpp, (unsigned long long)seed);
        }
  // the planted mistake must be caught by some interleaving
  int caught = 0;
  for (uint64_t seed = 1; seed <= 40 && !caught; ++seed) {
    const int N[4] = {4, 4, 4, 4};
    if (!emulate(N, 0, 1, seed, true)) ++caught;
  }
  CHECK(caught > 0, "the emulation did not detect a dropped dependency");
  std::printf("emulation %s (%d interleavings)\n", failures == f0 ? "ok" : "FAILED", runs);
  return failures == 0 ? 0 : 1;
}
