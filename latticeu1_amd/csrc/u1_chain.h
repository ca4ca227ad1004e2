// u1_chain.h -- schedule of a fused parity pass (k_chain): the four phases
// (parity, 0), ..., (parity, 3) of one Metropolis or overrelaxation pass in
// one launch, ordered by per-(phase, plane) completion counters instead of
// launch boundaries.
//
// Dependencies.  Phase (p, mu) at site x (parity p) reads, of the links its
// pass writes (parity p), only theta_nu(x) and theta_nu(x + mu - nu) for
// nu != mu (the c and d angles of the staple, staple.h:85-92).  With planes
// Q = (z, t):
//   mu = 1 reads theta_0 at Q only (x + e1 - e0 stays in the plane);
//   mu = 2 reads theta_0, theta_1 at Q and Q + e_z;
//   mu = 3 reads theta_0, theta_1 at Q, Q + e_t and theta_2 at Q, Q - e_z + e_t.
// Those reads must see the new values (read after write).  Conversely, phase
// mu writes theta_mu at Q, which phases nu < mu read (old values) at exactly
// the same pairs of planes (write after read), so one wait covers both: the
// workgroups of (mu, Q) wait until the phase-(mu - 1) work of the planes in
// chain_deps is complete.  The remaining cases follow by transitivity:
// (2, Q) complete implies (1, Q), (1, Q + e_z) and so (0, Q), (0, Q + e_z)
// complete; (2, Q - e_z + e_t) complete implies (1, Q - e_z + e_t) and
// (1, Q + e_t), and so (0, Q + e_t).
//
// Forward progress.  Workgroup b of a launch runs entry b >> 3 of the list of
// XCD b & 7; the eight lists have the same layout (XCD x covers tile x of a
// tz x (8 / tz) tiling of the (z, t) planes into tiles of Nz/tz x Nt/(8/tz)
// planes), and entries are sorted by a schedule time, time(mu, w) = w +
// off[mu], w the plane's index in its tile (z fastest).  The offsets are
// chosen so that every dependency has a strictly smaller time.  Then every dependency of a workgroup has a smaller
// workgroup index, and the unfinished workgroup with the smallest index never
// waits on anything unfinished: whether the dispatcher hands out workgroups in
// index order globally or per XCD, that workgroup is resident (every earlier
// one on its XCD is, or has finished) and completes.  The waits are bounded
// all the same (kChainSpin polls), and a timeout is reported, not hung on.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <vector>

#if defined(__HIPCC__)
#define U1C_HD __host__ __device__ inline
#else
#define U1C_HD inline
#endif

namespace u1d {

// Polls of a completion counter before a wavefront gives up waiting, records
// a timeout and goes on: with the poll's round trip and sleep about 0.1 s,
// against waits of microseconds when the dependencies run.
constexpr uint32_t kChainSpin = 1u << 16;
// List entries: mu in bits 0-1, the chunk of the plane in bits 2-9, the
// plane's index in its tile in bits 10-31.
constexpr int kChainMaxChunks = 255;
constexpr uint32_t kChainMaxTilePlanes = 1u << 22;

// The (z, t) planes whose phase-(mu - 1) workgroups must be complete before
// phase mu runs on plane (z, t); returns their number (0 for mu = 0).
U1C_HD int chain_deps(int mu, int z, int t, int Nz, int Nt, int (&dz)[2], int (&dt)[2]) {
  if (mu == 0) return 0;
  dz[0] = z;
  dt[0] = t;
  if (mu == 1) return 1;
  if (mu == 2) {
    dz[1] = z + 1 == Nz ? 0 : z + 1;
    dt[1] = t;
  } else {
    dz[1] = z == 0 ? Nz - 1 : z - 1;
    dt[1] = t + 1 == Nt ? 0 : t + 1;
  }
  return 2;
}

// Tile of XCD x in a tz x (8 / tz) tiling: z in [(x % tz) bz, + bz), t in
// [(x / tz) bt, + bt); plane w of the tile, z fastest.
U1C_HD void chain_tile_plane(int x, int w, int tz, int bz, int bt, int* z, int* t) {
  *z = (x % tz) * bz + w % bz;
  *t = (x / tz) * bt + w / bz;
}
U1C_HD uint32_t chain_entry(int mu, int chunk, int w) {
  return (uint32_t)mu | ((uint32_t)chunk << 2) | ((uint32_t)w << 10);
}

// A fused pass needs a tz x (8 / tz) tiling (tz in {1, 2, 4, 8} dividing Nz,
// 8 / tz dividing Nt) and at most kChainMaxChunks workgroups per plane.
U1C_HD bool chain_possible(int Nz, int Nt, int bpp, int tz) {
  return (tz == 1 || tz == 2 || tz == 4 || tz == 8) && Nz >= tz && Nz % tz == 0 && Nt >= 8 / tz &&
         Nt % (8 / tz) == 0 && bpp >= 1 && bpp <= kChainMaxChunks &&
         (uint32_t)(Nz / tz) * (uint32_t)(Nt / (8 / tz)) < kChainMaxTilePlanes;
}

// Host side: the schedule list.
struct ChainSchedule {
  int tz = 0, bz = 0, bt = 0, bpp = 0, margin = 0;
  int off[4] = {0, 0, 0, 0};  // schedule time of phase mu on tile plane w: w + off[mu]
  std::vector<uint32_t> list;  // one XCD's list; the grid is 8 * list.size() workgroups
};

// Index of plane (z, t) within its tile and the tile's XCD.
inline int chain_tile_index(int z, int t, int tz, int bz, int bt, int* x) {
  *x = (z / bz) + tz * (t / bt);
  return (z % bz) + bz * (t % bt);
}

// off[mu] = off[mu - 1] + need[mu] + 1 + margin, need[mu] the largest
// w_dep - w over the phase-mu dependencies: every dependency lies at least
// margin + 1 schedule steps back.  A dependency fewer steps back than the
// launch holds in flight is often still running, and its dependant spins;
// the margin keeps them apart.  need[3] grows with the tile's z extent (a
// plane's phase-3 work depends on the phase-2 work of the plane at z - 1,
// t + 1, up to 2 bz - 1 places later in tile order), and the planes an XCD
// touches between a plane's phase 0 and its phase 3 number off[3]: narrow
// tiles and small margins keep that window in the XCD's L2.
inline ChainSchedule chain_schedule(int Nz, int Nt, int bpp, int tz, int margin) {
  ChainSchedule s;
  s.tz = tz;
  s.bz = Nz / tz;
  s.bt = Nt / (8 / tz);
  s.bpp = bpp;
  s.margin = margin;
  const int A = s.bz * s.bt;
  int need[4] = {0, 0, 0, 0};
  for (int x = 0; x < 8; ++x)
    for (int w = 0; w < A; ++w)
      for (int mu = 1; mu < 4; ++mu) {
        int z, t, dz[2], dt[2];
        chain_tile_plane(x, w, tz, s.bz, s.bt, &z, &t);
        const int n = chain_deps(mu, z, t, Nz, Nt, dz, dt);
        for (int i = 0; i < n; ++i) {
          int xd;
          need[mu] = std::max(need[mu], chain_tile_index(dz[i], dt[i], tz, s.bz, s.bt, &xd) - w);
        }
      }
  for (int mu = 1; mu < 4; ++mu) s.off[mu] = s.off[mu - 1] + need[mu] + 1 + margin;
  for (int tm = 0; tm < A + s.off[3]; ++tm)
    for (int mu = 0; mu < 4; ++mu) {
      const int w = tm - s.off[mu];
      if (w < 0 || w >= A) continue;
      for (int c = 0; c < bpp; ++c) s.list.push_back(chain_entry(mu, c, w));
    }
  return s;
}

}  // namespace u1d
