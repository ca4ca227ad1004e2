// CPU check of the device-side indexing (u1_device.h: make_magic, mdiv,
// site_of): for every even/odd site of several geometries, the coordinates and
// the twelve staple neighbour positions the update kernels form
//   opp + fwd[mu], opp + fwd[nu], same, same + fwd[mu] + bwd[nu], opp + bwd[nu]
// must equal the brute-force positions from explicit coordinates and lie inside
// the half-plane [0, V/2).  Also checks the magic division exhaustively for
// small divisors.  Host-only program (no kernel launch).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../latticeu1_amd/csrc/u1_device.h"

static int lexi(const int* N, const int* x) { return x[0] + N[0] * (x[1] + N[1] * (x[2] + N[2] * x[3])); }

int main() {
  // magic division
  for (uint32_t d = 1; d <= 300; ++d) {
    const u1d::MagicDiv md = u1d::make_magic(d);
    for (uint32_t n = 0; n < 200000; n += 7)
      if (u1d::mdiv(n, md.m, md.s) != n / d) { printf("mdiv fail d=%u n=%u\n", d, n); return 1; }
    for (uint32_t n = (1u << 31) - 5000; n < (1u << 31); ++n)
      if (u1d::mdiv(n, md.m, md.s) != n / d) { printf("mdiv fail d=%u n=%u\n", d, n); return 1; }
  }
  const int geoms[][4] = {{2, 2, 2, 2}, {4, 4, 4, 4}, {8, 8, 8, 8}, {6, 10, 4, 12}, {4, 4, 4, 8},
                          {12, 12, 12, 16}, {24, 24, 24, 32}, {16, 16, 16, 32}, {48, 48, 48, 48}};
  long checked = 0;
  for (const auto& G : geoms) {
    u1d::RepDesc r{};
    for (int i = 0; i < 4; ++i) r.N[i] = G[i];
    r.nxh = G[0] / 2;
    r.V = G[0] * G[1] * G[2] * G[3];
    r.HV = r.V / 2;
    r.sy = r.nxh; r.sz = r.nxh * G[1]; r.st = r.nxh * G[1] * G[2];
    u1d::MagicDiv a = u1d::make_magic(r.nxh), b = u1d::make_magic(G[1]), c = u1d::make_magic(G[2]);
    r.div_nxh_m = a.m; r.div_nxh_s = a.s; r.div_ny_m = b.m; r.div_ny_s = b.s; r.div_nz_m = c.m; r.div_nz_s = c.s;
    const int step = r.V > 1000000 ? 13 : 1;
    for (int parity = 0; parity < 2; ++parity)
      for (int id = 0; id < r.HV; id += step) {
        u1d::Site s;
        u1d::site_of(r, id, parity, s);
        for (int k = 0; k < 4; ++k)
          if (s.x[k] < 0 || s.x[k] >= G[k]) { printf("coord out of range\n"); return 1; }
        if ((lexi(G, s.x) >> 1) != id || ((s.x[0] + s.x[1] + s.x[2] + s.x[3]) & 1) != parity) {
          printf("coords wrong %d %d\n", id, parity); return 1;
        }
        for (int mu = 0; mu < 4; ++mu)
          for (int nu = 0; nu < 4; ++nu) {
            if (nu == mu) continue;
            auto pos = [&](int dm, int dn) {
              int y[4] = {s.x[0], s.x[1], s.x[2], s.x[3]};
              y[mu] = (y[mu] + dm + G[mu]) % G[mu];
              y[nu] = (y[nu] + dn + G[nu]) % G[nu];
              return lexi(G, y) >> 1;
            };
            const int got[5] = {id + s.fwd[mu], id + s.fwd[nu], id, id + s.fwd[mu] + s.bwd[nu], id + s.bwd[nu]};
            const int want[5] = {pos(1, 0), pos(0, 1), id, pos(1, -1), pos(0, -1)};
            for (int k = 0; k < 5; ++k) {
              if (got[k] != want[k] || got[k] < 0 || got[k] >= r.HV) {
                printf("neighbour %d wrong: G=%dx%dx%dx%d id=%d parity=%d mu=%d nu=%d got=%d want=%d\n", k, G[0], G[1],
                       G[2], G[3], id, parity, mu, nu, got[k], want[k]);
                return 1;
              }
            }
            ++checked;
          }
      }
  }
  // XCD-aware workgroup order is a bijection of [0, n)
  for (uint32_t n = 1; n < 3000; n += (n < 64 ? 1 : 37)) {
    std::vector<unsigned char> hit(n, 0);
    for (uint32_t b = 0; b < n; ++b) {
      const uint32_t L = u1d::xcd_remap(b, n);
      if (L >= n || hit[L]) { printf("xcd_remap not a bijection n=%u b=%u\n", n, b); return 1; }
      hit[L] = 1;
    }
  }
  // update-launch decomposition (phase_site): every site exactly once per
  // parity, right z/t, neighbour offsets equal to brute force
  long psites = 0;
  const int blocks[] = {64, 128, 192, 256, 384, 512, 768, 1024};  // block * sites-per-thread
  for (const auto& G : geoms) {
    u1d::RepDesc r{};
    for (int i = 0; i < 4; ++i) r.N[i] = G[i];
    r.nxh = G[0] / 2;
    r.V = G[0] * G[1] * G[2] * G[3];
    r.HV = r.V / 2;
    r.sy = r.nxh; r.sz = r.nxh * G[1]; r.st = r.nxh * G[1] * G[2];
    u1d::MagicDiv a = u1d::make_magic(r.nxh), b = u1d::make_magic(G[1]), c = u1d::make_magic(G[2]);
    r.div_nxh_m = a.m; r.div_nxh_s = a.s; r.div_ny_m = b.m; r.div_ny_s = b.s; r.div_nz_m = c.m; r.div_nz_s = c.s;
    for (int tiled = 0; tiled < 2; ++tiled) {
      for (int B : blocks) {
        if (r.V > 1000000 && B != 192 && B != 256 && B != 384) continue;
        u1d::set_phase_blocks(r, B);
        u1d::set_phase_tiles(r, tiled != 0);
        if (tiled && !r.tile) continue;  // Nt % 4 != 0: no tiling
        for (int parity = 0; parity < 2; ++parity) {
          std::vector<unsigned char> seen(r.HV, 0);
          for (int blk = 0; blk < r.nblk; ++blk)
            for (int tid = 0; tid < B; ++tid) {
              u1d::PhaseSite ps;
              u1d::phase_site(r, blk, tid, B, parity, ps);
              if (!ps.valid) {
                // tail threads load (and never store): every address in bounds
                for (int d = 0; d < 4; ++d) {
                  const long f = (long)ps.id + ps.fwd[d], bk = (long)ps.id + ps.bwd[d];
                  const long ff = f + ps.fwd[(d + 1) % 4], fb = f + ps.bwd[(d + 1) % 4];
                  if (ps.id >= (uint32_t)r.HV || f < 0 || f >= r.HV || bk < 0 || bk >= r.HV || ff < 0 ||
                      ff >= r.HV || fb < 0 || fb >= r.HV) {
                    printf("phase_site: tail thread out of bounds\n");
                    return 1;
                  }
                }
                continue;
              }
              if (ps.id >= (uint32_t)r.HV || seen[ps.id]) { printf("phase_site: bad/duplicate id %u\n", ps.id); return 1; }
              seen[ps.id] = 1;
              // brute-force coordinates of (id, parity): indexEO restated
              int x[4];
              const int za = ps.id / (G[0] / 2), zb = za / G[1];
              x[1] = za - zb * G[1]; x[3] = zb / G[2]; x[2] = zb - x[3] * G[2];
              x[0] = 2 * (int)ps.id + ((x[1] + x[2] + x[3] + parity) & 1) - za * G[0];
              if ((int)ps.t != x[3] || ps.z != x[2]) { printf("phase_site: z/t wrong\n"); return 1; }
              for (int d = 0; d < 4; ++d)
                for (int sgn = -1; sgn <= 1; sgn += 2) {
                  int y[4] = {x[0], x[1], x[2], x[3]};
                  y[d] = (y[d] + sgn + G[d]) % G[d];
                  const int want = lexi(G, y) >> 1;
                  const int got = (int)ps.id + (sgn > 0 ? ps.fwd[d] : ps.bwd[d]);
                  if (got != want) { printf("phase_site: neighbour d=%d sgn=%d wrong\n", d, sgn); return 1; }
                }
              ++psites;
            }
          for (int k = 0; k < r.HV; ++k)
            if (!seen[k]) { printf("phase_site: site %d not covered (B=%d)\n", k, B); return 1; }
        }
      }
    }
  }
  // fused-pair decomposition: one chunk (workgroup) per plane, the plane
  // covered in steps of spt * block sites (tail threads in bounds); the LDS
  // exchange of phase 0's theta_0: in-plane index id - pbase and that of
  // x + e1 - e0 (id + fwd[1] + bwd[0]) inside [0, P) for every thread, equal
  // to the brute-force in-plane position of the neighbour for valid sites
  long pairsites = 0;
  for (const auto& G : geoms) {
    u1d::RepDesc r{};
    for (int i = 0; i < 4; ++i) r.N[i] = G[i];
    r.nxh = G[0] / 2;
    r.V = G[0] * G[1] * G[2] * G[3];
    r.HV = r.V / 2;
    r.sy = r.nxh; r.sz = r.nxh * G[1]; r.st = r.nxh * G[1] * G[2];
    u1d::MagicDiv a = u1d::make_magic(r.nxh), b = u1d::make_magic(G[1]), c = u1d::make_magic(G[2]);
    r.div_nxh_m = a.m; r.div_nxh_s = a.s; r.div_ny_m = b.m; r.div_ny_s = b.s; r.div_nz_m = c.m; r.div_nz_s = c.s;
    const uint32_t P = (uint32_t)(r.nxh * G[1]);
    if (r.V > 1000000) continue;
    for (int tiled = 0; tiled < 2; ++tiled)
      for (int step : {64, 128, 576, 1024, 2048}) {
        u1d::set_phase_blocks(r, (int)P);
        u1d::set_phase_tiles(r, tiled != 0);
        if (tiled && !r.tile) continue;
        if (r.bpp != 1 || r.nblk != G[2] * G[3]) { printf("pair: not one chunk per plane\n"); return 1; }
        for (int parity = 0; parity < 2; ++parity) {
          std::vector<unsigned char> seen(r.HV, 0);
          for (int blk = 0; blk < r.nblk; ++blk) {
            uint32_t t0 = 0;
            int32_t z0 = 0;
            for (uint32_t base = 0; base < P; base += (uint32_t)step)
              for (uint32_t tid = 0; tid < (uint32_t)step; ++tid) {
                u1d::PhaseSite ps;
                u1d::phase_site(r, blk, base + tid, (uint32_t)step, parity, ps);
                if (ps.id >= (uint32_t)r.HV) { printf("pair: id out of range\n"); return 1; }
                if (base == 0 && tid == 0) {
                  t0 = ps.t;
                  z0 = ps.z;
                }
                if (ps.t != t0 || ps.z != z0) { printf("pair: z/t not uniform over the workgroup\n"); return 1; }
                const uint32_t pbase = ((uint32_t)ps.z + (uint32_t)G[2] * ps.t) * P;
                const uint32_t i = ps.id - pbase, j = i + (uint32_t)(ps.fwd[1] + ps.bwd[0]);
                if (i >= P || j >= P) { printf("pair: LDS index out of the plane\n"); return 1; }
                if (!ps.valid) continue;
                if (seen[ps.id]) { printf("pair: duplicate site\n"); return 1; }
                seen[ps.id] = 1;
                // x + e1 - e0 from explicit coordinates, as an in-plane index
                int x[4];
                const int za = ps.id / (G[0] / 2), zb = za / G[1];
                x[1] = za - zb * G[1]; x[3] = zb / G[2]; x[2] = zb - x[3] * G[2];
                x[0] = 2 * (int)ps.id + ((x[1] + x[2] + x[3] + parity) & 1) - za * G[0];
                int y[4] = {(x[0] - 1 + G[0]) % G[0], (x[1] + 1) % G[1], x[2], x[3]};
                if ((uint32_t)(lexi(G, y) >> 1) - pbase != j) { printf("pair: LDS neighbour wrong\n"); return 1; }
                ++pairsites;
              }
          }
          for (int k = 0; k < r.HV; ++k)
            if (!seen[k]) { printf("pair: site %d not covered\n", k); return 1; }
        }
      }
  }
  printf("pair decomposition ok (%ld site visits)\n", pairsites);
  printf("phase decomposition ok (%ld site visits)\n", psites);
  printf("indexing ok (%ld site/direction pairs)\n", checked);
  return 0;
}
