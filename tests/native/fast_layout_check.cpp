// Host-side check of the launch layout of the specialised SU(2) d=4 step
// (su2_md_fast.hpp): for each geometry the (block, thread) -> (s, mu) map must
// hit every link exactly once, and the neighbour offsets must equal the
// periodic lexicographic neighbours ((t Lx + x) Ly + y) Lz + z.
// Prints one line per geometry and "ALL OK" when every check passed.
#include <cstdio>
#include <vector>

#include "../../su2_amd/csrc/su2_md_tw.hpp"

using namespace su2k;

static int check(const int L[4], bool tblock, bool xwalk, bool wmu = false) {
  const int V = L[0] * L[1] * L[2] * L[3];
  FastGeo g;
  if (!make_fast_geo(L, V, tblock, g, xwalk, wmu)) {
    printf("%dx%dx%dx%d tblock=%d: not applicable\n", L[0], L[1], L[2], L[3], (int)tblock);
    return 0;
  }
  const uint32_t nb = 4u * (uint32_t)(V / NTF);
  if (g.tblock && nb % 32 != 0) {
    printf("FAIL grid %u not a multiple of 32\n", nb);
    return 1;
  }
  std::vector<unsigned char> seen((size_t)4 * V, 0);
  const int str[4] = {L[1] * L[2] * L[3], L[2] * L[3], L[3], 1};
  for (uint32_t b = 0; b < nb; ++b) {
    for (uint32_t tid = 0; tid < (uint32_t)NTF; ++tid) {
      const FastSite f = fast_site(g, b, nb, tid);
      if (wmu && f.mu != (int)(tid >> 6)) {
        printf("FAIL wmu: wave %u of block %u has mu %d\n", tid >> 6, b, f.mu);
        return 1;
      }
      if (f.mu < 0 || f.mu > 3 || f.s < 0 || f.s >= V) {
        printf("FAIL out of range b=%u tid=%u mu=%d s=%d\n", b, tid, f.mu, f.s);
        return 1;
      }
      if (seen[(size_t)f.mu * V + f.s]++) {
        printf("FAIL link (s=%d, mu=%d) visited twice\n", f.s, f.mu);
        return 1;
      }
      int c[4], rest = f.s;
      for (int d = 0; d < 4; ++d) {
        c[d] = rest / str[d];
        rest -= c[d] * str[d];
      }
      for (int d = 0; d < 4; ++d) {
        const int up = f.s + ((c[d] + 1) % L[d] - c[d]) * str[d];
        const int dn = f.s + ((c[d] + L[d] - 1) % L[d] - c[d]) * str[d];
        if (f.s + f.up[d] != up || f.s + f.dn[d] != dn) {
          printf("FAIL neighbours of s=%d along %d\n", f.s, d);
          return 1;
        }
      }
    }
  }
  for (size_t i = 0; i < seen.size(); ++i)
    if (seen[i] != 1) {
      printf("FAIL link %zu not visited\n", i);
      return 1;
    }
  // the t-walk shape (t-blocked layouts): 32 nb3 blocks, each thread one
  // (3-volume site, mu) walking the Tc slices of its XCD's chunk
  // (wmu: the mapping of the general-exchange t-walk, one direction per wave)
  // (lc < 3: t-chunks of Lt / 2^lc slices, XCD k walking chunk k % 2^lc of
  // 3-volume part k / 2^lc)
  for (int lc = 3; lc >= 0 && g.tblock; --lc) {
    const int C = 1 << lc, P = 8 / C;
    if (lc < 3 && (L[0] % C != 0 || (wmu ? g.nb64 : g.nb3) % P != 0)) continue;
    FastGeo gw = g;
    gw.tw_lc = lc;
    gw.Tc = L[0] / C;
    std::vector<unsigned char> seen_t((size_t)4 * V, 0);
    const uint32_t nbt = wmu ? tw_blocks_w(gw) : tw_blocks(gw);
    for (uint32_t b = 0; b < nbt; ++b)
      for (uint32_t tid = 0; tid < (uint32_t)NTF; ++tid) {
        const TwSite f = wmu ? tw_site_w(gw, b, tid) : tw_site(gw, b, tid);
        if (f.mu < 0 || f.mu > 3 || f.o3 < 0 || f.o3 >= g.vol3 || f.t0 < 0 || f.t0 + gw.Tc > L[0]) {
          printf("FAIL t-walk range b=%u tid=%u mu=%d o3=%d t0=%d\n", b, tid, f.mu, f.o3, f.t0);
          return 1;
        }
        // the whole-row exchange (64 % Lz == 0) on either mapping: lane l's
        // z is l mod Lz, and the lane the exchange reads holds the +z site
        if (64 % L[3] == 0) {
          const uint32_t lane = tid & 63u, zl = lane & (uint32_t)(L[3] - 1);
          const uint32_t nb_lane = zl == (uint32_t)(L[3] - 1) ? lane - (L[3] - 1) : lane + 1;
          const TwSite n = wmu ? tw_site_w(gw, b, tid - lane + nb_lane) : tw_site(gw, b, tid - lane + nb_lane);
          if ((uint32_t)(f.o3 % L[3]) != zl || n.o3 != f.o3 + f.up[3] || n.mu != f.mu) {
            printf("FAIL t-walk +z lane of b=%u tid=%u\n", b, tid);
            return 1;
          }
        }
        for (int kt = 0; kt < gw.Tc; ++kt) {
          const int t = f.t0 + kt, s = t * g.vol3 + f.o3;
          if (seen_t[(size_t)f.mu * V + s]++) {
            printf("FAIL t-walk link (s=%d, mu=%d) visited twice\n", s, f.mu);
            return 1;
          }
          int up[4], dn[4];
          tw_t_offsets(g, t, up[0], dn[0]);
          for (int d = 1; d < 4; ++d) {
            up[d] = f.up[d];
            dn[d] = f.dn[d];
          }
          int c[4], rest = s;
          for (int d = 0; d < 4; ++d) {
            c[d] = rest / str[d];
            rest -= c[d] * str[d];
          }
          for (int d = 0; d < 4; ++d) {
            const int upn = s + ((c[d] + 1) % L[d] - c[d]) * str[d];
            const int dnn = s + ((c[d] + L[d] - 1) % L[d] - c[d]) * str[d];
            if (s + up[d] != upn || s + dn[d] != dnn) {
              printf("FAIL t-walk neighbours of s=%d along %d\n", s, d);
              return 1;
            }
          }
        }
      }
    for (size_t i = 0; i < seen_t.size(); ++i)
      if (seen_t[i] != 1) {
        printf("FAIL t-walk link %zu not visited\n", i);
        return 1;
      }
    if (lc == 3)
      printf("%dx%dx%dx%d t-walk%s: ok (%u blocks)\n", L[0], L[1], L[2], L[3], wmu ? " wmu" : "", nbt);
    else
      printf("%dx%dx%dx%d t-walk%s %d chunks: ok (%u blocks)\n", L[0], L[1], L[2], L[3],
             wmu ? " wmu" : "", C, nbt);
  }
  // the plaquette t-walk: 8 nb3 blocks, each thread one 3-volume site walking
  // the Tc slices of its XCD's chunk; with 64 % Lz == 0 the +z neighbour's
  // site is held by lane plaq_tw_zup of the same wave
  if (g.tblock && !wmu && !xwalk) {
    std::vector<unsigned char> seen_p((size_t)V, 0);
    const uint32_t nbp = plaq_tw_blocks(g);
    const bool zx = 64 % L[3] == 0;
    for (uint32_t b = 0; b < nbp; ++b)
      for (uint32_t tid = 0; tid < (uint32_t)NTF; ++tid) {
        const PlaqTwSite f = plaq_tw_site(g, b, tid);
        if (f.o3 < 0 || f.o3 >= g.vol3 || f.t0 < 0 || f.t0 + g.Tc > L[0]) {
          printf("FAIL plaquette t-walk range b=%u tid=%u o3=%d t0=%d\n", b, tid, f.o3, f.t0);
          return 1;
        }
        if (zx) {
          const uint32_t lane = tid & 63u, zl = (uint32_t)plaq_tw_zup(g, (int)lane);
          if (zl >= 64u || plaq_tw_site(g, b, tid - lane + zl).o3 != f.o3 + f.upz) {
            printf("FAIL plaquette t-walk +z lane of b=%u tid=%u\n", b, tid);
            return 1;
          }
        }
        int s = f.t0 * g.vol3 + f.o3;
        for (int kt = 0; kt < g.Tc; ++kt) {
          if (s < 0 || s >= V || seen_p[s]++) {
            printf("FAIL plaquette t-walk site s=%d\n", s);
            return 1;
          }
          int up0, dn0;
          tw_t_offsets(g, f.t0 + kt, up0, dn0);
          const int up[4] = {up0, f.upx, f.upy, f.upz};
          int c[4], rest = s;
          for (int d = 0; d < 4; ++d) {
            c[d] = rest / str[d];
            rest -= c[d] * str[d];
          }
          for (int d = 0; d < 4; ++d)
            if (s + up[d] != s + ((c[d] + 1) % L[d] - c[d]) * str[d]) {
              printf("FAIL plaquette t-walk neighbours of s=%d along %d\n", s, d);
              return 1;
            }
          s += up0;
        }
      }
    for (size_t i = 0; i < seen_p.size(); ++i)
      if (seen_p[i] != 1) {
        printf("FAIL plaquette t-walk site %zu not visited\n", i);
        return 1;
      }
    printf("%dx%dx%dx%d plaquette t-walk%s: ok (%u blocks)\n", L[0], L[1], L[2], L[3],
           zx ? " z-exchange" : "", nbp);
  }
  // the thread-per-site shape: V / NTF blocks, every site exactly once
  const uint32_t nbs = (uint32_t)(V / NTF);
  if (g.tblock && nbs % 8 != 0) {
    printf("FAIL site grid %u not a multiple of 8\n", nbs);
    return 1;
  }
  std::vector<unsigned char> seen_s((size_t)V, 0);
  for (uint32_t b = 0; b < nbs; ++b) {
    for (uint32_t tid = 0; tid < (uint32_t)NTF; ++tid) {
      int s, up[4], dn[4];
      fast_site_s(g, b, nbs, tid, s, up, dn);
      if (s < 0 || s >= V || seen_s[s]++) {
        printf("FAIL site map b=%u tid=%u s=%d\n", b, tid, s);
        return 1;
      }
      int c[4], rest = s;
      for (int d = 0; d < 4; ++d) {
        c[d] = rest / str[d];
        rest -= c[d] * str[d];
      }
      for (int d = 0; d < 4; ++d) {
        const int upn = s + ((c[d] + 1) % L[d] - c[d]) * str[d];
        const int dnn = s + ((c[d] + L[d] - 1) % L[d] - c[d]) * str[d];
        if (s + up[d] != upn || s + dn[d] != dnn) {
          printf("FAIL site neighbours of s=%d along %d\n", s, d);
          return 1;
        }
      }
    }
  }
  printf("%dx%dx%dx%d tblock=%d xwalk=%d wmu=%d: ok (%u blocks)\n", L[0], L[1], L[2], L[3], g.tblock,
         g.xwalk, g.wmu, nb);
  return 0;
}

int main() {
  // extents (t, x, y, z)
  const int geoms[][4] = {{32, 8, 8, 8},  {8, 8, 8, 8},     {16, 16, 16, 16}, {32, 32, 32, 32}, {24, 24, 24, 24},
                          {48, 48, 48, 48}, {16, 8, 8, 8},  {12, 8, 8, 8},    {4, 8, 8, 8},
                          {8, 4, 8, 16},  {8, 16, 16, 1},   {40, 16, 4, 4},   {8, 12, 12, 12},
                          {64, 16, 16, 16}, {40, 16, 8, 8}, {32, 8, 16, 16}};
  int fails = 0;
  for (const auto &L : geoms)
    for (int tb = 0; tb < 2; ++tb)
      for (int xw = 0; xw < 3; ++xw) fails += check(L, tb != 0, xw == 1, xw == 2);
  // Regression: a divisor of 1 has no 32-bit multiply-high constant
  // (floor(2^32 / 1) + 1 = 2^32 + 1 truncates to 1, and umulhi(n, 1) = 0).  With
  // Lt = 8 the t-chunk per XCD is Tc = 1, so that constant would give q = 0 for
  // r = 1 and t = xcd + 1 = 8 = Lt for the last XCD: loads past the link field
  // and momentum / link stores up to 512 elements past the end of the mu = 3
  // planes (the 8^4 flow-stage launches of the specialised step).  fast_div
  // codes d = 1 as m = 0; the truncated constant must still be caught here.
  {
    const int L8[4] = {8, 8, 8, 8};
    FastGeo g;
    if (!make_fast_geo(L8, 4096, true, g) || g.Tc != 1 || g.m_Tc != 0) {
      printf("FAIL 8^4 layout: Tc = %d, m_Tc = %u\n", g.Tc, g.m_Tc);
      return 1;
    }
    g.m_Tc = (uint32_t)((((uint64_t)1) << 32) / 1 + 1);  // the truncated constant: 1
    int max_s = -1;
    for (uint32_t b = 0; b < 64; ++b)
      for (uint32_t tid = 0; tid < (uint32_t)NTF; ++tid) {
        const FastSite f = fast_site(g, b, 64, tid);
        if (f.s > max_s) max_s = f.s;
      }
    if (max_s < 4096) {
      printf("FAIL the truncated d = 1 constant stays in range (max s %d)\n", max_s);
      return 1;
    }
    printf("8x8x8x8 truncated d=1 multiplier: site %d >= V = 4096 (caught)\n", max_s);
    if (fast_div(12345u, fast_div_magic(1)) != 12345u) {
      printf("FAIL fast_div by 1\n");
      return 1;
    }
  }
  // the multiply-high division against the exact quotient over its stated range
  for (uint32_t d = 1; d <= 2048; ++d) {
    const uint32_t m = fast_div_magic(d);
    const uint32_t nmax = (uint32_t)((1ull << 32) / d);
    const uint32_t step = nmax / 4099 + 1;
    for (uint64_t n = 0; n < nmax; n += step)
      if (fast_div((uint32_t)n, m) != (uint32_t)n / d) {
        printf("FAIL fast_div %u / %u\n", (uint32_t)n, d);
        return 1;
      }
    if (fast_div(nmax - 1, m) != (nmax - 1) / d) {
      printf("FAIL fast_div %u / %u\n", nmax - 1, d);
      return 1;
    }
  }
  printf(fails ? "FAILED\n" : "ALL OK\n");
  return fails ? 1 : 0;
}
