// su2_md_tw.hpp -- the specialised SU(2) d = 4 "P then U" step walking t
// (get_staples.hh:26-80, flat-gaugemonomial.hh:105-131, update_momenta.hh:27-30,
// update_gauge.hh:11-14): the arithmetic of md_su2_core (su2_md_fast.hpp), bit
// for bit, with fewer link loads.
//
// A thread keeps one 3-volume site and one direction mu and walks the Tc
// t-slices of a chunk upwards.  The staple links that the link at t + 1 shares
// with the link at t stay in registers:
//   * mu != t, plane t: U_t(s+mu-t), U_mu(s-t), U_t(s-t) of the link at t + 1
//     are U_t(s+mu), U_mu(s), U_t(s) of the link at t, and its own link
//     U_mu(s+t) is the link at t's U_mu(s+t): 2 loads instead of 6 + 1;
//   * mu = t, planes nu: U_nu(s) and U_nu(s-nu) of the link at t + 1 are
//     U_nu(s+t) and U_nu(s+t-nu) of the link at t: 4 loads instead of 6.
// With the +-z lane exchange (ZS = 1) that leaves 10-11 link loads per link
// instead of 14-15; the chunk's first slice loads what the slice below would
// have passed on (its t-1 neighbours come from the neighbouring XCD's chunk).
//
// Chunks: C = 2^tw_lc t-chunks of Tc = Lt / C slices; hardware block b runs on
// XCD k = b % 8, which walks chunk k mod C of the 3-volume part k / C (8 / C
// parts; C = 8: one chunk of Lt / 8 slices per XCD and the whole 3-volume).
// Needs the t-blocked layout (Lt % 8 == 0).  k_md_su2_4d_tw: one direction per
// block, 32 nb3 / (8 / C) blocks of NTF threads (nb3 = vol3 / NTF), local =
// b / 8 = 4 q + mu, site o3 = (part nq + q) NTF + tid.
//
// One direction per wave (tw_site_w): a block holds 64 sites (o3 = (part nq +
// q) 64 + lane) and its wave w the direction w, so the four directions of the
// same sites walk t together on one CU.  k_md_su2_4d_twz takes it with the
// general +-z exchange (ZS = 2, any Lz: the neighbour of lane l is lane l +- 1,
// or l -+ (Lz - 1) at a row's wrap, where that lane is in the wave; the lanes
// whose neighbour is in another wave load it in the same memory round trip),
// k_md_su2_4d_tww with the whole-row exchange (ZS = 1).  The _acc kernels are
// the first and last steps of a trajectory on either mapping; k_plaq_su2_4d_tw
// is the plaquette sum walking t.
#pragma once
#include "su2_md_fast.hpp"

namespace su2k {

struct TwSite {
  int mu, o3, t0;
  int up[4], dn[4];  // x, y, z neighbour offsets of o3 in [1..3]; [0] per slice
};

__host__ __device__ __forceinline__ TwSite tw_site(const FastGeo &g, uint32_t b, uint32_t tid) {
  TwSite f;
  const uint32_t xcd = b & 7u, local = b >> 3;
  f.mu = (int)(local & 3u);
  const uint32_t lc = (uint32_t)g.tw_lc;
  const uint32_t chunk = xcd & ((1u << lc) - 1u), part = xcd >> lc;
  const uint32_t nq = (uint32_t)g.nb3 >> (3u - lc);
  const uint32_t o3 = (part * nq + (local >> 2)) * (uint32_t)NTF + tid;
  f.o3 = (int)o3;
  f.t0 = (int)chunk * g.Tc;
  const uint32_t x = fast_div(o3, g.m_plane);
  const uint32_t rem3 = o3 - x * (uint32_t)g.plane;
  const uint32_t y = fast_div(rem3, g.m_Lz);
  const uint32_t z = rem3 - y * (uint32_t)g.Lz;
  f.up[0] = f.dn[0] = 0;
  f.up[1] = ((int)x == g.Lx - 1) ? -(g.Lx - 1) * g.plane : g.plane;
  f.dn[1] = (x == 0) ? (g.Lx - 1) * g.plane : -g.plane;
  f.up[2] = ((int)y == g.Ly - 1) ? -(g.Ly - 1) * g.Lz : g.Lz;
  f.dn[2] = (y == 0) ? (g.Ly - 1) * g.Lz : -g.Lz;
  f.up[3] = ((int)z == g.Lz - 1) ? -(g.Lz - 1) : 1;
  f.dn[3] = (z == 0) ? (g.Lz - 1) : -1;
  return f;
}

// the mapping of k_md_su2_4d_twz: q = b / 8, mu = wave, o3 = (part nq + q) 64 + lane
__host__ __device__ __forceinline__ TwSite tw_site_w(const FastGeo &g, uint32_t b, uint32_t tid) {
  TwSite f;
  const uint32_t xcd = b & 7u, q = b >> 3;
  f.mu = (int)(tid >> 6);
  const uint32_t lc = (uint32_t)g.tw_lc;
  const uint32_t chunk = xcd & ((1u << lc) - 1u), part = xcd >> lc;
  const uint32_t nq = (uint32_t)g.nb64 >> (3u - lc);
  const uint32_t o3 = (part * nq + q) * 64u + (tid & 63u);
  f.o3 = (int)o3;
  f.t0 = (int)chunk * g.Tc;
  const uint32_t x = fast_div(o3, g.m_plane);
  const uint32_t rem3 = o3 - x * (uint32_t)g.plane;
  const uint32_t y = fast_div(rem3, g.m_Lz);
  const uint32_t z = rem3 - y * (uint32_t)g.Lz;
  f.up[0] = f.dn[0] = 0;
  f.up[1] = ((int)x == g.Lx - 1) ? -(g.Lx - 1) * g.plane : g.plane;
  f.dn[1] = (x == 0) ? (g.Lx - 1) * g.plane : -g.plane;
  f.up[2] = ((int)y == g.Ly - 1) ? -(g.Ly - 1) * g.Lz : g.Lz;
  f.dn[2] = (y == 0) ? (g.Ly - 1) * g.Lz : -g.Lz;
  f.up[3] = ((int)z == g.Lz - 1) ? -(g.Lz - 1) : 1;
  f.dn[3] = (z == 0) ? (g.Lz - 1) : -1;
  return f;
}
inline uint32_t tw_blocks_w(const FastGeo &g) { return (8u * (uint32_t)g.nb64) >> (3 - g.tw_lc); }

// periodic t-neighbour offsets of slice t
__host__ __device__ __forceinline__ void tw_t_offsets(const FastGeo &g, int t, int &up0,
                                                      int &dn0) {
  up0 = (t == g.Lt - 1) ? -(g.Lt - 1) * g.vol3 : g.vol3;
  dn0 = (t == 0) ? (g.Lt - 1) * g.vol3 : -g.vol3;
}

// blocks of a t-walk launch
inline uint32_t tw_blocks(const FastGeo &g) { return (32u * (uint32_t)g.nb3) >> (3 - g.tw_lc); }

// W: write the new link (false: the momentum update alone); ACC: add each
// slice's squared old and new momenta and ReTr(U K) to acc (the step's
// kinetic-energy and action partial sums); GEN: the old momenta are not
// loaded but drawn, link (s, mu) from the Philox counter of
// k_momenta_philox (s 4 + mu, 0, traj), so they equal that kernel's bit for bit
template <typename T, int MU, int ZS, int NT, class UI, class UO, bool W = true, bool ACC = false,
          bool GEN = false>
__device__ __forceinline__ void md_su2_tw_body(UI Uin, UO Uout, T *__restrict__ P,
                                               const FastGeo &g, const FusedCoef &c,
                                               const TwSite &f, double (&acc)[3]) {
  static_assert(ZS >= 0 && ZS <= 2, "t-walk: ZS 0, 1 or 2");
  const int V = g.V;
  int up[4] = {0, f.up[1], f.up[2], f.up[3]}, dn[4] = {0, f.dn[1], f.dn[2], f.dn[3]};
  const UI Um = fld_add(Uin, (size_t)MU * V);
  const UI Ut = Uin;  // direction 0 (t)
  int zup = 0, zdn = 0;
  bool vup = true, vdn = true;  // ZS = 2: the neighbour lane is in the wave
  if (ZS == 1) {
    const int lane = threadIdx.x & 63, zl = lane & (g.Lz - 1);
    zup = (zl == g.Lz - 1) ? lane - (g.Lz - 1) : lane + 1;
    zdn = (zl == 0) ? lane + (g.Lz - 1) : lane - 1;
  } else if (ZS == 2) {
    const int lane = threadIdx.x & 63;
    zup = lane + up[3];
    zdn = lane + dn[3];
    vup = (unsigned)zup < 64u;
    vdn = (unsigned)zdn < 64u;
    if (!vup) zup = lane;
    if (!vdn) zdn = lane;
  }
  int s = f.t0 * g.vol3 + f.o3;
  tw_t_offsets(g, f.t0, up[0], dn[0]);
  // carried links
  quat<T> u;            // own link U_mu(s)
  quat<T> D0, E0, F0;   // mu != t, plane t: U_t(s+mu-t), U_mu(s-t), U_t(s-t)
  quat<T> Cc[3], Fc[3]; // mu = t, plane nu = k + 1: U_nu(s), U_nu(s-nu)
  if constexpr (MU != 0) {
    u = ldq(Um, (uint32_t)s);
    E0 = ldq(Um, (uint32_t)(s + dn[0]));
    F0 = ldq(Ut, (uint32_t)(s + dn[0]));
    if (ZS && MU == 3) {
      if (ZS == 2 && !vup) D0 = ldq(Ut, (uint32_t)(s + up[3] + dn[0]));
      const quat<T> d = qshfl(F0, zup);  // U_t(s+z-t)
      if (ZS == 1 || vup) D0 = d;
    } else {
      D0 = ldq(Ut, (uint32_t)(s + up[MU] + dn[0]));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int nu = k + 1;
      const UI Un = fld_add(Uin, (size_t)nu * V);
      Cc[k] = ldq(Un, (uint32_t)s);
      if (ZS && nu == 3) {
        if (ZS == 2 && !vdn) Fc[k] = ldq(Un, (uint32_t)(s + dn[3]));
        const quat<T> f = qshfl(Cc[k], zdn);  // U_z(s-z)
        if (ZS == 1 || vdn) Fc[k] = f;
      } else {
        Fc[k] = ldq(Un, (uint32_t)(s + dn[nu]));
      }
    }
  }
  for (int kt = 0; kt < g.Tc; ++kt) {
    if (kt) {
      s += g.vol3;
      tw_t_offsets(g, f.t0 + kt, up[0], dn[0]);
    }
    quat<T> A[3], B[3], C[3], D[3], E[3], F[3];
    if constexpr (MU != 0) {
      // plane t (k = 0)
      B[0] = ldq(Um, (uint32_t)(s + up[0]));  // U_mu(s+t): the next slice's own link
      C[0] = ldq(Ut, (uint32_t)s);            // U_t(s)
      if (!(ZS && MU == 3)) A[0] = ldq(Ut, (uint32_t)(s + up[MU]));  // U_t(s+mu)
      else if (ZS == 2 && !vup) A[0] = ldq(Ut, (uint32_t)(s + up[3]));
      D[0] = D0;
      E[0] = E0;
      F[0] = F0;
      // planes nu != t, mu: the loads of md_su2_core
#pragma unroll
      for (int k = 1; k < 3; ++k) {
        const int nu = k + (k >= MU ? 1 : 0);
        const UI Un = fld_add(Uin, (size_t)nu * V);
        if (ZS && nu == 3) {
          A[k] = ldq(Un, (uint32_t)(s + up[MU]));  // U_z(s+mu)
          C[k] = ldq(Un, (uint32_t)s);             // U_z(s)
          if (ZS == 2) {
            // neighbours in another wave, loaded by the lanes that need them
            if (!vup) B[k] = ldq(Um, (uint32_t)(s + up[3]));
            if (!vdn) {
              E[k] = ldq(Um, (uint32_t)(s + dn[3]));
              D[k] = ldq(Un, (uint32_t)(s + up[MU] + dn[3]));
              F[k] = ldq(Un, (uint32_t)(s + dn[3]));
            }
          }
        } else if (ZS && MU == 3) {
          B[k] = ldq(Um, (uint32_t)(s + up[nu]));  // U_z(s+nu)
          C[k] = ldq(Un, (uint32_t)s);             // U_nu(s)
          E[k] = ldq(Um, (uint32_t)(s + dn[nu]));  // U_z(s-nu)
          F[k] = ldq(Un, (uint32_t)(s + dn[nu]));  // U_nu(s-nu)
          if (ZS == 2 && !vup) {
            A[k] = ldq(Un, (uint32_t)(s + up[3]));
            D[k] = ldq(Un, (uint32_t)(s + up[3] + dn[nu]));
          }
        } else {
          A[k] = ldq(Un, (uint32_t)(s + up[MU]));
          B[k] = ldq(Um, (uint32_t)(s + up[nu]));
          C[k] = ldq(Un, (uint32_t)s);
          D[k] = ldq(Un, (uint32_t)(s + up[MU] + dn[nu]));
          E[k] = ldq(Um, (uint32_t)(s + dn[nu]));
          F[k] = ldq(Un, (uint32_t)(s + dn[nu]));
        }
      }
      // every lane takes part in each exchange; with ZS = 2 a lane keeps its
      // loaded value where its neighbour is in another wave
      if (ZS && MU == 3) {
        const quat<T> a = qshfl(C[0], zup);  // U_t(s+z)
        if (ZS == 1 || vup) A[0] = a;
      }
      if (ZS) {
#pragma unroll
        for (int k = 1; k < 3; ++k) {
          const int nu = k + (k >= MU ? 1 : 0);
          if (nu == 3) {
            const quat<T> b = qshfl(u, zup);     // U_mu(s+z)
            const quat<T> e = qshfl(u, zdn);     // U_mu(s-z)
            const quat<T> d = qshfl(A[k], zdn);  // U_z(s+mu-z)
            const quat<T> f = qshfl(C[k], zdn);  // U_z(s-z)
            if (ZS == 1 || vup) B[k] = b;
            if (ZS == 1 || vdn) {
              E[k] = e;
              D[k] = d;
              F[k] = f;
            }
          } else if (MU == 3) {
            const quat<T> a = qshfl(C[k], zup);  // U_nu(s+z)
            const quat<T> d = qshfl(F[k], zup);  // U_nu(s+z-nu)
            if (ZS == 1 || vup) {
              A[k] = a;
              D[k] = d;
            }
          }
        }
      }
    } else {
      u = ldq(Um, (uint32_t)s);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int nu = k + 1;
        const UI Un = fld_add(Uin, (size_t)nu * V);
        C[k] = Cc[k];  // U_nu(s)
        F[k] = Fc[k];  // U_nu(s-nu)
        A[k] = ldq(Un, (uint32_t)(s + up[0]));  // U_nu(s+t)
        if (!(ZS && nu == 3)) {
          B[k] = ldq(Um, (uint32_t)(s + up[nu]));          // U_t(s+nu)
          D[k] = ldq(Un, (uint32_t)(s + up[0] + dn[nu]));  // U_nu(s+t-nu)
          E[k] = ldq(Um, (uint32_t)(s + dn[nu]));          // U_t(s-nu)
        } else if (ZS == 2) {
          if (!vup) B[k] = ldq(Um, (uint32_t)(s + up[3]));
          if (!vdn) {
            E[k] = ldq(Um, (uint32_t)(s + dn[3]));
            D[k] = ldq(Un, (uint32_t)(s + up[0] + dn[3]));
          }
        }
      }
      if (ZS) {
        const quat<T> b = qshfl(u, zup);     // U_t(s+z)
        const quat<T> e = qshfl(u, zdn);     // U_t(s-z)
        const quat<T> d = qshfl(A[2], zdn);  // U_z(s+t-z)
        if (ZS == 1 || vup) B[2] = b;
        if (ZS == 1 || vdn) {
          E[2] = e;
          D[2] = d;
        }
      }
    }
    // -0.0 is the identity of FP64 addition: the first staple's addition folds away
    double K[4] = {-0.0, -0.0, -0.0, -0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
      staple_pair<T, false>(K, 1.0, A[k], B[k], C[k], D[k], E[k], F[k]);
    T *__restrict__ P0 = P + (size_t)(MU * 3 + 0) * V;
    T *__restrict__ P1 = P + (size_t)(MU * 3 + 1) * V;
    T *__restrict__ P2 = P + (size_t)(MU * 3 + 2) * V;
    T pi0, pi1, pi2;
    if constexpr (GEN) {
      T z[4];
      philox_normals4<T>((uint32_t)(s * 4 + MU), (uint32_t)(c.gtraj & 0xffffffffu),
                         (uint32_t)(c.gtraj >> 32), c.gseed, z);
      pi0 = z[0];
      pi1 = z[1];
      pi2 = z[2];
    } else {
      pi0 = ldm<NT == 2>(P0, (uint32_t)s);
      pi1 = ldm<NT == 2>(P1, (uint32_t)s);
      pi2 = ldm<NT == 2>(P2, (uint32_t)s);
    }
    T pn[3];
    const double mar = md_su2_update<T, W, NT, UO>(u, K, pi0, pi1, pi2, P0, P1, P2,
                                                   fld_add(Uout, (size_t)MU * V), (uint32_t)s, c,
                                                   pn);
    if (ACC) {
      acc[0] += sq3(pi0, pi1, pi2);
      acc[1] += 2.0 * mar;
      acc[2] += sq3(pn[0], pn[1], pn[2]);
    }
    // what the next slice's link shares with this one
    if constexpr (MU != 0) {
      E0 = u;
      u = B[0];
      D0 = A[0];
      F0 = C[0];
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        Cc[k] = A[k];
        Fc[k] = D[k];
      }
    }
  }
}

// ZS: 1 the +-z lane exchange (64 % Lz == 0), 0 none; NT as k_md_su2_4d
template <typename T, int ZS, int NT, class UI = const quat<T> *, class UO = quat<T> *>
__global__ __launch_bounds__(NTF) void k_md_su2_4d_tw(UI Uin, UO Uout, T *__restrict__ P,
                                                    FastGeo g, FusedCoef c) {
  const TwSite f = tw_site(g, blockIdx.x, threadIdx.x);
  double acc[3];
  switch (__builtin_amdgcn_readfirstlane(f.mu)) {
    case 0: md_su2_tw_body<T, 0, ZS, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    case 1: md_su2_tw_body<T, 1, ZS, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    case 2: md_su2_tw_body<T, 2, ZS, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    default: md_su2_tw_body<T, 3, ZS, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
  }
}

// the t-walk with the general +-z exchange, one direction per wave.  FP32:
// at least 4 waves per SIMD (128 instead of 132 VGPRs, no spills; 48^4 383.7
// against 395.4 us, 40^4 187.8 against 194.6 us, 56^4 725 against 743 us)
template <typename T, int NT, class UI = const quat<T> *, class UO = quat<T> *>
__global__ __launch_bounds__(NTF) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
void k_md_su2_4d_twz(UI Uin, UO Uout, T *__restrict__ P, FastGeo g, FusedCoef c) {
  const TwSite f = tw_site_w(g, blockIdx.x, threadIdx.x);
  double acc[3];
  switch (__builtin_amdgcn_readfirstlane(f.mu)) {
    case 0: md_su2_tw_body<T, 0, 2, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    case 1: md_su2_tw_body<T, 1, 2, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    case 2: md_su2_tw_body<T, 2, 2, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    default: md_su2_tw_body<T, 3, 2, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
  }
}

// the t-walk with the whole-row +-z exchange and one direction per wave (the
// mapping of k_md_su2_4d_twz: a block holds 64 sites, its four waves their
// four directions, so the links the directions share are requested by waves
// of one CU); the FP64 step where 64 % Lz == 0
template <typename T, int NT, class UI = const quat<T> *, class UO = quat<T> *>
__global__ __launch_bounds__(NTF) void k_md_su2_4d_tww(UI Uin, UO Uout, T *__restrict__ P,
                                                     FastGeo g, FusedCoef c) {
  const TwSite f = tw_site_w(g, blockIdx.x, threadIdx.x);
  double acc[3];
  switch (__builtin_amdgcn_readfirstlane(f.mu)) {
    case 0: md_su2_tw_body<T, 0, 1, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    case 1: md_su2_tw_body<T, 1, 1, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    case 2: md_su2_tw_body<T, 2, 1, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
    default: md_su2_tw_body<T, 3, 1, NT, UI, UO>(Uin, Uout, P, g, c, f, acc); break;
  }
}

// The block's partial sums of a t-walk with ACC: one block does the work of
// Tc blocks of the one-launch step (Tc slices of NTF links), so its sums go
// to entry Tc b of each row and entries Tc b + 1 .. Tc b + Tc - 1 are written
// 0: the partial count, 4 V / NTF, is the one-launch step's.
__device__ __forceinline__ void tw_partials(double (&acc)[3], double *lds, int Tc,
                                            double *__restrict__ partials, int pstride) {
  block_sum3<NTF>(acc[0], acc[1], acc[2], lds);
  const size_t e0 = (size_t)Tc * blockIdx.x;
  for (int i = (int)threadIdx.x; i < Tc; i += NTF) {
    partials[0 * (size_t)pstride + e0 + i] = i == 0 ? acc[0] : 0.0;
    partials[1 * (size_t)pstride + e0 + i] = i == 0 ? acc[1] : 0.0;
    partials[2 * (size_t)pstride + e0 + i] = i == 0 ? acc[2] : 0.0;
  }
}

// The first and the last step of a trajectory as t-walks: the step with the
// kinetic-energy and action partial sums (ACCUM), and with W = false the
// momentum update alone (the last step); GEN: the first step draws the
// trajectory's momenta itself.
// FP32: at least 4 waves per SIMD, as k_md_su2_4d_twz
template <typename T, bool W, int ZS, int NT, class UI, class UO, bool GEN = false>
__global__ __launch_bounds__(NTF) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
void k_md_su2_4d_tw_acc(UI Uin, UO Uout, T *__restrict__ P, FastGeo g, FusedCoef c,
                        double *__restrict__ partials, int pstride) {
  __shared__ double lds[3 * NTF / 64];
  const TwSite f = tw_site(g, blockIdx.x, threadIdx.x);
  double acc[3] = {0.0, 0.0, 0.0};
  switch (__builtin_amdgcn_readfirstlane(f.mu)) {
    case 0: md_su2_tw_body<T, 0, ZS, NT, UI, UO, W, true, GEN>(Uin, Uout, P, g, c, f, acc); break;
    case 1: md_su2_tw_body<T, 1, ZS, NT, UI, UO, W, true, GEN>(Uin, Uout, P, g, c, f, acc); break;
    case 2: md_su2_tw_body<T, 2, ZS, NT, UI, UO, W, true, GEN>(Uin, Uout, P, g, c, f, acc); break;
    default: md_su2_tw_body<T, 3, ZS, NT, UI, UO, W, true, GEN>(Uin, Uout, P, g, c, f, acc); break;
  }
  tw_partials(acc, lds, g.Tc, partials, pstride);
}
// (FP32 with W: 3 waves per SIMD; at 4 its 143 VGPRs would spill)
template <typename T, bool W, int NT, class UI, class UO, bool GEN = false>
__global__ __launch_bounds__(NTF) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && !W ? 4 : 1)))
void k_md_su2_4d_twz_acc(UI Uin, UO Uout, T *__restrict__ P, FastGeo g, FusedCoef c,
                         double *__restrict__ partials, int pstride) {
  __shared__ double lds[3 * NTF / 64];
  const TwSite f = tw_site_w(g, blockIdx.x, threadIdx.x);
  double acc[3] = {0.0, 0.0, 0.0};
  switch (__builtin_amdgcn_readfirstlane(f.mu)) {
    case 0: md_su2_tw_body<T, 0, 2, NT, UI, UO, W, true, GEN>(Uin, Uout, P, g, c, f, acc); break;
    case 1: md_su2_tw_body<T, 1, 2, NT, UI, UO, W, true, GEN>(Uin, Uout, P, g, c, f, acc); break;
    case 2: md_su2_tw_body<T, 2, 2, NT, UI, UO, W, true, GEN>(Uin, Uout, P, g, c, f, acc); break;
    default: md_su2_tw_body<T, 3, 2, NT, UI, UO, W, true, GEN>(Uin, Uout, P, g, c, f, acc); break;
  }
  tw_partials(acc, lds, g.Tc, partials, pstride);
}

// Plaquette sum (retr_sum_Wplaquettes, flat-gauge_energy.hpp:31-94) walking t:
// the terms of k_plaq_su2_4d (su2_md_fast.hpp), a site's six planes in the
// same order and arithmetic, with fewer link loads.  A thread keeps one
// 3-volume site (o3 = q NTF + tid, the chunk of XCD b % 8) and walks the
// chunk's Tc slices upwards: U_nu(s+t) (nu = x, y, z) of slice t are the next
// slice's own links and stay in registers, and with ZS = 1 (64 % Lz == 0)
// U_mu(s+z) (mu = t, x, y) comes from lane l + 1's own links — 10 link loads
// per site instead of 16 (13 with ZS = 0).  Launch: 8 nb3 blocks of NTF
// threads, one partial per block (the thread sums its slices' planes, then
// the block sum).  RESTORE as k_plaq_su2_4d: every link reunitarised on load,
// each site's own restored links written to Uout (record layout) once, by the
// thread that owns the site.
struct PlaqTwSite {
  int o3, t0;
  int upx, upy, upz;  // +x, +y, +z neighbour offsets of o3
};
// block b, thread tid of the plaquette t-walk: site o3 = (b / 8) NTF + tid of
// the chunk of XCD b % 8
__host__ __device__ __forceinline__ PlaqTwSite plaq_tw_site(const FastGeo &g, uint32_t b,
                                                            uint32_t tid) {
  PlaqTwSite f;
  const uint32_t xcd = b & 7u, q = b >> 3;
  const uint32_t o3 = q * (uint32_t)NTF + tid;
  f.o3 = (int)o3;
  f.t0 = (int)xcd * g.Tc;
  const uint32_t x = fast_div(o3, g.m_plane);
  const uint32_t rem3 = o3 - x * (uint32_t)g.plane;
  const uint32_t y = fast_div(rem3, g.m_Lz);
  const uint32_t z = rem3 - y * (uint32_t)g.Lz;
  f.upx = ((int)x == g.Lx - 1) ? -(g.Lx - 1) * g.plane : g.plane;
  f.upy = ((int)y == g.Ly - 1) ? -(g.Ly - 1) * g.Lz : g.Lz;
  f.upz = ((int)z == g.Lz - 1) ? -(g.Lz - 1) : 1;
  return f;
}
// the lane holding the +z neighbour of lane `lane`'s site (64 % Lz == 0: a
// wave holds 64 / Lz whole z rows)
__host__ __device__ __forceinline__ int plaq_tw_zup(const FastGeo &g, int lane) {
  const int zl = lane & (g.Lz - 1);
  return (zl == g.Lz - 1) ? lane - (g.Lz - 1) : lane + 1;
}

template <typename T, bool RESTORE, int ZS, class UI = const quat<T> *>
__global__ __launch_bounds__(NTF) void k_plaq_su2_4d_tw(UI Uin, quat<T> *__restrict__ Uout,
                                                       FastGeo g, double w_t, double w_s,
                                                       int mu_start, int aniso,
                                                       double *__restrict__ partials) {
  static_assert(ZS == 0 || ZS == 1, "plaquette t-walk: ZS 0 or 1");
  __shared__ double lds[3 * NTF / 64];
  const int V = g.V;
  const PlaqTwSite f = plaq_tw_site(g, blockIdx.x, threadIdx.x);
  const int upx = f.upx, upy = f.upy, upz = f.upz;
  const int zup = (ZS == 1) ? plaq_tw_zup(g, (int)(threadIdx.x & 63)) : 0;
  // link (m, i), reunitarised where RESTORE
  auto lr = [&](int m, int i) {
    quat<T> u = ldq(fld_add(Uin, (size_t)m * V), (uint32_t)i);
    if (RESTORE) u = qrestore(u);
    return u;
  };
  // one plane's term, as in k_plaq_su2_4d: 2 (U_mu(s) a) . (U_nu(s) b)
  auto term = [&](const quat<T> &um, const quat<T> &a, const quat<T> &un, const quat<T> &b,
                  int mu) {
    const quat<double> X = qmul(to_d(um), to_d(a));
    const quat<double> Y = qmul(to_d(un), to_d(b));
    const double v = 2.0 * (X.ar * Y.ar + X.ai * Y.ai + X.br * Y.br + X.bi * Y.bi);
    return aniso ? ((mu == 0) ? w_t : w_s) * v : v;
  };
  int s = f.t0 * g.vol3 + f.o3;
  quat<T> own[4];
#pragma unroll
  for (int m = 1; m < 4; ++m) own[m] = lr(m, s);
  double sum = 0.0;
  for (int tt = 0; tt < g.Tc; ++tt) {
    int up0, dn0;
    tw_t_offsets(g, f.t0 + tt, up0, dn0);
    own[0] = lr(0, s);
    quat<T> nx[3];  // U_nu(s+t), nu = x, y, z: the next slice's own links
#pragma unroll
    for (int k = 0; k < 3; ++k) nx[k] = lr(k + 1, s + up0);
    if (RESTORE) {
#pragma unroll
      for (int m = 0; m < 4; ++m) sto(Uout + (size_t)m * V, (uint32_t)s, own[m]);
    }
    const quat<T> b01 = lr(0, s + upx);  // U_t(s+x)
    const quat<T> b02 = lr(0, s + upy);  // U_t(s+y)
    const quat<T> a12 = lr(2, s + upx);  // U_y(s+x)
    const quat<T> b12 = lr(1, s + upy);  // U_x(s+y)
    const quat<T> a13 = lr(3, s + upx);  // U_z(s+x)
    const quat<T> a23 = lr(3, s + upy);  // U_z(s+y)
    quat<T> b03, b13, b23;               // U_t(s+z), U_x(s+z), U_y(s+z)
    if (ZS == 1) {
      b03 = qshfl(own[0], zup);
      b13 = qshfl(own[1], zup);
      b23 = qshfl(own[2], zup);
    } else {
      b03 = lr(0, s + upz);
      b13 = lr(1, s + upz);
      b23 = lr(2, s + upz);
    }
    if (mu_start <= 0) {
      sum += term(own[0], nx[0], own[1], b01, 0);
      sum += term(own[0], nx[1], own[2], b02, 0);
      sum += term(own[0], nx[2], own[3], b03, 0);
    }
    if (mu_start <= 1) {
      sum += term(own[1], a12, own[2], b12, 1);
      sum += term(own[1], a13, own[3], b13, 1);
    }
    if (mu_start <= 2) sum += term(own[2], a23, own[3], b23, 2);
#pragma unroll
    for (int k = 0; k < 3; ++k) own[k + 1] = nx[k];
    s += up0;
  }
  const double r = block_sum<NTF>(sum, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// blocks (and partials) of a plaquette t-walk launch
inline uint32_t plaq_tw_blocks(const FastGeo &g) { return 8u * (uint32_t)g.nb3; }

}  // namespace su2k
