// su2_md_fast.hpp -- the fused force / momentum / link update of the HMC hot
// loop, specialised for SU(2) in d = 4 (get_staples.hh:26-80,
// flat-gaugemonomial.hh:105-131, update_momenta.hh:27-30, update_gauge.hh:11-14).
//
// The generic kernels (su2_md_site.hpp) are limited by VALU issue, not by HBM:
// a thread per link spends ~800 vector instructions on integer division for
// its coordinates, 64-bit address arithmetic, TwoSum-compensated staple sums
// and the library sin/cos.  This body keeps the arithmetic of the step and
// removes the rest:
//   * coordinates come from the launch layout (scalar block decomposition,
//     one multiply-high division per coordinate), neighbour offsets along t
//     and x are wave-uniform;
//   * every load is a per-direction base pointer (scalar) plus a 32-bit
//     element offset;
//   * FP32 staples (two quaternion products each, in FP32) are accumulated in
//     FP64: the sum of the six staples is exact to FP64 rounding, which is a
//     stronger compensation than TwoSum in FP32 at a third of its cost;
//   * the force projection and the momentum update run in FP64 on the FP32
//     operands and round the new momentum once;
//   * FP32: exp(eps P) = (cos a, sin(a)/a eps P) from the Taylor series of cos
//     and sin(x)/x in x = a^2 (Horner; for a^2 <= 0.05 truncated after x^3,
//     remainder below 2e-10; else truncated at x^7, remainder below 7e-11 for a^2 <= 2.5, larger angles by
//     halving and double-angle formulas);
//   * FP32: U' = E U in FP64 and reunitarised in FP64 before the single
//     rounding to FP32 (unbiased; see qrestore<float>).
// FP64 links keep the generic exp (sincos) and the reference's order of
// operations for the staple sum, so the FP64 form stays within 1e-10 of the
// reference per link.
#pragma once
#include "su2_md_site.hpp"

namespace su2k {

constexpr int NTF = 256;  // threads per workgroup of the specialised step

// cos a and sin(a)/a for x = a^2.  The MD step's range x <= 0.05 (|eps P| <=
// 0.22): series truncated after x^3 (remainders x^4/8! and x^4/9! below 2e-10,
// far below the FP32 rounding of the updated link); larger x: cos_sinc.
__device__ __forceinline__ void cos_sinc_step(double x, double &c, double &s) {
  if (x <= 0.05) {
    // c = 1 - x/2 + x^2/24 - x^3/720, s = 1 - x/6 + x^2/120 - x^3/5040, in a
    // factored Horner form whose non-inline constants are all multiplicands
    // (scalar-register operands, no vector copies of FP64 literals)
    double t = fma(x, -1.0 / 30, 1.0);  // 1 - x/30
    t *= 1.0 / 24;                      // 1/24 - x/720
    t = fma(x, t, -0.5);
    c = fma(x, t, 1.0);
    double w = fma(x, -1.0 / 42, 1.0);  // 1 - x/42
    w = fma(x * (-1.0 / 20), w, 1.0);   // 1 - x/20 + x^2/840
    s = fma(x * (-1.0 / 6), w, 1.0);    // 1 - x/6 + x^2/120 - x^3/5040
    return;
  }
  cos_sinc(x, c, s);
}

// Launch layout of the specialised step.  nb = 4 V / NT blocks; the hardware
// dispatches block b to XCD b % 8.  With tblock, XCD k owns the t-slices
// [k Lt/8, (k+1) Lt/8) and walks (3-volume block, t, mu) with mu fastest, so
// the t +- 1 neighbours of a block were loaded by the blocks just before it on
// the same XCD and are still in its L2.  Without tblock each XCD owns one
// contiguous eighth of the (site block, mu) sequence.
struct FastGeo {
  int Lt, Lx, Ly, Lz;
  int plane;      // Ly Lz
  int vol3;       // Lx Ly Lz
  int V;
  int nb3;        // vol3 / NTF
  int Tc;         // Lt / 8 (tblock)
  int tblock;
  int xwalk;      // tblock walk: 3-volume blocks x-fastest (bx blocks per x plane)
  int bx;         // plane / NTF
  uint32_t m_plane, m_Lz, m_Tc, m_nb3, m_Lx;  // multiply-high divisors (fast_div)
  // wmu: a block holds 64 sites and its four waves the four directions mu of
  // them (wave w: mu = w), so the links the directions share are requested by
  // waves of one CU at about the same time
  int wmu;
  int nb64;       // vol3 / 64
  uint32_t m_nb64;
  int tw_lc;      // t-walk: log2 of the t-chunks (3: one per XCD)
};

// floor(n / d) = umulhi(n, m) with m = floor(2^32 / d) + 1, exact whenever
// n * d < 2^32 (m d = 2^32 + e with 0 < e <= d, so the error term n e / 2^32
// stays below 1 / d).  d = 1 has no 32-bit multiplier (2^32 + 1); it is coded
// as m = 0 and divides by passing n through.
__host__ __device__ __forceinline__ uint32_t fast_div(uint32_t n, uint32_t m) {
#ifdef __HIP_DEVICE_COMPILE__
  return m ? __umulhi(n, m) : n;
#else
  return m ? (uint32_t)(((uint64_t)n * m) >> 32) : n;
#endif
}
inline uint32_t fast_div_magic(uint32_t d) {
  return d == 1 ? 0u : (uint32_t)((((uint64_t)1) << 32) / d + 1);
}

// the link (s, mu) and its periodic neighbour offsets handled by thread tid of
// hardware block b in a grid of nb = 4 V / NTF blocks
struct FastSite {
  int mu, s;
  int up[4], dn[4];
};
__host__ __device__ __forceinline__ FastSite fast_site(const FastGeo &g, uint32_t b, uint32_t nb,
                                                       uint32_t tid) {
  FastSite f;
  int t, blk3;
  uint32_t o3;
  if (g.wmu) {
    // block: 64 sites, wave w: direction w; XCD k walks (64-site block, t)
    // with t fastest inside its t-chunk (tblock), or one eighth of the
    // (t, block) sequence
    f.mu = (int)(tid >> 6);
    const uint32_t xcd = b & 7u, local = b >> 3;
    int bl;
    if (g.tblock) {
      const uint32_t q = fast_div(local, g.m_Tc);
      bl = (int)q;
      t = (int)xcd * g.Tc + (int)(local - q * (uint32_t)g.Tc);
    } else {
      const uint32_t per = nb >> 3, rem = nb & 7u;
      const uint32_t lb = xcd * per + (xcd < rem ? xcd : rem) + local;
      const uint32_t q = fast_div(lb, g.m_nb64);
      t = (int)q;
      bl = (int)(lb - q * (uint32_t)g.nb64);
    }
    o3 = (uint32_t)bl * 64u + (tid & 63u);
  } else {
  if (g.tblock) {
    const uint32_t xcd = b & 7u, local = b >> 3;
    f.mu = (int)(local & 3u);
    const uint32_t r = local >> 2;
    const uint32_t q = fast_div(r, g.m_Tc);  // r = q Tc + tt
    blk3 = (int)q;
    if (g.xwalk) {
      // q = w Lx + x: consecutive q are x neighbours (block w of the x plane)
      const uint32_t w = fast_div(q, g.m_Lx);
      blk3 = (int)((q - w * (uint32_t)g.Lx) * (uint32_t)g.bx + w);
    }
    t = (int)xcd * g.Tc + (int)(r - q * (uint32_t)g.Tc);
  } else {
    // each XCD one contiguous eighth of the logical blocks (xcd_block)
    const uint32_t xcd = b & 7u, local = b >> 3, per = nb >> 3, rem = nb & 7u;
    const uint32_t lb = xcd * per + (xcd < rem ? xcd : rem) + local;
    f.mu = (int)(lb & 3u);
    const uint32_t sb = lb >> 2;
    const uint32_t q = fast_div(sb, g.m_nb3);
    t = (int)q;
    blk3 = (int)(sb - q * (uint32_t)g.nb3);
  }
  o3 = (uint32_t)blk3 * NTF + tid;
  }
  const uint32_t x = fast_div(o3, g.m_plane);
  const uint32_t rem3 = o3 - x * (uint32_t)g.plane;
  const uint32_t y = fast_div(rem3, g.m_Lz);
  const uint32_t z = rem3 - y * (uint32_t)g.Lz;
  f.s = t * g.vol3 + (int)o3;
  f.up[0] = (t == g.Lt - 1) ? -(g.Lt - 1) * g.vol3 : g.vol3;
  f.dn[0] = (t == 0) ? (g.Lt - 1) * g.vol3 : -g.vol3;
  f.up[1] = ((int)x == g.Lx - 1) ? -(g.Lx - 1) * g.plane : g.plane;
  f.dn[1] = (x == 0) ? (g.Lx - 1) * g.plane : -g.plane;
  f.up[2] = ((int)y == g.Ly - 1) ? -(g.Ly - 1) * g.Lz : g.Lz;
  f.dn[2] = (y == 0) ? (g.Ly - 1) * g.Lz : -g.Lz;
  f.up[3] = ((int)z == g.Lz - 1) ? -(g.Lz - 1) : 1;
  f.dn[3] = (z == 0) ? (g.Lz - 1) : -1;
  return f;
}

// layout of the specialised step for a d = 4 geometry (extents t, x, y, z), or
// false where it does not apply: the 3-volume a multiple of the block size,
// 32-bit byte offsets within one direction plane, and every multiply-high
// division within its exact range (n d < 2^32)
inline bool make_fast_geo(const int L[4], int V, bool allow_tblock, FastGeo &f,
                          bool allow_xwalk = false, bool wmu = false) {
  f.Lt = L[0];
  f.Lx = L[1];
  f.Ly = L[2];
  f.Lz = L[3];
  f.plane = f.Ly * f.Lz;
  f.vol3 = f.Lx * f.plane;
  f.V = V;
  if (f.vol3 % NTF != 0 || (int64_t)f.Lt * f.vol3 != V) return false;
  f.nb3 = f.vol3 / NTF;
  f.tblock = (allow_tblock && f.Lt % 8 == 0) ? 1 : 0;
  f.Tc = f.tblock ? f.Lt / 8 : 1;
  const uint64_t two32 = 1ull << 32;
  if ((uint64_t)V * sizeof(quat<double>) >= two32) return false;
  if ((uint64_t)f.vol3 * f.plane >= two32 || (uint64_t)f.plane * f.Lz >= two32) return false;
  const uint64_t nsb = (uint64_t)V / NTF;  // site blocks
  if (nsb * f.nb3 >= two32 || nsb * f.Tc >= two32) return false;
  f.m_plane = fast_div_magic(f.plane);
  f.m_Lz = fast_div_magic(f.Lz);
  f.m_Tc = fast_div_magic(f.Tc);
  f.m_nb3 = fast_div_magic(f.nb3);
  f.xwalk = (allow_xwalk && f.tblock && f.plane % NTF == 0) ? 1 : 0;
  f.bx = f.plane / NTF;
  f.m_Lx = fast_div_magic(f.Lx);
  f.wmu = wmu ? 1 : 0;
  f.nb64 = f.vol3 / 64;
  // block counts divided: local < V / 64 / 8 by Tc, lb < V / 64 by nb64
  if (wmu && ((uint64_t)V / 64 * f.nb64 >= two32 || (uint64_t)V / 64 * f.Tc >= two32)) return false;
  f.m_nb64 = fast_div_magic(f.nb64);
  f.tw_lc = 3;
  if (wmu) f.xwalk = 0;
  return true;
}

// element i of an array as scalar base + 32-bit byte offset (global_load with
// an SGPR base: no 64-bit address arithmetic per load)
template <typename E> __device__ __forceinline__ E ldo(const E *__restrict__ base, uint32_t i) {
  return *(const E *)((const char *)base + i * (uint32_t)sizeof(E));
}
template <typename E> __device__ __forceinline__ void sto(E *__restrict__ base, uint32_t i, E v) {
  *(E *)((char *)base + i * (uint32_t)sizeof(E)) = v;
}
// the same store with the non-temporal hint (streamed past the caches): for
// lattices whose fields exceed the Infinity Cache
typedef float su2_f4 __attribute__((ext_vector_type(4)));
typedef double su2_d2 __attribute__((ext_vector_type(2)));
template <typename E> __device__ __forceinline__ void sto_nt(E *__restrict__ base, uint32_t i, E v) {
  __builtin_nontemporal_store(v, (E *)((char *)base + i * (uint32_t)sizeof(E)));
}
__device__ __forceinline__ void sto_nt(quat<float> *__restrict__ base, uint32_t i, quat<float> v) {
  const su2_f4 w = {v.ar, v.ai, v.br, v.bi};
  __builtin_nontemporal_store(w, (su2_f4 *)((char *)base + i * 16u));
}
__device__ __forceinline__ void sto_nt(quat<double> *__restrict__ base, uint32_t i, quat<double> v) {
  const su2_d2 a = {v.ar, v.ai}, b = {v.br, v.bi};
  __builtin_nontemporal_store(a, (su2_d2 *)((char *)base + i * 32u));
  __builtin_nontemporal_store(b, (su2_d2 *)((char *)base + i * 32u + 16u));
}
template <int NT, typename E> __device__ __forceinline__ void sto_h(E *__restrict__ base, uint32_t i, E v) {
  if constexpr (NT != 0) sto_nt(base, i, v);
  else sto(base, i, v);
}
// a momentum component, with the non-temporal hint (NTL) or without
template <bool NTL, typename E> __device__ __forceinline__ E ldm(const E *__restrict__ base, uint32_t i) {
  if constexpr (NTL) return __builtin_nontemporal_load((const E *)((const char *)base + i * (uint32_t)sizeof(E)));
  else return ldo(base, i);
}

// Half-field layout of an FP64 SU(2) link field of n links: the 16-B halves
// (Re a, Im a) of every link first, then the halves (Re b, Im b), each in the
// link order of the record layout (link (s, mu) at mu V + s).  A wave's load
// of 64 consecutive links is then two contiguous 1 KB runs instead of two
// instructions that each touch every other 16 B of 2 KB.  The MD steps of a
// trajectory keep their intermediate links in it (QHalf / QHalfW: the two
// half-field bases, already offset to a direction plane where one is taken);
// the start and final links stay in the record layout.
struct QHalf {
  const su2_d2 *a, *b;
};
struct QHalfW {
  su2_d2 *a, *b;
};
template <typename T> __device__ __forceinline__ const quat<T> *fld_add(const quat<T> *p, size_t k) {
  return p + k;
}
template <typename T> __device__ __forceinline__ quat<T> *fld_add(quat<T> *p, size_t k) {
  return p + k;
}
__device__ __forceinline__ QHalf fld_add(QHalf f, size_t k) { return QHalf{f.a + k, f.b + k}; }
__device__ __forceinline__ QHalfW fld_add(QHalfW f, size_t k) { return QHalfW{f.a + k, f.b + k}; }
template <typename T> __device__ __forceinline__ quat<T> ldq(const quat<T> *__restrict__ p, uint32_t i) {
  return ldo(p, i);
}
__device__ __forceinline__ quat<double> ldq(QHalf f, uint32_t i) {
  const su2_d2 x = ldo(f.a, i), y = ldo(f.b, i);
  return quat<double>{x.x, x.y, y.x, y.y};
}
template <int NT, typename T>
__device__ __forceinline__ void stq(quat<T> *__restrict__ p, uint32_t i, const quat<T> &v) {
  sto_h<NT>(p, i, v);
}
template <int NT>
__device__ __forceinline__ void stq(QHalfW f, uint32_t i, const quat<double> &v) {
  const su2_d2 x = {v.ar, v.ai}, y = {v.br, v.bi};
  if constexpr (NT != 0) {
    __builtin_nontemporal_store(x, (su2_d2 *)((char *)f.a + i * 16u));
    __builtin_nontemporal_store(y, (su2_d2 *)((char *)f.b + i * 16u));
  } else {
    sto(f.a, i, x);
    sto(f.b, i, y);
  }
}

// u v as four FMA chains (16 FMA / MUL), conjugations of the operands passed in
// through qdag and folded into the source negations of the FMAs
template <typename T> __device__ __forceinline__ quat<T> qmul_fma(const quat<T> &u, const quat<T> &v) {
  quat<T> r;
  r.ar = fma(u.ar, v.ar, -fma(u.ai, v.ai, fma(u.br, v.br, u.bi * v.bi)));
  r.ai = fma(u.ar, v.ai, fma(u.ai, v.ar, fma(-u.bi, v.br, u.br * v.bi)));
  r.br = fma(u.ar, v.br, fma(-u.ai, v.bi, fma(u.br, v.ar, u.bi * v.ai)));
  r.bi = fma(u.ar, v.bi, fma(u.ai, v.br, fma(u.bi, v.ar, -(u.br * v.ai))));
  return r;
}

// one staple accumulated into K (FP64)
template <typename T>
__device__ __forceinline__ void acc_d(double (&K)[4], const quat<T> &q) {
  K[0] += double(q.ar);
  K[1] += double(q.ai);
  K[2] += double(q.br);
  K[3] += double(q.bi);
}
// K += w q in one rounding per component: the anisotropy weight of the plane
// (get_staples.hh:58-79) costs nothing over the plain sum (an FMA for an add;
// w = 1 gives the plain sum bit for bit)
template <typename T>
__device__ __forceinline__ void acc_dw(double (&K)[4], double w, const quat<T> &q) {
  K[0] = fma(w, double(q.ar), K[0]);
  K[1] = fma(w, double(q.ai), K[1]);
  K[2] = fma(w, double(q.br), K[2]);
  K[3] = fma(w, double(q.bi), K[3]);
}


// Momentum and link update of one link from its staple sum K (FP64):
// M = U K, P <- P - ep coef 2 Im(M) rounded once, U' = exp(eu P) U.
// Returns Re(M) and the new momentum for the block sums.
template <typename T, bool WRITE_U, int NT = 0, class UO = quat<T> *>
__device__ __forceinline__ double md_su2_update(const quat<T> &u, const double (&K)[4], T pi0,
                                                T pi1, T pi2, T *__restrict__ P0,
                                                T *__restrict__ P1, T *__restrict__ P2,
                                                UO Uo, uint32_t s,
                                                const FusedCoef &c, T (&pn)[3]) {
  const double ua = u.ar, ub = u.ai, uc = u.br, ud = u.bi;
  const double Mai = ua * K[1] + ub * K[0] - ud * K[2] + uc * K[3];
  const double Mbr = ua * K[2] - ub * K[3] + uc * K[0] + ud * K[1];
  const double Mbi = ua * K[3] + ub * K[2] + ud * K[0] - uc * K[1];
  const double f = 2.0 * c.coef * c.ep;
  const T p0 = T(double(pi0) - f * Mbi);
  const T p1 = T(double(pi1) - f * Mbr);
  const T p2 = T(double(pi2) - f * Mai);
  sto_h<NT>(P0, s, p0);
  sto_h<NT>(P1, s, p1);
  sto_h<NT>(P2, s, p2);
  pn[0] = p0;
  pn[1] = p1;
  pn[2] = p2;
  if (WRITE_U) {
    quat<T> un;
    if (sizeof(T) == 4) {
      // exp(eu P) U in FP64 from the stored FP32 momentum, reunitarised in FP64
      const double e0 = c.eu * double(p0), e1 = c.eu * double(p1), e2 = c.eu * double(p2);
      const double xx = e0 * e0 + e1 * e1 + e2 * e2;
      double cs, sn;
      cos_sinc_step(xx, cs, sn);
      // E = (cos, sinc e2, sinc e1, sinc e0) in (Re a, Im a, Re b, Im b)
      const double Ea = cs, Eb = sn * e2, Ec = sn * e1, Ed = sn * e0;
      double na = Ea * ua - Eb * ub - Ec * uc - Ed * ud;
      double nb = Ea * ub + Eb * ua - Ed * uc + Ec * ud;
      double nc = Ea * uc - Eb * ud + Ec * ua + Ed * ub;
      double nd = Ea * ud + Eb * uc + Ed * ua - Ec * ub;
      if (c.flags & F_RESTORE) {
        // 1/sqrt(1 + d) = 1 - d/2 + 3 d^2/8 - ..., remainder below 3e-13 for
        // |d| < 1e-4 (an FP32 link times the series exp has |d| ~ 1e-7)
        const double d = fma(na, na, fma(nb, nb, fma(nc, nc, fma(nd, nd, -1.0))));
        const double r = (fabs(d) < 1e-4) ? 1.0 + d * (-0.5 + 0.375 * d) : 1.0 / sqrt(1.0 + d);
        na *= r;
        nb *= r;
        nc *= r;
        nd *= r;
      }
      un = quat<T>{T(na), T(nb), T(nc), T(nd)};
    } else {
      un = GSU2<T>::expmul(pn, T(c.eu), u);
      if (c.flags & F_RESTORE) un = GSU2<T>::restore(un);
    }
    stq<NT>(Uo, s, un);
  }
  return ua * K[0] - ub * K[1] - uc * K[2] - ud * K[3];
}

template <typename T> __device__ __forceinline__ double sq3(T a, T b, T c) {
  return double(a) * double(a) + double(b) * double(b) + double(c) * double(c);
}

// upper + lower staple of one plane, weighted by w, into K (FP64): the
// reference's U_nu(s+mu) U_mu(s+nu)^+ U_nu(s)^+ and U_nu(s+mu-nu)^+ U_mu(s-nu)^+ U_nu(s-nu)
// (AN = false: w is 1 and unused, the plain sum)
template <typename T, bool AN>
__device__ __forceinline__ void staple_pair(double (&K)[4], double w, const quat<T> &A,
                                            const quat<T> &B, const quat<T> &C, const quat<T> &D,
                                            const quat<T> &E, const quat<T> &F) {
  quat<T> up, lo;
  if constexpr (sizeof(T) == 4) {
    up = qmul_fma(qmul_fma(A, qdag(B)), qdag(C));
    lo = qmul_fma(qmul_fma(qdag(D), qdag(E)), F);
  } else {
    up = GSU2<T>::upper(A, B, C);
    lo = GSU2<T>::lower(D, E, F);
  }
  if (AN) {
    acc_dw(K, w, up);
    acc_dw(K, w, lo);
  } else {
    acc_d(K, up);
    acc_d(K, lo);
  }
}

// The step for link (s, mu).  MU >= 0: mu = MU is a compile-time constant, the
// three planes nu != mu are known at compile time and all 18 staple links are
// loaded before any arithmetic (one memory round trip per thread).  MU < 0:
// mu at run time, one round trip per plane.
// md_su2_core: the loads, staple sum and update of one link in ND = 3 or 4
// dimensions (direction ND - 1 the fastest); returns the input momentum (pi),
// the new one (pn) and Re(U K) (mar) for the sums.  Of the geometry it reads
// g.V and, with ZS, g.Lz (the extent of direction ND - 1).
// NT: 0 ordinary stores, 1 non-temporal stores of the new links and
// momenta, 2 also non-temporal loads of the momenta
template <typename T, int MU, bool WRITE_U, int ZS = 0, bool AN = false, int ND = 4,
          int NT = 0, bool POUT = false, class UI = const quat<T> *, class UO = quat<T> *>
__device__ __forceinline__ void md_su2_core(UI Uin, UO Uout, T *__restrict__ P,
                                            const FastGeo &g, const FusedCoef &c,
                                            const FastSite &fs, T (&pi)[3], T (&pn)[3],
                                            double &mar, T *__restrict__ Pw = nullptr) {
  const int mu = MU >= 0 ? MU : fs.mu;
  const int s = fs.s;
  const int *up = fs.up, *dn = fs.dn;
  const int V = g.V;
  const UI Um = fld_add(Uin, (size_t)mu * V);
  T *__restrict__ P0 = P + (size_t)(mu * 3 + 0) * V;
  T *__restrict__ P1 = P + (size_t)(mu * 3 + 1) * V;
  T *__restrict__ P2 = P + (size_t)(mu * 3 + 2) * V;
  const quat<T> u = ldq(Um, (uint32_t)s);
  // NT == 2: the momenta, read once per step, also load with the
  // non-temporal hint (they do not displace the links from the caches)
  const T pi0 = ldm<NT == 2>(P0, (uint32_t)s), pi1 = ldm<NT == 2>(P1, (uint32_t)s),
          pi2 = ldm<NT == 2>(P2, (uint32_t)s);
  // -0.0 is the identity of FP64 addition (+0.0 is not: +0.0 + -0.0 = +0.0),
  // so the first staple's addition folds away
  double K[4] = {-0.0, -0.0, -0.0, -0.0};
  if constexpr (MU >= 0) {
    quat<T> A[ND - 1], B[ND - 1], C[ND - 1], D[ND - 1], E[ND - 1], F[ND - 1];
    // ZS: the +-z neighbours of the wave's sites sit in the neighbouring lanes,
    // so they are exchanged through ds_bpermute instead of loaded: 4 of the 6
    // loads of the plane nu = z, 2 of 6 in every plane when mu = z.
    //   ZS = 1: the wave holds whole z-rows (64 % Lz == 0), every neighbour is
    //           in the wave;
    //   ZS = 2: any Lz; the wave holds 64 consecutive sites, so the neighbour
    //           s +- z of lane l is lane l + up_z / l + dn_z where that lies in
    //           [0, 64), and the few lanes whose neighbour is in another wave
    //           (a row's wrap, the wave's ends) load it.
    int zup = 0, zdn = 0;
    bool vup = true, vdn = true;
    if (ZS == 1) {
      const int lane = threadIdx.x & 63, zl = lane & (g.Lz - 1);
      zup = (zl == g.Lz - 1) ? lane - (g.Lz - 1) : lane + 1;
      zdn = (zl == 0) ? lane + (g.Lz - 1) : lane - 1;
    } else if (ZS == 2) {
      const int lane = threadIdx.x & 63;
      zup = lane + up[ND - 1];
      zdn = lane + dn[ND - 1];
      vup = (unsigned)zup < 64u;
      vdn = (unsigned)zdn < 64u;
      if (!vup) zup = lane;
      if (!vdn) zdn = lane;
    }
#pragma unroll
    for (int k = 0; k < ND - 1; ++k) {
      const int nu = k + (k >= MU ? 1 : 0);
      const UI Un = fld_add(Uin, (size_t)nu * V);
      if (ZS && nu == ND - 1) {
        A[k] = ldq(Un, (uint32_t)(s + up[MU]));         // U_z(s+mu)
        C[k] = ldq(Un, (uint32_t)s);                    // U_z(s)
      } else if (ZS && MU == ND - 1) {
        B[k] = ldq(Um, (uint32_t)(s + up[nu]));         // U_z(s+nu)
        C[k] = ldq(Un, (uint32_t)s);                    // U_nu(s)
        E[k] = ldq(Um, (uint32_t)(s + dn[nu]));         // U_z(s-nu)
        F[k] = ldq(Un, (uint32_t)(s + dn[nu]));         // U_nu(s-nu)
      } else {
        A[k] = ldq(Un, (uint32_t)(s + up[MU]));           // U_nu(s+mu)
        B[k] = ldq(Um, (uint32_t)(s + up[nu]));           // U_mu(s+nu)
        C[k] = ldq(Un, (uint32_t)s);                      // U_nu(s)
        D[k] = ldq(Un, (uint32_t)(s + up[MU] + dn[nu]));  // U_nu(s+mu-nu)
        E[k] = ldq(Um, (uint32_t)(s + dn[nu]));           // U_mu(s-nu)
        F[k] = ldq(Un, (uint32_t)(s + dn[nu]));           // U_nu(s-nu)
      }
    }
    if (ZS == 2) {
      // neighbours in another wave: loaded by the lanes that need them, in
      // the same memory round trip as the other loads (the exchange below
      // keeps them)
#pragma unroll
      for (int k = 0; k < ND - 1; ++k) {
        const int nu = k + (k >= MU ? 1 : 0);
        const UI Un = fld_add(Uin, (size_t)nu * V);
        if (nu == ND - 1) {
          if (!vup) B[k] = ldq(Um, (uint32_t)(s + up[ND - 1]));
          if (!vdn) {
            E[k] = ldq(Um, (uint32_t)(s + dn[ND - 1]));
            D[k] = ldq(Un, (uint32_t)(s + up[MU] + dn[ND - 1]));
            F[k] = ldq(Un, (uint32_t)(s + dn[ND - 1]));
          }
        } else if (MU == ND - 1) {
          if (!vup) {
            A[k] = ldq(Un, (uint32_t)(s + up[ND - 1]));
            D[k] = ldq(Un, (uint32_t)(s + up[ND - 1] + dn[nu]));
          }
        }
      }
    }
    if (ZS) {
      // every lane takes part in each exchange; with ZS = 2 a lane keeps its
      // loaded value where its neighbour is in another wave
#pragma unroll
      for (int k = 0; k < ND - 1; ++k) {
        const int nu = k + (k >= MU ? 1 : 0);
        if (nu == ND - 1) {
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
        } else if (MU == ND - 1) {
          const quat<T> a = qshfl(C[k], zup);  // U_nu(s+z)
          const quat<T> d = qshfl(F[k], zup);  // U_nu(s+z-nu)
          if (ZS == 1 || vup) {
            A[k] = a;
            D[k] = d;
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < ND - 1; ++k) {
      const int nu = k + (k >= MU ? 1 : 0);
      staple_pair<T, AN>(K, (MU == 0 || nu == 0) ? c.w_t : c.w_s, A[k], B[k], C[k], D[k], E[k],
                         F[k]);
    }
  } else {
#pragma unroll
    for (int nu = 0; nu < ND; ++nu) {
      if (nu == mu) continue;
      const UI Un = fld_add(Uin, (size_t)nu * V);
      staple_pair<T, AN>(K, (mu == 0 || nu == 0) ? c.w_t : c.w_s, ldq(Un, (uint32_t)(s + up[mu])),
                  ldq(Um, (uint32_t)(s + up[nu])),
                  ldq(Un, (uint32_t)s), ldq(Un, (uint32_t)(s + up[mu] + dn[nu])),
                  ldq(Um, (uint32_t)(s + dn[nu])), ldq(Un, (uint32_t)(s + dn[nu])));
    }
  }
  pi[0] = pi0;
  pi[1] = pi1;
  pi[2] = pi2;
  // POUT: the new momentum goes to Pw (same layout), P is only read
  T *__restrict__ Q = POUT ? Pw : P;
  mar = md_su2_update<T, WRITE_U, NT, UO>(u, K, pi0, pi1, pi2, Q + (size_t)(mu * 3 + 0) * V,
                                          Q + (size_t)(mu * 3 + 1) * V, Q + (size_t)(mu * 3 + 2) * V,
                                          fld_add(Uout, (size_t)mu * V), (uint32_t)s, c, pn);
}

template <typename T, int MU, bool WRITE_U, bool ACCUM, int ZS = 0, bool AN = false,
          int NT = 0, bool POUT = false, class UI = const quat<T> *, class UO = quat<T> *>
__device__ __forceinline__ void md_su2_body(UI Uin, UO Uout, T *__restrict__ P,
                                            const FastGeo &g, const FusedCoef &c,
                                            const FastSite &fs, double *__restrict__ partials,
                                            int pstride, double *lds,
                                            T *__restrict__ Pw = nullptr) {
  T pi[3], pn[3];
  double mar;
  md_su2_core<T, MU, WRITE_U, ZS, AN, 4, NT, POUT, UI, UO>(Uin, Uout, P, g, c, fs, pi, pn, mar, Pw);
  if (ACCUM) {
    double r0 = sq3(pi[0], pi[1], pi[2]), r1 = 2.0 * mar, r2 = sq3(pn[0], pn[1], pn[2]);
    block_sum3<NTF>(r0, r1, r2, lds);
    if (threadIdx.x == 0) {
      partials[0 * pstride + blockIdx.x] = r0;
      partials[1 * pstride + blockIdx.x] = r1;
      partials[2 * pstride + blockIdx.x] = r2;
    }
  }
}

// SPEC: 1 = the block's mu selects one of four bodies with mu fixed at compile
// time (mu is uniform per block); 0 = one body with mu at run time
// ZS (with SPEC): +-z neighbours by lane exchange, 1: 64 % Lz == 0, 2: any Lz
// AN: anisotropic weights w_t / w_s on the planes (FMA accumulation)
// UI / UO: the link field types read and written (record-layout pointers, or
// QHalf / QHalfW for the FP64 half-field layout)
template <typename T, bool WRITE_U, bool ACCUM, int SPEC = 1, int ZS = 0, bool AN = false,
          int NT = 0, class UI = const quat<T> *, class UO = quat<T> *>
__global__ __launch_bounds__(NTF) void k_md_su2_4d(UI Uin, UO Uout,
                                                 T *__restrict__ P, FastGeo g, FusedCoef c,
                                                 double *__restrict__ partials, int pstride) {
  __shared__ double lds[3 * NTF / 64];
  const FastSite fs = fast_site(g, blockIdx.x, gridDim.x, threadIdx.x);
  if (SPEC) {
    // mu is uniform per wave (per block, or wave index with wmu): read it as
    // such so that the bodies are selected by a scalar branch
    switch (__builtin_amdgcn_readfirstlane(fs.mu)) {
      case 0: md_su2_body<T, 0, WRITE_U, ACCUM, ZS, AN, NT, false, UI, UO>(Uin, Uout, P, g, c, fs, partials, pstride, lds); break;
      case 1: md_su2_body<T, 1, WRITE_U, ACCUM, ZS, AN, NT, false, UI, UO>(Uin, Uout, P, g, c, fs, partials, pstride, lds); break;
      case 2: md_su2_body<T, 2, WRITE_U, ACCUM, ZS, AN, NT, false, UI, UO>(Uin, Uout, P, g, c, fs, partials, pstride, lds); break;
      default: md_su2_body<T, 3, WRITE_U, ACCUM, ZS, AN, NT, false, UI, UO>(Uin, Uout, P, g, c, fs, partials, pstride, lds); break;
    }
  } else {
    md_su2_body<T, -1, WRITE_U, ACCUM, 0, AN, NT, false, UI, UO>(Uin, Uout, P, g, c, fs, partials, pstride, lds);
  }
}

// The same step reading the momenta from P and writing the new ones to Pw (a
// separate buffer of the same layout): the gradient-flow recomputation keeps
// every stage's Z' without copying it first.
template <typename T, bool WRITE_U, int ZS>
__global__ __launch_bounds__(NTF) void k_md_su2_4d_pout(const quat<T> *__restrict__ Uin,
                                                      quat<T> *__restrict__ Uout,
                                                      T *__restrict__ P, T *__restrict__ Pw,
                                                      FastGeo g, FusedCoef c) {
  const FastSite fs = fast_site(g, blockIdx.x, gridDim.x, threadIdx.x);
  switch (__builtin_amdgcn_readfirstlane(fs.mu)) {
    case 0: md_su2_body<T, 0, WRITE_U, false, ZS, false, false, true>(Uin, Uout, P, g, c, fs, nullptr, 0, nullptr, Pw); break;
    case 1: md_su2_body<T, 1, WRITE_U, false, ZS, false, false, true>(Uin, Uout, P, g, c, fs, nullptr, 0, nullptr, Pw); break;
    case 2: md_su2_body<T, 2, WRITE_U, false, ZS, false, false, true>(Uin, Uout, P, g, c, fs, nullptr, 0, nullptr, Pw); break;
    default: md_su2_body<T, 3, WRITE_U, false, ZS, false, false, true>(Uin, Uout, P, g, c, fs, nullptr, 0, nullptr, Pw); break;
  }
}

// The same step with one thread per SITE updating its four links.  The upper
// staples of links mu and nu in plane (mu, nu) share W = U_nu(s+mu) U_mu(s+nu)^+:
//   upper_mu = W U_nu(s)^+,  upper_nu = (U_mu(s) W)^+,
// so a site needs 42 products (10.5 per link instead of 12) and 52 loads
// (13 per link instead of 19).  Launch: V / NTF blocks, sites from fast_site_s
// (the link layout without mu); the block sums of block b go to partial entry
// 4 b and entries 4 b + 1..3 are written 0, so the partial count (4 V / NTF)
// equals the link shape's.
__host__ __device__ __forceinline__ void fast_site_s(const FastGeo &g, uint32_t b, uint32_t nb,
                                                     uint32_t tid, int &s, int (&up)[4],
                                                     int (&dn)[4]) {
  int t, blk3;
  if (g.tblock) {
    const uint32_t xcd = b & 7u, local = b >> 3;
    const uint32_t q = fast_div(local, g.m_Tc);
    blk3 = (int)q;
    t = (int)xcd * g.Tc + (int)(local - q * (uint32_t)g.Tc);
  } else {
    const uint32_t xcd = b & 7u, local = b >> 3, per = nb >> 3, rem = nb & 7u;
    const uint32_t lb = xcd * per + (xcd < rem ? xcd : rem) + local;
    const uint32_t q = fast_div(lb, g.m_nb3);
    t = (int)q;
    blk3 = (int)(lb - q * (uint32_t)g.nb3);
  }
  const uint32_t o3 = (uint32_t)blk3 * NTF + tid;
  const uint32_t x = fast_div(o3, g.m_plane);
  const uint32_t rem3 = o3 - x * (uint32_t)g.plane;
  const uint32_t y = fast_div(rem3, g.m_Lz);
  const uint32_t z = rem3 - y * (uint32_t)g.Lz;
  s = t * g.vol3 + (int)o3;
  up[0] = (t == g.Lt - 1) ? -(g.Lt - 1) * g.vol3 : g.vol3;
  dn[0] = (t == 0) ? (g.Lt - 1) * g.vol3 : -g.vol3;
  up[1] = ((int)x == g.Lx - 1) ? -(g.Lx - 1) * g.plane : g.plane;
  dn[1] = (x == 0) ? (g.Lx - 1) * g.plane : -g.plane;
  up[2] = ((int)y == g.Ly - 1) ? -(g.Ly - 1) * g.Lz : g.Lz;
  dn[2] = (y == 0) ? (g.Ly - 1) * g.Lz : -g.Lz;
  up[3] = ((int)z == g.Lz - 1) ? -(g.Lz - 1) : 1;
  dn[3] = (z == 0) ? (g.Lz - 1) : -1;
}

template <typename T, bool WRITE_U, bool ACCUM>
__device__ __forceinline__ void md_su2_site_body(const quat<T> *__restrict__ Uin,
                                                 quat<T> *__restrict__ Uout, T *__restrict__ P,
                                                 const FastGeo &g, const FusedCoef &c,
                                                 double *__restrict__ partials, int pstride,
                                                 double *lds) {
  int s, up[4], dn[4];
  fast_site_s(g, blockIdx.x, gridDim.x, threadIdx.x, s, up, dn);
  const int V = g.V;
  quat<T> own[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) own[m] = ldo(Uin + (size_t)m * V, (uint32_t)s);
  double K[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) K[m][0] = K[m][1] = K[m][2] = K[m][3] = -0.0;  // additive identity
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
#pragma unroll
    for (int nu = mu + 1; nu < 4; ++nu) {
      const quat<T> *__restrict__ Um = Uin + (size_t)mu * V;
      const quat<T> *__restrict__ Un = Uin + (size_t)nu * V;
      const quat<T> A = ldo(Un, (uint32_t)(s + up[mu]));            // U_nu(s+mu)
      const quat<T> B = ldo(Um, (uint32_t)(s + up[nu]));            // U_mu(s+nu)
      const quat<T> Dm = ldo(Un, (uint32_t)(s + up[mu] + dn[nu]));  // U_nu(s+mu-nu)
      const quat<T> Em = ldo(Um, (uint32_t)(s + dn[nu]));           // U_mu(s-nu)
      const quat<T> Fm = ldo(Un, (uint32_t)(s + dn[nu]));           // U_nu(s-nu)
      const quat<T> Dn = ldo(Um, (uint32_t)(s + up[nu] + dn[mu]));  // U_mu(s+nu-mu)
      const quat<T> En = ldo(Un, (uint32_t)(s + dn[mu]));           // U_nu(s-mu)
      const quat<T> Fn = ldo(Um, (uint32_t)(s + dn[mu]));           // U_mu(s-mu)
      const quat<T> W = qmul_fma(A, qdag(B));
      acc_d(K[mu], qmul_fma(W, qdag(own[nu])));
      acc_d(K[nu], qdag(qmul_fma(own[mu], W)));
      acc_d(K[mu], qmul_fma(qmul_fma(qdag(Dm), qdag(Em)), Fm));
      acc_d(K[nu], qmul_fma(qmul_fma(qdag(Dn), qdag(En)), Fn));
      // keep the scheduler from hoisting the next plane's loads above this
      // plane's arithmetic (register pressure)
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  double ko = 0.0, act = 0.0, kn = 0.0;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    T *__restrict__ P0 = P + (size_t)(mu * 3 + 0) * V;
    T *__restrict__ P1 = P + (size_t)(mu * 3 + 1) * V;
    T *__restrict__ P2 = P + (size_t)(mu * 3 + 2) * V;
    const T pi0 = ldo(P0, (uint32_t)s), pi1 = ldo(P1, (uint32_t)s), pi2 = ldo(P2, (uint32_t)s);
    T pn[3];
    const double mar = md_su2_update<T, WRITE_U>(own[mu], K[mu], pi0, pi1, pi2, P0, P1, P2,
                                                 Uout + (size_t)mu * V, (uint32_t)s, c, pn);
    if (ACCUM) {
      ko += sq3(pi0, pi1, pi2);
      act += 2.0 * mar;
      kn += sq3(pn[0], pn[1], pn[2]);
    }
  }
  if (ACCUM) {
    double r0 = ko, r1 = act, r2 = kn;
    block_sum3<NTF>(r0, r1, r2, lds);
    if (threadIdx.x < 4) {
      const uint32_t e = 4u * blockIdx.x + threadIdx.x;
      partials[0 * pstride + e] = threadIdx.x == 0 ? r0 : 0.0;
      partials[1 * pstride + e] = threadIdx.x == 0 ? r1 : 0.0;
      partials[2 * pstride + e] = threadIdx.x == 0 ? r2 : 0.0;
    }
  }
}

template <typename T, bool WRITE_U, bool ACCUM>
__global__ __launch_bounds__(NTF) void k_md_su2_4d_site(const quat<T> *__restrict__ Uin,
                                                      quat<T> *__restrict__ Uout,
                                                      T *__restrict__ P, FastGeo g, FusedCoef c,
                                                      double *__restrict__ partials, int pstride) {
  __shared__ double lds[3 * NTF / 64];
  md_su2_site_body<T, WRITE_U, ACCUM>(Uin, Uout, P, g, c, partials, pstride, lds);
}

// Any other geometry (FP32 form): the same step body with the neighbour offsets
// of the lattice's own geometry, for lattices whose 3-volume is not a multiple of NTF
// (no launch-layout coordinates, no t-blocking).  Blocks of NTF links of one
// mu, nb = ceil(V / NTF) blocks per mu in XCD-aware order (the partial-sum
// count 4 nb of the generic thread-per-link kernel); lanes beyond V idle.
// ZS needs whole waves (V % 64 == 0) and 64 % Lz == 0.
template <typename T, bool WRITE_U, bool ACCUM, bool ZS, bool AN>
__global__ __launch_bounds__(NTF) void k_md_su2_geo(const quat<T> *__restrict__ Uin,
                                                  quat<T> *__restrict__ Uout,
                                                  T *__restrict__ P, Geo geo, FusedCoef c,
                                                  double *__restrict__ partials, int pstride) {
  __shared__ double lds[3 * NTF / 64];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int nb = (geo.V + NTF - 1) / NTF;
  const int mu = lb / nb;
  const int s = (lb - mu * nb) * NTF + threadIdx.x;
  T pi[3] = {T(0), T(0), T(0)}, pn[3] = {T(0), T(0), T(0)};
  double mar = 0.0;
  if (s < geo.V) {
    FastGeo g{};
    g.V = geo.V;
    g.Lz = geo.L[3];
    FastSite fs{};
    fs.s = s;
    fs.mu = mu;
    neighbour_offsets<4>(geo, s, fs.up, fs.dn);
    switch (mu) {
      case 0: md_su2_core<T, 0, WRITE_U, ZS, AN>(Uin, Uout, P, g, c, fs, pi, pn, mar); break;
      case 1: md_su2_core<T, 1, WRITE_U, ZS, AN>(Uin, Uout, P, g, c, fs, pi, pn, mar); break;
      case 2: md_su2_core<T, 2, WRITE_U, ZS, AN>(Uin, Uout, P, g, c, fs, pi, pn, mar); break;
      default: md_su2_core<T, 3, WRITE_U, ZS, AN>(Uin, Uout, P, g, c, fs, pi, pn, mar); break;
    }
  }
  if (ACCUM) {
    double r0 = sq3(pi[0], pi[1], pi[2]), r1 = 2.0 * mar, r2 = sq3(pn[0], pn[1], pn[2]);
    block_sum3<NTF>(r0, r1, r2, lds);
    if (threadIdx.x == 0) {
      partials[0 * pstride + blockIdx.x] = r0;
      partials[1 * pstride + blockIdx.x] = r1;
      partials[2 * pstride + blockIdx.x] = r2;
    }
  }
}

// Plaquette sum (retr_sum_Wplaquettes, flat-gauge_energy.hpp:31-94) for SU(2)
// in d = 4, one thread per site on the site layout of fast_site_s (V / NTF
// blocks), FP64 whatever the storage precision:
//   ReTr(U_mu(s) U_nu(s+mu) U_mu(s+nu)^+ U_nu(s)^+) = 2 (U_mu(s) U_nu(s+mu)) . (U_nu(s) U_mu(s+nu))
// (4-vector dot product of two quaternion products).  RESTORE: every link is
// reunitarised on load and the site's own restored links are written to Uout.
// Planes with mu < mu_start are skipped; aniso weights w_t (mu = 0) / w_s.
// (Measured at 32^4 FP32: 19 us for this shape against 43 us for one thread
// per link with one or two planes each.)
// UI: the input field's type (record layout or QHalf); Uout (RESTORE) is in
// the record layout
template <typename T, bool RESTORE, class UI = const quat<T> *>
__global__ __launch_bounds__(NTF) void k_plaq_su2_4d(UI Uin,
                                                   quat<T> *__restrict__ Uout, FastGeo g,
                                                   double w_t, double w_s, int mu_start,
                                                   int aniso, double *__restrict__ partials) {
  __shared__ double lds[3 * NTF / 64];
  int s, up[4], dn[4];
  fast_site_s(g, blockIdx.x, gridDim.x, threadIdx.x, s, up, dn);
  const int V = g.V;
  quat<double> own[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    quat<T> q = ldq(fld_add(Uin, (size_t)m * V), (uint32_t)s);
    if (RESTORE) {
      q = qrestore(q);
      sto(Uout + (size_t)m * V, (uint32_t)s, q);
    }
    own[m] = to_d(q);
  }
  double sum = 0.0;
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    if (mu < mu_start) continue;
#pragma unroll
    for (int nu = mu + 1; nu < 4; ++nu) {
      quat<T> a = ldq(fld_add(Uin, (size_t)nu * V), (uint32_t)(s + up[mu]));  // U_nu(s+mu)
      quat<T> b = ldq(fld_add(Uin, (size_t)mu * V), (uint32_t)(s + up[nu]));  // U_mu(s+nu)
      if (RESTORE) {
        a = qrestore(a);
        b = qrestore(b);
      }
      const quat<double> X = qmul(own[mu], to_d(a));
      const quat<double> Y = qmul(own[nu], to_d(b));
      const double v = 2.0 * (X.ar * Y.ar + X.ai * Y.ai + X.br * Y.br + X.bi * Y.bi);
      sum += aniso ? ((mu == 0) ? w_t : w_s) * v : v;
    }
  }
  const double r = block_sum<NTF>(sum, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

}  // namespace su2k
