// su2_device.hpp -- device-side SU(2) / U(1) algebra and lattice geometry for
// the MI355X (gfx950) hot path.
//
// SU(2) links are quaternion-packed structs in the reference's component order
// (Re a, Im a, Re b, Im b) for [[a,b],[-b*,a*]] (su2.hh:11-70), so host <-> device
// conversion is a pure transpose of the link direction.  U(1) links are phases
// (u1.hh:11-60).  The group traits below bundle what the staple force, the
// momentum update and the link update need from a group, so one kernel body
// serves SU(2) and U(1) in FP64 and FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace su2dev {

template <typename T> struct alignas(4 * sizeof(T)) quat {
  T ar, ai, br, bi;
};
template <typename T> struct cplx {
  T re, im;
};

// ---------------------------------------------------------------- SU(2) ---
// operator* (su2.hh:87-92): a = a1 a2 - b1 conj(b2), b = a1 b2 + b1 conj(a2)
template <typename T> __device__ __forceinline__ quat<T> qmul(const quat<T> &u, const quat<T> &v) {
  quat<T> r;
  r.ar = (u.ar * v.ar - u.ai * v.ai) - (u.br * v.br + u.bi * v.bi);
  r.ai = (u.ar * v.ai + u.ai * v.ar) - (u.bi * v.br - u.br * v.bi);
  r.br = (u.ar * v.br - u.ai * v.bi) + (u.br * v.ar + u.bi * v.ai);
  r.bi = (u.ar * v.bi + u.ai * v.br) + (u.bi * v.ar - u.br * v.ai);
  return r;
}
// dagger (su2.hh:53-55)
template <typename T> __device__ __forceinline__ quat<T> qdag(const quat<T> &u) {
  return quat<T>{u.ar, -u.ai, -u.br, -u.bi};
}
template <typename T> __device__ __forceinline__ quat<T> qadd(const quat<T> &u, const quat<T> &v) {
  return quat<T>{u.ar + v.ar, u.ai + v.ai, u.br + v.br, u.bi + v.bi};
}
template <typename T> __device__ __forceinline__ quat<T> qscale(T f, const quat<T> &u) {
  return quat<T>{f * u.ar, f * u.ai, f * u.br, f * u.bi};
}

__device__ __forceinline__ void dsincos(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ void dsincos(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }

// exp(adjointsu2) (exp_gauge.cc:8-18) as (cos|p|, sin|p| p/|p|) mapped to
// a = cos + i s p_c, b = s p_b + i s p_a.  sin(al)/al is evaluated as a
// series below al = 1e-4 (where it is exact to rounding), so p = 0 gives the
// identity instead of the reference's 0/0.
template <typename T> __device__ __forceinline__ quat<T> qexp(T pa, T pb, T pc) {
  const T a2 = pa * pa + pb * pb + pc * pc;
  const T al = dsqrt(a2);
  T s, c;
  dsincos(al, &s, &c);
  const T sinc = (al > T(1e-4)) ? s / al : T(1) - a2 * T(1.0 / 6.0);
  return quat<T>{c, sinc * pc, sinc * pb, sinc * pa};
}

// restoreSU (su2.hh:62-66): q / |q|
template <typename T> __device__ __forceinline__ quat<T> qrestore(const quat<T> &u) {
  const T r = dsqrt(u.ar * u.ar + u.ai * u.ai + u.br * u.br + u.bi * u.bi);
  return quat<T>{u.ar / r, u.ai / r, u.br / r, u.bi / r};
}
// FP32 links are normalised in FP64 and rounded once per component.  An FP32
// |q| rounds to 1 for every true |q| in (1, 1 + 2^-24), where the FP32 grid
// is twice as coarse as below 1, so q / |q| in FP32 leaves |q|^2 - 1 biased
// by about +3e-8 per link; over the 4 V links of a 32^4 lattice that shifts the
// Wilson action by O(1) between a start configuration and a reunitarised one.
// Rounding the FP64-normalised components is unbiased (mean |q|^2 - 1 ~ 1e-11).
template <> __device__ __forceinline__ quat<float> qrestore(const quat<float> &u) {
  const double a = u.ar, b = u.ai, c = u.br, d = u.bi;
  const double s = 1.0 / sqrt(a * a + b * b + c * c + d * d);
  return quat<float>{float(a * s), float(b * s), float(c * s), float(d * s)};
}

// promotion of stored links to FP64 for the FP64 sums of the FP32 form
__device__ __forceinline__ quat<double> to_d(const quat<float> &q) {
  return quat<double>{double(q.ar), double(q.ai), double(q.br), double(q.bi)};
}
__device__ __forceinline__ quat<double> to_d(const quat<double> &q) { return q; }
__device__ __forceinline__ double to_d(float x) { return double(x); }
__device__ __forceinline__ double to_d(double x) { return x; }

// error-free transformation a + b = s + e (Knuth TwoSum)
template <typename T> __device__ __forceinline__ void two_sum(T a, T b, T &s, T &e) {
  s = a + b;
  const T bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// cos(a) and sin(a)/a in x = a^2: Taylor series in Horner form (remainder
// below 7e-11 for x <= 2.5), larger angles by halving and the double-angle
// formulas cos 2a = 2 cos^2 a - 1, sin(2a)/(2a) = cos a sin(a)/a.  Used where
// the result is rounded to FP32.
__device__ __forceinline__ void cos_sinc(double x, double &c, double &s) {
  if (x <= 0.05) {
    // the MD step's usual range (|eps P| <= 0.22): truncation after x^4,
    // remainders x^5/10! and x^5/11! below 1e-13
    c = 1.0 + x * (-1.0 / 2 + x * (1.0 / 24 + x * (-1.0 / 720 + x * (1.0 / 40320))));
    s = 1.0 + x * (-1.0 / 6 + x * (1.0 / 120 + x * (-1.0 / 5040 + x * (1.0 / 362880))));
    return;
  }
  // at most 520 halvings: every finite x <= 1.8e308 is below 2.5 by then, and
  // an infinite angle (a diverged momentum) ends the loop with a NaN result
  // instead of spinning
  int k = 0;
  while (x > 2.5 && k < 520) {
    x *= 0.25;
    ++k;
  }
  c = 1.0 + x * (-1.0 / 2 + x * (1.0 / 24 + x * (-1.0 / 720 + x * (1.0 / 40320 + x * (-1.0 / 3628800 +
          x * (1.0 / 479001600 + x * (-1.0 / 87178291200.0)))))));
  s = 1.0 + x * (-1.0 / 6 + x * (1.0 / 120 + x * (-1.0 / 5040 + x * (1.0 / 362880 + x * (-1.0 / 39916800 +
          x * (1.0 / 6227020800.0 + x * (-1.0 / 1307674368000.0)))))));
  for (; k > 0; --k) {
    s *= c;
    c = 2.0 * c * c - 1.0;
  }
}

// -------------------------------------------------------- group traits ----
// SU(2): momentum = 3 reals (adjointsu2, adjointfield.hh:12-51); the force
// projection is get_deriv (adjointfield.hh:53-56) = 2 (Im b, Re b, Im a) of U K.
// FP32 links: the staples (two FP32 products each) are summed in FP64 and the
// force is projected in FP64 (exact staple sum; no TwoSum needed), and
// exp(eps P) U is formed in FP64 from the series above and rounded once.
// FP64 links keep the reference's order of operations throughout.
template <typename T> struct GSU2 {
  using real = T;
  using link = quat<T>;
  using accr = typename std::conditional<sizeof(T) == 4, double, T>::type;
  using acc = quat<accr>;
  static constexpr int NMOM = 3;
  static constexpr int NCOMP = 4;
  static constexpr double Nc = 2.0;
  __device__ static acc zero() { return acc{0, 0, 0, 0}; }
  __device__ static acc widen(const quat<T> &q) {
    return acc{accr(q.ar), accr(q.ai), accr(q.br), accr(q.bi)};
  }
  // U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+   (get_staples.hh:38-45)
  __device__ static acc upper(const link &a, const link &b, const link &c) {
    return widen(qmul(qmul(a, qdag(b)), qdag(c)));
  }
  // U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)   (get_staples.hh:46-54)
  __device__ static acc lower(const link &a, const link &b, const link &c) {
    return widen(qmul(qmul(qdag(a), qdag(b)), c));
  }
  __device__ static acc add(const acc &x, const acc &y) { return qadd(x, y); }
  __device__ static acc scale(T f, const acc &x) { return qscale(accr(f), x); }
  // compensated accumulation: s + c holds the running sum (FP32 links: the
  // FP64 sum is already exact to FP64 rounding)
  __device__ static void add_comp(acc &s, acc &c, const acc &x) {
    if (sizeof(T) == 4) {
      s = qadd(s, x);
    } else {
      accr e;
      two_sum(s.ar, x.ar, s.ar, e); c.ar += e;
      two_sum(s.ai, x.ai, s.ai, e); c.ai += e;
      two_sum(s.br, x.br, s.br, e); c.br += e;
      two_sum(s.bi, x.bi, s.bi, e); c.bi += e;
    }
  }
  // M = U K; deriv = 2 Im-quat(M); retr = ReTr(M)
  __device__ static void project(const link &u, const acc &K, T *d, T &retr) {
    const acc M = qmul(widen(u), K);
    d[0] = T(2 * M.bi);
    d[1] = T(2 * M.br);
    d[2] = T(2 * M.ai);
    retr = T(2 * M.ar);
  }
  __device__ static link expmul(const T *p, T eps, const link &u) {
    if (sizeof(T) == 4) {
      const double e0 = double(eps) * double(p[0]), e1 = double(eps) * double(p[1]),
                   e2 = double(eps) * double(p[2]);
      double cs, sn;
      cos_sinc(e0 * e0 + e1 * e1 + e2 * e2, cs, sn);
      const quat<double> n = qmul(quat<double>{cs, sn * e2, sn * e1, sn * e0}, to_d(u));
      return link{T(n.ar), T(n.ai), T(n.br), T(n.bi)};
    }
    return qmul(qexp(eps * p[0], eps * p[1], eps * p[2]), u);
  }
  __device__ static link restore(const link &u) { return qrestore(u); }
  __device__ static link round(const link &u, T dn) {
    // su2::round (su2.hh:29-33), std::round = half away from zero
    return link{::round(u.ar * dn) / dn, ::round(u.ai * dn) / dn, ::round(u.br * dn) / dn,
                ::round(u.bi * dn) / dn};
  }
  // ReTr of the plaquette U1 U2 U3^+ U4^+ (flat-gauge_energy.hpp:46-48)
  __device__ static T plaq(const link &a, const link &b, const link &c, const link &d) {
    const quat<T> p = qmul(qmul(qmul(a, b), qdag(c)), qdag(d));
    return T(2) * p.ar;
  }
};

// U(1): phases; staples accumulate exp(i phase) as a complex number
// (u1.hh:95-119, accum_type<_u1> = Complex); get_deriv = Im (adjointfield.hh:79-81).
template <typename T> struct GU1 {
  using real = T;
  using link = T;
  using acc = cplx<T>;
  static constexpr int NMOM = 1;
  static constexpr int NCOMP = 1;
  static constexpr double Nc = 1.0;
  __device__ static acc zero() { return acc{0, 0}; }
  __device__ static acc expi(T th) {
    T s, c;
    dsincos(th, &s, &c);
    return acc{c, s};
  }
  __device__ static acc upper(const link &a, const link &b, const link &c) {
    return expi((a - b) - c);
  }
  __device__ static acc lower(const link &a, const link &b, const link &c) {
    return expi((-a - b) + c);
  }
  __device__ static acc add(const acc &x, const acc &y) { return acc{x.re + y.re, x.im + y.im}; }
  __device__ static acc scale(T f, const acc &x) { return acc{f * x.re, f * x.im}; }
  __device__ static void add_comp(acc &s, acc &c, const acc &x) {
    T e;
    two_sum(s.re, x.re, s.re, e); c.re += e;
    two_sum(s.im, x.im, s.im, e); c.im += e;
  }
  __device__ static void project(const link &u, const acc &K, T *d, T &retr) {
    const acc e = expi(u);
    d[0] = e.re * K.im + e.im * K.re;
    retr = e.re * K.re - e.im * K.im;
  }
  __device__ static link expmul(const T *p, T eps, const link &u) { return eps * p[0] + u; }
  __device__ static link restore(const link &u) { return u; }
  __device__ static link round(const link &u, T dn) { return (::round(u) * dn) / dn; }
  __device__ static T plaq(const link &a, const link &b, const link &c, const link &d) {
    T s, co;
    dsincos(((a + b) - c) - d, &s, &co);
    return co;
  }
};

// ------------------------------------------------------------ geometry ----
// Sites in the reference's lexicographic order s = ((t*Lx + x)*Ly + y)*Lz + z
// (gaugeconfig.hh:211-222), direction mu = 0..3 is (t, x, y, z).
struct Geo {
  int L[4];    // extent along direction mu
  int str[4];  // site stride along direction mu
  int V;
  int nd;
  // multiply-high divisors of the extents (geo_set_div): floor(n / L[mu]) =
  // umulhi(n, mL[mu]) with mL = floor(2^32 / L) + 1, exact for n L < 2^32;
  // L = 1 is coded as 0 (pass-through).  fdiv = 0: integer division instead.
  uint32_t mL[4];
  int fdiv;
};

// host: fill the divisors; exact for every n <= V when V * max L < 2^32
inline void geo_set_div(Geo &g) {
  uint64_t maxL = 1;
  for (int mu = 0; mu < 4; ++mu) {
    const uint32_t d = (uint32_t)g.L[mu];
    g.mL[mu] = d <= 1 ? 0u : (uint32_t)((((uint64_t)1) << 32) / d + 1);
    if ((uint64_t)d > maxL) maxL = d;
  }
  g.fdiv = ((uint64_t)g.V * maxL < (((uint64_t)1) << 32)) ? 1 : 0;
}

// periodic neighbour offsets of site s: up[mu] = (s+mu) - s, dn[mu] = (s-mu) - s
template <int ND>
__device__ __forceinline__ void neighbour_offsets(const Geo &g, int s, int *up, int *dn) {
  int c[ND];
  if (g.fdiv) {
    // coordinates from the fastest direction down, one multiply-high each
    uint32_t q = (uint32_t)s;
#pragma unroll
    for (int mu = ND - 1; mu > 0; --mu) {
      const uint32_t qn = g.mL[mu] ? __umulhi(q, g.mL[mu]) : q;
      c[mu] = (int)(q - qn * (uint32_t)g.L[mu]);
      q = qn;
    }
    c[0] = (int)q;
  } else {
#pragma unroll
    for (int mu = 0; mu < ND; ++mu) c[mu] = (s / g.str[mu]) % g.L[mu];
  }
#pragma unroll
  for (int mu = 0; mu < ND; ++mu) {
    up[mu] = (c[mu] == g.L[mu] - 1) ? -(g.L[mu] - 1) * g.str[mu] : g.str[mu];
    dn[mu] = (c[mu] == 0) ? (g.L[mu] - 1) * g.str[mu] : -g.str[mu];
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; this
// maps hardware block b to a logical block so that each XCD walks one
// contiguous chunk of the lattice and neighbour rows stay in its own L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int nx = 8;
  const int xcd = b % nx, local = b / nx;
  const int per = nb / nx, rem = nb % nx;
  return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// three wave64 + workgroup sums at once (one LDS round, two barriers); results
// valid in thread 0.  lds: 3 NT / 64 doubles.
template <int NT>
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
    c += __shfl_down(c, o, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    lds[w] = a;
    lds[NT / 64 + w] = b;
    lds[2 * (NT / 64) + w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = b = c = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
      a += lds[i];
      b += lds[NT / 64 + i];
      c += lds[2 * (NT / 64) + i];
    }
  }
}

// wave64 + workgroup sum of a double, result valid in thread 0
template <int NT> __device__ __forceinline__ double block_sum(double v, double *lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) r += lds[i];
  }
  return r;
}

}  // namespace su2dev
