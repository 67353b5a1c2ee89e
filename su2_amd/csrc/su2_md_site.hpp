// su2_md_site.hpp -- the per-site body of the fused force / momentum / link
// update, shared by the single-lattice kernel (su2_kernels.hip) and the batched
// ensemble kernel (su2_batch.hip).  For every link (s, mu) of site s:
//   K  = sum_{nu != mu} w_{mu nu} [upper + lower staple]   (get_staples.hh:26-80)
//   F  = coef * get_deriv(U_mu(s) K)                       (flat-gaugemonomial.hh:105-131)
//   P <- P - ep * F                                        (update_momenta.hh:27-30)
//   U' = exp(eu * P) U_mu(s) -> Uout                       (update_gauge.hh:11-14)
#pragma once
#include "su2_device.hpp"
#include "su2_launch.hpp"

namespace su2k {

using namespace su2dev;

// FP32 storage emulation (F_R32L / F_R32P): the value an FP32 field would hold
template <typename T> __device__ __forceinline__ T r32(T x) { return T(float(x)); }
template <typename T> __device__ __forceinline__ quat<T> r32(const quat<T> &q) {
  return quat<T>{T(float(q.ar)), T(float(q.ai)), T(float(q.br)), T(float(q.bi))};
}

template <class G, int ND, bool WRITE_U>
__device__ __forceinline__ void md_site(const typename G::link *__restrict__ Uin,
                                        typename G::link *__restrict__ Uout,
                                        typename G::real *__restrict__ P, const Geo &geo,
                                        const FusedCoef &c, int s, double &kin_old, double &act,
                                        double &kin_new) {
  using T = typename G::real;
  using link = typename G::link;
  using acc = typename G::acc;
  constexpr int NM = G::NMOM;
  const int V = geo.V;
  int up[ND], dn[ND];
  neighbour_offsets<ND>(geo, s, up, dn);
  link own[ND];
#pragma unroll
  for (int nu = 0; nu < ND; ++nu) own[nu] = Uin[nu * V + s];
  const T ep = T(c.ep), eu = T(c.eu), coef = T(c.coef);
  const T rn = T(c.round_n);
#pragma unroll
  for (int mu = 0; mu < ND; ++mu) {
    acc K = G::zero(), Kc = G::zero();
#pragma unroll
    for (int nu = 0; nu < ND; ++nu) {
      if (nu == mu) continue;
      const T w = T((mu == 0 || nu == 0) ? c.w_t : c.w_s);
      acc t1 = G::upper(Uin[nu * V + s + up[mu]], Uin[mu * V + s + up[nu]], own[nu]);
      acc t2 = G::lower(Uin[nu * V + s + up[mu] + dn[nu]], Uin[mu * V + s + dn[nu]],
                        Uin[nu * V + s + dn[nu]]);
      if (c.flags & F_ANISO) {
        t1 = G::scale(w, t1);
        t2 = G::scale(w, t2);
      }
      if (c.flags & F_COMPENSATED) {
        G::add_comp(K, Kc, t1);
        G::add_comp(K, Kc, t2);
      } else {
        K = G::add(G::add(K, t1), t2);
      }
    }
    if (c.flags & F_COMPENSATED) K = G::add(K, Kc);
    T d[NM];
    T retr;
    G::project(own[mu], K, d, retr);
    act += double(retr);
    T p[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      p[k] = (c.flags & F_PZERO) ? T(0) : P[(mu * NM + k) * V + s];
      kin_old += double(p[k]) * double(p[k]);
      T f = coef * d[k];
      if (c.flags & F_ROUND) f = ::round(f * rn) / rn;
      p[k] -= ep * f;
      if (c.flags & F_R32P) p[k] = r32(p[k]);
      kin_new += double(p[k]) * double(p[k]);
      P[(mu * NM + k) * V + s] = p[k];
    }
    if (WRITE_U) {
      link u = own[mu];
      if (c.flags & F_ROUND) u = G::round(u, rn);
      link un = G::expmul(p, eu, u);
      if (c.flags & F_RESTORE) un = G::restore(un);
      if (c.flags & F_R32L) un = r32(un);
      Uout[mu * V + s] = un;
    }
  }
}

template <typename T> __device__ __forceinline__ quat<T> qshfl(const quat<T> &q, int src) {
  return quat<T>{__shfl(q.ar, src, 64), __shfl(q.ai, src, 64), __shfl(q.br, src, 64),
                 __shfl(q.bi, src, 64)};
}
template <typename T> __device__ __forceinline__ T qshfl(const T &q, int src) {
  return __shfl(q, src, 64);
}

// The same update for ONE link (s, mu): a thread per link needs a third of the
// registers of a thread per site (all nd links), which doubles the waves in
// flight for this gather-bound stencil.  With ZS (requires 64 % Lz == 0 and
// every lane of the wave active) the +-z neighbours of the wave's sites are
// taken from neighbouring lanes instead of memory: for nu = z the staple
// needs 2 loads + 4 shuffles instead of 6 loads, for mu = z 4 loads + 2.
template <class G, int ND, bool WRITE_U, bool ZS>
__device__ __forceinline__ void md_link(const typename G::link *__restrict__ Uin,
                                        typename G::link *__restrict__ Uout,
                                        typename G::real *__restrict__ P, const Geo &geo,
                                        const FusedCoef &c, int s, int mu, double &kin_old,
                                        double &act, double &kin_new) {
  using T = typename G::real;
  using link = typename G::link;
  using acc = typename G::acc;
  constexpr int NM = G::NMOM;
  constexpr int Z = ND - 1;  // the fastest coordinate direction
  const int V = geo.V;
  int up[ND], dn[ND];
  neighbour_offsets<ND>(geo, s, up, dn);
  int src_up = 0, src_dn = 0;
  if (ZS) {
    const int lane = threadIdx.x & 63, Lz = geo.L[Z];
    const int zl = lane % Lz;
    src_up = (zl == Lz - 1) ? lane - (Lz - 1) : lane + 1;
    src_dn = (zl == 0) ? lane + (Lz - 1) : lane - 1;
  }
  const link u = Uin[mu * V + s];
  acc K = G::zero(), Kc = G::zero();
#pragma unroll
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    link A, B, C, D, E, F;
    if (ZS && nu == Z) {
      A = Uin[nu * V + s + up[mu]];
      C = Uin[nu * V + s];
      B = qshfl(u, src_up);
      D = qshfl(A, src_dn);
      E = qshfl(u, src_dn);
      F = qshfl(C, src_dn);
    } else if (ZS && mu == Z) {
      C = Uin[nu * V + s];
      F = Uin[nu * V + s + dn[nu]];
      A = qshfl(C, src_up);
      D = qshfl(F, src_up);
      B = Uin[mu * V + s + up[nu]];
      E = Uin[mu * V + s + dn[nu]];
    } else {
      A = Uin[nu * V + s + up[mu]];
      B = Uin[mu * V + s + up[nu]];
      C = Uin[nu * V + s];
      D = Uin[nu * V + s + up[mu] + dn[nu]];
      E = Uin[mu * V + s + dn[nu]];
      F = Uin[nu * V + s + dn[nu]];
    }
    acc t1 = G::upper(A, B, C);
    acc t2 = G::lower(D, E, F);
    if (c.flags & F_ANISO) {
      const T w = T((mu == 0 || nu == 0) ? c.w_t : c.w_s);
      t1 = G::scale(w, t1);
      t2 = G::scale(w, t2);
    }
    if (c.flags & F_COMPENSATED) {
      G::add_comp(K, Kc, t1);
      G::add_comp(K, Kc, t2);
    } else {
      K = G::add(G::add(K, t1), t2);
    }
  }
  if (c.flags & F_COMPENSATED) K = G::add(K, Kc);
  T d[NM];
  T retr;
  G::project(u, K, d, retr);
  act += double(retr);
  const T ep = T(c.ep), eu = T(c.eu), coef = T(c.coef), rn = T(c.round_n);
  T p[NM];
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    p[k] = (c.flags & F_PZERO) ? T(0) : P[(mu * NM + k) * V + s];
    kin_old += double(p[k]) * double(p[k]);
    T f = coef * d[k];
    if (c.flags & F_ROUND) f = ::round(f * rn) / rn;
    p[k] -= ep * f;
    if (c.flags & F_R32P) p[k] = r32(p[k]);
    kin_new += double(p[k]) * double(p[k]);
    P[(mu * NM + k) * V + s] = p[k];
  }
  if (WRITE_U) {
    link uu = u;
    if (c.flags & F_ROUND) uu = G::round(uu, rn);
    link un = G::expmul(p, eu, uu);
    if (c.flags & F_RESTORE) un = G::restore(un);
    if (c.flags & F_R32L) un = r32(un);
    Uout[mu * V + s] = un;
  }
}

}  // namespace su2k
