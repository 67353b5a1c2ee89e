// su2_flow.hip -- adjoint (backward) kernels through the Wilson gradient flow
// (Luescher RK3 of flat-gradient_flow.hh:17-50) for SU(2).
//
// Conventions.  A link perturbation is left-trivialised, U -> exp(X) U with X a
// pure quaternion whose vector part is the adjointsu2 triple (a, b, c) (the
// same frame as get_deriv, adjointfield.hh:53-56, and exp, exp_gauge.cc:8-18).
// A cotangent is a triple per link lambda with dO = sum lambda . x.
//
// One RK3 stage is  Z' = a Z + b D(W),  W' = exp(c Z') W,  D(W) = 2 Im(W K)
// (the force without beta/Nc).  Its adjoint:
//   lambda_Z' += c J(cZ')^T lambda_W'      J(A) = sum ad_A^k/(k+1)!, J^T = J(-A)
//   lambda_W   = Ad(exp(cZ'))^T lambda_W'  (= Ad(exp(-cZ')))
//   lambda_W  += b (dD/dW)^T lambda_Z'     (Hessian-vector product, k_flow_vjp_*)
//   lambda_Z   = a lambda_Z'
//
// (dD/dW)^T lambda collects, per link, one term from each of its 2(d-1)
// plaquettes; every plaquette is evaluated once and gives the terms of its four
// links (k_flow_vjp_plaq, then k_flow_vjp_gather or, fused with the pointwise
// adjoint of the next stage to run, k_flow_gather_stage_adj).  For the
// plaquette P = a b c^+ d^+ (a = U_mu(x), b = U_nu(x+mu), c = U_mu(x+nu),
// d = U_nu(x)) sum_m lambda_m . D_m restricted to P equals -2 Re(P Omega) with
// Omega = L_a + a L_b a^+ - d L_c d^+ - L_d (L = pure quaternion of lambda), and
// the gradient of -Re(P Omega) w.r.t. the four links is
//   a: Im(P Om) + Im([a L_b a^+, P])     b: Im(a^+ P Om a)
//   c: -Im(d^+ Om P d)                   d: -Im(Om P) + Im([-d L_c d^+, P])
// The clover energy density E = sum_{x,mu<nu} |Im C_{mu nu}(x)|^2 / (8V)
// (flat-energy_density.hh:57-164 for SU(2)) has the same structure with the
// insertions Gamma = Im C at the four corners of each plaquette:
//   Om = G_x + a G_{x+mu} a^+ + (ab) G_{x+mu+nu} (ab)^+ + d G_{x+nu} d^+.
#include "su2_device.hpp"
#include "su2_launch.hpp"

using namespace su2dev;

namespace su2k {

constexpr int NTF = 256;

template <typename T> struct v3 {
  T x, y, z;  // adjointsu2 (a, b, c)
};

// (a,b,c) <-> pure quaternion in storage order (Re a, Im a, Re b, Im b) = (q0, q3, q2, q1)
template <typename T> __device__ __forceinline__ quat<T> vq(const v3<T> &v) {
  return quat<T>{T(0), v.z, v.y, v.x};
}
template <typename T> __device__ __forceinline__ v3<T> im(const quat<T> &q) {
  return v3<T>{q.bi, q.br, q.ai};
}
template <typename T> __device__ __forceinline__ quat<T> qsub(const quat<T> &u, const quat<T> &v) {
  return quat<T>{u.ar - v.ar, u.ai - v.ai, u.br - v.br, u.bi - v.bi};
}
template <typename T> __device__ __forceinline__ quat<T> comm(const quat<T> &u, const quat<T> &v) {
  return qsub(qmul(u, v), qmul(v, u));
}
template <typename T> __device__ __forceinline__ quat<T> conj_by(const quat<T> &g, const quat<T> &y) {
  return qmul(qmul(g, y), qdag(g));  // g y g^+
}
template <typename T> __device__ __forceinline__ void vadd(v3<T> &a, const v3<T> &b, T s = T(1)) {
  a.x += s * b.x;
  a.y += s * b.y;
  a.z += s * b.z;
}
template <typename T> __device__ __forceinline__ v3<T> ld3(const T *P, int plane, int V, int s) {
  return v3<T>{P[(plane * 3 + 0) * V + s], P[(plane * 3 + 1) * V + s], P[(plane * 3 + 2) * V + s]};
}

// ---------------------------------------------------------------------------
// pointwise part of one stage adjoint (see header) for one link: z = Z', y =
// lambda_W' in; lz (lambda_Z, accumulated) and the returned lambda_W out
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ v3<T> stage_adj_link(const v3<T> &z, const v3<T> &y, v3<T> &lz, T c) {
  const v3<T> a{c * z.x, c * z.y, c * z.z};
  const quat<T> A = vq(a), Y = vq(y);
  // J(-A) y = y - f1 [A, y] + f2 [A, [A, y]],  theta = 2|a| (|ad_A| = 2|a|)
  const T al2 = a.x * a.x + a.y * a.y + a.z * a.z;
  const T th2 = T(4) * al2;
  T f1, f2;
  if (th2 < T(1e-6)) {
    f1 = T(0.5) - th2 * T(1.0 / 24.0);
    f2 = T(1.0 / 6.0) - th2 * T(1.0 / 120.0);
  } else {
    const T th = dsqrt(th2);
    T sn, cs;
    dsincos(th, &sn, &cs);
    f1 = (T(1) - cs) / th2;
    f2 = (th - sn) / (th2 * th);
  }
  const quat<T> k1 = comm(A, Y);
  const quat<T> k2 = comm(A, k1);
  const v3<T> jy = im(quat<T>{T(0), Y.ai - f1 * k1.ai + f2 * k2.ai, Y.br - f1 * k1.br + f2 * k2.br,
                              Y.bi - f1 * k1.bi + f2 * k2.bi});
  vadd(lz, jy, c);
  // Ad(exp(A))^T y = g^+ y g
  const quat<T> g = qexp(a.x, a.y, a.z);
  return im(qmul(qmul(qdag(g), Y), g));
}

template <typename T> __device__ __forceinline__ void st3(T *P, int plane, int V, int s, const v3<T> &v) {
  P[(plane * 3 + 0) * V + s] = v.x;
  P[(plane * 3 + 1) * V + s] = v.y;
  P[(plane * 3 + 2) * V + s] = v.z;
}

// one stage adjoint, one thread per link, in place on lamW / lamZ
template <typename T>
__global__ __launch_bounds__(NTF) void k_flow_stage_adj(const T *__restrict__ Zs, T *__restrict__ lamW,
                                                        T *__restrict__ lamZ, int V, int nd,
                                                        double c, int zero_lamZ) {
  const int i = blockIdx.x * NTF + threadIdx.x;
  if (i >= V * nd) return;
  const int mu = i / V, s = i - mu * V;
  v3<T> lz = zero_lamZ ? v3<T>{0, 0, 0} : ld3(lamZ, mu, V, s);
  const v3<T> yw = stage_adj_link<T>(ld3(Zs, mu, V, s), ld3(lamW, mu, V, s), lz, T(c));
  st3(lamZ, mu, V, s, lz);
  st3(lamW, mu, V, s, yw);
}

// the four role gradients of -Re(P Omega) for the HVP insertion pattern of one
// plaquette (a, b, c, d and their cotangents La..Ld as in the header); P, Omega,
// P Omega and Omega P are shared by the roles
template <typename T> struct roles4 {
  v3<T> ga, gb, gc, gd;
};
template <typename T>
__device__ __forceinline__ roles4<T> hvp_roles(const quat<T> &a, const quat<T> &b, const quat<T> &c,
                                               const quat<T> &d, const v3<T> &La, const v3<T> &Lb,
                                               const v3<T> &Lc, const v3<T> &Ld) {
  const quat<T> P = qmul(qmul(qmul(a, b), qdag(c)), qdag(d));
  const quat<T> Tb = conj_by(a, vq(Lb));
  const quat<T> Tc = qscale(T(-1), conj_by(d, vq(Lc)));
  const quat<T> Om = qsub(qadd(qadd(vq(La), Tb), Tc), vq(Ld));
  const quat<T> PO = qmul(P, Om), OP = qmul(Om, P);
  roles4<T> r;
  r.ga = im(PO);  // a: Im(P Om) + Im([Tb, P])
  vadd(r.ga, im(comm(Tb, P)));
  r.gb = im(qmul(qmul(qdag(a), PO), a));  // b: Im(a^+ P Om a)
  const v3<T> c3 = im(qmul(qmul(qdag(d), OP), d));  // c: -Im(d^+ Om P d)
  r.gc = v3<T>{-c3.x, -c3.y, -c3.z};
  const v3<T> d3 = im(OP);  // d: -Im(Om P) + Im([Tc, P])
  r.gd = v3<T>{-d3.x, -d3.y, -d3.z};
  vadd(r.gd, im(comm(Tc, P)));
  return r;
}

// role-specific gradient of -Re(P Omega_E) for the clover insertions
template <typename T>
__device__ __forceinline__ v3<T> clover_role(int role, const quat<T> &a, const quat<T> &b,
                                             const quat<T> &c, const quat<T> &d, const v3<T> &Gx,
                                             const v3<T> &Gxm, const v3<T> &Gxmn, const v3<T> &Gxn) {
  const quat<T> P = qmul(qmul(qmul(a, b), qdag(c)), qdag(d));
  const quat<T> ab = qmul(a, b);
  const quat<T> Tb = conj_by(a, vq(Gxm));
  const quat<T> Tc = conj_by(ab, vq(Gxmn));
  const quat<T> Td = conj_by(d, vq(Gxn));
  const quat<T> Om = qadd(qadd(qadd(vq(Gx), Tb), Tc), Td);
  switch (role) {
    case 0: {
      v3<T> r = im(qmul(P, Om));
      vadd(r, im(comm(Tb, P)));
      vadd(r, im(comm(Tc, P)));
      return r;
    }
    case 1: {
      const quat<T> aPa = qmul(qmul(qdag(a), P), a);
      v3<T> r = im(qmul(qmul(qmul(qdag(a), P), Om), a));
      vadd(r, im(comm(conj_by(b, vq(Gxmn)), aPa)));
      return r;
    }
    case 2: {
      v3<T> r = im(qmul(qmul(qmul(qdag(d), Om), P), d));
      return v3<T>{-r.x, -r.y, -r.z};
    }
    default: {
      v3<T> r = im(qmul(Om, P));
      r = v3<T>{-r.x, -r.y, -r.z};
      vadd(r, im(comm(Td, P)));
      return r;
    }
  }
}

__device__ __forceinline__ int plane_id(int mu, int nu) {
  // mu < nu: (0,1)=0 (0,2)=1 (0,3)=2 (1,2)=3 (1,3)=4 (2,3)=5
  return mu == 0 ? nu - 1 : (mu == 1 ? nu + 1 : 5);
}

// lamW += bcoef * (dD/dW)^T lamZ in two passes.  Pass 1, one thread per site x:
// each plaquette P_{mu nu}(x), mu < nu, is evaluated once and gives its four
// role gradients; roles a and d belong to the site's own links (x, mu) and
// (x, nu) and are summed in registers (one read-modify-write of lamW per link);
// roles b and c belong to (x+mu, nu) and (x+nu, mu) and go to the scratch rows
// R[(p*2 + r)*3 + k][x] (p = plane_id, r = 0: b, 1: c), which pass 2 gathers.
// Every plaquette is evaluated once instead of once per link (four times).
template <typename T, int ND>
__global__ __launch_bounds__(NTF) void k_flow_vjp_plaq(const quat<T> *__restrict__ U,
                                                       const T *__restrict__ lamZ,
                                                       T *__restrict__ lamW, T *__restrict__ R,
                                                       Geo geo, double bcoef, double w_t,
                                                       double w_s) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = lb * NTF + threadIdx.x;
  const int V = geo.V;
  if (s >= V) return;
  int up[ND], dn[ND];
  neighbour_offsets<ND>(geo, s, up, dn);
  v3<T> acc[ND];
#pragma unroll
  for (int m = 0; m < ND; ++m) acc[m] = v3<T>{0, 0, 0};
#pragma unroll
  for (int mu = 0; mu < ND; ++mu) {
#pragma unroll
    for (int nu = mu + 1; nu < ND; ++nu) {
      const T w = T(mu == 0 ? w_t : w_s);
      const roles4<T> r = hvp_roles<T>(
          U[mu * V + s], U[nu * V + s + up[mu]], U[mu * V + s + up[nu]], U[nu * V + s],
          ld3(lamZ, mu, V, s), ld3(lamZ, nu, V, s + up[mu]), ld3(lamZ, mu, V, s + up[nu]),
          ld3(lamZ, nu, V, s));
      vadd(acc[mu], r.ga, w);
      vadd(acc[nu], r.gd, w);
      const int p = plane_id(mu, nu);
      T *Rb = R + (size_t)(p * 2 + 0) * 3 * V, *Rc = R + (size_t)(p * 2 + 1) * 3 * V;
      Rb[0 * V + s] = w * r.gb.x;
      Rb[1 * V + s] = w * r.gb.y;
      Rb[2 * V + s] = w * r.gb.z;
      Rc[0 * V + s] = w * r.gc.x;
      Rc[1 * V + s] = w * r.gc.y;
      Rc[2 * V + s] = w * r.gc.z;
    }
  }
  const T f = T(2.0 * bcoef);
#pragma unroll
  for (int rho = 0; rho < ND; ++rho) {
    lamW[(rho * 3 + 0) * V + s] += f * acc[rho].x;
    lamW[(rho * 3 + 1) * V + s] += f * acc[rho].y;
    lamW[(rho * 3 + 2) * V + s] += f * acc[rho].z;
  }
}

// pass 2, one thread per site x, for each link (x, rho) and sigma != rho: the
// role c of P_{rho sigma}(x - sigma) (rho < sigma) or the role b of
// P_{sigma rho}(x - sigma) (rho > sigma), added in the fixed sigma order, so the
// result is deterministic
template <typename T, int ND>
__global__ __launch_bounds__(NTF) void k_flow_vjp_gather(const T *__restrict__ R,
                                                         T *__restrict__ lamW, Geo geo,
                                                         double bcoef) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = lb * NTF + threadIdx.x;
  const int V = geo.V;
  if (s >= V) return;
  int up[ND], dn[ND];
  neighbour_offsets<ND>(geo, s, up, dn);
  const T f = T(2.0 * bcoef);
#pragma unroll
  for (int rho = 0; rho < ND; ++rho) {
    v3<T> g{0, 0, 0};
#pragma unroll
    for (int sg = 0; sg < ND; ++sg) {
      if (sg == rho) continue;
      const int p = rho < sg ? plane_id(rho, sg) : plane_id(sg, rho);
      const T *Rr = R + (size_t)(p * 2 + (rho < sg ? 1 : 0)) * 3 * V;
      const int sm = s + dn[sg];
      vadd(g, v3<T>{Rr[sm], Rr[V + sm], Rr[2 * V + sm]});
    }
    lamW[(rho * 3 + 0) * V + s] += f * g.x;
    lamW[(rho * 3 + 1) * V + s] += f * g.y;
    lamW[(rho * 3 + 2) * V + s] += f * g.z;
  }
}

// pass 2 of one stage's Hessian-vector product fused with the pointwise
// adjoint of the stage before it: per site x and link (x, rho), the b/c role
// terms are gathered as in k_flow_vjp_gather (same order, same rounding), which
// completes lambda_W of this stage, and stage_adj_link then runs on it with the
// earlier stage's Z' and coefficient c.  Saves one pass over lamW and lamZ.
template <typename T, int ND>
__global__ __launch_bounds__(NTF) void k_flow_gather_stage_adj(const T *__restrict__ R,
                                                               T *__restrict__ lamW,
                                                               T *__restrict__ lamZ,
                                                               const T *__restrict__ Zs, Geo geo,
                                                               double bcoef, double c,
                                                               int zero_lamZ) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = lb * NTF + threadIdx.x;
  const int V = geo.V;
  if (s >= V) return;
  int up[ND], dn[ND];
  neighbour_offsets<ND>(geo, s, up, dn);
  const T f = T(2.0 * bcoef);
#pragma unroll
  for (int rho = 0; rho < ND; ++rho) {
    v3<T> g{0, 0, 0};
#pragma unroll
    for (int sg = 0; sg < ND; ++sg) {
      if (sg == rho) continue;
      const int p = rho < sg ? plane_id(rho, sg) : plane_id(sg, rho);
      const T *Rr = R + (size_t)(p * 2 + (rho < sg ? 1 : 0)) * 3 * V;
      const int sm = s + dn[sg];
      vadd(g, v3<T>{Rr[sm], Rr[V + sm], Rr[2 * V + sm]});
    }
    v3<T> y = ld3(lamW, rho, V, s);
    y.x += f * g.x;
    y.y += f * g.y;
    y.z += f * g.z;
    v3<T> lz = zero_lamZ ? v3<T>{0, 0, 0} : ld3(lamZ, rho, V, s);
    const v3<T> yw = stage_adj_link<T>(ld3(Zs, rho, V, s), y, lz, T(c));
    st3(lamZ, rho, V, s, lz);
    st3(lamW, rho, V, s, yw);
  }
}

// pass 1 of grad E: Gamma_{mu nu}(x) = Im C_{mu nu}(x) (clover, flat-energy_density.hh:85-119)
template <typename T, int ND>
__global__ __launch_bounds__(NTF) void k_clover_im(const quat<T> *__restrict__ U, T *__restrict__ Gam,
                                                   Geo geo) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = lb * NTF + threadIdx.x;
  const int V = geo.V;
  if (s >= V) return;
  int up[ND], dn[ND];
  neighbour_offsets<ND>(geo, s, up, dn);
  auto L = [&](int site, int d) { return U[d * V + site]; };
#pragma unroll
  for (int mu = 0; mu < ND - 1; ++mu)
#pragma unroll
    for (int nu = mu + 1; nu < ND; ++nu) {
      const int x1 = s + up[mu], x2 = s + up[nu];
      quat<T> leaf = qmul(qmul(qmul(L(s, mu), L(x1, nu)), qdag(L(x2, mu))), qdag(L(s, nu)));
      const int a1 = s + dn[mu] + up[nu], a2 = s + dn[mu];
      leaf = qadd(leaf, qmul(qmul(qmul(L(s, nu), qdag(L(a1, mu))), qdag(L(a2, nu))), L(a2, mu)));
      const int b2 = s + dn[mu] + dn[nu], b3 = s + dn[nu];
      leaf = qadd(leaf, qmul(qmul(qmul(qdag(L(a2, mu)), qdag(L(b2, nu))), L(b2, mu)), L(b3, nu)));
      const int c2 = s + dn[nu] + up[mu];
      leaf = qadd(leaf, qmul(qmul(qmul(qdag(L(b3, nu)), L(b3, mu)), L(c2, nu)), qdag(L(s, mu))));
      const int p = plane_id(mu, nu);
      const v3<T> g = im(leaf);
      Gam[(p * 3 + 0) * V + s] = g.x;
      Gam[(p * 3 + 1) * V + s] = g.y;
      Gam[(p * 3 + 2) * V + s] = g.z;
    }
}

// pass 2 of grad E: lamW = scale * sum over the plaquettes of each link
template <typename T, int ND>
__global__ __launch_bounds__(NTF) void k_clover_grad(const quat<T> *__restrict__ U,
                                                     const T *__restrict__ Gam, T *__restrict__ lamW,
                                                     Geo geo, double scale) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = lb * NTF + threadIdx.x;
  const int V = geo.V;
  if (s >= V) return;
  int up[ND], dn[ND];
  neighbour_offsets<ND>(geo, s, up, dn);
  auto Ul = [&](int d, int site) { return U[d * V + site]; };
  auto Gl = [&](int m, int n, int site) { return ld3(Gam, plane_id(m, n), V, site); };
#pragma unroll
  for (int rho = 0; rho < ND; ++rho) {
    v3<T> g{0, 0, 0};
#pragma unroll
    for (int sg = 0; sg < ND; ++sg) {
      if (sg == rho) continue;
      const int sm = s + dn[sg];
      if (rho < sg) {
        const int m = rho, n = sg;
        // x = s: corners s, s+rho, s+rho+sg, s+sg; link is a
        vadd(g, clover_role<T>(0, Ul(m, s), Ul(n, s + up[m]), Ul(m, s + up[n]), Ul(n, s), Gl(m, n, s),
                               Gl(m, n, s + up[m]), Gl(m, n, s + up[m] + up[n]), Gl(m, n, s + up[n])));
        // x = s - sg: corners sm, sm+rho, s+rho, s; link is c
        vadd(g, clover_role<T>(2, Ul(m, sm), Ul(n, sm + up[m]), Ul(m, s), Ul(n, sm), Gl(m, n, sm),
                               Gl(m, n, sm + up[m]), Gl(m, n, s + up[m]), Gl(m, n, s)));
      } else {
        const int m = sg, n = rho;
        // x = s (mu = sg, nu = rho): corners s, s+sg, s+sg+rho, s+rho; link is d
        vadd(g, clover_role<T>(3, Ul(m, s), Ul(n, s + up[m]), Ul(m, s + up[n]), Ul(n, s), Gl(m, n, s),
                               Gl(m, n, s + up[m]), Gl(m, n, s + up[m] + up[n]), Gl(m, n, s + up[n])));
        // x = s - sg: corners sm, s, s+rho, sm+rho; link is b
        vadd(g, clover_role<T>(1, Ul(m, sm), Ul(n, s), Ul(m, sm + up[n]), Ul(n, sm), Gl(m, n, sm),
                               Gl(m, n, s), Gl(m, n, s + up[n]), Gl(m, n, sm + up[n])));
      }
    }
    const T f = T(scale);
    lamW[(rho * 3 + 0) * V + s] = f * g.x;
    lamW[(rho * 3 + 1) * V + s] = f * g.y;
    lamW[(rho * 3 + 2) * V + s] = f * g.z;
  }
}

// ------------------------------------------------------------- launchers ----
static inline int nblkf(long n) { return (int)((n + NTF - 1) / NTF); }

void launch_flow_stage_adj(int prec, const void *Zs, void *lamW, void *lamZ, int V, int nd, double c,
                           int zero_lamZ, hipStream_t st) {
  dim3 grid(nblkf((long)V * nd));
  if (prec == 0)
    hipLaunchKernelGGL((k_flow_stage_adj<double>), grid, dim3(NTF), 0, st, (const double *)Zs,
                       (double *)lamW, (double *)lamZ, V, nd, c, zero_lamZ);
  else
    hipLaunchKernelGGL((k_flow_stage_adj<float>), grid, dim3(NTF), 0, st, (const float *)Zs,
                       (float *)lamW, (float *)lamZ, V, nd, c, zero_lamZ);
}

template <typename T, int ND>
static void vjp_nd(const void *U, const void *lamZ, void *lamW, void *R, const Geo &g, double b,
                   double w_t, double w_s, bool gather, hipStream_t st) {
  hipLaunchKernelGGL((k_flow_vjp_plaq<T, ND>), dim3(nblkf(g.V)), dim3(NTF), 0, st,
                     (const quat<T> *)U, (const T *)lamZ, (T *)lamW, (T *)R, g, b, w_t, w_s);
  if (gather)
    hipLaunchKernelGGL((k_flow_vjp_gather<T, ND>), dim3(nblkf(g.V)), dim3(NTF), 0, st,
                       (const T *)R, (T *)lamW, g, b);
}
template <typename T>
static void vjp_t(const void *U, const void *lamZ, void *lamW, void *R, const Geo &g, double b,
                  double w_t, double w_s, bool gather, hipStream_t st) {
  switch (g.nd) {
    case 2: vjp_nd<T, 2>(U, lamZ, lamW, R, g, b, w_t, w_s, gather, st); break;
    case 3: vjp_nd<T, 3>(U, lamZ, lamW, R, g, b, w_t, w_s, gather, st); break;
    default: vjp_nd<T, 4>(U, lamZ, lamW, R, g, b, w_t, w_s, gather, st); break;
  }
}
void launch_flow_vjp(int prec, const void *U, const void *lamZ, void *lamW, void *R, const Geo &g,
                     double b, double w_t, double w_s, bool gather, hipStream_t st) {
  if (prec == 0) vjp_t<double>(U, lamZ, lamW, R, g, b, w_t, w_s, gather, st);
  else vjp_t<float>(U, lamZ, lamW, R, g, b, w_t, w_s, gather, st);
}

template <typename T>
static void gather_t(const void *R, void *lamW, const Geo &g, double b, hipStream_t st) {
  const dim3 grid(nblkf(g.V));
  switch (g.nd) {
    case 2: hipLaunchKernelGGL((k_flow_vjp_gather<T, 2>), grid, dim3(NTF), 0, st, (const T *)R, (T *)lamW, g, b); break;
    case 3: hipLaunchKernelGGL((k_flow_vjp_gather<T, 3>), grid, dim3(NTF), 0, st, (const T *)R, (T *)lamW, g, b); break;
    default: hipLaunchKernelGGL((k_flow_vjp_gather<T, 4>), grid, dim3(NTF), 0, st, (const T *)R, (T *)lamW, g, b); break;
  }
}
void launch_flow_gather(int prec, const void *R, void *lamW, const Geo &g, double b, hipStream_t st) {
  if (prec == 0) gather_t<double>(R, lamW, g, b, st);
  else gather_t<float>(R, lamW, g, b, st);
}

template <typename T>
static void gather_adj_t(const void *R, void *lamW, void *lamZ, const void *Zs, const Geo &g,
                         double b, double c, int zero_lamZ, hipStream_t st) {
  const dim3 grid(nblkf(g.V));
#define SU2_GADJ(ND)                                                                            \
  hipLaunchKernelGGL((k_flow_gather_stage_adj<T, ND>), grid, dim3(NTF), 0, st, (const T *)R,   \
                     (T *)lamW, (T *)lamZ, (const T *)Zs, g, b, c, zero_lamZ)
  switch (g.nd) {
    case 2: SU2_GADJ(2); break;
    case 3: SU2_GADJ(3); break;
    default: SU2_GADJ(4); break;
  }
#undef SU2_GADJ
}
void launch_flow_gather_stage_adj(int prec, const void *R, void *lamW, void *lamZ, const void *Zs,
                                  const Geo &g, double b, double c, int zero_lamZ,
                                  hipStream_t st) {
  if (prec == 0) gather_adj_t<double>(R, lamW, lamZ, Zs, g, b, c, zero_lamZ, st);
  else gather_adj_t<float>(R, lamW, lamZ, Zs, g, b, c, zero_lamZ, st);
}

template <typename T>
static void clover_t(const void *U, void *Gam, void *lamW, const Geo &g, double scale, hipStream_t st) {
  dim3 grid(nblkf(g.V));
  switch (g.nd) {
    case 2:
      hipLaunchKernelGGL((k_clover_im<T, 2>), grid, dim3(NTF), 0, st, (const quat<T> *)U, (T *)Gam, g);
      hipLaunchKernelGGL((k_clover_grad<T, 2>), grid, dim3(NTF), 0, st, (const quat<T> *)U, (const T *)Gam, (T *)lamW, g, scale);
      break;
    case 3:
      hipLaunchKernelGGL((k_clover_im<T, 3>), grid, dim3(NTF), 0, st, (const quat<T> *)U, (T *)Gam, g);
      hipLaunchKernelGGL((k_clover_grad<T, 3>), grid, dim3(NTF), 0, st, (const quat<T> *)U, (const T *)Gam, (T *)lamW, g, scale);
      break;
    default:
      hipLaunchKernelGGL((k_clover_im<T, 4>), grid, dim3(NTF), 0, st, (const quat<T> *)U, (T *)Gam, g);
      hipLaunchKernelGGL((k_clover_grad<T, 4>), grid, dim3(NTF), 0, st, (const quat<T> *)U, (const T *)Gam, (T *)lamW, g, scale);
      break;
  }
}
void launch_clover_grad(int prec, const void *U, void *Gam, void *lamW, const Geo &g, double scale,
                        hipStream_t st) {
  if (prec == 0) clover_t<double>(U, Gam, lamW, g, scale, st);
  else clover_t<float>(U, Gam, lamW, g, scale, st);
}

}  // namespace su2k
