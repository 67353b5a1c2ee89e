// su2_kernels.hip -- hand-written gfx950 kernels for the pure-gauge HMC
// molecular-dynamics step (staples, Wilson force, momentum/link updates,
// plaquette / kinetic sums, clover energy density, layout conversion, device
// momenta).  No MFMA: 2x2 complex products with a different matrix per link are
// not a dense contraction.  The generic kernels here serve every group,
// dimension and geometry; the SU(2) d=4 MD step and plaquette sum have
// specialised forms in su2_md_fast.hpp.
//
// Field layout in HBM (one lattice):
//   U : nd planes of V links, plane mu holds link (s, mu) at U[mu*V + s];
//       SU(2) link = quat<T> (16 B FP32 / 32 B FP64, one or two 128-bit loads),
//       U(1) link = T phase.
//   P : nd*NMOM planes of V reals, component c of link (s,mu) at P[(mu*NMOM+c)*V+s].
// Sites s follow the reference's lexicographic order (gaugeconfig.hh:211-222).
#include <cstdlib>
#include <cstring>

#include "su2_device.hpp"
#include "su2_launch.hpp"
#include "su2_md_site.hpp"
#include "su2_md_fast.hpp"
#include "su2_md_tw.hpp"

#include <type_traits>

using namespace su2dev;

namespace su2k {

constexpr int NT = 256;  // threads per workgroup (4 wave64)

// ----------------------------------------------------------------------------
// Fused force + momentum update (+ link update).  One thread per site, all nd
// link directions.  For every link (s, mu):
//   K    = sum_{nu != mu} w_{mu nu} [upper + lower staple]   (get_staples.hh:26-80)
//   F    = coef * get_deriv(U_mu(s) K)        (flat-gaugemonomial.hh:105-131)
//   P   <- P - ep * F                          (update_momenta.hh:27-30)
//   U'   = exp(eu * P) U_mu(s)  -> Uout        (update_gauge.hh:11-14)
// Uin is read only; the link update goes to a different buffer (ping-pong), so
// one launch is one "P then U" pair of the integrator with no grid barrier.
// Optional block partial sums: |P_in|^2, ReTr(U K) and |P_out|^2 in FP64.
// ----------------------------------------------------------------------------
template <class G, int ND, bool WRITE_U>
__global__ __launch_bounds__(NT) void k_md_fused(const typename G::link *__restrict__ Uin,
                                                 typename G::link *__restrict__ Uout,
                                                 typename G::real *__restrict__ P, Geo geo,
                                                 FusedCoef c, double *__restrict__ partials,
                                                 int pstride) {
  __shared__ double lds[NT / 64];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = lb * NT + threadIdx.x;
  double kin_old = 0.0, act = 0.0, kin_new = 0.0;
  if (s < geo.V) md_site<G, ND, WRITE_U>(Uin, Uout, P, geo, c, s, kin_old, act, kin_new);
  if (c.flags & F_ACCUM) {
    double r0 = block_sum<NT>(kin_old, lds);
    double r1 = block_sum<NT>(act, lds);
    double r2 = block_sum<NT>(kin_new, lds);
    if (threadIdx.x == 0) {
      partials[0 * pstride + lb] = r0;
      partials[1 * pstride + lb] = r1;
      partials[2 * pstride + lb] = r2;
    }
  }
}

// The same step with one thread per link: logical block lb -> (site block, mu)
// with mu fastest, so the nd blocks of one site range run together on one XCD.
template <class G, int ND, bool WRITE_U, bool ZS>
__global__ __launch_bounds__(NT) void k_md_link(const typename G::link *__restrict__ Uin,
                                                typename G::link *__restrict__ Uout,
                                                typename G::real *__restrict__ P, Geo geo,
                                                FusedCoef c, double *__restrict__ partials,
                                                int pstride) {
  __shared__ double lds[NT / 64];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int mu = lb % ND, s = (lb / ND) * NT + threadIdx.x;
  double kin_old = 0.0, act = 0.0, kin_new = 0.0;
  // with ZS every lane is active (V % NT == 0 is checked by the launcher)
  if (ZS || s < geo.V)
    md_link<G, ND, WRITE_U, ZS>(Uin, Uout, P, geo, c, s, mu, kin_old, act, kin_new);
  if (c.flags & F_ACCUM) {
    double r0 = block_sum<NT>(kin_old, lds);
    double r1 = block_sum<NT>(act, lds);
    double r2 = block_sum<NT>(kin_new, lds);
    if (threadIdx.x == 0) {
      partials[0 * pstride + lb] = r0;
      partials[1 * pstride + lb] = r1;
      partials[2 * pstride + lb] = r2;
    }
  }
}

// U <- exp(eps P) U in place, per link (update_gauge.hh:8-16; round variant :18-27)
template <class G>
__global__ __launch_bounds__(NT) void k_update_gauge(typename G::link *__restrict__ U,
                                                     const typename G::real *__restrict__ P,
                                                     int V, int nd, double eps, double round_n,
                                                     int flags) {
  using T = typename G::real;
  constexpr int NM = G::NMOM;
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= V * nd) return;
  const int mu = i / V, s = i - mu * V;
  T p[NM];
#pragma unroll
  for (int k = 0; k < NM; ++k) p[k] = P[(mu * NM + k) * V + s];
  typename G::link u = U[i];
  if (flags & F_ROUND) u = G::round(u, T(round_n));
  u = G::expmul(p, T(eps), u);
  if (flags & F_RESTORE) u = G::restore(u);
  U[i] = u;
}

// restoreSU (gaugeconfig.hh:115-120) in place
template <class G>
__global__ __launch_bounds__(NT) void k_restore(typename G::link *__restrict__ U, int n) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) U[i] = G::restore(U[i]);
}

// FP64 SU(2) half-field layout (su2_md_fast.hpp QHalf) -> record layout
__global__ __launch_bounds__(NT) void k_half_to_record(const su2_d2 *__restrict__ src,
                                                       quat<double> *__restrict__ dst, int n) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) {
    const su2_d2 a = src[i], b = src[(size_t)n + i];
    dst[i] = quat<double>{a.x, a.y, b.x, b.y};
  }
}

// Plaquette sum (retr_sum_Wplaquettes, flat-gauge_energy.hpp:31-94) computed in
// FP64 whatever the storage precision.  With F_RESTORE each link is restored on
// load and the restored links of the site are written to Uout (the end of an
// integrate() followed by the accept step's action, md_update.hh:47-51).
template <class GS, class GD, int ND>
__global__ __launch_bounds__(NT) void k_plaq_sum(const typename GS::link *__restrict__ Uin,
                                                 typename GS::link *__restrict__ Uout, Geo geo,
                                                 double w_t, double w_s, int mu_start, int flags,
                                                 double *__restrict__ partials) {
  using linkd = typename GD::link;
  __shared__ double lds[NT / 64];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = lb * NT + threadIdx.x;
  const int V = geo.V;
  double sum = 0.0;
  if (s < V) {
    int up[ND], dn[ND];
    neighbour_offsets<ND>(geo, s, up, dn);
    const bool rs = (flags & F_RESTORE) != 0;
    linkd own[ND];
#pragma unroll
    for (int nu = 0; nu < ND; ++nu) {
      typename GS::link u = Uin[nu * V + s];
      if (rs) {
        u = GS::restore(u);
        Uout[nu * V + s] = u;
      }
      own[nu] = to_d(u);
    }
#pragma unroll
    for (int mu = 0; mu < ND - 1; ++mu) {
      if (mu < mu_start) continue;
#pragma unroll
      for (int nu = mu + 1; nu < ND; ++nu) {
        typename GS::link a = Uin[nu * V + s + up[mu]], b = Uin[mu * V + s + up[nu]];
        if (rs) {
          a = GS::restore(a);
          b = GS::restore(b);
        }
        const double v = GD::plaq(own[mu], to_d(a), to_d(b), own[nu]);
        const double eta = (mu == 0) ? w_t : w_s;  // mu < nu: temporal iff mu == 0
        sum += (flags & F_ANISO) ? eta * v : v;
      }
    }
  }
  const double r = block_sum<NT>(sum, lds);
  if (threadIdx.x == 0) partials[lb] = r;
}

// 1/2 sum p^2 partials (kineticmonomial, monomial.hh:33-47), FP64 accumulation
template <typename T>
__global__ __launch_bounds__(NT) void k_sumsq(const T *__restrict__ P, int n,
                                              double *__restrict__ partials) {
  __shared__ double lds[NT / 64];
  double v = 0.0;
  for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
    const double x = double(P[i]);
    v += x * x;
  }
  const double r = block_sum<NT>(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <typename T> __global__ __launch_bounds__(NT) void k_scale(T *__restrict__ P, int n, T f) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) P[i] = f * P[i];
}

// Deterministic final reduction, one block per listed row of block partials:
// TwoSum-compensated strided accumulation, then a fixed-order LDS tree.
// out[2*r] + out[2*r+1] is the sum of row r.
constexpr int NTR = 1024;
__global__ __launch_bounds__(NTR) void k_reduce(const double *__restrict__ partials,
                                                size_t row_stride, RowList rl,
                                                double *__restrict__ out) {
  __shared__ double hs[NTR], ls[NTR];
  const int r = rl.row[blockIdx.x], nb = rl.nb[blockIdx.x];
  const double *row = partials + (size_t)r * row_stride;
  // four independent compensated accumulators per thread keep four loads in
  // flight; they are combined in a fixed order
  double h4[4] = {0.0, 0.0, 0.0, 0.0}, l4[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += 4 * NTR) {
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = (i + k * NTR < nb) ? row[i + k * NTR] : 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double e;
      two_sum(h4[k], v[k], h4[k], e);
      l4[k] += e;
    }
  }
  double hi = h4[0], lo = l4[0];
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    double e;
    two_sum(hi, h4[k], hi, e);
    lo += l4[k] + e;
  }
  hs[threadIdx.x] = hi;
  ls[threadIdx.x] = lo;
  __syncthreads();
  for (int w = NTR / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double e, h;
      two_sum(hs[threadIdx.x], hs[threadIdx.x + w], h, e);
      hs[threadIdx.x] = h;
      ls[threadIdx.x] = ls[threadIdx.x] + ls[threadIdx.x + w] + e;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double h, e;
    two_sum(hs[0], ls[0], h, e);
    out[2 * r] = h;
    out[2 * r + 1] = e;
  }
}

// ---------------------------------------------------------------------------
// clover energy density and topological charge (flat-energy_density.hh:57-164)
// per site: G_{mu nu} = traceless-antihermitian part of the 4-leaf clover.
// partials rows: 0 = sum ReTr(G G), 1 = sum eps ReTr(G G) (d=4) / -Im tr G01 (d=2)
// ---------------------------------------------------------------------------
template <class GS, int ND>
__global__ __launch_bounds__(NT) void k_energy_density(const typename GS::link *__restrict__ Uin,
                                                       Geo geo, int clover, int mu_start,
                                                       double *__restrict__ partials, int pstride) {
  __shared__ double lds[NT / 64];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int s = lb * NT + threadIdx.x;
  const int V = geo.V;
  double e_sum = 0.0, q_sum = 0.0;
  if (s < V) {
    int up[ND], dn[ND];
    neighbour_offsets<ND>(geo, s, up, dn);
    auto L = [&](int site, int d) { return to_d(Uin[d * V + site]); };
    if constexpr (GS::NCOMP == 4) {
      quat<double> Gm[ND][ND];
#pragma unroll
      for (int mu = 0; mu < ND - 1; ++mu)
#pragma unroll
        for (int nu = mu + 1; nu < ND; ++nu) {
          if (mu < mu_start) { Gm[mu][nu] = quat<double>{0, 0, 0, 0}; continue; }
          const int x1 = s + up[mu], x2 = s + up[nu];
          quat<double> leaf = qmul(qmul(qmul(L(s, mu), L(x1, nu)), qdag(L(x2, mu))), qdag(L(s, nu)));
          if (clover) {
            const int a1 = s + dn[mu] + up[nu], a2 = s + dn[mu];
            leaf = qadd(leaf, qmul(qmul(qmul(L(s, nu), qdag(L(a1, mu))), qdag(L(a2, nu))), L(a2, mu)));
            const int b1 = s + dn[mu], b2 = s + dn[mu] + dn[nu], b3 = s + dn[nu];
            leaf = qadd(leaf, qmul(qmul(qmul(qdag(L(b1, mu)), qdag(L(b2, nu))), L(b2, mu)), L(b3, nu)));
            const int c1 = s + dn[nu], c2 = s + dn[nu] + up[mu];
            leaf = qadd(leaf, qmul(qmul(qmul(qdag(L(c1, nu)), L(c1, mu)), L(c2, nu)), qdag(L(s, mu))));
          }
          // traceless_antiherm (su2.hh:83-85): (i Im a, b)
          const quat<double> g = {0.0, leaf.ai, leaf.br, leaf.bi};
          Gm[mu][nu] = g;
          e_sum += 2.0 * qmul(g, g).ar;
        }
      if constexpr (ND == 4) {
        // eps_{0123}-ordered sum over mu<nu, rho<sigma (tensors.hh:23-58):
        // pairs (01,23), (02,13), (03,12) and their swaps, signs + - + + - +
        const double t0123 = 2.0 * qmul(Gm[0][1], Gm[2][3]).ar;
        const double t0213 = 2.0 * qmul(Gm[0][2], Gm[1][3]).ar;
        const double t0312 = 2.0 * qmul(Gm[0][3], Gm[1][2]).ar;
        const double t1203 = 2.0 * qmul(Gm[1][2], Gm[0][3]).ar;
        const double t1302 = 2.0 * qmul(Gm[1][3], Gm[0][2]).ar;
        const double t2301 = 2.0 * qmul(Gm[2][3], Gm[0][1]).ar;
        q_sum = ((((t0123 - t0213) + t0312) + t1203) - t1302) + t2301;
      }
    } else {
      double Gi[ND][ND];
#pragma unroll
      for (int mu = 0; mu < ND - 1; ++mu)
#pragma unroll
        for (int nu = mu + 1; nu < ND; ++nu) {
          if (mu < mu_start) { Gi[mu][nu] = 0.0; continue; }
          const int x1 = s + up[mu], x2 = s + up[nu];
          double sn, cs, im = 0.0;
          sincos(((L(s, mu) + L(x1, nu)) - L(x2, mu)) - L(s, nu), &sn, &cs);
          im += sn;
          if (clover) {
            const int a1 = s + dn[mu] + up[nu], a2 = s + dn[mu];
            im += sin(((L(s, nu) - L(a1, mu)) - L(a2, nu)) + L(a2, mu));
            const int b1 = s + dn[mu], b2 = s + dn[mu] + dn[nu], b3 = s + dn[nu];
            im += sin(((-L(b1, mu) - L(b2, nu)) + L(b2, mu)) + L(b3, nu));
            const int c1 = s + dn[nu], c2 = s + dn[nu] + up[mu];
            im += sin(((-L(c1, nu) + L(c1, mu)) + L(c2, nu)) - L(s, mu));
          }
          Gi[mu][nu] = im;
          e_sum += -im * im;  // Re((i g)(i g))
        }
      if constexpr (ND == 4) {
        const double t0123 = -Gi[0][1] * Gi[2][3], t0213 = -Gi[0][2] * Gi[1][3];
        const double t0312 = -Gi[0][3] * Gi[1][2], t1203 = -Gi[1][2] * Gi[0][3];
        const double t1302 = -Gi[1][3] * Gi[0][2], t2301 = -Gi[2][3] * Gi[0][1];
        q_sum = ((((t0123 - t0213) + t0312) + t1203) - t1302) + t2301;
      } else if constexpr (ND == 2) {
        q_sum = -Gi[0][1];
      }
    }
  }
  const double r0 = block_sum<NT>(e_sum, lds);
  const double r1 = block_sum<NT>(q_sum, lds);
  if (threadIdx.x == 0) {
    partials[lb] = r0;
    partials[pstride + lb] = r1;
  }
}

// ---------------------------------------------------------------------------
// layout conversion between the reference AoS order (site-major, mu fastest,
// gaugeconfig.hh:211-222) held as doubles with a per-link stride, and the
// direction-planar device layout.
// ---------------------------------------------------------------------------
template <typename T, int NC>
__global__ __launch_bounds__(NT) void k_links_from_aos(const double *__restrict__ src, int stride,
                                                       int offset, T *__restrict__ dst, int V,
                                                       int nd) {
  const int i = blockIdx.x * NT + threadIdx.x;  // AoS link index s*nd + mu
  if (i >= V * nd) return;
  const int s = i / nd, mu = i - s * nd;
  const double *p = src + (size_t)i * stride + offset;
#pragma unroll
  for (int k = 0; k < NC; ++k) dst[((size_t)mu * V + s) * NC + k] = T(p[k]);
}
template <typename T, int NC>
__global__ __launch_bounds__(NT) void k_links_to_aos(const T *__restrict__ src, double *__restrict__ dst,
                                                     int stride, int offset, int V, int nd) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= V * nd) return;
  const int s = i / nd, mu = i - s * nd;
  double *p = dst + (size_t)i * stride + offset;
#pragma unroll
  for (int k = 0; k < NC; ++k) p[k] = double(src[((size_t)mu * V + s) * NC + k]);
}
// momenta: AoS (s, mu, c) <-> planar ((mu*NM + c)*V + s)
template <typename T>
__global__ __launch_bounds__(NT) void k_mom_from_aos(const double *__restrict__ src, T *__restrict__ dst,
                                                     int V, int nd, int nm) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= V * nd) return;
  const int s = i / nd, mu = i - s * nd;
  for (int k = 0; k < nm; ++k) dst[((size_t)mu * nm + k) * V + s] = T(src[(size_t)i * nm + k]);
}
template <typename T>
__global__ __launch_bounds__(NT) void k_mom_to_aos(const T *__restrict__ src, double *__restrict__ dst,
                                                   int V, int nd, int nm) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= V * nd) return;
  const int s = i / nd, mu = i - s * nd;
  for (int k = 0; k < nm; ++k) dst[(size_t)i * nm + k] = double(src[((size_t)mu * nm + k) * V + s]);
}

// Gaussian momenta from the counter-based Philox4x32-10 stream (device
// throughput mode): link (s, mu) draws from counter (s*nd+mu, 0, traj_lo,
// traj_hi) with key = seed; Box-Muller on 32-bit uniforms in the field's
// precision.  Threads run direction-major (thread mu*V + s), so each store
// instruction of a wave writes one contiguous run of a component plane; the
// counter, and with it every value, does not depend on that order.
template <typename T>
__global__ __launch_bounds__(NT) void k_momenta_philox(T *__restrict__ P, int V, int nd, int nm,
                                                       uint64_t seed, uint64_t traj) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= V * nd) return;
  const int mu = i / V, s = i - mu * V;
  T z[4];
  philox_normals4<T>((uint32_t)(s * nd + mu), (uint32_t)(traj & 0xffffffffu),
                     (uint32_t)(traj >> 32), seed, z);
  for (int k = 0; k < nm; ++k) P[((size_t)mu * nm + k) * V + s] = z[k];
}

// plain streaming copy (16 B per lane per iteration): the measured HBM copy
// bandwidth and the calibration of the FETCH_SIZE / WRITE_SIZE counters
__global__ __launch_bounds__(NT) void k_copy16(const float4 *__restrict__ src,
                                               float4 *__restrict__ dst, long n) {
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT)
    dst[i] = src[i];
}

// ------------------------------------------------------------- launchers ----
// one float4 per thread (the fastest of the copy shapes measured on MI355X:
// 6.2 TB/s for 1 GiB, against 4.7 TB/s for a grid-stride loop over 8192 blocks)
void launch_copy16(const void *src, void *dst, long n16, hipStream_t st) {
  const long nb = (n16 + NT - 1) / NT;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)(nb < 1 ? 1 : nb)), dim3(NT), 0, st,
                     (const float4 *)src, (float4 *)dst, n16);
}
static inline int nblk(long n) { return (int)((n + NT - 1) / NT); }

// kernel shape of the fused step: SU2GPU_FUSED=site|link|linkz (default: linkz
// where 64 % L_fastest == 0 and V % 256 == 0, else link)
static int fused_mode_env() {
  static int mode = -1;
  if (mode < 0) {
    const char *e = getenv("SU2GPU_FUSED");
    mode = !e ? 3 : (!strcmp(e, "site") ? 0 : (!strcmp(e, "link") ? 1 : (!strcmp(e, "linkz") ? 2 : 3)));
  }
  return mode;
}
int fused_mode(const Geo &g) {
  const int m = fused_mode_env();
  const bool zs_ok = (64 % g.L[g.nd - 1] == 0) && (g.V % NT == 0);
  if (m == 0) return 0;
  if (m == 1) return 1;
  return zs_ok ? 2 : 1;
}

// SU2GPU_FAST = 0 (generic kernels only) | 1 (specialised SU(2) d=4 step for
// the FP32 form only) | 2 (also for FP64, default); SU2GPU_TBLOCK = 0 disables
// the t-blocked XCD layout of the specialised step
static int env_int(const char *name, int def) {
  const char *e = getenv(name);
  return e ? atoi(e) : def;
}
static int fast_level() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_FAST", 2);
  return v;
}
static int tblock_pref() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_TBLOCK", 1);
  return v;
}
// SU2GPU_FASTSHAPE = 0: one thread per link (default), 1: one thread per site
static int fast_shape() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_FASTSHAPE", 0);
  return v;
}
// SU2GPU_XWALK = 1 / 0: the t-blocked walk takes the 3-volume blocks x-fastest
// / in memory order; default: x-fastest when an x plane holds at least 16
// blocks (64^4: 1436-1441 against 1496-1501 us in FP32, 3006-3024 against
// 3180 us in FP64; 48^4, 9 blocks per plane: 487 against 480 us)
static int xwalk_pref() {
  static int v = -2;
  if (v == -2) v = env_int("SU2GPU_XWALK", -1);
  return v;
}
// SU2GPU_NT = 1 / 0: non-temporal stores of the new links and momenta in the
// specialised step; default: where the two link fields and the momenta exceed
// the 256 MB Infinity Cache (48^4 FP32 466 against 496 us, 64^4 1331 against
// 1446 us; at 32^4, which fits, 79 against 67 us).  2: also non-temporal
// momentum loads in the FP64 half-field step (default beyond 1 GiB)
static int nt_pref() {
  static int v = -2;
  if (v == -2) v = env_int("SU2GPU_NT", -1);
  return v;
}
// SU2GPU_HALF = 0 keeps the FP64 trajectory's intermediate links in the record
// layout instead of the half-field layout
static int half_pref() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_HALF", 1);
  return v;
}
// SU2GPU_ZG = 1 / 0: the general +-z lane exchange in the FP32 specialised
// step where Lz does not divide 64; default: where the fields exceed the
// Infinity Cache (48^4: 388-397 against 454-478 us without; 24^4: 24.5 against
// 23.2 us)
static int zg_pref() {
  static int v = -2;
  if (v == -2) v = env_int("SU2GPU_ZG", -1);
  return v;
}
// SU2GPU_TW = 0 keeps the lane-exchange step on one launch per link instead
// of the t-walk (k_md_su2_4d_tw) where that applies
static int tw_pref() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_TW", 1);
  return v;
}
// the t-walk's chunks (SU2GPU_TWC = C in 1, 2, 4, 8; default 2): C t-chunks of
// Lt / C slices, XCD k walks chunk k % C of the 3-volume part k / C (8 / C
// parts), so a thread's first slice, which loads what a slice below would
// have passed on, comes once per Lt / C slices instead of once per Lt / 8
// (32^4 FP64 141-142.5 against 143.4-145.9 us, FP32 53.6-54.2 against
// 54.7-55.3 us, 64^4 FP64 2484 against 2508 us; alternating runs, one box).
// Falls back to one chunk per XCD where the parts do not divide the blocks.
// The general-exchange walk (wmu) keeps one chunk per XCD by default
// (SU2GPU_TWCZ; two chunks: 48^4 FP64 805-808 against 785-786 us).
static FastGeo tw_geo(const FastGeo &f, bool wmu) {
  static int c = -1, cz = -1;
  if (c < 0) c = env_int("SU2GPU_TWC", 2);
  if (cz < 0) cz = env_int("SU2GPU_TWCZ", 8);
  FastGeo g = f;
  const int cc = wmu ? cz : c;
  const int lc = cc == 1 ? 0 : cc == 2 ? 1 : cc == 4 ? 2 : 3;
  const int C = 1 << lc, P = 8 / C;
  if (lc == 3 || f.Lt % C != 0 || (wmu ? f.nb64 : f.nb3) % P != 0 || f.Lt / C < 4) return g;
  g.tw_lc = lc;
  g.Tc = f.Lt / C;
  return g;
}
// SU2GPU_TWW = 0: the FP64 lane-exchange walk with one direction per block
// (k_md_su2_4d_tw) instead of one per wave (k_md_su2_4d_tww)
static int tww_pref() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_TWW", 1);
  return v;
}
// SU2GPU_TWA = 0: the first and the last step of a trajectory (kinetic and
// action partials, momentum-only last step) as one launch per link instead of
// the t-walks
static int twa_pref() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_TWA", 1);
  return v;
}
// SU2GPU_ZS = 0 disables the +-z lane exchange of the specialised step
static int zs_pref() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_ZS", 1);
  return v;
}

// The specialised step on FP64 links in the half-field layout: input UI
// (QHalf or a record-layout pointer), output QHalfW (a P-only step passes a
// null QHalfW it never writes)
// nt: 0 ordinary stores, 1 non-temporal stores, 2 also non-temporal
// momentum loads
template <bool W, bool A, class UI>
static void launch_half(bool zs, int nt, UI uin, QHalfW uout, const FusedLaunch &L,
                        const FastGeo &f, hipStream_t st) {
  // P-then-U steps with the lane exchange and >= 4 t-slices per XCD: the
  // t-walk (32^4: 140-144 against 145-147 us in the same process; the last
  // bit of ~0.6 % of its values differs from this step's, the update's FMA
  // contraction, within the oracle parity either way)
  if constexpr (W && !A) {
    // one direction per wave (k_md_su2_4d_tww, the mapping of the
    // general-exchange walk; SU2GPU_TWW=0: one per block, k_md_su2_4d_tw)
    // below 1 GiB of fields: 32^4 138.3-138.6 against 139.6-139.7 us
    // (alternating runs on one box); beyond it one per block in two chunks
    // (64^4: 2462 against 2511-2513 us)
    if (zs && f.tblock && f.Tc >= 4 && tw_pref() && tww_pref() && nt < 2) {
      const FastGeo g = tw_geo(f, true);
      const dim3 gt(tw_blocks_w(g)), bt(NTF);
      if (nt == 2)
        hipLaunchKernelGGL((k_md_su2_4d_tww<double, 2, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      else if (nt == 1)
        hipLaunchKernelGGL((k_md_su2_4d_tww<double, 1, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      else
        hipLaunchKernelGGL((k_md_su2_4d_tww<double, 0, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      return;
    }
    if (zs && f.tblock && f.Tc >= 4 && tw_pref()) {
      const FastGeo g = tw_geo(f, false);
      const dim3 gt(tw_blocks(g)), bt(NTF);
      if (nt == 2)
        hipLaunchKernelGGL((k_md_su2_4d_tw<double, 1, 2, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      else if (nt == 1)
        hipLaunchKernelGGL((k_md_su2_4d_tw<double, 1, 1, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      else
        hipLaunchKernelGGL((k_md_su2_4d_tw<double, 1, 0, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      return;
    }
    // without the whole-row exchange (Lz does not divide 64): the t-walk with
    // the general exchange, one direction per wave (48^4: 776 against 832 us
    // in the same process; within the last bit of this step, as above)
    if (!zs && zs_pref() && zg_pref() != 0 && f.tblock && f.Tc >= 4 && tw_pref()) {
      const FastGeo g = tw_geo(f, true);
      const dim3 gt(tw_blocks_w(g)), bt(NTF);
      if (nt == 2)
        hipLaunchKernelGGL((k_md_su2_4d_twz<double, 2, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      else if (nt == 1)
        hipLaunchKernelGGL((k_md_su2_4d_twz<double, 1, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      else
        hipLaunchKernelGGL((k_md_su2_4d_twz<double, 0, UI, QHalfW>), gt, bt, 0, st, uin, uout,
                           (double *)L.P, g, L.c);
      return;
    }
  }
  // the first step (kinetic-energy and action partials) and the last,
  // momentum-only step of a trajectory take the same walks
  if constexpr (A) {
    if (f.tblock && f.Tc >= 4 && tw_pref() && twa_pref()) {
      if (zs) {
        const FastGeo g = tw_geo(f, false);
        const dim3 gt(tw_blocks(g)), bt(NTF);
#define SU2_TWA(N)                                                                              \
  hipLaunchKernelGGL((k_md_su2_4d_tw_acc<double, W, 1, N, UI, QHalfW>), gt, bt, 0, st, uin, uout, \
                     (double *)L.P, g, L.c, L.partials, L.pstride)
        if (nt == 2) SU2_TWA(2);
        else if (nt == 1) SU2_TWA(1);
        else SU2_TWA(0);
#undef SU2_TWA
        return;
      }
      if (zs_pref() && zg_pref() != 0) {
        const FastGeo g = tw_geo(f, true);
        const dim3 gt(tw_blocks_w(g)), bt(NTF);
#define SU2_TWA(N)                                                                              \
  hipLaunchKernelGGL((k_md_su2_4d_twz_acc<double, W, N, UI, QHalfW>), gt, bt, 0, st, uin, uout,   \
                     (double *)L.P, g, L.c, L.partials, L.pstride)
        if (nt == 2) SU2_TWA(2);
        else if (nt == 1) SU2_TWA(1);
        else SU2_TWA(0);
#undef SU2_TWA
        return;
      }
    }
  }
  const dim3 grid(4 * (L.geo.V / NTF)), block(NTF);
#define SU2_HALF(Z, N)                                                                          \
  hipLaunchKernelGGL((k_md_su2_4d<double, W, A, 1, Z, false, N, UI, QHalfW>), grid, block, 0, st, \
                     uin, uout, (double *)L.P, f, L.c, L.partials, L.pstride)
  if (zs) {
    if (nt == 2) SU2_HALF(1, 2);
    else if (nt == 1) SU2_HALF(1, 1);
    else SU2_HALF(1, 0);
  } else {
    if (nt == 2) SU2_HALF(0, 2);
    else if (nt == 1) SU2_HALF(0, 1);
    else SU2_HALF(0, 0);
  }
#undef SU2_HALF
}

// the conditions under which launch_fast takes the specialised thread-per-link
// step for a launch (shared by launch_fast and fused_half_ok)
static bool fast_geo_for(const Geo &g, FastGeo &f, bool wmu = false) {
  const int xw = xwalk_pref();
  const bool xwalk = xw >= 0 ? xw != 0 : (g.L[2] * g.L[3]) / NTF >= 16;
  return make_fast_geo(g.L, g.V, tblock_pref() != 0, f, xwalk, wmu);
}

bool fused_half_ok(int group, int prec, const Geo &g, int flags) {
  if (group != 1 || prec != 0 || g.nd != 4) return false;
  if (fast_level() < 2 || fast_shape() == 1 || half_pref() == 0) return false;
  if (flags & (F_PZERO | F_ROUND | F_ANISO | F_R32L | F_R32P)) return false;
  if (fused_mode(g) == 0) return false;
  FastGeo f;
  return fast_geo_for(g, f);
}

template <class G, int ND>
static bool launch_fast(const FusedLaunch &L, hipStream_t st) {
  using T = typename G::real;
  if constexpr (ND != 4 || !std::is_same<G, GSU2<T>>::value) {
    return false;
  } else {
    const int lvl = fast_level();
    if (lvl == 0 || (sizeof(T) == 8 && lvl < 2)) return false;
    if (L.c.flags & (F_PZERO | F_ROUND | F_R32L | F_R32P)) return false;
    // the thread-per-site shape keeps its own partial-sum count (fused_blocks)
    if (fused_mode(L.geo) == 0) return false;
    FastGeo f;
    const bool acc = (L.c.flags & F_ACCUM) != 0;
    if (!fast_geo_for(L.geo, f)) {
      // FP32: the same step body over the lattice's own geometry (FP64 keeps
      // the generic kernel there: 9.9 against 11.1 us at 12^4, 28.5 against
      // 29.8 us at 20^4; FP32 8.5 -> 8.1 and 17.6 -> 16.4 us)
      if (sizeof(T) == 8 || fast_shape() == 1) return false;
      const bool zs = zs_pref() && (64 % L.geo.L[3] == 0) && (L.geo.V % 64 == 0);
      const bool an = (L.c.flags & F_ANISO) != 0;
      const dim3 grid(4 * ((L.geo.V + NTF - 1) / NTF)), block(NTF);
#define SU2_GEO(W, A, Z, N)                                                                 \
  hipLaunchKernelGGL((k_md_su2_geo<T, W, A, Z, N>), grid, block, 0, st, (const quat<T> *)L.Uin, \
                     (quat<T> *)L.Uout, (T *)L.P, L.geo, L.c, L.partials, L.pstride)
#define SU2_GEO2(W, A)                      \
  if (an && zs) SU2_GEO(W, A, true, true);  \
  else if (an) SU2_GEO(W, A, false, true);  \
  else if (zs) SU2_GEO(W, A, true, false);  \
  else SU2_GEO(W, A, false, false)
      if (L.Uout) {
        if (acc) SU2_GEO2(true, true);
        else SU2_GEO2(true, false);
      } else {
        if (acc) SU2_GEO2(false, true);
        else SU2_GEO2(false, false);
      }
#undef SU2_GEO2
#undef SU2_GEO
      return true;
    }
    const bool site = fast_shape() == 1;
    // the thread-per-link shape weights each plane (w_t, w_s); the per-site one
    // sums unweighted
    if (site && (L.c.flags & F_ANISO)) return false;
    // +-z neighbours by lane exchange where every wave holds whole z-rows
    const bool zs = !site && zs_pref() && (64 % f.Lz == 0);
    const bool an = (L.c.flags & F_ANISO) != 0;
    const int ntp = nt_pref();
    const double fields = 4.0 * L.geo.V * (2.0 * sizeof(quat<T>) + 3.0 * sizeof(T));
    const bool nt = ntp >= 0 ? ntp != 0 : fields > 256.0 * 1024 * 1024;
    // FP32 lattices beyond the Infinity Cache whose Lz does not divide 64: the
    // general +-z lane exchange (neighbours in another wave loaded by the
    // lanes that need them, in the first memory round trip), the four
    // directions of 64 sites per block, and the momenta loaded non-temporally
    // (48^4: 454-478 -> 388-397 us; FP64 keeps the loads, 850 against 943 us,
    // where the exchange's registers cost a wave per SIMD)
    FastGeo fw;
    const int zg = zg_pref();
    if constexpr (sizeof(T) == 4) {
    if (!site && !an && !zs && zs_pref() && (zg >= 0 ? zg != 0 : nt) && !L.in_half &&
        !L.out_half && fast_geo_for(L.geo, fw, true)) {
      // P-then-U steps with >= 4 t-slices per XCD: the t-walk with the same
      // exchange, bit for bit (48^4: 386.6 against 395.6 us, 56^4: 723.5
      // against 745.2 us)
      if ((acc ? twa_pref() != 0 : L.Uout != nullptr) && fw.tblock && fw.Tc >= 4 && tw_pref()) {
        const FastGeo g = tw_geo(fw, true);
        const dim3 gt(tw_blocks_w(g)), bt(NTF);
        const quat<T> *ui = (const quat<T> *)L.Uin;
        quat<T> *uo = (quat<T> *)L.Uout;
        if (!acc) {
          if (nt)
            hipLaunchKernelGGL((k_md_su2_4d_twz<T, 2>), gt, bt, 0, st, ui, uo, (T *)L.P, g, L.c);
          else
            hipLaunchKernelGGL((k_md_su2_4d_twz<T, 0>), gt, bt, 0, st, ui, uo, (T *)L.P, g, L.c);
          return true;
        }
        // the first and the last step of a trajectory
#define SU2_TWA(W, N)                                                                          \
  hipLaunchKernelGGL((k_md_su2_4d_twz_acc<T, W, N, const quat<T> *, quat<T> *>), gt, bt, 0, st, \
                     ui, uo, (T *)L.P, g, L.c, L.partials, L.pstride)
        if (L.Uout) {
          if (nt) SU2_TWA(true, 2);
          else SU2_TWA(true, 0);
        } else {
          if (nt) SU2_TWA(false, 2);
          else SU2_TWA(false, 0);
        }
#undef SU2_TWA
        return true;
      }
      const dim3 grid(4 * (L.geo.V / NTF)), block(NTF);
#define SU2_ZG(W, A, N)                                                                         \
  hipLaunchKernelGGL((k_md_su2_4d<T, W, A, 1, 2, false, N>), grid, block, 0, st,               \
                     (const quat<T> *)L.Uin, (quat<T> *)L.Uout, (T *)L.P, fw, L.c, L.partials, \
                     L.pstride)
#define SU2_ZG2(W, A) \
  if (nt) SU2_ZG(W, A, 2); \
  else SU2_ZG(W, A, 0)
      if (L.Uout) {
        if (acc) { SU2_ZG2(true, true); }
        else { SU2_ZG2(true, false); }
      } else {
        if (acc) { SU2_ZG2(false, true); }
        else { SU2_ZG2(false, false); }
      }
#undef SU2_ZG2
#undef SU2_ZG
      return true;
    }
    }
    // FP32 P-then-U steps with the lane exchange whose fields fit in the
    // Infinity Cache and whose XCD t-chunks hold at least 4 slices: the
    // t-walk, the same arithmetic bit for bit with 10.75 instead of 14.5
    // link loads per link (32^4: 52.5 against 55.2 us)
    if constexpr (sizeof(T) == 4) {
      if (zs && !an && !nt && (acc ? twa_pref() != 0 : L.Uout != nullptr) && !L.in_half &&
          !L.out_half && f.tblock && f.Tc >= 4 && tw_pref()) {
        const FastGeo g = tw_geo(f, false);
        const dim3 gt(tw_blocks(g)), bt(NTF);
        const quat<T> *ui = (const quat<T> *)L.Uin;
        quat<T> *uo = (quat<T> *)L.Uout;
        if (!acc)
          hipLaunchKernelGGL((k_md_su2_4d_tw<T, 1, 0>), gt, bt, 0, st, ui, uo, (T *)L.P, g, L.c);
        else if (L.Uout)
          hipLaunchKernelGGL((k_md_su2_4d_tw_acc<T, true, 1, 0, const quat<T> *, quat<T> *>), gt,
                             bt, 0, st, ui, uo, (T *)L.P, g, L.c, L.partials, L.pstride);
        else
          hipLaunchKernelGGL((k_md_su2_4d_tw_acc<T, false, 1, 0, const quat<T> *, quat<T> *>), gt,
                             bt, 0, st, ui, uo, (T *)L.P, g, L.c, L.partials, L.pstride);
        return true;
      }
    }
    if (L.in_half || L.out_half) {
      // the host asked for it only where fused_half_ok holds
      if constexpr (sizeof(T) == 8) {
        // momenta loaded non-temporally too where the fields exceed 1 GiB
        // (48^4: 832 against 850 us; 32^4, 368 MB: 161 against 148 us)
        int ntl = nt ? ((ntp < 0 && fields > 1024.0 * 1024 * 1024) || ntp == 2 ? 2 : 1) : 0;
        // the last, momentum-only step loads the momenta non-temporally too:
        // more of the links that the final restore + action launch reads next
        // stay in the Infinity Cache (32^4: that launch 71-72 against 81-84 us,
        // the trajectory 5.851-5.860 against 5.884-5.886 ms on one box)
        if (!L.Uout && ntl == 1) ntl = 2;
        const size_t nl = 4 * (size_t)L.geo.V;
        const su2_d2 *ia = (const su2_d2 *)L.Uin;
        su2_d2 *oa = (su2_d2 *)L.Uout;
        const QHalfW hout{oa, oa ? oa + nl : nullptr};
        if (L.Uout && !L.out_half) return false;  // half -> record steps are not built
        if (L.in_half) {
          const QHalf hin{ia, ia + nl};
          if (L.Uout) acc ? launch_half<true, true>(zs, ntl, hin, hout, L, f, st)
                          : launch_half<true, false>(zs, ntl, hin, hout, L, f, st);
          else acc ? launch_half<false, true>(zs, ntl, hin, hout, L, f, st)
                   : launch_half<false, false>(zs, ntl, hin, hout, L, f, st);
        } else {
          const quat<double> *rin = (const quat<double> *)L.Uin;
          acc ? launch_half<true, true>(zs, ntl, rin, hout, L, f, st)
              : launch_half<true, false>(zs, ntl, rin, hout, L, f, st);
        }
        return true;
      } else {
        return false;
      }
    }
    // with the lane exchange beyond 1 GiB of fields the momenta also load
    // non-temporally (FP32 64^4: 1285-1290 against 1315-1350 us)
    const bool nt2 = nt && (ntp == 2 || (ntp < 0 && fields > 1024.0 * 1024 * 1024));
    dim3 grid((site ? 1 : 4) * (L.geo.V / NTF)), block(NTF);
#define SU2_FAST(W, A)                                                                          \
  if (site)                                                                                     \
    hipLaunchKernelGGL((k_md_su2_4d_site<T, W, A>), grid, block, 0, st, (const quat<T> *)L.Uin, \
                       (quat<T> *)L.Uout, (T *)L.P, f, L.c, L.partials, L.pstride);             \
  else if (an && zs)                                                                            \
    hipLaunchKernelGGL((k_md_su2_4d<T, W, A, 1, true, true>), grid, block, 0, st,              \
                       (const quat<T> *)L.Uin, (quat<T> *)L.Uout, (T *)L.P, f, L.c,            \
                       L.partials, L.pstride);                                                  \
  else if (an)                                                                                  \
    hipLaunchKernelGGL((k_md_su2_4d<T, W, A, 1, false, true>), grid, block, 0, st,             \
                       (const quat<T> *)L.Uin, (quat<T> *)L.Uout, (T *)L.P, f, L.c,            \
                       L.partials, L.pstride);                                                  \
  else if (zs && nt2)                                                                           \
    hipLaunchKernelGGL((k_md_su2_4d<T, W, A, 1, 1, false, 2>), grid, block, 0, st,             \
                       (const quat<T> *)L.Uin, (quat<T> *)L.Uout, (T *)L.P, f, L.c,            \
                       L.partials, L.pstride);                                                  \
  else if (zs && nt)                                                                            \
    hipLaunchKernelGGL((k_md_su2_4d<T, W, A, 1, true, false, true>), grid, block, 0, st,       \
                       (const quat<T> *)L.Uin, (quat<T> *)L.Uout, (T *)L.P, f, L.c,            \
                       L.partials, L.pstride);                                                  \
  else if (zs)                                                                                  \
    hipLaunchKernelGGL((k_md_su2_4d<T, W, A, 1, true>), grid, block, 0, st,                    \
                       (const quat<T> *)L.Uin, (quat<T> *)L.Uout, (T *)L.P, f, L.c,            \
                       L.partials, L.pstride);                                                  \
  else if (nt)                                                                                  \
    hipLaunchKernelGGL((k_md_su2_4d<T, W, A, 1, false, false, true>), grid, block, 0, st,      \
                       (const quat<T> *)L.Uin, (quat<T> *)L.Uout, (T *)L.P, f, L.c,            \
                       L.partials, L.pstride);                                                  \
  else                                                                                          \
    hipLaunchKernelGGL((k_md_su2_4d<T, W, A>), grid, block, 0, st, (const quat<T> *)L.Uin,      \
                       (quat<T> *)L.Uout, (T *)L.P, f, L.c, L.partials, L.pstride)
    if (L.Uout) {
      if (acc) SU2_FAST(true, true);
      else SU2_FAST(true, false);
    } else {
      if (acc) SU2_FAST(false, true);
      else SU2_FAST(false, false);
    }
#undef SU2_FAST
    return true;
  }
}

// the step with a separate momentum output (FusedLaunch::Pout) in the
// specialised kernel, where it applies (isotropic, no partial sums)
template <class G, int ND>
static bool launch_fast_pout(const FusedLaunch &L, hipStream_t st) {
  using T = typename G::real;
  if constexpr (ND != 4 || !std::is_same<G, GSU2<T>>::value) {
    return false;
  } else {
    const int lvl = fast_level();
    if (lvl == 0 || (sizeof(T) == 8 && lvl < 2)) return false;
    if (L.c.flags & (F_PZERO | F_ROUND | F_ACCUM | F_ANISO | F_R32L | F_R32P)) return false;
    if (fused_mode(L.geo) == 0 || fast_shape() == 1) return false;
    FastGeo f;
    const int xw = xwalk_pref();
    const bool xwalk = xw >= 0 ? xw != 0 : (L.geo.L[2] * L.geo.L[3]) / NTF >= 16;
    if (!make_fast_geo(L.geo.L, L.geo.V, tblock_pref() != 0, f, xwalk)) return false;
    const bool zs = zs_pref() && (64 % f.Lz == 0);
    const dim3 grid(4 * (L.geo.V / NTF)), block(NTF);
#define SU2_POUT(W, Z)                                                                       \
  hipLaunchKernelGGL((k_md_su2_4d_pout<T, W, Z>), grid, block, 0, st, (const quat<T> *)L.Uin, \
                     (quat<T> *)L.Uout, (T *)L.P, (T *)L.Pout, f, L.c)
    if (L.Uout) {
      if (zs) SU2_POUT(true, true);
      else SU2_POUT(true, false);
    } else {
      if (zs) SU2_POUT(false, true);
      else SU2_POUT(false, false);
    }
#undef SU2_POUT
    return true;
  }
}

template <class G, int ND>
static bool launch_fused_nd(const FusedLaunch &L, hipStream_t st) {
  using link = typename G::link;
  using T = typename G::real;
  if (L.in_half || L.out_half) {
    // only the specialised step reads / writes the half-field layout
    if (L.Pout) return false;
    return launch_fast<G, ND>(L, st);
  }
  if (L.Pout) {
    if (launch_fast_pout<G, ND>(L, st)) return true;
    // elsewhere: copy the momenta, then the in-place step on the copy
    const size_t bytes = (size_t)L.geo.V * ND * G::NMOM * sizeof(T);
    (void)hipMemcpyAsync(L.Pout, L.P, bytes, hipMemcpyDeviceToDevice, st);
    FusedLaunch L2 = L;
    L2.P = L.Pout;
    L2.Pout = nullptr;
    return launch_fused_nd<G, ND>(L2, st);
  }
  if (launch_fast<G, ND>(L, st)) return true;
  const int mode = fused_mode(L.geo);
  if (mode != 0) {
    dim3 grid(ND * nblk(L.geo.V)), block(NT);
#define SU2_LINK(W, Z)                                                                     \
  hipLaunchKernelGGL((k_md_link<G, ND, W, Z>), grid, block, 0, st, (const link *)L.Uin,   \
                     (link *)L.Uout, (T *)L.P, L.geo, L.c, L.partials, L.pstride)
    if (L.Uout) {
      if (mode == 2) SU2_LINK(true, true);
      else SU2_LINK(true, false);
    } else {
      if (mode == 2) SU2_LINK(false, true);
      else SU2_LINK(false, false);
    }
#undef SU2_LINK
    return true;
  }
  dim3 grid(nblk(L.geo.V)), block(NT);
  if (L.Uout)
    hipLaunchKernelGGL((k_md_fused<G, ND, true>), grid, block, 0, st, (const link *)L.Uin,
                       (link *)L.Uout, (T *)L.P, L.geo, L.c, L.partials, L.pstride);
  else
    hipLaunchKernelGGL((k_md_fused<G, ND, false>), grid, block, 0, st, (const link *)L.Uin,
                       (link *)nullptr, (T *)L.P, L.geo, L.c, L.partials, L.pstride);
  return true;
}

template <class G> static bool launch_fused_g(const FusedLaunch &L, hipStream_t st) {
  switch (L.geo.nd) {
    case 2: return launch_fused_nd<G, 2>(L, st);
    case 3: return launch_fused_nd<G, 3>(L, st);
    default: return launch_fused_nd<G, 4>(L, st);
  }
}

// SU2GPU_GENFUSE = 0: device momenta always drawn by their own launch
static int genfuse_pref() {
  static int v = -1;
  if (v < 0) v = env_int("SU2GPU_GENFUSE", 1);
  return v;
}

// The first P-then-U step of a trajectory that draws its momenta itself
// (FusedLaunch::gen): the walks of launch_half / launch_fast with GEN, under
// the conditions under which those take the first step's walk.  Returns false
// (nothing launched) elsewhere; the caller then draws the momenta first.
static bool launch_first_step_gen(int group, int prec, const FusedLaunch &L, hipStream_t st) {
  if (!genfuse_pref() || group != 1 || L.geo.nd != 4 || !L.Uout || L.Pout) return false;
  if (!(L.c.flags & F_ACCUM)) return false;
  if (L.c.flags & (F_PZERO | F_ROUND | F_R32L | F_R32P | F_ANISO)) return false;
  if (fast_level() < (prec == 0 ? 2 : 1) || fast_shape() == 1 || fused_mode(L.geo) == 0) return false;
  if (!tw_pref() || !twa_pref()) return false;
  FastGeo f;
  if (!fast_geo_for(L.geo, f) || !f.tblock || f.Tc < 4) return false;
  const bool zs = zs_pref() && (64 % f.Lz == 0);
  const int ntp = nt_pref();
  const double fields =
      4.0 * L.geo.V * (2.0 * (prec == 0 ? sizeof(quat<double>) : sizeof(quat<float>)) +
                       3.0 * (prec == 0 ? 8.0 : 4.0));
  const bool nt = ntp >= 0 ? ntp != 0 : fields > 256.0 * 1024 * 1024;
  const dim3 bt(NTF);
  if (prec == 0) {
    // FP64: the walks write the half-field layout
    if (!L.out_half) return false;
    const int ntl = nt ? ((ntp < 0 && fields > 1024.0 * 1024 * 1024) || ntp == 2 ? 2 : 1) : 0;
    const size_t nl = 4 * (size_t)L.geo.V;
    su2_d2 *oa = (su2_d2 *)L.Uout;
    const QHalfW hout{oa, oa + nl};
    const su2_d2 *ia = (const su2_d2 *)L.Uin;
    const QHalf hin{ia, ia + nl};
    const quat<double> *rin = (const quat<double> *)L.Uin;
    double *P = (double *)L.P;
    if (zs) {
      const FastGeo g = tw_geo(f, false);
      const dim3 gt(tw_blocks(g));
#define SU2_GEN(UI, u, N)                                                                      \
  hipLaunchKernelGGL((k_md_su2_4d_tw_acc<double, true, 1, N, UI, QHalfW, true>), gt, bt, 0, st, \
                     u, hout, P, g, L.c, L.partials, L.pstride)
      if (L.in_half) {
        if (ntl == 2) SU2_GEN(QHalf, hin, 2);
        else if (ntl == 1) SU2_GEN(QHalf, hin, 1);
        else SU2_GEN(QHalf, hin, 0);
      } else {
        if (ntl == 2) SU2_GEN(const quat<double> *, rin, 2);
        else if (ntl == 1) SU2_GEN(const quat<double> *, rin, 1);
        else SU2_GEN(const quat<double> *, rin, 0);
      }
#undef SU2_GEN
      return true;
    }
    if (!zs_pref() || zg_pref() == 0) return false;
    const FastGeo g = tw_geo(f, true);
    const dim3 gt(tw_blocks_w(g));
#define SU2_GEN(UI, u, N)                                                                       \
  hipLaunchKernelGGL((k_md_su2_4d_twz_acc<double, true, N, UI, QHalfW, true>), gt, bt, 0, st, u, \
                     hout, P, g, L.c, L.partials, L.pstride)
    if (L.in_half) {
      if (ntl == 2) SU2_GEN(QHalf, hin, 2);
      else if (ntl == 1) SU2_GEN(QHalf, hin, 1);
      else SU2_GEN(QHalf, hin, 0);
    } else {
      if (ntl == 2) SU2_GEN(const quat<double> *, rin, 2);
      else if (ntl == 1) SU2_GEN(const quat<double> *, rin, 1);
      else SU2_GEN(const quat<double> *, rin, 0);
    }
#undef SU2_GEN
    return true;
  }
  // FP32: record-layout links
  if (L.in_half || L.out_half) return false;
  const quat<float> *ui = (const quat<float> *)L.Uin;
  quat<float> *uo = (quat<float> *)L.Uout;
  float *P = (float *)L.P;
  if (zs) {
    if (nt) return false;
    const FastGeo g = tw_geo(f, false);
    hipLaunchKernelGGL((k_md_su2_4d_tw_acc<float, true, 1, 0, const quat<float> *, quat<float> *, true>),
                       dim3(tw_blocks(g)), bt, 0, st, ui, uo, P, g, L.c, L.partials, L.pstride);
    return true;
  }
  const int zg = zg_pref();
  FastGeo fw;
  if (!zs_pref() || !(zg >= 0 ? zg != 0 : nt) || !fast_geo_for(L.geo, fw, true) || !fw.tblock ||
      fw.Tc < 4)
    return false;
  const FastGeo g = tw_geo(fw, true);
  if (nt)
    hipLaunchKernelGGL((k_md_su2_4d_twz_acc<float, true, 2, const quat<float> *, quat<float> *, true>),
                       dim3(tw_blocks_w(g)), bt, 0, st, ui, uo, P, g, L.c, L.partials, L.pstride);
  else
    hipLaunchKernelGGL((k_md_su2_4d_twz_acc<float, true, 0, const quat<float> *, quat<float> *, true>),
                       dim3(tw_blocks_w(g)), bt, 0, st, ui, uo, P, g, L.c, L.partials, L.pstride);
  return true;
}

bool launch_fused(int group, int prec, const FusedLaunch &L, hipStream_t st) {
  if (L.gen) {
    if (launch_first_step_gen(group, prec, L, st)) return true;
    launch_momenta_philox(prec, L.P, L.geo.V, L.geo.nd, group == 1 ? 3 : 1, L.c.gseed, L.c.gtraj,
                          st);
    FusedLaunch L2 = L;
    L2.gen = 0;
    return launch_fused(group, prec, L2, st);
  }
  if (group == 1) {
    if (prec == 0) return launch_fused_g<GSU2<double>>(L, st);
    return launch_fused_g<GSU2<float>>(L, st);
  }
  if (L.in_half || L.out_half) return false;
  if (prec == 0) return launch_fused_g<GU1<double>>(L, st);
  return launch_fused_g<GU1<float>>(L, st);
}

void launch_half_to_record(const void *src, void *dst, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_half_to_record, dim3(nblk(n)), dim3(NT), 0, st, (const su2_d2 *)src,
                     (quat<double> *)dst, (int)n);
}

int fused_blocks(const Geo &g) { return fused_mode(g) == 0 ? nblk(g.V) : g.nd * nblk(g.V); }
int site_blocks(const Geo &g) { return nblk(g.V); }

template <class G>
static void launch_update_gauge_g(void *U, const void *P, int V, int nd, double eps,
                                  double round_n, int flags, hipStream_t st) {
  hipLaunchKernelGGL((k_update_gauge<G>), dim3(nblk((long)V * nd)), dim3(NT), 0, st,
                     (typename G::link *)U, (const typename G::real *)P, V, nd, eps, round_n, flags);
}
void launch_update_gauge(int group, int prec, void *U, const void *P, int V, int nd, double eps,
                         double round_n, int flags, hipStream_t st) {
  if (group == 1) {
    if (prec == 0) launch_update_gauge_g<GSU2<double>>(U, P, V, nd, eps, round_n, flags, st);
    else launch_update_gauge_g<GSU2<float>>(U, P, V, nd, eps, round_n, flags, st);
  } else {
    if (prec == 0) launch_update_gauge_g<GU1<double>>(U, P, V, nd, eps, round_n, flags, st);
    else launch_update_gauge_g<GU1<float>>(U, P, V, nd, eps, round_n, flags, st);
  }
}

__global__ __launch_bounds__(NT) void k_round_f32(double *__restrict__ x, size_t n) {
  const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
  if (i < n) x[i] = (double)(float)x[i];
}
void launch_round_f32(double *x, size_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_round_f32, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, st, x, n);
}

void launch_restore(int group, int prec, void *U, int n, hipStream_t st) {
  if (group != 1) return;
  if (prec == 0)
    hipLaunchKernelGGL((k_restore<GSU2<double>>), dim3(nblk(n)), dim3(NT), 0, st, (quat<double> *)U, n);
  else
    hipLaunchKernelGGL((k_restore<GSU2<float>>), dim3(nblk(n)), dim3(NT), 0, st, (quat<float> *)U, n);
}

template <class GS, class GD>
static void launch_plaq_g(const void *Uin, void *Uout, const Geo &g, double w_t, double w_s,
                          int mu_start, int flags, double *partials, hipStream_t st) {
  using link = typename GS::link;
  dim3 grid(nblk(g.V));
  switch (g.nd) {
    case 2:
      hipLaunchKernelGGL((k_plaq_sum<GS, GD, 2>), grid, dim3(NT), 0, st, (const link *)Uin,
                         (link *)Uout, g, w_t, w_s, mu_start, flags, partials);
      break;
    case 3:
      hipLaunchKernelGGL((k_plaq_sum<GS, GD, 3>), grid, dim3(NT), 0, st, (const link *)Uin,
                         (link *)Uout, g, w_t, w_s, mu_start, flags, partials);
      break;
    default:
      hipLaunchKernelGGL((k_plaq_sum<GS, GD, 4>), grid, dim3(NT), 0, st, (const link *)Uin,
                         (link *)Uout, g, w_t, w_s, mu_start, flags, partials);
      break;
  }
}
// the plaquette sum of the specialised geometry walks t (k_plaq_su2_4d_tw) on
// the t-blocked layout (SU2GPU_TW=0: one thread per site and slice), with the
// +-z lane exchange where 64 % Lz == 0
static bool plaq_tw_ok(const FastGeo &f) { return f.tblock && tw_pref(); }
static bool plaq_tw_zs(const FastGeo &f) { return 64 % f.Lz == 0 && zs_pref(); }
int launch_plaq(int group, int prec, const void *Uin, void *Uout, const Geo &g, double w_t,
                double w_s, int mu_start, int flags, double *partials, hipStream_t st,
                int in_half) {
  FastGeo f;
  if (in_half) {
    // FP64 SU(2) links in the half-field layout (the specialised geometry,
    // fused_half_ok): restored links, if asked for, go out in the record layout
    if (!(group == 1 && prec == 0 && make_fast_geo(g.L, g.V, tblock_pref() != 0, f))) return -1;
    const dim3 grid(g.V / NTF), block(NTF);
    const int an = (flags & F_ANISO) ? 1 : 0;
    const size_t nl = 4 * (size_t)g.V;
    const QHalf hin{(const su2_d2 *)Uin, (const su2_d2 *)Uin + nl};
    if (plaq_tw_ok(f)) {
      const bool rs = (flags & F_RESTORE) != 0, zx = plaq_tw_zs(f);
      const dim3 gt(plaq_tw_blocks(f));
#define SU2_PLAQ_TW(R, Z)                                                                  \
  hipLaunchKernelGGL((k_plaq_su2_4d_tw<double, R, Z, QHalf>), gt, block, 0, st, hin,       \
                     (quat<double> *)Uout, f, w_t, w_s, mu_start, an, partials)
      if (rs) {
        if (zx) SU2_PLAQ_TW(true, 1);
        else SU2_PLAQ_TW(true, 0);
      } else {
        if (zx) SU2_PLAQ_TW(false, 1);
        else SU2_PLAQ_TW(false, 0);
      }
#undef SU2_PLAQ_TW
      return (int)plaq_tw_blocks(f);
    }
    if (flags & F_RESTORE)
      hipLaunchKernelGGL((k_plaq_su2_4d<double, true, QHalf>), grid, block, 0, st, hin,
                         (quat<double> *)Uout, f, w_t, w_s, mu_start, an, partials);
    else
      hipLaunchKernelGGL((k_plaq_su2_4d<double, false, QHalf>), grid, block, 0, st, hin,
                         (quat<double> *)Uout, f, w_t, w_s, mu_start, an, partials);
    return g.V / NTF;
  }
  if (group == 1 && g.nd == 4 && fast_level() > 0 &&
      make_fast_geo(g.L, g.V, tblock_pref() != 0, f)) {
    // specialised SU(2) d=4 sum: one thread per site, V / NTF blocks
    const dim3 grid(g.V / NTF), block(NTF);
    const bool rs = (flags & F_RESTORE) != 0;
    const int an = (flags & F_ANISO) ? 1 : 0;
    if (plaq_tw_ok(f)) {
      const bool zx = plaq_tw_zs(f);
      const dim3 gt(plaq_tw_blocks(f));
#define SU2_PLAQ_TW(T, R, Z)                                                               \
  hipLaunchKernelGGL((k_plaq_su2_4d_tw<T, R, Z>), gt, block, 0, st, (const quat<T> *)Uin,  \
                     (quat<T> *)Uout, f, w_t, w_s, mu_start, an, partials)
#define SU2_PLAQ_TW2(T, R)                                                                 \
  do {                                                                                     \
    if (zx) SU2_PLAQ_TW(T, R, 1);                                                          \
    else SU2_PLAQ_TW(T, R, 0);                                                             \
  } while (0)
      if (prec == 0) {
        if (rs) SU2_PLAQ_TW2(double, true);
        else SU2_PLAQ_TW2(double, false);
      } else {
        if (rs) SU2_PLAQ_TW2(float, true);
        else SU2_PLAQ_TW2(float, false);
      }
#undef SU2_PLAQ_TW2
#undef SU2_PLAQ_TW
      return (int)plaq_tw_blocks(f);
    }
#define SU2_PLAQ(T, R)                                                                     \
  hipLaunchKernelGGL((k_plaq_su2_4d<T, R>), grid, block, 0, st, (const quat<T> *)Uin,     \
                     (quat<T> *)Uout, f, w_t, w_s, mu_start, an, partials)
    if (prec == 0) {
      if (rs) SU2_PLAQ(double, true);
      else SU2_PLAQ(double, false);
    } else {
      if (rs) SU2_PLAQ(float, true);
      else SU2_PLAQ(float, false);
    }
#undef SU2_PLAQ
    return g.V / NTF;
  }
  if (group == 1) {
    if (prec == 0) launch_plaq_g<GSU2<double>, GSU2<double>>(Uin, Uout, g, w_t, w_s, mu_start, flags, partials, st);
    else launch_plaq_g<GSU2<float>, GSU2<double>>(Uin, Uout, g, w_t, w_s, mu_start, flags, partials, st);
  } else {
    if (prec == 0) launch_plaq_g<GU1<double>, GU1<double>>(Uin, Uout, g, w_t, w_s, mu_start, flags, partials, st);
    else launch_plaq_g<GU1<float>, GU1<double>>(Uin, Uout, g, w_t, w_s, mu_start, flags, partials, st);
  }
  return nblk(g.V);
}

int sumsq_blocks(long n) {
  int b = nblk(n);
  return b < 2048 ? b : 2048;
}
void launch_sumsq(int prec, const void *P, long n, double *partials, hipStream_t st) {
  const int b = sumsq_blocks(n);
  if (prec == 0)
    hipLaunchKernelGGL((k_sumsq<double>), dim3(b), dim3(NT), 0, st, (const double *)P, (int)n, partials);
  else
    hipLaunchKernelGGL((k_sumsq<float>), dim3(b), dim3(NT), 0, st, (const float *)P, (int)n, partials);
}

void launch_scale(int prec, void *P, long n, double f, hipStream_t st) {
  if (prec == 0)
    hipLaunchKernelGGL((k_scale<double>), dim3(nblk(n)), dim3(NT), 0, st, (double *)P, (int)n, f);
  else
    hipLaunchKernelGGL((k_scale<float>), dim3(nblk(n)), dim3(NT), 0, st, (float *)P, (int)n, (float)f);
}

void launch_reduce(const double *partials, size_t row_stride, const RowList &rl, double *out,
                   hipStream_t st) {
  if (rl.n > 0)
    hipLaunchKernelGGL(k_reduce, dim3(rl.n), dim3(NTR), 0, st, partials, row_stride, rl, out);
}

template <class GS>
static void launch_ed_g(const void *U, const Geo &g, int clover, int mu_start, double *partials,
                        int pstride, hipStream_t st) {
  using link = typename GS::link;
  dim3 grid(nblk(g.V));
  switch (g.nd) {
    case 2:
      hipLaunchKernelGGL((k_energy_density<GS, 2>), grid, dim3(NT), 0, st, (const link *)U, g, clover, mu_start, partials, pstride);
      break;
    case 3:
      hipLaunchKernelGGL((k_energy_density<GS, 3>), grid, dim3(NT), 0, st, (const link *)U, g, clover, mu_start, partials, pstride);
      break;
    default:
      hipLaunchKernelGGL((k_energy_density<GS, 4>), grid, dim3(NT), 0, st, (const link *)U, g, clover, mu_start, partials, pstride);
      break;
  }
}
void launch_energy_density(int group, int prec, const void *U, const Geo &g, int clover,
                           int mu_start, double *partials, int pstride, hipStream_t st) {
  if (group == 1) {
    if (prec == 0) launch_ed_g<GSU2<double>>(U, g, clover, mu_start, partials, pstride, st);
    else launch_ed_g<GSU2<float>>(U, g, clover, mu_start, partials, pstride, st);
  } else {
    if (prec == 0) launch_ed_g<GU1<double>>(U, g, clover, mu_start, partials, pstride, st);
    else launch_ed_g<GU1<float>>(U, g, clover, mu_start, partials, pstride, st);
  }
}

void launch_links_from_aos(int group, int prec, const double *src, int stride, int offset, void *dst,
                           int V, int nd, hipStream_t st) {
  dim3 grid(nblk((long)V * nd));
  if (group == 1) {
    if (prec == 0) hipLaunchKernelGGL((k_links_from_aos<double, 4>), grid, dim3(NT), 0, st, src, stride, offset, (double *)dst, V, nd);
    else hipLaunchKernelGGL((k_links_from_aos<float, 4>), grid, dim3(NT), 0, st, src, stride, offset, (float *)dst, V, nd);
  } else {
    if (prec == 0) hipLaunchKernelGGL((k_links_from_aos<double, 1>), grid, dim3(NT), 0, st, src, stride, offset, (double *)dst, V, nd);
    else hipLaunchKernelGGL((k_links_from_aos<float, 1>), grid, dim3(NT), 0, st, src, stride, offset, (float *)dst, V, nd);
  }
}
void launch_links_to_aos(int group, int prec, const void *src, double *dst, int stride, int offset,
                         int V, int nd, hipStream_t st) {
  dim3 grid(nblk((long)V * nd));
  if (group == 1) {
    if (prec == 0) hipLaunchKernelGGL((k_links_to_aos<double, 4>), grid, dim3(NT), 0, st, (const double *)src, dst, stride, offset, V, nd);
    else hipLaunchKernelGGL((k_links_to_aos<float, 4>), grid, dim3(NT), 0, st, (const float *)src, dst, stride, offset, V, nd);
  } else {
    if (prec == 0) hipLaunchKernelGGL((k_links_to_aos<double, 1>), grid, dim3(NT), 0, st, (const double *)src, dst, stride, offset, V, nd);
    else hipLaunchKernelGGL((k_links_to_aos<float, 1>), grid, dim3(NT), 0, st, (const float *)src, dst, stride, offset, V, nd);
  }
}
void launch_mom_from_aos(int prec, const double *src, void *dst, int V, int nd, int nm, hipStream_t st) {
  dim3 grid(nblk((long)V * nd));
  if (prec == 0) hipLaunchKernelGGL((k_mom_from_aos<double>), grid, dim3(NT), 0, st, src, (double *)dst, V, nd, nm);
  else hipLaunchKernelGGL((k_mom_from_aos<float>), grid, dim3(NT), 0, st, src, (float *)dst, V, nd, nm);
}
void launch_mom_to_aos(int prec, const void *src, double *dst, int V, int nd, int nm, hipStream_t st) {
  dim3 grid(nblk((long)V * nd));
  if (prec == 0) hipLaunchKernelGGL((k_mom_to_aos<double>), grid, dim3(NT), 0, st, (const double *)src, dst, V, nd, nm);
  else hipLaunchKernelGGL((k_mom_to_aos<float>), grid, dim3(NT), 0, st, (const float *)src, dst, V, nd, nm);
}
void launch_momenta_philox(int prec, void *P, int V, int nd, int nm, uint64_t seed, uint64_t traj,
                           hipStream_t st) {
  dim3 grid(nblk((long)V * nd));
  if (prec == 0) hipLaunchKernelGGL((k_momenta_philox<double>), grid, dim3(NT), 0, st, (double *)P, V, nd, nm, seed, traj);
  else hipLaunchKernelGGL((k_momenta_philox<float>), grid, dim3(NT), 0, st, (float *)P, V, nd, nm, seed, traj);
}

}  // namespace su2k
