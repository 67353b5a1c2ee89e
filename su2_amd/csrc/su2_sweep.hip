// su2_sweep.hip -- N_hit Metropolis updates of every link
// (flat_spacetime::sweep, flat-sweep.hh:64-164) for SU(2) and U(1), d = 2, 3, 4.
//
// Order.  The reference walks the even t-slices in parallel and, inside a
// slice, the sites and directions one after another; a GPU needs independent
// links.  The staple of U_mu(x) holds, of the links of direction mu, only
// U_mu(x +- nu) with nu != mu, so for a fixed mu the links of one class of a
// proper colouring of the sites under those neighbour steps are independent.
// Every extent in use even: the two classes of the site parity (sum of the
// coordinates mod 2), 2 nd launches per sweep over V/2 sites each, sites
// computed from the launch index.  Any odd extent: three classes
// c(x) = sum_{nu != mu} h(x_nu) mod 3 with h(k) = k + min(k, a),
// a = (3 - L_nu mod 3) mod 3 (every step along nu changes h by 1 or 2 mod 3,
// the periodic wrap by 1), 3 nd launches over host-built site lists.
//
// One hit (flat-sweep.hh:95-110): R = random_element(delta)
// (random_element.hh:9-25, :27-34), dS = beta/Nc (ReTr(U K) - ReTr(U R K)),
// accept if dS < 0 or v < exp(-dS), then U <- restoreSU(U R); K (the staples,
// get_staples.hh:26-80 with the anisotropy weights) is formed once per link.
//
// Random numbers.  Hit n of link (s, mu) in sweep number w takes one
// Philox4x32-10 block with counter (s nd + mu, n, w_lo, w_hi) and key = seed:
// its words give alpha, u, theta and the acceptance uniform, each as
// (word + 1/2) 2^-32.  The chain therefore depends on (seed, sweep number) only,
// not on the launch shape; it is not the reference's per-thread mt19937 stream.
#include "su2_device.hpp"
#include "su2_launch.hpp"

using namespace su2dev;

namespace su2k {

constexpr int NTS = 256;

__device__ __forceinline__ double u01_32(uint32_t w) {
  return ((double)w + 0.5) * (1.0 / 4294967296.0);
}

// ReTr(u K) in FP64 (su2.hh:56-58, 72-74; u1.hh:61-67)
__device__ __forceinline__ double retr_mul(const quat<double> &u, const quat<double> &K) {
  return 2.0 * (u.ar * K.ar - u.ai * K.ai - u.br * K.br - u.bi * K.bi);
}
__device__ __forceinline__ double retr_mul_u1(double th, double kre, double kim) {
  double s, c;
  sincos(th, &s, &c);
  return c * kre - s * kim;
}

// list == nullptr: the sites of parity `parity` (every extent even);
// otherwise the nlist sites list[0 .. nlist) of one colour class
template <class G, int ND>
__global__ __launch_bounds__(NTS) void k_sweep(typename G::link *__restrict__ U, Geo geo, int mu,
                                               int parity, const int *__restrict__ list,
                                               int nlist, uint64_t seed, uint64_t sweep,
                                               double delta, int n_hit, double beta_nc,
                                               double w_t, double w_s, int aniso,
                                               double *__restrict__ partials) {
  using T = typename G::real;
  using link = typename G::link;
  using acc = typename G::acc;
  __shared__ double lds[NTS / 64];
  const int V = geo.V;
  const int j = blockIdx.x * NTS + threadIdx.x;
  double accepted = 0.0;
  if (list ? j < nlist : j < V / 2) {
    int s;
    if (list) {
      s = list[j];
    } else {
      // site j of this parity: row of the fastest direction, then every second site
      const int Lf = geo.L[ND - 1];  // fastest direction in use (even)
      const int half = Lf / 2;
      const int row = j / half, k = j - row * half;
      int rp = 0, q = row;
#pragma unroll
      for (int d = ND - 2; d >= 0; --d) {
        const int Ld = geo.L[d];
        const int qn = q / Ld;
        rp += q - qn * Ld;
        q = qn;
      }
      s = row * Lf + 2 * k + ((parity + rp) & 1);
    }
    int up[ND], dn[ND];
    neighbour_offsets<ND>(geo, s, up, dn);
    // staples (get_staples.hh:26-80): all upper, then all lower
    acc K = G::zero();
#pragma unroll
    for (int nu = 0; nu < ND; ++nu) {
      if (nu == mu) continue;
      acc t = G::upper(U[nu * V + s + up[mu]], U[mu * V + s + up[nu]], U[nu * V + s]);
      if (aniso) t = G::scale(T((mu == 0 || nu == 0) ? w_t : w_s), t);
      K = G::add(K, t);
    }
#pragma unroll
    for (int nu = 0; nu < ND; ++nu) {
      if (nu == mu) continue;
      acc t = G::lower(U[nu * V + s + up[mu] + dn[nu]], U[mu * V + s + dn[nu]],
                       U[nu * V + s + dn[nu]]);
      if (aniso) t = G::scale(T((mu == 0 || nu == 0) ? w_t : w_s), t);
      K = G::add(K, t);
    }
    const uint32_t lid = (uint32_t)s * (uint32_t)ND + (uint32_t)mu;
    link u = U[mu * V + s];
    for (int n = 0; n < n_hit; ++n) {
      uint32_t c[4] = {lid, (uint32_t)n, (uint32_t)(sweep & 0xffffffffu), (uint32_t)(sweep >> 32)};
      philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
      const double v = u01_32(c[3]);
      if constexpr (G::NCOMP == 4) {
        // random_element (random_element.hh:9-25): (cos a + i sin a u, sin a r (sin th + i cos th))
        // alpha = 2 pi delta u0 and theta = 2 pi u2 through sincospi (no
        // argument reduction by pi)
        const double uu = -1.0 + 2.0 * u01_32(c[1]);
        const double r = sqrt(1.0 - uu * uu);
        double sa, ca, st, ct;
        sincospi(2.0 * delta * u01_32(c[0]), &sa, &ca);
        sincospi(2.0 * u01_32(c[2]), &st, &ct);
        const quat<double> R{ca, sa * uu, sa * r * st, sa * r * ct};
        const quat<double> ud{double(u.ar), double(u.ai), double(u.br), double(u.bi)};
        const quat<double> Kd{double(K.ar), double(K.ai), double(K.br), double(K.bi)};
        const quat<double> uR = qmul(ud, R);
        const double dS = beta_nc * (retr_mul(ud, Kd) - retr_mul(uR, Kd));
        if (dS < 0.0 || v < exp(-dS)) {
          // restoreSU (su2.hh:62-66) in FP64, one rounding to the storage precision
          const double rn =
              1.0 / sqrt(uR.ar * uR.ar + uR.ai * uR.ai + uR.br * uR.br + uR.bi * uR.bi);
          u = link{T(uR.ar * rn), T(uR.ai * rn), T(uR.br * rn), T(uR.bi * rn)};
          accepted += 1.0;
        }
      } else {
        // random_element (random_element.hh:27-34): phase in [-pi delta, pi delta)
        const double th = -M_PI * delta + (2.0 * M_PI * delta) * u01_32(c[0]);
        const double ud = double(u), un = ud + th;
        const double dS = beta_nc * (retr_mul_u1(ud, double(K.re), double(K.im)) -
                                     retr_mul_u1(un, double(K.re), double(K.im)));
        if (dS < 0.0 || v < exp(-dS)) {
          u = T(un);
          accepted += 1.0;
        }
      }
    }
    U[mu * V + s] = u;
  }
  const double r = block_sum<NTS>(accepted, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

int sweep_blocks(const Geo &g, int nlist) { return ((nlist >= 0 ? nlist : g.V / 2) + NTS - 1) / NTS; }

template <class G>
static void sweep_g(void *U, const Geo &g, int mu, int parity, const int *list, int nlist,
                    uint64_t seed, uint64_t sweep, double delta, int n_hit, double beta_nc,
                    double w_t, double w_s, int aniso, double *partials, hipStream_t st) {
  using link = typename G::link;
  const dim3 grid(sweep_blocks(g, list ? nlist : -1));
  if (grid.x == 0) return;
  switch (g.nd) {
    case 2:
      hipLaunchKernelGGL((k_sweep<G, 2>), grid, dim3(NTS), 0, st, (link *)U, g, mu, parity, list,
                         nlist, seed, sweep, delta, n_hit, beta_nc, w_t, w_s, aniso, partials);
      break;
    case 3:
      hipLaunchKernelGGL((k_sweep<G, 3>), grid, dim3(NTS), 0, st, (link *)U, g, mu, parity, list,
                         nlist, seed, sweep, delta, n_hit, beta_nc, w_t, w_s, aniso, partials);
      break;
    default:
      hipLaunchKernelGGL((k_sweep<G, 4>), grid, dim3(NTS), 0, st, (link *)U, g, mu, parity, list,
                         nlist, seed, sweep, delta, n_hit, beta_nc, w_t, w_s, aniso, partials);
      break;
  }
}

void launch_sweep(int group, int prec, void *U, const Geo &g, int mu, int parity, const int *list,
                  int nlist, uint64_t seed, uint64_t sweep, double delta, int n_hit,
                  double beta_nc, double w_t, double w_s, int aniso, double *partials,
                  hipStream_t st) {
#define SWEEP_ARGS U, g, mu, parity, list, nlist, seed, sweep, delta, n_hit, beta_nc, w_t, w_s, aniso, partials, st
  if (group == 1) {
    if (prec == 0) sweep_g<GSU2<double>>(SWEEP_ARGS);
    else sweep_g<GSU2<float>>(SWEEP_ARGS);
  } else {
    if (prec == 0) sweep_g<GU1<double>>(SWEEP_ARGS);
    else sweep_g<GU1<float>>(SWEEP_ARGS);
  }
#undef SWEEP_ARGS
}

}  // namespace su2k
