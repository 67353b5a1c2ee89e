// su2_launch.hpp -- launch interface between the C-ABI host code and the
// gfx950 kernels, plus the counter-based Philox4x32-10 generator shared by host
// and device (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11: 10 rounds, multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "su2_device.hpp"

namespace su2k {

using su2dev::Geo;

// flags of the fused MD kernel and of the sum kernels
enum : int {
  F_PZERO = 1,         // momentum input treated as zero (gradient-flow stage 0)
  F_ACCUM = 2,         // write block partials of |P_in|^2, ReTr(U K), |P_out|^2
  F_RESTORE = 4,       // restoreSU on the updated link / on load
  F_COMPENSATED = 8,   // TwoSum-compensated staple accumulation
  F_ROUND = 16,        // lp_* integrators: round force and link (update_*.hh round variants)
  F_ANISO = 32,        // anisotropic staple / plaquette weights
  F_R32L = 64,         // FP32 storage emulation: updated links rounded to FP32
  F_R32P = 128,        // FP32 storage emulation: updated momenta rounded to FP32
};

struct FusedCoef {
  double ep;       // P <- P - ep * F
  double eu;       // U' = exp(eu P) U
  double coef;     // fac * beta / Nc
  double w_t;      // weight of staples/plaquettes touching direction 0 (1/xi)
  double w_s;      // weight of purely spatial ones (xi)
  double round_n;  // rounding multiplier 10^exponent (F_ROUND)
  int flags;
  // the first step drawing its momenta (FusedLaunch::gen): the Philox stream
  // of launch_momenta_philox with this seed and trajectory number
  uint64_t gseed = 0, gtraj = 0;
};

struct FusedLaunch {
  const void *Uin;
  void *Uout;  // nullptr: momentum-only step
  void *P;
  void *Pout = nullptr;  // non-null: new momenta go here, P is only read
  Geo geo;
  FusedCoef c;
  double *partials;  // 3 rows of fused_blocks(geo) doubles (F_ACCUM) ...
  int pstride;       // ... pstride doubles apart
  // FP64 SU(2) half-field layout (su2_md_fast.hpp QHalf) of the links read /
  // written; only where fused_half_ok says the step supports it
  int in_half = 0, out_half = 0;
  // 1: P holds no momenta yet; they are the Philox momenta of c.gseed /
  // c.gtraj (launch_momenta_philox), drawn inside the step where it walks t,
  // else by launch_momenta_philox before it
  int gen = 0;
};

// false if a half-field layout was asked for where the step does not support it
bool launch_fused(int group, int prec, const FusedLaunch &L, hipStream_t st);
// whether a fused launch of this lattice with these flags runs the specialised
// step that reads and writes the half-field layout (FP64 SU(2) d = 4,
// isotropic, specialised launch layout)
bool fused_half_ok(int group, int prec, const Geo &g, int flags);
// half-field links (n links) -> record layout, out-of-place
void launch_half_to_record(const void *src, void *dst, long n, hipStream_t st);
int fused_blocks(const Geo &g);  // blocks (= partial-sum entries) of one fused launch
int site_blocks(const Geo &g);   // blocks of the one-thread-per-site sum kernels
int fused_mode(const Geo &g);    // 0 thread per site, 1 thread per link, 2 link + z shuffles
void launch_update_gauge(int group, int prec, void *U, const void *P, int V, int nd, double eps,
                         double round_n, int flags, hipStream_t st);
void launch_restore(int group, int prec, void *U, int n, hipStream_t st);
// every element of an FP64 array rounded to the nearest FP32 value, in place
void launch_round_f32(double *x, size_t n, hipStream_t st);
// returns the number of block partials written (one row, at most fused_blocks(g));
// in_half: Uin in the half-field layout (where fused_half_ok holds), Uout always
// in the record layout
int launch_plaq(int group, int prec, const void *Uin, void *Uout, const Geo &g, double w_t,
                double w_s, int mu_start, int flags, double *partials, hipStream_t st,
                int in_half = 0);
int sumsq_blocks(long n);
void launch_sumsq(int prec, const void *P, long n, double *partials, hipStream_t st);
void launch_scale(int prec, void *P, long n, double f, hipStream_t st);
// rows to reduce in one launch: row k of the partials (stride row_stride
// doubles) holds nb[k] block partials; its sum goes to out[2 row[k] + {0,1}]
struct RowList {
  static constexpr int MAX = 16;
  int n = 0;
  int row[MAX];
  int nb[MAX];
};
void launch_reduce(const double *partials, size_t row_stride, const RowList &rl, double *out,
                   hipStream_t st);
void launch_energy_density(int group, int prec, const void *U, const Geo &g, int clover,
                           int mu_start, double *partials, int pstride, hipStream_t st);
void launch_links_from_aos(int group, int prec, const double *src, int stride, int offset,
                           void *dst, int V, int nd, hipStream_t st);
void launch_links_to_aos(int group, int prec, const void *src, double *dst, int stride,
                         int offset, int V, int nd, hipStream_t st);
void launch_mom_from_aos(int prec, const double *src, void *dst, int V, int nd, int nm,
                         hipStream_t st);
void launch_mom_to_aos(int prec, const void *src, double *dst, int V, int nd, int nm,
                       hipStream_t st);
void launch_momenta_philox(int prec, void *P, int V, int nd, int nm, uint64_t seed,
                           uint64_t traj, hipStream_t st);

void launch_copy16(const void *src, void *dst, long n16, hipStream_t st);

// gradient-flow adjoint (su2_flow.hip)
void launch_flow_stage_adj(int prec, const void *Zs, void *lamW, void *lamZ, int V, int nd,
                           double c, int zero_lamZ, hipStream_t st);
// Metropolis sub-sweep: the links U_mu(x), x of the given parity (list ==
// nullptr, every extent even) or the nlist sites of a device site list, n_hit
// hits each; one accepted-hit count per block to partials[0 .. sweep_blocks)
int sweep_blocks(const Geo &g, int nlist);  // nlist < 0: the parity form
void launch_sweep(int group, int prec, void *U, const Geo &g, int mu, int parity, const int *list,
                  int nlist, uint64_t seed, uint64_t sweep, double delta, int n_hit,
                  double beta_nc, double w_t, double w_s, int aniso, double *partials,
                  hipStream_t st);
// lamW += b (dD/dW)^T lamZ; R: scratch of 36 V reals (the b and c role
// gradients of the 6 planes).  gather = false leaves the R terms to a later
// launch_flow_gather or launch_flow_gather_stage_adj.
void launch_flow_vjp(int prec, const void *U, const void *lamZ, void *lamW, void *R, const Geo &g,
                     double b, double w_t, double w_s, bool gather, hipStream_t st);
void launch_flow_gather(int prec, const void *R, void *lamW, const Geo &g, double b,
                        hipStream_t st);
// the gather of one stage, then the pointwise adjoint of the stage before it
// (Z' = Zs, coefficient c) on every link
void launch_flow_gather_stage_adj(int prec, const void *R, void *lamW, void *lamZ, const void *Zs,
                                  const Geo &g, double b, double c, int zero_lamZ,
                                  hipStream_t st);
void launch_clover_grad(int prec, const void *U, void *Gam, void *lamW, const Geo &g,
                        double scale, hipStream_t st);

}  // namespace su2k

// ----------------------------------------------------------------- Philox ---
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c[0];
    const uint64_t p1 = (uint64_t)M1 * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

// 53-bit uniform in (0, 1) from two 32-bit words
__host__ __device__ inline double philox_u01(uint32_t hi, uint32_t lo) {
  const uint64_t m = (((uint64_t)hi << 32) | lo) >> 11;
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Four N(0,1) draws for counter (i, 0, t0, t1) and key = seed: one
// Philox4x32-10 block, Box-Muller on the word pairs (w0, w1) and (w2, w3) with
// radius uniform (w0 + 1) 2^-32 in (0, 1] and angle 2 pi w1 2^-32 (sincospi
// needs no range reduction).  T = double: exact uniforms, FP64 log and
// sincospi.  T = float: the same uniforms rounded to FP32 and FP32 log and
// sincospi, so both forms draw the same momenta to FP32 accuracy (the radius
// differs by at most ~3e-4 in the 2^-24-probability corner u -> 1, r -> 0).
template <typename T>
__device__ inline void philox_normals4(uint32_t i, uint32_t t0, uint32_t t1, uint64_t seed,
                                       T z[4]) {
  uint32_t c[4] = {i, 0u, t0, t1};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const T u1 = ((T)c[2 * h] + T(1)) * T(0x1p-32);
    const T u2 = (T)c[2 * h + 1] * T(0x1p-32);
    const T r = sqrt(T(-2) * log(u1));
    T sn, cs;
    if constexpr (sizeof(T) == 4) sincospif(2.0f * u2, &sn, &cs);
    else sincospi(T(2) * u2, &sn, &cs);
    z[2 * h] = r * cs;
    z[2 * h + 1] = r * sn;
  }
}

// the accept/reject uniform of trajectory traj in the device-momenta mode
__host__ __device__ inline double philox_accept_uniform(uint64_t seed, uint64_t traj) {
  uint32_t c[4] = {0xffffffffu, 0xffffffffu, (uint32_t)traj, (uint32_t)(traj >> 32)};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return philox_u01(c[0], c[1]);
}
