// su2_batch.hip -- many independent lattices of uneven sizes (U(1) / SU(2),
// d = 2, 3, 4, own beta, extents, integrator step count) advanced together.
//
// All lattices live in three link arenas and one momentum arena.  A lattice is
// cut into tiles of `tile` sites; one MD tick (one "P then U" pair of every
// lattice's integrator, or its final momentum step) is three launches of a
// grid-stride kernel: over the tiles of the SU(2) lattices in d = 4, in d = 3
// (the specialised step body of su2_md_fast.hpp), and over the rest, whatever the
// lattice's group and dimension: a tile selects its (group, d) code path
// uniformly for the block.
// Lattice l's k-th link update reads arena r(k) and writes w(k) with
//   r(0) = a0, r(k>0) = (k odd ? a1 : a2), w(k) = (k even ? a1 : a2),
//   a0 = cur_l, a1 = cur_l+1 mod 3, a2 = cur_l+2 mod 3,
// so the start configuration a0 survives the trajectory and a rejected
// trajectory costs nothing (cur_l is kept) while an accepted one moves cur_l to
// the last written arena.
#include "su2_batch.hpp"
#include "su2_md_fast.hpp"

using namespace su2dev;

namespace su2k {

// integrator pair k of lattice l: (P coefficient, U coefficient) in units of dtau
__device__ __forceinline__ void pair_coef(int kind, int k, double &cp, double &cu) {
  if (kind == 6) {  // omf2 (integrator.hh:83-115): P(l) [U(1/2) P(1-2l) U(1/2) P(2l)]... P(l)
    const double lam = 0.1938;
    cp = (k == 0) ? lam : ((k & 1) ? 1. - 2. * lam : 2. * lam);
    cu = 0.5;
  } else {  // leapfrog (integrator.hh:26-50): P(1/2) U(1) [P(1) U(1)]... P(1/2)
    cp = (k == 0) ? 0.5 : 1.0;
    cu = 1.0;
  }
}
__device__ __forceinline__ double final_coef(int kind) { return kind == 6 ? 0.1938 : 0.5; }

// SU(2) in d = 3, 4: the specialised step body (mu selects one of ND compiled
// bodies; it is uniform over the tile), neighbour offsets from the lattice's
// own geometry
template <typename T, int ND, bool WRITE_U, bool ZS>
__device__ __forceinline__ void spec_step(const BLat &L, char *const *arena, T *Parena, int rd, int wr,
                                          const FusedCoef &c, int s, int mu, double &k0, double &a,
                                          double &k1) {
  const quat<T> *Uin = (const quat<T> *)(arena[rd] + L.uoff);
  quat<T> *Uout = WRITE_U ? (quat<T> *)(arena[wr] + L.uoff) : nullptr;
  T *P = Parena + L.poff;
  FastGeo g{};
  g.V = L.geo.V;
  g.Lz = L.geo.L[ND - 1];
  FastSite fs{};
  fs.s = s;
  fs.mu = mu;
  neighbour_offsets<ND>(L.geo, s, fs.up, fs.dn);
  T pi[3], pn[3];
  double mar;
  switch (mu) {
    case 0: md_su2_core<T, 0, WRITE_U, ZS, false, ND>(Uin, Uout, P, g, c, fs, pi, pn, mar); break;
    case 1: md_su2_core<T, 1, WRITE_U, ZS, false, ND>(Uin, Uout, P, g, c, fs, pi, pn, mar); break;
    case 2: md_su2_core<T, 2, WRITE_U, ZS, false, ND>(Uin, Uout, P, g, c, fs, pi, pn, mar); break;
    default:
      if constexpr (ND == 4) md_su2_core<T, 3, WRITE_U, ZS, false, ND>(Uin, Uout, P, g, c, fs, pi, pn, mar);
      break;
  }
  k0 += sq3(pi[0], pi[1], pi[2]);
  a += 2.0 * mar;
  k1 += sq3(pn[0], pn[1], pn[2]);
}

template <typename T, class G, int ND, bool WRITE_U>
__device__ __forceinline__ void tile_step(const BLat &L, char *const *arena, T *Parena, int rd, int wr,
                                          const FusedCoef &c, int s, int mu, double &k0, double &a,
                                          double &k1) {
  using link = typename G::link;
  const link *Uin = (const link *)(arena[rd] + L.uoff);
  link *Uout = WRITE_U ? (link *)(arena[wr] + L.uoff) : nullptr;
  if (L.zs) md_link<G, ND, WRITE_U, true>(Uin, Uout, Parena + L.poff, L.geo, c, s, mu, k0, a, k1);
  else md_link<G, ND, WRITE_U, false>(Uin, Uout, Parena + L.poff, L.geo, c, s, mu, k0, a, k1);
}

// SPEC = 4, 3: the tile belongs to a specialised SU(2) lattice in that
// dimension (first and second part of the tile list); 0: any other lattice
template <typename T, bool WRITE_U, int SPEC>
__device__ __forceinline__ void dispatch_step(const BLat &L, char *const *arena, T *Parena, int rd,
                                              int wr, const FusedCoef &c, int s, int mu,
                                              double &k0, double &a, double &k1) {
  if constexpr (SPEC != 0) {
    if (L.zs) spec_step<T, SPEC, WRITE_U, true>(L, arena, Parena, rd, wr, c, s, mu, k0, a, k1);
    else spec_step<T, SPEC, WRITE_U, false>(L, arena, Parena, rd, wr, c, s, mu, k0, a, k1);
    return;
  }
  const int code = L.group * 8 + L.geo.nd;
  switch (code) {
    case 0 * 8 + 2: tile_step<T, GU1<T>, 2, WRITE_U>(L, arena, Parena, rd, wr, c, s, mu, k0, a, k1); break;
    case 0 * 8 + 3: tile_step<T, GU1<T>, 3, WRITE_U>(L, arena, Parena, rd, wr, c, s, mu, k0, a, k1); break;
    case 0 * 8 + 4: tile_step<T, GU1<T>, 4, WRITE_U>(L, arena, Parena, rd, wr, c, s, mu, k0, a, k1); break;
    case 1 * 8 + 2: tile_step<T, GSU2<T>, 2, WRITE_U>(L, arena, Parena, rd, wr, c, s, mu, k0, a, k1); break;
    case 1 * 8 + 3: tile_step<T, GSU2<T>, 3, WRITE_U>(L, arena, Parena, rd, wr, c, s, mu, k0, a, k1); break;
    default: tile_step<T, GSU2<T>, 4, WRITE_U>(L, arena, Parena, rd, wr, c, s, mu, k0, a, k1); break;
  }
}

// One MD tick of every lattice.  partials rows (ntiles each): 0 |P_in|^2 and
// 1 ReTr(U K) at a lattice's first pair, 2 |P_out|^2 at its final momentum step.
// Tiles [t0, t0 + count) of the list; partial rows are ntiles long.
template <typename T, int SPEC>
__global__ __launch_bounds__(256) void k_batch_tick(const BLat *__restrict__ lats,
                                                    const int2 *__restrict__ tiles, int t0,
                                                    int count, int ntiles, BArenas ar, int tick,
                                                    int flags, double *__restrict__ partials) {
  const int nt = blockDim.x;
  for (int t = blockIdx.x; t < count; t += gridDim.x) {
    const int lt = t0 + xcd_block(t, count);
    const int2 tl = tiles[lt];
    const BLat &L = lats[tl.x];
    if (tick > L.npairs) continue;  // lattice finished (uniform over the block)
    const int s = (tl.y >> 2) + threadIdx.x, mu = tl.y & 3;
    const int a0 = L.cur, a1 = (L.cur + 1) % 3, a2 = (L.cur + 2) % 3;
    FusedCoef c{};
    c.coef = L.coef;
    c.w_t = 1.0;
    c.w_s = 1.0;
    c.flags = flags;
    double k0 = 0.0, act = 0.0, k1 = 0.0;
    const bool final_p = (tick == L.npairs);
    int rd, wr;
    if (!final_p) {
      double cp, cu;
      pair_coef(L.kind, tick, cp, cu);
      c.ep = cp * L.dtau;
      c.eu = cu * L.dtau;
      rd = (tick == 0) ? a0 : ((tick & 1) ? a1 : a2);
      wr = (tick & 1) ? a2 : a1;
    } else {
      c.ep = final_coef(L.kind) * L.dtau;
      rd = L.npairs == 0 ? a0 : (((L.npairs - 1) & 1) ? a2 : a1);
      wr = rd;
    }
    if (s < L.geo.V) {  // always true on z-shuffle lattices (V % tile == 0)
      if (!final_p) dispatch_step<T, true, SPEC>(L, ar.U, (T *)ar.P, rd, wr, c, s, mu, k0, act, k1);
      else dispatch_step<T, false, SPEC>(L, ar.U, (T *)ar.P, rd, wr, c, s, mu, k0, act, k1);
    }
    if (tick == 0 || final_p) {
      // uniform over the block: every thread of the tile takes this branch
      double r0 = 0.0, r1 = 0.0, r2 = 0.0;
      {
        double v = k0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        double w = act;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) w += __shfl_down(w, o, 64);
        double x = k1;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        __shared__ double l3[3][4];
        __syncthreads();
        if (lane == 0) {
          l3[0][wv] = v;
          l3[1][wv] = w;
          l3[2][wv] = x;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          for (int i = 0; i < nt / 64; ++i) {
            r0 += l3[0][i];
            r1 += l3[1][i];
            r2 += l3[2][i];
          }
          if (tick == 0) {
            partials[0 * ntiles + lt] = r0;
            partials[1 * ntiles + lt] = r1;
          }
          if (final_p) partials[2 * ntiles + lt] = r2;
        }
        __syncthreads();
      }
    }
  }
}

// FP64 sum of ReTr P per tile (plaquettes of flat-gauge_energy.hpp:31-94) on
// each lattice's start (use_final = 0) or final arena, into partial row `row`.
template <typename T, class GS, int ND>
__device__ __forceinline__ double plaq_site(const BLat &L, const char *base, int s) {
  using link = typename GS::link;
  using GD = typename std::conditional<GS::NCOMP == 4, GSU2<double>, GU1<double>>::type;
  const link *U = (const link *)(base + L.uoff);
  const int V = L.geo.V;
  int up[ND], dn[ND];
  neighbour_offsets<ND>(L.geo, s, up, dn);
  double sum = 0.0;
#pragma unroll
  for (int mu = 0; mu < ND - 1; ++mu)
#pragma unroll
    for (int nu = mu + 1; nu < ND; ++nu)
      sum += GD::plaq(to_d(U[mu * V + s]), to_d(U[nu * V + s + up[mu]]), to_d(U[mu * V + s + up[nu]]),
                      to_d(U[nu * V + s]));
  return sum;
}

template <typename T>
__global__ __launch_bounds__(256) void k_batch_plaq(const BLat *__restrict__ lats,
                                                    const int2 *__restrict__ tiles, int ntiles,
                                                    BArenas ar, int use_final, int row, int rowlen,
                                                    double *__restrict__ partials) {
  __shared__ double l1[4];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int lt = xcd_block(t, ntiles);
    const int2 tl = tiles[lt];
    const BLat &L = lats[tl.x];
    const int s = tl.y + threadIdx.x;
    const int arena = use_final ? L.fin : L.cur;
    double v = 0.0;
    if (s < L.geo.V) {
      const char *base = ar.U[arena];
      switch (L.group * 8 + L.geo.nd) {
        case 2: v = plaq_site<T, GU1<T>, 2>(L, base, s); break;
        case 3: v = plaq_site<T, GU1<T>, 3>(L, base, s); break;
        case 4: v = plaq_site<T, GU1<T>, 4>(L, base, s); break;
        case 10: v = plaq_site<T, GSU2<T>, 2>(L, base, s); break;
        case 11: v = plaq_site<T, GSU2<T>, 3>(L, base, s); break;
        default: v = plaq_site<T, GSU2<T>, 4>(L, base, s); break;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) l1[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = 0.0;
      for (int i = 0; i < (int)blockDim.x / 64; ++i) r += l1[i];
      partials[(size_t)row * rowlen + lt] = r;
    }
  }
}

// in-place restoreSU of every SU(2) lattice's final arena (integrator.hh:48,111)
template <typename T>
__global__ __launch_bounds__(256) void k_batch_restore(const BLat *__restrict__ lats,
                                                       const int2 *__restrict__ tiles, int ntiles,
                                                       BArenas ar) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int2 tl = tiles[t];
    const BLat &L = lats[tl.x];
    if (L.group != 1) continue;
    const int s = tl.y + threadIdx.x;
    if (s >= L.geo.V) continue;
    quat<T> *U = (quat<T> *)(ar.U[L.fin] + L.uoff);
    for (int mu = 0; mu < L.geo.nd; ++mu) U[mu * L.geo.V + s] = qrestore(U[mu * L.geo.V + s]);
  }
}

// Gaussian momenta of every lattice from Philox keyed by the lattice's seed and
// the trajectory counter (device scalar, so a captured graph can be replayed)
template <typename T>
__global__ __launch_bounds__(256) void k_batch_momenta(const BLat *__restrict__ lats,
                                                       const int2 *__restrict__ tiles, int ntiles,
                                                       BArenas ar, const uint64_t *__restrict__ traj) {
  const uint64_t tr = *traj;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int2 tl = tiles[t];
    const BLat &L = lats[tl.x];
    const int s = tl.y + threadIdx.x;
    if (s >= L.geo.V) continue;
    const int nm = L.group == 1 ? 3 : 1;
    T *P = (T *)ar.P + L.poff;
    for (int mu = 0; mu < L.geo.nd; ++mu) {
      T z[4];
      philox_normals4<T>((uint32_t)(s * L.geo.nd + mu), (uint32_t)(tr & 0xffffffffu),
                         (uint32_t)(tr >> 32), L.seed, z);
      for (int k = 0; k < nm; ++k) P[((size_t)mu * nm + k) * L.geo.V + s] = z[k];
    }
  }
}

// per-lattice reduction of the 5 tile rows (rows 0-2 over link tiles, 3-4 over
// site tiles); out[l*10 + 2*r + {0,1}] = hi, lo
__global__ __launch_bounds__(256) void k_batch_reduce(const BLat *__restrict__ lats,
                                                      const double *__restrict__ partials,
                                                      int rowlen, double *__restrict__ out) {
  __shared__ double hs[256], ls[256];
  const BLat &L = lats[blockIdx.x];
  for (int r = 0; r < 5; ++r) {
    const int t0 = r < 3 ? L.ltile0 : L.tile0, nt = r < 3 ? L.nltiles : L.ntiles;
    double hi = 0.0, lo = 0.0;
    for (int i = threadIdx.x; i < nt; i += blockDim.x) {
      double e;
      two_sum(hi, partials[(size_t)r * rowlen + t0 + i], hi, e);
      lo += e;
    }
    hs[threadIdx.x] = hi;
    ls[threadIdx.x] = lo;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        double e, h;
        two_sum(hs[threadIdx.x], hs[threadIdx.x + w], h, e);
        hs[threadIdx.x] = h;
        ls[threadIdx.x] = ls[threadIdx.x] + ls[threadIdx.x + w] + e;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      out[blockIdx.x * 10 + 2 * r] = hs[0];
      out[blockIdx.x * 10 + 2 * r + 1] = ls[0];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- launchers ----
void launch_batch_tick(int prec, const BLat *lats, const int2 *tiles, int ntiles, int n4, int n3,
                       const BArenas &ar, int tick, int flags, double *partials, int grid,
                       int block, hipStream_t st, int only, const int *active) {
  const int parts[3][2] = {{0, n4}, {n4, n3}, {n4 + n3, ntiles - n4 - n3}};
  for (int k = 0; k < 3; ++k) {
    const int t0 = parts[k][0];
    const int n = active ? (active[k] < parts[k][1] ? active[k] : parts[k][1]) : parts[k][1];
    if (n <= 0 || (only >= 0 && k != only)) continue;
    const dim3 g(grid < n ? grid : n), bl(block);
#define SU2_TICK(T, S) \
  hipLaunchKernelGGL((k_batch_tick<T, S>), g, bl, 0, st, lats, tiles, t0, n, ntiles, ar, tick, flags, partials)
#define SU2_TICK3(T)            \
  if (k == 0) SU2_TICK(T, 4);   \
  else if (k == 1) SU2_TICK(T, 3); \
  else SU2_TICK(T, 0)
    if (prec == 0) {
      SU2_TICK3(double);
    } else {
      SU2_TICK3(float);
    }
#undef SU2_TICK3
#undef SU2_TICK
  }
}
void launch_batch_plaq(int prec, const BLat *lats, const int2 *tiles, int ntiles, const BArenas &ar,
                       int use_final, int row, int rowlen, double *partials, int grid, int block,
                       hipStream_t st) {
  if (prec == 0)
    hipLaunchKernelGGL((k_batch_plaq<double>), dim3(grid), dim3(block), 0, st, lats, tiles, ntiles, ar, use_final, row, rowlen, partials);
  else
    hipLaunchKernelGGL((k_batch_plaq<float>), dim3(grid), dim3(block), 0, st, lats, tiles, ntiles, ar, use_final, row, rowlen, partials);
}
void launch_batch_restore(int prec, const BLat *lats, const int2 *tiles, int ntiles, const BArenas &ar,
                          int grid, int block, hipStream_t st) {
  if (prec == 0)
    hipLaunchKernelGGL((k_batch_restore<double>), dim3(grid), dim3(block), 0, st, lats, tiles, ntiles, ar);
  else
    hipLaunchKernelGGL((k_batch_restore<float>), dim3(grid), dim3(block), 0, st, lats, tiles, ntiles, ar);
}
void launch_batch_momenta(int prec, const BLat *lats, const int2 *tiles, int ntiles, const BArenas &ar,
                          const uint64_t *traj, int grid, int block, hipStream_t st) {
  if (prec == 0)
    hipLaunchKernelGGL((k_batch_momenta<double>), dim3(grid), dim3(block), 0, st, lats, tiles, ntiles, ar, traj);
  else
    hipLaunchKernelGGL((k_batch_momenta<float>), dim3(grid), dim3(block), 0, st, lats, tiles, ntiles, ar, traj);
}
void launch_batch_reduce(const BLat *lats, int nlat, const double *partials, int rowlen, double *out,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_batch_reduce, dim3(nlat), dim3(256), 0, st, lats, partials, rowlen, out);
}

}  // namespace su2k
