// su2_batch.hpp -- descriptors and launchers of the batched-ensemble kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "su2_launch.hpp"

namespace su2k {

// one lattice of the batch (device copy)
struct BLat {
  Geo geo;
  int group;      // 0 U(1), 1 SU(2)
  int kind;       // integrator id: 0 leapfrog, 6 omf2
  int npairs;     // number of "P then U" pairs; tick npairs is the final momentum step
  int cur;        // arena holding the configuration at the start of the trajectory
  int fin;        // arena holding the last link update
  int tile0, ntiles;    // site tiles (sums, momenta, restore)
  int ltile0, nltiles;  // link tiles (MD ticks): one per (site tile, mu)
  int zs;               // z-shuffle path allowed (64 % L_fastest == 0, V % tile == 0)
  int spec;             // 4, 3: SU(2) in that dimension through the specialised step body (md_su2_core); 0: generic
  long uoff;      // byte offset of the lattice's links in every link arena
  long poff;      // element offset of its momenta in the momentum arena
  double coef;    // beta / Nc
  double dtau;    // tau / nsteps
  uint64_t seed;  // Philox key of its momenta
};

struct BArenas {
  char *U[3];
  void *P;
};

// ltiles: (lattice, site_start * 4 + mu): those of the specialised SU(2)
// lattices in d = 4 ([0, n4)), then in d = 3 ([n4, n4 + n3)), then the rest;
// partial rows are nltiles long.  One launch per part (at most `grid` blocks
// each), so that each part has its own register budget; only >= 0 launches
// that part alone (0: d = 4, 1: d = 3, 2: the rest).  active (optional): the
// number of leading tiles of each part that belong to lattices still running
// at this tick (each part's lattices are ordered by step count, longest
// first); the launch covers only those.
void launch_batch_tick(int prec, const BLat *lats, const int2 *ltiles, int nltiles, int n4, int n3,
                       const BArenas &ar, int tick, int flags, double *partials, int grid,
                       int block, hipStream_t st, int only = -1, const int *active = nullptr);
void launch_batch_plaq(int prec, const BLat *lats, const int2 *tiles, int ntiles, const BArenas &ar,
                       int use_final, int row, int rowlen, double *partials, int grid, int block,
                       hipStream_t st);
void launch_batch_restore(int prec, const BLat *lats, const int2 *tiles, int ntiles,
                          const BArenas &ar, int grid, int block, hipStream_t st);
void launch_batch_momenta(int prec, const BLat *lats, const int2 *tiles, int ntiles,
                          const BArenas &ar, const uint64_t *traj, int grid, int block,
                          hipStream_t st);
void launch_batch_reduce(const BLat *lats, int nlat, const double *partials, int rowlen,
                         double *out, hipStream_t st);

}  // namespace su2k
