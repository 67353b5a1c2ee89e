// su2_batch.cpp -- host side of the batched ensemble (C ABI in include/su2_mi355x.h):
// arena layout, tile list, per-trajectory launch sequence (optionally replayed
// from a captured hipGraph), Metropolis per lattice, and the launch-shape
// autotuner with its JSON cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/su2_mi355x.h"
#include "su2_batch.hpp"

namespace su2k {
void launch_links_from_aos(int group, int prec, const double *src, int stride, int offset,
                           void *dst, int V, int nd, hipStream_t st);
void launch_links_to_aos(int group, int prec, const void *src, double *dst, int stride,
                         int offset, int V, int nd, hipStream_t st);
void launch_mom_from_aos(int prec, const double *src, void *dst, int V, int nd, int nm,
                         hipStream_t st);
}  // namespace su2k

extern "C" int su2gpu_internal_set_error(int code, const char *msg);

namespace {

int bfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return su2gpu_internal_set_error(code, buf);
}

#define BH(expr)                                                                          \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return bfail(SU2GPU_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
#define BK(expr)                    \
  do {                              \
    int r_ = (expr);                \
    if (r_ != SU2GPU_OK) return r_; \
  } while (0)

constexpr int NSUM = 5;  // rows: |P_in|^2, ReTr(UK), |P_out|^2, S_new plaquettes, S_old plaquettes

}  // namespace

struct su2gpu_batch {
  int prec = 1;
  std::vector<su2gpu_batch_member> mem;
  std::vector<su2k::BLat> lat;  // host copy of the descriptors
  size_t ubytes = 0, pelems = 0;
  su2k::BArenas ar{};
  su2k::BLat *dlat = nullptr;
  int2 *dtiles = nullptr, *dltiles = nullptr;
  int ntiles = 0, nltiles = 0, nspec4 = 0, nspec3 = 0, block = 256, grid = 0, grid_site = 0,
      maxpairs = 0;
  // active[3 t + k]: leading link tiles of part k whose lattices still run at tick t
  std::vector<int> active;
  double *partials = nullptr, *dsums = nullptr, *hsums = nullptr;
  uint64_t *dtraj = nullptr, *htraj = nullptr;
  double *staging = nullptr, *hstage = nullptr;
  size_t stage_cap = 0;
  hipStream_t stream = nullptr;
  // the d = 3 and generic parts' tick chains run on streams of their own,
  // forked from and joined back into `stream` (captured as parallel branches
  // of the trajectory's graph); each part reads and writes only its own
  // lattices, so the chains are independent
  hipStream_t pstream[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  bool concurrent = true;
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_devmom = -1;
  int ncu = 256;
};

namespace {

using B = su2gpu_batch;

int npairs_of(int kind, size_t n) { return kind == SU2GPU_OMF2 ? (int)(2 * n) : (int)n; }

// site tiles (tile size = block) and link tiles (one per site tile and mu, mu
// fastest, site start packed as 4*s + mu); lattices contiguous
// The link tiles of the SU(2) lattices in d = 4, then in d = 3 (the specialised
// step body) come first, so that each MD tick is three launches with their own
// register budgets: one per specialised dimension and the generic one.
int build_tiles(B *b) {
  std::vector<int2> tiles, ltiles;
  // SU2GPU_BATCH_SPEC = 0 keeps the SU(2) d = 3, 4 lattices on the generic step
  const char *e = std::getenv("SU2GPU_BATCH_SPEC");
  const bool spec_on = !(e && std::atoi(e) == 0);
  for (auto &L : b->lat) {
    L.zs = (64 % L.geo.L[L.geo.nd - 1] == 0) && (L.geo.V % b->block == 0);
    L.spec = (spec_on && L.group == 1 && L.geo.nd >= 3) ? L.geo.nd : 0;
  }
  for (size_t l = 0; l < b->lat.size(); ++l) {
    auto &L = b->lat[l];
    L.tile0 = (int)tiles.size();
    for (int s = 0; s < L.geo.V; s += b->block) tiles.push_back(int2{(int)l, s});
    L.ntiles = (int)tiles.size() - L.tile0;
  }
  // within each part the lattices with more steps first, so that the tiles
  // still running at tick t are a prefix of the part's tiles
  std::vector<size_t> order(b->lat.size());
  for (size_t l = 0; l < order.size(); ++l) order[l] = l;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t x, size_t y) { return b->lat[x].npairs > b->lat[y].npairs; });
  b->active.assign((size_t)3 * (b->maxpairs + 1), 0);
  for (int pass : {4, 3, 0}) {
    const int before = (int)ltiles.size();
    const int part = pass == 4 ? 0 : pass == 3 ? 1 : 2;
    for (size_t l : order) {
      auto &L = b->lat[l];
      if (L.spec != pass) continue;
      L.ltile0 = (int)ltiles.size();
      for (int s = 0; s < L.geo.V; s += b->block)
        for (int mu = 0; mu < L.geo.nd; ++mu) ltiles.push_back(int2{(int)l, 4 * s + mu});
      L.nltiles = (int)ltiles.size() - L.ltile0;
      for (int t = 0; t <= L.npairs && t <= b->maxpairs; ++t) b->active[3 * t + part] += L.nltiles;
    }
    if (pass == 4) b->nspec4 = (int)ltiles.size() - before;
    if (pass == 3) b->nspec3 = (int)ltiles.size() - before;
  }
  b->ntiles = (int)tiles.size();
  b->nltiles = (int)ltiles.size();
  for (int2 **p : {&b->dtiles, &b->dltiles}) {
    if (*p) BH(hipFree(*p));
    *p = nullptr;
  }
  BH(hipMalloc(&b->dtiles, tiles.size() * sizeof(int2)));
  BH(hipMemcpy(b->dtiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice));
  BH(hipMalloc(&b->dltiles, ltiles.size() * sizeof(int2)));
  BH(hipMemcpy(b->dltiles, ltiles.data(), ltiles.size() * sizeof(int2), hipMemcpyHostToDevice));
  if (b->partials) BH(hipFree(b->partials));
  b->partials = nullptr;
  BH(hipMalloc(&b->partials, (size_t)NSUM * b->nltiles * sizeof(double)));
  // site kernels: one pass over the site tiles; MD ticks: `grid` blocks (default
  // one pass over the link tiles), multiples of the 8 XCDs
  b->grid_site = ((b->ntiles + 7) / 8) * 8;
  if (b->grid <= 0 || b->grid > b->nltiles) b->grid = ((b->nltiles + 7) / 8) * 8;
  if (b->gexec) {
    hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  if (b->graph) {
    hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return SU2GPU_OK;
}

int upload_desc(B *b) {
  BH(hipMemcpyAsync(b->dlat, b->lat.data(), b->lat.size() * sizeof(su2k::BLat),
                    hipMemcpyHostToDevice, b->stream));
  return SU2GPU_OK;
}

int fin_arena(const su2k::BLat &L) {
  const int a1 = (L.cur + 1) % 3, a2 = (L.cur + 2) % 3;
  return ((L.npairs - 1) & 1) ? a2 : a1;
}

// HIP events around the launch groups of one eager trajectory
// (su2gpu_batch_profile): kinds of su2gpu_batch_profile's output
struct Rec {
  std::vector<int> kind;
  std::vector<hipEvent_t> ev;  // two per entry
  hipStream_t st = nullptr;
  void begin(int k) {
    hipEvent_t a, e;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&e);
    (void)hipEventRecord(a, st);
    kind.push_back(k);
    ev.push_back(a);
    ev.push_back(e);
  }
  void end() { (void)hipEventRecord(ev.back(), st); }
  ~Rec() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};
enum { RK_MOMENTA = 0, RK_PLAQ = 1, RK_TICK4 = 2, RK_TICK3 = 3, RK_TICK0 = 4, RK_RESTORE = 5,
       RK_REDUCE = 6, RK_N = 7 };

// the launches of one trajectory (captured into a graph when use_graph;
// rec: eager, each launch group between two events)
int enqueue(B *b, bool devmom, Rec *rec = nullptr) {
  const int n = (int)b->lat.size();
  auto beg = [&](int k) {
    if (rec) rec->begin(k);
  };
  auto fin = [&]() {
    if (rec) rec->end();
  };
  if (devmom) {
    beg(RK_MOMENTA);
    su2k::launch_batch_momenta(b->prec, b->dlat, b->dtiles, b->ntiles, b->ar, b->dtraj,
                               b->grid_site, b->block, b->stream);
    fin();
  }
  beg(RK_PLAQ);
  su2k::launch_batch_plaq(b->prec, b->dlat, b->dtiles, b->ntiles, b->ar, 0, 4, b->nltiles,
                          b->partials, b->grid_site, b->block, b->stream);
  fin();
  const int flags = (b->prec == 1) ? (su2k::F_COMPENSATED | su2k::F_RESTORE) : 0;
  if (!rec && b->concurrent) {
    // three independent tick chains: d = 4 on `stream`, d = 3 and the rest on
    // the part streams, joined before the final sums
    BH(hipEventRecord(b->ev_fork, b->stream));
    for (int k = 0; k < 2; ++k) BH(hipStreamWaitEvent(b->pstream[k], b->ev_fork, 0));
    for (int t = 0; t <= b->maxpairs; ++t)
      for (int part = 0; part < 3; ++part)
        su2k::launch_batch_tick(b->prec, b->dlat, b->dltiles, b->nltiles, b->nspec4, b->nspec3,
                                b->ar, t, flags, b->partials, b->grid, b->block,
                                part == 0 ? b->stream : b->pstream[part - 1], part,
                                &b->active[3 * t]);
    for (int k = 0; k < 2; ++k) {
      BH(hipEventRecord(b->ev_join[k], b->pstream[k]));
      BH(hipStreamWaitEvent(b->stream, b->ev_join[k], 0));
    }
  }
  for (int t = 0; t <= b->maxpairs && (rec || !b->concurrent); ++t) {
    if (!rec) {
      su2k::launch_batch_tick(b->prec, b->dlat, b->dltiles, b->nltiles, b->nspec4, b->nspec3,
                              b->ar, t, flags, b->partials, b->grid, b->block, b->stream, -1,
                              &b->active[3 * t]);
      continue;
    }
    for (int part = 0; part < 3; ++part) {
      beg(RK_TICK4 + part);
      su2k::launch_batch_tick(b->prec, b->dlat, b->dltiles, b->nltiles, b->nspec4, b->nspec3,
                              b->ar, t, flags, b->partials, b->grid, b->block, b->stream, part,
                              &b->active[3 * t]);
      fin();
    }
  }
  if (b->prec == 0) {
    beg(RK_RESTORE);
    su2k::launch_batch_restore(b->prec, b->dlat, b->dtiles, b->ntiles, b->ar, b->grid_site,
                               b->block, b->stream);
    fin();
  }
  beg(RK_PLAQ);
  su2k::launch_batch_plaq(b->prec, b->dlat, b->dtiles, b->ntiles, b->ar, 1, 3, b->nltiles,
                          b->partials, b->grid_site, b->block, b->stream);
  fin();
  beg(RK_REDUCE);
  su2k::launch_batch_reduce(b->dlat, n, b->partials, b->nltiles, b->dsums, b->stream);
  fin();
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(SU2GPU_EHIP, "batch launch: %s", hipGetErrorString(e));
  BH(hipMemcpyAsync(b->hsums, b->dsums, (size_t)n * 2 * NSUM * sizeof(double),
                    hipMemcpyDeviceToHost, b->stream));
  return SU2GPU_OK;
}

long double bsum(const B *b, int l, int r) {
  return (long double)b->hsums[l * 2 * NSUM + 2 * r] + (long double)b->hsums[l * 2 * NSUM + 2 * r + 1];
}

int run_trajectory(B *b, bool devmom, uint64_t traj, int metropolis, su2gpu_md_result *res,
                   Rec *rec = nullptr) {
  *b->htraj = traj;
  BH(hipMemcpyAsync(b->dtraj, b->htraj, sizeof(uint64_t), hipMemcpyHostToDevice, b->stream));
  for (auto &L : b->lat) L.fin = fin_arena(L);
  BK(upload_desc(b));
  if (rec) {
    BK(enqueue(b, devmom, rec));
  } else if (b->use_graph) {
    if (!b->gexec || b->graph_devmom != (int)devmom) {
      if (b->gexec) BH(hipGraphExecDestroy(b->gexec));
      if (b->graph) BH(hipGraphDestroy(b->graph));
      b->gexec = nullptr;
      b->graph = nullptr;
      BH(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue(b, devmom);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc != SU2GPU_OK) return rc;
      if (e != hipSuccess) return bfail(SU2GPU_EHIP, "graph capture: %s", hipGetErrorString(e));
      b->graph = g;
      BH(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
      b->graph_devmom = (int)devmom;
    }
    BH(hipGraphLaunch(b->gexec, b->stream));
  } else {
    BK(enqueue(b, devmom));
  }
  BH(hipStreamSynchronize(b->stream));
  for (size_t l = 0; l < b->lat.size(); ++l) {
    auto &L = b->lat[l];
    const double Nc = L.group == 1 ? 2.0 : 1.0;
    const long double sg = -(long double)b->mem[l].beta / Nc;
    const long double Ko = 0.5L * bsum(b, (int)l, 0), Kn = 0.5L * bsum(b, (int)l, 2);
    const long double So = sg * bsum(b, (int)l, 4), Sn = sg * bsum(b, (int)l, 3);
    su2gpu_md_result &r = res[l];
    std::memset(&r, 0, sizeof(r));
    r.kin_old = (double)Ko;
    r.kin_new = (double)Kn;
    r.gauge_old = (double)So;
    r.gauge_new = (double)Sn;
    r.deltaH = (double)((Kn - Ko) + (Sn - So));
    r.deltadeltaH = NAN;
    r.uniform = -1.0;
    int accept = 1;
    if (metropolis && r.deltaH > 0) {
      r.uniform = philox_accept_uniform(L.seed, traj);
      if (r.uniform > std::exp(-r.deltaH)) accept = 0;
    }
    r.accept = accept;
    if (accept) L.cur = L.fin;
  }
  return SU2GPU_OK;
}

int ensure_stage(B *b, size_t n) {
  if (b->stage_cap < n) {
    if (b->staging) BH(hipFree(b->staging));
    if (b->hstage) BH(hipHostFree(b->hstage));
    b->staging = nullptr;
    b->hstage = nullptr;
    BH(hipMalloc(&b->staging, n * sizeof(double)));
    BH(hipHostMalloc(&b->hstage, n * sizeof(double), hipHostMallocDefault));
    b->stage_cap = n;
  }
  return SU2GPU_OK;
}

std::string signature(const B *b) {
  std::ostringstream ss;
  ss << "prec" << b->prec << "_n" << b->mem.size();
  unsigned long long hsh = 1469598103934665603ull;
  for (auto &m : b->mem) {
    const unsigned long long v[6] = {(unsigned long long)m.group, m.Lx, m.Ly, m.Lz, m.Lt, m.ndims};
    for (auto x : v) hsh = (hsh ^ x) * 1099511628211ull;
    hsh = (hsh ^ (unsigned long long)npairs_of(m.integrator, m.nsteps)) * 1099511628211ull;
  }
  ss << "_" << std::hex << hsh;
  return ss.str();
}

}  // namespace

extern "C" {

int su2gpu_batch_create(su2gpu_batch **out, const su2gpu_batch_member *members, size_t n,
                        int precision) {
  if (!out || !members || n == 0) return bfail(SU2GPU_EINVAL, "bad batch arguments");
  *out = nullptr;
  if (precision != SU2GPU_FP64 && precision != SU2GPU_FP32) return bfail(SU2GPU_EINVAL, "bad precision");
  B *b = new B();
  b->prec = precision;
  b->mem.assign(members, members + n);
  const size_t rb = precision == SU2GPU_FP64 ? 8 : 4;
  for (size_t i = 0; i < n; ++i) {
    const auto &m = members[i];
    if ((m.group != SU2GPU_U1 && m.group != SU2GPU_SU2) || m.ndims < 2 || m.ndims > 4 ||
        !m.Lx || !m.Ly || !m.Lz || !m.Lt || (m.ndims < 4 && m.Lz != 1) ||
        (m.ndims < 3 && m.Ly != 1) || m.nsteps < 1 ||
        (m.integrator != SU2GPU_LEAPFROG && m.integrator != SU2GPU_OMF2)) {
      delete b;
      return bfail(SU2GPU_EINVAL, "batch member %zu: bad geometry / integrator (leapfrog or omf2)", i);
    }
    su2k::BLat L{};
    const int Ls[4] = {(int)m.Lt, (int)m.Lx, (int)m.Ly, (int)m.Lz};
    for (int k = 0; k < 4; ++k) L.geo.L[k] = Ls[k];
    L.geo.str[3] = 1;
    L.geo.str[2] = Ls[3];
    L.geo.str[1] = Ls[3] * Ls[2];
    L.geo.str[0] = Ls[3] * Ls[2] * Ls[1];
    L.geo.V = Ls[0] * Ls[1] * Ls[2] * Ls[3];
    L.geo.nd = (int)m.ndims;
    su2dev::geo_set_div(L.geo);
    L.group = m.group;
    L.kind = m.integrator;
    L.npairs = npairs_of(m.integrator, m.nsteps);
    L.cur = 0;
    L.coef = 1.0 * m.beta / (m.group == SU2GPU_SU2 ? 2.0 : 1.0);
    L.dtau = m.tau / double(m.nsteps);
    L.seed = m.seed;
    const size_t nc = m.group == SU2GPU_SU2 ? 4 : 1, nm = m.group == SU2GPU_SU2 ? 3 : 1;
    L.uoff = (long)b->ubytes;
    b->ubytes += ((size_t)L.geo.V * m.ndims * nc * rb + 255) / 256 * 256;
    L.poff = (long)b->pelems;
    b->pelems += ((size_t)L.geo.V * m.ndims * nm + 63) / 64 * 64;
    b->maxpairs = std::max(b->maxpairs, L.npairs);
    b->lat.push_back(L);
  }
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  for (int k = 0; k < 2 && e == hipSuccess; ++k)
    e = hipStreamCreateWithFlags(&b->pstream[k], hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
  for (int k = 0; k < 2 && e == hipSuccess; ++k)
    e = hipEventCreateWithFlags(&b->ev_join[k], hipEventDisableTiming);
  {
    // SU2GPU_BATCH_CONCURRENT = 0: the three parts one after another on one stream
    const char *ev = std::getenv("SU2GPU_BATCH_CONCURRENT");
    b->concurrent = !(ev && std::atoi(ev) == 0);
  }
  for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipMalloc(&b->ar.U[k], b->ubytes);
  if (e == hipSuccess) e = hipMalloc(&b->ar.P, b->pelems * rb);
  if (e == hipSuccess) e = hipMalloc(&b->dlat, n * sizeof(su2k::BLat));
  if (e == hipSuccess) e = hipMalloc(&b->dsums, n * 2 * NSUM * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&b->hsums, n * 2 * NSUM * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&b->dtraj, sizeof(uint64_t));
  if (e == hipSuccess) e = hipHostMalloc(&b->htraj, sizeof(uint64_t), hipHostMallocDefault);
  int dev = 0;
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) {
    bfail(SU2GPU_EHIP, "batch create: %s", hipGetErrorString(e));
    su2gpu_batch_destroy(b);
    return SU2GPU_EHIP;
  }
  int rc = build_tiles(b);
  if (rc == SU2GPU_OK) rc = upload_desc(b);
  if (rc == SU2GPU_OK) {
    // cold start: unit links
    for (size_t i = 0; i < n && rc == SU2GPU_OK; ++i) {
      const auto &L = b->lat[i];
      const size_t nc = L.group == 1 ? 4 : 1;
      std::vector<double> one((size_t)L.geo.V * L.geo.nd * nc, 0.0);
      if (nc == 4)
        for (size_t k = 0; k < one.size(); k += 4) one[k] = 1.0;
      rc = su2gpu_batch_set_links(b, i, one.data());
    }
  }
  if (rc != SU2GPU_OK) {
    su2gpu_batch_destroy(b);
    return rc;
  }
  *out = b;
  return SU2GPU_OK;
}

int su2gpu_batch_destroy(su2gpu_batch *b) {
  if (!b) return SU2GPU_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  for (hipStream_t ps : b->pstream)
    if (ps) hipStreamSynchronize(ps);
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (int k = 0; k < 3; ++k)
    if (b->ar.U[k]) hipFree(b->ar.U[k]);
  for (void *p : {(void *)b->ar.P, (void *)b->dlat, (void *)b->dtiles, (void *)b->dltiles, (void *)b->partials,
                  (void *)b->dsums, (void *)b->dtraj, (void *)b->staging})
    if (p) hipFree(p);
  for (void *p : {(void *)b->hsums, (void *)b->htraj, (void *)b->hstage})
    if (p) hipHostFree(p);
  for (hipStream_t ps : b->pstream)
    if (ps) hipStreamDestroy(ps);
  for (hipEvent_t ev : {b->ev_fork, b->ev_join[0], b->ev_join[1]})
    if (ev) hipEventDestroy(ev);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return SU2GPU_OK;
}

int su2gpu_batch_size(const su2gpu_batch *b, size_t *n) {
  if (!b || !n) return bfail(SU2GPU_EINVAL, "null pointer");
  *n = b->mem.size();
  return SU2GPU_OK;
}

int su2gpu_batch_set_links(su2gpu_batch *b, size_t i, const double *src) {
  if (!b || !src || i >= b->lat.size()) return bfail(SU2GPU_EINVAL, "bad batch member");
  const auto &L = b->lat[i];
  const int nc = L.group == 1 ? 4 : 1;
  const size_t n = (size_t)L.geo.V * L.geo.nd * nc;
  BK(ensure_stage(b, n));
  BH(hipMemcpyAsync(b->staging, src, n * sizeof(double), hipMemcpyHostToDevice, b->stream));
  su2k::launch_links_from_aos(L.group, b->prec, b->staging, nc, 0, b->ar.U[L.cur] + L.uoff, L.geo.V,
                              L.geo.nd, b->stream);
  BH(hipGetLastError());
  BH(hipStreamSynchronize(b->stream));
  return SU2GPU_OK;
}

int su2gpu_batch_get_links(su2gpu_batch *b, size_t i, double *dst) {
  if (!b || !dst || i >= b->lat.size()) return bfail(SU2GPU_EINVAL, "bad batch member");
  const auto &L = b->lat[i];
  const int nc = L.group == 1 ? 4 : 1;
  const size_t n = (size_t)L.geo.V * L.geo.nd * nc;
  BK(ensure_stage(b, n));
  su2k::launch_links_to_aos(L.group, b->prec, b->ar.U[L.cur] + L.uoff, b->staging, nc, 0, L.geo.V,
                            L.geo.nd, b->stream);
  BH(hipGetLastError());
  BH(hipMemcpyAsync(dst, b->staging, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BH(hipStreamSynchronize(b->stream));
  return SU2GPU_OK;
}

int su2gpu_batch_set_momenta(su2gpu_batch *b, size_t i, const double *src) {
  if (!b || !src || i >= b->lat.size()) return bfail(SU2GPU_EINVAL, "bad batch member");
  const auto &L = b->lat[i];
  const int nm = L.group == 1 ? 3 : 1;
  const size_t n = (size_t)L.geo.V * L.geo.nd * nm;
  BK(ensure_stage(b, n));
  BH(hipMemcpyAsync(b->staging, src, n * sizeof(double), hipMemcpyHostToDevice, b->stream));
  const size_t rb = b->prec == 0 ? 8 : 4;
  su2k::launch_mom_from_aos(b->prec, b->staging, (char *)b->ar.P + (size_t)L.poff * rb, L.geo.V,
                            L.geo.nd, nm, b->stream);
  BH(hipGetLastError());
  BH(hipStreamSynchronize(b->stream));
  return SU2GPU_OK;
}

int su2gpu_batch_trajectory(su2gpu_batch *b, int momenta_source, uint64_t traj, int metropolis,
                            su2gpu_md_result *results) {
  if (!b || !results) return bfail(SU2GPU_EINVAL, "null pointer");
  if (momenta_source != SU2GPU_MOMENTA_DEVICE_PHILOX && momenta_source != SU2GPU_MOMENTA_HOST_MT19937)
    return bfail(SU2GPU_EINVAL, "bad momenta source");
  // host source: the momenta set with su2gpu_batch_set_momenta are used as they are
  return run_trajectory(b, momenta_source == SU2GPU_MOMENTA_DEVICE_PHILOX, traj, metropolis, results);
}

int su2gpu_batch_set_launch(su2gpu_batch *b, int block, int grid) {
  if (!b || (block != 64 && block != 128 && block != 256) || grid < 0)
    return bfail(SU2GPU_EINVAL, "block must be 64, 128 or 256, grid >= 0");
  b->block = block;
  b->grid = grid;
  BK(build_tiles(b));
  return SU2GPU_OK;
}

int su2gpu_batch_get_launch(const su2gpu_batch *b, int *block, int *grid) {
  if (!b || !block || !grid) return bfail(SU2GPU_EINVAL, "null pointer");
  *block = b->block;
  *grid = b->grid;
  return SU2GPU_OK;
}

int su2gpu_batch_use_graph(su2gpu_batch *b, int on) {
  if (!b) return bfail(SU2GPU_EINVAL, "null batch");
  b->use_graph = on != 0;
  return SU2GPU_OK;
}

// One trajectory of device momenta (Metropolis as su2gpu_batch_trajectory),
// eager, with HIP events around every launch group.  out[0 .. 7]: ms of the
// whole trajectory (first to last event), of the momentum draw, the two
// plaquette sums, the MD ticks of the SU(2) d = 4 part, of the d = 3 part, of
// the remaining (generic) part, the final restore and the reduction;
// out[8 .. 10]: link updates of the three parts; out[11 .. 13]: their active
// tile fraction (tile-ticks of lattices not yet finished over the part's
// tiles times its ticks: what the step-count spread leaves of a full launch
// per tick; the launches cover only the active tiles).
int su2gpu_batch_profile(su2gpu_batch *b, uint64_t traj, double *out, size_t nout) {
  if (!b || !out) return bfail(SU2GPU_EINVAL, "null pointer");
  if (nout < 14) return bfail(SU2GPU_EINVAL, "batch profile: out needs 14 entries");
  Rec rec;
  rec.st = b->stream;
  std::vector<su2gpu_md_result> res(b->lat.size());
  BK(run_trajectory(b, true, traj, 1, res.data(), &rec));
  for (size_t i = 0; i < 14; ++i) out[i] = 0.0;
  for (size_t i = 0; i < rec.kind.size(); ++i) {
    float ms = 0.f;
    BH(hipEventElapsedTime(&ms, rec.ev[2 * i], rec.ev[2 * i + 1]));
    out[1 + rec.kind[i]] += ms;
  }
  if (!rec.ev.empty()) {
    float ms = 0.f;
    BH(hipEventElapsedTime(&ms, rec.ev.front(), rec.ev.back()));
    out[0] = ms;
  }
  double act[3] = {0, 0, 0}, all[3] = {0, 0, 0};
  for (size_t i = 0; i < b->lat.size(); ++i) {
    const su2k::BLat &L = b->lat[i];
    const int part = L.spec == 4 ? 0 : L.spec == 3 ? 1 : 2;
    out[8 + part] += (double)L.geo.V * L.geo.nd * (double)b->mem[i].nsteps;
    act[part] += (double)L.nltiles * (L.npairs + 1);
    all[part] += (double)L.nltiles * (b->maxpairs + 1);
  }
  for (int k = 0; k < 3; ++k) out[11 + k] = all[k] > 0 ? act[k] / all[k] : 0.0;
  return SU2GPU_OK;
}

int su2gpu_batch_link_updates(const su2gpu_batch *b, double *per_trajectory) {
  if (!b || !per_trajectory) return bfail(SU2GPU_EINVAL, "null pointer");
  double s = 0;
  for (size_t i = 0; i < b->mem.size(); ++i)
    s += (double)b->lat[i].geo.V * b->lat[i].geo.nd * (double)b->mem[i].nsteps;
  *per_trajectory = s;
  return SU2GPU_OK;
}

// Times every candidate (block in {64,128,256}) x (grid in {1,2,4,8} x CUs, one
// pass over the tiles) on a trajectory of device momenta with every lattice's
// links saved and restored around it, keeps the fastest, and records it in the
// JSON cache at cache_path (NULL: $SU2GPU_TUNE_CACHE or no cache) keyed by the
// batch signature.  A cache hit applies the recorded shape without timing.
int su2gpu_batch_autotune(su2gpu_batch *b, const char *cache_path, int *block_out, int *grid_out) {
  if (!b) return bfail(SU2GPU_EINVAL, "null batch");
  std::string path = cache_path ? cache_path : "";
  if (path.empty() && std::getenv("SU2GPU_TUNE_CACHE")) path = std::getenv("SU2GPU_TUNE_CACHE");
  const std::string key = signature(b);
  std::map<std::string, std::pair<int, int>> cache;
  if (!path.empty()) {
    std::ifstream in(path);
    std::string line;
    // one entry per line: "key": [block, grid],
    while (std::getline(in, line)) {
      const auto q1 = line.find('"'), q2 = line.find('"', q1 + 1), lb = line.find('[');
      if (q1 == std::string::npos || q2 == std::string::npos || lb == std::string::npos) continue;
      int bl = 0, gr = 0;
      if (std::sscanf(line.c_str() + lb, "[%d, %d]", &bl, &gr) == 2)
        cache[line.substr(q1 + 1, q2 - q1 - 1)] = {bl, gr};
    }
  }
  auto it = cache.find(key);
  if (it != cache.end()) {
    BK(su2gpu_batch_set_launch(b, it->second.first, it->second.second));
  } else {
    // a trajectory never writes a lattice's start arena, so restoring the
    // cursors after the timing runs leaves the configurations untouched
    std::vector<int> saved_cur;
    for (auto &L : b->lat) saved_cur.push_back(L.cur);
    auto restore_cur = [&]() {
      for (size_t l = 0; l < b->lat.size(); ++l) b->lat[l].cur = saved_cur[l];
    };
    std::vector<su2gpu_md_result> res(b->lat.size());
    const bool g = b->use_graph;
    b->use_graph = false;
    hipEvent_t e0, e1;
    BH(hipEventCreate(&e0));
    BH(hipEventCreate(&e1));
    float best = 1e30f;
    int bb = 256, bg = 0;
    for (int bl : {64, 128, 256}) {
      b->block = bl;
      b->grid = 0;
      BK(build_tiles(b));
      const int full = b->grid;
      for (int mult : {1, 2, 4, 8, 0}) {
        const int gr = mult == 0 ? full : std::min(full, ((mult * b->ncu + 7) / 8) * 8);
        b->grid = gr;
        restore_cur();
        BK(run_trajectory(b, true, 0x7f7f7f7full, 0, res.data()));  // warm-up
        restore_cur();
        BH(hipEventRecord(e0, b->stream));
        BK(run_trajectory(b, true, 0x7f7f7f7full, 0, res.data()));
        BH(hipEventRecord(e1, b->stream));
        BH(hipEventSynchronize(e1));
        float ms = 0;
        BH(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          bb = bl;
          bg = gr;
        }
      }
    }
    BH(hipEventDestroy(e0));
    BH(hipEventDestroy(e1));
    restore_cur();
    b->use_graph = g;
    b->block = bb;
    b->grid = bg;
    BK(build_tiles(b));
    BK(upload_desc(b));
    if (!path.empty()) {
      cache[key] = {bb, bg};
      std::ofstream o(path);
      o << "{\n";
      size_t k = 0;
      for (auto &kv : cache)
        o << "  \"" << kv.first << "\": [" << kv.second.first << ", " << kv.second.second << "]"
          << (++k < cache.size() ? "," : "") << "\n";
      o << "}\n";
    }
  }
  if (block_out) *block_out = b->block;
  if (grid_out) *grid_out = b->grid;
  return SU2GPU_OK;
}

}  // extern "C"
