// su2_capi.cpp -- host side of the C ABI declared in include/su2_mi355x.h.
//
// Owns the device fields of one lattice, sequences the integrators of
// integrator.hh as launches of the fused kernels, runs the HMC trajectory of
// md_update.hh, and reproduces the reference's host random streams (libstdc++
// std::mt19937 / normal_distribution / uniform_real_distribution) wherever the
// reference draws on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/su2_mi355x.h"
#include "../../include/su2_mi355x/normals.hh"
#include "su2_launch.hpp"
#include "su2_handle.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HC(expr)                                                                         \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(SU2GPU_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
#define CK(expr)                  \
  do {                            \
    int r_ = (expr);              \
    if (r_ != SU2GPU_OK) return r_; \
  } while (0)

constexpr int NBUF = 4;
constexpr int NROWS = 36;

}  // namespace

struct su2gpu_lattice {
  int group = 1, prec = 0, nd = 4, nm = 3, ncomp = 4;
  size_t Lx = 0, Ly = 0, Lz = 0, Lt = 0;
  double beta = 0, xi = 1;
  bool aniso = false;
  su2k::Geo geo{};
  int V = 0;
  long nlinks = 0;
  size_t real_bytes = 8, U_bytes = 0, P_bytes = 0;
  void *U[NBUF] = {nullptr, nullptr, nullptr, nullptr};
  // 1: U[i] holds FP64 SU(2) links in the half-field layout (an integrator's
  // intermediate links); 0: the record layout every other path reads
  int half[NBUF] = {0, 0, 0, 0};
  int cur = 0;
  void *P = nullptr, *deriv = nullptr;
  double *partials = nullptr;
  int row_stride = 0;
  double *dsums = nullptr, *hsums = nullptr;
  double *staging = nullptr;
  size_t staging_cap = 0;
  double *hstage = nullptr;
  std::vector<double> hpairs;  // accepted polar-method pairs of host_normals
  size_t hstage_cap = 0;
  hipStream_t stream = nullptr;
  int device = 0;
  // split md_update state
  bool pending = false;
  int pend_old = -1, pend_new = -1;
  // device momenta of md_update not drawn yet: the trajectory's first fused
  // step draws them (FusedLaunch::gen), or flush_gen before anything else
  // reads P
  bool gen_pending = false;
  uint64_t gen_seed = 0, gen_traj = 0;
  // FP32 storage emulation of the FP64 form (su2gpu_set_fp32_storage_emulation):
  // bit 0 links, bit 1 momenta
  int r32 = 0;
  // profiling
  bool prof = false;
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> evpool;
  std::map<std::string, std::pair<double, long>> prof_acc;
  // Metropolis sweep on lattices with an odd extent: per direction mu the
  // sites of each of the three colour classes (sweep_colour), on the device,
  // class (mu, c) at sweep_list[sweep_off[3 mu + c] ..)
  int *sweep_list = nullptr;
  std::vector<int> sweep_off;
  // gradient-flow forward state kept for the backward pass
  struct Flow {
    bool valid = false;
    double eps = 0, value = 0;
    size_t nsteps = 0, ckpt = 0;
    int obs = 0;
    std::vector<void *> ckpts;          // W at steps 0, ckpt, 2 ckpt, ...
    void *W[2] = {nullptr, nullptr};    // forward ping-pong; W[fin] = W(t)
    int fin = 0;
    void *Z = nullptr;                  // RK3 Z field
    std::vector<void *> seg;            // recomputed W_k of one segment
    void *W1 = nullptr, *W2 = nullptr;  // stage states of one step
    void *Zs[2] = {nullptr, nullptr};   // Z after stages 0 and 1
    void *lamW = nullptr, *lamZ = nullptr, *Gam = nullptr;
    void *R = nullptr;                  // scratch of the adjoint's plaquette pass
    std::vector<void *> stU, stZ;       // stage states of a segment: W1, W2 and Z0', Z1', Z2' per step
    bool kept_last = false;             // the forward left the last segment's states in seg/stU/stZ
  } flow;
};

namespace {

using L = su2gpu_lattice;

// ---------------------------------------------------------------- helpers --
int ensure_buf(L *h, int i) {
  if (!h->U[i]) HC(hipMalloc(&h->U[i], h->U_bytes));
  return SU2GPU_OK;
}
int free_buf(L *h, std::initializer_list<int> excl, int *out) {
  for (int i = 0; i < NBUF; ++i) {
    bool used = false;
    for (int e : excl)
      if (e == i) used = true;
    if (!used) {
      CK(ensure_buf(h, i));
      *out = i;
      return SU2GPU_OK;
    }
  }
  return fail(SU2GPU_ESTATE, "no free link buffer");
}
int ensure_staging(L *h, size_t ndoubles) {
  if (h->staging_cap < ndoubles) {
    if (h->staging) HC(hipFree(h->staging));
    h->staging = nullptr;
    HC(hipMalloc(&h->staging, ndoubles * sizeof(double)));
    h->staging_cap = ndoubles;
  }
  if (h->hstage_cap < ndoubles) {
    if (h->hstage) HC(hipHostFree(h->hstage));
    h->hstage = nullptr;
    HC(hipHostMalloc(&h->hstage, ndoubles * sizeof(double), hipHostMallocDefault));
    h->hstage_cap = ndoubles;
  }
  return SU2GPU_OK;
}
int ensure_deriv(L *h) {
  if (!h->deriv) {
    HC(hipMalloc(&h->deriv, h->P_bytes));
    HC(hipMemsetAsync(h->deriv, 0, h->P_bytes, h->stream));
  }
  return SU2GPU_OK;
}

// HIP-event timing around one launch
struct Prof {
  L *h;
  size_t idx = (size_t)-1;
  Prof(L *hh, const char *name) : h(hh) {
    if (!h->prof) return;
    hipEvent_t a, b;
    if (h->evpool.size() >= 2) {
      a = h->evpool.back();
      h->evpool.pop_back();
      b = h->evpool.back();
      h->evpool.pop_back();
    } else {
      hipEventCreate(&a);
      hipEventCreate(&b);
    }
    hipEventRecord(a, h->stream);
    h->recs.push_back({name, a, b});
    idx = h->recs.size() - 1;
  }
  ~Prof() {
    if (idx != (size_t)-1) hipEventRecord(h->recs[idx].b, h->stream);
  }
};

int collect_prof(L *h) {
  if (h->recs.empty()) return SU2GPU_OK;
  HC(hipStreamSynchronize(h->stream));
  for (auto &r : h->recs) {
    float ms = 0;
    HC(hipEventElapsedTime(&ms, r.a, r.b));
    auto &acc = h->prof_acc[r.name];
    acc.first += ms;
    acc.second += 1;
    h->evpool.push_back(r.a);
    h->evpool.push_back(r.b);
  }
  h->recs.clear();
  return SU2GPU_OK;
}

int last_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SU2GPU_EHIP, "launch %s: %s", what, hipGetErrorString(e));
  return SU2GPU_OK;
}

double Nc(const L *h) { return h->group == 1 ? 2.0 : 1.0; }

su2k::FusedCoef base_coef(const L *h) {
  su2k::FusedCoef c{};
  c.coef = 1.0 * h->beta / Nc(h);
  c.w_t = h->aniso ? 1.0 / h->xi : 1.0;
  c.w_s = h->aniso ? h->xi : 1.0;
  c.flags = (h->aniso ? su2k::F_ANISO : 0) | (h->prec == 1 ? su2k::F_COMPENSATED : 0);
  return c;
}

double *row(L *h, int r) { return h->partials + (size_t)r * h->row_stride; }

// fused "P then U" (Uout != nullptr) or momentum-only step on raw buffers
int fused_ptr(L *h, const void *Uin, void *Uout, void *P, su2k::FusedCoef c, int accum_row,
              void *Pout = nullptr) {
  su2k::FusedLaunch fl;
  fl.Uin = Uin;
  fl.Uout = Uout;
  fl.P = P;
  fl.Pout = Pout;
  fl.geo = h->geo;
  fl.c = c;
  fl.partials = nullptr;
  fl.pstride = h->row_stride;
  if (accum_row >= 0) {
    fl.c.flags |= su2k::F_ACCUM;
    fl.partials = h->partials + (size_t)accum_row * h->row_stride;
  }
  {
    Prof pr(h, Uout ? "md_fused" : "md_p_only");
    su2k::launch_fused(h->group, h->prec, fl, h->stream);
  }
  return last_launch("md_fused");
}

// fused "P then U" (Uout >= 0) or momentum-only (Uout < 0) step.  allow_half
// (the integrators only): where the step supports it, the new links go out in
// the half-field layout; links already in it are read in it.
int flush_gen(L *h);
int fused(L *h, int uin, int uout, void *P, su2k::FusedCoef c, int accum_row,
          bool allow_half = false) {
  // pending device momenta: drawn inside a first P-then-U step with partial
  // sums, else by their own launch now
  const bool gen = h->gen_pending && P == h->P && uout >= 0 && accum_row >= 0;
  if (h->gen_pending && !gen) CK(flush_gen(h));
  su2k::FusedLaunch fl;
  fl.Uin = h->U[uin];
  fl.Uout = uout >= 0 ? h->U[uout] : nullptr;
  fl.P = P;
  fl.geo = h->geo;
  fl.c = c;
  fl.partials = nullptr;
  fl.pstride = h->row_stride;
  if (accum_row >= 0) {
    fl.c.flags |= su2k::F_ACCUM;
    fl.partials = row(h, accum_row);
  }
  fl.in_half = h->half[uin];
  if (allow_half && su2k::fused_half_ok(h->group, h->prec, h->geo, fl.c.flags))
    fl.out_half = uout >= 0 ? 1 : 0;
  else if (fl.in_half)
    return fail(SU2GPU_ESTATE, "links in the half-field layout where the step cannot read them");
  if (gen) {
    fl.gen = 1;
    fl.c.gseed = h->gen_seed;
    fl.c.gtraj = h->gen_traj;
    h->gen_pending = false;
  }
  bool ok;
  {
    Prof pr(h, uout >= 0 ? "md_fused" : "md_p_only");
    ok = su2k::launch_fused(h->group, h->prec, fl, h->stream);
  }
  if (!ok) return fail(SU2GPU_ESTATE, "fused step: link layout not supported for this lattice");
  if (uout >= 0) h->half[uout] = fl.out_half;
  return last_launch("md_fused");
}

// pending device momenta drawn by their own launch
int flush_gen(L *h) {
  if (!h->gen_pending) return SU2GPU_OK;
  h->gen_pending = false;
  {
    Prof pr(h, "momenta");
    su2k::launch_momenta_philox(h->prec, h->P, h->V, h->nd, h->nm, h->gen_seed, h->gen_traj,
                                h->stream);
  }
  return last_launch("momenta");
}

// links of buffer b in the record layout: a half-field buffer is converted
// into a free buffer (not one of keep), which becomes the new b
int to_record(L *h, int *b, std::initializer_list<int> keep) {
  if (!h->half[*b]) return SU2GPU_OK;
  int out = -1;
  for (int i = 0; i < NBUF && out < 0; ++i) {
    bool used = (i == *b);
    for (int k : keep) used = used || (k == i);
    if (!used) out = i;
  }
  if (out < 0) return fail(SU2GPU_ESTATE, "no free link buffer");
  CK(ensure_buf(h, out));
  {
    Prof pr(h, "convert");
    su2k::launch_half_to_record(h->U[*b], h->U[out], h->nlinks, h->stream);
  }
  CK(last_launch("half_to_record"));
  h->half[out] = 0;
  *b = out;
  return SU2GPU_OK;
}

// one reduce launch for every row of rl
int reduce_rows(L *h, const su2k::RowList &rl) {
  {
    Prof pr(h, "reduce");
    su2k::launch_reduce(h->partials, (size_t)h->row_stride, rl, h->dsums, h->stream);
  }
  return last_launch("reduce");
}
// reduce row r (nb partials) now, or append it to *defer for a later reduce_rows
int reduce_row(L *h, int r, int nb, su2k::RowList *defer = nullptr) {
  su2k::RowList one;
  su2k::RowList &rl = defer ? *defer : one;
  if (rl.n >= su2k::RowList::MAX) return fail(SU2GPU_ESTATE, "reduce row list full");
  rl.row[rl.n] = r;
  rl.nb[rl.n] = nb;
  ++rl.n;
  return defer ? SU2GPU_OK : reduce_rows(h, rl);
}
int fetch_sums(L *h) {
  HC(hipMemcpyAsync(h->hsums, h->dsums, 2 * NROWS * sizeof(double), hipMemcpyDeviceToHost,
                    h->stream));
  HC(hipStreamSynchronize(h->stream));
  return SU2GPU_OK;
}
long double sum_of(const L *h, int r) {
  return (long double)h->hsums[2 * r] + (long double)h->hsums[2 * r + 1];
}

int plaq_launch(L *h, int uin, int uout, double w_t, double w_s, int mu_start, int flags, int r,
                su2k::RowList *defer = nullptr) {
  int nb;
  {
    Prof pr(h, "plaq");
    nb = su2k::launch_plaq(h->group, h->prec, h->U[uin], uout >= 0 ? h->U[uout] : nullptr,
                           h->geo, w_t, w_s, mu_start, flags, row(h, r), h->stream, h->half[uin]);
  }
  if (nb < 0) return fail(SU2GPU_ESTATE, "plaquette sum: link layout not supported");
  CK(last_launch("plaq"));
  // restored links (F_RESTORE) are written in the record layout
  if (uout >= 0) h->half[uout] = 0;
  return reduce_row(h, r, nb, defer);
}

int sumsq_launch(L *h, const void *P, int r, su2k::RowList *defer = nullptr) {
  const long n = h->nlinks * h->nm;
  {
    Prof pr(h, "sumsq");
    su2k::launch_sumsq(h->prec, P, n, row(h, r), h->stream);
  }
  CK(last_launch("sumsq"));
  return reduce_row(h, r, su2k::sumsq_blocks(n), defer);
}

// ------------------------------------------------------------- integrators --
struct Op {
  char kind;  // 'P' or 'U'
  double c;   // step = c * dtau
};

// the operation sequences of integrator.hh (leapfrog :26-50, omf2 :83-115,
// omf4 :118-155 with its loop from i = 1, euler :201-217, ruth :220-241)
int integrator_ops(int kind, size_t n, std::vector<Op> &ops) {
  ops.clear();
  if (n < 1) return fail(SU2GPU_EINVAL, "nsteps must be >= 1");
  switch (kind) {
    case SU2GPU_LEAPFROG:
    case SU2GPU_LP_LEAPFROG:
      ops.push_back({'P', 0.5});
      ops.push_back({'U', 1.0});
      for (size_t i = 0; i + 1 < n; ++i) {
        ops.push_back({'P', 1.0});
        ops.push_back({'U', 1.0});
      }
      ops.push_back({'P', 0.5});
      break;
    case SU2GPU_OMF2: {
      const double lambda = 0.1938, om = 1. - 2. * lambda;
      ops.push_back({'P', lambda});
      for (size_t i = 0; i + 1 < n; ++i) {
        ops.push_back({'U', 0.5});
        ops.push_back({'P', om});
        ops.push_back({'U', 0.5});
        ops.push_back({'P', 2 * lambda});
      }
      ops.push_back({'U', 0.5});
      ops.push_back({'P', om});
      ops.push_back({'U', 0.5});
      ops.push_back({'P', lambda});
      break;
    }
    case SU2GPU_OMF4:
    case SU2GPU_LP_OMF4: {
      const double rho = 0.2539785108410595, theta = -0.03230286765269967,
                   vartheta = 0.08398315262876693, lambda = 0.6822365335719091;
      const double e[10] = {rho, lambda, theta, 0.5 * (1 - 2. * (lambda + vartheta)),
                            (1 - 2. * (theta + rho)), 0.5 * (1 - 2. * (lambda + vartheta)),
                            theta, lambda, rho, 2 * vartheta};
      ops.push_back({'P', 0.5 * e[9]});
      for (size_t i = 1; i + 1 < n; ++i)
        for (int j = 0; j < 5; ++j) {
          ops.push_back({'U', e[2 * j]});
          ops.push_back({'P', e[2 * j + 1]});
        }
      for (int j = 0; j < 4; ++j) {
        ops.push_back({'U', e[2 * j]});
        ops.push_back({'P', e[2 * j + 1]});
      }
      ops.push_back({'U', e[8]});
      ops.push_back({'P', 0.5 * e[9]});
      break;
    }
    case SU2GPU_EULER:
      for (size_t i = 0; i < n; ++i) {
        ops.push_back({'P', 1.0});
        ops.push_back({'U', 1.0});
      }
      break;
    case SU2GPU_RUTH:
      for (size_t i = 0; i < n; ++i) {
        ops.push_back({'P', 1.0});
        ops.push_back({'U', -1. / 24.});
        ops.push_back({'P', -2. / 3.});
        ops.push_back({'U', 3. / 4.});
        ops.push_back({'P', 2. / 3.});
        ops.push_back({'U', 7. / 24.});
      }
      break;
    default:
      // set_integrator falls back to leapfrog for an unknown id (integrator.hh:274-277)
      return integrator_ops(SU2GPU_LEAPFROG, n, ops);
  }
  return SU2GPU_OK;
}

// The coefficients are multiplied by dtau exactly as the reference does:
// leapfrog dtau/2. and dtau, omf2 lambda*dtau etc.  The products below use
// the same association as integrator.hh so that FP64 steps match bit for bit
// on the host side of the computation.
struct IntegrateOut {
  bool first_accum = false;  // rows 0..2 hold kin_old/act/kin_new of the first step
  bool last_p_only = false;  // rows 3..5 hold the last momentum-only step sums
  int final_buf = -1;
};

int run_integrator(L *h, int kind, size_t nsteps, double tau, size_t exponent,
                   std::initializer_list<int> protect, bool accum, int row_base,
                   IntegrateOut *io) {
  std::vector<Op> ops;
  CK(integrator_ops(kind, nsteps, ops));
  const double dtau = tau / double(nsteps);
  su2k::FusedCoef c0 = base_coef(h);
  if (kind == SU2GPU_LP_LEAPFROG || kind == SU2GPU_LP_OMF4) {
    c0.flags |= su2k::F_ROUND;
    c0.round_n = (double)(size_t)std::pow(10, (double)exponent);
  }
  if (h->prec == 1 && h->group == 1) c0.flags |= su2k::F_RESTORE;  // FP32: reunitarise every step
  if (h->r32 & 1) c0.flags |= su2k::F_R32L;
  if (h->r32 & 2) c0.flags |= su2k::F_R32P;
  int cur = h->cur;
  auto is_prot = [&](int b) {
    for (int p : protect)
      if (p == b) return true;
    return false;
  };
  bool first = true;
  size_t i = 0;
  while (i < ops.size()) {
    const Op &o = ops[i];
    if (o.kind == 'P') {
      su2k::FusedCoef c = c0;
      c.ep = o.c * dtau;
      const int acc_row = (accum && first) ? row_base + 0 : -1;
      const bool last_p = (i + 1 == ops.size());
      if (i + 1 < ops.size() && ops[i + 1].kind == 'U') {
        c.eu = ops[i + 1].c * dtau;
        int out;
        CK(free_buf(h, {cur, protect.size() > 0 ? *protect.begin() : cur,
                        protect.size() > 1 ? *(protect.begin() + 1) : cur},
                    &out));
        CK(fused(h, cur, out, h->P, c, acc_row, true));
        cur = out;
        i += 2;
      } else {
        const int r = (accum && first) ? row_base + 0 : (accum && last_p) ? row_base + 3 : -1;
        CK(fused(h, cur, -1, h->P, c, r, true));
        if (accum && last_p && !first) io->last_p_only = true;
        i += 1;
      }
      if (first && accum) io->first_accum = true;
      first = false;
    } else {
      // a link update without a momentum step before it (no integrator of
      // integrator.hh has one; kept general): the generic kernel, record layout
      CK(flush_gen(h));
      {
        std::initializer_list<int> keep = protect;
        CK(to_record(h, &cur, keep));
      }
      if (is_prot(cur)) {
        int out;
        CK(free_buf(h, {cur, protect.size() > 0 ? *protect.begin() : cur,
                        protect.size() > 1 ? *(protect.begin() + 1) : cur},
                    &out));
        HC(hipMemcpyAsync(h->U[out], h->U[cur], h->U_bytes, hipMemcpyDeviceToDevice, h->stream));
        h->half[out] = 0;
        cur = out;
      }
      {
        Prof pr(h, "update_gauge");
        su2k::launch_update_gauge(h->group, h->prec, h->U[cur], h->P, h->V, h->nd, o.c * dtau,
                                  c0.round_n, c0.flags & (su2k::F_ROUND | su2k::F_RESTORE),
                                  h->stream);
      }
      CK(last_launch("update_gauge"));
      first = false;
      i += 1;
    }
  }
  io->final_buf = cur;
  return SU2GPU_OK;
}

// host-side random start configurations with the reference's exact streams
constexpr double ref_pi() { return 3.14159265358979311599796346854; }  // atan(1)*4

// uniform_real_distribution<double>(a, b)(e) of libstdc++: u * (b - a) + a with
// u = generate_canonical<double, 53>(e) (canonical53, normals.hh)
inline double uniform_ab(std::mt19937 &e, double a, double b) {
  return (su2_mi355x::detail::canonical53(e) * (b - a)) + a;
}
void random_su2(std::mt19937 &e, double delta, double *q) {
  // random_element (random_element.hh:9-25)
  const double alpha = uniform_ab(e, 0., delta * 2 * ref_pi());
  const double u = uniform_ab(e, -1., 1.);
  const double theta = uniform_ab(e, 0., 2 * ref_pi());
  const double r = std::sqrt(1 - u * u);
  // the reference build evaluates each sin/cos pair with one glibc sincos call
  // (GCC merges them); sincos and separate sin/cos can differ in the last bit
  double salpha, calpha, st, ct;
  ::sincos(alpha, &salpha, &calpha);
  ::sincos(theta, &st, &ct);
  const double sr = salpha * r;
  q[0] = calpha;
  q[1] = salpha * u;
  q[2] = sr * st;
  q[3] = sr * ct;
}
double random_u1(std::mt19937 &e, double delta) {
  return uniform_ab(e, -ref_pi() * delta, ref_pi() * delta);
}

// (a1 + i b1)(a2 + i b2) in std::complex order
inline void cm(double ar, double ai, double br, double bi, double &rr, double &ri) {
  rr = ar * br - ai * bi;
  ri = ar * bi + ai * br;
}
void su2_mul_host(const double *u, const double *v, double *o) {
  double t1r, t1i, t2r, t2i;
  cm(u[0], u[1], v[0], v[1], t1r, t1i);
  cm(u[2], u[3], v[2], -v[3], t2r, t2i);
  double ar = t1r - t2r, ai = t1i - t2i;
  cm(u[0], u[1], v[2], v[3], t1r, t1i);
  cm(u[2], u[3], v[0], -v[1], t2r, t2i);
  o[0] = ar;
  o[1] = ai;
  o[2] = t1r + t2r;
  o[3] = t1i + t2i;
}

int upload_links_packed(L *h, const double *src_dev_staging, int stride, int offset) {
  int b;
  CK(ensure_buf(h, h->cur));
  b = h->cur;
  h->half[b] = 0;
  {
    Prof pr(h, "convert");
    su2k::launch_links_from_aos(h->group, h->prec, src_dev_staging, stride, offset, h->U[b], h->V,
                                h->nd, h->stream);
  }
  return last_launch("links_from_aos");
}

// host AoS (doubles, per-link stride / offset) -> device links
int set_links_host(L *h, const double *src, int stride, int offset) {
  if (h->pending) return fail(SU2GPU_ESTATE, "md_trajectory pending: call su2gpu_md_finish first");
  const size_t n = (size_t)h->nlinks * stride;
  CK(ensure_staging(h, n));
  HC(hipMemcpyAsync(h->staging, src, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  CK(upload_links_packed(h, h->staging, stride, offset));
  HC(hipStreamSynchronize(h->stream));
  return SU2GPU_OK;
}
int get_links_host(L *h, double *dst, int stride, int offset) {
  const size_t n = (size_t)h->nlinks * stride;
  CK(ensure_staging(h, n));
  if (stride != h->ncomp) HC(hipMemsetAsync(h->staging, 0, n * sizeof(double), h->stream));
  {
    Prof pr(h, "convert");
    su2k::launch_links_to_aos(h->group, h->prec, h->U[h->cur], h->staging, stride, offset, h->V,
                              h->nd, h->stream);
  }
  CK(last_launch("links_to_aos"));
  HC(hipMemcpyAsync(dst, h->staging, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HC(hipStreamSynchronize(h->stream));
  return SU2GPU_OK;
}
int set_mom_host(L *h, void *dstP, const double *src) {
  const size_t n = (size_t)h->nlinks * h->nm;
  CK(ensure_staging(h, n));
  HC(hipMemcpyAsync(h->staging, src, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  {
    Prof pr(h, "convert");
    su2k::launch_mom_from_aos(h->prec, h->staging, dstP, h->V, h->nd, h->nm, h->stream);
  }
  CK(last_launch("mom_from_aos"));
  HC(hipStreamSynchronize(h->stream));
  return SU2GPU_OK;
}
int get_mom_host(L *h, const void *srcP, double *dst) {
  const size_t n = (size_t)h->nlinks * h->nm;
  CK(ensure_staging(h, n));
  {
    Prof pr(h, "convert");
    su2k::launch_mom_to_aos(h->prec, srcP, h->staging, h->V, h->nd, h->nm, h->stream);
  }
  CK(last_launch("mom_to_aos"));
  HC(hipMemcpyAsync(dst, h->staging, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HC(hipStreamSynchronize(h->stream));
  return SU2GPU_OK;
}

// initnormal (adjointfield.hh:240-258) on the host into hstage
// initnormal (adjointfield.hh:240-258) of the handle's momenta on the host:
// the reference's normal stream bit for bit (two passes, normals.hh)
void host_normals(L *h, std::mt19937 &engine) {
  su2_mi355x::detail::polar_normals(engine, h->hstage, (size_t)h->nlinks * h->nm, h->hpairs);
}

int check(const L *h) {
  if (!h) return fail(SU2GPU_EINVAL, "null handle");
  return SU2GPU_OK;
}

}  // namespace

// =========================================================================== C ABI

extern "C" {

const char *su2gpu_last_error(void) { return g_err.c_str(); }

// error channel shared with the batch translation unit (not part of the header)
int su2gpu_internal_set_error(int code, const char *msg) {
  g_err = msg;
  return code;
}

}  // extern "C"

// what the entry points of other translation units need from a handle
// (declared in su2_handle.hpp; not part of the C ABI)
int su2_internal_lattice_view(su2gpu_lattice *h, su2k::LatticeView *v) {
  CK(check(h));
  v->group = h->group;
  v->prec = h->prec;
  v->geo = h->geo;
  v->U = h->U[h->cur];
  v->stream = h->stream;
  v->device = h->device;
  v->pending = h->pending;
  return SU2GPU_OK;
}

extern "C" {

int su2gpu_device_count(int *n) {
  if (!n) return fail(SU2GPU_EINVAL, "null pointer");
  HC(hipGetDeviceCount(n));
  return SU2GPU_OK;
}

int su2gpu_create(su2gpu_lattice **out, int group, size_t Lx, size_t Ly, size_t Lz, size_t Lt,
                  size_t ndims, double beta, double xi, int precision) {
  if (!out) return fail(SU2GPU_EINVAL, "null output handle");
  *out = nullptr;
  if (group != SU2GPU_U1 && group != SU2GPU_SU2) return fail(SU2GPU_EINVAL, "group must be U1 or SU2");
  if (ndims < 2 || ndims > 4) return fail(SU2GPU_EINVAL, "ndims must be 2, 3 or 4");
  if (precision != SU2GPU_FP64 && precision != SU2GPU_FP32) return fail(SU2GPU_EINVAL, "bad precision");
  if (!Lx || !Ly || !Lz || !Lt) return fail(SU2GPU_EINVAL, "extents must be >= 1");
  if (ndims < 4 && Lz != 1) return fail(SU2GPU_EINVAL, "ndims < 4 needs Lz = 1");
  if (ndims < 3 && Ly != 1) return fail(SU2GPU_EINVAL, "ndims < 3 needs Ly = 1");
  if (!(xi > 0)) return fail(SU2GPU_EINVAL, "xi must be > 0");
  const unsigned long long V = (unsigned long long)Lx * Ly * Lz * Lt;
  const int nm = group == SU2GPU_SU2 ? 3 : 1;
  if (V * ndims * (nm > 4 ? nm : 4) >= (1ull << 31))
    return fail(SU2GPU_EINVAL, "lattice too large for 32-bit indexing");
  int ndev = 0;
  HC(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(SU2GPU_EHIP, "no HIP device");
  L *h = new L();
  h->group = group;
  h->prec = precision;
  h->nd = (int)ndims;
  h->nm = nm;
  h->ncomp = group == SU2GPU_SU2 ? 4 : 1;
  h->Lx = Lx; h->Ly = Ly; h->Lz = Lz; h->Lt = Lt;
  h->beta = beta;
  h->xi = xi;
  h->aniso = (xi != 1.0);
  h->V = (int)V;
  h->nlinks = (long)V * (long)ndims;
  h->real_bytes = precision == SU2GPU_FP64 ? 8 : 4;
  h->U_bytes = (size_t)h->nlinks * h->ncomp * h->real_bytes;
  h->P_bytes = (size_t)h->nlinks * h->nm * h->real_bytes;
  // direction mu = (t, x, y, z)
  const int Ls[4] = {(int)Lt, (int)Lx, (int)Ly, (int)Lz};
  h->geo.L[0] = Ls[0]; h->geo.L[1] = Ls[1]; h->geo.L[2] = Ls[2]; h->geo.L[3] = Ls[3];
  h->geo.str[3] = 1;
  h->geo.str[2] = Ls[3];
  h->geo.str[1] = Ls[3] * Ls[2];
  h->geo.str[0] = Ls[3] * Ls[2] * Ls[1];
  h->geo.V = h->V;
  h->geo.nd = h->nd;
  su2dev::geo_set_div(h->geo);
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  const int nbf = su2k::fused_blocks(h->geo), nbs = su2k::sumsq_blocks((long)h->nlinks * h->nm);
  h->row_stride = nbf > nbs ? nbf : nbs;
  if (e == hipSuccess) e = hipMalloc(&h->U[0], h->U_bytes);
  if (e == hipSuccess) e = hipMalloc(&h->U[1], h->U_bytes);
  if (e == hipSuccess) e = hipMalloc(&h->U[2], h->U_bytes);
  if (e == hipSuccess) e = hipMalloc(&h->P, h->P_bytes);
  if (e == hipSuccess) e = hipMalloc(&h->partials, (size_t)NROWS * h->row_stride * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->dsums, 2 * NROWS * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&h->hsums, 2 * NROWS * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemsetAsync(h->P, 0, h->P_bytes, h->stream);
  if (e != hipSuccess) {
    fail(SU2GPU_EHIP, "create: %s", hipGetErrorString(e));
    su2gpu_destroy(h);
    return SU2GPU_EHIP;
  }
  *out = h;
  int r = su2gpu_coldstart(h);
  if (r != SU2GPU_OK) {
    su2gpu_destroy(h);
    *out = nullptr;
  }
  return r;
}

int su2gpu_destroy(su2gpu_lattice *h) {
  if (!h) return SU2GPU_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (auto &r : h->recs) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto ev : h->evpool) hipEventDestroy(ev);
  {
    auto &f = h->flow;
    for (void *p : f.ckpts) hipFree(p);
    for (void *p : f.seg) hipFree(p);
    for (void *p : f.stU) hipFree(p);
    for (void *p : f.stZ) hipFree(p);
    for (void *p : {f.W[0], f.W[1], f.Z, f.W1, f.W2, f.Zs[0], f.Zs[1], f.lamW, f.lamZ, f.Gam, f.R})
      if (p) hipFree(p);
  }
  for (int i = 0; i < NBUF; ++i)
    if (h->U[i]) hipFree(h->U[i]);
  if (h->P) hipFree(h->P);
  if (h->deriv) hipFree(h->deriv);
  if (h->partials) hipFree(h->partials);
  if (h->dsums) hipFree(h->dsums);
  if (h->hsums) hipHostFree(h->hsums);
  if (h->staging) hipFree(h->staging);
  if (h->hstage) hipHostFree(h->hstage);
  if (h->sweep_list) hipFree(h->sweep_list);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return SU2GPU_OK;
}

int su2gpu_set_beta(su2gpu_lattice *h, double beta) {
  CK(check(h));
  h->beta = beta;
  return SU2GPU_OK;
}
int su2gpu_get_beta(const su2gpu_lattice *h, double *beta) {
  CK(check(h));
  if (!beta) return fail(SU2GPU_EINVAL, "null pointer");
  *beta = h->beta;
  return SU2GPU_OK;
}
int su2gpu_get_geometry(const su2gpu_lattice *h, size_t out[8]) {
  CK(check(h));
  if (!out) return fail(SU2GPU_EINVAL, "null pointer");
  out[0] = h->Lx; out[1] = h->Ly; out[2] = h->Lz; out[3] = h->Lt;
  out[4] = (size_t)h->nd; out[5] = (size_t)h->V; out[6] = (size_t)h->group; out[7] = (size_t)h->prec;
  return SU2GPU_OK;
}

int su2gpu_set_links(su2gpu_lattice *h, const void *src, int layout) {
  CK(check(h));
  if (!src) return fail(SU2GPU_EINVAL, "null source");
  if (layout == SU2GPU_LAYOUT_PACKED) return set_links_host(h, (const double *)src, h->ncomp, 0);
  if (layout == SU2GPU_LAYOUT_OBJECT) return set_links_host(h, (const double *)src, h->ncomp + 1, 1);
  return fail(SU2GPU_EINVAL, "bad layout");
}

int su2gpu_get_links(su2gpu_lattice *h, void *dst, int layout) {
  CK(check(h));
  if (!dst) return fail(SU2GPU_EINVAL, "null destination");
  if (layout == SU2GPU_LAYOUT_PACKED) return get_links_host(h, (double *)dst, h->ncomp, 0);
  if (layout == SU2GPU_LAYOUT_OBJECT) {
    CK(get_links_host(h, (double *)dst, h->ncomp + 1, 1));
    // the objects' leading size_t N_c (su2.hh:13, u1.hh:12)
    const size_t nc = h->group == SU2GPU_SU2 ? 2 : 1;
    char *p = (char *)dst;
    for (long i = 0; i < h->nlinks; ++i)
      std::memcpy(p + (size_t)i * (h->ncomp + 1) * sizeof(double), &nc, sizeof(size_t));
    return SU2GPU_OK;
  }
  return fail(SU2GPU_EINVAL, "bad layout");
}

int su2gpu_set_momenta(su2gpu_lattice *h, const double *src) {
  CK(check(h));
  if (!src) return fail(SU2GPU_EINVAL, "null source");
  return set_mom_host(h, h->P, src);
}
int su2gpu_get_momenta(su2gpu_lattice *h, double *dst) {
  CK(check(h));
  if (!dst) return fail(SU2GPU_EINVAL, "null destination");
  return get_mom_host(h, h->P, dst);
}
int su2gpu_set_fp32_storage_emulation(su2gpu_lattice *h, int links, int momenta) {
  CK(check(h));
  if (h->prec != SU2GPU_FP64) return fail(SU2GPU_EINVAL, "FP32 storage emulation needs the FP64 form");
  if (h->pending) return fail(SU2GPU_ESTATE, "md_trajectory pending: call su2gpu_md_finish first");
  h->r32 = (links ? 1 : 0) | (momenta ? 2 : 0);
  return SU2GPU_OK;
}
int su2gpu_flip_momenta(su2gpu_lattice *h) {
  CK(check(h));
  {
    Prof pr(h, "scale");
    su2k::launch_scale(h->prec, h->P, h->nlinks * h->nm, -1.0, h->stream);
  }
  return last_launch("scale");
}
int su2gpu_initnormal(su2gpu_lattice *h, uint64_t seed) {
  CK(check(h));
  CK(ensure_staging(h, (size_t)h->nlinks * h->nm));
  std::mt19937 engine(seed);
  host_normals(h, engine);
  return set_mom_host(h, h->P, h->hstage);
}

int su2gpu_hotstart(su2gpu_lattice *h, uint64_t seed, double delta) {
  CK(check(h));
  if (delta < 0.) delta = 0.;
  if (delta > 1.) delta = 1.;
  const size_t n = (size_t)h->nlinks * h->ncomp;
  CK(ensure_staging(h, n));
  std::mt19937 engine((int)seed);
  for (long i = 0; i < h->nlinks; ++i) {
    if (h->group == SU2GPU_SU2) random_su2(engine, delta, h->hstage + 4 * i);
    else h->hstage[i] = random_u1(engine, delta);
  }
  return set_links_host(h, h->hstage, h->ncomp, 0);
}

int su2gpu_coldstart(su2gpu_lattice *h) {
  CK(check(h));
  const size_t n = (size_t)h->nlinks * h->ncomp;
  CK(ensure_staging(h, n));
  for (long i = 0; i < h->nlinks; ++i) {
    if (h->group == SU2GPU_SU2) {
      h->hstage[4 * i] = 1.0;
      h->hstage[4 * i + 1] = h->hstage[4 * i + 2] = h->hstage[4 * i + 3] = 0.0;
    } else {
      h->hstage[i] = 0.0;
    }
  }
  return set_links_host(h, h->hstage, h->ncomp, 0);
}

int su2gpu_haarstart(su2gpu_lattice *h, uint64_t seed) {
  CK(check(h));
  const size_t n = (size_t)h->nlinks * h->ncomp;
  CK(ensure_staging(h, n));
  std::mt19937 engine(seed);
  if (h->group == SU2GPU_SU2) {
    // four normals per link in stream order (the one-by-one normal_distribution
    // draws, normals.hh), normalised
    su2_mi355x::detail::polar_normals(engine, h->hstage, n, h->hpairs);
    for (long i = 0; i < h->nlinks; ++i) {
      double *q = h->hstage + 4 * i;
      double r2 = 0.0;
      for (int k = 0; k < 4; ++k) r2 += q[k] * q[k];
      const double r = std::sqrt(r2);
      for (int k = 0; k < 4; ++k) q[k] = q[k] / r;
    }
  } else {
    for (long i = 0; i < h->nlinks; ++i) h->hstage[i] = uniform_ab(engine, -ref_pi(), ref_pi());
  }
  return set_links_host(h, h->hstage, h->ncomp, 0);
}

int su2gpu_random_gauge_trafo(su2gpu_lattice *h, uint64_t seed) {
  CK(check(h));
  // sequential over sites in the reference order (random_gauge_trafo.hh:15-31):
  // done on the host on the AoS links, then uploaded
  std::vector<double> U((size_t)h->nlinks * h->ncomp);
  CK(get_links_host(h, U.data(), h->ncomp, 0));
  std::mt19937 engine((int)seed);
  const long Ls[4] = {(long)h->Lt, (long)h->Lx, (long)h->Ly, (long)h->Lz};
  const int nd = h->nd;
  long s = 0;
  for (long t = 0; t < Ls[0]; ++t)
    for (long x = 0; x < Ls[1]; ++x)
      for (long y = 0; y < Ls[2]; ++y)
        for (long z = 0; z < Ls[3]; ++z, ++s) {
          long c[4] = {t, x, y, z};
          double rq[4], ru = 0.0;
          if (h->group == SU2GPU_SU2) random_su2(engine, 1.0, rq);
          else ru = random_u1(engine, 1.0);
          for (int mu = 0; mu < nd; ++mu) {
            long cm_[4] = {c[0], c[1], c[2], c[3]};
            cm_[mu] = (cm_[mu] - 1 + Ls[mu]) % Ls[mu];
            const long sm = ((cm_[0] * Ls[1] + cm_[1]) * Ls[2] + cm_[2]) * Ls[3] + cm_[3];
            const size_t i = (size_t)s * nd + mu, j = (size_t)sm * nd + mu;
            if (h->group == SU2GPU_SU2) {
              double o[4];
              su2_mul_host(rq, &U[4 * i], o);
              std::memcpy(&U[4 * i], o, sizeof(o));
              const double rd[4] = {rq[0], -rq[1], -rq[2], -rq[3]};
              su2_mul_host(&U[4 * j], rd, o);
              std::memcpy(&U[4 * j], o, sizeof(o));
            } else {
              U[i] = ru + U[i];
              U[j] = U[j] + (-ru);
            }
          }
        }
  return set_links_host(h, U.data(), h->ncomp, 0);
}

int su2gpu_restoreSU(su2gpu_lattice *h) {
  CK(check(h));
  {
    Prof pr(h, "restore");
    su2k::launch_restore(h->group, h->prec, h->U[h->cur], (int)h->nlinks, h->stream);
  }
  return last_launch("restore");
}

int su2gpu_retr_sum_plaquettes(su2gpu_lattice *h, double xi, int anisotropic, int spatial_only,
                               double *res) {
  CK(check(h));
  if (!res) return fail(SU2GPU_EINVAL, "null pointer");
  const int flags = anisotropic ? su2k::F_ANISO : 0;
  CK(plaq_launch(h, h->cur, -1, anisotropic ? 1.0 / xi : 1.0, anisotropic ? xi : 1.0,
                 spatial_only ? 1 : 0, flags, 6));
  CK(fetch_sums(h));
  *res = (double)sum_of(h, 6);
  return SU2GPU_OK;
}

int su2gpu_gauge_energy(su2gpu_lattice *h, int spatial_only, double *res) {
  return su2gpu_retr_sum_plaquettes(h, 1.0, 0, spatial_only, res);
}

int su2gpu_gauge_action(su2gpu_lattice *h, double *S) {
  CK(check(h));
  if (!S) return fail(SU2GPU_EINVAL, "null pointer");
  double r;
  CK(su2gpu_retr_sum_plaquettes(h, h->xi, h->aniso ? 1 : 0, 0, &r));
  *S = -h->beta * r / Nc(h);
  return SU2GPU_OK;
}

int su2gpu_kinetic_energy(su2gpu_lattice *h, double *K) {
  CK(check(h));
  if (!K) return fail(SU2GPU_EINVAL, "null pointer");
  CK(sumsq_launch(h, h->P, 7));
  CK(fetch_sums(h));
  *K = (double)(0.5L * sum_of(h, 7));
  return SU2GPU_OK;
}

int su2gpu_zero_derivative(su2gpu_lattice *h) {
  CK(check(h));
  CK(ensure_deriv(h));
  HC(hipMemsetAsync(h->deriv, 0, h->P_bytes, h->stream));
  return SU2GPU_OK;
}

int su2gpu_derivative_device(su2gpu_lattice *h, double fac) {
  CK(check(h));
  CK(ensure_deriv(h));
  su2k::FusedCoef c = base_coef(h);
  c.coef = fac * h->beta / Nc(h);  // fac * beta / Nc * get_deriv (flat-gaugemonomial.hh:123-124)
  c.ep = -1.0;                     // deriv += coef * d
  return fused(h, h->cur, -1, h->deriv, c, -1);
}

int su2gpu_get_derivative(su2gpu_lattice *h, double *deriv) {
  CK(check(h));
  if (!deriv) return fail(SU2GPU_EINVAL, "null destination");
  CK(ensure_deriv(h));
  return get_mom_host(h, h->deriv, deriv);
}

int su2gpu_derivative(su2gpu_lattice *h, double fac, double *deriv) {
  CK(check(h));
  if (!deriv) return fail(SU2GPU_EINVAL, "null deriv");
  CK(ensure_deriv(h));
  CK(set_mom_host(h, h->deriv, deriv));
  CK(su2gpu_derivative_device(h, fac));
  return get_mom_host(h, h->deriv, deriv);
}

int su2gpu_update_momenta(su2gpu_lattice *h, double dtau) {
  CK(check(h));
  su2k::FusedCoef c = base_coef(h);
  c.ep = dtau;
  return fused(h, h->cur, -1, h->P, c, -1);
}

int su2gpu_update_gauge(su2gpu_lattice *h, double dtau) {
  CK(check(h));
  if (h->pending) return fail(SU2GPU_ESTATE, "md_trajectory pending");
  const int flags = (h->prec == 1 && h->group == 1) ? su2k::F_RESTORE : 0;
  {
    Prof pr(h, "update_gauge");
    su2k::launch_update_gauge(h->group, h->prec, h->U[h->cur], h->P, h->V, h->nd, dtau, 0.0, flags,
                              h->stream);
  }
  return last_launch("update_gauge");
}

int su2gpu_integrate(su2gpu_lattice *h, int integrator, size_t nsteps, double tau, int restore,
                     size_t exponent) {
  CK(check(h));
  if (h->pending) return fail(SU2GPU_ESTATE, "md_trajectory pending");
  IntegrateOut io;
  CK(run_integrator(h, integrator, nsteps, tau, exponent, {}, false, 0, &io));
  int fin = io.final_buf;
  CK(to_record(h, &fin, {}));
  h->cur = fin;
  if (restore) CK(su2gpu_restoreSU(h));
  return SU2GPU_OK;
}

// ---------------------------------------------------------------- md_update --
// partial rows, rbase + : 0..2 first fused step, 3..5 last momentum step,
// 6 S_old (FP32), 7 S_new, 8 K_new (fallback); rbase 0 for the trajectory,
// 12 for the reversibility run
static int trajectory_sums(L *h, int kind, size_t nsteps, double tau, size_t exponent,
                           std::initializer_list<int> protect, int rbase, int u_start,
                           long double *K_old, long double *S_old, long double *K_new,
                           long double *S_new, int *final_buf) {
  const int nbf = su2k::fused_blocks(h->geo);
  const double w_t = h->aniso ? 1.0 / h->xi : 1.0, w_s = h->aniso ? h->xi : 1.0;
  const int pflags = h->aniso ? su2k::F_ANISO : 0;
  h->cur = u_start;
  if (h->r32 & 2) CK(flush_gen(h));
  // FP32 storage emulation: the fields an FP32 store would hold at the start
  if (h->r32 & 1) su2k::launch_round_f32((double *)h->U[u_start], (size_t)h->nlinks * h->ncomp, h->stream);
  if (h->r32 & 2) su2k::launch_round_f32((double *)h->P, (size_t)h->nlinks * h->nm, h->stream);
  if (h->r32) CK(last_launch("round_f32"));
  su2k::RowList rl;  // every sum of the trajectory is reduced in one launch at the end
  // the FP32 form takes S_old as the FP64 plaquette sum of its FP32 links
  if (h->prec == 1) CK(plaq_launch(h, u_start, -1, w_t, w_s, 0, pflags, rbase + 6, &rl));
  IntegrateOut io;
  CK(run_integrator(h, kind, nsteps, tau, exponent, protect, true, rbase, &io));
  int cur = io.final_buf;
  // restoreSU at the end of integrate (integrator.hh:48,111) fused with the
  // accept step's action (md_update.hh:47-51): restored links go to a free
  // buffer.  The FP32 form reunitarises every link update already (F_RESTORE),
  // so its final links need no further restore.
  int out;
  if (h->group == 1 && h->prec == 0) {
    CK(free_buf(h, {cur, protect.size() > 0 ? *protect.begin() : cur,
                    protect.size() > 1 ? *(protect.begin() + 1) : cur},
                &out));
    CK(plaq_launch(h, cur, out, w_t, w_s, 0, pflags | su2k::F_RESTORE, rbase + 7, &rl));
    cur = out;
    if (h->r32 & 1) {
      // the restored links as FP32 would store them, and their action
      su2k::launch_round_f32((double *)h->U[cur], (size_t)h->nlinks * h->ncomp, h->stream);
      CK(last_launch("round_f32"));
      CK(plaq_launch(h, cur, -1, w_t, w_s, 0, pflags, rbase + 7, &rl));
    }
  } else {
    CK(plaq_launch(h, cur, -1, w_t, w_s, 0, pflags, rbase + 7, &rl));
  }
  if (!io.last_p_only) CK(sumsq_launch(h, h->P, rbase + 8, &rl));
  if (io.first_accum) CK(reduce_row(h, rbase + 0, nbf, &rl));
  if (io.first_accum && h->prec == 0) CK(reduce_row(h, rbase + 1, nbf, &rl));
  if (io.last_p_only) CK(reduce_row(h, rbase + 5, nbf, &rl));
  CK(reduce_rows(h, rl));
  CK(fetch_sums(h));
  *K_old = 0.5L * sum_of(h, rbase + 0);
  const long double sg = -(long double)h->beta / (long double)Nc(h);
  *S_old = (h->prec == 0) ? sg * (sum_of(h, rbase + 1) / 4.0L) : sg * sum_of(h, rbase + 6);
  *K_new = 0.5L * (io.last_p_only ? sum_of(h, rbase + 5) : sum_of(h, rbase + 8));
  *S_new = sg * sum_of(h, rbase + 7);
  *final_buf = cur;
  return SU2GPU_OK;
}

int su2gpu_md_trajectory(su2gpu_lattice *h, const double *momenta, int integrator, size_t nsteps,
                         double tau, size_t exponent, int revtest, su2gpu_md_result *res) {
  CK(check(h));
  if (!res) return fail(SU2GPU_EINVAL, "null result");
  if (h->pending) return fail(SU2GPU_ESTATE, "previous trajectory not finished");
  if (momenta) CK(set_mom_host(h, h->P, momenta));
  const int u0 = h->cur;
  long double Ko, So, Kn, Sn;
  int unew;
  CK(trajectory_sums(h, integrator, nsteps, tau, exponent, {u0}, 0, u0, &Ko, &So, &Kn, &Sn, &unew));
  std::memset(res, 0, sizeof(*res));
  res->kin_old = (double)Ko;
  res->gauge_old = (double)So;
  res->kin_new = (double)Kn;
  res->gauge_new = (double)Sn;
  res->deltaH = (double)((Kn - Ko) + (Sn - So));
  res->deltadeltaH = NAN;
  res->uniform = -1.0;
  res->accept = 1;
  if (revtest) {
    // md_update.hh:62-72: flip P, integrate back from U_new, H - H_old
    CK(su2gpu_flip_momenta(h));
    long double Kr_o, Sr_o, Kr, Sr;
    int urev;
    CK(trajectory_sums(h, integrator, nsteps, tau, exponent, {u0, unew}, 12, unew, &Kr_o, &Sr_o,
                       &Kr, &Sr, &urev));
    res->deltadeltaH = (double)((Kr - Ko) + (Sr - So));
    res->revtest = 1;
  }
  h->cur = u0;
  h->pending = true;
  h->pend_old = u0;
  h->pend_new = unew;
  return SU2GPU_OK;
}

int su2gpu_md_finish(su2gpu_lattice *h, int accept) {
  CK(check(h));
  if (!h->pending) return fail(SU2GPU_ESTATE, "no trajectory pending");
  h->cur = accept ? h->pend_new : h->pend_old;
  h->pending = false;
  return SU2GPU_OK;
}

int su2gpu_md_update(su2gpu_lattice *h, uint64_t seed, int integrator, size_t nsteps, double tau,
                     size_t exponent, int revtest, int momenta_source, su2gpu_md_result *res) {
  CK(check(h));
  if (!res) return fail(SU2GPU_EINVAL, "null result");
  std::mt19937 engine(seed);
  if (momenta_source == SU2GPU_MOMENTA_HOST_MT19937) {
    CK(ensure_staging(h, (size_t)h->nlinks * h->nm));
    host_normals(h, engine);
    CK(set_mom_host(h, h->P, h->hstage));
  } else if (momenta_source == SU2GPU_MOMENTA_DEVICE_PHILOX) {
    // drawn by the trajectory's first step where that walks t (the Philox
    // stream of launch_momenta_philox either way), else by their own launch
    h->gen_pending = true;
    h->gen_seed = seed;
    h->gen_traj = 0;
  } else {
    return fail(SU2GPU_EINVAL, "bad momenta source");
  }
  const int rt = su2gpu_md_trajectory(h, nullptr, integrator, nsteps, tau, exponent, revtest, res);
  if (h->gen_pending) {
    // not consumed (an error before the first step): P holds what it would
    h->gen_pending = false;
    CK(rt);
    return fail(SU2GPU_ESTATE, "device momenta not drawn");
  }
  CK(rt);
  int accept = 1;
  if (res->deltaH > 0) {
    double u;
    if (momenta_source == SU2GPU_MOMENTA_HOST_MT19937) {
      std::uniform_real_distribution<double> uniform(0., 1.);
      u = uniform(engine);
    } else {
      u = philox_accept_uniform(seed, 0);
    }
    res->uniform = u;
    if (u > std::exp(-res->deltaH)) accept = 0;
  }
  res->accept = accept;
  return su2gpu_md_finish(h, accept);
}

int su2gpu_energy_density(su2gpu_lattice *h, int clover, int ss, double *E, double *Q) {
  CK(check(h));
  if (!E || !Q) return fail(SU2GPU_EINVAL, "null pointer");
  const int nbf = su2k::site_blocks(h->geo);
  {
    Prof pr(h, "energy_density");
    su2k::launch_energy_density(h->group, h->prec, h->U[h->cur], h->geo, clover ? 1 : 0,
                                ss ? 1 : 0, row(h, 20), h->row_stride, h->stream);
  }
  CK(last_launch("energy_density"));
  su2k::RowList rl;
  CK(reduce_row(h, 20, nbf, &rl));
  CK(reduce_row(h, 21, nbf, &rl));
  CK(reduce_rows(h, rl));
  CK(fetch_sums(h));
  long double e = sum_of(h, 20), q = sum_of(h, 21);
  // flat-energy_density.hh:152-163
  double res = -(double)e / (double)h->V;
  if (clover) res /= 16.;
  double qq = -4. * (double)q / (32.0 * M_PI * M_PI);
  if (clover) qq /= 16.;
  if (h->nd == 3) qq = 0.0;
  *E = res;
  *Q = qq;
  return SU2GPU_OK;
}

}  // extern "C"

namespace {
// colour class of site (t, x, y, z) for the links of direction mu on a lattice
// with an odd extent: sum over nu != mu of h(x_nu) mod 3, h(k) = k + min(k, a),
// a = (3 - L_nu mod 3) mod 3 (each step along nu changes h by 1 or 2 mod 3
// and the periodic wrap by 1, so neighbours never share a class)
int sweep_colour3(const su2k::Geo &g, const int c[4], int mu) {
  int col = 0;
  for (int d = 0; d < g.nd; ++d) {
    if (d == mu) continue;
    const int a = (3 - g.L[d] % 3) % 3;
    col += c[d] + (c[d] < a ? c[d] : a);
  }
  return col % 3;
}

int ensure_sweep_lists(L *h) {
  if (h->sweep_list) return SU2GPU_OK;
  const su2k::Geo &g = h->geo;
  const int nd = h->nd, V = h->V;
  std::vector<int> cnt((size_t)3 * nd, 0);
  std::vector<int> col((size_t)nd * V);
  int s = 0;
  for (int t = 0; t < g.L[0]; ++t)
    for (int x = 0; x < g.L[1]; ++x)
      for (int y = 0; y < g.L[2]; ++y)
        for (int z = 0; z < g.L[3]; ++z, ++s) {
          const int c[4] = {t, x, y, z};
          for (int mu = 0; mu < nd; ++mu) {
            const int k = sweep_colour3(g, c, mu);
            col[(size_t)mu * V + s] = k;
            ++cnt[3 * mu + k];
          }
        }
  h->sweep_off.assign((size_t)3 * nd + 1, 0);
  for (int i = 0; i < 3 * nd; ++i) h->sweep_off[i + 1] = h->sweep_off[i] + cnt[i];
  std::vector<int> list((size_t)nd * V), pos(h->sweep_off.begin(), h->sweep_off.end() - 1);
  for (int mu = 0; mu < nd; ++mu)
    for (int site = 0; site < V; ++site) list[pos[3 * mu + col[(size_t)mu * V + site]]++] = site;
  HC(hipMalloc(&h->sweep_list, list.size() * sizeof(int)));
  HC(hipMemcpy(h->sweep_list, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice));
  return SU2GPU_OK;
}
}  // namespace

extern "C" {

int su2gpu_sweep(su2gpu_lattice *h, uint64_t seed, uint64_t sweep, double delta, size_t n_hit,
                 double *rate, double *rate_time) {
  CK(check(h));
  if (!rate || !rate_time) return fail(SU2GPU_EINVAL, "null pointer");
  if (h->pending) return fail(SU2GPU_ESTATE, "md_trajectory pending");
  if (n_hit > (size_t)INT32_MAX) return fail(SU2GPU_EINVAL, "sweep: n_hit too large");
  // every extent in use even: the two parity classes; else three colour classes
  int ncol = 2;
  for (int mu = 0; mu < h->nd; ++mu)
    if (h->geo.L[mu] % 2 != 0) ncol = 3;
  if (ncol == 3) CK(ensure_sweep_lists(h));
  // one row of accepted-hit counts per sub-sweep (mu, class), reduced together
  constexpr int R0 = 22;
  static_assert(R0 + 12 <= NROWS, "partial rows");
  const double w_t = h->aniso ? 1.0 / h->xi : 1.0, w_s = h->aniso ? h->xi : 1.0;
  su2k::RowList rl;
  for (int mu = 0; mu < h->nd; ++mu)
    for (int par = 0; par < ncol; ++par) {
      const int r = R0 + ncol * mu + par;
      const int *list = nullptr;
      int nlist = -1;
      if (ncol == 3) {
        list = h->sweep_list + h->sweep_off[3 * mu + par];
        nlist = h->sweep_off[3 * mu + par + 1] - h->sweep_off[3 * mu + par];
      }
      const int nb = su2k::sweep_blocks(h->geo, nlist);
      if (nb == 0) continue;
      {
        Prof pr(h, "sweep");
        su2k::launch_sweep(h->group, h->prec, h->U[h->cur], h->geo, mu, par, list, nlist, seed,
                           sweep, delta, (int)n_hit, h->beta / Nc(h), w_t, w_s, h->aniso ? 1 : 0,
                           row(h, r), h->stream);
      }
      CK(last_launch("sweep"));
      CK(reduce_row(h, r, nb, &rl));
    }
  CK(reduce_rows(h, rl));
  CK(fetch_sums(h));
  long double all = 0.0L, temporal = 0.0L;
  for (int i = 0; i < rl.n; ++i) {
    const long double a = sum_of(h, rl.row[i]);
    all += a;
    if (rl.row[i] < R0 + ncol) temporal += a;  // mu = 0
  }
  const long double nh = n_hit > 0 ? (long double)n_hit : 1.0L;
  *rate = (double)(all / nh / (long double)h->nlinks);
  *rate_time = (double)(temporal / nh / (long double)h->V);
  return SU2GPU_OK;
}

int su2gpu_runge_kutta(su2gpu_lattice *h, double eps) {
  CK(check(h));
  if (h->pending) return fail(SU2GPU_ESTATE, "md_trajectory pending");
  CK(ensure_deriv(h));
  // flat-gradient_flow.hh:17-50 with Z in the deriv field
  const double zfac[3] = {(-17.0) / (36.0), (8.0) / (9.0), (-3.0) / (4.0)};
  const double expfac[3] = {-36.0 / 4. / 17.0, 1., -1.};
  for (int f = 0; f < 3; ++f) {
    su2k::FusedCoef c = base_coef(h);
    if (h->prec == 1 && h->group == 1) c.flags |= su2k::F_RESTORE;
    const double fac = 2. * Nc(h) * zfac[f] / h->beta;
    c.coef = fac * h->beta / Nc(h);
    c.ep = -1.0;
    c.eu = -eps * expfac[f];
    if (f == 0) c.flags |= su2k::F_PZERO;
    int out;
    CK(free_buf(h, {h->cur}, &out));
    CK(fused(h, h->cur, out, h->deriv, c, -1));
    h->cur = out;
  }
  return SU2GPU_OK;
}

int su2gpu_gradient_flow(su2gpu_lattice *h, double tmax, double eps, double tstart, double *rows,
                         size_t max_rows, size_t *nrows) {
  CK(check(h));
  if (!rows || !nrows) return fail(SU2GPU_EINVAL, "null pointer");
  if (h->pending) return fail(SU2GPU_ESTATE, "md_trajectory pending");
  // the flow runs on a copy Vt of the links (flat-gradient_flow.hh:116-118)
  const int orig = h->cur;
  int vt;
  CK(free_buf(h, {orig}, &vt));
  HC(hipMemcpyAsync(h->U[vt], h->U[orig], h->U_bytes, hipMemcpyDeviceToDevice, h->stream));
  h->half[vt] = 0;
  // ping-pong on the two buffers other than orig
  const int d = h->nd;
  const double ndf = d * (d - 1) / 2.0, ndf_ss = (d - 1) * (d - 2) / 2.0;
  const double den = (double)h->V * Nc(h);
  double t[3], P[3] = {0, 0, 0}, E[3] = {0, 0, 0}, Q[3] = {0, 0, 0}, Pss[3] = {0, 0, 0},
               Ess[3] = {0, 0, 0};
  for (int i = 0; i < 3; ++i) t[i] = tstart;
  int rc = SU2GPU_OK;
  auto measure = [&](int k, bool withQ) -> int {
    double g, gss, e, q, ess, qss;
    CK(su2gpu_gauge_energy(h, 0, &g));
    CK(su2gpu_gauge_energy(h, 1, &gss));
    CK(su2gpu_energy_density(h, 1, 0, &e, &q));
    CK(su2gpu_energy_density(h, 1, 1, &ess, &qss));
    P[k] = g / den;
    Pss[k] = gss / den;
    E[k] = e;
    Ess[k] = ess;
    if (withQ) Q[k] = q;
    return SU2GPU_OK;
  };
  h->cur = vt;
  size_t n = 0;
  rc = measure(2, false);
  while (rc == SU2GPU_OK && t[1] < tmax - eps) {
    t[0] = t[2]; P[0] = P[2]; E[0] = E[2]; Q[0] = Q[2]; Pss[0] = Pss[2]; Ess[0] = Ess[2];
    for (int k = 1; k < 3 && rc == SU2GPU_OK; ++k) {
      t[k] = t[k - 1] + eps;
      // keep orig untouched: runge_kutta picks free buffers other than cur; make
      // sure orig is never chosen by temporarily marking it as pending-protected
      rc = [&]() -> int {
        CK(ensure_deriv(h));
        const double zfac[3] = {(-17.0) / (36.0), (8.0) / (9.0), (-3.0) / (4.0)};
        const double expfac[3] = {-36.0 / 4. / 17.0, 1., -1.};
        for (int f = 0; f < 3; ++f) {
          su2k::FusedCoef c = base_coef(h);
          if (h->prec == 1 && h->group == 1) c.flags |= su2k::F_RESTORE;
          const double fac = 2. * Nc(h) * zfac[f] / h->beta;
          c.coef = fac * h->beta / Nc(h);
          c.ep = -1.0;
          c.eu = -eps * expfac[f];
          if (f == 0) c.flags |= su2k::F_PZERO;
          int out;
          CK(free_buf(h, {h->cur, orig}, &out));
          CK(fused(h, h->cur, out, h->deriv, c, -1));
          h->cur = out;
        }
        return SU2GPU_OK;
      }();
      if (rc == SU2GPU_OK) rc = measure(k, true);
    }
    if (rc != SU2GPU_OK) break;
    const double Ep = ndf - P[1], Ep_ss = ndf_ss - Pss[1];
    const double xi2 = ((Ep - Ep_ss) / (d - 1)) / (Ep_ss / ndf_ss);
    if (n < max_rows) {
      double *r = rows + 9 * n;
      r[0] = t[1]; r[1] = std::sqrt(xi2); r[2] = P[1]; r[3] = Pss[1]; r[4] = Ep; r[5] = Ep_ss;
      r[6] = E[1]; r[7] = Ess[1]; r[8] = Q[1];
    }
    ++n;
  }
  h->cur = orig;
  *nrows = n < max_rows ? n : max_rows;
  return rc;
}

// ------------------------------------------------------ gradient-flow adjoint --
namespace {

int flow_alloc(L *h, void **p, size_t bytes) {
  if (!*p) HC(hipMalloc(p, bytes));
  return SU2GPU_OK;
}

// coefficients of one RK3 step (flat-gradient_flow.hh:17-50): stage f adds
// b_f D(W) to Z (b_f = fac_f * beta / Nc with fac_f = 2 Nc z_f / beta) and
// multiplies W by exp(c_f Z), c_f = -eps * expfac_f
void rk_coef(const L *h, double eps, double b[3], double c[3]) {
  const double zfac[3] = {(-17.0) / (36.0), (8.0) / (9.0), (-3.0) / (4.0)};
  const double expfac[3] = {-36.0 / 4. / 17.0, 1., -1.};
  for (int f = 0; f < 3; ++f) {
    const double fac = 2. * Nc(h) * zfac[f] / h->beta;
    b[f] = fac * h->beta / Nc(h);
    c[f] = -eps * expfac[f];
  }
}

// one RK3 stage on raw buffers: Z (+)= b D(Win); Wout = exp(c Z) Win (Wout may be
// null); with Zout the new Z goes there and Z is only read
int rk_stage(L *h, const void *Win, void *Wout, void *Z, double b, double c, bool first,
             void *Zout = nullptr) {
  su2k::FusedCoef k = base_coef(h);
  if (h->prec == 1 && h->group == 1) k.flags |= su2k::F_RESTORE;
  k.coef = b;
  k.ep = -1.0;
  k.eu = c;
  if (first) k.flags |= su2k::F_PZERO;
  return fused_ptr(h, Win, Wout, Z, k, -1, Zout);
}

// full RK3 step Win -> Wout using tmp as the intermediate buffer
int rk_step(L *h, const void *Win, void *Wout, void *tmp, void *Z, double eps) {
  double b[3], c[3];
  rk_coef(h, eps, b, c);
  CK(rk_stage(h, Win, Wout, Z, b[0], c[0], true));
  CK(rk_stage(h, Wout, tmp, Z, b[1], c[1], false));
  CK(rk_stage(h, tmp, Wout, Z, b[2], c[2], false));
  return SU2GPU_OK;
}

}  // namespace

namespace {
// buffers for the stage states of one segment (seg: W_k, stU: W1 W2, stZ: the
// three Z' per step), kept when they fit in half of the free device memory
// and SU2GPU_FLOW_RECOMPUTE is not set; false: recompute per step instead
int flow_keep_buffers(L *h, bool *keep) {
  auto &f = h->flow;
  const char *rc = std::getenv("SU2GPU_FLOW_RECOMPUTE");
  *keep = !(rc && std::atoi(rc) != 0);
  if (f.seg.size() < f.ckpt) {
    size_t old = f.seg.size();
    f.seg.resize(f.ckpt, nullptr);
    for (size_t i = old; i < f.ckpt; ++i) HC(hipMalloc(&f.seg[i], h->U_bytes));
  }
  if (*keep && f.stU.size() < 2 * f.ckpt) {
    size_t fr = 0, tot = 0;
    HC(hipMemGetInfo(&fr, &tot));
    // the backward pass's own fields (W2, Zs[2], lamW, lamZ, the 36 V scratch
    // R and the 18 V clover cotangent Gam) count against the same free memory
    // when it has not allocated them yet
    size_t bwd = 0;
    if (!f.W2) bwd += h->U_bytes;
    for (void *p : {f.Zs[0], f.Zs[1], f.lamW, f.lamZ})
      if (!p) bwd += h->P_bytes;
    if (!f.R) bwd += (size_t)h->V * 36 * h->real_bytes;
    if (!f.Gam) bwd += (size_t)h->V * 18 * h->real_bytes;
    const size_t need = (2 * f.ckpt - f.stU.size()) * h->U_bytes +
                        (3 * f.ckpt - f.stZ.size()) * h->P_bytes + bwd;
    *keep = need < fr / 2;
    if (*keep) {
      while (f.stU.size() < 2 * f.ckpt) {
        f.stU.push_back(nullptr);
        HC(hipMalloc(&f.stU.back(), h->U_bytes));
      }
      while (f.stZ.size() < 3 * f.ckpt) {
        f.stZ.push_back(nullptr);
        HC(hipMalloc(&f.stZ.back(), h->P_bytes));
      }
    }
  }
  return SU2GPU_OK;
}

// step k0 + j of a segment stage by stage from seg[j], keeping W1, W2 and the
// three Z'; the new W goes to Wout
int flow_kept_step(L *h, size_t j, void *Wout, const double b[3], const double c[3]) {
  auto &f = h->flow;
  // stage 0 ignores its Z input and writes Z0' in place; stages 1 and 2 read
  // the previous Z' and write the new one to its own buffer
  CK(rk_stage(h, f.seg[j], f.stU[2 * j], f.stZ[3 * j], b[0], c[0], true));
  CK(rk_stage(h, f.stU[2 * j], f.stU[2 * j + 1], f.stZ[3 * j], b[1], c[1], false,
              f.stZ[3 * j + 1]));
  CK(rk_stage(h, f.stU[2 * j + 1], Wout, f.stZ[3 * j + 1], b[2], c[2], false, f.stZ[3 * j + 2]));
  return SU2GPU_OK;
}
}  // namespace

int su2gpu_flow_forward(su2gpu_lattice *h, double eps, size_t nsteps, size_t ckpt_every,
                        int observable, double *value) {
  CK(check(h));
  if (!value) return fail(SU2GPU_EINVAL, "null value");
  if (h->group != SU2GPU_SU2) return fail(SU2GPU_EINVAL, "the flow adjoint is implemented for SU(2)");
  if (observable != SU2GPU_OBS_CLOVER_E && observable != SU2GPU_OBS_WILSON_ACTION)
    return fail(SU2GPU_EINVAL, "bad observable");
  if (ckpt_every < 1) return fail(SU2GPU_EINVAL, "ckpt_every must be >= 1");
  if (h->pending) return fail(SU2GPU_ESTATE, "md_trajectory pending");
  auto &f = h->flow;
  const size_t nck = (nsteps + ckpt_every - 1) / ckpt_every;
  for (void *p : f.ckpts) HC(hipFree(p));
  f.ckpts.assign(nck, nullptr);
  for (auto &p : f.ckpts) HC(hipMalloc(&p, h->U_bytes));
  CK(flow_alloc(h, &f.W[0], h->U_bytes));
  CK(flow_alloc(h, &f.W[1], h->U_bytes));
  CK(flow_alloc(h, &f.W1, h->U_bytes));
  CK(flow_alloc(h, &f.Z, h->P_bytes));
  HC(hipMemcpyAsync(f.W[0], h->U[h->cur], h->U_bytes, hipMemcpyDeviceToDevice, h->stream));
  f.valid = false;
  f.kept_last = false;
  f.eps = eps;
  f.nsteps = nsteps;
  f.ckpt = ckpt_every;
  // the last segment, the first one the backward pass needs, runs stage by
  // stage into the kept-state buffers, so the backward does not recompute it
  bool keep = false;
  if (nsteps > 0) CK(flow_keep_buffers(h, &keep));
  const size_t klast = nck > 0 ? (nck - 1) * ckpt_every : 0;
  double rb[3], rc3[3];
  rk_coef(h, eps, rb, rc3);
  int a = 0;
  for (size_t k = 0; k < nsteps; ++k) {
    if (k % ckpt_every == 0)
      HC(hipMemcpyAsync(f.ckpts[k / ckpt_every], f.W[a], h->U_bytes, hipMemcpyDeviceToDevice, h->stream));
    if (keep && k >= klast) {
      const size_t j = k - klast;
      if (j == 0)
        HC(hipMemcpyAsync(f.seg[0], f.W[a], h->U_bytes, hipMemcpyDeviceToDevice, h->stream));
      if (k + 1 < nsteps) {
        CK(flow_kept_step(h, j, f.seg[j + 1], rb, rc3));
      } else {
        CK(flow_kept_step(h, j, f.W[1 - a], rb, rc3));
        a = 1 - a;
      }
    } else {
      CK(rk_step(h, f.W[a], f.W[1 - a], f.W1, f.Z, eps));
      a = 1 - a;
    }
  }
  f.kept_last = keep;
  f.fin = a;
  f.obs = observable;
  // observable at W(t): t^2 E_clover(t) or the Wilson action S (get_S_G_hmc)
  const int save = h->cur;
  int tmp;
  CK(free_buf(h, {save}, &tmp));
  HC(hipMemcpyAsync(h->U[tmp], f.W[a], h->U_bytes, hipMemcpyDeviceToDevice, h->stream));
  h->half[tmp] = 0;
  h->cur = tmp;
  int rc;
  const double t = eps * (double)nsteps;
  if (observable == SU2GPU_OBS_CLOVER_E) {
    double E, Q;
    rc = su2gpu_energy_density(h, 1, 0, &E, &Q);
    f.value = t * t * E;
  } else {
    double S;
    rc = su2gpu_gauge_action(h, &S);
    f.value = S;
  }
  h->cur = save;
  CK(rc);
  f.valid = true;
  *value = f.value;
  return SU2GPU_OK;
}

int su2gpu_flow_backward(su2gpu_lattice *h, double *grad) {
  CK(check(h));
  if (!grad) return fail(SU2GPU_EINVAL, "null grad");
  auto &f = h->flow;
  if (!f.valid) return fail(SU2GPU_ESTATE, "no flow_forward to differentiate");
  CK(flow_alloc(h, &f.W2, h->U_bytes));
  CK(flow_alloc(h, &f.Zs[0], h->P_bytes));
  CK(flow_alloc(h, &f.Zs[1], h->P_bytes));
  CK(flow_alloc(h, &f.lamW, h->P_bytes));
  CK(flow_alloc(h, &f.lamZ, h->P_bytes));
  CK(flow_alloc(h, &f.R, (size_t)h->V * 36 * h->real_bytes));
  const double w_t = h->aniso ? 1.0 / h->xi : 1.0, w_s = h->aniso ? h->xi : 1.0;
  // cotangent of W(t)
  const double t = f.eps * (double)f.nsteps;
  if (f.obs == SU2GPU_OBS_CLOVER_E) {
    CK(flow_alloc(h, &f.Gam, (size_t)h->V * 6 * 3 * h->real_bytes));
    {
      Prof pr(h, "flow_clover_grad");
      su2k::launch_clover_grad(h->prec, f.W[f.fin], f.Gam, f.lamW, h->geo,
                               t * t / (4.0 * (double)h->V), h->stream);
    }
    CK(last_launch("clover_grad"));
  } else {
    HC(hipMemsetAsync(f.lamW, 0, h->P_bytes, h->stream));
    su2k::FusedCoef k = base_coef(h);
    k.ep = -1.0;  // lamW += (beta/Nc) D(W) = grad S
    CK(fused_ptr(h, f.W[f.fin], nullptr, f.lamW, k, -1));
  }
  double b[3], c[3];
  rk_coef(h, f.eps, b, c);
  // The stage states of a whole segment (per step W1, W2 and Z0', Z1', Z2') are
  // kept from its forward recomputation when they fit in half of the free
  // device memory; otherwise every step recomputes its stages from W_k (the
  // same kernels on the same inputs: identical results, one more forward pass).
  // SU2GPU_FLOW_RECOMPUTE=1 forces the recomputation.  The forward pass left
  // the last segment's states in place (kept_last).
  bool keep = false;
  CK(flow_keep_buffers(h, &keep));
  // The gather of a stage's Hessian-vector product (the b/c role terms in R)
  // runs fused with the pointwise adjoint of the stage before it; the stage
  // recomputation in between touches neither R nor lamW / lamZ.
  bool pending = false;
  double pending_b = 0.0;
  const size_t nseg = f.ckpts.size();
  for (size_t sg = nseg; sg-- > 0;) {
    const size_t k0 = sg * f.ckpt, k1 = std::min(f.nsteps, k0 + f.ckpt);
    // recompute W_k, k0 <= k < k1, from the checkpoint
    const bool have = keep && f.kept_last && sg + 1 == nseg;
    if (!have)
      HC(hipMemcpyAsync(f.seg[0], f.ckpts[sg], h->U_bytes, hipMemcpyDeviceToDevice, h->stream));
    if (have) {
      // the forward pass left this segment's states in place
    } else if (keep) {
      // stage by stage, keeping W1, W2 and the three Z'
      for (size_t k = k0; k < k1; ++k) {
        const size_t j = k - k0;
        CK(flow_kept_step(h, j, k + 1 < k1 ? f.seg[j + 1] : nullptr, b, c));
      }
    } else {
      for (size_t k = k0; k + 1 < k1; ++k)
        CK(rk_step(h, f.seg[k - k0], f.seg[k - k0 + 1], f.W1, f.Z, f.eps));
    }
    for (size_t k = k1; k-- > k0;) {
      const size_t j = k - k0;
      const void *W0 = f.seg[j];
      const void *Wst[3] = {W0, f.W1, f.W2};
      const void *Zst[3] = {f.Zs[0], f.Zs[1], f.Z};
      if (keep) {
        Wst[1] = f.stU[2 * j];
        Wst[2] = f.stU[2 * j + 1];
        Zst[0] = f.stZ[3 * j];
        Zst[1] = f.stZ[3 * j + 1];
        Zst[2] = f.stZ[3 * j + 2];
      } else {
        // stage states: Z0' = b0 D(W0), W1 = exp(c0 Z0') W0, Z1' = Z0' + b1 D(W1),
        // W2 = exp(c1 Z1') W1, Z2' = Z1' + b2 D(W2)
        CK(rk_stage(h, W0, f.W1, f.Zs[0], b[0], c[0], true));
        CK(rk_stage(h, f.W1, f.W2, f.Zs[0], b[1], c[1], false, f.Zs[1]));
        CK(rk_stage(h, f.W2, nullptr, f.Zs[1], b[2], c[2], false, f.Z));
      }
      for (int st = 2; st >= 0; --st) {
        {
          Prof pr(h, "flow_stage_adj");
          if (pending)
            su2k::launch_flow_gather_stage_adj(h->prec, f.R, f.lamW, f.lamZ, Zst[st], h->geo,
                                               pending_b, c[st], st == 2 ? 1 : 0, h->stream);
          else
            su2k::launch_flow_stage_adj(h->prec, Zst[st], f.lamW, f.lamZ, h->V, h->nd, c[st],
                                        st == 2 ? 1 : 0, h->stream);
        }
        CK(last_launch("flow_stage_adj"));
        {
          Prof pr(h, "flow_vjp");
          su2k::launch_flow_vjp(h->prec, Wst[st], f.lamZ, f.lamW, f.R, h->geo, b[st], w_t, w_s,
                                false, h->stream);
        }
        CK(last_launch("flow_vjp"));
        pending = true;
        pending_b = b[st];
      }
    }
  }
  if (pending) {
    su2k::launch_flow_gather(h->prec, f.R, f.lamW, h->geo, pending_b, h->stream);
    CK(last_launch("flow_gather"));
  }
  // With more than one segment the buffers now hold the first segment's
  // states, not the last one's: a further backward of the same forward must
  // recompute the last segment.
  if (nseg > 1) f.kept_last = false;
  return get_mom_host(h, f.lamW, grad);
}

int su2gpu_save(su2gpu_lattice *h, const char *path) {
  CK(check(h));
  if (!path) return fail(SU2GPU_EINVAL, "null path");
  const size_t rec = (size_t)(h->ncomp + 1);
  std::vector<double> buf((size_t)h->nlinks * rec);
  CK(su2gpu_get_links(h, buf.data(), SU2GPU_LAYOUT_OBJECT));
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(SU2GPU_EIO, "cannot open %s for writing", path);
  const size_t w = std::fwrite(buf.data(), sizeof(double), buf.size(), f);
  std::fclose(f);
  if (w != buf.size()) return fail(SU2GPU_EIO, "short write to %s", path);
  return SU2GPU_OK;
}

int su2gpu_load(su2gpu_lattice *h, const char *path) {
  CK(check(h));
  if (!path) return fail(SU2GPU_EINVAL, "null path");
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(SU2GPU_EIO, "could not read file from %s", path);
  const size_t rec = (size_t)(h->ncomp + 1);
  std::vector<double> buf((size_t)h->nlinks * rec);
  const size_t r = std::fread(buf.data(), sizeof(double), buf.size(), f);
  std::fclose(f);
  if (r != buf.size()) return fail(SU2GPU_EIO, "%s: expected %zu bytes, read %zu", path,
                                   buf.size() * sizeof(double), r * sizeof(double));
  return su2gpu_set_links(h, buf.data(), SU2GPU_LAYOUT_OBJECT);
}

int su2gpu_copy_probe(size_t nbytes, int reps, double *ms_per_copy) {
  if (!ms_per_copy || nbytes < 16 || reps < 1) return fail(SU2GPU_EINVAL, "bad copy probe arguments");
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const long n16 = (long)(nbytes / 16);
  int rc = SU2GPU_OK;
  hipError_t e = hipMalloc(&a, (size_t)n16 * 16);
  if (e == hipSuccess) e = hipMalloc(&b, (size_t)n16 * 16);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemsetAsync(a, 1, (size_t)n16 * 16, st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    su2k::launch_copy16(a, b, n16, st);  // warm-up
    e = hipEventRecord(e0, st);
  }
  if (e == hipSuccess) {
    for (int r = 0; r < reps; ++r) su2k::launch_copy16(a, b, n16, st);
    e = hipEventRecord(e1, st);
  }
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e != hipSuccess) rc = fail(SU2GPU_EHIP, "copy probe: %s", hipGetErrorString(e));
  else *ms_per_copy = ms / reps;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  if (a) hipFree(a);
  if (b) hipFree(b);
  return rc;
}

int su2gpu_synchronize(su2gpu_lattice *h) {
  CK(check(h));
  HC(hipStreamSynchronize(h->stream));
  return SU2GPU_OK;
}

int su2gpu_profile_enable(su2gpu_lattice *h, int on) {
  CK(check(h));
  CK(collect_prof(h));
  h->prof = on != 0;
  return SU2GPU_OK;
}
int su2gpu_profile_reset(su2gpu_lattice *h) {
  CK(check(h));
  CK(collect_prof(h));
  h->prof_acc.clear();
  return SU2GPU_OK;
}
int su2gpu_profile_read(su2gpu_lattice *h, const char *name, double *ms, long *count) {
  CK(check(h));
  if (!name || !ms || !count) return fail(SU2GPU_EINVAL, "null pointer");
  CK(collect_prof(h));
  auto it = h->prof_acc.find(name);
  *ms = it == h->prof_acc.end() ? 0.0 : it->second.first;
  *count = it == h->prof_acc.end() ? 0 : it->second.second;
  return SU2GPU_OK;
}

}  // extern "C"
