// su2_noalloc.hip -- the plaquette sum of a lattice handle into device memory,
// with no device allocation: for callers that cannot allocate (code under HIP
// graph capture, tight serving loops).
//
// The sum is the one of su2gpu_retr_sum_plaquettes (retr_sum_Wplaquettes,
// flat-gauge_energy.hpp:31-94): the same plaquette kernels write one partial
// per block and the same compensated reduction adds them.  Where that entry
// point keeps its partials in the handle's buffers and copies the result to
// the host (synchronising the stream), this one keeps them in `scratch`, a
// file-scope device array of SCRATCH_DOUBLES doubles (2 MiB, held by every GPU
// the library runs on until the process ends), and leaves the result in the
// caller's device memory, stream-ordered on the handle's stream.
#include <hip/hip_runtime.h>

#include <mutex>

#include "su2_handle.hpp"

namespace {

constexpr int SCRATCH_DOUBLES = 1 << 18;  // block partials + 2 reduction outputs
constexpr int MAX_PARTIALS = SCRATCH_DOUBLES - 2;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

// the device address of `scratch` on each device (host code must not pass the
// array's name itself to a kernel: on the host it is only a stand-in)
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int scratch_on(int device, double **out) {
  if (device < 0 || device >= MAX_DEVICES)
    return su2gpu_internal_set_error(SU2GPU_EINVAL, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return su2gpu_internal_set_error(SU2GPU_EHIP, "noalloc: hipGetSymbolAddress(scratch) failed");
    g_scratch[device] = (double *)p;
  }
  *out = g_scratch[device];
  return SU2GPU_OK;
}

// sum = hi + lo of the compensated reduction
__global__ void k_hi_lo(const double *__restrict__ hl, double *__restrict__ sum) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sum[0] = hl[0] + hl[1];
}

}  // namespace

extern "C" int su2gpu_retr_sum_plaquettes_noalloc(su2gpu_lattice *h, double xi, int anisotropic,
                                                  int spatial_only, double *sum_dev) {
  su2k::LatticeView v;
  int r = su2_internal_lattice_view(h, &v);
  if (r != SU2GPU_OK) return r;
  if (!sum_dev) return su2gpu_internal_set_error(SU2GPU_EINVAL, "null output");
  if (!(xi > 0)) return su2gpu_internal_set_error(SU2GPU_EINVAL, "xi must be > 0");
  const int nb_max = su2k::site_blocks(v.geo);
  if (nb_max > MAX_PARTIALS)
    return su2gpu_internal_set_error(
        SU2GPU_EINVAL, "noalloc: lattice too large for the scratch array (at most 67,108,352 sites)");
  double *sp = nullptr;
  r = scratch_on(v.device, &sp);
  if (r != SU2GPU_OK) return r;
  const int flags = anisotropic ? su2k::F_ANISO : 0;
  const double w_t = anisotropic ? 1.0 / xi : 1.0, w_s = anisotropic ? xi : 1.0;
  // partials in sp[0, nb), the reduction's hi / lo in sp[MAX_PARTIALS, +2);
  // every entry read is written first in the same call, so nothing is cleared
  const int nb = su2k::launch_plaq(v.group, v.prec, v.U, nullptr, v.geo, w_t, w_s,
                                   spatial_only ? 1 : 0, flags, sp, v.stream);
  su2k::RowList rl;
  rl.n = 1;
  rl.row[0] = 0;
  rl.nb[0] = nb;
  // the reduction writes out[2 row + {0,1}] for row 0
  su2k::launch_reduce(sp, 0, rl, sp + MAX_PARTIALS, v.stream);
  hipLaunchKernelGGL(k_hi_lo, dim3(1), dim3(64), 0, v.stream, sp + MAX_PARTIALS, sum_dev);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return su2gpu_internal_set_error(SU2GPU_EHIP, hipGetErrorString(e));
  return SU2GPU_OK;
}
