// su2_handle.hpp -- the part of a lattice handle (struct su2gpu_lattice,
// su2_capi.cpp) that entry points in other translation units read.  Internal:
// not part of the C ABI.
#pragma once
#include "../../include/su2_mi355x.h"
#include "su2_launch.hpp"

namespace su2k {
struct LatticeView {
  int group, prec;
  Geo geo;
  const void *U;  // the current link field (device layout)
  hipStream_t stream;
  int device;
  bool pending;  // an md_trajectory awaits md_finish
};
}  // namespace su2k

int su2_internal_lattice_view(su2gpu_lattice *h, su2k::LatticeView *v);
extern "C" int su2gpu_internal_set_error(int code, const char *msg);
