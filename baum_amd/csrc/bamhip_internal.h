// bamhip_internal.h -- shared declarations of the HIP implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bamhip.h"

namespace bamhip {

// error state (bamhip_last_error)
int set_error(int code, const char *fmt, ...);
int check_hip(hipError_t e, const char *what);

#define BH_HIP(call)                                     \
  do {                                                   \
    hipError_t _e = (call);                              \
    if (_e != hipSuccess) return check_hip(_e, #call);   \
  } while (0)

hipStream_t stream();

// device-side box description (one BAM box)
struct DevBox {
  int m, n, o;
  long noffset;
  double x0, y0, z0;
  double dx, dy, dz;
  int skip[6];
  int bflag[6];
};

}  // namespace bamhip

struct bamhip_level {
  int l;
  int nboxes;
  int nvariables;
  long npoints;
  std::vector<bamhip::DevBox> boxes;
  std::vector<double *> v;   // device pointer per variable index
  int *d_flag;               // device scalar used for non-finite detection
  int check_mode = 1;        // 1: every RHS call synchronises and reports
};

namespace bamhip {

// event timers keyed by kernel family
struct TimerScope {
  TimerScope(const char *name);
  ~TimerScope();
  const char *name_;
  hipEvent_t a_, b_;
  bool on_;
};

// collect the 25 device pointers of an index list (returns error if missing)
int gather_ptrs(bamhip_level *lev, const int *idx, int n, double **out, const char *what);

}  // namespace bamhip
