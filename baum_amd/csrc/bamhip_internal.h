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
hipStream_t side_stream();
int fork_side();
int join_side();

// device-side box description (one BAM box)
struct DevBox {
  int m, n, o;
  long noffset;
  double x0, y0, z0;
  double dx, dy, dz;
  int skip[6];
  int bflag[6];
};

// Index regions of one fine box for the FMR transfers, derived from the
// flag fields as prolong_varlist_boxtobox_N / restrict_varlist_boxtobox_N
// derive them (n == 0: nothing to do for this box).
struct ProlongRegion {
  int n;
  int ifbox[6], ifbox0[6], icbox[6];
};
struct RestrictRegion {
  int n;
  int icbox[6], ifbox[6];
};

// parent relation of a fine level (tL::prlocal, tB::pr) and the regions
// recorded by bamhip_set_rp_flags
struct FmrLink {
  struct bamhip_level *lc = nullptr;
  std::vector<int> parent;  // coarse box of each fine box
  int nbuffer = 0;          // amr_nbuffer: flag depth and prolongation zone
  int fp = -1, fr = -1;     // variable indices of flagprolong (fine), flagrestrict (coarse)
  // prolongation for nbuffer >= 3, == 1, < 0; restriction for Rflag 3, 2
  std::vector<ProlongRegion> pro[3];
  std::vector<RestrictRegion> res[2];
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
  bamhip::FmrLink fmr;       // set on a fine level by bamhip_set_rp_flags
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
// looks up the device address of the fixed reduction scratch of output.hip
// on `device` (current); called once per device by bamhip_init
int noalloc_scratch_init(int device);

}  // namespace bamhip
