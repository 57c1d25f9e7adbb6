// output.hip -- reductions and point gathers for 0d / 1d output.
//
// Reference: the reductions of write_level_0d (src/utility/output/
// output0d.c:28-118) through bampi_allreduce (src/main/bampi/reduce.c:36-131):
// max, min, the l2 norm, the l-infinity norm and the sum over the points
// whose boundary flag is 0 (the inner points of every box), and the point
// values that bufferorderedpoints (src/main/amr/points.c:278-360) collects for
// 1d output.  A reduction is deterministic: every box is split into a fixed
// number of blocks, each thread accumulates a fixed strided sequence, the
// block combines its threads in a fixed tree, and the host adds the block
// partials in block order.
#include <cfloat>
#include <cmath>
#include <vector>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr int kRBlock = 256;
constexpr int kRMaxBlocks = 1024;
constexpr int kRVals = 5;  // max, min, sum, sum of squares, max |u|

__global__ void __launch_bounds__(kRBlock) reduce_kernel(const double *u, DevBox B, long ninner, double *partial) {
  __shared__ double s[kRVals][kRBlock];
  const int ni = B.m - B.skip[0] - B.skip[1], nj = B.n - B.skip[2] - B.skip[3];
  double mx = -DBL_MAX, mn = DBL_MAX, sum = 0., sq = 0., ma = 0.;
  for (long t = (long)blockIdx.x * kRBlock + threadIdx.x; t < ninner; t += (long)gridDim.x * kRBlock) {
    const int i = (int)(t % ni) + B.skip[0], j = (int)((t / ni) % nj) + B.skip[2], k = (int)(t / ((long)ni * nj)) + B.skip[4];
    const double v = u[B.noffset + i + (long)B.m * j + (long)B.m * B.n * k];
    mx = v > mx ? v : mx;
    mn = v < mn ? v : mn;
    sum += v;
    sq += v * v;
    ma = fabs(v) > ma ? fabs(v) : ma;
  }
  const int tid = threadIdx.x;
  s[0][tid] = mx, s[1][tid] = mn, s[2][tid] = sum, s[3][tid] = sq, s[4][tid] = ma;
  __syncthreads();
  for (int w = kRBlock / 2; w > 0; w >>= 1) {
    if (tid < w) {
      s[0][tid] = fmax(s[0][tid], s[0][tid + w]);
      s[1][tid] = fmin(s[1][tid], s[1][tid + w]);
      s[2][tid] += s[2][tid + w];
      s[3][tid] += s[3][tid + w];
      s[4][tid] = fmax(s[4][tid], s[4][tid + w]);
    }
    __syncthreads();
  }
  if (tid < kRVals) partial[kRVals * blockIdx.x + tid] = s[tid][0];
}

__global__ void gather_kernel(const double *u, const long *idx, long n, double *out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = u[idx[t]];
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_level_info(const bamhip_level *lev, int *l, int *nboxes) {
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_level_info: NULL level");
  if (l) *l = lev->l;
  if (nboxes) *nboxes = (int)lev->boxes.size();
  return BAMHIP_OK;
}

extern "C" int bamhip_level_box(const bamhip_level *lev, int b, bamhip_box *box, long *noffset) {
  if (!lev || !box || b < 0 || b >= (int)lev->boxes.size())
    return set_error(BAMHIP_EINVAL, "bamhip_level_box: bad arguments");
  const DevBox &B = lev->boxes[b];
  box->m = B.m, box->n = B.n, box->o = B.o;
  box->x0 = B.x0, box->y0 = B.y0, box->z0 = B.z0;
  box->dx = B.dx, box->dy = B.dy, box->dz = B.dz;
  for (int f = 0; f < 6; f++) box->skip[f] = B.skip[f], box->bflag[f] = B.bflag[f];
  if (noffset) *noffset = B.noffset;
  return BAMHIP_OK;
}

extern "C" int bamhip_level_reduce(bamhip_level *lev, int nv, const int *vars, int type, double *result) {
  const char *what = "bamhip_level_reduce";
  if (!lev || nv < 1 || !vars || !result) return set_error(BAMHIP_EINVAL, "%s: bad arguments", what);
  if (type < 0 || type > 4) return set_error(BAMHIP_EINVAL, "%s: type %d (0 max, 1 min, 2 norm, 3 norminf, 4 integral)", what, type);
  std::vector<double *> u(nv);
  int e;
  if ((e = gather_ptrs(lev, vars, nv, u.data(), what))) return e;
  // block layout per box (fixed, so the sums are reproducible)
  std::vector<long> ninner(lev->boxes.size()), nblk(lev->boxes.size());
  long totblk = 0, npts = 0;
  for (size_t b = 0; b < lev->boxes.size(); b++) {
    const DevBox &B = lev->boxes[b];
    const long ni = B.m - B.skip[0] - B.skip[1], nj = B.n - B.skip[2] - B.skip[3], nk = B.o - B.skip[4] - B.skip[5];
    ninner[b] = (ni > 0 && nj > 0 && nk > 0) ? ni * nj * nk : 0;
    nblk[b] = ninner[b] ? std::min<long>(kRMaxBlocks, (ninner[b] + kRBlock - 1) / kRBlock) : 0;
    totblk += nblk[b];
    npts += ninner[b];
  }
  if (totblk == 0) return set_error(BAMHIP_EINVAL, "%s: the level has no inner points", what);
  double *partial = nullptr;
  BH_HIP(hipMallocAsync((void **)&partial, sizeof(double) * kRVals * totblk, stream()));
  std::vector<double> hp(kRVals * totblk);
  for (int v = 0; v < nv; v++) {
    long off = 0;
    for (size_t b = 0; b < lev->boxes.size(); b++) {
      if (!nblk[b]) continue;
      reduce_kernel<<<(unsigned)nblk[b], kRBlock, 0, stream()>>>(u[v], lev->boxes[b], ninner[b],
                                                                  partial + kRVals * off);
      off += nblk[b];
    }
    BH_HIP(hipGetLastError());
    BH_HIP(hipMemcpyAsync(hp.data(), partial, sizeof(double) * kRVals * totblk, hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    double mx = -DBL_MAX, mn = DBL_MAX, sum = 0., sq = 0., ma = 0.;
    for (long k = 0; k < totblk; k++) {
      const double *p = &hp[kRVals * k];
      mx = p[0] > mx ? p[0] : mx;
      mn = p[1] < mn ? p[1] : mn;
      sum += p[2];
      sq += p[3];
      ma = p[4] > ma ? p[4] : ma;
    }
    // reduce.c:36-131 with write_level_0d's corrections (output0d.c:70-95):
    // the l2 norm divides by the count over all nv variables, then * sqrt(nv);
    // integrals are sums times dx^3 of the level
    const double n = (double)npts * nv;
    if (type == 0) result[v] = mx;
    if (type == 1) result[v] = mn;
    if (type == 2) result[v] = sqrt(sq / n) * sqrt((double)nv);
    if (type == 3) result[v] = ma;
    if (type == 4) result[v] = sum * pow(lev->boxes[0].dx, 3);
  }
  BH_HIP(hipFreeAsync(partial, stream()));
  BH_HIP(hipStreamSynchronize(stream()));
  return BAMHIP_OK;
}

// ---- the reduction without allocation (bamhip_level_reduce_noalloc)
// Block partials live in `scratch`, a fixed device array (one copy per GPU,
// held until the process ends); its device address is looked up once per
// device when the library is initialised (noalloc_scratch_init, called by
// bamhip_init with the device current), and the host never passes the
// array's name itself to a kernel or a copy.  The host-side copy of the
// partials is a fixed array too.  Every slot a call reads is written by the
// same call's kernels (one slot per launched block), so nothing needs
// clearing; calls that overlap on one GPU share the array and race.
constexpr int kNoallocBlocks = 4096;   // at most 4096 reduction blocks per level
constexpr int kNoallocVars = 64;
constexpr int kNoallocBoxes = 64;
constexpr int kNoallocDevices = 16;
static __device__ __align__(256) double scratch[kRVals * kNoallocBlocks];
static double *g_scratch_dev[kNoallocDevices];
static double g_scratch_host[kRVals * kNoallocBlocks];

namespace bamhip {
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNoallocDevices)
    return set_error(BAMHIP_EINVAL, "bamhip_init: device %d beyond the %d supported by the reduction scratch", device,
                     kNoallocDevices);
  if (g_scratch_dev[device]) return BAMHIP_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_error(BAMHIP_EDEVICE, "bamhip_init: address of the reduction scratch: %s", hipGetErrorString(e));
  g_scratch_dev[device] = (double *)p;
  return BAMHIP_OK;
}
}  // namespace bamhip

extern "C" int bamhip_level_reduce_noalloc(bamhip_level *lev, int nv, const int *vars, int type, double *result) {
  const char *what = "bamhip_level_reduce_noalloc";
  if (!lev || nv < 1 || !vars || !result) return set_error(BAMHIP_EINVAL, "%s: bad arguments", what);
  if (type < 0 || type > 4) return set_error(BAMHIP_EINVAL, "%s: type %d (0 max, 1 min, 2 norm, 3 norminf, 4 integral)", what, type);
  if (nv > kNoallocVars) return set_error(BAMHIP_EINVAL, "%s: %d variables, at most %d", what, nv, kNoallocVars);
  const int nbox = (int)lev->boxes.size();
  if (nbox > kNoallocBoxes) return set_error(BAMHIP_EINVAL, "%s: %d boxes, at most %d", what, nbox, kNoallocBoxes);
  int dev = -1;
  BH_HIP(hipGetDevice(&dev));
  double *part = (dev >= 0 && dev < kNoallocDevices) ? g_scratch_dev[dev] : nullptr;
  if (!part) return set_error(BAMHIP_EINVAL, "%s: no reduction scratch on device %d (call bamhip_init)", what, dev);
  double *u[kNoallocVars];
  int e;
  if ((e = gather_ptrs(lev, vars, nv, u, what))) return e;
  // the block layout of bamhip_level_reduce, so both give the same sums
  long ninner[kNoallocBoxes], nblk[kNoallocBoxes];
  long totblk = 0, npts = 0;
  for (int b = 0; b < nbox; b++) {
    const DevBox &B = lev->boxes[b];
    const long ni = B.m - B.skip[0] - B.skip[1], nj = B.n - B.skip[2] - B.skip[3], nk = B.o - B.skip[4] - B.skip[5];
    ninner[b] = (ni > 0 && nj > 0 && nk > 0) ? ni * nj * nk : 0;
    nblk[b] = ninner[b] ? std::min<long>(kRMaxBlocks, (ninner[b] + kRBlock - 1) / kRBlock) : 0;
    totblk += nblk[b];
    npts += ninner[b];
  }
  if (totblk == 0) return set_error(BAMHIP_EINVAL, "%s: the level has no inner points", what);
  if (totblk > kNoallocBlocks)
    return set_error(BAMHIP_EINVAL, "%s: the level needs %ld reduction blocks, the fixed scratch holds %d", what, totblk,
                     kNoallocBlocks);
  for (int v = 0; v < nv; v++) {
    long off = 0;
    for (int b = 0; b < nbox; b++) {
      if (!nblk[b]) continue;
      reduce_kernel<<<(unsigned)nblk[b], kRBlock, 0, stream()>>>(u[v], lev->boxes[b], ninner[b], part + kRVals * off);
      off += nblk[b];
    }
    BH_HIP(hipGetLastError());
    BH_HIP(hipMemcpyAsync(g_scratch_host, part, sizeof(double) * kRVals * totblk, hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    double mx = -DBL_MAX, mn = DBL_MAX, sum = 0., sq = 0., ma = 0.;
    for (long k = 0; k < totblk; k++) {
      const double *p = &g_scratch_host[kRVals * k];
      mx = p[0] > mx ? p[0] : mx;
      mn = p[1] < mn ? p[1] : mn;
      sum += p[2];
      sq += p[3];
      ma = p[4] > ma ? p[4] : ma;
    }
    const double n = (double)npts * nv;
    if (type == 0) result[v] = mx;
    if (type == 1) result[v] = mn;
    if (type == 2) result[v] = sqrt(sq / n) * sqrt((double)nv);
    if (type == 3) result[v] = ma;
    if (type == 4) result[v] = sum * pow(lev->boxes[0].dx, 3);
  }
  return BAMHIP_OK;
}

extern "C" int bamhip_level_gather(bamhip_level *lev, int var, long n, const long *idx, double *out) {
  const char *what = "bamhip_level_gather";
  if (!lev || n < 0 || (n > 0 && (!idx || !out))) return set_error(BAMHIP_EINVAL, "%s: bad arguments", what);
  if (n == 0) return BAMHIP_OK;
  double *u;
  int e;
  if ((e = gather_ptrs(lev, &var, 1, &u, what))) return e;
  for (long t = 0; t < n; t++)
    if (idx[t] < 0 || idx[t] >= lev->npoints) return set_error(BAMHIP_EINVAL, "%s: index %ld out of range", what, idx[t]);
  long *didx = nullptr;
  double *dout = nullptr;
  BH_HIP(hipMallocAsync((void **)&didx, sizeof(long) * n, stream()));
  BH_HIP(hipMallocAsync((void **)&dout, sizeof(double) * n, stream()));
  BH_HIP(hipMemcpyAsync(didx, idx, sizeof(long) * n, hipMemcpyHostToDevice, stream()));
  gather_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream()>>>(u, didx, n, dout);
  BH_HIP(hipGetLastError());
  BH_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, stream()));
  BH_HIP(hipFreeAsync(didx, stream()));
  BH_HIP(hipFreeAsync(dout, stream()));
  BH_HIP(hipStreamSynchronize(stream()));
  return BAMHIP_OK;
}
