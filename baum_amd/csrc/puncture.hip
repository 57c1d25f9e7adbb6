// puncture.hip -- the moving-puncture tracker on the device.
//
// Reference: track_moving_puncture_integrate (src/projects/Gauge/
// moving_punctures_integrate.c:29-85): three ICN iterations
//   x0 = x1 - dt * (beta(x0) + beta_p(x1)) / 2,
// with the shift interpolated by interpolate_xyz_scalar
// (src/utility/interpolate/interpolate.c:419-470): a cube of `order` points
// per direction with origin xm + dx (floor((x - xm + dequaleps)/dx) -
// (order-1)/2), xm the level's lowest coordinate, value 0 when no box holds
// the whole cube; Lagrange sums in z, then y, then x (interpolate_TriN,
// interpolate.c:121-147; lagrange.c:13-39).  track_puncture
// (moving_punctures.c:174-185) around it: z kept when moving_puncture_fixz is
// BH or BHNS (flag bit 0), and a new position that lies in no box's bounding
// box (xyzinsidelevel, amr/points.c:96-133) is dropped for the old one.
//
// One workgroup per puncture: its lanes load each interpolation cube together
// through LDS.  The positions stay in a device array, so the tracker runs
// inside a captured coarse step with no host synchronisation.
// FP contraction is off: every product and sum rounds as in the reference's C.
#pragma clang fp contract(off)

#include <cmath>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr int kMaxPBoxes = 16;
constexpr int kMaxOrder = 8;
constexpr double kDequalEps = 1e-10;  // bam_main.h:19

struct PunctureArgs {
  int nb, order, np, flags;
  double xm, ym, zm, dx, dy, dz, dt;
  DevBox box[kMaxPBoxes];
  const double *beta[3], *beta_p[3];  // level arrays of the shift components
};

__device__ void lagrange_at(int n, double x, double xmin, double h, double *c) {
  for (int i = 0; i < n; i++) {
    double d = 1;
    for (int j = 0; j < n; j++)
      if (j != i) d *= (x - xmin - j * h) / ((i - j) * h);
    c[i] = d;
  }
}

constexpr int kThreads = 64;
constexpr int kCube = kMaxOrder * kMaxOrder * kMaxOrder;

struct TrackLds {
  double cube[3][kCube];                  // the interpolation cube of the three components
  double v[3][kMaxOrder * kMaxOrder];     // the sums over z
  double out[3];
};

// interpolate_xyz_scalar for the three components var[0..2] at (x, y, z),
// by the whole workgroup: the lanes load the cube together, the sums keep the
// reference's order (z for every (i, j), then y, then x, each from 0 upwards).
// Every lane returns the three values.
__device__ void interp3(const PunctureArgs &A, const double *const *var, double x, double y, double z, TrackLds &S,
                        double r[3]) {
  const int order = A.order, n = (order - 1) / 2, tid = threadIdx.x;
  const double xc = A.xm + A.dx * (floor((x - A.xm + kDequalEps) / A.dx) - n);
  const double yc = A.ym + A.dy * (floor((y - A.ym + kDequalEps) / A.dy) - n);
  const double zc = A.zm + A.dz * (floor((z - A.zm + kDequalEps) / A.dz) - n);
  int hit = -1, i0 = 0, j0 = 0, k0 = 0;
  for (int b = 0; b < A.nb && hit < 0; b++) {
    const DevBox &B = A.box[b];
    const double fi = (xc - B.x0) / A.dx, fj = (yc - B.y0) / A.dy, fk = (zc - B.z0) / A.dz;
    i0 = (int)floor(fi + 0.5);
    j0 = (int)floor(fj + 0.5);
    k0 = (int)floor(fk + 0.5);
    if (fabs(fi - i0) > 1e-6 || fabs(fj - j0) > 1e-6 || fabs(fk - k0) > 1e-6) continue;
    if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + order > B.m || j0 + order > B.n || k0 + order > B.o) continue;
    hit = b;
  }
  if (hit < 0) {  // uniform over the workgroup
    r[0] = r[1] = r[2] = 0.0;
    return;
  }
  const DevBox &B = A.box[hit];
  const long mn = (long)B.m * B.n, base = B.noffset + i0 + (long)B.m * j0 + mn * k0;
  const int o2 = order * order, o3 = o2 * order;
  for (int t = tid; t < o3; t += kThreads) {
    const int i = t % order, j = (t / order) % order, k = t / o2;
    const long idx = base + i + (long)B.m * j + mn * k;
    S.cube[0][t] = var[0][idx];
    S.cube[1][t] = var[1][idx];
    S.cube[2][t] = var[2][idx];
  }
  __syncthreads();
  double c[kMaxOrder];
  lagrange_at(order, z, zc, A.dz, c);
  for (int t = tid; t < 3 * o2; t += kThreads) {
    const int a = t / o2, ij = t % o2;
    const double *u = &S.cube[a][ij];
    double sum = 0;
    for (int k = 0; k < order; k++) sum += c[k] * u[k * o2];
    S.v[a][ij] = sum;  // ij = i + order j
  }
  __syncthreads();
  if (tid < 3) {
    double cy[kMaxOrder], cx[kMaxOrder];
    lagrange_at(order, y, yc, A.dy, cy);
    lagrange_at(order, x, xc, A.dx, cx);
    double res = 0;
    for (int i = 0; i < order; i++) {
      double w = 0;
      for (int j = 0; j < order; j++) w += cy[j] * S.v[tid][i + order * j];
      res += cx[i] * w;
    }
    S.out[tid] = res;
  }
  __syncthreads();
  r[0] = S.out[0];
  r[1] = S.out[1];
  r[2] = S.out[2];
  __syncthreads();  // S is reused by the next call
}

__device__ bool inside_level(const PunctureArgs &A, double x, double y, double z) {
  for (int b = 0; b < A.nb; b++) {
    const DevBox &B = A.box[b];
    const double x1 = B.x0 + (B.m - 1) * A.dx, y1 = B.y0 + (B.n - 1) * A.dy, z1 = B.z0 + (B.o - 1) * A.dz;
    if (x < B.x0 - kDequalEps || y < B.y0 - kDequalEps || z < B.z0 - kDequalEps || x1 < x - kDequalEps ||
        y1 < y - kDequalEps || z1 < z - kDequalEps)
      continue;
    return true;
  }
  return false;
}

// one workgroup per puncture
__global__ void __launch_bounds__(kThreads) track_kernel(PunctureArgs A, double *pos) {
  __shared__ TrackLds S;
  const int p = blockIdx.x;
  const double x1 = pos[3 * p], y1 = pos[3 * p + 1], z1 = pos[3 * p + 2];
  double b1[3], b0[3];
  interp3(A, A.beta_p, x1, y1, z1, S, b1);
  double x0 = x1, y0 = y1, z0 = z1;
  for (int j = 0; j < 3; j++) {
    interp3(A, A.beta, x0, y0, z0, S, b0);
    x0 = x1 - A.dt * 0.5 * (b0[0] + b1[0]);
    y0 = y1 - A.dt * 0.5 * (b0[1] + b1[1]);
    z0 = z1 - A.dt * 0.5 * (b0[2] + b1[2]);
  }
  if (A.flags & 1) z0 = z1;
  if (!inside_level(A, x0, y0, z0)) {
    x0 = x1;
    y0 = y1;
    z0 = z1;
  }
  if (threadIdx.x == 0) {
    pos[3 * p] = x0;
    pos[3 * p + 1] = y0;
    pos[3 * p + 2] = z0;
  }
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_puncture_track(bamhip_level *lev, const int *beta, const int *beta_p, double dt, int order,
                                     int flags, int np, double *pos) {
  const char *what = "bamhip_puncture_track";
  if (!lev || !beta || !beta_p || !pos) return set_error(BAMHIP_EINVAL, "%s: NULL argument", what);
  if (np < 1 || np > 64) return set_error(BAMHIP_EINVAL, "%s: %d punctures (1..64)", what, np);
  if (flags & ~BAMHIP_PUNCTURE_PIN_Z) return set_error(BAMHIP_EINVAL, "%s: unknown flags 0x%x", what, flags);
  if (order < 2 || order > kMaxOrder) return set_error(BAMHIP_EINVAL, "%s: order %d (2..%d)", what, order, kMaxOrder);
  const int nb = (int)lev->boxes.size();
  if (nb < 1 || nb > kMaxPBoxes) return set_error(BAMHIP_EINVAL, "%s: %d boxes (1..%d)", what, nb, kMaxPBoxes);
  PunctureArgs A{};
  A.nb = nb;
  A.order = order;
  A.np = np;
  A.flags = flags;
  A.dt = dt;
  A.dx = lev->boxes[0].dx;
  A.dy = lev->boxes[0].dy;
  A.dz = lev->boxes[0].dz;
  A.xm = lev->boxes[0].x0;
  A.ym = lev->boxes[0].y0;
  A.zm = lev->boxes[0].z0;
  for (int b = 0; b < nb; b++) {
    A.box[b] = lev->boxes[b];
    A.xm = fmin(A.xm, lev->boxes[b].x0);
    A.ym = fmin(A.ym, lev->boxes[b].y0);
    A.zm = fmin(A.zm, lev->boxes[b].z0);
  }
  for (int a = 0; a < 3; a++) {
    A.beta[a] = bamhip_level_ptr(lev, beta[a]);
    A.beta_p[a] = bamhip_level_ptr(lev, beta_p[a]);
    if (!A.beta[a] || !A.beta_p[a]) return set_error(BAMHIP_EINVAL, "%s: shift variable not enabled", what);
  }
  TimerScope t("puncture");
  track_kernel<<<np, kThreads, 0, stream()>>>(A, pos);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}
