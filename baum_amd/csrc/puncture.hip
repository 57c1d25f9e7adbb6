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
// interpolate.c:121-147; lagrange.c:13-39).
//
// One thread per puncture.  The positions stay in a device array, so the
// tracker runs inside a captured coarse step with no host synchronisation.
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
  int nb, order, np;
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

__device__ double lagrange_sum(int n, const double *c, const double *u) {
  double sum = 0;
  for (int i = 0; i < n; i++) sum += c[i] * u[i];
  return sum;
}

__device__ double interp_point(const PunctureArgs &A, const double *var, double x, double y, double z) {
  const int order = A.order, n = (order - 1) / 2;
  const double xc = A.xm + A.dx * (floor((x - A.xm + kDequalEps) / A.dx) - n);
  const double yc = A.ym + A.dy * (floor((y - A.ym + kDequalEps) / A.dy) - n);
  const double zc = A.zm + A.dz * (floor((z - A.zm + kDequalEps) / A.dz) - n);
  for (int b = 0; b < A.nb; b++) {
    const DevBox &B = A.box[b];
    const double fi = (xc - B.x0) / A.dx, fj = (yc - B.y0) / A.dy, fk = (zc - B.z0) / A.dz;
    const int i0 = (int)floor(fi + 0.5), j0 = (int)floor(fj + 0.5), k0 = (int)floor(fk + 0.5);
    if (fabs(fi - i0) > 1e-6 || fabs(fj - j0) > 1e-6 || fabs(fk - k0) > 1e-6) continue;
    if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + order > B.m || j0 + order > B.n || k0 + order > B.o) continue;
    const double *u = var + B.noffset;
    double c[kMaxOrder], v[kMaxOrder][kMaxOrder], w[kMaxOrder], cube[kMaxOrder];
    lagrange_at(order, z, zc, A.dz, c);
    for (int i = 0; i < order; i++)
      for (int j = 0; j < order; j++) {
        for (int k = 0; k < order; k++) cube[k] = u[(i0 + i) + (long)B.m * (j0 + j) + (long)B.m * B.n * (k0 + k)];
        v[i][j] = lagrange_sum(order, c, cube);
      }
    lagrange_at(order, y, yc, A.dy, c);
    for (int i = 0; i < order; i++) w[i] = lagrange_sum(order, c, v[i]);
    lagrange_at(order, x, xc, A.dx, c);
    return lagrange_sum(order, c, w);
  }
  return 0.0;
}

__global__ void track_kernel(PunctureArgs A, double *pos) {
  const int p = threadIdx.x;
  if (p >= A.np) return;
  const double x1 = pos[3 * p], y1 = pos[3 * p + 1], z1 = pos[3 * p + 2];
  const double bx1 = interp_point(A, A.beta_p[0], x1, y1, z1);
  const double by1 = interp_point(A, A.beta_p[1], x1, y1, z1);
  const double bz1 = interp_point(A, A.beta_p[2], x1, y1, z1);
  double x0 = x1, y0 = y1, z0 = z1;
  for (int j = 0; j < 3; j++) {
    const double bx0 = interp_point(A, A.beta[0], x0, y0, z0);
    const double by0 = interp_point(A, A.beta[1], x0, y0, z0);
    const double bz0 = interp_point(A, A.beta[2], x0, y0, z0);
    x0 = x1 - A.dt * 0.5 * (bx0 + bx1);
    y0 = y1 - A.dt * 0.5 * (by0 + by1);
    z0 = z1 - A.dt * 0.5 * (bz0 + bz1);
  }
  pos[3 * p] = x0;
  pos[3 * p + 1] = y0;
  pos[3 * p + 2] = z0;
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_puncture_track(bamhip_level *lev, const int *beta, const int *beta_p, double dt, int order,
                                     int np, double *pos) {
  const char *what = "bamhip_puncture_track";
  if (!lev || !beta || !beta_p || !pos) return set_error(BAMHIP_EINVAL, "%s: NULL argument", what);
  if (np < 1 || np > 64) return set_error(BAMHIP_EINVAL, "%s: %d punctures (1..64)", what, np);
  if (order < 2 || order > kMaxOrder) return set_error(BAMHIP_EINVAL, "%s: order %d (2..%d)", what, order, kMaxOrder);
  const int nb = (int)lev->boxes.size();
  if (nb < 1 || nb > kMaxPBoxes) return set_error(BAMHIP_EINVAL, "%s: %d boxes (1..%d)", what, nb, kMaxPBoxes);
  PunctureArgs A{};
  A.nb = nb;
  A.order = order;
  A.np = np;
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
  track_kernel<<<1, 64, 0, stream()>>>(A, pos);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}
