/* puncture_oracle.c -- CPU restatement of BAM's moving-puncture tracker.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * interpolate_xyz_scalar (src/utility/interpolate/interpolate.c:419-470):
 * the cube of `order` points per direction whose origin is
 * xm + dx * (floor((x - xm + dequaleps) / dx) - (order - 1) / 2), xm being the
 * level's lowest coordinate; the value is 0 when no box of the level holds the
 * whole cube.  interpolate_TriN (interpolate.c:121-147): Lagrange sums in z,
 * then y, then x with coefficients_lagrange_N (lagrange.c:13-27) and
 * interpolate_lagrange_N (lagrange.c:30-39).
 *
 * track_moving_puncture_integrate (src/projects/Gauge/
 * moving_punctures_integrate.c:29-85): three ICN iterations
 * x0 = x1 - dt * (beta(x0) + beta_p(x1)) / 2, beta_p the shift of the previous
 * time level.  track_puncture (moving_punctures.c:174-185) around it: with
 * moving_puncture_fixz BH or BHNS the z coordinate stays; a new position that
 * lies in no box's bounding box (xyzinsidelevel, amr/points.c:96-133, dless
 * with dequaleps) is dropped and the old one kept.
 */
#include <math.h>

#include "oracle.h"

#define OB_DEQUALEPS 1e-10 /* bam_main.h:19 */

static void lagrange_coeffs_at(int n, double x, double xmin, double h, double *c)
{
  for (int i = 0; i < n; i++) {
    double d = 1;
    for (int j = 0; j < n; j++)
      if (j != i) d *= (x - xmin - j * h) / ((i - j) * h);
    c[i] = d;
  }
}

static double lagrange_sum(int n, const double *c, const double *u)
{
  double sum = 0;
  for (int i = 0; i < n; i++) sum += c[i] * u[i];
  return sum;
}

double ob_interp_point(int nb, const ob_box *boxes, const double *const *var, double x, double y, double z, int order,
                       int *found)
{
  double xm = boxes[0].x0, ym = boxes[0].y0, zm = boxes[0].z0;
  for (int b = 1; b < nb; b++) {
    if (boxes[b].x0 < xm) xm = boxes[b].x0;
    if (boxes[b].y0 < ym) ym = boxes[b].y0;
    if (boxes[b].z0 < zm) zm = boxes[b].z0;
  }
  const double dx = boxes[0].dx, dy = boxes[0].dy, dz = boxes[0].dz;
  const int n = (order - 1) / 2;
  const double xc = xm + dx * (floor((x - xm + OB_DEQUALEPS) / dx) - n);
  const double yc = ym + dy * (floor((y - ym + OB_DEQUALEPS) / dy) - n);
  const double zc = zm + dz * (floor((z - zm + OB_DEQUALEPS) / dz) - n);
  *found = 0;
  for (int b = 0; b < nb; b++) {
    const ob_box *B = &boxes[b];
    const double fi = (xc - B->x0) / dx, fj = (yc - B->y0) / dy, fk = (zc - B->z0) / dz;
    const int i0 = (int)floor(fi + 0.5), j0 = (int)floor(fj + 0.5), k0 = (int)floor(fk + 0.5);
    if (fabs(fi - i0) > 1e-6 || fabs(fj - j0) > 1e-6 || fabs(fk - k0) > 1e-6) continue;
    if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + order > B->m || j0 + order > B->n || k0 + order > B->o) continue;
    const double *u = var[b];
    double c[8], v[8][8], w[8], cube[8];
    lagrange_coeffs_at(order, z, zc, dz, c);
    for (int i = 0; i < order; i++)
      for (int j = 0; j < order; j++) {
        for (int k = 0; k < order; k++)
          cube[k] = u[(i0 + i) + (long)B->m * (j0 + j) + (long)B->m * B->n * (k0 + k)];
        v[i][j] = lagrange_sum(order, c, cube);
      }
    lagrange_coeffs_at(order, y, yc, dy, c);
    for (int i = 0; i < order; i++) w[i] = lagrange_sum(order, c, v[i]);
    lagrange_coeffs_at(order, x, xc, dx, c);
    *found = 1;
    return lagrange_sum(order, c, w);
  }
  return 0.0;
}

static int inside_level(int nb, const ob_box *boxes, double x, double y, double z)
{
  for (int b = 0; b < nb; b++) {
    const ob_box *B = &boxes[b];
    const double x1 = B->x0 + (B->m - 1) * B->dx, y1 = B->y0 + (B->n - 1) * B->dy, z1 = B->z0 + (B->o - 1) * B->dz;
    if (x < B->x0 - OB_DEQUALEPS || y < B->y0 - OB_DEQUALEPS || z < B->z0 - OB_DEQUALEPS || x1 < x - OB_DEQUALEPS ||
        y1 < y - OB_DEQUALEPS || z1 < z - OB_DEQUALEPS)
      continue;
    return 1;
  }
  return 0;
}

void ob_track_puncture(int nb, const ob_box *boxes, const double *const *beta, const double *const *beta_p, double dt,
                       int order, int flags, const double pold[3], double pnew[3])
{
  int f;
  const double x1 = pold[0], y1 = pold[1], z1 = pold[2];
  const double bx1 = ob_interp_point(nb, boxes, beta_p + 0 * nb, x1, y1, z1, order, &f);
  const double by1 = ob_interp_point(nb, boxes, beta_p + 1 * nb, x1, y1, z1, order, &f);
  const double bz1 = ob_interp_point(nb, boxes, beta_p + 2 * nb, x1, y1, z1, order, &f);
  double x0 = x1, y0 = y1, z0 = z1;
  for (int j = 0; j < 3; j++) {
    const double bx0 = ob_interp_point(nb, boxes, beta + 0 * nb, x0, y0, z0, order, &f);
    const double by0 = ob_interp_point(nb, boxes, beta + 1 * nb, x0, y0, z0, order, &f);
    const double bz0 = ob_interp_point(nb, boxes, beta + 2 * nb, x0, y0, z0, order, &f);
    x0 = x1 - dt * 0.5 * (bx0 + bx1);
    y0 = y1 - dt * 0.5 * (by0 + by1);
    z0 = z1 - dt * 0.5 * (bz0 + bz1);
  }
  if (flags & 1) z0 = z1;
  if (!inside_level(nb, boxes, x0, y0, z0)) {
    x0 = x1;
    y0 = y1;
    z0 = z1;
  }
  pnew[0] = x0;
  pnew[1] = y0;
  pnew[2] = z0;
}
