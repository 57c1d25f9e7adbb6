/* output_oracle.c -- CPU restatement of the reductions and 1d point lists of
 * BAM's 0d / 1d output.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * bampi_allreduce (src/main/bampi/reduce.c:36-131): over the points whose
 * boundary flag is 0, variable by variable inside the point loop; the count
 * n grows once per point and variable.  write_level_0d (src/utility/output/
 * output0d.c:58-95): type 0 max, 1 min, 2 norm = sqrt(sum/n) * sqrt(nv),
 * 3 norminf = max |u|, 4 integral = sum * dx^3.
 * makeorderedlist1d (output1d.c:18-103) with the snap of 1doutinterpolate =
 * no, and the node search of bufferorderedpoints (src/main/amr/points.c:
 * 304-335): for every point of the line the first node at or after the
 * previous hit whose coordinates equal it within dequaleps.
 */
#include <float.h>
#include <math.h>

#include "oracle.h"

#define OB_EPS 1e-10

void ob_reduce_0d(int nb, const ob_box *boxes, const double *const *var, int nv, int type, double *result)
{
  double sum[64], sq[64], mx[64], mn[64], ma[64];
  double n = 0;
  for (int v = 0; v < nv; v++) sum[v] = sq[v] = 0, mx[v] = -DBL_MAX, mn[v] = DBL_MAX, ma[v] = -DBL_MAX;
  for (int b = 0; b < nb; b++) {
    const ob_box *B = &boxes[b];
    for (int k = 0; k < B->o; k++)
      for (int j = 0; j < B->n; j++)
        for (int i = 0; i < B->m; i++) {
          if (i < B->skip[0] || i >= B->m - B->skip[1] || j < B->skip[2] || j >= B->n - B->skip[3] ||
              k < B->skip[4] || k >= B->o - B->skip[5])
            continue;
          const long q = i + (long)B->m * j + (long)B->m * B->n * k;
          for (int v = 0; v < nv; v++) {
            const double u = var[v * nb + b][q];
            if (u > mx[v]) mx[v] = u;
            if (u < mn[v]) mn[v] = u;
            if (fabs(u) > ma[v]) ma[v] = fabs(u);
            sum[v] += u;
            sq[v] += u * u;
            n++;
          }
        }
  }
  for (int v = 0; v < nv; v++) {
    if (type == 0) result[v] = mx[v];
    if (type == 1) result[v] = mn[v];
    if (type == 2) result[v] = (n ? sqrt(sq[v] / n) : -1) * sqrt(nv);
    if (type == 3) result[v] = ma[v];
    if (type == 4) result[v] = sum[v] * pow(boxes[0].dx, 3);
  }
}

static int dless(double a, double b) { return a < b - OB_EPS; }
static int dequal(double a, double b) { return !(dless(a, b) || dless(b, a)); }

int ob_line_1d(int nb, const ob_box *boxes, int line, double x0, double y0, double z0, int maxp, double *coord,
               int *box_of, long *index)
{
  double bbox[6] = {1e300, -1e300, 1e300, -1e300, 1e300, -1e300};
  for (int b = 0; b < nb; b++) {
    const ob_box *B = &boxes[b];
    const double lo[3] = {B->x0, B->y0, B->z0};
    const double hi[3] = {B->x0 + (B->m - 1) * B->dx, B->y0 + (B->n - 1) * B->dy, B->z0 + (B->o - 1) * B->dz};
    for (int d = 0; d < 3; d++) {
      if (lo[d] < bbox[2 * d]) bbox[2 * d] = lo[d];
      if (hi[d] > bbox[2 * d + 1]) bbox[2 * d + 1] = hi[d];
    }
  }
  const double dx = boxes[0].dx, dy = boxes[0].dy, dz = boxes[0].dz;
  x0 = bbox[0] + floor((x0 - bbox[0]) / dx - 0.001) * dx + dx;
  y0 = bbox[2] + floor((y0 - bbox[2]) / dy - 0.001) * dy + dy;
  z0 = bbox[4] + floor((z0 - bbox[4]) / dz - 0.001) * dz + dz;
  const double s[3] = {line == 1 ? dx : 0, line == 2 ? dy : 0, line == 3 ? dz : 0};
  double p0[3] = {x0, y0, z0};
  int inside = 1;
  for (int d = 0; d < 3; d++) inside &= !dless(p0[d], bbox[2 * d]) && !dless(bbox[2 * d + 1], p0[d]);
  if (!inside) {
    for (int d = 0; d < 3; d++)
      if (s[d]) p0[d] = bbox[2 * d];
    inside = 1;
    for (int d = 0; d < 3; d++) inside &= !dless(p0[d], bbox[2 * d]) && !dless(bbox[2 * d + 1], p0[d]);
    if (!inside) return 0;
  }
  long m = 0, n = 0;
  for (;; m--) {
    int out = 0;
    for (int d = 0; d < 3; d++) out |= dless(p0[d] + m * s[d], bbox[2 * d]);
    if (out) break;
  }
  m++;
  for (;; n++) {
    int out = 0;
    for (int d = 0; d < 3; d++) out |= dless(bbox[2 * d + 1], p0[d] + n * s[d]);
    if (out) break;
  }
  n--;
  int np = (int)(n - m + 1);
  if (np > maxp) np = maxp;
  long last = 0;
  for (int p = 0; p < np; p++) {
    double x[3];
    for (int d = 0; d < 3; d++) x[d] = p0[d] + (m + p) * s[d];
    coord[p] = x[line - 1];
    box_of[p] = -1;
    index[p] = -1;
    /* scan the nodes forward from the last hit: boxes in order, ijk order */
    long base = 0, found = -1;
    int fb = -1;
    for (int b = 0; b < nb && found < 0; b++) {
      const ob_box *B = &boxes[b];
      const long nn = (long)B->m * B->n * B->o;
      for (long q = (base > last ? 0 : last - base); q < nn; q++) {
        const int i = (int)(q % B->m), j = (int)((q / B->m) % B->n), k = (int)(q / ((long)B->m * B->n));
        if (dequal(x[0], B->x0 + i * B->dx) && dequal(x[1], B->y0 + j * B->dy) && dequal(x[2], B->z0 + k * B->dz)) {
          found = base + q;
          fb = b;
          index[p] = q;
          break;
        }
      }
      base += nn;
    }
    if (found >= 0) {
      last = found;
      box_of[p] = fb;
    }
  }
  return np;
}
