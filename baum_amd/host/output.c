/* output.c -- 0d and 1d output of device-resident levels in the reference's
 * file formats.
 *
 * 0d (write_level_0d and write_scalar_N, src/utility/output/output0d.c:
 * 11-118): per variable one line "time  value  " (%22.15e each, two spaces
 * after every number) appended to <outdir>/<name>_<type>.l<l>, type one of
 * max, min, norm, norminf, integral (output.c:219), the value reduced over
 * the inner points of the level on the device.
 *
 * 1d (write_level_1d, output1d.c:300-420, with 1doutinterpolate = no, its
 * default, bam_output.c:40): the x, y or z line through the output origin
 * snapped to the next grid point (makeorderedlist1d, output1d.c:18-103),
 * spanning the level's bounding box; every point is looked up among the
 * level's nodes by its coordinates, scanning forward from the previous hit
 * as bufferorderedpoints does (src/main/amr/points.c:278-335), and points no
 * box holds are written as -1.  The block "Time = <t>" / "x  value" rows / a
 * blank line is appended to <outdir>/<name>.<x|y|z><l>.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/bamhip.h"

#define OUT_EPS 1e-10 /* dequaleps, bam_main.h:19 */

static const char *const k0d[5] = {"max", "min", "norm", "norminf", "integral"};
static const char *const k1d[4] = {"d", "x", "y", "z"};

int bamhip_output_0d(bamhip_level *lev, int nv, const int *vars, const char *const *names, int type, double time,
                     const char *outdir)
{
  int l, e, j;
  char file[1024];
  if (!lev || nv < 1 || !vars || !names || !outdir || type < 0 || type > 4) return BAMHIP_EINVAL;
  double *res = malloc(sizeof(double) * nv);
  if (!res) return BAMHIP_EINVAL;
  if ((e = bamhip_level_reduce(lev, nv, vars, type, res)) || (e = bamhip_level_info(lev, &l, 0))) {
    free(res);
    return e;
  }
  for (j = 0; j < nv; j++) {
    snprintf(file, sizeof file, "%s/%s_%s.l%d", outdir, names[j], k0d[type], l);
    FILE *fp = fopen(file, "a");
    if (!fp) {
      free(res);
      return BAMHIP_EINVAL;
    }
    fprintf(fp, "%22.15e  %22.15e  \n", time, res[j]);
    fclose(fp);
  }
  free(res);
  return BAMHIP_OK;
}

static int dless(double a, double b) { return a < b - OUT_EPS; }
static int dequal(double a, double b) { return !(dless(a, b) || dless(b, a)); }

/* the node index of (x, y, z) in box B (offset noff) or -1 */
static long node_of(const bamhip_box *B, long noff, double x, double y, double z)
{
  const long i = lround((x - B->x0) / B->dx), j = lround((y - B->y0) / B->dy), k = lround((z - B->z0) / B->dz);
  if (i < 0 || j < 0 || k < 0 || i >= B->m || j >= B->n || k >= B->o) return -1;
  if (!dequal(x, B->x0 + i * B->dx) || !dequal(y, B->y0 + j * B->dy) || !dequal(z, B->z0 + k * B->dz)) return -1;
  return noff + i + (long)B->m * j + (long)B->m * B->n * k;
}

int bamhip_output_1d(bamhip_level *lev, int nv, const int *vars, const char *const *names, int line, double x0,
                     double y0, double z0, double time, const char *outdir)
{
  int l, nb, b, e = BAMHIP_OK, j;
  long p, np, ilast = 0;
  char file[1024];
  if (!lev || nv < 1 || !vars || !names || !outdir || line < 1 || line > 3) return BAMHIP_EINVAL;
  if ((e = bamhip_level_info(lev, &l, &nb))) return e;
  bamhip_box *boxes = malloc(sizeof(bamhip_box) * nb);
  long *noff = malloc(sizeof(long) * nb);
  double bbox[6] = {1e300, -1e300, 1e300, -1e300, 1e300, -1e300};
  for (b = 0; b < nb; b++) {
    bamhip_level_box(lev, b, &boxes[b], &noff[b]);
    const bamhip_box *B = &boxes[b];
    const double lo[3] = {B->x0, B->y0, B->z0};
    const double hi[3] = {B->x0 + (B->m - 1) * B->dx, B->y0 + (B->n - 1) * B->dy, B->z0 + (B->o - 1) * B->dz};
    for (int d = 0; d < 3; d++) {
      if (lo[d] < bbox[2 * d]) bbox[2 * d] = lo[d];
      if (hi[d] > bbox[2 * d + 1]) bbox[2 * d + 1] = hi[d];
    }
  }
  const double dx = boxes[0].dx, dy = boxes[0].dy, dz = boxes[0].dz;
  /* snap to the next larger grid point (output1d.c:32-43) */
  x0 = bbox[0] + floor((x0 - bbox[0]) / dx - 0.001) * dx + dx;
  y0 = bbox[2] + floor((y0 - bbox[2]) / dy - 0.001) * dy + dy;
  z0 = bbox[4] + floor((z0 - bbox[4]) / dz - 0.001) * dz + dz;
  const double sx = line == 1 ? dx : 0, sy = line == 2 ? dy : 0, sz = line == 3 ? dz : 0;
#define INSIDE(x, y, z)                                                                                            \
  (!(dless((x), bbox[0]) || dless((y), bbox[2]) || dless((z), bbox[4]) || dless(bbox[1], (x)) ||                  \
     dless(bbox[3], (y)) || dless(bbox[5], (z))))
  if (!INSIDE(x0, y0, z0)) { /* off-centre box: the line starts at the box edge */
    if (sx) x0 = bbox[0];
    if (sy) y0 = bbox[2];
    if (sz) z0 = bbox[4];
    if (!INSIDE(x0, y0, z0)) goto done;
  }
  {
    const long nmax = lround((bbox[1] - bbox[0]) / dx + (bbox[3] - bbox[2]) / dy + (bbox[5] - bbox[4]) / dz) + 3;
    long m, n;
    for (m = 0; m > -nmax; m--)
      if (dless(x0 + m * sx, bbox[0]) || dless(y0 + m * sy, bbox[2]) || dless(z0 + m * sz, bbox[4])) break;
    m++;
    for (n = 0; n < nmax; n++)
      if (dless(bbox[1], x0 + n * sx) || dless(bbox[3], y0 + n * sy) || dless(bbox[5], z0 + n * sz)) break;
    n--;
    np = n - m + 1;
    if (np <= 0) goto done;
    double *crd = malloc(sizeof(double) * np), *val = malloc(sizeof(double) * np * nv);
    long *idx = malloc(sizeof(long) * np), *hit = malloc(sizeof(long) * np);
    long nhit = 0;
    for (p = 0; p < np; p++) {
      const double x = x0 + (m + p) * sx, y = y0 + (m + p) * sy, z = z0 + (m + p) * sz;
      crd[p] = line == 1 ? x : (line == 2 ? y : z);
      /* the first node at or after the previous hit with these coordinates */
      long best = -1;
      for (b = 0; b < nb; b++) {
        const long q = node_of(&boxes[b], noff[b], x, y, z);
        if (q >= ilast && (best < 0 || q < best)) best = q;
      }
      idx[p] = best;
      if (best >= 0) {
        ilast = best;
        hit[nhit++] = p;
      }
    }
    long *hidx = malloc(sizeof(long) * (nhit ? nhit : 1));
    double *hval = malloc(sizeof(double) * (nhit ? nhit : 1));
    for (p = 0; p < nhit; p++) hidx[p] = idx[hit[p]];
    for (j = 0; j < nv && !e; j++) {
      for (p = 0; p < np; p++) val[p * nv + j] = -1; /* not found */
      if ((e = bamhip_level_gather(lev, vars[j], nhit, hidx, hval))) break;
      for (p = 0; p < nhit; p++) val[hit[p] * nv + j] = hval[p];
      snprintf(file, sizeof file, "%s/%s.%s%d", outdir, names[j], k1d[line], l);
      FILE *fp = fopen(file, "a");
      if (!fp) {
        e = BAMHIP_EINVAL;
        break;
      }
      fprintf(fp, "\"Time = %f\"\n", time);
      for (p = 0; p < np; p++) fprintf(fp, "%22.15e  %22.15e\n", crd[p], val[p * nv + j]);
      fprintf(fp, "\n");
      fclose(fp);
    }
    free(crd), free(val), free(idx), free(hit), free(hidx), free(hval);
  }
#undef INSIDE
done:
  free(boxes);
  free(noff);
  return e;
}
