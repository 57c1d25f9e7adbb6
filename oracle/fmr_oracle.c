/* fmr_oracle.c -- TEST INFRASTRUCTURE: CPU restatement of BAM's fixed mesh
 * refinement coupling: R/P flags, prolongation and restriction regions, and
 * the box-to-box Lagrange prolongation / restriction.  Checker only.
 *
 * Follows src/main/amr/flag.c:122-262 (setRPflags),
 * src/main/amr/box.c:284-309 (box_parent_index: nearest, box_point_index:
 * snap down), src/utility/interpolate/interpolate_boxtobox_N.c:585-865
 * (prolong_boxtobox_N: three 1-D sweeps z, y, x with weights at N/2-1/4 and,
 * shifted by one coarse point, N/2-3/4; the fbox0 interior is skipped per
 * x pair), :870-1050 (prolong_varlist_boxtobox_N: region from flagprolong),
 * :1055-1220 (restrict_boxtobox_N: three sweeps, weights at N/2-1/2, two fine
 * points per coarse point) and :1224-1370 (restrict_varlist_boxtobox_N).
 * The sweeps keep the reference's intermediate arrays, loop order and
 * summation order (sum = 0; sum += c[p]*u[p]).
 *
 * Levels are stored box after box (box b starts at the sum of the sizes of
 * boxes 0..b-1), ijk = i + m*j + m*n*k as in BAM.  Fine box b lies in coarse
 * box parent[b]; the reference's ghost parent box of a fine box is the part
 * of that coarse box the fine box covers, so the restriction region is
 * scanned over that footprint.
 */
#include <limits.h>
#include <stdlib.h>
#include "oracle.h"

static void lag(int n, double x, double *c)
{
  for (int i = 0; i < n; i++) {
    double d = 1;
    for (int j = 0; j < n; j++)
      if (j != i) d *= (x - j) / (double)(i - j);
    c[i] = d;
  }
}

static double interp(int n, const double *c, const double *u)
{
  double sum = 0;
  for (int i = 0; i < n; i++) sum += c[i] * u[i];
  return sum;
}

static long box_size(const ob_box *b) { return (long)b->m * b->n * b->o; }

static long level_offset(const ob_box *b, int ib)
{
  long off = 0;
  for (int i = 0; i < ib; i++) off += box_size(&b[i]);
  return off;
}

/* box_parent_index (box.c:284-287) and box_point_index (box.c:304-307) */
static int parent_index(double cx0, double cdx, double fx0, double fdx, int fi)
{
  return (int)(0.5 + (fx0 + fi * fdx - cx0) / cdx);
}

static int point_index(double X0, double dX, double x0, double dx, int i)
{
  return (int)((x0 + i * dx - X0) / dX);
}

/* ------------------------------------------------------------------------ */
/* prolong_boxtobox_N: cv, fv are box data with strides (1, cdj, cdk) */
void ob_prolong_box(int N, const int *cbox, const int *fbox, const int *fbox0, const double *cv, double *fv,
                    int cdj, int cdk, int fdj, int fdk)
{
  double c1[16], c2[16], u[16];
  lag(N, N / 2.0 - 0.25, c1);
  lag(N, N / 2.0 - 0.75, c2);
  /* z sweep: every (x, y) column of cbox to fine z */
  const int nx = cbox[1] - cbox[0] + 1, ny = cbox[3] - cbox[2] + 1;
  const int nzf = (cbox[5] - cbox[4] - N) * 2 + 2;
  double *dv = malloc(sizeof(double) * nx * ny * nzf);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const double *col = cv + (i + cbox[0]) + (long)(j + cbox[2]) * cdj + (long)cbox[4] * cdk;
      for (int t = 0; 2 * t < nzf; t++) {
        for (int p = 0; p < N; p++) u[p] = col[(long)(t + p) * cdk];
        dv[i + nx * (j + ny * (2 * t))] = interp(N, c1, u);
        for (int p = 0; p < N; p++) u[p] = col[(long)(t + 1 + p) * cdk];
        dv[i + nx * (j + ny * (2 * t + 1))] = interp(N, c2, u);
      }
    }
  /* y sweep */
  const int nyf = (cbox[3] - cbox[2] - N) * 2 + 2;
  double *ev = malloc(sizeof(double) * nx * nyf * nzf);
  for (int k = 0; k < nzf; k++)
    for (int i = 0; i < nx; i++)
      for (int t = 0; 2 * t < nyf; t++) {
        for (int p = 0; p < N; p++) u[p] = dv[i + nx * (t + p + ny * k)];
        ev[i + nx * (2 * t + nyf * k)] = interp(N, c1, u);
        for (int p = 0; p < N; p++) u[p] = dv[i + nx * (t + 1 + p + ny * k)];
        ev[i + nx * (2 * t + 1 + nyf * k)] = interp(N, c2, u);
      }
  /* x sweep into the fine box, skipping pairs whose first point is in fbox0 */
  for (int k = fbox[4]; k <= fbox[5]; k++)
    for (int j = fbox[2]; j <= fbox[3]; j++) {
      const double *row = ev + nx * ((j - fbox[2]) + nyf * (k - fbox[4]));
      for (int i = fbox[0], t = 0; i <= fbox[1]; i += 2, t++) {
        if (i >= fbox0[0] && i <= fbox0[1] && j >= fbox0[2] && j <= fbox0[3] && k >= fbox0[4] &&
            k <= fbox0[5])
          continue;
        const long q = i + (long)j * fdj + (long)k * fdk;
        for (int p = 0; p < N; p++) u[p] = row[t + p];
        fv[q] = interp(N, c1, u);
        for (int p = 0; p < N; p++) u[p] = row[t + 1 + p];
        fv[q + 1] = interp(N, c2, u);
      }
    }
  free(dv);
  free(ev);
}

/* restrict_boxtobox_N */
void ob_restrict_box(int N, const int *fbox, const int *cbox, const double *fv, double *cv, int fdj, int fdk,
                     int cdj, int cdk)
{
  double c[16], u[16];
  lag(N, N / 2.0 - 0.5, c);
  const int nx = fbox[1] - fbox[0] + 1, ny = fbox[3] - fbox[2] + 1;
  const int nzc = (fbox[5] - fbox[4] - N + 2) / 2 + 1;
  double *ev = malloc(sizeof(double) * nx * ny * nzc);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const double *col = fv + (i + fbox[0]) + (long)(j + fbox[2]) * fdj + (long)fbox[4] * fdk;
      for (int k = 0; k < nzc; k++) {
        for (int p = 0; p < N; p++) u[p] = col[(long)(2 * k + p) * fdk];
        ev[i + nx * (j + ny * k)] = interp(N, c, u);
      }
    }
  const int nyc = (fbox[3] - fbox[2] - N + 2) / 2 + 1;
  double *dv = malloc(sizeof(double) * nx * nyc * nzc);
  for (int k = 0; k < nzc; k++)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < nyc; j++) {
        for (int p = 0; p < N; p++) u[p] = ev[i + nx * (2 * j + p + ny * k)];
        dv[i + nx * (j + nyc * k)] = interp(N, c, u);
      }
  const int nxc = (fbox[1] - fbox[0] - N + 2) / 2 + 1;
  for (int k = 0; k < nzc; k++)
    for (int j = 0; j < nyc; j++)
      for (int i = 0; i < nxc; i++) {
        for (int p = 0; p < N; p++) u[p] = dv[2 * i + p + nx * (j + nyc * k)];
        cv[(cbox[0] + i) + (long)(cbox[2] + j) * cdj + (long)(cbox[4] + k) * cdk] = interp(N, c, u);
      }
  free(ev);
  free(dv);
}

/* ------------------------------------------------------------------------ */
/* setRPflags for a fine level (nf boxes) and its coarse level (nc boxes).
 * fp: fine level data, fr: coarse level data.  Fine boxes have refinement
 * boundaries on every face and no symmetry or process ghosts. */
void ob_rp_flags(int nf, const ob_box *fb, int nc, const ob_box *cb, const int *parent, int nbuffer, double *fp,
                 double *fr)
{
  for (int b = 0; b < nf; b++) {
    const ob_box *f = &fb[b];
    double *p = fp + level_offset(fb, b);
    const int np = (int)box_size(f);
    const double bb[6] = {f->x0, f->x0 + (f->m - 1) * f->dx, f->y0, f->y0 + (f->n - 1) * f->dy,
                          f->z0, f->z0 + (f->o - 1) * f->dz};
    for (int k = 0; k < f->o; k++)
      for (int j = 0; j < f->n; j++)
        for (int i = 0; i < f->m; i++) {
          const double x = f->x0 + i * f->dx, y = f->y0 + j * f->dy, z = f->z0 + k * f->dz;
          int ii = np, jj = np, kk = np, n;
          n = 1.5 + (x - bb[0]) / f->dx;
          if (n <= nbuffer) ii = n;
          n = 1.5 - (x - bb[1]) / f->dx;
          if (n <= nbuffer) ii = n;
          n = 1.5 + (y - bb[2]) / f->dy;
          if (n <= nbuffer) jj = n;
          n = 1.5 - (y - bb[3]) / f->dy;
          if (n <= nbuffer) jj = n;
          n = 1.5 + (z - bb[4]) / f->dz;
          if (n <= nbuffer) kk = n;
          n = 1.5 - (z - bb[5]) / f->dz;
          if (n <= nbuffer) kk = n;
          if (ii > jj) ii = jj;
          if (ii > kk) ii = kk;
          p[i + (long)f->m * (j + (long)f->n * k)] = (ii < np) ? ii : 0;
        }
  }
  long ncp = 0;
  for (int b = 0; b < nc; b++) ncp += box_size(&cb[b]);
  for (long q = 0; q < ncp; q++) fr[q] = 0;
  for (int pass = 0; pass < 2; pass++)
    for (int b = 0; b < nf; b++) {
      const ob_box *f = &fb[b], *c = &cb[parent[b]];
      const double *p = fp + level_offset(fb, b);
      double *r = fr + level_offset(cb, parent[b]);
      for (int k = 0; k < f->o; k++)
        for (int j = 0; j < f->n; j++)
          for (int i = 0; i < f->m; i++) {
            const double pv = p[i + (long)f->m * (j + (long)f->n * k)];
            const long q = parent_index(c->x0, c->dx, f->x0, f->dx, i) +
                           (long)c->m * (parent_index(c->y0, c->dy, f->y0, f->dy, j) +
                                         (long)c->n * parent_index(c->z0, c->dz, f->z0, f->dz, k));
            if (pass == 0 && pv == 0) r[q] = 3;
            if (pass == 1) {
              if (pv >= 3) r[q] = 2;
              else if (pv == 2) r[q] = 1;
            }
          }
    }
}

/* prolong_varlist_boxtobox_N region of one fine box (fp: that box's flags).
 * nbuffer >= 3: 0 < flag <= zone; nbuffer == 1: flag == 1; nbuffer < 0: all.
 * Returns the number of points to prolong. */
int ob_prolong_region(const ob_box *f, const ob_box *c, const double *fp, int nbuffer, int zone, int order,
                      int *ifbox, int *ifbox0, int *icbox)
{
  int n = 0;
  for (int i = 0; i < 6; i += 2) ifbox[i] = ifbox0[i] = INT_MAX;
  for (int i = 1; i < 6; i += 2) ifbox[i] = ifbox0[i] = INT_MIN;
  for (int k = 0; k < f->o; k++)
    for (int j = 0; j < f->n; j++)
      for (int i = 0; i < f->m; i++) {
        const double p = fp[i + (long)f->m * (j + (long)f->n * k)];
        const int sel = (nbuffer >= 3 && p > 0 && zone >= p) || (nbuffer == 1 && p == 1) || nbuffer < 0;
        int *bb = sel ? ifbox : ifbox0;
        n += sel;
        if (i < bb[0]) bb[0] = i;
        if (i > bb[1]) bb[1] = i;
        if (j < bb[2]) bb[2] = j;
        if (j > bb[3]) bb[3] = j;
        if (k < bb[4]) bb[4] = k;
        if (k > bb[5]) bb[5] = k;
      }
  if (!n) return 0;
  const int b = order / 2;
  icbox[0] = point_index(c->x0, c->dx, f->x0, f->dx, ifbox[0]) - b + 1;
  icbox[2] = point_index(c->y0, c->dy, f->y0, f->dy, ifbox[2]) - b + 1;
  icbox[4] = point_index(c->z0, c->dz, f->z0, f->dz, ifbox[4]) - b + 1;
  icbox[1] = point_index(c->x0, c->dx, f->x0, f->dx, ifbox[1]) + b;
  icbox[3] = point_index(c->y0, c->dy, f->y0, f->dy, ifbox[3]) + b;
  icbox[5] = point_index(c->z0, c->dz, f->z0, f->dz, ifbox[5]) + b;
  return n;
}

/* the coarse index range covered by fine box f: parents of its corner points */
static void footprint(const ob_box *f, const ob_box *c, int *lo, int *hi)
{
  lo[0] = parent_index(c->x0, c->dx, f->x0, f->dx, 0);
  hi[0] = parent_index(c->x0, c->dx, f->x0, f->dx, f->m - 1);
  lo[1] = parent_index(c->y0, c->dy, f->y0, f->dy, 0);
  hi[1] = parent_index(c->y0, c->dy, f->y0, f->dy, f->n - 1);
  lo[2] = parent_index(c->z0, c->dz, f->z0, f->dz, 0);
  hi[2] = parent_index(c->z0, c->dz, f->z0, f->dz, f->o - 1);
}

/* restrict_varlist_boxtobox_N region of one fine box (fr: the parent box's
 * flags); Rflag 3 for nbuffer >= 3, 2 for nbuffer == 1.  Returns the number
 * of flagged coarse points, -1 for an invalid nbuffer. */
int ob_restrict_region(const ob_box *f, const ob_box *c, const double *fr, int nbuffer, int order, int *icbox,
                       int *ifbox)
{
  int Rflag, n = 0, lo[3], hi[3];
  if (nbuffer >= 3) Rflag = 3;
  else if (nbuffer == 1) Rflag = 2;
  else return -1;
  for (int i = 0; i < 6; i += 2) icbox[i] = INT_MAX;
  for (int i = 1; i < 6; i += 2) icbox[i] = INT_MIN;
  footprint(f, c, lo, hi);
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++)
        if (fr[i + (long)c->m * (j + (long)c->n * k)] >= Rflag) {
          n++;
          if (i < icbox[0]) icbox[0] = i;
          if (i > icbox[1]) icbox[1] = i;
          if (j < icbox[2]) icbox[2] = j;
          if (j > icbox[3]) icbox[3] = j;
          if (k < icbox[4]) icbox[4] = k;
          if (k > icbox[5]) icbox[5] = k;
        }
  if (!n) return 0;
  const int b = order / 2;
  ifbox[0] = point_index(f->x0, f->dx, c->x0, c->dx, icbox[0]) - b + 1;
  ifbox[2] = point_index(f->y0, f->dy, c->y0, c->dy, icbox[2]) - b + 1;
  ifbox[4] = point_index(f->z0, f->dz, c->z0, c->dz, icbox[4]) - b + 1;
  ifbox[1] = point_index(f->x0, f->dx, c->x0, c->dx, icbox[1]) + b;
  ifbox[3] = point_index(f->y0, f->dy, c->y0, c->dy, icbox[3]) + b;
  ifbox[5] = point_index(f->z0, f->dz, c->z0, c->dz, icbox[5]) + b;
  return n;
}

static int inside(const int *bb, const ob_box *b)
{
  return bb[0] >= 0 && bb[1] < b->m && bb[2] >= 0 && bb[3] < b->n && bb[4] >= 0 && bb[5] < b->o;
}

/* prolong_varlist_boxtobox_N over a fine level: nvar variables, fine data
 * fv[v*nfp ...], coarse data cv[v*ncp ...].  Returns 0, or -(b+1) when the
 * coarse stencil of fine box b leaves its parent box ("boxes are too small"). */
int ob_prolong_level(int nf, const ob_box *fb, int nc, const ob_box *cb, const int *parent, const double *fp,
                     int nbuffer, int zone, int order, int nvar, const double *cv, double *fv)
{
  const long nfp = level_offset(fb, nf), ncp = level_offset(cb, nc);
  for (int b = 0; b < nf; b++) {
    const ob_box *f = &fb[b], *c = &cb[parent[b]];
    int ifbox[6], ifbox0[6], icbox[6];
    if (!ob_prolong_region(f, c, fp + level_offset(fb, b), nbuffer, zone, order, ifbox, ifbox0, icbox))
      continue;
    if (!inside(icbox, c)) return -(b + 1);
    for (int v = 0; v < nvar; v++)
      ob_prolong_box(order, icbox, ifbox, ifbox0, cv + v * ncp + level_offset(cb, parent[b]),
                     fv + v * nfp + level_offset(fb, b), c->m, c->m * c->n, f->m, f->m * f->n);
  }
  return 0;
}

/* restrict_varlist_boxtobox_N over a fine level */
int ob_restrict_level(int nf, const ob_box *fb, int nc, const ob_box *cb, const int *parent, const double *fr,
                      int nbuffer, int order, int nvar, const double *fv, double *cv)
{
  const long nfp = level_offset(fb, nf), ncp = level_offset(cb, nc);
  for (int b = 0; b < nf; b++) {
    const ob_box *f = &fb[b], *c = &cb[parent[b]];
    int ifbox[6], icbox[6];
    const int n = ob_restrict_region(f, c, fr + level_offset(cb, parent[b]), nbuffer, order, icbox, ifbox);
    if (n < 0) return n;
    if (!n) continue;
    if (!inside(ifbox, f)) return -(b + 1);
    for (int v = 0; v < nvar; v++)
      ob_restrict_box(order, ifbox, icbox, fv + v * nfp + level_offset(fb, b),
                      cv + v * ncp + level_offset(cb, parent[b]), f->m, f->m * f->n, c->m, c->m * c->n);
  }
  return 0;
}
