/* z4c_aux_oracle.c -- TEST INFRASTRUCTURE: CPU restatement of the pieces
 * around the Z4c RHS: radiative and periodic boundaries, algebraic
 * constraints, ADM<->Z4c conversion, the 2-register RK4 driver and the
 * Hamiltonian-constraint monitor.  Checker only; never the product path.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static const int SY[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

static double det3(const double g[3][3])
{
  return g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1])
       - g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0])
       + g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
}
static void inv3(const double g[3][3], double gi[3][3])
{
  const double id = 1.0 / det3(g);
  gi[0][0] = id * (g[1][1] * g[2][2] - g[1][2] * g[2][1]);
  gi[0][1] = id * (g[0][2] * g[2][1] - g[0][1] * g[2][2]);
  gi[0][2] = id * (g[0][1] * g[1][2] - g[0][2] * g[1][1]);
  gi[1][0] = id * (g[1][2] * g[2][0] - g[1][0] * g[2][2]);
  gi[1][1] = id * (g[0][0] * g[2][2] - g[0][2] * g[2][0]);
  gi[1][2] = id * (g[0][2] * g[1][0] - g[0][0] * g[1][2]);
  gi[2][0] = id * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
  gi[2][1] = id * (g[0][1] * g[2][0] - g[0][0] * g[2][1]);
  gi[2][2] = id * (g[0][0] * g[1][1] - g[0][1] * g[1][0]);
}

/* ---------------------------------------------------------------------------
 * Radiative boundary: set_boundary_radiative / set_boundary_rot_rad without
 * rotation (src/utility/boundary/radiative.c:14-100,105-236) on every
 * PHYBOUND point = outermost point layer of a single box (flag.c:23-35).
 * Neighbour existence follows set_nbinds_check_7 / check2_6
 * (bam_amr_loops_box.h:338-386): a missing neighbour index equals ccc.
 */
static double rad_del(const double *u, double vx, long c, long mm, long m, long p, long pp)
{
  /* computedel(): radiative.c:18-40 */
#define CEN (u[p] - u[m])
#define RGT (4 * u[p] - u[pp] - 3 * u[c])
#define LFT (3 * u[c] + u[mm] - 4 * u[m])
  if (vx > 0) {
    if (mm != c) return LFT;
    if (m != c) return CEN;
    return RGT;
  } else if (vx < 0) {
    if (pp != c) return RGT;
    if (m != c) return CEN;
    return LFT;
  }
  if (p != c && m != c) return CEN;
  if (m == c) return RGT;
  return LFT;
#undef CEN
#undef RGT
#undef LFT
}

void ob_radiative_bc(const ob_box *b, double *unew, const double *upre, double c,
                     const double *ucur, double var0, double v)
{
  const int M = b->m, N = b->n, O = b->o;
  const long di = 1, dj = M, dk = (long)M * N;
  for (int k = 0; k < O; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < M; i++) {
        if (!(i == 0 || i == M - 1 || j == 0 || j == N - 1 || k == 0 || k == O - 1)) continue;
        const long q = i + dj * j + dk * k;
        const long mcc = (i > 0) ? q - di : q, pcc = (i + 1 < M) ? q + di : q;
        const long cmc = (j > 0) ? q - dj : q, cpc = (j + 1 < N) ? q + dj : q;
        const long ccm = (k > 0) ? q - dk : q, ccp = (k + 1 < O) ? q + dk : q;
        const long Mcc = (i > 1) ? mcc - di : q, Pcc = (i + 2 < M) ? pcc + di : q;
        const long cMc = (j > 1) ? cmc - dj : q, cPc = (j + 2 < N) ? cpc + dj : q;
        const long ccM = (k > 1) ? ccm - dk : q, ccP = (k + 2 < O) ? ccp + dk : q;
        const double x = b->x0 + i * b->dx, y = b->y0 + j * b->dy, z = b->z0 + k * b->dz;
        const double r = sqrt(x * x + y * y + z * z);
        const double vx = v * x / r, vy = v * y / r, vz = v * z / r;
        double rhs = -v * (ucur[q] - var0) / r;
        const double delx = rad_del(ucur, vx, q, Mcc, mcc, pcc, Pcc);
        const double dely = rad_del(ucur, vy, q, cMc, cmc, cpc, cPc);
        const double delz = rad_del(ucur, vz, q, ccM, ccm, ccp, ccP);
        rhs -= vx * delx / (2 * b->dx) + vy * dely / (2 * b->dy) + vz * delz / (2 * b->dz);
        unew[q] = (c != 0.0) ? upre[q] + c * rhs : rhs;
      }
}

/* far limits and propagation speeds set in z4_startup (z4.c:200-228) */
void ob_z4c_radiative_bc(const ob_box *b, const ob_z4c_params *p,
                         double *const *unew, const double *const *upre, double c,
                         const double *const *ucur)
{
  double vg;
  if (p->oploglapse) vg = sqrt(p->lapseharmonicf);
  else if (p->oploglapse2) vg = sqrt(4.0 / 3.0);
  else vg = 1;
  for (int v = 0; v < OV_NVARS; v++) {
    double far = 0, speed = 1;
    if (v == OV_GXX || v == OV_GYY || v == OV_GZZ) far = 1;
    if (v == OV_CHI || v == OV_ALPHA) { far = 1; speed = vg; }
    if (v == OV_KHAT) speed = vg;
    ob_radiative_bc(b, unew[v], upre ? upre[v] : NULL, c, ucur[v], far, speed);
  }
}

/* Periodic copy with g layers (src/utility/boundary/periodic.c:476-560 uses
 * g = 2): receive layers [0,g) from [n-2g, n-g) and [n-g, n) from [g, 2g). */
void ob_periodic_fill(const ob_box *b, const int per[3], int g, double *u)
{
  const int dims[3] = {b->m, b->n, b->o};
  const long st[3] = {1, b->m, (long)b->m * b->n};
  for (int d = 0; d < 3; d++) {
    if (!per[d]) continue;
    const int nd = dims[d];
    const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
    for (int a = 0; a < dims[e1]; a++)
      for (int c = 0; c < dims[e2]; c++) {
        const long base = a * st[e1] + c * st[e2];
        for (int l = 0; l < g; l++) {
          u[base + l * st[d]] = u[base + (nd - 2 * g + l) * st[d]];
          u[base + (nd - g + l) * st[d]] = u[base + (g + l) * st[d]];
        }
      }
  }
}

/* ---------------------------------------------------------------------------
 * z4_algcon (src/projects/z4/z4_algcon.m:14-48) on all points */
void ob_z4c_algcon(const ob_box *b, int subtractA, int normalizedetg,
                   double *const *g6, double *const *A6)
{
  const long np = (long)b->m * b->n * b->o;
  const double oot = 1.0 / 3.0;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < np; q++) {
    double g[3][3], A[3][3], gi[3][3];
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) {
        g[a][c] = g6[SY[a][c]][q];
        A[a][c] = A6[SY[a][c]][q];
      }
    double detg = det3(g);
    if (detg <= 0.) detg = 1.;
    if (normalizedetg) {
      const double eps = detg - 1.;
      const double aux = (fabs(eps) < 1e-12) ? 1. - oot * eps : pow(detg, -oot);
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) g[a][c] *= aux;
    }
    detg = det3(g);
    if (detg <= 0.) detg = 1.;
    /* matrixinvdet / detg, but with the guarded determinant */
    {
      double adj[3][3];
      adj[0][0] = g[1][1] * g[2][2] - g[1][2] * g[2][1];
      adj[0][1] = g[0][2] * g[2][1] - g[0][1] * g[2][2];
      adj[0][2] = g[0][1] * g[1][2] - g[0][2] * g[1][1];
      adj[1][0] = adj[0][1];
      adj[1][1] = g[0][0] * g[2][2] - g[0][2] * g[2][0];
      adj[1][2] = g[0][2] * g[1][0] - g[0][0] * g[1][2];
      adj[2][0] = adj[0][2];
      adj[2][1] = adj[1][2];
      adj[2][2] = g[0][0] * g[1][1] - g[0][1] * g[1][0];
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) gi[a][c] = adj[a][c] / detg;
    }
    double trA = 0;
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) trA += gi[a][c] * A[a][c];
    if (subtractA)
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) A[a][c] += -trA * oot * g[a][c];
    for (int a = 0; a < 3; a++)
      for (int c = a; c < 3; c++) {
        g6[SY[a][c]][q] = g[a][c];
        A6[SY[a][c]][q] = A[a][c];
      }
  }
}

/* ---------------------------------------------------------------------------
 * ADM -> Z4c (src/projects/z4/z4_init.m:14-66): gt = detg^-1/3 g,
 * At = detg^-1/3 (K - g trK/3), chi = psi^chipsipower with psi = detg^1/12,
 * Khat = trK, Theta = 0, G^i = -d_j gtinv^ji with 2nd-order stencils at inner
 * points (zero on the outermost layer, which z4_setbound_G extrapolates). */
/* set_boundary_extrapolate (src/utility/boundary/extrapolate.c:10-47) on the
 * outermost point layer of a box whose faces are all PHYBOUND: three passes
 * in point order, in place, each point from linear extrapolation along the
 * axes on which a neighbour is missing (set_nbinds_check_7 / check2_6), the
 * mean of those, or of the three centred means when none is missing */
static void extrapolate_outer_layer(const ob_box *b, double *v)
{
  const long di = 1, dj = b->m, dk = (long)b->m * b->n;
  for (int rep = 0; rep < 3; rep++)
    for (int k = 0; k < b->o; k++)
      for (int j = 0; j < b->n; j++)
        for (int i = 0; i < b->m; i++) {
          if (i > 0 && j > 0 && k > 0 && i < b->m - 1 && j < b->n - 1 && k < b->o - 1) continue;
          const long c = i + dj * j + dk * k;
          const long nb[3][4] = {
              {i > 0 ? c - di : c, i > 1 ? c - 2 * di : c, i + 1 < b->m ? c + di : c, i + 2 < b->m ? c + 2 * di : c},
              {j > 0 ? c - dj : c, j > 1 ? c - 2 * dj : c, j + 1 < b->n ? c + dj : c, j + 2 < b->n ? c + 2 * dj : c},
              {k > 0 ? c - dk : c, k > 1 ? c - 2 * dk : c, k + 1 < b->o ? c + dk : c, k + 2 < b->o ? c + 2 * dk : c}};
          double val[3];
          int use[3];
          for (int d = 0; d < 3; d++) {
            const long m1 = nb[d][0], M2 = nb[d][1], p1 = nb[d][2], P2 = nb[d][3];
            if (p1 == c) { /* extrapolate from the left */
              val[d] = 2.0 * v[m1] - v[M2];
              use[d] = 1;
            } else if (m1 == c) { /* from the right */
              val[d] = 2.0 * v[p1] - v[P2];
              use[d] = 1;
            } else {
              val[d] = 0.5 * (v[p1] + v[m1]);
              use[d] = 0;
            }
          }
          if (use[0] + use[1] + use[2] > 0)
            v[c] = (use[0] * val[0] + use[1] * val[1] + use[2] * val[2]) / (use[0] + use[1] + use[2]);
          else
            v[c] = (val[0] + val[1] + val[2]) / 3.0;
        }
}

/* ADM -> Z4c: z4_init (src/projects/z4/z4_init.m:9-80) as adm_to_z4 runs it
 * (z4.c:142-171): at every point detg, psi^-4 = detg^(-1/3), psi =
 * detg^(1/12), gt = psi^-4 g, K = gt^ij psi^-4 K_ij, At = psi^-4 (K_ij -
 * g_ij K / 3), chi = psi^chipsipower, Theta = 0; at the inner points (flag
 * layers excluded) G^a = -d_b gt^{ba} with the generated 2nd-order form
 * oo2dx (-u[-1] + u[1]); then z4_setbound_G (z4.c:33-39): the outermost layer
 * extrapolated when phybound is set (a box with PHYBOUND faces). */
void ob_adm_to_z4c(const ob_box *b, double chipsipower, const double *const *gadm, const double *const *kadm,
                   double *const *u, int phybound)
{
  const long np = (long)b->m * b->n * b->o;
  const long st[3] = {1, b->m, (long)b->m * b->n};
  const double oo2h[3] = {1 / (2 * b->dx), 1 / (2 * b->dy), 1 / (2 * b->dz)};
  double *gtinv = malloc(sizeof(double) * 6 * np);
  for (long q = 0; q < np; q++) {
    double g[3][3], K[3][3], gt[3][3], gti[3][3];
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) {
        g[a][c] = gadm[SY[a][c]][q];
        K[a][c] = kadm[SY[a][c]][q];
      }
    const double detg = det3(g);
    const double psim4 = pow(detg, -1.0 / 3.0), psi = pow(detg, 1.0 / 12.0);
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) gt[a][c] = psim4 * g[a][c];
    inv3(gt, gti);
    double trK = 0;
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) trK += gti[a][c] * psim4 * K[a][c];
    for (int a = 0; a < 3; a++)
      for (int c = a; c < 3; c++) {
        u[OV_GXX + SY[a][c]][q] = gt[a][c];
        u[OV_AXX + SY[a][c]][q] = psim4 * (K[a][c] - g[a][c] * trK / 3);
        gtinv[SY[a][c] * np + q] = gti[a][c];
      }
    u[OV_KHAT][q] = trK;
    u[OV_CHI][q] = pow(psi, chipsipower);
    u[OV_THETA][q] = 0;
  }
  for (int k = b->skip[4]; k < b->o - b->skip[5]; k++)
    for (int j = b->skip[2]; j < b->n - b->skip[3]; j++)
      for (int i = b->skip[0]; i < b->m - b->skip[1]; i++) {
        const long q = i + st[1] * j + st[2] * k;
        for (int a = 0; a < 3; a++) {
          double s = 0;
          for (int c = 0; c < 3; c++) {
            const double *gi = gtinv + SY[c][a] * np;
            s -= oo2h[c] * (-gi[q - st[c]] + gi[q + st[c]]);
          }
          u[OV_GX + a][q] = s;
        }
      }
  if (phybound)
    for (int a = 0; a < 3; a++) extrapolate_outer_layer(b, u[OV_GX + a]);
  free(gtinv);
}

/* Brill-Lindquist data: SetMWBL with mwblpsiofxyz (src/projects/punctures/
 * ConformalFactor.c:84-200; psi = sum_i (m_i (1/r_i)) 0.5 + 1, r = EPSZERO
 * where r equals 0 within dequaleps), K_ij = 0, and the conformal factor
 * absorbed into the metric (adm_undo_conftrans, adm.c:96-111): g_ij =
 * psi^4 delta_ij.  pos[3 n]: puncture positions; psi may be NULL. */
void ob_punctures_bl(const ob_box *b, int n, const double *mass, const double *pos, double *const *gadm,
                     double *const *kadm, double *psi_out)
{
  for (int k = 0; k < b->o; k++)
    for (int j = 0; j < b->n; j++)
      for (int i = 0; i < b->m; i++) {
        const long q = i + (long)b->m * j + (long)b->m * b->n * k;
        const double x = b->x0 + i * b->dx, y = b->y0 + j * b->dy, z = b->z0 + k * b->dz;
        double p = 0;
        for (int c = 0; c < n; c++) {
          const double xi = x - pos[3 * c], yi = y - pos[3 * c + 1], zi = z - pos[3 * c + 2];
          double r = sqrt(xi * xi + yi * yi + zi * zi);
          if (!(r < 0.0 - 1e-10 || 0.0 < r - 1e-10)) r = 1e-10 / 1000.0;
          const double ri = 1 / r;
          p += mass[c] * ri * (0.5);
        }
        p += 1.0;
        const double psi4 = pow(p, 4.0);
        for (int a = 0; a < 6; a++) {
          gadm[a][q] = (a == 0 || a == 3 || a == 5 ? 1.0 : 0.0) * psi4;
          kadm[a][q] = 0.0 * pow(p, -2.0);
        }
        if (psi_out) psi_out[q] = p;
      }
}

/* Z4c -> ADM (z4.c:93-124): psi = chi^(1/chipsipower), g = psi^4 gt,
 * K_ij = psi^4 At + Khat g /3 (the stored bssn_K is Khat) */
void ob_z4c_to_adm(const ob_box *b, double chipsipower, const double *const *u,
                   double *const *gadm, double *const *kadm)
{
  const long np = (long)b->m * b->n * b->o;
  for (long q = 0; q < np; q++) {
    const double psi = pow(u[OV_CHI][q], 1. / chipsipower);
    const double psi4 = pow(psi, 4.);
    for (int c = 0; c < 6; c++) {
      gadm[c][q] = psi4 * u[OV_GXX + c][q];
      kadm[c][q] = psi4 * u[OV_AXX + c][q] + 1. / 3. * u[OV_KHAT][q] * gadm[c][q];
    }
  }
}

/* ---------------------------------------------------------------------------
 * SoA vector ops (src/main/main/variables.c:650-745) */
void ob_vladd(long npts, double *r, double ca, const double *a, double cb, const double *b)
{
#pragma omp parallel for schedule(static)
  for (long i = 0; i < npts; i++) {
    double s = 0;
    if (ca != 0) s += ca * a[i];
    if (cb != 0) s += cb * b[i];
    r[i] = s;
  }
}

static void rhs_and_bc(const ob_box *b, const ob_z4c_params *p, int bc, int pgh,
                       double *const *r, const double *const *v, int *nbad)
{
  *nbad += ob_z4c_rhs(b, p, r, NULL, 0.0, v);
  if (bc == 1) {
    ob_z4c_radiative_bc(b, p, r, NULL, 0.0, v);
  } else if (bc == 2) {
    const int per[3] = {1, 1, 1};
    for (int q = 0; q < OV_NVARS; q++) ob_periodic_fill(b, per, pgh, r[q]);
  }
}

/* evolve_rk_rk4 (src/utility/evolve/rk.c:220-257) + z4_algcon_wrapper. */
int ob_z4c_rk4_step(const ob_box *b, const ob_z4c_params *p, double dt, int bc,
                    int pgh, double *const *u, double *const *work, int algcon)
{
  const long np = (long)b->m * b->n * b->o;
  double *const *up = work, *const *r = work + OV_NVARS, *const *v = work + 2 * OV_NVARS;
  int nbad = 0;
  for (int q = 0; q < OV_NVARS; q++) memcpy(up[q], u[q], sizeof(double) * np);
  /* stage 1 */
  rhs_and_bc(b, p, bc, pgh, r, (const double *const *)u, &nbad);
  for (int q = 0; q < OV_NVARS; q++) ob_vladd(np, u[q], 1.0, u[q], dt / 6, r[q]);
  /* stage 2 */
  for (int q = 0; q < OV_NVARS; q++) ob_vladd(np, v[q], 1.0, up[q], dt / 2, r[q]);
  rhs_and_bc(b, p, bc, pgh, r, (const double *const *)v, &nbad);
  for (int q = 0; q < OV_NVARS; q++) ob_vladd(np, u[q], 1.0, u[q], dt / 3, r[q]);
  /* stage 3 */
  for (int q = 0; q < OV_NVARS; q++) ob_vladd(np, v[q], 1.0, up[q], dt / 2, r[q]);
  rhs_and_bc(b, p, bc, pgh, r, (const double *const *)v, &nbad);
  for (int q = 0; q < OV_NVARS; q++) ob_vladd(np, u[q], 1.0, u[q], dt / 3, r[q]);
  /* stage 4 */
  for (int q = 0; q < OV_NVARS; q++) ob_vladd(np, v[q], 1.0, up[q], dt, r[q]);
  rhs_and_bc(b, p, bc, pgh, r, (const double *const *)v, &nbad);
  for (int q = 0; q < OV_NVARS; q++) ob_vladd(np, u[q], 1.0, u[q], dt / 6, r[q]);
  if (algcon) ob_z4c_algcon(b, p->subtractA, p->normalizedetg, u + OV_GXX, u + OV_AXX);
  return nbad;
}

/* ---------------------------------------------------------------------------
 * Hamiltonian and momentum constraints (src/projects/adm/adm_constraints_N.m:
 * 14-116 with usepsi off, no matter) from the ADM fields of z4_to_adm,
 * 4th-order centred (2nd order one point from the box edge); L2 norms of
 * reduce.c:25-30 over the inner points.  l2[0] = |ham|, l2[1..3] = |mom_i|;
 * ham and mom[0..2] receive the fields when not NULL. */
void ob_z4c_constraints_l2(const ob_box *b, const double *const *u, double chipsipower, double *ham,
                           double *const *mom, double l2[4])
{
  const long np = (long)b->m * b->n * b->o;
  const int M = b->m, N = b->n, O = b->o;
  const long st[3] = {1, M, (long)M * N};
  const double h[3] = {b->dx, b->dy, b->dz};
  double *buf = malloc(sizeof(double) * 12 * np);
  double *gadm[6], *kadm[6];
  for (int c = 0; c < 6; c++) {
    gadm[c] = buf + c * np;
    kadm[c] = buf + (6 + c) * np;
  }
  ob_z4c_to_adm(b, chipsipower, u, gadm, kadm);
  double sum = 0, sm0 = 0, sm1 = 0, sm2 = 0;
  long cnt = 0;
#pragma omp parallel for collapse(2) reduction(+ : sum, sm0, sm1, sm2, cnt) schedule(static)
  for (int k = b->skip[4]; k < O - b->skip[5]; k++)
    for (int j = b->skip[2]; j < N - b->skip[3]; j++)
      for (int i = b->skip[0]; i < M - b->skip[1]; i++) {
        const long q = i + st[1] * j + st[2] * k;
        const int away1 = (i == 1 || j == 1 || k == 1 || i == M - 2 || j == N - 2 || k == O - 2);
        const int ord = away1 ? 2 : 4;
        double g[3][3], K[3][3], gi[3][3], dg[3][3][3], ddg[3][3][3][3], dK[3][3][3];
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) {
            g[a][c] = gadm[SY[a][c]][q];
            K[a][c] = kadm[SY[a][c]][q];
          }
        for (int e = 0; e < 3; e++)
          for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) {
              const double *f = gadm[SY[a][c]], *kf = kadm[SY[a][c]];
              const long s = st[e];
              dg[e][a][c] = (ord == 2) ? (f[q + s] - f[q - s]) / (2 * h[e])
                          : (f[q - 2 * s] - 8 * f[q - s] + 8 * f[q + s] - f[q + 2 * s]) / (12 * h[e]);
              dK[e][a][c] = (ord == 2) ? (kf[q + s] - kf[q - s]) / (2 * h[e])
                          : (kf[q - 2 * s] - 8 * kf[q - s] + 8 * kf[q + s] - kf[q + 2 * s]) / (12 * h[e]);
              for (int e2 = 0; e2 < 3; e2++) {
                const long s2 = st[e2];
                double v;
                if (e == e2) {
                  v = (ord == 2) ? (f[q - s] - 2 * f[q] + f[q + s]) / (h[e] * h[e])
                    : (-f[q - 2 * s] + 16 * f[q - s] - 30 * f[q] + 16 * f[q + s] - f[q + 2 * s]) / (12 * h[e] * h[e]);
                } else if (ord == 2) {
                  v = (f[q + s + s2] - f[q + s - s2] - f[q - s + s2] + f[q - s - s2]) / (4 * h[e] * h[e2]);
                } else {
                  static const int of[4] = {-2, -1, 1, 2};
                  static const double w[4] = {1, -8, 8, -1};
                  double t = 0;
                  for (int x1 = 0; x1 < 4; x1++)
                    for (int x2 = 0; x2 < 4; x2++) t += w[x1] * w[x2] * f[q + of[x1] * s + of[x2] * s2];
                  v = t / (144 * h[e] * h[e2]);
                }
                ddg[e][e2][a][c] = v;
              }
            }
        inv3(g, gi);
        double gdo[3][3][3], gam[3][3][3];
        for (int c = 0; c < 3; c++)
          for (int a = 0; a < 3; a++)
            for (int bb = 0; bb < 3; bb++) gdo[c][a][bb] = 0.5 * (dg[a][bb][c] + dg[bb][a][c] - dg[c][a][bb]);
        for (int c = 0; c < 3; c++)
          for (int a = 0; a < 3; a++)
            for (int bb = 0; bb < 3; bb++) {
              double s = 0;
              for (int d = 0; d < 3; d++) s += gi[c][d] * gdo[d][a][bb];
              gam[c][a][bb] = s;
            }
        double R = 0;
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            double Rab = 0;
            for (int c = 0; c < 3; c++)
              for (int d = 0; d < 3; d++) {
                double t = 0.5 * (-ddg[c][d][a][bb] - ddg[a][bb][c][d] + ddg[a][c][bb][d] + ddg[bb][c][a][d]);
                for (int e = 0; e < 3; e++) t += gam[e][a][c] * gdo[e][bb][d] - gam[e][a][bb] * gdo[e][c][d];
                Rab += gi[c][d] * t;
              }
            R += gi[a][bb] * Rab;
          }
        double Kud[3][3], trK = 0, KK = 0;
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            double s = 0;
            for (int c = 0; c < 3; c++) s += gi[a][c] * K[c][bb];
            Kud[a][bb] = s;
          }
        for (int a = 0; a < 3; a++) {
          trK += Kud[a][a];
          for (int bb = 0; bb < 3; bb++) KK += Kud[a][bb] * Kud[bb][a];
        }
        const double H = R + trK * trK - KK;
        /* covariant derivative D_a K_bc and D^a K_bc */
        double cdK[3][3][3], cdKu[3][3][3];
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++)
            for (int c = 0; c < 3; c++) {
              double t = dK[a][bb][c];
              for (int d = 0; d < 3; d++) t -= gam[d][a][bb] * K[d][c] + gam[d][a][c] * K[bb][d];
              cdK[a][bb][c] = t;
            }
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++)
            for (int c = 0; c < 3; c++) {
              double t = 0;
              for (int d = 0; d < 3; d++) t += gi[a][d] * cdK[d][bb][c];
              cdKu[a][bb][c] = t;
            }
        /* mom^a = g^ab D^c K_bc - g^bc D^a K_bc; mom_a = g_ab mom^b */
        double momu[3], M[3];
        for (int a = 0; a < 3; a++) {
          double t = 0;
          for (int bb = 0; bb < 3; bb++)
            for (int c = 0; c < 3; c++) t += gi[a][bb] * cdKu[c][bb][c] - gi[bb][c] * cdKu[a][bb][c];
          momu[a] = t;
        }
        for (int a = 0; a < 3; a++) M[a] = g[a][0] * momu[0] + g[a][1] * momu[1] + g[a][2] * momu[2];
        if (ham) ham[q] = H;
        if (mom)
          for (int a = 0; a < 3; a++) mom[a][q] = M[a];
        sum += H * H;
        sm0 += M[0] * M[0];
        sm1 += M[1] * M[1];
        sm2 += M[2] * M[2];
        cnt += 1;
      }
  free(buf);
  l2[0] = cnt ? sqrt(sum / cnt) : 0.0;
  l2[1] = cnt ? sqrt(sm0 / cnt) : 0.0;
  l2[2] = cnt ? sqrt(sm1 / cnt) : 0.0;
  l2[3] = cnt ? sqrt(sm2 / cnt) : 0.0;
}

double ob_z4c_ham_l2(const ob_box *b, const double *const *u, double chipsipower, double *ham)
{
  double l2[4];
  ob_z4c_constraints_l2(b, u, chipsipower, ham, NULL, l2);
  return l2[0];
}
