/* z4c_oracle.c -- TEST INFRASTRUCTURE: CPU restatement of the BAM Z4c RHS.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file; it is the checker for the HIP kernels, never the product path.
 *
 * Follows src/projects/z4/z4_rhs_movpunc_N.m:17-483 (equations and branch
 * structure), with the stencils of src/math/MathToC/discretize.m:18-46,97-135,
 * the advection stencils of src/main/amr/bam_amr_advection.h:28-155 and the
 * box-edge predicates of src/main/amr/bam_amr_loops_box.h:186-215.
 * Written as plain index loops over full 3x3 tensors.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* stencil4 of bam_amr_advection.h:28-36 (5-point, rows = offset) */
static const double ST4[5][5] = {
  {+25. / 12, -48. / 12, +36. / 12, -16. / 12, +3. / 12},
  {+3. / 12, +10. / 12, -18. / 12, +6. / 12, -1. / 12},
  {-1. / 12, +8. / 12, 0. / 12, -8. / 12, +1. / 12},
  {+1. / 12, -6. / 12, +18. / 12, -10. / 12, -3. / 12},
  {-3. / 12, +16. / 12, -36. / 12, +48. / 12, -25. / 12}};
/* stencil2 of bam_amr_advection.h:28-30 */
static const double ST2[3][3] = {{+1.5, -2.0, +0.5}, {+0.5, 0.0, -0.5}, {-0.5, +2.0, -1.5}};

/* first derivative along stride s (discretize.m:21-22 d1stencil2nd/4th) */
static double D1(const double *u, long q, long s, double h, int ord)
{
  if (ord == 2) return (u[q + s] - u[q - s]) / (2 * h);
  return (u[q - 2 * s] - 8 * u[q - s] + 8 * u[q + s] - u[q + 2 * s]) / (12 * h);
}
/* pure second derivative (d2stencil2nd/4th) */
static double D2(const double *u, long q, long s, double h, int ord)
{
  if (ord == 2) return (u[q - s] - 2 * u[q] + u[q + s]) / (h * h);
  return (-u[q - 2 * s] + 16 * u[q - s] - 30 * u[q] + 16 * u[q + s] - u[q + 2 * s]) / (12 * h * h);
}
/* mixed derivative: tensor product of first-derivative stencils (findif2) */
static double D11(const double *u, long q, long s1, double h1, long s2, double h2, int ord)
{
  static const int o2[2] = {-1, 1};
  static const double w2[2] = {-1, 1};
  static const int o4[4] = {-2, -1, 1, 2};
  static const double w4[4] = {1, -8, 8, -1};
  double sum = 0;
  if (ord == 2) {
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) sum += w2[a] * w2[b] * u[q + o2[a] * s1 + o2[b] * s2];
    return sum / (2 * h1) / (2 * h2);
  }
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) sum += w4[a] * w4[b] * u[q + o4[a] * s1 + o4[b] * s2];
  return sum / (12 * h1) / (12 * h2);
}

/* per-point, per-direction advection stencil (set_advectionN_onedir) */
typedef struct {
  int n;          /* number of points */
  long idx[5];    /* offsets relative to q */
  double w[5];    /* weights, already multiplied by beta/h */
} advdir;

static void set_adv_dir(advdir *a, int i, int imax, long s, double wx, int ord, int lopsided)
{
  int rstencil = (ord == 2) ? 2 : 4 - lopsided;
  int wsign, offset;
  if (wx < 0) {
    wsign = -1;
    offset = (i >= rstencil) ? 0 : rstencil - i;
  } else if (wx > 0) {
    wsign = +1;
    offset = (imax - i >= rstencil) ? 0 : rstencil - (imax - i);
  } else {
    wsign = 0;
    offset = 0;
  }
  if (ord == 2) {
    a->n = 3;
    for (int k = 0; k < 3; k++) {
      a->idx[k] = (long)wsign * (k - offset) * s;
      a->w[k] = -wsign * wx * ST2[offset][k];
    }
  } else {
    offset += lopsided;
    a->n = 5;
    for (int k = 0; k < 5; k++) {
      a->idx[k] = (long)wsign * (k - offset) * s;
      a->w[k] = -wsign * wx * ST4[offset][k];
    }
  }
}

static double ADV(const advdir ad[3], const double *u, long q)
{
  double s = 0;
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < ad[d].n; k++) s += ad[d].w[k] * u[q + ad[d].idx[k]];
  return s;
}

/* Kreiss-Oliger term dissipation4 (discretize.m:131-135) */
static double KO4(const double *u, long q, const long st[3], const double h[3])
{
  double s = 0;
  for (int d = 0; d < 3; d++) {
    long t = st[d];
    s += (u[q - 2 * t] - 4 * u[q - t] + 6 * u[q] - 4 * u[q + t] + u[q + 2 * t]) / h[d];
  }
  return -s;
}

/* symmetric index map for 6-component storage xx xy xz yy yz zz */
static const int SY[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

static double det3(const double g[3][3])
{
  return g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1])
       - g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0])
       + g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
}
/* adjugate ("matrixinvdet") of a symmetric matrix */
static void adj3(const double g[3][3], double a[3][3])
{
  a[0][0] = g[1][1] * g[2][2] - g[1][2] * g[2][1];
  a[0][1] = g[0][2] * g[2][1] - g[0][1] * g[2][2];
  a[0][2] = g[0][1] * g[1][2] - g[0][2] * g[1][1];
  a[1][0] = g[1][2] * g[2][0] - g[1][0] * g[2][2];
  a[1][1] = g[0][0] * g[2][2] - g[0][2] * g[2][0];
  a[1][2] = g[0][2] * g[1][0] - g[0][0] * g[1][2];
  a[2][0] = g[1][0] * g[2][1] - g[1][1] * g[2][0];
  a[2][1] = g[0][1] * g[2][0] - g[0][0] * g[2][1];
  a[2][2] = g[0][0] * g[1][1] - g[0][1] * g[1][0];
}

int ob_z4c_rhs(const ob_box *b, const ob_z4c_params *p,
               double *const *unew, const double *const *upre, double c,
               const double *const *ucur)
{
  const int m = b->m, n = b->n, o = b->o;
  const int imax = m - 1, jmax = n - 1, kmax = o - 1;
  const long st[3] = {1, m, (long)m * n};
  const double h[3] = {b->dx, b->dy, b->dz};
  const int addlinear = (c != 0.0);
  int nbad = 0;

#pragma omp parallel for collapse(2) schedule(static) reduction(+ : nbad)
  for (int k = b->skip[4]; k <= kmax - b->skip[5]; k++)
    for (int j = b->skip[2]; j <= jmax - b->skip[3]; j++)
      for (int i = b->skip[0]; i <= imax - b->skip[1]; i++) {
        const long q = i + st[1] * j + st[2] * k;
        const int pos[3] = {i, j, k}, pmax[3] = {imax, jmax, kmax};
        const int away1 = (i == 1 || j == 1 || k == 1 || i == imax - 1 || j == jmax - 1 || k == kmax - 1);
        const int twoormore = (i >= 2 && j >= 2 && k >= 2 && i <= imax - 2 && j <= jmax - 2 && k <= kmax - 2);
        const int ordc = (p->order_centered == 2 || away1) ? 2 : 4;
        const int orda = (p->order_advection == 2 || away1) ? 2 : 4;

        /* point values */
        double g[3][3], A[3][3], G[3], beta[3], B[3];
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            g[a][bb] = ucur[OV_GXX + SY[a][bb]][q];
            A[a][bb] = ucur[OV_AXX + SY[a][bb]][q];
          }
        for (int a = 0; a < 3; a++) {
          G[a] = ucur[OV_GX + a][q];
          beta[a] = ucur[OV_BETAX + a][q];
          B[a] = ucur[OV_BX + a][q];
        }
        const double chi = ucur[OV_CHI][q], Khat = ucur[OV_KHAT][q];
        const double Theta = ucur[OV_THETA][q], alpha = ucur[OV_ALPHA][q];

        /* centred derivatives (z4_rhs_movpunc_N.m:35-63); dB and dA of the
           reference feed only the non-principal terms, which are off */
        double da[3], dda[3][3], db[3][3], ddb[3][3][3], delg[3][3][3];
        double deldelg[3][3][3][3], delG[3][3], dKhat[3], dchi[3], ddchi[3][3], dTheta[3];
        for (int a = 0; a < 3; a++) {
          const long s = st[a];
          da[a] = D1(ucur[OV_ALPHA], q, s, h[a], ordc);
          dKhat[a] = D1(ucur[OV_KHAT], q, s, h[a], ordc);
          dchi[a] = D1(ucur[OV_CHI], q, s, h[a], ordc);
          dTheta[a] = D1(ucur[OV_THETA], q, s, h[a], ordc);
          for (int bb = 0; bb < 3; bb++) {
            db[a][bb] = D1(ucur[OV_BETAX + bb], q, s, h[a], ordc);
            delG[a][bb] = D1(ucur[OV_GX + bb], q, s, h[a], ordc);
            for (int cc = 0; cc < 3; cc++) delg[a][bb][cc] = D1(ucur[OV_GXX + SY[bb][cc]], q, s, h[a], ordc);
          }
          for (int bb = 0; bb < 3; bb++) {
#define DD(u) ((a == bb) ? D2((u), q, st[a], h[a], ordc) : D11((u), q, st[a], h[a], st[bb], h[bb], ordc))
            dda[a][bb] = DD(ucur[OV_ALPHA]);
            ddchi[a][bb] = DD(ucur[OV_CHI]);
            for (int cc = 0; cc < 3; cc++) ddb[a][bb][cc] = DD(ucur[OV_BETAX + cc]);
            for (int cc = 0; cc < 3; cc++)
              for (int dd = 0; dd < 3; dd++) deldelg[a][bb][cc][dd] = DD(ucur[OV_GXX + SY[cc][dd]]);
#undef DD
          }
        }

        /* advection (z4_rhs_movpunc_N.m:160-215) */
        advdir ad[3];
        for (int d = 0; d < 3; d++)
          set_adv_dir(&ad[d], pos[d], pmax[d], st[d], beta[d] / h[d], orda, p->advection_lopsided);
        double lieg[3][3], lieA[3][3], advG[3], advbeta[3], advB[3];
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            lieg[a][bb] = ADV(ad, ucur[OV_GXX + SY[a][bb]], q);
            lieA[a][bb] = ADV(ad, ucur[OV_AXX + SY[a][bb]], q);
          }
        double lieKhat = ADV(ad, ucur[OV_KHAT], q);
        double liechi = ADV(ad, ucur[OV_CHI], q);
        double liealpha = ADV(ad, ucur[OV_ALPHA], q);
        double advTheta = ADV(ad, ucur[OV_THETA], q);
        for (int a = 0; a < 3; a++) {
          advG[a] = ADV(ad, ucur[OV_GX + a], q);
          advbeta[a] = ADV(ad, ucur[OV_BETAX + a], q);
          advB[a] = ADV(ad, ucur[OV_BX + a], q);
        }

        /* algebra (z4_rhs_movpunc_N.m:225-319) */
        const double K = Khat + 2 * Theta;
        double ginv[3][3], adjg[3][3];
        adj3(g, adjg);
        const double detginv = 1.0 / det3(g);
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) ginv[a][bb] = detginv * adjg[a][bb];

        double gammado[3][3][3], gamma[3][3][3], Gfromg[3];
        for (int cc = 0; cc < 3; cc++)
          for (int a = 0; a < 3; a++)
            for (int bb = 0; bb < 3; bb++)
              gammado[cc][a][bb] = 0.5 * (delg[a][bb][cc] + delg[bb][a][cc] - delg[cc][a][bb]);
        for (int cc = 0; cc < 3; cc++)
          for (int a = 0; a < 3; a++)
            for (int bb = 0; bb < 3; bb++) {
              double s = 0;
              for (int d = 0; d < 3; d++) s += ginv[cc][d] * gammado[d][a][bb];
              gamma[cc][a][bb] = s;
            }
        for (int a = 0; a < 3; a++) {
          double s = 0;
          for (int bb = 0; bb < 3; bb++)
            for (int cc = 0; cc < 3; cc++) s += ginv[bb][cc] * gamma[a][bb][cc];
          Gfromg[a] = s;
        }

        double R[3][3];
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            double s = 0;
            for (int cc = 0; cc < 3; cc++)
              for (int d = 0; d < 3; d++) {
                double t = -0.5 * deldelg[cc][d][a][bb];
                for (int e = 0; e < 3; e++)
                  t += gamma[e][cc][a] * gammado[bb][e][d] + gamma[e][cc][bb] * gammado[a][e][d]
                     + gamma[e][a][d] * gammado[e][cc][bb];
                s += ginv[cc][d] * t;
              }
            double t2 = 0;
            for (int cc = 0; cc < 3; cc++)
              t2 += g[cc][a] * delG[bb][cc] + g[cc][bb] * delG[a][cc]
                  + Gfromg[cc] * (gammado[a][bb][cc] + gammado[bb][a][cc]);
            R[a][bb] = s + 0.5 * t2;
          }

        double chiguarded = chi;
        if (chiguarded < p->chiDivFloor) chiguarded = p->chiDivFloor;
        const double oochipsipower = 1.0 / p->chipsipower;
        const double f = oochipsipower * log(chiguarded);
        const double psim4 = exp(-4 * f);
        double df[3], ddf[3][3], cddf[3][3];
        for (int a = 0; a < 3; a++) df[a] = oochipsipower * dchi[a] / chiguarded;
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++)
            ddf[a][bb] = oochipsipower * ddchi[a][bb] / chiguarded - p->chipsipower * df[a] * df[bb];
        double trcddf = 0, dfdf = 0, dfda = 0;
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            double s = ddf[a][bb];
            for (int cc = 0; cc < 3; cc++) s -= gamma[cc][a][bb] * df[cc];
            cddf[a][bb] = s;
          }
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            trcddf += ginv[a][bb] * cddf[a][bb];
            dfdf += ginv[a][bb] * df[a] * df[bb];
            dfda += ginv[a][bb] * df[a] * da[bb];
          }
        double Rphi[3][3], cdda[3][3];
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            Rphi[a][bb] = 4 * df[a] * df[bb] - 2 * cddf[a][bb] - 2 * g[a][bb] * trcddf - 4 * g[a][bb] * dfdf;
            double s = dda[a][bb];
            for (int cc = 0; cc < 3; cc++) s -= gamma[cc][a][bb] * da[cc];
            cdda[a][bb] = s - 2 * (df[a] * da[bb] + df[bb] * da[a] - g[a][bb] * dfda);
          }
        double trcdda = 0;
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) trcdda += ginv[a][bb] * cdda[a][bb];
        trcdda *= psim4;

        double AA[3][3], Ainv[3][3], cAA = 0;
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            double s = 0, t = 0;
            for (int cc = 0; cc < 3; cc++)
              for (int d = 0; d < 3; d++) {
                s += ginv[cc][d] * A[a][cc] * A[d][bb];
                t += ginv[a][cc] * ginv[bb][d] * A[cc][d];
              }
            AA[a][bb] = s;
            Ainv[a][bb] = t;
          }
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) cAA += ginv[a][bb] * AA[a][bb];
        double divAinv[3];
        for (int a = 0; a < 3; a++) {
          double s = 0;
          for (int bb = 0; bb < 3; bb++) {
            for (int cc = 0; cc < 3; cc++) s += gamma[a][bb][cc] * Ainv[bb][cc];
            s -= 1.5 * Ainv[a][bb] * dchi[bb] / chiguarded;
            s -= (1.0 / 3.0) * ginv[a][bb] * (2 * dKhat[bb] + dTheta[bb]);
          }
          divAinv[a] = s;
        }
        double Rhat = 0;
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) Rhat += ginv[a][bb] * (R[a][bb] + Rphi[a][bb]);
        Rhat *= psim4;
        const double Hhat = Rhat + 2 * K * K / 3 - cAA;

        const double divbeta = db[0][0] + db[1][1] + db[2][2];
        const double totdivbeta = 2.0 / 3.0 * divbeta;
        double ootddivbeta[3];
        for (int a = 0; a < 3; a++) ootddivbeta[a] = (1.0 / 3.0) * (ddb[a][0][0] + ddb[a][1][1] + ddb[a][2][2]);
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            double sg = -g[a][bb] * totdivbeta, sA = -A[a][bb] * totdivbeta;
            for (int cc = 0; cc < 3; cc++) {
              sg += db[a][cc] * g[bb][cc] + db[bb][cc] * g[a][cc];
              sA += db[a][cc] * A[bb][cc] + db[bb][cc] * A[a][cc];
            }
            lieg[a][bb] += sg;
            lieA[a][bb] += sA;
          }
        liechi += (p->chipsipower / 6) * chiguarded * divbeta;
        double pseudolieG[3];
        for (int a = 0; a < 3; a++) {
          double s = 0;
          for (int bb = 0; bb < 3; bb++) {
            for (int cc = 0; cc < 3; cc++) s += ginv[bb][cc] * ddb[bb][cc][a];
            s += ginv[a][bb] * ootddivbeta[bb] - Gfromg[bb] * db[bb][a];
          }
          pseudolieG[a] = s + Gfromg[a] * totdivbeta + advG[a];
        }

        double rg[3][3], rA[3][3], rG[3], rbeta[3], rB[3];
        for (int a = 0; a < 3; a++)
          for (int bb = 0; bb < 3; bb++) {
            rg[a][bb] = -2 * alpha * A[a][bb] + lieg[a][bb];
            rA[a][bb] = psim4 * (-cdda[a][bb] + alpha * (R[a][bb] + Rphi[a][bb]))
                      - (1.0 / 3.0) * g[a][bb] * (-trcdda + alpha * Rhat)
                      + alpha * (K * A[a][bb] - 2 * AA[a][bb]) + lieA[a][bb];
          }
        for (int a = 0; a < 3; a++) {
          double s = 0;
          for (int bb = 0; bb < 3; bb++) s += -2 * Ainv[a][bb] * da[bb];
          rG[a] = s + 2 * alpha * divAinv[a] - 2 * alpha * p->kappa1 * (G[a] - Gfromg[a]) + pseudolieG[a];
        }
        double rKhat = -trcdda + alpha * (cAA + K * K / 3) + lieKhat + p->kappa1 * (1 - p->kappa2) * alpha * Theta;
        double rchi = (-p->chipsipower / 6) * chiguarded * alpha * K + liechi;
        double rTheta = (alpha * (Hhat / 2 - (2 + p->kappa2) * p->kappa1 * Theta) + advTheta) * p->zeroThetaRHS;

        /* gauge (z4_rhs_movpunc_N.m:375-394) */
        const double falpha = p->oploglapse * p->lapseharmonicf + p->oploglapse2 * 8.0 / 3.0 / (3 - alpha)
                            + p->harmoniclapse * alpha;
        double ralpha = p->oplogwithshift * liealpha - falpha * alpha * Khat;
        if (p->withB != 0) {
          const double betaF = p->shiftgammacoeff * pow(alpha, p->shiftalphapower);
          for (int a = 0; a < 3; a++) {
            rbeta[a] = (p->gamma2factor + p->gamma0factor * betaF) * B[a] + p->withShiftadv * advbeta[a];
            rB[a] = (p->gamma0factor + p->gamma2factor * betaF) * (rG[a] - p->withGadv * advG[a])
                  - p->shiftdriver * B[a] + p->withBadv * advB[a];
          }
        } else {
          for (int a = 0; a < 3; a++) {
            rB[a] = 0;
            rbeta[a] = G[a] + p->withShiftadv * advbeta[a] - p->shiftdriver * beta[a];
          }
        }

        /* assemble the 25 right-hand sides in list order */
        double r[OV_NVARS];
        for (int a = 0; a < 3; a++)
          for (int bb = a; bb < 3; bb++) {
            r[OV_GXX + SY[a][bb]] = rg[a][bb];
            r[OV_AXX + SY[a][bb]] = rA[a][bb];
          }
        r[OV_CHI] = rchi;
        r[OV_KHAT] = rKhat;
        r[OV_THETA] = rTheta;
        r[OV_ALPHA] = ralpha;
        for (int a = 0; a < 3; a++) {
          r[OV_GX + a] = rG[a];
          r[OV_BETAX + a] = rbeta[a];
          r[OV_BX + a] = rB[a];
        }

        /* Kreiss-Oliger dissipation (z4_rhs_movpunc_N.m:404-414) */
        if (p->order_dissipation == 4 && twoormore)
          for (int v = 0; v < OV_NVARS; v++) r[v] += p->dissfactor * KO4(ucur[v], q, st, h);

        for (int v = 0; v < OV_NVARS; v++)
          if (!isfinite(r[v])) {
            nbad++;
            break;
          }

        if (addlinear) {
          /* unew = upre + c rhs, then trA removal and det normalisation
             (z4_rhs_movpunc_N.m:446-467) */
          double ng[3][3], nA[3][3], adjn[3][3];
          for (int v = 0; v < OV_NVARS; v++) unew[v][q] = upre[v][q] + c * r[v];
          for (int a = 0; a < 3; a++)
            for (int bb = 0; bb < 3; bb++) {
              ng[a][bb] = unew[OV_GXX + SY[a][bb]][q];
              nA[a][bb] = unew[OV_AXX + SY[a][bb]][q];
            }
          const double detnginv = 1.0 / det3(ng);
          if (p->subtractA) {
            adj3(ng, adjn);
            double traceA = 0;
            for (int a = 0; a < 3; a++)
              for (int bb = 0; bb < 3; bb++) traceA += detnginv * adjn[a][bb] * nA[a][bb];
            const double aux = -traceA / 3;
            for (int a = 0; a < 3; a++)
              for (int bb = a; bb < 3; bb++) unew[OV_AXX + SY[a][bb]][q] = nA[a][bb] + aux * ng[a][bb];
          }
          if (p->normalizedetg) {
            const double aux = pow(detnginv, 1.0 / 3.0);
            for (int a = 0; a < 3; a++)
              for (int bb = a; bb < 3; bb++) unew[OV_GXX + SY[a][bb]][q] = aux * ng[a][bb];
          }
        } else {
          for (int v = 0; v < OV_NVARS; v++) unew[v][q] = r[v];
        }
      }
  return nbad;
}
