/* grhd_flux_oracle.c -- CPU restatement of BAM's GRHD flux divergence.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * matter_fluxes (src/projects/matter/matter_fluxes.c:20-86): for x, then y,
 * then z, every line of the box is copied into 1-D buffers (compute_flux1d,
 * :492-647: lapse, the shift component along the line, the ADM metric and
 * the primitives, tensor and vector components permuted so that the line is
 * "x", dir_perm :91-213, dir_perm_b :215-248), extended by flat
 * extrapolation (extrapolate, :467-477), and the interface fluxes
 * F[0..N] are formed by grhd_flx1d_fvpr (src/projects/grhd/grhd_flx1d.c:
 * 25-250); add_flx1d_rhs (:251-465) zeroes every interface flux with a
 * non-finite component and subtracts (F[l] - F[l-1]) / h (times c when
 * c != 0) from the right-hand side at all N points of the line.
 *
 * grhd_flx1d_fvpr: left state rec_p(w, i), right state rec_m(w, i+1) for
 * the primitives (hrsc_rec WENOZ or WENO5, hrsc_rec1d.c:283-363, 502-581),
 * the metric by rec1d_p_lag4 on both sides (hrsc_rec_metric LAG4, :259-263,
 * 479-482); EOS p and cs2 (eos_ideal / eos_poly, eos_analytic.c:12-72,
 * 423-434); v_i and v^2 capped at grhd_vmax, W capped at grhd_Wlor_max and
 * the flux zeroed when the cap is hit (CureFlxVmax); conserved variables and
 * physical fluxes (grhd_p2c.c:11-106); eigenvalues (grhd_eig.c:20-60); LLF or
 * HLL (hrsc_numflx1d.c:14-110); the tau flux is 0 for a cold EOS.
 * The atmosphere mask (grhd_use_atmosphere_mask, default no) is not used.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define NG 4 /* hrsc_nghosts of the reference's parameter files */

static const int vperm[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
static const int tperm[3][6] = {{0, 1, 2, 3, 4, 5}, {3, 4, 1, 5, 2, 0}, {5, 2, 4, 0, 1, 3}};

#define SQ(a) ((a) * (a))

static const double c13o12 = 13. / 12.;
static const double weno_eps = 1e-40;

static double weno_weights_apply(const double b[3], const double uk[3], int z)
{
  static const double ow[3] = {1. / 10., 3. / 5., 3. / 10.};
  double a[3], w[3];
  for (int j = 0; j < 3; j++)
    a[j] = z ? ow[j] * (1. + fabs(b[0] - b[2]) / (weno_eps + b[j])) : ow[j] / SQ(weno_eps + b[j]);
  const double dsa = 1. / (a[0] + a[1] + a[2]);
  for (int j = 0; j < 3; j++) w[j] = a[j] * dsa;
  return w[0] * uk[0] + w[1] * uk[1] + w[2] * uk[2];
}

/* rec1d_p_weno5 / rec1d_p_wenoz: the value at i + 1/2 from the left */
static double rec_p(const double *u, int i, int z)
{
  const double umt = u[i - 2], umo = u[i - 1], ui = u[i], upo = u[i + 1], upt = u[i + 2];
  double b[3], uk[3];
  b[0] = c13o12 * SQ(umt - 2. * umo + ui) + (1. / 4.) * SQ(umt - 4. * umo + 3. * ui);
  b[1] = c13o12 * SQ(umo - 2. * ui + upo) + (1. / 4.) * SQ(umo - upo);
  b[2] = c13o12 * SQ(ui - 2. * upo + upt) + (1. / 4.) * SQ(3. * ui - 4. * upo + upt);
  uk[0] = (1. / 6.) * (2. * umt - 7. * umo + 11. * ui);
  uk[1] = (1. / 6.) * (-1. * umo + 5. * ui + 2. * upo);
  uk[2] = (1. / 6.) * (2. * ui + 5. * upo - upt);
  return weno_weights_apply(b, uk, z);
}

/* rec1d_m_weno5 / rec1d_m_wenoz: the value at i - 1/2 from the right */
static double rec_m(const double *u, int i, int z)
{
  const double umt = u[i - 2], umo = u[i - 1], ui = u[i], upo = u[i + 1], upt = u[i + 2];
  double b[3], uk[3];
  b[0] = c13o12 * SQ(upt - 2. * upo + ui) + (1. / 4.) * SQ(upt - 4. * upo + 3. * ui);
  b[1] = c13o12 * SQ(upo - 2. * ui + umo) + (1. / 4.) * SQ(upo - umo);
  b[2] = c13o12 * SQ(ui - 2. * umo + umt) + (1. / 4.) * SQ(3. * ui - 4. * umo + umt);
  uk[0] = (1. / 6.) * (2. * upt - 7. * upo + 11. * ui);
  uk[1] = (1. / 6.) * (-upo + 5. * ui + 2. * umo);
  uk[2] = (1. / 6.) * (2. * ui + 5. * umo - umt);
  return weno_weights_apply(b, uk, z);
}

/* rec1d_p_lag4: 4-point Lagrange value at i + 1/2 */
static double lag4(const double *u, int i) { return (9. * (u[i] + u[i + 1]) - u[i - 1] - u[i + 2]) * (1. / 16.); }

double ob_hrsc_rec(int kind, int side, const double *u, int i)
{
  /* kind 0 WENOZ, 1 WENO5, 2 LAG4; side 0: rec_p(u, i), 1: rec_m(u, i) */
  if (kind == 2) return side ? lag4(u, i - 1) : lag4(u, i);
  return side ? rec_m(u, i, kind == 0) : rec_p(u, i, kind == 0);
}

/* EOS.comp("re", ..., "pc"): eos_ideal / eos_poly with eos_cs2_rep */
static void eos_pc(const ob_hrsc_params *P, double rho, double epsl, double *p, double *cs2)
{
  double dpdrho, dpdepsl;
  if (P->eos == 1) {
    const double krg = P->K * pow(rho, P->gamma - 1.);
    *p = rho * krg;
    dpdepsl = 0.;
    dpdrho = P->gamma * krg;
    if (epsl <= -1.) epsl = krg / (P->gamma - 1.);
  } else {
    *p = (P->gamma - 1.) * rho * epsl;
    dpdepsl = (P->gamma - 1.) * rho;
    dpdrho = (P->gamma - 1.) * epsl;
  }
  *cs2 = rho == 0. ? 0. : (dpdrho + *p * dpdepsl / (rho * rho)) / (1.0 + epsl + *p / rho);
}

static void eig(double alpha, double betax, double vx, double v2, double gxxup, double cs2, double *lam0,
                double *lamm, double *lamp)
{
  const double Csqrt = sqrt(cs2 * ((1. - v2) * (gxxup * (1. - v2 * cs2) - vx * vx * (1. - cs2))));
  *lam0 = alpha * vx - betax;
  *lamm = alpha * (vx * (1. - cs2) - Csqrt) / (1. - v2 * cs2) - betax;
  *lamp = alpha * (vx * (1. - cs2) + Csqrt) / (1. - v2 * cs2) - betax;
  if (!isfinite(*lamm)) *lamm = 0.;
  if (!isfinite(*lamp)) *lamp = 0.;
  if (!isfinite(*lam0)) *lam0 = 0.;
}

/* one side: conserved variables, physical flux, eigenvalues; returns 1 when W hit the cap */
static void side_state(const ob_hrsc_params *P, const double g[8], double detg, double gupxx, double rho,
                       double epsl, double vx, double vy, double vz, int right, double q[5], double f[5],
                       double lam[5], double *Wlor)
{
  double p, cs2;
  eos_pc(P, rho, epsl, &p, &cs2);
  const double gxx = g[2], gxy = g[3], gxz = g[4], gyy = g[5], gyz = g[6], gzz = g[7];
  const double vlx = gxx * vx + gxy * vy + gxz * vz;
  const double vly = gxy * vx + gyy * vy + gyz * vz;
  const double vlz = gxz * vx + gyz * vy + gzz * vz;
  double v2 = vlx * vx + vly * vy + vlz * vz;
  v2 = fmin(v2, P->vmax);
  if (right ? v2 >= P->vmax : v2 == P->vmax)
    *Wlor = P->wlormax;
  else
    *Wlor = 1. / sqrt(1. - v2);
  const double W = *Wlor;
  const double sq = sqrt(detg);
  const double W2hrho = W * W * (rho + rho * epsl + p);
  q[0] = sq * W * rho;
  q[1] = sq * (W2hrho - p - W * rho);
  q[2] = sq * W2hrho * vlx;
  q[3] = sq * W2hrho * vly;
  q[4] = sq * W2hrho * vlz;
  const double alpha = g[0], betax = g[1];
  const double sqrtg = sqrt(detg);
  const double avb = alpha * vx - betax;
  f[0] = q[0] * avb;
  f[1] = q[1] * avb + alpha * sqrtg * p * vx;
  f[2] = q[2] * avb + alpha * sqrtg * p;
  f[3] = q[3] * avb;
  f[4] = q[4] * avb;
  eig(alpha, betax, vx, v2, gupxx, cs2, &lam[1], &lam[0], &lam[4]);
  lam[2] = lam[3] = lam[1];
}

/* the interface flux between the reconstructed states (grhd_flx1d_fvpr body) */
void ob_grhd_interface_flux(const ob_hrsc_params *P, const double g[8], const double wl[5], const double wr[5],
                            double F[5])
{
  const double gxx = g[2], gxy = g[3], gxz = g[4], gyy = g[5], gyz = g[6], gzz = g[7];
  const double gi11 = gyy * gzz - gyz * gyz, gi12 = gxz * gyz - gxy * gzz, gi13 = gxy * gyz - gxz * gyy;
  const double detg = gxx * gi11 + gxy * gi12 + gxz * gi13;
  const double detginv = 1. / detg;
  const double gupxx = detginv * (gyy * gzz - gyz * gyz);
  double qL[5], qR[5], fL[5], fR[5], lamL[5], lamR[5], WL, WR;
  side_state(P, g, detg, gupxx, wl[0], wl[1], wl[2], wl[3], wl[4], 0, qL, fL, lamL, &WL);
  side_state(P, g, detg, gupxx, wr[0], wr[1], wr[2], wr[3], wr[4], 1, qR, fR, lamR, &WR);
  if (WL == P->wlormax || WR == P->wlormax) {
    for (int k = 0; k < 5; k++) F[k] = 0.;
    return;
  }
  double aplus = 0., amins = 0.;
  for (int v = 0; v < 5; v++) {
    aplus = fmax(aplus, fmax(lamR[v], lamL[v]));
    amins = fmin(amins, fmin(lamR[v], lamL[v]));
  }
  aplus = fabs(aplus);
  amins = fabs(amins);
  if (P->riemann == 1) {
    const double oda = 1. / (aplus + amins + 1e-32);
    const double apm = aplus * amins;
    for (int v = 0; v < 5; v++) F[v] = oda * (aplus * fL[v] + amins * fR[v] - apm * (qR[v] - qL[v]));
  } else {
    const double amax = fmax(amins, aplus);
    for (int v = 0; v < 5; v++) F[v] = 0.5 * (fR[v] + fL[v] - amax * (qR[v] - qL[v]));
  }
  if (P->eos == 1) F[1] = 0.;
}

void ob_grhd_fluxes(const ob_box *B, const ob_hrsc_params *P, double c, const double *const *metric,
                    const double *const *prim, double *const *rhs)
{
  const int nmax = B->m > B->n ? (B->m > B->o ? B->m : B->o) : (B->n > B->o ? B->n : B->o);
  const int len = nmax + 2 * NG;
  double *buf = malloc(sizeof(double) * 13 * len);
  double *F = malloc(sizeof(double) * 5 * (nmax + 1));
  const int z = P->rec == 0;
  for (int dir = 0; dir < 3; dir++) {
    const int N = dir == 0 ? B->m : (dir == 1 ? B->n : B->o);
    const int na = dir == 0 ? B->n : B->m, nb = dir == 2 ? B->n : B->o;
    const double h = dir == 0 ? B->dx : (dir == 1 ? B->dy : B->dz);
    const double ooh = 1. / h;
    const double *src[13];
    src[0] = metric[0];
    src[1] = metric[1 + vperm[dir][0]];
    for (int k = 0; k < 6; k++) src[2 + k] = metric[4 + tperm[dir][k]];
    src[8] = prim[0];
    src[9] = prim[1];
    for (int k = 0; k < 3; k++) src[10 + k] = prim[2 + vperm[dir][k]];
    double *dst[5] = {rhs[0], rhs[1], rhs[2 + vperm[dir][0]], rhs[2 + vperm[dir][1]], rhs[2 + vperm[dir][2]]};
    for (int b = 0; b < nb; b++)
      for (int a = 0; a < na; a++) {
        long base, stride;
        if (dir == 0) {
          base = (long)B->m * a + (long)B->m * B->n * b;
          stride = 1;
        } else if (dir == 1) {
          base = a + (long)B->m * B->n * b;
          stride = B->m;
        } else {
          base = a + (long)B->m * b;
          stride = (long)B->m * B->n;
        }
        /* 1-based line index l = 1..N lives at buf[f][l + NG - 1] */
        for (int f = 0; f < 13; f++) {
          double *u = buf + f * len + NG - 1;
          for (int l = 1; l <= N; l++) u[l] = src[f][base + stride * (l - 1)];
          for (int l = -NG + 1; l < 1; l++) u[l] = u[1];
          for (int l = N + 1; l <= N + NG; l++) u[l] = u[N];
        }
        for (int i = 0; i <= N; i++) {
          double g[8], wl[5], wr[5];
          for (int f = 0; f < 8; f++) g[f] = lag4(buf + f * len + NG - 1, i);
          for (int f = 0; f < 5; f++) {
            const double *u = buf + (8 + f) * len + NG - 1;
            wl[f] = rec_p(u, i, z);
            wr[f] = rec_m(u, i + 1, z);
          }
          double Fi[5];
          ob_grhd_interface_flux(P, g, wl, wr, Fi);
          int bad = 0;
          for (int v = 0; v < 5; v++) bad |= !isfinite(Fi[v]);
          for (int v = 0; v < 5; v++) F[v * (nmax + 1) + i] = bad ? 0. : Fi[v];
        }
        for (int l = 1; l <= N; l++) {
          const long ijk = base + stride * (l - 1);
          for (int v = 0; v < 5; v++) {
            const double dF = ooh * (F[v * (nmax + 1) + l] - F[v * (nmax + 1) + l - 1]);
            if (c != 0.)
              dst[v][ijk] -= c * dF;
            else
              dst[v][ijk] -= dF;
          }
        }
      }
  }
  free(buf);
  free(F);
}
