/* grhd_source_oracle.c -- CPU restatement of BAM's GRHD source terms.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * grhd_sources (src/projects/grhd/grhd_sources.m:8-190, generated as
 * grhd_sources.c) at the inner points of every box, for the reference's
 * defaults grhd_source_computation = prim and grhd_cooling = 0:
 *   - detg <= 0 or alpha <= 0: no source, matter ADM variables 0 (:21-31);
 *   - first derivatives of alpha, beta^i, g_ij: 2nd order at distance 1 from
 *     a box face or with order_centered 2, else 4th order (:46-60);
 *   - W from v^i (capped at grhd_Wlor_max when v^2 >= 1), T^00, T^0i, T^ij,
 *     T^0_j from the primitives (:66-105);
 *   - s_tau = alpha sqrt(g) (T^00 (beta beta K - beta d alpha)
 *     + T^0i (2 beta^j K_ij - d_i alpha) + T^ij K_ij),
 *     s_S_c = alpha sqrt(g) (T^00 (beta beta d_c g / 2 - alpha d_c alpha)
 *     + T^0i beta^j d_c g_ij + T^0_j d_c beta^j + T^ij d_c g_ij / 2) (:110-124);
 *   - matter ADM variables rho, S_i, S_ij, S (:128-132);
 *   - no momentum source below the atmosphere density, no tau source for a
 *     cold EOS, everything 0 when any of the 16 values is not finite
 *     (:153-176); then unew += source (times c when c != 0, :180-190).
 * Sums run over full 3x3 index ranges in the written order.
 */
#include <math.h>

#include "oracle.h"

static const int S6[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

/* first derivative of u along stride s, centred, 2nd or 4th order, in the
 * form of the generated code (grhd_sources.c: oo2dx = 1/(2 dx),
 * oo2dx (-u[-1] + u[1]) and 1/6 oo2dx (u[-2] + 8 (-u[-1] + u[1]) - u[2])) */
static double d1(const double *u, long q, long s, double oo2h, int ord)
{
  if (ord == 2) return oo2h * (-u[q - s] + u[q + s]);
  return 0.16666666666666666667 * oo2h * (u[q - 2 * s] + 8. * (-u[q - s] + u[q + s]) - u[q + 2 * s]);
}

void ob_grhd_sources(const ob_box *B, const ob_hrsc_params *P, double rhoatm, int order_centered, double c,
                     const double *const *metric, const double *const *Kij, const double *const *prim,
                     double *const *rhs, double *const *madm)
{
  const long m = B->m, mn = (long)B->m * B->n;
  const long str[3] = {1, m, mn};
  const double oo2h[3] = {1 / (2 * B->dx), 1 / (2 * B->dy), 1 / (2 * B->dz)};
  const int imax = B->m - 1, jmax = B->n - 1, kmax = B->o - 1;
  for (int k = B->skip[4]; k <= kmax - B->skip[5]; k++)
    for (int j = B->skip[2]; j <= jmax - B->skip[3]; j++)
      for (int i = B->skip[0]; i <= imax - B->skip[1]; i++) {
        const long q = i + m * j + mn * k;
        const double al = metric[0][q];
        const double be[3] = {metric[1][q], metric[2][q], metric[3][q]};
        double g[6], K[6];
        for (int a = 0; a < 6; a++) {
          g[a] = metric[4 + a][q];
          K[a] = Kij[a][q];
        }
        const double rho = prim[0][q], eps = prim[1][q], p = prim[5][q];
        const double vu[3] = {prim[2][q], prim[3][q], prim[4][q]};
        const double adj[6] = {g[3] * g[5] - g[4] * g[4], g[2] * g[4] - g[1] * g[5], g[1] * g[4] - g[2] * g[3],
                               g[0] * g[5] - g[2] * g[2], g[1] * g[2] - g[0] * g[4], g[0] * g[3] - g[1] * g[1]};
        const double detg = g[0] * adj[0] + g[1] * adj[1] + g[2] * adj[2];
        double sD = 0., sT = 0., sS[3] = {0., 0., 0.}, ad[11];
        for (int a = 0; a < 11; a++) ad[a] = 0.;
        if (detg > 0. && al > 0.) {
          const int ord = (order_centered == 2 || i == 1 || j == 1 || k == 1 || i == imax - 1 || j == jmax - 1 ||
                           k == kmax - 1)
                              ? 2
                              : 4;
          double dal[3], dbe[3][3], dg[3][6];
          for (int d = 0; d < 3; d++) {
            dal[d] = d1(metric[0], q, str[d], oo2h[d], ord);
            for (int a = 0; a < 3; a++) dbe[d][a] = d1(metric[1 + a], q, str[d], oo2h[d], ord);
            for (int a = 0; a < 6; a++) dg[d][a] = d1(metric[4 + a], q, str[d], oo2h[d], ord);
          }
          const double detginv = 1. / detg;
          double gi[6];
          for (int a = 0; a < 6; a++) gi[a] = detginv * adj[a];
          const double sqrtdetg = sqrt(detg), sqrtmdetg = al * sqrtdetg;
          double vl[3], vsq = 0.;
          for (int a = 0; a < 3; a++) {
            vl[a] = 0.;
            for (int b = 0; b < 3; b++) vl[a] += g[S6[a][b]] * vu[b];
          }
          for (int a = 0; a < 3; a++) vsq += vu[a] * vl[a];
          double W = 1. / sqrt(1. - vsq);
          if (vsq >= 1.) W = P->wlormax;
          const double W2hrho = W * W * (rho + rho * eps + p);
          const double al2 = al * al;
          double vs[3], T0i[3], T0j[3], Tij[6];
          for (int a = 0; a < 3; a++) vs[a] = vu[a] - be[a] / al;
          const double T00 = (W2hrho - p) / al2;
          for (int a = 0; a < 3; a++) {
            T0i[a] = W2hrho * vs[a] / al + p * be[a] / al2;
            T0j[a] = W2hrho * vl[a] / al;
          }
          for (int a = 0; a < 3; a++)
            for (int b = a; b < 3; b++) Tij[S6[a][b]] = W2hrho * vs[a] * vs[b] + p * (gi[S6[a][b]] - be[a] * be[b] / al2);
          /* tau */
          double bbK = 0., bda = 0., t0 = 0., tk = 0.;
          for (int a = 0; a < 3; a++) {
            bda += be[a] * dal[a];
            double bK = 0.;
            for (int b = 0; b < 3; b++) {
              bbK += be[a] * be[b] * K[S6[a][b]];
              bK += be[b] * K[S6[a][b]];
              tk += Tij[S6[a][b]] * K[S6[a][b]];
            }
            t0 += T0i[a] * (2. * bK - dal[a]);
          }
          sT = sqrtmdetg * (T00 * (bbK - bda) + t0 + tk);
          /* momentum */
          for (int d = 0; d < 3; d++) {
            double bbdg = 0., t0dg = 0., tdg = 0., tdb = 0.;
            for (int a = 0; a < 3; a++) {
              tdb += T0j[a] * dbe[d][a];
              for (int b = 0; b < 3; b++) {
                bbdg += be[a] * be[b] * dg[d][S6[a][b]];
                t0dg += T0i[a] * be[b] * dg[d][S6[a][b]];
                tdg += Tij[S6[a][b]] * dg[d][S6[a][b]];
              }
            }
            sS[d] = sqrtmdetg * (T00 * (0.5 * bbdg - al * dal[d]) + t0dg + tdb + 0.5 * tdg);
          }
          /* matter ADM variables: rho, S_i, S_ij, S */
          double vv = 0.;
          for (int a = 0; a < 3; a++) vv += vl[a] * vu[a];
          ad[0] = W2hrho - p;
          for (int a = 0; a < 3; a++) ad[1 + a] = W2hrho * vu[a];
          for (int a = 0; a < 3; a++)
            for (int b = a; b < 3; b++) ad[4 + S6[a][b]] = W2hrho * vl[a] * vl[b] + g[S6[a][b]] * p;
          ad[10] = W2hrho * vv + 3. * p;
          if (rho < rhoatm) sS[0] = sS[1] = sS[2] = 0.;
          if (P->eos == 1) sT = 0.;
          int bad = !isfinite(sD) || !isfinite(sT);
          for (int a = 0; a < 3; a++) bad |= !isfinite(sS[a]);
          for (int a = 0; a < 11; a++) bad |= !isfinite(ad[a]);
          if (bad) {
            sD = sT = sS[0] = sS[1] = sS[2] = 0.;
            for (int a = 0; a < 11; a++) ad[a] = 0.;
          }
        }
        for (int a = 0; a < 11; a++) madm[a][q] = ad[a];
        const double s5[5] = {sD, sT, sS[0], sS[1], sS[2]};
        for (int v = 0; v < 5; v++) rhs[v][q] = c != 0. ? rhs[v][q] + c * s5[v] : rhs[v][q] + s5[v];
      }
}
