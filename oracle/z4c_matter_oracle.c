/* z4c_matter_oracle.c -- CPU restatement of the matter terms of the Z4c
 * right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * z4_rhs_add_source (src/projects/z4/z4_rhs_add_source.m:14-77, generated
 * z4_rhs_add_source.c; registered as "geometry_source" at z4.c:247 and called
 * after the matter RHS in evolve_rhs, evolve.c:72-76) at the inner points:
 *   psi^4 = exp(4 log(max(chi, z4_chi_div_floor)) / z4_chi_psipower),
 *   A_ij  += -8 pi alpha psi^-4 (S_ij - psi^4 g~_ij S / 3),
 *   Khat  += 4 pi alpha (rho + S),
 *   G^i   += -16 pi alpha psi^4 S^i,
 *   Theta += -8 pi alpha rho,
 *   B^i   += withB (gamma0factor + gamma2factor shiftgammacoeff
 *            alpha^shiftalphapower) (the G^i term),
 * all of them 0 when any (or the unused -16/3 pi alpha g~_ij rho) is not
 * finite; times c when c != 0.
 */
#include <math.h>

#include "oracle.h"

#define OB_PI 3.1415926535897932 /* bam_main.h:14 */

void ob_z4c_matter_source(const ob_box *B, const ob_z4c_params *P, double c, const double *const *u,
                          const double *const *madm, double *const *r)
{
  const long m = B->m, mn = (long)B->m * B->n;
  for (int k = B->skip[4]; k < B->o - B->skip[5]; k++)
    for (int j = B->skip[2]; j < B->n - B->skip[3]; j++)
      for (int i = B->skip[0]; i < B->m - B->skip[1]; i++) {
        const long q = i + m * j + mn * k;
        double g[6];
        for (int a = 0; a < 6; a++) g[a] = u[a][q];
        const double al = u[18][q];
        double chig = u[6][q];
        if (chig < P->chiDivFloor) chig = P->chiDivFloor;
        const double f = (1.0 / P->chipsipower) * log(chig);
        const double psi4 = exp(4. * f), psim4 = 1. / psi4;
        const double rho = madm[0][q], ST = madm[10][q];
        double rA[6], r2A[6], rG[3], rB[3];
        for (int a = 0; a < 6; a++) {
          const double sstf = madm[4 + a][q] - (1. / 3.) * (psi4 * g[a]) * ST;
          rA[a] = -psim4 * 8. * OB_PI * al * sstf;
          r2A[a] = -(1. / 3.) * g[a] * 16. * OB_PI * al * rho;
        }
        double rKhat = 4. * OB_PI * al * (rho + ST);
        double rTheta = al * (-16. * OB_PI * rho) / 2.;
        const double betaF = P->shiftgammacoeff * pow(al, P->shiftalphapower);
        for (int a = 0; a < 3; a++) {
          rG[a] = 2. * al * (-8. * OB_PI * psi4 * madm[1 + a][q]);
          rB[a] = P->withB * (P->gamma0factor + P->gamma2factor * betaF) * rG[a];
        }
        int bad = !isfinite(rKhat) || !isfinite(rTheta);
        for (int a = 0; a < 6; a++) bad |= !isfinite(rA[a]) || !isfinite(r2A[a]);
        for (int a = 0; a < 3; a++) bad |= !isfinite(rG[a]) || !isfinite(rB[a]);
        if (bad) {
          rKhat = rTheta = 0.;
          for (int a = 0; a < 6; a++) rA[a] = 0.;
          for (int a = 0; a < 3; a++) rG[a] = rB[a] = 0.;
        }
        const double cc = c != 0. ? c : 1.;
        for (int a = 0; a < 6; a++) r[7 + a][q] += cc * rA[a];
        r[13][q] += cc * rKhat;
        for (int a = 0; a < 3; a++) r[14 + a][q] += cc * rG[a];
        r[17][q] += cc * rTheta;
        for (int a = 0; a < 3; a++) r[22 + a][q] += cc * rB[a];
      }
}
