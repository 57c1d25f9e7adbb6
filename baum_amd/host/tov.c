/* tov.c -- the TOV star of a polytrope, integrated in the log enthalpy.
 *
 * tov_h (src/projects/BBIDdataReader/tov.c:104-236, the default
 * BBID_tov_integrate = h, Lindblom 1992): from the centre value hc down to
 * zero in npts - 1 RK4 steps of -hc/npts for r(h), m(h) and the isotropic
 * integral I(h) (tov_h_rhs, tov.c:79-101), started from the series
 * expansion one step below the centre.  Then per point the potential
 * phi = phi(R) + h(R) - h with phi(R) = ln(1 - 2M/R)/2, the matter from the
 * enthalpy (eos_poly_H, src/projects/eos/eos_analytic.c:76-92) and the
 * isotropic radius r C exp(I) with C fixed by the exterior Schwarzschild
 * solution at the surface; the baryonic mass as computeMb (tov.c:10-24).
 * Polytrope p = K rho^Gamma only.
 */
#include <math.h>
#include <stdlib.h>

#include "../../include/bamhip.h"

typedef struct {
  double K, G;
} poly_t;

/* eos_poly_H: p, rho and epsl of the log enthalpy h (negative h as 0) */
static void eos_h(const poly_t *E, double h, double *p, double *rho, double *epsl)
{
  if (h < 0.) h = 0.;
  const double G1 = E->G - 1.;
  const double H1 = 2. * sinh(0.5 * h) * exp(0.5 * h); /* exp(h) - 1 */
  const double n = pow(G1 / (E->G * E->K) * H1, 1.0 / G1);
  const double krhotog1 = E->K * pow(n, G1);
  *p = n * krhotog1;
  *rho = n;
  *epsl = krhotog1 / G1;
}

/* total energy density and pressure */
static void eos_ep(const poly_t *E, double h, double *e, double *p)
{
  double rho, epsl;
  eos_h(E, h, p, &rho, &epsl);
  *e = rho * (1. + epsl);
}

/* d(h, r, m, I)/dh */
static void rhs(const poly_t *E, const double *u, double *k)
{
  const double r = u[1], m = u[2];
  double e, p;
  eos_ep(E, u[0], &e, &p);
  const double r3 = r * r * r;
  const double tmp = -(r - 2. * m) / (m + 4 * M_PI * r3 * p);
  const double drdh = r * tmp;
  const double f = sqrt(1. - 2. * m / r);
  k[0] = 0;
  k[1] = drdh;
  k[2] = 4 * M_PI * e * r3 * tmp;
  k[3] = (1. - f) / (r * f) * drdh;
}

int bamhip_tov_h(double hc, double K, double Gamma, int npts, double *r, double *m, double *h, double *rho,
                 double *pre, double *phi, double *riso, double *out)
{
  if (!(hc > 0) || !(K > 0) || !(Gamma > 1) || npts < 4 || !r || !m || !h || !rho || !pre || !phi || !riso)
    return BAMHIP_EINVAL;
  const poly_t E = {K, Gamma};
  double(*u)[4] = malloc(sizeof(double[4]) * (size_t)npts);
  if (!u) return BAMHIP_ENOMEM;
  const double dh = hc / npts;
  double ec, pc, e, p;
  eos_ep(&E, hc, &ec, &pc);
  /* series start one step below the centre */
  const double h0 = hc - dh;
  eos_ep(&E, h0, &e, &p);
  const double decdh = (e - ec) / (h0 - hc);
  const double r0 = sqrt(3. * dh / (2. * M_PI * (ec + 3. * pc))) *
                    (1. - 0.25 * (ec - 3. * pc - 3. / 5. * decdh) * (dh / (ec + 3. * pc)));
  u[0][0] = h0;
  u[0][1] = r0;
  u[0][2] = 4. / 3. * M_PI * ec * r0 * r0 * r0 * (1. - 3. / 5. * decdh * dh / ec);
  u[0][3] = 0;
  const double stp = -dh;
  for (int n = 0; n < npts - 1; n++) {
    double u1[4], u2[4], u3[4], k[4];
    int v;
    rhs(&E, u[n], k);
    for (v = 0; v < 4; v++) u1[v] = u[n][v] + 0.5 * stp * k[v];
    rhs(&E, u1, k);
    for (v = 0; v < 4; v++) u2[v] = u[n][v] + 0.5 * stp * k[v];
    rhs(&E, u2, k);
    for (v = 0; v < 4; v++) u3[v] = u[n][v] + stp * k[v];
    rhs(&E, u3, k);
    /* u = (2 (u1 + u3 - u) + 4 u2 + stp k) / 6, the classical RK4 sum */
    for (v = 0; v < 4; v++) u[n + 1][v] = (1. / 6.) * (2. * (-u[n][v] + u1[v] + u3[v]) + 4. * u2[v] + stp * k[v]);
    u[n + 1][0] += stp;
  }
  const double hR = u[npts - 1][0], R = u[npts - 1][1], M = u[npts - 1][2], IR = u[npts - 1][3];
  const double phiR = 0.5 * log(1. - 2. * M / R);
  const double C = 1 / (2 * R) * (sqrt(R * R - 2 * M * R) + R - M) * exp(-IR);
  for (int n = 0; n < npts; n++) {
    double epsl;
    h[n] = u[n][0];
    r[n] = u[n][1];
    m[n] = u[n][2];
    phi[n] = phiR + hR - h[n];
    eos_h(&E, h[n], &pre[n], &rho[n], &epsl);
    riso[n] = r[n] * C * exp(u[n][3]);
  }
  free(u);
  if (out) {
    double sum = 0.;
    for (int i = 1; i < npts - 1; i++) /* computeMb(..., npts - 1) */
      sum += r[i] * r[i] * rho[i] * (1. / sqrt(1. - 2. * m[i] / r[i])) * (r[i] - r[i - 1]);
    out[0] = M;
    out[1] = R;
    out[2] = riso[npts - 1];
    out[3] = 4. * M_PI * sum;
  }
  return BAMHIP_OK;
}
