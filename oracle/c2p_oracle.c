/* c2p_oracle.c -- TEST INFRASTRUCTURE: CPU restatement of BAM's GRHD
 * conservative-to-primitive recovery and its helpers.  Checker only.
 *
 * Follows src/projects/grhd/grhd_c2p_v2.c:24-253 (try_hot_part,
 * try_cold_part), :265-464 (grhd_c2p_proot_ColdStaticAtm_hybrid point loop),
 * src/projects/grhd/grhd_p2c.c:11-86 (prim -> cons), the analytic EOS of
 * src/projects/eos/eos_analytic.c:12-72,423-434 and invg of
 * src/projects/adm/adm_utility.c:27-46, and grhd_c2p_vanal of
 * src/projects/grhd/grhd_c2p.c:239-379.
 */
#include <math.h>
#include "oracle.h"

#define DMAX(a, b) ((a) > (b) ? (a) : (b))
#define BIG 1e+15

double ob_invg(double g11, double g12, double g13, double g22, double g23, double g33,
               double *i11, double *i12, double *i13, double *i22, double *i23, double *i33)
{
  const double c11 = g22 * g33 - g23 * g23;
  const double c12 = g13 * g23 - g12 * g33;
  const double c13 = g12 * g23 - g13 * g22;
  const double c22 = g11 * g33 - g13 * g13;
  const double c23 = g12 * g13 - g11 * g23;
  const double c33 = g11 * g22 - g12 * g12;
  const double det = g11 * c11 + g12 * c12 + g13 * c13;
  *i11 = c11 / det;
  *i12 = c12 / det;
  *i13 = c13 / det;
  *i22 = c22 / det;
  *i23 = c23 / det;
  *i33 = c33 / det;
  return det;
}

/* EOS.comp("re",...,"pc","re"): p(rho,eps) and its partial derivatives.
 * Ideal gas: p = (G-1) rho eps (eos_ideal); polytrope: p = K rho^G (eos_poly).
 * Returns the "problem" flag (non-finite input/output). */
static int eos_re(const ob_c2p_params *P, double rho, double eps, double *p, double *chi, double *kappa)
{
  double cs2;
  if (P->eos == 0) {
    *p = (P->gamma - 1) * rho * eps;
    *kappa = (P->gamma - 1) * rho;
    *chi = (P->gamma - 1) * eps;
  } else {
    const double kr = P->K * pow(rho, P->gamma - 1.);
    *p = rho * kr;
    *kappa = 0.;
    *chi = P->gamma * kr;
  }
  /* eos_cs2_rep is only used for its finiteness check */
  cs2 = (rho == 0.) ? 0. : (*chi + *p * *kappa / (rho * rho)) / (1.0 + eps + *p / rho);
  return (!isfinite(rho) || !isfinite(eps) || !isfinite(cs2)) ? 1 : 0;
}

/* cold part EOS.comp("r",...,"pe","r"): polytrope with eps reset */
static void eos_cold(const ob_c2p_params *P, double rho, double *p, double *eps, double *chi)
{
  const double kr = P->K * pow(rho, P->gamma - 1.);
  *p = rho * kr;
  *chi = P->gamma * kr;
  *eps = kr / (P->gamma - 1.);
}

int ob_try_hot(const ob_c2p_params *P, double conD, double conS2, double conT,
               double *rho, double *epsl, double *pres, double init, int *iters)
{
  const double errormax = P->errormax;
  const int countmax = P->countmax;
  double pmin = sqrt(conS2) - conD - conT;
  pmin = DMAX(pmin, P->patm);
  double pold = DMAX(init, pmin);
  *iters = 0;
  if (!isfinite(pold)) return OB_C2P_FAIL;
  double temp = conT + pold + conD;
  double tempSQRTarg = temp * temp - conS2;
  double tempSQRT = sqrt(tempSQRTarg);
  double Wlor = temp / tempSQRT;
  *rho = conD / Wlor;
  *epsl = (tempSQRT - Wlor * pold) / conD - 1.;
  if (tempSQRTarg <= 0.0) return OB_C2P_COLD;

  int count;
  double error = 2 * errormax;
  for (count = 0; count <= countmax; count++) {
    double peos, chi, kappa;
    const int problem = eos_re(P, *rho, *epsl, &peos, &chi, &kappa);
    const double f = pold - peos;
    const double drhodp = conD * conS2 / (tempSQRT * temp * temp);
    const double depsldp = pold * conS2 / (conD * tempSQRTarg * tempSQRT);
    const double df = 1. - chi * drhodp - kappa * depsldp;
    *pres = DMAX(pmin, pold - f / df);
    error = fabs(1. - pold / (*pres));
    pold = *pres;
    temp = conT + pold + conD;
    tempSQRTarg = temp * temp - conS2;
    tempSQRT = sqrt(tempSQRTarg);
    Wlor = temp / tempSQRT;
    *rho = conD / Wlor;
    *epsl = (tempSQRT - Wlor * pold) / conD - 1.;
    *iters = count + 1;
    if (*rho < P->fatm * P->rhoatm) return OB_C2P_ATM;
    if (problem) return OB_C2P_ATM;
    if (df == 0.) return OB_C2P_NR;
    if (error < errormax) break;
  }
  if (count >= countmax) return OB_C2P_NR;
  if (!isfinite(*pres) || !isfinite(*rho) || !isfinite(*epsl)) return OB_C2P_COLD;
  if (*epsl < 0.) return OB_C2P_COLD;
  const double v2 = conS2 / ((conT + conD + *pres) * (conT + conD + *pres));
  if (v2 > P->vmax) return OB_C2P_ATM;
  return OB_C2P_OK;
}

int ob_try_cold(const ob_c2p_params *P, double conD, double conS2, double *conT,
                double *rho, double *epsl, double *pres, double init, int *iters)
{
  const double errormax = P->errormax;
  const int countmax = P->countmax;
  double h, dhdrho, chi, tempSQRTarg, Wlor, g, dg;
  *iters = 0;
  if (conD <= 0.) return OB_C2P_ATM;
  const double rhomin = P->rhoatm;
  double rhold = DMAX(init, rhomin);
  if (!isfinite(rhold)) return OB_C2P_FAIL;
  int count;
  double error = 2 * errormax;
  for (count = 0; count <= countmax; count++) {
    eos_cold(P, rhold, pres, epsl, &chi);
    h = 1.0 + *epsl + *pres / rhold;
    dhdrho = chi / rhold;
    if (conD * h == 0.) tempSQRTarg = 1.0 + conS2 * BIG;
    else tempSQRTarg = 1.0 + conS2 / ((conD * h) * (conD * h));
    Wlor = sqrt(tempSQRTarg);
    g = Wlor * rhold - conD;
    dg = Wlor - (rhold * conS2 * dhdrho) / (Wlor * conD * conD * h * h * h);
    *iters = count + 1;
    if ((!isfinite(dg)) || (dg == 0.0)) return OB_C2P_NR;
    *rho = DMAX(rhomin, rhold - g / dg);
    error = fabs(1. - rhold / (*rho));
    rhold = *rho;
    if (error < errormax) break;
  }
  if (count >= countmax) return OB_C2P_NR;
  eos_cold(P, rhold, pres, epsl, &chi);
  h = 1.0 + *epsl + *pres / rhold;
  if (conD * h == 0.) tempSQRTarg = 1.0 + conS2 * BIG;
  else tempSQRTarg = 1.0 + conS2 / ((conD * h) * (conD * h));
  Wlor = sqrt(tempSQRTarg);
  if (*rho < P->fatm * P->rhoatm) return OB_C2P_ATM;
  *conT = Wlor * Wlor * (*rho) * h - *pres - conD;
  const double v2 = conS2 / ((*conT + conD + *pres) * (*conT + conD + *pres));
  if (v2 > P->vmax) return OB_C2P_ATM;
  return OB_C2P_OK;
}

/* one point of grhd_c2p_proot_ColdStaticAtm_hybrid (grhd_c2p_v2.c:303-460);
 * the NaN diagnostics of the reference only print, so they are omitted
 * except where they change the control flow (result = FAIL). */
void ob_c2p_hybrid(const ob_c2p_params *P, long n, double *cons, const double *metric,
                   double *prim, int *status, int *iters)
{
#pragma omp parallel for schedule(dynamic, 4096)
  for (long q = 0; q < n; q++) {
    double *MD = cons + 0 * n, *MT = cons + 1 * n, *MSx = cons + 2 * n, *MSy = cons + 3 * n, *MSz = cons + 4 * n;
    const double *gxx = metric + 0 * n, *gxy = metric + 1 * n, *gxz = metric + 2 * n;
    const double *gyy = metric + 3 * n, *gyz = metric + 4 * n, *gzz = metric + 5 * n;
    double *Mrho = prim + 0 * n, *Mepsl = prim + 1 * n, *Mvx = prim + 2 * n, *Mvy = prim + 3 * n;
    double *Mvz = prim + 4 * n, *Mp = prim + 5 * n, *Mv2 = prim + 6 * n, *detg = prim + 7 * n;
    double gixx, gixy, gixz, giyy, giyz, gizz;
    double rho = 0, epsl = 0, pres = 0;
    int it = 0, it2 = 0;

    detg[q] = ob_invg(gxx[q], gxy[q], gxz[q], gyy[q], gyz[q], gzz[q], &gixx, &gixy, &gixz, &giyy, &giyz, &gizz);
    double SQRTdetg = sqrt(fabs(detg[q]));
    if (!isfinite(SQRTdetg)) {
      detg[q] = 1.;
      SQRTdetg = 1.;
    }
    const double oo = 1. / SQRTdetg;
    double conD = MD[q] * oo, conSx = MSx[q] * oo, conSy = MSy[q] * oo, conSz = MSz[q] * oo, conT = MT[q] * oo;
    double conSix = gixx * conSx + gixy * conSy + gixz * conSz;
    double conSiy = gixy * conSx + giyy * conSy + giyz * conSz;
    double conSiz = gixz * conSx + giyz * conSy + gizz * conSz;
    double conS2 = conSx * conSix + conSy * conSiy + conSz * conSiz;
    double SQRTconS2 = sqrt(fabs(conS2));
    int result = OB_C2P_OK;
    if (!isfinite(conD) || !isfinite(conSx) || !isfinite(conSy) || !isfinite(conSz) || !isfinite(conT))
      result = OB_C2P_FAIL;
    if (!isfinite(gxx[q]) || !isfinite(gxy[q]) || !isfinite(gxz[q]) || !isfinite(gyy[q]) || !isfinite(gyz[q])
        || !isfinite(gzz[q]) || !isfinite(detg[q]) || !isfinite(gixx) || !isfinite(gixy) || !isfinite(gixz)
        || !isfinite(giyy) || !isfinite(giyz) || !isfinite(gizz))
      result = OB_C2P_FAIL;
    if (!isfinite(conSix) || !isfinite(conSiy) || !isfinite(conSiz) || !isfinite(conS2) || !isfinite(SQRTconS2)) {
      conSx = conSix = 0.;
      conSy = conSiy = 0.;
      conSz = conSiz = 0.;
      conS2 = SQRTconS2 = 0.;
      result = OB_C2P_FAIL;
    }
    if (result != OB_C2P_FAIL) result = ob_try_hot(P, conD, conS2, conT, &rho, &epsl, &pres, Mp[q], &it);
    if (result == OB_C2P_COLD) {
      result = ob_try_cold(P, conD, conS2, &conT, &rho, &epsl, &pres, Mrho[q], &it2);
      MT[q] = SQRTdetg * conT;
    }
    if (result != OB_C2P_OK && result != OB_C2P_ATM) {
      /* failures fall back to the atmosphere; the recorded status keeps the cause */
      status[q] = result;
      result = OB_C2P_ATM;
    } else {
      status[q] = result;
    }
    if (result == OB_C2P_ATM) {
      pres = P->patm;
      rho = P->rhoatm;
      epsl = P->epslatm;
      conD = P->rhoatm;
      conSx = conSix = 0.;
      conSy = conSiy = 0.;
      conSz = conSiz = 0.;
      conS2 = 0.;
      conT = P->rhoatm * P->epslatm;
      MD[q] = SQRTdetg * conD;
      MSx[q] = SQRTdetg * conSx;
      MSy[q] = SQRTdetg * conSy;
      MSz[q] = SQRTdetg * conSz;
      MT[q] = SQRTdetg * conT;
    }
    const double ooDpT = 1.0 / (conD + pres + conT);
    Mrho[q] = rho;
    Mepsl[q] = epsl;
    Mp[q] = pres;
    Mvx[q] = conSix * ooDpT;
    Mvy[q] = conSiy * ooDpT;
    Mvz[q] = conSiz * ooDpT;
    {
      const double vlx = gxx[q] * Mvx[q] + gxy[q] * Mvy[q] + gxz[q] * Mvz[q];
      const double vly = gxy[q] * Mvx[q] + gyy[q] * Mvy[q] + gyz[q] * Mvz[q];
      const double vlz = gxz[q] * Mvx[q] + gyz[q] * Mvy[q] + gzz[q] * Mvz[q];
      Mv2[q] = vlx * Mvx[q] + vly * Mvy[q] + vlz * Mvz[q];
    }
    iters[q] = it + it2;
  }
}

/* prim (rho, eps, vx, vy, vz with contravariant v^i) -> densitised cons
 * (D, T, Sx, Sy, Sz): grhd_compute_v2_pt + grhd_compute_q_pt, with the
 * pressure from the EOS and W = 1/sqrt(1-v^2). */
void ob_p2c(const ob_c2p_params *P, long n, const double *prim5, const double *metric, double *cons)
{
  for (long q = 0; q < n; q++) {
    const double rho = prim5[q], eps = prim5[n + q];
    const double vx = prim5[2 * n + q], vy = prim5[3 * n + q], vz = prim5[4 * n + q];
    const double gxx = metric[q], gxy = metric[n + q], gxz = metric[2 * n + q];
    const double gyy = metric[3 * n + q], gyz = metric[4 * n + q], gzz = metric[5 * n + q];
    double p, chi, kappa;
    eos_re(P, rho, eps, &p, &chi, &kappa);
    const double vlx = gxx * vx + gxy * vy + gxz * vz;
    const double vly = gxy * vx + gyy * vy + gyz * vz;
    const double vlz = gxz * vx + gyz * vy + gzz * vz;
    const double v2 = vlx * vx + vly * vy + vlz * vz;
    const double W = 1.0 / sqrt(1.0 - v2);
    double i11, i12, i13, i22, i23, i33;
    const double detg = ob_invg(gxx, gxy, gxz, gyy, gyz, gzz, &i11, &i12, &i13, &i22, &i23, &i33);
    const double sq = sqrt(detg);
    const double W2hrho = W * W * (rho + rho * eps + p);
    cons[q] = sq * W * rho;
    cons[n + q] = sq * (W2hrho - p - W * rho);
    cons[2 * n + q] = sq * W2hrho * vlx;
    cons[3 * n + q] = sq * W2hrho * vly;
    cons[4 * n + q] = sq * W2hrho * vlz;
  }
}

void ob_c2p_proot(const ob_c2p_params *P, long n, const double *cons, const double *metric,
                  double *prim, int *status, int *iters)
{
#pragma omp parallel for schedule(dynamic, 4096)
  for (long q = 0; q < n; q++) {
    const double *gm = metric;
    double gixx, gixy, gixz, giyy, giyz, gizz;
    const double detg = ob_invg(gm[q], gm[n + q], gm[2 * n + q], gm[3 * n + q], gm[4 * n + q], gm[5 * n + q],
                                &gixx, &gixy, &gixz, &giyy, &giyz, &gizz);
    prim[7 * n + q] = detg;
    status[q] = OB_C2P_OK;
    iters[q] = 0;
    if (detg <= 0. || !isfinite(detg)) {
      status[q] = OB_C2P_FAIL;
      continue;
    }
    const double oo = 1. / sqrt(detg);
    const double conD = cons[q] * oo, conT = cons[n + q] * oo;
    const double conSx = cons[2 * n + q] * oo, conSy = cons[3 * n + q] * oo, conSz = cons[4 * n + q] * oo;
    const double conSix = gixx * conSx + gixy * conSy + gixz * conSz;
    const double conSiy = gixy * conSx + giyy * conSy + giyz * conSz;
    const double conSiz = gixz * conSx + giyz * conSy + gizz * conSz;
    const double conS2 = conSx * conSix + conSy * conSiy + conSz * conSiz;
    const double SQRTconS2 = sqrt(conS2);
    if (!isfinite(conD) || !isfinite(conSx) || !isfinite(conSy) || !isfinite(conSz) || !isfinite(conT)
        || !isfinite(SQRTconS2)) {
      status[q] = OB_C2P_FAIL;
      continue;
    }
    const double pmin = SQRTconS2 - conD - conT;
    double pold = DMAX(prim[5 * n + q], pmin);
    double temp = conT + pold + conD, arg = temp * temp - conS2, sq = sqrt(arg), W = temp / sq;
    double rho = conD / W, eps = (sq - W * pold) / conD - 1.;
    if (!isfinite(pold) || arg < 0.0) {
      status[q] = OB_C2P_FAIL;
      continue;
    }
    double pnew = pold, error = 2 * P->errormax;
    int count;
    for (count = 0; count <= P->countmax; count++) {
      double peos, chi, kappa;
      eos_re(P, rho, eps, &peos, &chi, &kappa);
      const double f = pold - peos;
      const double drhodp = conD * conS2 / (sq * temp * temp);
      const double depsldp = pold * conS2 / (conD * arg * sq);
      const double df = 1. - chi * drhodp - kappa * depsldp;
      if (df == 0.0) {
        status[q] = OB_C2P_NR;
        break;
      }
      pnew = DMAX(pmin, pold - f / df);
      error = fabs(1. - pold / pnew);
      pold = pnew;
      temp = conT + pold + conD;
      arg = temp * temp - conS2;
      sq = sqrt(arg);
      W = temp / sq;
      rho = conD / W;
      eps = (sq - W * pold) / conD - 1.;
      iters[q] = count + 1;
      if (error < P->errormax) break;
    }
    if (status[q] == OB_C2P_OK && count >= P->countmax) status[q] = OB_C2P_NR;
    if (status[q] == OB_C2P_OK && eps < 0.) status[q] = OB_C2P_FAIL;
    prim[q] = rho;
    prim[n + q] = eps;
    prim[5 * n + q] = pnew;
    prim[2 * n + q] = conSix / (conD + pnew + conT);
    prim[3 * n + q] = conSiy / (conD + pnew + conT);
    prim[4 * n + q] = conSiz / (conD + pnew + conT);
    const double vx = prim[2 * n + q], vy = prim[3 * n + q], vz = prim[4 * n + q];
    const double vlx = gm[q] * vx + gm[n + q] * vy + gm[2 * n + q] * vz;
    const double vly = gm[n + q] * vx + gm[3 * n + q] * vy + gm[4 * n + q] * vz;
    const double vlz = gm[2 * n + q] * vx + gm[4 * n + q] * vy + gm[5 * n + q] * vz;
    prim[6 * n + q] = vlx * vx + vly * vy + vlz * vz;
  }
}

/* Bracketed Newton-Raphson on the pressure (this build's safeguarded solver;
 * no reference counterpart).  Same f(p) = p - P(rho(p), eps(p)) and
 * derivative as try_hot_part, with a bracket [plo, phi] that always holds
 * the sign change: plo = max(|S|-D-tau, 0), phi = (Gamma-1)(tau+D) for the
 * ideal gas (rho eps <= tau + D).  A Newton step leaving the bracket is
 * replaced by bisection.  Stops when the relative update or the relative
 * bracket width is below errormax, or when |f| is within its rounding floor
 * (4 eps_mach times the magnitudes entering f, amplified by W^2 where
 * sqrt(temp^2 - S^2) cancels, so that cold fast flows, whose
 * pressure the conserved variables fix only to ~1e-8, do not iterate on
 * noise), or after countmax iterations (status NR).
 * guess <= 0 starts from the bracket's lower end. */
int ob_c2p_bracketed_pt(const ob_c2p_params *P, double conD, double conS2, double conT, double guess,
                        double *rho, double *eps, double *pres, int *iters)
{
  double plo = sqrt(conS2) - conD - conT;
  if (plo < 0.) plo = 0.;
  double phi = (P->gamma - 1.) * (conT + conD);
  if (!(phi > plo)) phi = plo + 1e-300;
  double p = (guess > plo && guess < phi) ? guess : plo;
  const double em = 2.220446049250313e-16;
  int count;
  *iters = 0;
  for (count = 0; count < P->countmax; count++) {
    const double temp = conT + p + conD;
    const double arg = temp * temp - conS2;
    const double sq = sqrt(arg);
    const double W = temp / sq;
    const double r = conD / W;
    const double e = (sq - W * p) / conD - 1.;
    double peos, chi, kappa;
    eos_re(P, r, e, &peos, &chi, &kappa);
    const double f = p - peos;
    /* rounding floor of f: sq = sqrt(temp^2 - S^2) carries a relative error
       ~ eps_mach W^2 (temp^2 / arg = W^2), which enters eps through sq / D */
    const double noise = 4. * em * (fabs(p) + fabs(peos) + fabs(chi) * r + fabs(kappa) * (W * W * sq + W * fabs(p)) / conD);
    *iters = count + 1;
    if (fabs(f) <= noise) break;
    if (f < 0.) plo = p; else phi = p;
    const double drhodp = conD * conS2 / (sq * temp * temp);
    const double depsldp = p * conS2 / (conD * arg * sq);
    const double df = 1. - chi * drhodp - kappa * depsldp;
    double pn = p - f / df;
    if (!(pn >= plo && pn <= phi)) pn = 0.5 * (plo + phi);
    const double step = fabs(pn - p);
    p = pn;
    if (step <= P->errormax * p || phi - plo <= P->errormax * p) break;
  }
  const double temp = conT + p + conD;
  const double sq = sqrt(temp * temp - conS2);
  const double W = temp / sq;
  *rho = conD / W;
  *eps = (sq - W * p) / conD - 1.;
  *pres = p;
  if (count >= P->countmax) return OB_C2P_NR;
  if (!isfinite(p) || !isfinite(*rho) || !isfinite(*eps)) return OB_C2P_FAIL;
  return OB_C2P_OK;
}

void ob_c2p_bracketed(const ob_c2p_params *P, long n, const double *cons, const double *metric,
                      double *prim, int *status, int *iters)
{
#pragma omp parallel for schedule(dynamic, 4096)
  for (long q = 0; q < n; q++) {
    const double *gm = metric;
    double gixx, gixy, gixz, giyy, giyz, gizz;
    const double detg = ob_invg(gm[q], gm[n + q], gm[2 * n + q], gm[3 * n + q], gm[4 * n + q], gm[5 * n + q],
                                &gixx, &gixy, &gixz, &giyy, &giyz, &gizz);
    prim[7 * n + q] = detg;
    const double oo = 1. / sqrt(detg);
    const double conD = cons[q] * oo, conT = cons[n + q] * oo;
    const double conSx = cons[2 * n + q] * oo, conSy = cons[3 * n + q] * oo, conSz = cons[4 * n + q] * oo;
    const double conSix = gixx * conSx + gixy * conSy + gixz * conSz;
    const double conSiy = gixy * conSx + giyy * conSy + giyz * conSz;
    const double conSiz = gixz * conSx + giyz * conSy + gizz * conSz;
    const double conS2 = conSx * conSix + conSy * conSiy + conSz * conSiz;
    double rho, eps, p;
    status[q] = ob_c2p_bracketed_pt(P, conD, conS2, conT, prim[5 * n + q], &rho, &eps, &p, &iters[q]);
    const double ooDpT = 1.0 / (conD + p + conT);
    const double vx = conSix * ooDpT, vy = conSiy * ooDpT, vz = conSiz * ooDpT;
    prim[q] = rho;
    prim[n + q] = eps;
    prim[2 * n + q] = vx;
    prim[3 * n + q] = vy;
    prim[4 * n + q] = vz;
    prim[5 * n + q] = p;
    const double vlx = gm[q] * vx + gm[n + q] * vy + gm[2 * n + q] * vz;
    const double vly = gm[n + q] * vx + gm[3 * n + q] * vy + gm[4 * n + q] * vz;
    const double vlz = gm[2 * n + q] * vx + gm[4 * n + q] * vy + gm[5 * n + q] * vz;
    prim[6 * n + q] = vlx * vx + vly * vy + vlz * vz;
  }
}

/* grhd_atm_set_u (src/projects/grhd/grhd_atm.c:217-262).  The levels
 * rhoatmlevel = RHOATM*FATM, Datmlevel = RHOATM, Tatmlevel = RHOATM*EPSLATM
 * are formed first (:234-236); DETGMIN is grhd.h's 1e-5 (grhd_atm.c includes
 * grhd.h); DMAX(detg, DETGMIN) keeps the macro's comparison, so a NaN detg
 * becomes DETGMIN. */
long ob_grhd_atm_set_u(const ob_c2p_params *P, long n, double *cons, const double *rho, double *detg)
{
  const double rhoatmlevel = P->rhoatm * P->fatm;
  const double Datmlevel = P->rhoatm;
  const double Tatmlevel = P->rhoatm * P->epslatm;
  double *MD = cons, *MT = cons + n, *MSx = cons + 2 * n, *MSy = cons + 3 * n, *MSz = cons + 4 * n;
  long bad = 0;
  for (long q = 0; q < n; q++) {
    if (rho[q] < rhoatmlevel) {
      double SQRTdetg;
      detg[q] = DMAX(detg[q], 1e-5);
      SQRTdetg = sqrt(detg[q]);
      if (!isfinite(SQRTdetg)) SQRTdetg = 1.;
      MD[q] = SQRTdetg * Datmlevel;
      MSx[q] = 0.;
      MSy[q] = 0.;
      MSz[q] = 0.;
      MT[q] = SQRTdetg * Tatmlevel;
    }
    if (!isfinite(MD[q]) || !isfinite(MSx[q]) || !isfinite(MSy[q]) || !isfinite(MSz[q]) || !isfinite(MT[q]) ||
        !isfinite(detg[q])) {
      MD[q] = MSx[q] = MSy[q] = MSz[q] = MT[q] = 0.;
      detg[q] = 1.;
      bad++;
    }
  }
  return bad;
}

/* Analytic Gamma-law recovery grhd_c2p_vanal (grhd_c2p.c:239-379): the
 * velocity magnitude v = |v| is the root of the quartic
 *   v^4 + a3 v^3 + a2 v^2 + a1 v + a0 = 0
 * taken through the resolvent cubic (Ferrari), clamped to [0, 1 - 1e-15]
 * (fmax/fmin, so a NaN root becomes 0), then rho = sqrt(1 - v^2) D,
 * p = (G-1) (E - S_i S_i v/|S| - rho), eps = p / ((G-1) rho).  Ideal gas
 * only (eos_Gamma).  The reference's errorexit cases (detg <= 0 or not
 * finite, non-finite conserved variables, eps < 0) become status FAIL;
 * v^2 (prim[6]) is not written, as in the reference.  coef (optional,
 * 4*n): a3, a2, a1, a0 per point (0 where |S| <= 1e-20). */
void ob_c2p_vanal(const ob_c2p_params *P, long n, const double *cons, const double *metric,
                  double *prim, int *status, double *coef)
{
  const double gammamo = P->gamma - 1.;
  const double oo3 = 1. / 3.;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < n; q++) {
    const double *gm = metric;
    double gixx, gixy, gixz, giyy, giyz, gizz;
    const double detg = ob_invg(gm[q], gm[n + q], gm[2 * n + q], gm[3 * n + q], gm[4 * n + q], gm[5 * n + q],
                                &gixx, &gixy, &gixz, &giyy, &giyz, &gizz);
    prim[7 * n + q] = detg;
    status[q] = OB_C2P_OK;
    if (coef) coef[q] = coef[n + q] = coef[2 * n + q] = coef[3 * n + q] = 0.;
    if (detg <= 0. || !isfinite(detg)) {
      status[q] = OB_C2P_FAIL;
      continue;
    }
    const double SQRTdetg = sqrt(detg);
    const double oo = 1. / SQRTdetg;
    const double conD = cons[q] * oo, conT = cons[n + q] * oo;
    const double conSx = cons[2 * n + q] * oo, conSy = cons[3 * n + q] * oo, conSz = cons[4 * n + q] * oo;
    const double conSix = gixx * conSx + gixy * conSy + gixz * conSz;
    const double conSiy = gixy * conSx + giyy * conSy + giyz * conSz;
    const double conSiz = gixz * conSx + giyz * conSy + gizz * conSz;
    const double conS2 = conSx * conSix + conSy * conSiy + conSz * conSiz;
    const double SQRTconS2 = sqrt(conS2);
    if (!isfinite(conD) || !isfinite(conSx) || !isfinite(conSy) || !isfinite(conSz) || !isfinite(conT) ||
        !isfinite(SQRTconS2)) {
      status[q] = OB_C2P_FAIL;
      continue;
    }
    const double D = conD, S2 = conS2, S = SQRTconS2, E = conT + conD;
    double v = 0.0;
    if (fabs(S) > 1e-20) {
      const double SE = S * E, D2 = D * D;
      const double Gamma = gammamo + 1, Gammamosq = gammamo * gammamo;
      const double div = 1.0 / (Gammamosq * (S2 + D2));
      const double a3 = (-2.0 * Gamma * gammamo * SE) * div;
      const double a2 = (Gamma * Gamma * E * E + 2.0 * gammamo * S2 - Gammamosq * D2) * div;
      const double a1 = (-2.0 * Gamma * SE) * div;
      const double a0 = S2 * div;
      if (coef) {
        coef[q] = a3;
        coef[n + q] = a2;
        coef[2 * n + q] = a1;
        coef[3 * n + q] = a0;
      }
      /* resolvent cubic y^3 + i1 y^2 + i2 y + i3, largest real root ix1 */
      const double i1 = -a2;
      const double i2 = a3 * a1 - 4.0 * a0;
      const double i3 = 4.0 * a2 * a0 - a1 * a1 - a3 * a3 * a0;
      const double iR = (9.0 * i1 * i2 - 27.0 * i3 - 2.0 * i1 * i1 * i1) / 54.0;
      const double iS = (3.0 * i2 - a2 * a2) / 9.0;
      const double iT = iR * iR + iS * iS * iS;
      double ix1;
      if (iT < 0.)
        ix1 = 2.0 * pow(sqrt(iR * iR + iT), oo3) * cos(atan2(sqrt(-iT), iR) * oo3) - i1 * oo3;
      else
        ix1 = pow(iR + sqrt(iT), oo3) + pow(iR - sqrt(iT), oo3) - i1 * oo3;
      /* the quadratic factor v^2 + iB v + iC holding the physical root */
      const double iB = 0.5 * (a3 + sqrt(a3 * a3 - 4.0 * a2 + 4.0 * ix1));
      const double iC = 0.5 * (ix1 - sqrt(ix1 * ix1 - 4.0 * a0));
      v = (-iB + sqrt(iB * iB - 4.0 * iC)) * 0.5;
    }
    v = fmax(v, 0.0);
    v = fmin(v, 1.0 - 1.0e-15);
    double div = v / S;
    if (fabs(S) < 1e-20) div = 0.0;
    const double rho = sqrt(1.0 - v * v) * D;
    const double vx = conSx * div, vy = conSy * div, vz = conSz * div;
    const double p = gammamo * ((E - conSx * vx - conSy * vy - conSz * vz) - rho);
    const double epsl = p / (gammamo * rho);
    if (epsl < 0.) {
      status[q] = OB_C2P_FAIL;
      continue;
    }
    div = 1. / (conD + p + conT);
    prim[q] = rho;
    prim[n + q] = epsl;
    prim[5 * n + q] = p;
    prim[2 * n + q] = conSix * div;
    prim[3 * n + q] = conSiy * div;
    prim[4 * n + q] = conSiz * div;
  }
}
