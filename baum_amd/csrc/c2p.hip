// c2p.hip -- GRHD conservative-to-primitive recovery on gfx950, FP64.
//
// Follows src/projects/grhd/grhd_c2p_v2.c:24-253 (try_hot_part: Newton-
// Raphson on the pressure with the general EOS; try_cold_part: Newton-
// Raphson on the density with the cold polytrope), :265-464 (the hybrid
// point loop with the cold static atmosphere, which rewrites the conserved
// variables on COLD/ATM outcomes) and src/projects/grhd/grhd_c2p.c:102-237
// (the plain p-root without atmosphere).  EOS: ideal gas / polytrope of
// src/projects/eos/eos_analytic.c:52-72,423-434.
//
// One thread per point, SoA arrays, no shared state: the kernel is bound by
// the 21 doubles per point it moves (HBM) and by the FP64 div/sqrt of each
// Newton iteration; iteration counts differ from point to point, so lanes of
// a wave finish at different times (the iteration histogram is returned).
// FP contraction is off so that every rounding matches the CPU oracle and
// status codes and iteration counts agree exactly.
#include <cmath>

#include "bamhip_internal.h"
#include "mlp_device.h"

#pragma clang fp contract(off)

namespace bamhip {

enum { C2P_COLD = 0, C2P_ATM, C2P_OK, C2P_FAIL, C2P_NR };

struct C2PK {
  int eos;  // 0 ideal gas (cold part: polytrope), 1 polytrope
  double gamma, K, rhoatm, fatm, epslatm, patm, vmax, errormax;
  int countmax;
};

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// EOS.comp("re",...,"pc","re"): p(rho, eps) with chi = dp/drho, kappa = dp/deps
__device__ __forceinline__ int eos_re(const C2PK &P, double rho, double eps, double *p, double *chi, double *kappa) {
  if (P.eos == 0) {
    *p = (P.gamma - 1) * rho * eps;
    *kappa = (P.gamma - 1) * rho;
    *chi = (P.gamma - 1) * eps;
  } else {
    const double kr = P.K * pow(rho, P.gamma - 1.);
    *p = rho * kr;
    *kappa = 0.;
    *chi = P.gamma * kr;
  }
  const double cs2 = (rho == 0.) ? 0. : (*chi + *p * *kappa / (rho * rho)) / (1.0 + eps + *p / rho);
  return (!isfinite(rho) || !isfinite(eps) || !isfinite(cs2)) ? 1 : 0;
}

// cold polytrope EOS.comp("r",...,"pe","r")
__device__ __forceinline__ void eos_cold(const C2PK &P, double rho, double *p, double *eps, double *chi) {
  const double kr = P.K * pow(rho, P.gamma - 1.);
  *p = rho * kr;
  *chi = P.gamma * kr;
  *eps = kr / (P.gamma - 1.);
}

__device__ int try_hot(const C2PK &P, double conD, double conS2, double conT, double *rho, double *epsl, double *pres,
                       double init, int *iters) {
  double pmin = sqrt(conS2) - conD - conT;
  pmin = dmax(pmin, P.patm);
  double pold = dmax(init, pmin);
  *iters = 0;
  if (!isfinite(pold)) return C2P_FAIL;
  double temp = conT + pold + conD;
  double arg = temp * temp - conS2;
  double sq = sqrt(arg);
  double W = temp / sq;
  *rho = conD / W;
  *epsl = (sq - W * pold) / conD - 1.;
  if (arg <= 0.0) return C2P_COLD;
  int count;
  double error = 2 * P.errormax;
  for (count = 0; count <= P.countmax; count++) {
    double peos, chi, kappa;
    const int problem = eos_re(P, *rho, *epsl, &peos, &chi, &kappa);
    const double f = pold - peos;
    const double drhodp = conD * conS2 / (sq * temp * temp);
    const double depsldp = pold * conS2 / (conD * arg * sq);
    const double df = 1. - chi * drhodp - kappa * depsldp;
    *pres = dmax(pmin, pold - f / df);
    error = fabs(1. - pold / (*pres));
    pold = *pres;
    temp = conT + pold + conD;
    arg = temp * temp - conS2;
    sq = sqrt(arg);
    W = temp / sq;
    *rho = conD / W;
    *epsl = (sq - W * pold) / conD - 1.;
    *iters = count + 1;
    if (*rho < P.fatm * P.rhoatm) return C2P_ATM;
    if (problem) return C2P_ATM;
    if (df == 0.) return C2P_NR;
    if (error < P.errormax) break;
  }
  if (count >= P.countmax) return C2P_NR;
  if (!isfinite(*pres) || !isfinite(*rho) || !isfinite(*epsl)) return C2P_COLD;
  if (*epsl < 0.) return C2P_COLD;
  const double v2 = conS2 / ((conT + conD + *pres) * (conT + conD + *pres));
  if (v2 > P.vmax) return C2P_ATM;
  return C2P_OK;
}

__device__ int try_cold(const C2PK &P, double conD, double conS2, double *conT, double *rho, double *epsl,
                        double *pres, double init, int *iters) {
  const double BIG = 1e+15;
  double h, dhdrho, chi, arg, W, g, dg;
  *iters = 0;
  if (conD <= 0.) return C2P_ATM;
  const double rhomin = P.rhoatm;
  double rhold = dmax(init, rhomin);
  if (!isfinite(rhold)) return C2P_FAIL;
  int count;
  double error = 2 * P.errormax;
  for (count = 0; count <= P.countmax; count++) {
    eos_cold(P, rhold, pres, epsl, &chi);
    h = 1.0 + *epsl + *pres / rhold;
    dhdrho = chi / rhold;
    if (conD * h == 0.) arg = 1.0 + conS2 * BIG;
    else arg = 1.0 + conS2 / ((conD * h) * (conD * h));
    W = sqrt(arg);
    g = W * rhold - conD;
    dg = W - (rhold * conS2 * dhdrho) / (W * conD * conD * h * h * h);
    *iters = count + 1;
    if ((!isfinite(dg)) || (dg == 0.0)) return C2P_NR;
    *rho = dmax(rhomin, rhold - g / dg);
    error = fabs(1. - rhold / (*rho));
    rhold = *rho;
    if (error < P.errormax) break;
  }
  if (count >= P.countmax) return C2P_NR;
  eos_cold(P, rhold, pres, epsl, &chi);
  h = 1.0 + *epsl + *pres / rhold;
  if (conD * h == 0.) arg = 1.0 + conS2 * BIG;
  else arg = 1.0 + conS2 / ((conD * h) * (conD * h));
  W = sqrt(arg);
  if (*rho < P.fatm * P.rhoatm) return C2P_ATM;
  *conT = W * W * (*rho) * h - *pres - conD;
  const double v2 = conS2 / ((*conT + conD + *pres) * (*conT + conD + *pres));
  if (v2 > P.vmax) return C2P_ATM;
  return C2P_OK;
}

__device__ __forceinline__ double invg(double g11, double g12, double g13, double g22, double g23, double g33,
                                       double *i11, double *i12, double *i13, double *i22, double *i23,
                                       double *i33) {
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

// SoA views: cons (D, T, Sx, Sy, Sz), metric (xx xy xz yy yz zz),
// prim (rho, eps, vx, vy, vz, p, v2, detg)
struct C2PArrays {
  double *cons[5];
  const double *g[6];
  double *prim[8];
  int *status;
  int *iters;
};

// The hybrid p-root with cold static atmosphere (grhd_c2p_v2.c:303-460) in
// two halves around its hot solve, shared by the reference's hybrid (hot
// solve started at the previous pressure) and the learned variant (hot
// solve started at the network's pressure).
struct HybPoint {
  double gxx, gxy, gxz, gyy, gyz, gzz;
  double detg, SQRTdetg;
  double conD, conT, conSx, conSy, conSz, conSix, conSiy, conSiz, conS2;
  int result;  // C2P_OK, or C2P_FAIL for unusable input
};

// metric inverse, undensitised conserved variables and the input checks
// (grhd_c2p_v2.c:303-370)
__device__ __forceinline__ void hybrid_prepare(const C2PArrays &A, long q, HybPoint &h) {
  h.gxx = A.g[0][q], h.gxy = A.g[1][q], h.gxz = A.g[2][q], h.gyy = A.g[3][q], h.gyz = A.g[4][q], h.gzz = A.g[5][q];
  double gixx, gixy, gixz, giyy, giyz, gizz;
  h.detg = invg(h.gxx, h.gxy, h.gxz, h.gyy, h.gyz, h.gzz, &gixx, &gixy, &gixz, &giyy, &giyz, &gizz);
  h.SQRTdetg = sqrt(fabs(h.detg));
  if (!isfinite(h.SQRTdetg)) {
    h.detg = 1.;
    h.SQRTdetg = 1.;
  }
  const double oo = 1. / h.SQRTdetg;
  h.conD = A.cons[0][q] * oo, h.conT = A.cons[1][q] * oo;
  h.conSx = A.cons[2][q] * oo, h.conSy = A.cons[3][q] * oo, h.conSz = A.cons[4][q] * oo;
  h.conSix = gixx * h.conSx + gixy * h.conSy + gixz * h.conSz;
  h.conSiy = gixy * h.conSx + giyy * h.conSy + giyz * h.conSz;
  h.conSiz = gixz * h.conSx + giyz * h.conSy + gizz * h.conSz;
  h.conS2 = h.conSx * h.conSix + h.conSy * h.conSiy + h.conSz * h.conSiz;
  const double SQRTconS2 = sqrt(fabs(h.conS2));
  h.result = C2P_OK;
  if (!isfinite(h.conD) || !isfinite(h.conSx) || !isfinite(h.conSy) || !isfinite(h.conSz) || !isfinite(h.conT))
    h.result = C2P_FAIL;
  if (!isfinite(h.gxx) || !isfinite(h.gxy) || !isfinite(h.gxz) || !isfinite(h.gyy) || !isfinite(h.gyz) ||
      !isfinite(h.gzz) || !isfinite(h.detg) || !isfinite(gixx) || !isfinite(gixy) || !isfinite(gixz) ||
      !isfinite(giyy) || !isfinite(giyz) || !isfinite(gizz))
    h.result = C2P_FAIL;
  if (!isfinite(h.conSix) || !isfinite(h.conSiy) || !isfinite(h.conSiz) || !isfinite(h.conS2) ||
      !isfinite(SQRTconS2)) {
    h.conSx = h.conSix = 0.;
    h.conSy = h.conSiy = 0.;
    h.conSz = h.conSiz = 0.;
    h.conS2 = 0.;
    h.result = C2P_FAIL;
  }
}

// after the hot solve: the cold branch (rewrites tau), the atmosphere (any
// other failure; rewrites every conserved variable) and the primitive stores
// (grhd_c2p_v2.c:372-460)
__device__ __forceinline__ void hybrid_finish(const C2PArrays &A, const C2PK &P, long q, HybPoint &h, int result,
                                              double rho, double epsl, double pres, int it) {
  int it2 = 0;
  if (result == C2P_COLD) {
    result = try_cold(P, h.conD, h.conS2, &h.conT, &rho, &epsl, &pres, A.prim[0][q], &it2);
    A.cons[1][q] = h.SQRTdetg * h.conT;
  }
  const int status = result;
  if (result != C2P_OK && result != C2P_ATM) result = C2P_ATM;
  if (result == C2P_ATM) {
    pres = P.patm;
    rho = P.rhoatm;
    epsl = P.epslatm;
    h.conD = P.rhoatm;
    h.conSx = h.conSix = 0.;
    h.conSy = h.conSiy = 0.;
    h.conSz = h.conSiz = 0.;
    h.conS2 = 0.;
    h.conT = P.rhoatm * P.epslatm;
    A.cons[0][q] = h.SQRTdetg * h.conD;
    A.cons[2][q] = h.SQRTdetg * h.conSx;
    A.cons[3][q] = h.SQRTdetg * h.conSy;
    A.cons[4][q] = h.SQRTdetg * h.conSz;
    A.cons[1][q] = h.SQRTdetg * h.conT;
  }
  const double ooDpT = 1.0 / (h.conD + pres + h.conT);
  const double vx = h.conSix * ooDpT, vy = h.conSiy * ooDpT, vz = h.conSiz * ooDpT;
  A.prim[0][q] = rho;
  A.prim[1][q] = epsl;
  A.prim[2][q] = vx;
  A.prim[3][q] = vy;
  A.prim[4][q] = vz;
  A.prim[5][q] = pres;
  const double vlx = h.gxx * vx + h.gxy * vy + h.gxz * vz;
  const double vly = h.gxy * vx + h.gyy * vy + h.gyz * vz;
  const double vlz = h.gxz * vx + h.gyz * vy + h.gzz * vz;
  A.prim[6][q] = vlx * vx + vly * vy + vlz * vz;
  A.prim[7][q] = h.detg;
  A.status[q] = status;
  A.iters[q] = it + it2;
}

// hybrid p-root with cold static atmosphere (grhd_c2p_v2.c:265-464): the hot
// solve starts at the previous pressure
__global__ void c2p_hybrid_kernel(C2PArrays A, C2PK P, long n) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    HybPoint h;
    hybrid_prepare(A, q, h);
    double rho = 0, epsl = 0, pres = 0;
    int it = 0;
    int result = h.result;
    if (result != C2P_FAIL) result = try_hot(P, h.conD, h.conS2, h.conT, &rho, &epsl, &pres, A.prim[5][q], &it);
    hybrid_finish(A, P, q, h, result, rho, epsl, pres, it);
  }
}

// plain p-root (grhd_c2p.c:102-237); errorexit cases become FAIL / NR
__global__ void c2p_plain_kernel(C2PArrays A, C2PK P, long n) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const double gxx = A.g[0][q], gxy = A.g[1][q], gxz = A.g[2][q], gyy = A.g[3][q], gyz = A.g[4][q],
                 gzz = A.g[5][q];
    double gixx, gixy, gixz, giyy, giyz, gizz;
    const double detg = invg(gxx, gxy, gxz, gyy, gyz, gzz, &gixx, &gixy, &gixz, &giyy, &giyz, &gizz);
    A.prim[7][q] = detg;
    int status = C2P_OK, iters = 0;
    if (detg <= 0. || !isfinite(detg)) {
      A.status[q] = C2P_FAIL;
      A.iters[q] = 0;
      continue;
    }
    const double oo = 1. / sqrt(detg);
    const double conD = A.cons[0][q] * oo, conT = A.cons[1][q] * oo;
    const double conSx = A.cons[2][q] * oo, conSy = A.cons[3][q] * oo, conSz = A.cons[4][q] * oo;
    const double conSix = gixx * conSx + gixy * conSy + gixz * conSz;
    const double conSiy = gixy * conSx + giyy * conSy + giyz * conSz;
    const double conSiz = gixz * conSx + giyz * conSy + gizz * conSz;
    const double conS2 = conSx * conSix + conSy * conSiy + conSz * conSiz;
    const double SQRTconS2 = sqrt(conS2);
    if (!isfinite(conD) || !isfinite(conSx) || !isfinite(conSy) || !isfinite(conSz) || !isfinite(conT) ||
        !isfinite(SQRTconS2)) {
      A.status[q] = C2P_FAIL;
      A.iters[q] = 0;
      continue;
    }
    const double pmin = SQRTconS2 - conD - conT;
    double pold = dmax(A.prim[5][q], pmin);
    double temp = conT + pold + conD, arg = temp * temp - conS2, sq = sqrt(arg), W = temp / sq;
    double rho = conD / W, eps = (sq - W * pold) / conD - 1.;
    if (!isfinite(pold) || arg < 0.0) {
      A.status[q] = C2P_FAIL;
      A.iters[q] = 0;
      continue;
    }
    double pnew = pold;
    int count;
    for (count = 0; count <= P.countmax; count++) {
      double peos, chi, kappa;
      eos_re(P, rho, eps, &peos, &chi, &kappa);
      const double f = pold - peos;
      const double drhodp = conD * conS2 / (sq * temp * temp);
      const double depsldp = pold * conS2 / (conD * arg * sq);
      const double df = 1. - chi * drhodp - kappa * depsldp;
      if (df == 0.0) {
        status = C2P_NR;
        break;
      }
      pnew = dmax(pmin, pold - f / df);
      const double error = fabs(1. - pold / pnew);
      pold = pnew;
      temp = conT + pold + conD;
      arg = temp * temp - conS2;
      sq = sqrt(arg);
      W = temp / sq;
      rho = conD / W;
      eps = (sq - W * pold) / conD - 1.;
      iters = count + 1;
      if (error < P.errormax) break;
    }
    if (status == C2P_OK && count >= P.countmax) status = C2P_NR;
    if (status == C2P_OK && eps < 0.) status = C2P_FAIL;
    // grhd_c2p.c:224-226 divides each component
    const double vx = conSix / (conD + pnew + conT), vy = conSiy / (conD + pnew + conT),
                 vz = conSiz / (conD + pnew + conT);
    A.prim[0][q] = rho;
    A.prim[1][q] = eps;
    A.prim[5][q] = pnew;
    A.prim[2][q] = vx;
    A.prim[3][q] = vy;
    A.prim[4][q] = vz;
    const double vlx = gxx * vx + gxy * vy + gxz * vz;
    const double vly = gxy * vx + gyy * vy + gyz * vz;
    const double vlz = gxz * vx + gyz * vy + gzz * vz;
    A.prim[6][q] = vlx * vx + vly * vy + vlz * vz;
    A.status[q] = status;
    A.iters[q] = iters;
  }
}

// Bracketed Newton-Raphson on the pressure: the safeguarded solver of this
// build (no reference counterpart).  f(p) and f'(p) are those of
// try_hot_part; the bracket [plo, phi] keeps the sign change (plo =
// max(|S|-D-tau, 0), phi = (Gamma-1)(tau+D) since rho eps <= tau + D), a
// Newton step leaving it is replaced by bisection, and the iteration stops
// when |f| is within its rounding floor (sqrt(temp^2 - S^2) has a relative
// error ~ eps_mach W^2), or the relative update / bracket is below
// errormax, or after maxit iterations (NR).  guess <= 0: start at plo.
__device__ int c2p_bracketed_pt(const C2PK &P, double conD, double conS2, double conT, double guess, double *rho,
                                double *eps, double *pres, int *iters, int maxit) {
  double plo = sqrt(conS2) - conD - conT;
  if (plo < 0.) plo = 0.;
  double phi = (P.gamma - 1.) * (conT + conD);
  if (!(phi > plo)) phi = plo + 1e-300;
  double p = (guess > plo && guess < phi) ? guess : plo;
  const double em = 2.220446049250313e-16;
  int count;
  *iters = 0;
  for (count = 0; count < maxit; count++) {
    const double temp = conT + p + conD;
    const double arg = temp * temp - conS2;
    const double sq = sqrt(arg);
    const double W = temp / sq;
    const double r = conD / W;
    const double e = (sq - W * p) / conD - 1.;
    double peos, chi, kappa;
    eos_re(P, r, e, &peos, &chi, &kappa);
    const double f = p - peos;
    const double noise =
        4. * em * (fabs(p) + fabs(peos) + fabs(chi) * r + fabs(kappa) * (W * W * sq + W * fabs(p)) / conD);
    *iters = count + 1;
    if (fabs(f) <= noise) break;
    if (f < 0.) plo = p;
    else phi = p;
    const double drhodp = conD * conS2 / (sq * temp * temp);
    const double depsldp = p * conS2 / (conD * arg * sq);
    const double df = 1. - chi * drhodp - kappa * depsldp;
    double pn = p - f / df;
    if (!(pn >= plo && pn <= phi)) pn = 0.5 * (plo + phi);
    const double step = fabs(pn - p);
    p = pn;
    if (step <= P.errormax * p || phi - plo <= P.errormax * p) break;
  }
  const double temp = conT + p + conD;
  const double sq = sqrt(temp * temp - conS2);
  const double W = temp / sq;
  *rho = conD / W;
  *eps = (sq - W * p) / conD - 1.;
  *pres = p;
  if (count >= maxit) return C2P_NR;
  if (!isfinite(p) || !isfinite(*rho) || !isfinite(*eps)) return C2P_FAIL;
  return C2P_OK;
}

// undensitise and raise S with the inverse metric: shared by the bracketed
// and learned solvers
struct Undens {
  double D, T, S2, Six, Siy, Siz, detg;
};
__device__ __forceinline__ Undens undensitise(const C2PArrays &A, long q) {
  Undens u;
  double gixx, gixy, gixz, giyy, giyz, gizz;
  u.detg = invg(A.g[0][q], A.g[1][q], A.g[2][q], A.g[3][q], A.g[4][q], A.g[5][q], &gixx, &gixy, &gixz, &giyy,
                &giyz, &gizz);
  const double oo = 1. / sqrt(u.detg);
  u.D = A.cons[0][q] * oo;
  u.T = A.cons[1][q] * oo;
  const double Sx = A.cons[2][q] * oo, Sy = A.cons[3][q] * oo, Sz = A.cons[4][q] * oo;
  u.Six = gixx * Sx + gixy * Sy + gixz * Sz;
  u.Siy = gixy * Sx + giyy * Sy + giyz * Sz;
  u.Siz = gixz * Sx + giyz * Sy + gizz * Sz;
  u.S2 = Sx * u.Six + Sy * u.Siy + Sz * u.Siz;
  return u;
}

__device__ __forceinline__ void store_prims(const C2PArrays &A, long q, const Undens &u, double rho, double eps,
                                            double p) {
  const double ooDpT = 1.0 / (u.D + p + u.T);
  const double vx = u.Six * ooDpT, vy = u.Siy * ooDpT, vz = u.Siz * ooDpT;
  A.prim[0][q] = rho;
  A.prim[1][q] = eps;
  A.prim[2][q] = vx;
  A.prim[3][q] = vy;
  A.prim[4][q] = vz;
  A.prim[5][q] = p;
  const double vlx = A.g[0][q] * vx + A.g[1][q] * vy + A.g[2][q] * vz;
  const double vly = A.g[1][q] * vx + A.g[3][q] * vy + A.g[4][q] * vz;
  const double vlz = A.g[2][q] * vx + A.g[4][q] * vy + A.g[5][q] * vz;
  A.prim[6][q] = vlx * vx + vly * vy + vlz * vz;
  A.prim[7][q] = u.detg;
}

__global__ void c2p_bracketed_kernel(C2PArrays A, C2PK P, long n) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const Undens u = undensitise(A, q);
    double rho, eps, p;
    int it;
    const int st = c2p_bracketed_pt(P, u.D, u.S2, u.T, A.prim[5][q], &rho, &eps, &p, &it, P.countmax);
    store_prims(A, q, u, rho, eps, p);
    A.status[q] = st;
    A.iters[q] = it;
  }
}

// MATTER.c2p with the bracketed solver: the hybrid of grhd_c2p_v2.c:265-464
// whose hot solve is the bracketed Newton iteration from the previous
// pressure.  A state with D < fatm rhoatm (atmosphere whatever the
// pressure) and a state whose bracketed solve is not OK, or lands below the
// atmosphere threshold, at eps < 0 or beyond vmax, take the reference's hot
// solve instead, so the cold branch, the atmosphere and the conserved
// rewrites are the reference's.
__global__ void c2p_bracketed_hybrid_kernel(C2PArrays A, C2PK P, long n) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    HybPoint h;
    hybrid_prepare(A, q, h);
    double rho = 0, epsl = 0, pres = 0;
    int it = 0;
    int result = h.result;
    if (result != C2P_FAIL) {
      bool done = false;
      if (h.conD >= P.fatm * P.rhoatm) {
        result = c2p_bracketed_pt(P, h.conD, h.conS2, h.conT, A.prim[5][q], &rho, &epsl, &pres, &it, P.countmax);
        const double s = h.conT + h.conD + pres;
        done = result == C2P_OK && rho >= P.fatm * P.rhoatm && epsl >= 0. && h.conS2 / (s * s) <= P.vmax;
      }
      if (!done) {
        int it2 = 0;
        result = try_hot(P, h.conD, h.conS2, h.conT, &rho, &epsl, &pres, A.prim[5][q], &it2);
        it += it2;
      }
    }
    hybrid_finish(A, P, q, h, result, rho, epsl, pres, it);
  }
}

// Learned guess + bracketed polish: each wavefront evaluates the network
// for its 64 points on the matrix cores (mlp_device.h), then every lane
// polishes its point with the bracketed Newton solver started at the
// predicted pressure, limited to polish_max iterations.  A point whose
// features are not finite or whose polish does not converge falls back to
// the cold-started bracketed solve (counted in *fallbacks).  The loop trip
// count is uniform over the block, so every wave reaches the barriers.
__global__ void __launch_bounds__(64) c2p_learned_kernel(C2PArrays A, C2PK P, MlpView M, long n, int polish_max,
                                                         int *fallbacks) {
  // one wavefront per block: the barriers only order this wave's LDS traffic
  __shared__ __attribute__((aligned(16))) __bf16 h1[64 * kHStr];
  __shared__ __attribute__((aligned(16))) float b2[kMlpH], w3[kMlpH], w1b1[4 * kMlpH];
  const int l = threadIdx.x;
  b2[l] = M.p[kOffB2 + l];
  w3[l] = M.wq[192 + l];
  for (int j = 0; j < 3; j++) w1b1[3 * l + j] = M.wq[3 * l + j];
  w1b1[192 + l] = M.p[kOffB1 + l];
  __syncthreads();
  int nfb = 0;
  for (long base = (long)blockIdx.x * 64; base < n; base += (long)gridDim.x * 64) {
    const long q = base + l;
    const bool act = q < n;
    Undens u = {};
    double th = 0.;
    float x[3] = {0.f, 0.f, 0.f};
    bool valid = false;
    if (act) {
      u = undensitise(A, q);
      th = mlp_thermal(u.D, u.S2, u.T);
      valid = mlp_features(M, u.D, u.S2, u.T, th, x);
      if (!valid) x[0] = x[1] = x[2] = 0.f;
    }
    mlp_layer1_from(w1b1, x, h1 + l * kHStr);
    __syncthreads();
    const float yn = mlp_infer23(M, h1, b2, w3);
    __syncthreads();
    if (act) {
      double rho = 0., eps = 0., p = 0.;
      int it = 0, it2 = 0, st = C2P_NR;
      if (valid) {
        const double guess = (P.gamma - 1.) * th * (double)__expf(M.ys + M.yk * yn);
        st = c2p_bracketed_pt(P, u.D, u.S2, u.T, guess, &rho, &eps, &p, &it, polish_max);
      }
      if (st != C2P_OK) {
        nfb++;
        st = c2p_bracketed_pt(P, u.D, u.S2, u.T, 0., &rho, &eps, &p, &it2, P.countmax);
      }
      store_prims(A, q, u, rho, eps, p);
      A.status[q] = st;
      A.iters[q] = it + it2;
    }
  }
  for (int o = 32; o > 0; o >>= 1) nfb += __shfl_xor(nfb, o);
  if (fallbacks && l == 0 && nfb) atomicAdd(fallbacks, nfb);
}

// The hybrid of grhd_c2p_v2.c:265-464 with the learned guess as the hot
// solve's initial pressure (MATTER.c2p with the network): each wavefront
// evaluates the network for its 64 points on the matrix cores, then every
// lane runs the reference's hot solve (try_hot, at most polish_max
// iterations) from the predicted pressure.  If that does not return OK the
// point falls back to the reference's hybrid from the previous pressure
// (counted in *fallbacks), so the cold branch and the atmosphere follow the
// reference's rules.  A state with D < fatm rhoatm (rho = D / W is then
// below the atmosphere threshold whatever the pressure) and a state with no
// thermal content skip the network.  The loop trip count is uniform over
// the block, so every wave reaches the barriers.
__global__ void __launch_bounds__(64) c2p_learned_hybrid_kernel(C2PArrays A, C2PK P, MlpView M, long n,
                                                                int polish_max, int *fallbacks) {
  __shared__ __attribute__((aligned(16))) __bf16 h1[64 * kHStr];
  __shared__ __attribute__((aligned(16))) float b2[kMlpH], w3[kMlpH], w1b1[4 * kMlpH];
  const int l = threadIdx.x;
  b2[l] = M.p[kOffB2 + l];
  w3[l] = M.wq[192 + l];
  for (int j = 0; j < 3; j++) w1b1[3 * l + j] = M.wq[3 * l + j];
  w1b1[192 + l] = M.p[kOffB1 + l];
  __syncthreads();
  C2PK Pl = P;
  Pl.countmax = polish_max;
  int nfb = 0;
  for (long base = (long)blockIdx.x * 64; base < n; base += (long)gridDim.x * 64) {
    const long q = base + l;
    const bool act = q < n;
    HybPoint h;
    double th = 0.;
    float x[3] = {0.f, 0.f, 0.f};
    bool valid = false;
    if (act) {
      hybrid_prepare(A, q, h);
      if (h.result != C2P_FAIL && h.conD >= P.fatm * P.rhoatm) {
        th = mlp_thermal(h.conD, h.conS2, h.conT);
        valid = mlp_features(M, h.conD, h.conS2, h.conT, th, x);
      }
      if (!valid) x[0] = x[1] = x[2] = 0.f;
    }
    mlp_layer1_from(w1b1, x, h1 + l * kHStr);
    __syncthreads();
    const float yn = mlp_infer23(M, h1, b2, w3);
    __syncthreads();
    if (act) {
      double rho = 0., epsl = 0., pres = 0.;
      int it = 0, result = h.result;
      if (result != C2P_FAIL) {
        if (valid) {
          const double guess = (P.gamma - 1.) * th * (double)__expf(M.ys + M.yk * yn);
          result = try_hot(Pl, h.conD, h.conS2, h.conT, &rho, &epsl, &pres, guess, &it);
        }
        if (!valid || result != C2P_OK) {
          int it2 = 0;
          if (valid) nfb++;
          result = try_hot(P, h.conD, h.conS2, h.conT, &rho, &epsl, &pres, A.prim[5][q], &it2);
          it += it2;
        }
      }
      hybrid_finish(A, P, q, h, result, rho, epsl, pres, it);
    }
  }
  for (int o = 32; o > 0; o >>= 1) nfb += __shfl_xor(nfb, o);
  if (fallbacks && l == 0 && nfb) atomicAdd(fallbacks, nfb);
}

// training data of the learned guess from solved points: raw features
// (log D, |S|/(tau+D), log(Delta/(tau+D))) and target
// log(p / ((Gamma-1) Delta)); the target is NaN where the solve did not
// succeed or the state has no thermal content
__global__ void mlp_prepare_kernel(C2PArrays A, C2PK P, long n, float *__restrict__ feat, float *__restrict__ target) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const Undens u = undensitise(A, q);
    const double th = mlp_thermal(u.D, u.S2, u.T);
    float r[3];
    const bool fok = mlp_raw_features(u.D, u.S2, u.T, th, r);
    for (int j = 0; j < 3; j++) feat[3 * q + j] = r[j];
    const bool ok = A.status[q] == C2P_OK && fok;
    target[q] = ok ? (float)log(A.prim[5][q] / ((P.gamma - 1.) * th)) : __builtin_nanf("");
  }
}

// primitives -> conserved variables: grhd_compute_v2_pt, EOS p(rho, eps),
// invg and grhd_compute_q_pt (grhd_p2c.c:11-25,68-86); prim5 = rho, eps, v^i
__global__ void p2c_kernel(C2PK P, long n, const double *__restrict__ prim5, const double *__restrict__ metric,
                           double *__restrict__ cons) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
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
    const double detg = invg(gxx, gxy, gxz, gyy, gyz, gzz, &i11, &i12, &i13, &i22, &i23, &i33);
    const double sq = sqrt(detg);
    const double W2hrho = W * W * (rho + rho * eps + p);
    cons[q] = sq * W * rho;
    cons[n + q] = sq * (W2hrho - p - W * rho);
    cons[2 * n + q] = sq * W2hrho * vlx;
    cons[3 * n + q] = sq * W2hrho * vly;
    cons[4 * n + q] = sq * W2hrho * vlz;
  }
}

// Cold static atmosphere on the evolved conserved variables, grhd_atm_set_u
// (src/projects/grhd/grhd_atm.c:217-262) as called by grhd_atm_set_prerhs
// (:268-278) on u_p: where rho < f_atm rho_atm, D = sqrt(detg) rho_atm,
// S_i = 0, tau = sqrt(detg) (rho_atm eps_atm), with detg raised to DETGMIN
// (grhd.h: 1e-5) in the primitives and sqrt(detg) = 1 if not finite.  A
// point still non-finite afterwards is zeroed (detg = 1) and raises *flag
// (the reference's errorexit "nans not cured by setting atmosphere").
__global__ void atm_set_u_kernel(C2PArrays A, C2PK P, long n, int *flag) {
  const double DETGMIN = 1e-5;
  const double rhoatmlevel = P.rhoatm * P.fatm;
  const double Datmlevel = P.rhoatm;
  const double Tatmlevel = P.rhoatm * P.epslatm;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    double D = A.cons[0][q], T = A.cons[1][q], Sx = A.cons[2][q], Sy = A.cons[3][q], Sz = A.cons[4][q];
    double detg = A.prim[7][q];
    if (A.prim[0][q] < rhoatmlevel) {
      detg = detg > DETGMIN ? detg : DETGMIN;
      double s = sqrt(detg);
      if (!isfinite(s)) s = 1.;
      D = s * Datmlevel;
      Sx = Sy = Sz = 0.;
      T = s * Tatmlevel;
      A.cons[0][q] = D;
      A.cons[1][q] = T;
      A.cons[2][q] = Sx;
      A.cons[3][q] = Sy;
      A.cons[4][q] = Sz;
      A.prim[7][q] = detg;
    }
    if (!isfinite(D) || !isfinite(Sx) || !isfinite(Sy) || !isfinite(Sz) || !isfinite(T) || !isfinite(detg)) {
      for (int k = 0; k < 5; k++) A.cons[k][q] = 0.;
      A.prim[7][q] = 1.;
      *flag = 1;
    }
  }
}

// analytic Gamma-law recovery (grhd_c2p_vanal, grhd_c2p.c:239-379): |v| is
// the physical root of v^4 + a3 v^3 + a2 v^2 + a1 v + a0, taken through the
// largest root of the resolvent cubic and the quadratic factor holding it;
// the reference's operation order throughout (fmax / fmin clamp, so a NaN
// root becomes 0).  Stops of the reference become status FAIL.
__global__ void c2p_vanal_kernel(C2PArrays A, C2PK P, long n) {
  const double gammamo = P.gamma - 1.;
  const double oo3 = 1. / 3.;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    double gixx, gixy, gixz, giyy, giyz, gizz;
    const double detg = invg(A.g[0][q], A.g[1][q], A.g[2][q], A.g[3][q], A.g[4][q], A.g[5][q], &gixx, &gixy,
                             &gixz, &giyy, &giyz, &gizz);
    A.prim[7][q] = detg;
    A.iters[q] = 0;
    if (detg <= 0. || !isfinite(detg)) {
      A.status[q] = C2P_FAIL;
      continue;
    }
    const double oo = 1. / sqrt(detg);
    const double conD = A.cons[0][q] * oo, conT = A.cons[1][q] * oo;
    const double conSx = A.cons[2][q] * oo, conSy = A.cons[3][q] * oo, conSz = A.cons[4][q] * oo;
    const double conSix = gixx * conSx + gixy * conSy + gixz * conSz;
    const double conSiy = gixy * conSx + giyy * conSy + giyz * conSz;
    const double conSiz = gixz * conSx + giyz * conSy + gizz * conSz;
    const double conS2 = conSx * conSix + conSy * conSiy + conSz * conSiz;
    const double S = sqrt(conS2);
    if (!isfinite(conD) || !isfinite(conSx) || !isfinite(conSy) || !isfinite(conSz) || !isfinite(conT) ||
        !isfinite(S)) {
      A.status[q] = C2P_FAIL;
      continue;
    }
    const double D = conD, E = conT + conD;
    double v = 0.0;
    if (fabs(S) > 1e-20) {
      const double SE = S * E, D2 = D * D;
      const double G = gammamo + 1, Gmosq = gammamo * gammamo;
      const double div = 1.0 / (Gmosq * (conS2 + D2));
      const double a3 = (-2.0 * G * gammamo * SE) * div;
      const double a2 = (G * G * E * E + 2.0 * gammamo * conS2 - Gmosq * D2) * div;
      const double a1 = (-2.0 * G * SE) * div;
      const double a0 = conS2 * div;
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
      A.status[q] = C2P_FAIL;
      continue;
    }
    div = 1. / (conD + p + conT);
    A.prim[0][q] = rho;
    A.prim[1][q] = epsl;
    A.prim[5][q] = p;
    A.prim[2][q] = conSix * div;
    A.prim[3][q] = conSiy * div;
    A.prim[4][q] = conSiz * div;
    A.status[q] = C2P_OK;
  }
}

static C2PK make_c2pk(const bamhip_c2p_params *p) {
  C2PK K;
  K.eos = p->eos;
  K.gamma = p->gamma;
  K.K = p->K;
  K.rhoatm = p->rhoatm;
  K.fatm = p->fatm;
  K.epslatm = p->epslatm;
  K.patm = p->patm;
  K.vmax = p->vmax;
  K.errormax = p->errormax;
  K.countmax = p->countmax;
  return K;
}

// level: MATTER.c2p on a level, where the bracketed method takes the
// hybrid's control flow (c2p_bracketed_hybrid_kernel); flat arrays run the
// bracketed solver alone
int c2p_launch(const bamhip_c2p_params *par, int method, long n, C2PArrays &A, bool level) {
  if (!par || n < 0) return set_error(BAMHIP_EINVAL, "bamhip_c2p: bad arguments");
  if (par->eos != 0 && par->eos != 1) return set_error(BAMHIP_EINVAL, "bamhip_c2p: eos=%d (0 ideal, 1 poly)", par->eos);
  if (!(par->gamma > 1.0)) return set_error(BAMHIP_EINVAL, "bamhip_c2p: gamma must be > 1");
  if (par->countmax < 1) return set_error(BAMHIP_EINVAL, "bamhip_c2p: countmax < 1");
  if (n == 0) return BAMHIP_OK;
  const C2PK K = make_c2pk(par);
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  TimerScope t("c2p");
  if (method == BAMHIP_C2P_HYBRID)
    c2p_hybrid_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(A, K, n);
  else if (method == BAMHIP_C2P_PLAIN)
    c2p_plain_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(A, K, n);
  else if (method == BAMHIP_C2P_BRACKETED && level)
    c2p_bracketed_hybrid_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(A, K, n);
  else if (method == BAMHIP_C2P_BRACKETED)
    c2p_bracketed_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(A, K, n);
  else if (method == BAMHIP_C2P_VANAL) {
    if (par->eos != 0) return set_error(BAMHIP_EINVAL, "bamhip_c2p: the analytic recovery needs the ideal gas");
    c2p_vanal_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(A, K, n);
  }
  else
    return set_error(BAMHIP_EINVAL, "bamhip_c2p: unknown method %d", method);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_c2p_batch(const bamhip_c2p_params *par, int method, long n, double *cons, const double *metric,
                                double *prim, int *status, int *iters) {
  if (!cons || !metric || !prim || !status || !iters) return set_error(BAMHIP_EINVAL, "bamhip_c2p_batch: NULL array");
  C2PArrays A;
  for (int k = 0; k < 5; k++) A.cons[k] = cons + k * n;
  for (int k = 0; k < 6; k++) A.g[k] = metric + k * n;
  for (int k = 0; k < 8; k++) A.prim[k] = prim + k * n;
  A.status = status;
  A.iters = iters;
  return c2p_launch(par, method, n, A, false);
}

static void c2p_flat(C2PArrays &A, long n, double *cons, const double *metric, double *prim, int *status,
                     int *iters) {
  for (int k = 0; k < 5; k++) A.cons[k] = cons + k * n;
  for (int k = 0; k < 6; k++) A.g[k] = metric + k * n;
  for (int k = 0; k < 8; k++) A.prim[k] = prim + k * n;
  A.status = status;
  A.iters = iters;
}

// hybrid: the reference hybrid's control flow (MATTER.c2p on a level);
// otherwise the bracketed polish and fallback of the flat-array solver
static int learned_launch(const bamhip_c2p_params *par, const bamhip_mlp *m, int polish_max, long n, C2PArrays &A,
                          int *fallbacks, bool hybrid) {
  if (!par || !m || n < 0 || polish_max < 1) return set_error(BAMHIP_EINVAL, "bamhip_c2p_learned: bad arguments");
  if (par->eos != 0) return set_error(BAMHIP_EINVAL, "bamhip_c2p_learned: the learned guess is trained for the ideal gas");
  if (par->countmax < 1) return set_error(BAMHIP_EINVAL, "bamhip_c2p_learned: countmax < 1");
  if (n == 0) return BAMHIP_OK;
  long blocks = (n + 63) / 64;
  if (blocks > 32768) blocks = 32768;
  TimerScope t("c2p");
  if (hybrid)
    c2p_learned_hybrid_kernel<<<(unsigned)blocks, 64, 0, stream()>>>(A, make_c2pk(par), mlp_view(m), n, polish_max,
                                                                     fallbacks);
  else
    c2p_learned_kernel<<<(unsigned)blocks, 64, 0, stream()>>>(A, make_c2pk(par), mlp_view(m), n, polish_max,
                                                              fallbacks);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

extern "C" int bamhip_c2p_learned_batch(const bamhip_c2p_params *par, const bamhip_mlp *m, int polish_max, long n,
                                        double *cons, const double *metric, double *prim, int *status, int *iters,
                                        int *fallbacks) {
  if (!cons || !metric || !prim || !status || !iters)
    return set_error(BAMHIP_EINVAL, "bamhip_c2p_learned_batch: NULL array");
  C2PArrays A;
  c2p_flat(A, n, cons, metric, prim, status, iters);
  return learned_launch(par, m, polish_max, n, A, fallbacks, false);
}

extern "C" int bamhip_grhd_c2p_learned(bamhip_level *lev, const bamhip_c2p_params *par, const bamhip_mlp *m,
                                       int polish_max, const int *cons5, const int *metric6, const int *prim8,
                                       int *status, int *iters, int *fallbacks) {
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_grhd_c2p_learned: NULL level");
  if (!status || !iters) return set_error(BAMHIP_EINVAL, "bamhip_grhd_c2p_learned: status/iters buffers are required");
  C2PArrays A;
  double *tmp[8];
  int e;
  if ((e = gather_ptrs(lev, cons5, 5, tmp, "bamhip_grhd_c2p_learned(cons)"))) return e;
  for (int k = 0; k < 5; k++) A.cons[k] = tmp[k];
  if ((e = gather_ptrs(lev, metric6, 6, tmp, "bamhip_grhd_c2p_learned(metric)"))) return e;
  for (int k = 0; k < 6; k++) A.g[k] = tmp[k];
  if ((e = gather_ptrs(lev, prim8, 8, tmp, "bamhip_grhd_c2p_learned(prim)"))) return e;
  for (int k = 0; k < 8; k++) A.prim[k] = tmp[k];
  A.status = status;
  A.iters = iters;
  return learned_launch(par, m, polish_max, lev->npoints, A, fallbacks, true);
}

extern "C" int bamhip_mlp_prepare(const bamhip_c2p_params *par, long n, const double *cons, const double *metric,
                                  const double *prim, const int *status, float *feat, float *target) {
  if (!par || n < 0 || !cons || !metric || !prim || !status || !feat || !target)
    return set_error(BAMHIP_EINVAL, "bamhip_mlp_prepare: bad arguments");
  if (n == 0) return BAMHIP_OK;
  C2PArrays A;
  c2p_flat(A, n, const_cast<double *>(cons), metric, const_cast<double *>(prim), const_cast<int *>(status), nullptr);
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  mlp_prepare_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(A, make_c2pk(par), n, feat, target);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

extern "C" int bamhip_p2c_batch(const bamhip_c2p_params *par, long n, const double *prim5, const double *metric,
                                double *cons) {
  if (!par || n < 0 || !prim5 || !metric || !cons) return set_error(BAMHIP_EINVAL, "bamhip_p2c_batch: bad arguments");
  if (n == 0) return BAMHIP_OK;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  p2c_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(make_c2pk(par), n, prim5, metric, cons);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

__global__ void atm_flag_set(int *p, int v) { *p = v; }

extern "C" int bamhip_grhd_atm_set_prerhs(bamhip_level *lev, const bamhip_c2p_params *par, const int *upre5,
                                          const int *prim8) {
  if (!lev || !par || !upre5 || !prim8) return set_error(BAMHIP_EINVAL, "bamhip_grhd_atm_set_prerhs: NULL argument");
  C2PArrays A = {};
  double *tmp[8];
  int e;
  if ((e = gather_ptrs(lev, upre5, 5, tmp, "bamhip_grhd_atm_set_prerhs(upre)"))) return e;
  for (int k = 0; k < 5; k++) A.cons[k] = tmp[k];
  if ((e = gather_ptrs(lev, prim8, 8, tmp, "bamhip_grhd_atm_set_prerhs(prim)"))) return e;
  for (int k = 0; k < 8; k++) A.prim[k] = tmp[k];
  const long n = lev->npoints;
  if (n == 0) return BAMHIP_OK;
  if (lev->check_mode) atm_flag_set<<<1, 1, 0, stream()>>>(lev->d_flag, 0);
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  {
    TimerScope t("c2p");
    atm_set_u_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(A, make_c2pk(par), n, lev->d_flag);
    BH_HIP(hipGetLastError());
  }
  if (lev->check_mode) {
    int h = 0;
    BH_HIP(hipMemcpyAsync(&h, lev->d_flag, sizeof(int), hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    if (h)
      return set_error(BAMHIP_ENONFINITE, "bamhip_grhd_atm_set_prerhs: nans not cured by setting atmosphere (level %d)",
                       lev->l);
  }
  return BAMHIP_OK;
}

extern "C" int bamhip_grhd_c2p(bamhip_level *lev, const bamhip_c2p_params *par, int method, const int *cons5,
                               const int *metric6, const int *prim8, int *status, int *iters) {
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_grhd_c2p: NULL level");
  if (!status || !iters) return set_error(BAMHIP_EINVAL, "bamhip_grhd_c2p: status/iters buffers are required");
  C2PArrays A;
  double *tmp[8];
  int e;
  if ((e = gather_ptrs(lev, cons5, 5, tmp, "bamhip_grhd_c2p(cons)"))) return e;
  for (int k = 0; k < 5; k++) A.cons[k] = tmp[k];
  if ((e = gather_ptrs(lev, metric6, 6, tmp, "bamhip_grhd_c2p(metric)"))) return e;
  for (int k = 0; k < 6; k++) A.g[k] = tmp[k];
  if ((e = gather_ptrs(lev, prim8, 8, tmp, "bamhip_grhd_c2p(prim)"))) return e;
  for (int k = 0; k < 8; k++) A.prim[k] = tmp[k];
  A.status = status;
  A.iters = iters;
  return c2p_launch(par, method, lev->npoints, A, true);
}
