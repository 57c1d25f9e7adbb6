// z4c_matter.hip -- the matter terms of the Z4c right-hand side on gfx950.
//
// Reference: z4_rhs_add_source (src/projects/z4/z4_rhs_add_source.m:14-77,
// generated z4_rhs_add_source.c), registered as "geometry_source"
// (z4.c:247) and called after the matter RHS in evolve_rhs
// (evolve.c:72-76).  At every inner point:
//   psi^4 = exp(4 log(max(chi, z4_chi_div_floor)) / z4_chi_psipower),
//   A_ij  += -8 pi alpha psi^-4 (S_ij - psi^4 g~_ij S / 3),
//   Khat  += 4 pi alpha (rho + S),
//   G^i   += -16 pi alpha psi^4 S^i,
//   Theta += -8 pi alpha rho,
//   B^i   += withB (gamma0factor + gamma2factor shiftgammacoeff
//            alpha^shiftalphapower) (the G^i term),
// all 0 when any of them (or the -16/3 pi alpha g~_ij rho that the
// reference also checks) is not finite; times c when c != 0.  One thread
// per point; the sums are those of the C restatement in the same order with
// FP contraction off (log, exp and pow round differently from glibc's, so
// the results agree to a few ulp).
#pragma clang fp contract(off)

#include <cmath>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr double kPi = 3.1415926535897932;  // bam_main.h:14
constexpr int kMBlock = 256;

struct MatterArgs {
  int m, n, o, skip[6];
  double c, chiDivFloor, oochipsipower, withB, gamma0factor, gamma2factor, shiftgammacoeff, shiftalphapower;
  const double *g[6], *chi, *alpha, *madm[11];
  double *rA[6], *rKhat, *rG[3], *rTheta, *rB[3];
};

__global__ void __launch_bounds__(kMBlock) z4c_matter_kernel(MatterArgs A) {
  const long npts = (long)A.m * A.n * A.o;
  const long q = (long)blockIdx.x * kMBlock + threadIdx.x;
  if (q >= npts) return;
  const long m = A.m, mn = (long)A.m * A.n;
  const int i = (int)(q % m), j = (int)((q / m) % A.n), k = (int)(q / mn);
  if (i < A.skip[0] || i >= A.m - A.skip[1] || j < A.skip[2] || j >= A.n - A.skip[3] || k < A.skip[4] ||
      k >= A.o - A.skip[5])
    return;
  const double al = A.alpha[q];
  double chig = A.chi[q];
  if (chig < A.chiDivFloor) chig = A.chiDivFloor;
  const double f = A.oochipsipower * log(chig);
  const double psi4 = exp(4. * f), psim4 = 1. / psi4;
  const double rho = A.madm[0][q], ST = A.madm[10][q];
  double rA[6], rG[3], rB[3];
  bool bad = false;
#pragma unroll
  for (int a = 0; a < 6; a++) {
    const double ga = A.g[a][q];
    const double sstf = A.madm[4 + a][q] - (1. / 3.) * (psi4 * ga) * ST;
    rA[a] = -psim4 * 8. * kPi * al * sstf;
    const double r2A = -(1. / 3.) * ga * 16. * kPi * al * rho;
    bad |= !isfinite(rA[a]) || !isfinite(r2A);
  }
  double rKhat = 4. * kPi * al * (rho + ST);
  double rTheta = al * (-16. * kPi * rho) / 2.;
  const double betaF = A.shiftgammacoeff * pow(al, A.shiftalphapower);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    rG[a] = 2. * al * (-8. * kPi * psi4 * A.madm[1 + a][q]);
    rB[a] = A.withB * (A.gamma0factor + A.gamma2factor * betaF) * rG[a];
    bad |= !isfinite(rG[a]) || !isfinite(rB[a]);
  }
  bad |= !isfinite(rKhat) || !isfinite(rTheta);
  if (bad) {
    rKhat = rTheta = 0.;
#pragma unroll
    for (int a = 0; a < 6; a++) rA[a] = 0.;
#pragma unroll
    for (int a = 0; a < 3; a++) rG[a] = rB[a] = 0.;
  }
  const double cc = A.c != 0. ? A.c : 1.;
#pragma unroll
  for (int a = 0; a < 6; a++) A.rA[a][q] += cc * rA[a];
  A.rKhat[q] += cc * rKhat;
#pragma unroll
  for (int a = 0; a < 3; a++) A.rG[a][q] += cc * rG[a];
  A.rTheta[q] += cc * rTheta;
#pragma unroll
  for (int a = 0; a < 3; a++) A.rB[a][q] += cc * rB[a];
}

// z4_to_adm_vl (z4.c:93-124) at every point: psi = chi^(1/chipsipower),
// g_ij = psi^4 g~_ij, K_ij = psi^4 A~_ij + Khat g_ij / 3 (the reference takes
// the stored bssn_K, which is Khat)
struct AdmArgs {
  const double *g[6], *A[6], *chi, *Kh;
  double *gadm[6], *kadm[6];
  double oopower;
  long n;
};

__global__ void __launch_bounds__(kMBlock) z4c_to_adm_kernel(AdmArgs A) {
  const long q = (long)blockIdx.x * kMBlock + threadIdx.x;
  if (q >= A.n) return;
  const double psi = pow(A.chi[q], A.oopower);
  const double psi4 = pow(psi, 4.);
  const double kh = A.Kh[q];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    const double gg = psi4 * A.g[c][q];
    A.gadm[c][q] = gg;
    A.kadm[c][q] = psi4 * A.A[c][q] + 1. / 3. * kh * gg;
  }
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_z4c_to_adm(bamhip_level *lev, const int *ucur, double chipsipower, const int *adm_g,
                                 const int *adm_K) {
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_z4c_to_adm: NULL level");
  double *u[25], *g[6], *k[6];
  int e;
  if ((e = gather_ptrs(lev, ucur, BAMHIP_Z4C_NVARS, u, "bamhip_z4c_to_adm(ucur)"))) return e;
  if ((e = gather_ptrs(lev, adm_g, 6, g, "bamhip_z4c_to_adm(adm_g)"))) return e;
  if ((e = gather_ptrs(lev, adm_K, 6, k, "bamhip_z4c_to_adm(adm_K)"))) return e;
  AdmArgs A{};
  for (int c = 0; c < 6; c++) A.g[c] = u[c], A.A[c] = u[7 + c], A.gadm[c] = g[c], A.kadm[c] = k[c];
  A.chi = u[6];
  A.Kh = u[13];
  A.oopower = 1. / chipsipower;
  A.n = lev->npoints;
  if (A.n > 0) z4c_to_adm_kernel<<<(unsigned)((A.n + kMBlock - 1) / kMBlock), kMBlock, 0, stream()>>>(A);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

extern "C" int bamhip_z4c_matter_source(bamhip_level *lev, const bamhip_z4c_params *par, const int *unew, double c,
                                        const int *ucur, const int *madm) {
  const char *what = "bamhip_z4c_matter_source";
  if (!lev || !par) return set_error(BAMHIP_EINVAL, "%s: NULL argument", what);
  double *u[25], *r[25], *ad[11];
  int e;
  if ((e = gather_ptrs(lev, ucur, BAMHIP_Z4C_NVARS, u, "bamhip_z4c_matter_source(ucur)"))) return e;
  if ((e = gather_ptrs(lev, unew, BAMHIP_Z4C_NVARS, r, "bamhip_z4c_matter_source(unew)"))) return e;
  if ((e = gather_ptrs(lev, madm, 11, ad, "bamhip_z4c_matter_source(madm)"))) return e;
  TimerScope ts("z4_matter_source");
  for (const DevBox &B : lev->boxes) {
    MatterArgs A{};
    A.m = B.m, A.n = B.n, A.o = B.o;
    for (int f = 0; f < 6; f++) A.skip[f] = B.skip[f];
    A.c = c;
    A.chiDivFloor = par->chiDivFloor;
    A.oochipsipower = 1.0 / par->chipsipower;
    A.withB = par->withB;
    A.gamma0factor = par->gamma0factor;
    A.gamma2factor = par->gamma2factor;
    A.shiftgammacoeff = par->shiftgammacoeff;
    A.shiftalphapower = par->shiftalphapower;
    const long o = B.noffset;
    for (int a = 0; a < 6; a++) A.g[a] = u[a] + o, A.rA[a] = r[7 + a] + o;
    A.chi = u[6] + o;
    A.alpha = u[18] + o;
    for (int a = 0; a < 11; a++) A.madm[a] = ad[a] + o;
    A.rKhat = r[13] + o;
    for (int a = 0; a < 3; a++) A.rG[a] = r[14 + a] + o, A.rB[a] = r[22 + a] + o;
    A.rTheta = r[17] + o;
    const long npts = (long)B.m * B.n * B.o;
    z4c_matter_kernel<<<(unsigned)((npts + kMBlock - 1) / kMBlock), kMBlock, 0, stream()>>>(A);
  }
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}
