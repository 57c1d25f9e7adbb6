// tov.hip -- a TOV star on the points of a level (set_TOV, src/projects/
// BBIDdataReader/BBIDdataReader.c:483-690, unboosted and unperturbed, with
// the enthalpy table of tov_h).
//
// One thread per point: the isotropic radius is located in the table by
// bisection (interp_locate, BBIDdataReader.c:248-277) and the Schwarzschild
// radius, the potential and the enthalpy are read by the 4-point Lagrange
// form of interp_lag4 (:280-318); the table (4 x npts doubles) is read from
// HBM through L2, a few cache lines per point.  Outside the surface the
// exterior Schwarzschild solution in isotropic coordinates, no matter.
#include <cmath>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

#pragma clang fp contract(off)

constexpr int kTovBlock = 256;

struct TovArgs {
  long np;
  int nb;
  DevBox box[16];
  int n;                    // table points
  const double *riso, *rsch, *phi, *h;
  double M, R, K, G, px, py, pz;
  double *g[6], *Kij[6], *alpha, *rho, *epsl, *vel[3], *pre;
};

__device__ int locate(const double *x, int n, double xv) {
  if (xv <= x[0]) return 0;
  if (xv >= x[n - 1]) return n - 1;
  const bool ascnd = x[n - 1] >= x[0];
  int jl = -1, ju = n;
  while (ju - jl > 1) {
    const int jm = (ju + jl) >> 1;
    if ((xv >= x[jm]) == ascnd) jl = jm;
    else ju = jm;
  }
  return jl;
}

__device__ double lag4(const double *f, const double *x, int n, int i, double xv) {
  const double ximo = x[i - 1], xi = x[i], xipo = x[i + 1], xipt = x[i + 2];
  const double C1 = (f[i] - f[i - 1]) / (xi - ximo);
  const double C2 = (-f[i] + f[i + 1]) / (-xi + xipo);
  const double C3 = (-f[i + 1] + f[i + 2]) / (-xipo + xipt);
  const double CC1 = (-C1 + C2) / (-ximo + xipo);
  const double CC2 = (-C2 + C3) / (-xi + xipt);
  const double CCC1 = (-CC1 + CC2) / (-ximo + xipt);
  return f[i - 1] + (-ximo + xv) * (C1 + (-xi + xv) * (CC1 + CCC1 * (-xipo + xv)));
}

__global__ void __launch_bounds__(kTovBlock) tov_kernel(TovArgs a) {
  const long ijk = (long)blockIdx.x * kTovBlock + threadIdx.x;
  if (ijk >= a.np) return;
  int b = 0;
  while (b + 1 < a.nb && ijk >= a.box[b + 1].noffset) b++;
  const DevBox &B = a.box[b];
  const long q = ijk - B.noffset;
  const int i = (int)(q % B.m), j = (int)((q / B.m) % B.n), k = (int)(q / ((long)B.m * B.n));
  const double x = B.x0 + i * B.dx - a.px, y = B.y0 + j * B.dy - a.py, z = B.z0 + k * B.dz - a.pz;
  const double r = sqrt(x * x + y * y + z * z);
  double alpha, psi4, rho, epsl, p;
  if (r < a.R) {
    int t = locate(a.riso, a.n, r);
    if (t < 1) t = 1;
    if (t > a.n - 3) t = a.n - 3;
    const double rsch = lag4(a.rsch, a.riso, a.n, t, r);
    const double phi = lag4(a.phi, a.riso, a.n, t, r);
    double h = lag4(a.h, a.riso, a.n, t, r);
    alpha = exp(phi);
    psi4 = pow(rsch / r, 2.);
    // eos_poly_H (src/projects/eos/eos_analytic.c:76-92), h < 0 as 0
    if (h < 0.) h = 0.;
    const double G1 = a.G - 1.;
    const double H1 = 2. * sinh(0.5 * h) * exp(0.5 * h);
    const double n = pow(G1 / (a.G * a.K) * H1, 1.0 / G1);
    const double krhotog1 = a.K * pow(n, G1);
    p = n * krhotog1;
    rho = n;
    epsl = krhotog1 / G1;
  } else {
    alpha = (1.0 - 0.5 * a.M / r) / (1.0 + 0.5 * a.M / r);
    psi4 = pow(1.0 + 0.5 * a.M / r, 4.);
    rho = 0.;
    epsl = 0.;
    p = 0.;
  }
  // adm metric psi4 delta_ij, no extrinsic curvature, matter at rest
  a.g[0][ijk] = psi4;
  a.g[1][ijk] = 0.;
  a.g[2][ijk] = 0.;
  a.g[3][ijk] = psi4;
  a.g[4][ijk] = 0.;
  a.g[5][ijk] = psi4;
  for (int c = 0; c < 6; c++) a.Kij[c][ijk] = 0.;
  a.alpha[ijk] = alpha;
  a.rho[ijk] = rho;
  a.epsl[ijk] = epsl;
  a.pre[ijk] = p;
  for (int c = 0; c < 3; c++) a.vel[c][ijk] = 0.;
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_tov_set(bamhip_level *lev, int n, const double *riso, const double *rsch, const double *phi,
                              const double *h, double M, double K, double Gamma, const double *center,
                              const int *adm_g, const int *adm_K, int alpha, const int *prim) {
  const char *what = "bamhip_tov_set";
  if (!lev || n < 4 || !riso || !rsch || !phi || !h || !center || !adm_g || !adm_K || !prim)
    return set_error(BAMHIP_EINVAL, "%s: bad arguments (table of %d points)", what, n);
  if (!(K > 0) || !(Gamma > 1)) return set_error(BAMHIP_EINVAL, "%s: polytrope K = %g, Gamma = %g", what, K, Gamma);
  if (lev->nboxes > 16) return set_error(BAMHIP_EINVAL, "%s: at most 16 boxes per level", what);
  TovArgs A{};
  int e;
  if ((e = gather_ptrs(lev, adm_g, 6, A.g, what)) || (e = gather_ptrs(lev, adm_K, 6, A.Kij, what))) return e;
  double *pp[6];
  if ((e = gather_ptrs(lev, prim, 6, pp, what))) return e;
  if ((e = gather_ptrs(lev, &alpha, 1, &A.alpha, what))) return e;
  A.rho = pp[0];
  A.epsl = pp[1];
  A.vel[0] = pp[2];
  A.vel[1] = pp[3];
  A.vel[2] = pp[4];
  A.pre = pp[5];
  A.np = lev->npoints;
  A.nb = lev->nboxes;
  for (int b = 0; b < lev->nboxes; b++) A.box[b] = lev->boxes[b];
  A.n = n;
  A.M = M;
  A.R = riso[n - 1];
  A.K = K;
  A.G = Gamma;
  A.px = center[0];
  A.py = center[1];
  A.pz = center[2];
  // the table on the device for the duration of the call
  double *d = nullptr;
  const size_t nb = sizeof(double) * (size_t)n;
  BH_HIP(hipMalloc(&d, 4 * nb));
  hipStream_t st = stream();
  const double *src[4] = {riso, rsch, phi, h};
  for (int c = 0; c < 4; c++)
    if (hipMemcpyAsync(d + (size_t)c * n, src[c], nb, hipMemcpyHostToDevice, st) != hipSuccess) {
      (void)hipFree(d);
      return set_error(BAMHIP_EINVAL, "%s: table copy failed", what);
    }
  A.riso = d;
  A.rsch = d + n;
  A.phi = d + 2 * (size_t)n;
  A.h = d + 3 * (size_t)n;
  if (A.np > 0) tov_kernel<<<(unsigned)((A.np + kTovBlock - 1) / kTovBlock), kTovBlock, 0, st>>>(A);
  hipError_t le = hipGetLastError();
  hipError_t se = hipStreamSynchronize(st);
  (void)hipFree(d);
  BH_HIP(le);
  BH_HIP(se);
  return BAMHIP_OK;
}
