// grhd_flux.hip -- the GRHD flux divergence on gfx950.
//
// Reference: matter_fluxes (src/projects/matter/matter_fluxes.c:20-86) with
// compute_flux1d (:492-647) and add_flx1d_rhs (:251-465), and the 1-D flux
// grhd_flx1d_fvpr (src/projects/grhd/grhd_flx1d.c:25-250) selected by
// hrsc_flux LLF / HLL (grhd.c:205-209): primitives reconstructed with WENOZ
// or WENO5 (hrsc_rec1d.c:283-363, 502-581), the metric with LAG4 (:259-263,
// 479-482), EOS p and cs2 (eos_analytic.c:12-72, 423-434), conserved
// variables and physical fluxes (grhd_p2c.c:11-106), eigenvalues
// (grhd_eig.c:20-60), LLF / HLL (hrsc_numflx1d.c:14-110).
//
// The reference copies every line of a box into 1-D buffers extended by flat
// extrapolation, forms the N+1 interface fluxes and subtracts their
// differences / h from the right-hand side at all N points, for x, then y,
// then z.  Here one launch per direction does a tile of lines: the line
// segments (plus the 3 points on either side that the stencils reach,
// clamped to the box = flat extrapolation) are staged in LDS, the interface
// fluxes of the tile go to LDS, and each thread then updates one point.
// The x, y, z launches run in order, so every right-hand side value sees the
// reference's sequence of subtractions.  FP contraction is off: products and
// sums round as in the reference's C; with the ideal-gas EOS the result is
// bitwise the reference's (the polytrope's pow is not correctly rounded).
#pragma clang fp contract(off)

#include <cmath>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr int kBlock = 256;
constexpr int kReach = 3;  // stencil points on each side of a segment

// tile of a direction: T points along the line, L lines across; the T + 1
// interfaces of the L lines are one per thread
template <int DIR> struct FluxTile {
  static constexpr int T = DIR == 0 ? 63 : 15;
  static constexpr int L = DIR == 0 ? 4 : 16;
};

struct HrscK {
  int eos, riemann, z;
  double gamma, K, vmax, wlormax;
};

struct FluxArgs {
  HrscK P;
  double c, ooh;
  long noff, sa, sc, st;  // index strides along the line, across, third axis
  int N, na, nt;          // points along, lines across, third-axis extent
  const double *src[13];  // alpha, beta (along), g (permuted 6), rho, epsl, v (permuted 3)
  double *dst[5];         // D, tau, S (permuted 3)
};

__device__ __forceinline__ double sq(double a) { return a * a; }

__device__ double weno_apply(const double b[3], const double uk[3], bool z) {
  const double ow[3] = {1. / 10., 3. / 5., 3. / 10.};
  double a[3], w[3];
  for (int j = 0; j < 3; j++) a[j] = z ? ow[j] * (1. + fabs(b[0] - b[2]) / (1e-40 + b[j])) : ow[j] / sq(1e-40 + b[j]);
  const double dsa = 1. / (a[0] + a[1] + a[2]);
  for (int j = 0; j < 3; j++) w[j] = a[j] * dsa;
  return w[0] * uk[0] + w[1] * uk[1] + w[2] * uk[2];
}

// value at i + 1/2 from the left, u points at i (rec1d_p_wenoz / _weno5)
__device__ double rec_p(const double *u, bool z) {
  const double umt = u[-2], umo = u[-1], ui = u[0], upo = u[1], upt = u[2];
  double b[3], uk[3];
  b[0] = (13. / 12.) * sq(umt - 2. * umo + ui) + (1. / 4.) * sq(umt - 4. * umo + 3. * ui);
  b[1] = (13. / 12.) * sq(umo - 2. * ui + upo) + (1. / 4.) * sq(umo - upo);
  b[2] = (13. / 12.) * sq(ui - 2. * upo + upt) + (1. / 4.) * sq(3. * ui - 4. * upo + upt);
  uk[0] = (1. / 6.) * (2. * umt - 7. * umo + 11. * ui);
  uk[1] = (1. / 6.) * (-1. * umo + 5. * ui + 2. * upo);
  uk[2] = (1. / 6.) * (2. * ui + 5. * upo - upt);
  return weno_apply(b, uk, z);
}

// value at i - 1/2 from the right, u points at i (rec1d_m_wenoz / _weno5)
__device__ double rec_m(const double *u, bool z) {
  const double umt = u[-2], umo = u[-1], ui = u[0], upo = u[1], upt = u[2];
  double b[3], uk[3];
  b[0] = (13. / 12.) * sq(upt - 2. * upo + ui) + (1. / 4.) * sq(upt - 4. * upo + 3. * ui);
  b[1] = (13. / 12.) * sq(upo - 2. * ui + umo) + (1. / 4.) * sq(upo - umo);
  b[2] = (13. / 12.) * sq(ui - 2. * umo + umt) + (1. / 4.) * sq(3. * ui - 4. * umo + umt);
  uk[0] = (1. / 6.) * (2. * upt - 7. * upo + 11. * ui);
  uk[1] = (1. / 6.) * (-upo + 5. * ui + 2. * umo);
  uk[2] = (1. / 6.) * (2. * ui + 5. * umo - umt);
  return weno_apply(b, uk, z);
}

// rec1d_p_lag4, u points at i
__device__ __forceinline__ double lag4(const double *u) { return (9. * (u[0] + u[1]) - u[-1] - u[2]) * (1. / 16.); }

struct Side {
  double q[5], f[5], lam[5], W;
};

__device__ void side_state(const HrscK &P, const double g[8], double detg, double gupxx, const double w[5], bool right,
                           Side &S) {
  const double rho = w[0], epsl = w[1], vx = w[2], vy = w[3], vz = w[4];
  double p, dpdrho, dpdepsl, e = epsl;
  if (P.eos == 1) {
    const double krg = P.K * pow(rho, P.gamma - 1.);
    p = rho * krg;
    dpdepsl = 0.;
    dpdrho = P.gamma * krg;
    if (e <= -1.) e = krg / (P.gamma - 1.);
  } else {
    p = (P.gamma - 1.) * rho * epsl;
    dpdepsl = (P.gamma - 1.) * rho;
    dpdrho = (P.gamma - 1.) * epsl;
  }
  const double cs2 = rho == 0. ? 0. : (dpdrho + p * dpdepsl / (rho * rho)) / (1.0 + e + p / rho);
  const double gxx = g[2], gxy = g[3], gxz = g[4], gyy = g[5], gyz = g[6], gzz = g[7];
  const double vlx = gxx * vx + gxy * vy + gxz * vz;
  const double vly = gxy * vx + gyy * vy + gyz * vz;
  const double vlz = gxz * vx + gyz * vy + gzz * vz;
  double v2 = vlx * vx + vly * vy + vlz * vz;
  v2 = fmin(v2, P.vmax);
  S.W = (right ? v2 >= P.vmax : v2 == P.vmax) ? P.wlormax : 1. / sqrt(1. - v2);
  const double W = S.W;
  const double sqg = sqrt(detg);
  const double W2hrho = W * W * (rho + rho * epsl + p);
  S.q[0] = sqg * W * rho;
  S.q[1] = sqg * (W2hrho - p - W * rho);
  S.q[2] = sqg * W2hrho * vlx;
  S.q[3] = sqg * W2hrho * vly;
  S.q[4] = sqg * W2hrho * vlz;
  const double alpha = g[0], betax = g[1];
  const double avb = alpha * vx - betax;
  S.f[0] = S.q[0] * avb;
  S.f[1] = S.q[1] * avb + alpha * sqg * p * vx;
  S.f[2] = S.q[2] * avb + alpha * sqg * p;
  S.f[3] = S.q[3] * avb;
  S.f[4] = S.q[4] * avb;
  const double Csqrt = sqrt(cs2 * ((1. - v2) * (gupxx * (1. - v2 * cs2) - vx * vx * (1. - cs2))));
  double l0 = alpha * vx - betax;
  double lm = alpha * (vx * (1. - cs2) - Csqrt) / (1. - v2 * cs2) - betax;
  double lp = alpha * (vx * (1. - cs2) + Csqrt) / (1. - v2 * cs2) - betax;
  if (!isfinite(lm)) lm = 0.;
  if (!isfinite(lp)) lp = 0.;
  if (!isfinite(l0)) l0 = 0.;
  S.lam[0] = lm;
  S.lam[1] = S.lam[2] = S.lam[3] = l0;
  S.lam[4] = lp;
}

// the interface flux of grhd_flx1d_fvpr between the reconstructed states
__device__ void interface_flux(const HrscK &P, const double g[8], const double wl[5], const double wr[5],
                               double F[5]) {
  const double gxx = g[2], gxy = g[3], gxz = g[4], gyy = g[5], gyz = g[6], gzz = g[7];
  const double gi11 = gyy * gzz - gyz * gyz, gi12 = gxz * gyz - gxy * gzz, gi13 = gxy * gyz - gxz * gyy;
  const double detg = gxx * gi11 + gxy * gi12 + gxz * gi13;
  const double detginv = 1. / detg;
  const double gupxx = detginv * (gyy * gzz - gyz * gyz);
  Side L, R;
  side_state(P, g, detg, gupxx, wl, false, L);
  side_state(P, g, detg, gupxx, wr, true, R);
  // W at grhd_Wlor_max on either side: the interface flux is 0 (CureFlxVmax)
  const bool capped = L.W == P.wlormax || R.W == P.wlormax;
  double aplus = 0., amins = 0.;
  for (int v = 0; v < 5; v++) {
    aplus = fmax(aplus, fmax(R.lam[v], L.lam[v]));
    amins = fmin(amins, fmin(R.lam[v], L.lam[v]));
  }
  aplus = fabs(aplus);
  amins = fabs(amins);
  if (P.riemann == 1) {
    const double oda = 1. / (aplus + amins + 1e-32);
    const double apm = aplus * amins;
    for (int v = 0; v < 5; v++) F[v] = oda * (aplus * L.f[v] + amins * R.f[v] - apm * (R.q[v] - L.q[v]));
  } else {
    const double amax = fmax(amins, aplus);
    for (int v = 0; v < 5; v++) F[v] = 0.5 * (R.f[v] + L.f[v] - amax * (R.q[v] - L.q[v]));
  }
  for (int v = 0; v < 5; v++) F[v] = capped ? 0. : F[v];
  if (P.eos == 1) F[1] = 0.;
}

template <int DIR>
__global__ void __launch_bounds__(kBlock) grhd_flux_kernel(FluxArgs A) {
  constexpr int T = FluxTile<DIR>::T, L = FluxTile<DIR>::L, S = T + 2 * kReach;
  __shared__ double seg[13][L][S];
  __shared__ double flx[5][L][T + 1];
  const int tid = threadIdx.x;
  const int s = blockIdx.x * T, c0 = blockIdx.y * L, t = blockIdx.z;
  const long base = A.noff + A.st * t;
  // stage the segments [s - 3, s + T + 2] of the tile's lines, clamped to the box
  for (int r = tid; r < L * S; r += kBlock) {
    const int q = DIR == 0 ? r % S : r / L, line = DIR == 0 ? r / S : r % L;
    const int c = c0 + line;
    if (c >= A.na) continue;
    const int a = min(max(s - kReach + q, 0), A.N - 1);
    const long ijk = base + A.sa * a + A.sc * c;
#pragma unroll
    for (int f = 0; f < 13; f++) seg[f][line][q] = A.src[f][ijk];
  }
  __syncthreads();
  // interface fluxes fi = s .. s + T (between 0-based points fi - 1 and fi)
  for (int idx = tid; idx < (T + 1) * L; idx += kBlock) {
    const int fl = DIR == 0 ? idx % (T + 1) : idx / L, line = DIR == 0 ? idx / (T + 1) : idx % L;
    if (c0 + line >= A.na || s + fl > A.N) continue;
    double g[8], wl[5], wr[5], F[5];
    for (int k = 0; k < 8; k++) g[k] = lag4(&seg[k][line][fl + 2]);
    for (int k = 0; k < 5; k++) {
      wl[k] = rec_p(&seg[8 + k][line][fl + 2], A.P.z);
      wr[k] = rec_m(&seg[8 + k][line][fl + 3], A.P.z);
    }
    interface_flux(A.P, g, wl, wr, F);
    bool bad = false;
    for (int k = 0; k < 5; k++) bad |= !isfinite(F[k]);
    for (int k = 0; k < 5; k++) flx[k][line][fl] = bad ? 0. : F[k];
  }
  __syncthreads();
  // one point per thread: rhs -= (F[p + 1] - F[p]) / h  (times c when c != 0)
  const int pl = DIR == 0 ? tid % T : tid / L, line = DIR == 0 ? tid / T : tid % L;
  const int p = s + pl, c = c0 + line;
  if (pl >= T || line >= L || p >= A.N || c >= A.na) return;  // 256 threads, T * L points
  const long ijk = base + A.sa * p + A.sc * c;
  for (int k = 0; k < 5; k++) {
    const double dF = A.ooh * (flx[k][line][pl + 1] - flx[k][line][pl]);
    double r = A.dst[k][ijk];
    if (A.c != 0.)
      r -= A.c * dF;
    else
      r -= dF;
    A.dst[k][ijk] = r;
  }
}

static_assert((FluxTile<0>::T + 1) * FluxTile<0>::L == kBlock && (FluxTile<1>::T + 1) * FluxTile<1>::L == kBlock,
              "one interface of the tile per thread");

constexpr int kVperm[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};                                   // dir_perm_b
constexpr int kTperm[3][6] = {{0, 1, 2, 3, 4, 5}, {3, 4, 1, 5, 2, 0}, {5, 2, 4, 0, 1, 3}};  // dir_perm, tensors

template <int DIR>
void launch_dir(FluxArgs &A, const DevBox &B, double *const met[10], double *const prim[5], double *const rhs[5]) {
  constexpr int T = FluxTile<DIR>::T, L = FluxTile<DIR>::L;
  const long m = B.m, mn = (long)B.m * B.n;
  A.noff = B.noffset;
  if (DIR == 0) {
    A.N = B.m, A.na = B.n, A.nt = B.o, A.sa = 1, A.sc = m, A.st = mn, A.ooh = 1. / B.dx;
  } else if (DIR == 1) {
    A.N = B.n, A.na = B.m, A.nt = B.o, A.sa = m, A.sc = 1, A.st = mn, A.ooh = 1. / B.dy;
  } else {
    A.N = B.o, A.na = B.m, A.nt = B.n, A.sa = mn, A.sc = 1, A.st = m, A.ooh = 1. / B.dz;
  }
  A.src[0] = met[0];
  A.src[1] = met[1 + kVperm[DIR][0]];
  for (int k = 0; k < 6; k++) A.src[2 + k] = met[4 + kTperm[DIR][k]];
  A.src[8] = prim[0];
  A.src[9] = prim[1];
  for (int k = 0; k < 3; k++) A.src[10 + k] = prim[2 + kVperm[DIR][k]];
  A.dst[0] = rhs[0];
  A.dst[1] = rhs[1];
  for (int k = 0; k < 3; k++) A.dst[2 + k] = rhs[2 + kVperm[DIR][k]];
  dim3 grid((A.N + T - 1) / T, (A.na + L - 1) / L, A.nt);
  grhd_flux_kernel<DIR><<<grid, kBlock, 0, stream()>>>(A);
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_grhd_fluxes(bamhip_level *lev, const bamhip_hrsc_params *par, const int *unew, double c,
                                  const int *metric, const int *prim) {
  const char *what = "bamhip_grhd_fluxes";
  if (!lev || !par) return set_error(BAMHIP_EINVAL, "%s: NULL argument", what);
  if (par->eos != 0 && par->eos != 1) return set_error(BAMHIP_EINVAL, "%s: eos=%d (0 ideal, 1 poly)", what, par->eos);
  if (par->riemann != 0 && par->riemann != 1)
    return set_error(BAMHIP_EINVAL, "%s: riemann=%d (0 LLF, 1 HLL)", what, par->riemann);
  if (par->rec != 0 && par->rec != 1) return set_error(BAMHIP_EINVAL, "%s: rec=%d (0 WENOZ, 1 WENO5)", what, par->rec);
  double *met[10], *pr[5], *rhs[5];
  int e;
  if ((e = gather_ptrs(lev, metric, 10, met, "bamhip_grhd_fluxes(metric)"))) return e;
  if ((e = gather_ptrs(lev, prim, 5, pr, "bamhip_grhd_fluxes(prim)"))) return e;
  if ((e = gather_ptrs(lev, unew, 5, rhs, "bamhip_grhd_fluxes(unew)"))) return e;
  for (const DevBox &B : lev->boxes)
    if (B.m < 1 || B.n < 1 || B.o < 1 || B.m > 65535 || B.n > 65535 || B.o > 65535)
      return set_error(BAMHIP_EINVAL, "%s: box %dx%dx%d outside the launch limits", what, B.m, B.n, B.o);
  TimerScope ts("matter_rhs_fluxes");
  FluxArgs A{};
  A.P = HrscK{par->eos, par->riemann, par->rec == 0, par->gamma, par->K, par->vmax, par->wlormax};
  A.c = c;
  for (const DevBox &B : lev->boxes) launch_dir<0>(A, B, met, pr, rhs);
  for (const DevBox &B : lev->boxes) launch_dir<1>(A, B, met, pr, rhs);
  for (const DevBox &B : lev->boxes) launch_dir<2>(A, B, met, pr, rhs);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}
