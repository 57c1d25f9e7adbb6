// grhd_source.hip -- GRHD source terms and matter ADM variables on gfx950.
//
// Reference: grhd_sources (src/projects/grhd/grhd_sources.m:8-190, generated
// as grhd_sources.c), MATTER.sources of matter_rhs (matter.c:438-442), with
// the reference's defaults grhd_source_computation = prim and grhd_cooling =
// 0 (bam_grhd.c:51-52).  At every inner point of every box:
//   - detg <= 0 or alpha <= 0: no source and matter ADM variables 0;
//   - centred first derivatives of alpha, beta^i, g_ij, 2nd order at distance
//     1 from a box face (or with order_centered 2), else 4th order;
//   - T^00, T^0i, T^ij, T^0_j from the primitives, W capped at grhd_Wlor_max
//     when v^2 >= 1;
//   - s_tau = alpha sqrt(g) (T^00 (beta beta K - beta d alpha)
//     + T^0i (2 beta^j K_ij - d_i alpha) + T^ij K_ij),
//     s_S_c = alpha sqrt(g) (T^00 (beta beta d_c g / 2 - alpha d_c alpha)
//     + T^0i beta^j d_c g_ij + T^0_j d_c beta^j + T^ij d_c g_ij / 2);
//   - matter ADM variables rho, S_i, S_ij, S for the Z4c matter terms;
//   - no momentum source below the atmosphere density, no tau source for a
//     cold EOS, all 16 values 0 when any is not finite;
//   - unew += source (times c when c != 0).
// One thread per point; the neighbours of the derivative stencils come
// through L1/L2.  The sums run over full 3x3 index ranges in a fixed order
// and FP contraction is off, so the results are bitwise those of the C
// restatement that the tests compare with.
#pragma clang fp contract(off)

#include <cmath>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr int kSrcBlock = 256;

struct SrcArgs {
  int m, n, o, skip[6];
  double h[3];
  double rhoatm, wlormax, c;
  int cold, order_centered;
  const double *met[10], *K[6], *prim[6];
  double *rhs[5], *madm[11];
};

__device__ constexpr int s6(int a, int b) { return a <= b ? (a == 0 ? b : (a == 1 ? 2 + b : 5)) : s6(b, a); }

__device__ __forceinline__ double d1(const double *u, long q, long s, double h, int ord) {
  if (ord == 2) return (u[q + s] - u[q - s]) / (2. * h);
  return (u[q - 2 * s] - 8. * u[q - s] + 8. * u[q + s] - u[q + 2 * s]) / (12. * h);
}

__global__ void __launch_bounds__(kSrcBlock) grhd_source_kernel(SrcArgs A) {
  const long npts = (long)A.m * A.n * A.o;
  const long q = (long)blockIdx.x * kSrcBlock + threadIdx.x;
  if (q >= npts) return;
  const long m = A.m, mn = (long)A.m * A.n;
  const int i = (int)(q % m), j = (int)((q / m) % A.n), k = (int)(q / mn);
  const int imax = A.m - 1, jmax = A.n - 1, kmax = A.o - 1;
  if (i < A.skip[0] || i > imax - A.skip[1] || j < A.skip[2] || j > jmax - A.skip[3] || k < A.skip[4] ||
      k > kmax - A.skip[5])
    return;
  const long str[3] = {1, m, mn};
  const double al = A.met[0][q];
  const double be[3] = {A.met[1][q], A.met[2][q], A.met[3][q]};
  double g[6], K[6];
#pragma unroll
  for (int a = 0; a < 6; a++) {
    g[a] = A.met[4 + a][q];
    K[a] = A.K[a][q];
  }
  const double rho = A.prim[0][q], eps = A.prim[1][q], p = A.prim[5][q];
  const double vu[3] = {A.prim[2][q], A.prim[3][q], A.prim[4][q]};
  const double adj[6] = {g[3] * g[5] - g[4] * g[4], g[2] * g[4] - g[1] * g[5], g[1] * g[4] - g[2] * g[3],
                         g[0] * g[5] - g[2] * g[2], g[1] * g[2] - g[0] * g[4], g[0] * g[3] - g[1] * g[1]};
  const double detg = g[0] * adj[0] + g[1] * adj[1] + g[2] * adj[2];
  double sD = 0., sT = 0., sS[3] = {0., 0., 0.}, ad[11];
#pragma unroll
  for (int a = 0; a < 11; a++) ad[a] = 0.;
  if (detg > 0. && al > 0.) {
    const int ord =
        (A.order_centered == 2 || i == 1 || j == 1 || k == 1 || i == imax - 1 || j == jmax - 1 || k == kmax - 1) ? 2
                                                                                                                : 4;
    const double detginv = 1. / detg;
    double gi[6];
#pragma unroll
    for (int a = 0; a < 6; a++) gi[a] = detginv * adj[a];
    const double sqrtdetg = sqrt(detg), sqrtmdetg = al * sqrtdetg;
    double vl[3], vsq = 0.;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      vl[a] = 0.;
#pragma unroll
      for (int b = 0; b < 3; b++) vl[a] += g[s6(a, b)] * vu[b];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) vsq += vu[a] * vl[a];
    double W = 1. / sqrt(1. - vsq);
    if (vsq >= 1.) W = A.wlormax;
    const double W2hrho = W * W * (rho + rho * eps + p);
    const double al2 = al * al;
    double vs[3], T0i[3], T0j[3], Tij[6];
#pragma unroll
    for (int a = 0; a < 3; a++) vs[a] = vu[a] - be[a] / al;
    const double T00 = (W2hrho - p) / al2;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      T0i[a] = W2hrho * vs[a] / al + p * be[a] / al2;
      T0j[a] = W2hrho * vl[a] / al;
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) Tij[s6(a, b)] = W2hrho * vs[a] * vs[b] + p * (gi[s6(a, b)] - be[a] * be[b] / al2);
    double dal[3];
#pragma unroll
    for (int d = 0; d < 3; d++) dal[d] = d1(A.met[0], q, str[d], A.h[d], ord);
    double bbK = 0., bda = 0., t0 = 0., tk = 0.;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      bda += be[a] * dal[a];
      double bK = 0.;
#pragma unroll
      for (int b = 0; b < 3; b++) {
        bbK += be[a] * be[b] * K[s6(a, b)];
        bK += be[b] * K[s6(a, b)];
        tk += Tij[s6(a, b)] * K[s6(a, b)];
      }
      t0 += T0i[a] * (2. * bK - dal[a]);
    }
    sT = sqrtmdetg * (T00 * (bbK - bda) + t0 + tk);
    // one direction at a time: its 9 derivatives feed its momentum source
    // (the barrier keeps the next direction's stencil loads from being
    // hoisted, which would hold 120 loads in registers at once)
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double dbe[3], dg[6];
#pragma unroll
      for (int a = 0; a < 3; a++) dbe[a] = d1(A.met[1 + a], q, str[d], A.h[d], ord);
#pragma unroll
      for (int a = 0; a < 6; a++) dg[a] = d1(A.met[4 + a], q, str[d], A.h[d], ord);
      double bbdg = 0., t0dg = 0., tdg = 0., tdb = 0.;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        tdb += T0j[a] * dbe[a];
#pragma unroll
        for (int b = 0; b < 3; b++) {
          bbdg += be[a] * be[b] * dg[s6(a, b)];
          t0dg += T0i[a] * be[b] * dg[s6(a, b)];
          tdg += Tij[s6(a, b)] * dg[s6(a, b)];
        }
      }
      sS[d] = sqrtmdetg * (T00 * (0.5 * bbdg - al * dal[d]) + t0dg + tdb + 0.5 * tdg);
      asm volatile("" ::: "memory");
    }
    double vv = 0.;
#pragma unroll
    for (int a = 0; a < 3; a++) vv += vl[a] * vu[a];
    ad[0] = W2hrho - p;
#pragma unroll
    for (int a = 0; a < 3; a++) ad[1 + a] = W2hrho * vu[a];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) ad[4 + s6(a, b)] = W2hrho * vl[a] * vl[b] + g[s6(a, b)] * p;
    ad[10] = W2hrho * vv + 3. * p;
    if (rho < A.rhoatm) sS[0] = sS[1] = sS[2] = 0.;
    if (A.cold) sT = 0.;
    bool bad = !isfinite(sD) || !isfinite(sT);
#pragma unroll
    for (int a = 0; a < 3; a++) bad |= !isfinite(sS[a]);
#pragma unroll
    for (int a = 0; a < 11; a++) bad |= !isfinite(ad[a]);
    if (bad) {
      sD = sT = sS[0] = sS[1] = sS[2] = 0.;
#pragma unroll
      for (int a = 0; a < 11; a++) ad[a] = 0.;
    }
  }
#pragma unroll
  for (int a = 0; a < 11; a++) A.madm[a][q] = ad[a];
  const double s5[5] = {sD, sT, sS[0], sS[1], sS[2]};
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const double r = A.rhs[v][q];
    A.rhs[v][q] = A.c != 0. ? r + A.c * s5[v] : r + s5[v];
  }
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_grhd_sources(bamhip_level *lev, const bamhip_hrsc_params *par, double rhoatm, int order_centered,
                                   const int *unew, double c, const int *metric, const int *Kij, const int *prim,
                                   const int *madm) {
  const char *what = "bamhip_grhd_sources";
  if (!lev || !par) return set_error(BAMHIP_EINVAL, "%s: NULL argument", what);
  if (par->eos != 0 && par->eos != 1) return set_error(BAMHIP_EINVAL, "%s: eos=%d (0 ideal, 1 poly)", what, par->eos);
  if (order_centered != 2 && order_centered != 4)
    return set_error(BAMHIP_EINVAL, "%s: order_centered=%d (2 or 4)", what, order_centered);
  double *met[10], *kk[6], *pr[6], *rhs[5], *ad[11];
  int e;
  if ((e = gather_ptrs(lev, metric, 10, met, "bamhip_grhd_sources(metric)"))) return e;
  if ((e = gather_ptrs(lev, Kij, 6, kk, "bamhip_grhd_sources(Kij)"))) return e;
  if ((e = gather_ptrs(lev, prim, 6, pr, "bamhip_grhd_sources(prim)"))) return e;
  if ((e = gather_ptrs(lev, unew, 5, rhs, "bamhip_grhd_sources(unew)"))) return e;
  if ((e = gather_ptrs(lev, madm, 11, ad, "bamhip_grhd_sources(madm)"))) return e;
  for (const DevBox &B : lev->boxes) {
    // the 4th-order stencil of an inner point reaches 2 points: inner points
    // at distance 0 from a face would read outside the box
    for (int f = 0; f < 6; f++)
      if (B.skip[f] < 1) return set_error(BAMHIP_EINVAL, "%s: box face %d has no boundary layer", what, f);
  }
  TimerScope ts("matter_rhs_source");
  for (const DevBox &B : lev->boxes) {
    SrcArgs A{};
    A.m = B.m, A.n = B.n, A.o = B.o;
    for (int f = 0; f < 6; f++) A.skip[f] = B.skip[f];
    A.h[0] = B.dx, A.h[1] = B.dy, A.h[2] = B.dz;
    A.rhoatm = rhoatm, A.wlormax = par->wlormax, A.c = c;
    A.cold = par->eos == 1, A.order_centered = order_centered;
    for (int a = 0; a < 10; a++) A.met[a] = met[a] + B.noffset;
    for (int a = 0; a < 6; a++) A.K[a] = kk[a] + B.noffset, A.prim[a] = pr[a] + B.noffset;
    for (int a = 0; a < 5; a++) A.rhs[a] = rhs[a] + B.noffset;
    for (int a = 0; a < 11; a++) A.madm[a] = ad[a] + B.noffset;
    const long npts = (long)B.m * B.n * B.o;
    grhd_source_kernel<<<(unsigned)((npts + kSrcBlock - 1) / kSrcBlock), kSrcBlock, 0, stream()>>>(A);
  }
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}
