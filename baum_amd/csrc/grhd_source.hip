// grhd_source.hip -- GRHD source terms and matter ADM variables on gfx950.
//
// Reference: grhd_sources (src/projects/grhd/grhd_sources.m:8-190, generated
// as grhd_sources.c), MATTER.sources of matter_rhs (matter.c:438-442), with
// the reference's defaults grhd_source_computation = prim and grhd_cooling =
// 0 (bam_grhd.c:51-52).  At every inner point of every box:
//   - detg <= 0 or alpha <= 0: no source and matter ADM variables 0;
//   - centred first derivatives of alpha, beta^i, g_ij, 2nd order at distance
//     1 from a box face (or with order_centered 2), else 4th order, in the
//     multiplied form of the generated code;
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
// A block computes a 16x4x4 tile of points.  The ten metric fields pass
// through LDS one at a time (the tile plus 2 points on every side, indices
// clamped to the box: only the stencils of inner points read the halo, and
// they stay inside it), two fields ahead prefetched into registers while the
// current one is differentiated; each thread keeps its 30 first derivatives
// in registers and then forms the sources.  The sums run over full 3x3 index
// ranges in a fixed order and FP contraction is off, so the results are
// bitwise those of the C restatement that the tests compare with.
#pragma clang fp contract(off)

#include <cmath>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr int kSrcBlock = 256;
constexpr int kTX = 16, kTY = 4, kTZ = 4, kHalo = 2;
constexpr int kSX = kTX + 2 * kHalo, kSY = kTY + 2 * kHalo, kSZ = kTZ + 2 * kHalo, kSN = kSX * kSY * kSZ;
constexpr int kPer = (kSN + kSrcBlock - 1) / kSrcBlock;  // staged values per thread and field
static_assert(kTX * kTY * kTZ == kSrcBlock, "one point of the tile per thread");

struct SrcArgs {
  int m, n, o, skip[6];
  double oo2h[3];  // 1 / (2 h)
  double rhoatm, wlormax, c;
  int cold, order_centered;
  const double *met[10], *K[6], *prim[6];
  double *rhs[5], *madm[11];
};

__device__ constexpr int s6(int a, int b) { return a <= b ? (a == 0 ? b : (a == 1 ? 2 + b : 5)) : s6(b, a); }

// first derivative, centred, in the form of the generated code
// (grhd_sources.c: oo2dx (-u[-1] + u[1]), 1/6 oo2dx (u[-2] + 8 (-u[-1] + u[1]) - u[2]))
__device__ __forceinline__ double d1(const double *u, int q, int s, double oo2h, int ord) {
  if (ord == 2) return oo2h * (-u[q - s] + u[q + s]);
  return 0.16666666666666666667 * oo2h * (u[q - 2 * s] + 8. * (-u[q - s] + u[q + s]) - u[q + 2 * s]);
}

// the staged values of a field for this thread: off[e] is the box index of
// staged element tid + e * 256 (clamped to the box), -1 past the tile
__device__ __forceinline__ void stage_load(const double *f, const long off[kPer], double r[kPer]) {
#pragma unroll
  for (int e = 0; e < kPer; e++) r[e] = off[e] >= 0 ? f[off[e]] : 0.;
}

__global__ void __launch_bounds__(kSrcBlock) grhd_source_kernel(SrcArgs A) {
  __shared__ double buf[3][kSN];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * kTX, j0 = blockIdx.y * kTY, k0 = blockIdx.z * kTZ;
  const int i = i0 + tid % kTX, j = j0 + (tid / kTX) % kTY, k = k0 + tid / (kTX * kTY);
  const int imax = A.m - 1, jmax = A.n - 1, kmax = A.o - 1;
  const bool inner = i >= A.skip[0] && i <= imax - A.skip[1] && j >= A.skip[2] && j <= jmax - A.skip[3] &&
                     k >= A.skip[4] && k <= kmax - A.skip[5];
  const int ord =
      (A.order_centered == 2 || i == 1 || j == 1 || k == 1 || i == imax - 1 || j == jmax - 1 || k == kmax - 1) ? 2 : 4;
  // position of the point in the staged tile and the staged strides
  const int c = (tid % kTX + kHalo) + kSX * ((tid / kTX) % kTY + kHalo) + kSX * kSY * (tid / (kTX * kTY) + kHalo);
  const int sst[3] = {1, kSX, kSX * kSY};
  long off[kPer];
#pragma unroll
  for (int e = 0; e < kPer; e++) {
    const int idx = tid + e * kSrcBlock;
    const int sx = idx % kSX, sy = (idx / kSX) % kSY, sz = idx / (kSX * kSY);
    const int gi = min(max(i0 - kHalo + sx, 0), A.m - 1), gj = min(max(j0 - kHalo + sy, 0), A.n - 1),
              gk = min(max(k0 - kHalo + sz, 0), A.o - 1);
    off[e] = idx < kSN ? gi + (long)A.m * gj + (long)A.m * A.n * gk : -1;
  }
  // point values read once, issued before the staging so their latency hides behind it
  long q = 0;
  double K[6], rho = 0., eps = 0., p = 0., vu[3] = {0., 0., 0.};
  if (inner) {
    q = (long)i + (long)A.m * j + (long)A.m * A.n * k;
#pragma unroll
    for (int a = 0; a < 6; a++) K[a] = A.K[a][q];
    rho = A.prim[0][q], eps = A.prim[1][q], p = A.prim[5][q];
    vu[0] = A.prim[2][q], vu[1] = A.prim[3][q], vu[2] = A.prim[4][q];
  }
  // the ten metric fields through three LDS buffers, two fields ahead: value
  // at the point and its 3 derivatives
  double val[10], der[3][10];
  double r0[kPer], r1[kPer];
  stage_load(A.met[0], off, r0);
  stage_load(A.met[1], off, r1);
#pragma unroll
  for (int e = 0; e < kPer; e++)
    if (tid + e * kSrcBlock < kSN) buf[0][tid + e * kSrcBlock] = r0[e];
  __syncthreads();
#pragma unroll
  for (int f = 0; f < 10; f++) {
    // r1 holds field f + 1; r0 receives field f + 2
    if (f + 2 < 10) stage_load(A.met[f + 2], off, r0);
    const double *u = buf[f % 3];
    val[f] = u[c];
#pragma unroll
    for (int d = 0; d < 3; d++) der[d][f] = inner ? d1(u, c, sst[d], A.oo2h[d], ord) : 0.;
    if (f + 1 < 10) {
#pragma unroll
      for (int e = 0; e < kPer; e++)
        if (tid + e * kSrcBlock < kSN) buf[(f + 1) % 3][tid + e * kSrcBlock] = r1[e];
    }
#pragma unroll
    for (int e = 0; e < kPer; e++) r1[e] = r0[e];
    __syncthreads();
  }
  if (!inner) return;
  const double al = val[0];
  const double be[3] = {val[1], val[2], val[3]};
  double g[6];
#pragma unroll
  for (int a = 0; a < 6; a++) g[a] = val[4 + a];
  const double adj[6] = {g[3] * g[5] - g[4] * g[4], g[2] * g[4] - g[1] * g[5], g[1] * g[4] - g[2] * g[3],
                         g[0] * g[5] - g[2] * g[2], g[1] * g[2] - g[0] * g[4], g[0] * g[3] - g[1] * g[1]};
  const double detg = g[0] * adj[0] + g[1] * adj[1] + g[2] * adj[2];
  double sD = 0., sT = 0., sS[3] = {0., 0., 0.}, ad[11];
#pragma unroll
  for (int a = 0; a < 11; a++) ad[a] = 0.;
  if (detg > 0. && al > 0.) {
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
    for (int d = 0; d < 3; d++) dal[d] = der[d][0];
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
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double dbe[3], dg[6];
#pragma unroll
      for (int a = 0; a < 3; a++) dbe[a] = der[d][1 + a];
#pragma unroll
      for (int a = 0; a < 6; a++) dg[a] = der[d][4 + a];
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
    A.oo2h[0] = 1 / (2 * B.dx), A.oo2h[1] = 1 / (2 * B.dy), A.oo2h[2] = 1 / (2 * B.dz);
    A.rhoatm = rhoatm, A.wlormax = par->wlormax, A.c = c;
    A.cold = par->eos == 1, A.order_centered = order_centered;
    for (int a = 0; a < 10; a++) A.met[a] = met[a] + B.noffset;
    for (int a = 0; a < 6; a++) A.K[a] = kk[a] + B.noffset, A.prim[a] = pr[a] + B.noffset;
    for (int a = 0; a < 5; a++) A.rhs[a] = rhs[a] + B.noffset;
    for (int a = 0; a < 11; a++) A.madm[a] = ad[a] + B.noffset;
    dim3 grid((B.m + kTX - 1) / kTX, (B.n + kTY - 1) / kTY, (B.o + kTZ - 1) / kTZ);
    grhd_source_kernel<<<grid, kSrcBlock, 0, stream()>>>(A);
  }
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}
