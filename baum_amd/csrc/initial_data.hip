// initial_data.hip -- puncture initial data and the ADM -> Z4c conversion on
// gfx950.
//
// Brill-Lindquist data: SetMWBL with mwblpsiofxyz (src/projects/punctures/
// ConformalFactor.c:84-200): psi = sum_i (m_i (1/r_i)) 0.5 + 1, r replaced by
// EPSZERO (punctures.h:85) where it equals 0 within dequaleps, conformal
// metric delta and K_ij = 0 (ConfToPhys, :234-260), the conformal factor
// absorbed into the metric (adm_undo_conftrans, src/projects/adm/adm.c:96-111).
//
// ADM -> Z4c: the z4 project's POST_INITIALDATA (src/projects/z4/z4.c:
// 280-330): lapse one or precollapsed (1/sqrt(adm_gxx)), zero shift, then
// adm_to_z4 (z4.c:142-171) = z4_init (z4_init.m:9-80) at every point, G^a =
// -d_b gt^{ba} at the inner points with the generated 2nd-order stencil, and
// z4_setbound_G (z4.c:33-39) = set_boundary_extrapolate (src/utility/boundary/
// extrapolate.c:10-47) on PHYBOUND points.  The extrapolation is the
// reference's three passes in point order, in place, with the updated values
// of earlier points feeding later ones; it runs on one thread per box (an
// initialisation step, a few hundred thousand points).
#pragma clang fp contract(off)

#include <cmath>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr int kIdBlock = 256;
constexpr int kMaxPunc = 16;
constexpr int kSY[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

struct BlArgs {
  DevBox B;
  int n;
  double mass[kMaxPunc], pos[3 * kMaxPunc];
  double *g[6], *K[6];
};

__global__ void __launch_bounds__(kIdBlock) punctures_bl_kernel(BlArgs A) {
  const long np = (long)A.B.m * A.B.n * A.B.o;
  const long q = (long)blockIdx.x * kIdBlock + threadIdx.x;
  if (q >= np) return;
  const int i = (int)(q % A.B.m), j = (int)((q / A.B.m) % A.B.n), k = (int)(q / ((long)A.B.m * A.B.n));
  const double x = A.B.x0 + i * A.B.dx, y = A.B.y0 + j * A.B.dy, z = A.B.z0 + k * A.B.dz;
  double p = 0;
  for (int c = 0; c < A.n; c++) {
    const double xi = x - A.pos[3 * c], yi = y - A.pos[3 * c + 1], zi = z - A.pos[3 * c + 2];
    double r = sqrt(xi * xi + yi * yi + zi * zi);
    if (!(r < 0.0 - 1e-10 || 0.0 < r - 1e-10)) r = 1e-10 / 1000.0;
    const double ri = 1 / r;
    p += A.mass[c] * ri * (0.5);
  }
  p += 1.0;
  const double psi4 = pow(p, 4.0), kf = pow(p, -2.0);
  const long o = A.B.noffset + q;
#pragma unroll
  for (int a = 0; a < 6; a++) {
    A.g[a][o] = (a == 0 || a == 3 || a == 5 ? 1.0 : 0.0) * psi4;
    A.K[a][o] = 0.0 * kf;
  }
}

struct InitArgs {
  const double *g[6], *K[6];
  double *u[25];
  double *gtinv[6];
  double oopower, chipsipower;
  int lapse, zero_shift;
  long n;
};

__device__ __forceinline__ double det3(const double g[3][3]) {
  return g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1]) - g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0]) +
         g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
}

// z4_init at every point (with the lapse and shift choices of z4.c:280-320)
__global__ void __launch_bounds__(kIdBlock) z4_init_point_kernel(InitArgs A) {
  const long q = (long)blockIdx.x * kIdBlock + threadIdx.x;
  if (q >= A.n) return;
  double g[3][3], K[3][3], gt[3][3], gi[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      g[a][c] = A.g[kSY[a][c]][q];
      K[a][c] = A.K[kSY[a][c]][q];
    }
  if (A.lapse == 1) A.u[18][q] = 1.;
  if (A.lapse == 2) A.u[18][q] = 1. / sqrt(g[0][0]);
  if (A.zero_shift) {
#pragma unroll
    for (int a = 0; a < 3; a++) A.u[19 + a][q] = 0., A.u[22 + a][q] = 0.;
  }
  const double detg = det3(g);
  const double psim4 = pow(detg, -1.0 / 3.0), psi = pow(detg, 1.0 / 12.0);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int c = 0; c < 3; c++) gt[a][c] = psim4 * g[a][c];
  const double id = 1.0 / det3(gt);
  gi[0][0] = id * (gt[1][1] * gt[2][2] - gt[1][2] * gt[2][1]);
  gi[0][1] = id * (gt[0][2] * gt[2][1] - gt[0][1] * gt[2][2]);
  gi[0][2] = id * (gt[0][1] * gt[1][2] - gt[0][2] * gt[1][1]);
  gi[1][0] = id * (gt[1][2] * gt[2][0] - gt[1][0] * gt[2][2]);
  gi[1][1] = id * (gt[0][0] * gt[2][2] - gt[0][2] * gt[2][0]);
  gi[1][2] = id * (gt[0][2] * gt[1][0] - gt[0][0] * gt[1][2]);
  gi[2][0] = id * (gt[1][0] * gt[2][1] - gt[1][1] * gt[2][0]);
  gi[2][1] = id * (gt[0][1] * gt[2][0] - gt[0][0] * gt[2][1]);
  gi[2][2] = id * (gt[0][0] * gt[1][1] - gt[0][1] * gt[1][0]);
  double trK = 0;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int c = 0; c < 3; c++) trK += gi[a][c] * psim4 * K[a][c];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int c = a; c < 3; c++) {
      A.u[kSY[a][c]][q] = gt[a][c];
      A.u[7 + kSY[a][c]][q] = psim4 * (K[a][c] - g[a][c] * trK / 3);
      A.gtinv[kSY[a][c]][q] = gi[a][c];
    }
  A.u[13][q] = trK;
  A.u[6][q] = pow(psi, A.chipsipower);
  A.u[17][q] = 0;
}

// G^a = -d_b gt^{ba} at the inner points of one box (generated form oo2dx (-u[-1] + u[1]))
__global__ void __launch_bounds__(kIdBlock) z4_init_gamma_kernel(InitArgs A, DevBox B) {
  const int ni = B.m - B.skip[0] - B.skip[1], nj = B.n - B.skip[2] - B.skip[3], nk = B.o - B.skip[4] - B.skip[5];
  const long t = (long)blockIdx.x * kIdBlock + threadIdx.x;
  if (ni <= 0 || nj <= 0 || nk <= 0 || t >= (long)ni * nj * nk) return;
  const int i = (int)(t % ni) + B.skip[0], j = (int)((t / ni) % nj) + B.skip[2], k = (int)(t / ((long)ni * nj)) + B.skip[4];
  const long st[3] = {1, B.m, (long)B.m * B.n};
  const double oo2h[3] = {1 / (2 * B.dx), 1 / (2 * B.dy), 1 / (2 * B.dz)};
  const long q = B.noffset + i + st[1] * j + st[2] * k;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double *gi = A.gtinv[kSY[c][a]];
      s -= oo2h[c] * (-gi[q - st[c]] + gi[q + st[c]]);
    }
    A.u[14 + a][q] = s;
  }
}

// set_boundary_extrapolate on the outermost layer of one box, in point order
__global__ void z4_init_extrapolate_kernel(double *v0, double *v1, double *v2, DevBox B) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double *vs[3] = {v0, v1, v2};
  const long di = 1, dj = B.m, dk = (long)B.m * B.n;
  for (int f = 0; f < 3; f++) {
    double *v = vs[f] + B.noffset;
    for (int rep = 0; rep < 3; rep++)
      for (int k = 0; k < B.o; k++)
        for (int j = 0; j < B.n; j++) {
          const bool row_out = j == 0 || k == 0 || j == B.n - 1 || k == B.o - 1;
          for (int i = 0; i < B.m; i += (row_out || i == B.m - 1) ? 1 : B.m - 1) {
            const long c = i + dj * j + dk * k;
            const long nb[3][4] = {
                {i > 0 ? c - di : c, i > 1 ? c - 2 * di : c, i + 1 < B.m ? c + di : c, i + 2 < B.m ? c + 2 * di : c},
                {j > 0 ? c - dj : c, j > 1 ? c - 2 * dj : c, j + 1 < B.n ? c + dj : c, j + 2 < B.n ? c + 2 * dj : c},
                {k > 0 ? c - dk : c, k > 1 ? c - 2 * dk : c, k + 1 < B.o ? c + dk : c, k + 2 < B.o ? c + 2 * dk : c}};
            double val[3];
            int use[3];
            for (int d = 0; d < 3; d++) {
              const long m1 = nb[d][0], M2 = nb[d][1], p1 = nb[d][2], P2 = nb[d][3];
              if (p1 == c) {
                val[d] = 2.0 * v[m1] - v[M2];
                use[d] = 1;
              } else if (m1 == c) {
                val[d] = 2.0 * v[p1] - v[P2];
                use[d] = 1;
              } else {
                val[d] = 0.5 * (v[p1] + v[m1]);
                use[d] = 0;
              }
            }
            if (use[0] + use[1] + use[2] > 0)
              v[c] = (use[0] * val[0] + use[1] * val[1] + use[2] * val[2]) / (use[0] + use[1] + use[2]);
            else
              v[c] = (val[0] + val[1] + val[2]) / 3.0;
          }
        }
  }
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" int bamhip_punctures_bl(bamhip_level *lev, int n, const double *mass, const double *pos, const int *adm_g,
                                   const int *adm_K) {
  const char *what = "bamhip_punctures_bl";
  if (!lev || !mass || !pos || n < 1 || n > kMaxPunc) return set_error(BAMHIP_EINVAL, "%s: bad arguments (n=%d)", what, n);
  double *g[6], *K[6];
  int e;
  if ((e = gather_ptrs(lev, adm_g, 6, g, "bamhip_punctures_bl(adm_g)"))) return e;
  if ((e = gather_ptrs(lev, adm_K, 6, K, "bamhip_punctures_bl(adm_K)"))) return e;
  for (const DevBox &B : lev->boxes) {
    BlArgs A{};
    A.B = B;
    A.n = n;
    for (int c = 0; c < n; c++) {
      A.mass[c] = mass[c];
      for (int d = 0; d < 3; d++) A.pos[3 * c + d] = pos[3 * c + d];
    }
    for (int a = 0; a < 6; a++) A.g[a] = g[a], A.K[a] = K[a];
    const long np = (long)B.m * B.n * B.o;
    punctures_bl_kernel<<<(unsigned)((np + kIdBlock - 1) / kIdBlock), kIdBlock, 0, stream()>>>(A);
  }
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

extern "C" int bamhip_adm_to_z4c(bamhip_level *lev, const int *adm_g, const int *adm_K, const int *ucur,
                                 double chipsipower, int initial_lapse, int zero_shift) {
  const char *what = "bamhip_adm_to_z4c";
  if (!lev) return set_error(BAMHIP_EINVAL, "%s: NULL level", what);
  if (initial_lapse < 0 || initial_lapse > 2)
    return set_error(BAMHIP_EINVAL, "%s: initial_lapse %d (0 unchanged, 1 one, 2 precollapsed)", what, initial_lapse);
  InitArgs A{};
  double *g[6], *K[6];
  int e;
  if ((e = gather_ptrs(lev, adm_g, 6, g, "bamhip_adm_to_z4c(adm_g)"))) return e;
  if ((e = gather_ptrs(lev, adm_K, 6, K, "bamhip_adm_to_z4c(adm_K)"))) return e;
  if ((e = gather_ptrs(lev, ucur, BAMHIP_Z4C_NVARS, A.u, "bamhip_adm_to_z4c(ucur)"))) return e;
  for (int a = 0; a < 6; a++) A.g[a] = g[a], A.K[a] = K[a];
  A.n = lev->npoints;
  A.oopower = 1. / chipsipower;
  A.chipsipower = chipsipower;
  A.lapse = initial_lapse;
  A.zero_shift = zero_shift;
  double *tmp = nullptr;  // the temporary inverse conformal metric (bssn_ginvtmpxx)
  BH_HIP(hipMallocAsync((void **)&tmp, sizeof(double) * 6 * A.n, stream()));
  for (int a = 0; a < 6; a++) A.gtinv[a] = tmp + a * A.n;
  z4_init_point_kernel<<<(unsigned)((A.n + kIdBlock - 1) / kIdBlock), kIdBlock, 0, stream()>>>(A);
  for (const DevBox &B : lev->boxes) {
    const long ni = B.m - B.skip[0] - B.skip[1], nj = B.n - B.skip[2] - B.skip[3], nk = B.o - B.skip[4] - B.skip[5];
    if (ni > 0 && nj > 0 && nk > 0)
      z4_init_gamma_kernel<<<(unsigned)((ni * nj * nk + kIdBlock - 1) / kIdBlock), kIdBlock, 0, stream()>>>(A, B);
  }
  for (const DevBox &B : lev->boxes) {
    bool phy = B.m > 2 && B.n > 2 && B.o > 2;
    for (int f = 0; f < 6; f++) phy = phy && B.bflag[f] == BAMHIP_PHYBOUND && B.skip[f] == 1;
    if (phy) z4_init_extrapolate_kernel<<<1, 64, 0, stream()>>>(A.u[14], A.u[15], A.u[16], B);
  }
  BH_HIP(hipGetLastError());
  BH_HIP(hipFreeAsync(tmp, stream()));
  return BAMHIP_OK;
}
