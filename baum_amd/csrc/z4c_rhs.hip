// z4c_rhs.hip -- Z4c right-hand side on gfx950, FP64.
//
// Equations: src/projects/z4/z4_rhs_movpunc_N.m:217-467 (conformal
// connection and Ricci, conformal-factor terms, 1+log / Gamma-driver gauge,
// Kreiss-Oliger dissipation, optional addlinear update with trace removal and
// det normalisation).  Stencils: src/math/MathToC/discretize.m:18-46,97-135
// and src/main/amr/bam_amr_advection.h:28-155.  Box-edge rules (order 2 one
// point from the edge, shifted advection stencils within rstencil points, KO
// only two or more points away): src/main/amr/bam_amr_loops_box.h:186-215.
//
// Layout: SoA, one double array per grid function, ijk = i + m*j + m*n*k
// within a box, boxes concatenated at noffset.  One thread per grid point;
// a block is a 64 x 4 x 1 tile (one wave per x-row of 64 points, so every
// x-neighbour load is a coalesced 512 B transaction and y/z neighbours are
// served from L1/L2).  The algebra is ordered so that derivative groups are
// consumed as soon as they are formed (second derivatives of gt go straight
// into the Ricci accumulators, those of beta into the Gamma-driver terms),
// keeping the live FP64 state within the VGPR budget of one wave per SIMD.
//
// Two derivative providers share the algebra:
//   Bulk4 : points >= 3 away from every box edge, 4th-order centred
//           stencils, lop-sided (offset 1) 4th-order advection;
//   Edge  : any point, with the reference's per-point order selection,
//           used for the shell of points 1 and 2 away from the box edge and
//           for parameter choices other than (4, 4, lop-sided 1).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "bamhip_internal.h"

// Contraction of a*b + c into an FMA only within one source expression: the
// fused operations are then fixed by the source, not by per-kernel backend
// choices, so every instantiation of the tile kernel (tile shape, staging,
// register budget) rounds identically and gives bitwise the same RHS.
#pragma clang fp contract(on)

// compile-time loop: f(integral_constant<int, I>) for I = 0..N-1; used for
// loops whose body holds a barrier, which the compiler does not unroll, so
// that per-field arrays keep constant indices and stay in registers
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

namespace bamhip {

// u: the state the stencils read; r: output (the RHS, unew = upre + c RHS, or
// in an RK4 stage the next stage's state); p: upre, or the RK4 step's start
// state in a stage.  In an RK4 stage (read and written at the point only):
// a: the accumulator read, c: the accumulator written (c = a + wc RHS), cp:
// where the start state p is copied to (or unused).
struct Z4cPtrs {
  const double *u[BAMHIP_Z4C_NVARS];
  double *r[BAMHIP_Z4C_NVARS];
  const double *p[BAMHIP_Z4C_NVARS];
  double *c[BAMHIP_Z4C_NVARS];
  const double *a[BAMHIP_Z4C_NVARS];
  double *cp[BAMHIP_Z4C_NVARS];
};

struct Z4cGeo {
  int m, n, o;      // box points
  long noffset;
  int lo[3], hi[3]; // computed index range [lo, hi] per direction
  long sj, sk;      // strides (di = 1)
  double h[3], ih[3];
  double i12h[3], i12hh[3], i144hh[3];  // 4th order: 1/(12h), 1/(12h^2), mixed 1/(144 ha hb) (xy, xz, yz)
  double i2h[3], ihh[3], i4hh[3];       // 2nd order: 1/(2h), 1/h^2, mixed 1/(4 ha hb)
  int edge_only;    // Edge kernel: skip points that the bulk kernel computes
};

struct Z4cK {
  int order_centered, order_advection, lopsided, order_dissipation;
  double dissfactor, kappa1, kappa2, zeroThetaRHS;
  int subtractA, normalizedetg;
  double oploglapse, oploglapse2, oplogwithshift, harmoniclapse, lapseharmonicf;
  double gamma0factor, gamma2factor, withGadv, withShiftadv, withBadv, withB;
  double shiftalphapower, shiftgammacoeff, shiftdriver;
  double chiDivFloor, chipsipower;
  double c;
  int addlinear;
  // RK4 stage (rk.c:231-252) fused into the store: c = a + wc RHS, when
  // has_v r = p + wv RHS, when has_cp cp = p
  int rk, has_v, has_cp;
  double wc, wv;
};

// scheduling fence: keeps the machine scheduler from hoisting the loads of
// later derivative groups above earlier ones, which would keep hundreds of
// FP64 stencil values live at once and spill
#define BH_PHASE() __builtin_amdgcn_sched_barrier(0)

// index of component (a,b) in xx xy xz yy yz zz storage; closed form (not
// recursive) so that it folds to a constant once the loops are unrolled and
// every tensor array stays in registers
__device__ __forceinline__ constexpr int sy(int a, int b) {
  return (a < b ? a : b) * (5 - (a < b ? a : b)) / 2 + (a < b ? b : a);
}
__device__ __forceinline__ constexpr int mixid(int a, int b) {  // xy=0 xz=1 yz=2
  return (a + b) - 1;
}

// ---------------------------------------------------------------- providers
// lop-sided advection rows: stencil4[offset] of order 4 and the order-2 rows
// (bam_amr_advection.h:28-36,90-146), row index = offset
__constant__ double kAdvSt4[5][5] = {{25. / 12, -48. / 12, 36. / 12, -16. / 12, 3. / 12},
                                     {3. / 12, 10. / 12, -18. / 12, 6. / 12, -1. / 12},
                                     {-1. / 12, 8. / 12, 0. / 12, -8. / 12, 1. / 12},
                                     {1. / 12, -6. / 12, 18. / 12, -10. / 12, -3. / 12},
                                     {-3. / 12, 16. / 12, -36. / 12, 48. / 12, -25. / 12}};
__constant__ double kAdvSt2[3][3] = {{1.5, -2.0, 0.5}, {0.5, 0.0, -0.5}, {-0.5, 2.0, -1.5}};

// runtime-order provider following the reference per point
struct Edge {
  long q;
  long s[3];
  const Z4cGeo &G;
  int ordc, orda, twoormore;
  int pos[3], pmax[3];
  __device__ Edge(const Z4cGeo &g, long q_, int i, int j, int k, const Z4cK &K) : q(q_), G(g) {
    s[0] = 1;
    s[1] = g.sj;
    s[2] = g.sk;
    pos[0] = i;
    pos[1] = j;
    pos[2] = k;
    pmax[0] = g.m - 1;
    pmax[1] = g.n - 1;
    pmax[2] = g.o - 1;
    const int away1 = (i == 1 || j == 1 || k == 1 || i == pmax[0] - 1 || j == pmax[1] - 1 || k == pmax[2] - 1);
    twoormore = (i >= 2 && j >= 2 && k >= 2 && i <= pmax[0] - 2 && j <= pmax[1] - 2 && k <= pmax[2] - 2);
    ordc = (K.order_centered == 2 || away1) ? 2 : 4;
    orda = (K.order_advection == 2 || away1) ? 2 : 4;
  }
  __device__ double d1(const double *__restrict__ f, int a) const {
    const long t = s[a];
    if (ordc == 2) return (f[q + t] - f[q - t]) * G.i2h[a];
    return ((f[q - 2 * t] - f[q + 2 * t]) + 8.0 * (f[q + t] - f[q - t])) * G.i12h[a];
  }
  __device__ double d2(const double *__restrict__ f, int a) const {
    const long t = s[a];
    if (ordc == 2) return (f[q - t] + f[q + t] - 2.0 * f[q]) * G.ihh[a];
    return (16.0 * (f[q - t] + f[q + t]) - (f[q - 2 * t] + f[q + 2 * t]) - 30.0 * f[q]) * G.i12hh[a];
  }
  __device__ double dd(const double *__restrict__ f, int a, int b) const {
    if (a == b) return d2(f, a);
    const long t = s[a], w = s[b];
    if (ordc == 2) {
      return ((f[q + t + w] - f[q + t - w]) - (f[q - t + w] - f[q - t - w])) * G.i4hh[mixid(a, b)];
    }
    auto row = [&](long o) { return (f[o - 2 * w] - f[o + 2 * w]) + 8.0 * (f[o + w] - f[o - w]); };
    return ((row(q - 2 * t) - row(q + 2 * t)) + 8.0 * (row(q + t) - row(q - t))) * G.i144hh[mixid(a, b)];
  }
  // advection weights over offsets -4..4 (slot 4 = centre), rows of
  // set_advection4 / stencil4 (bam_amr_advection.h:28-155)
  double W[3][9];
  __device__ void set_adv(const double beta[3], const Z4cK &K) {
    for (int a = 0; a < 3; a++) {
      for (int k = 0; k < 9; k++) W[a][k] = 0.0;
      const double wx = beta[a] * G.ih[a];
      const int i = pos[a], imax = pmax[a];
      const int rst = (orda == 2) ? 2 : 4 - K.lopsided;
      int wsign, offset;
      if (wx < 0) {
        wsign = -1;
        offset = (i >= rst) ? 0 : rst - i;
      } else if (wx > 0) {
        wsign = 1;
        offset = (imax - i >= rst) ? 0 : rst - (imax - i);
      } else {
        wsign = 0;
        offset = 0;
      }
      if (wsign == 0) continue;
      if (orda == 2) {
        for (int k = 0; k < 3; k++) W[a][4 + wsign * (k - offset)] += -wsign * wx * kAdvSt2[offset][k];
      } else {
        offset += K.lopsided;
        for (int k = 0; k < 5; k++) W[a][4 + wsign * (k - offset)] += -wsign * wx * kAdvSt4[offset][k];
      }
    }
  }
  __device__ double adv(const double *__restrict__ f) const {
    double acc = 0;
    for (int a = 0; a < 3; a++)
      for (int k = 0; k < 9; k++)
        if (W[a][k] != 0.0) acc += W[a][k] * f[q + (k - 4) * s[a]];
    return acc;
  }
  __device__ bool ko_on(const Z4cK &K) const { return K.order_dissipation == 4 && twoormore; }
  __device__ double ko(const double *__restrict__ f) const {
    double acc = 0;
    for (int a = 0; a < 3; a++) {
      const long t = s[a];
      acc += ((f[q - 2 * t] + f[q + 2 * t]) - 4.0 * (f[q - t] + f[q + t]) + 6.0 * f[q]) * G.ih[a];
    }
    return -acc;
  }
};

// write one point: raw RHS, or upre + c*RHS with trace removal and det
// normalisation (z4_rhs_movpunc_N.m:446-467)
__device__ __forceinline__ void store_point(const Z4cPtrs &P, const Z4cK &K, long q, double *rhs) {
  if (K.rk) {
    // all loads before any store: the lists may alias (the accumulator read
    // and the next state written at the point, for one), and loads written
    // after a store would each wait for it
    double av[25], pv[25];
#pragma unroll
    for (int v = 0; v < 25; v++) av[v] = P.a[v][q];
    if (K.has_v || K.has_cp) {
#pragma unroll
      for (int v = 0; v < 25; v++) pv[v] = P.p[v][q];
    }
#pragma unroll
    for (int v = 0; v < 25; v++) P.c[v][q] = av[v] + K.wc * rhs[v];
    if (K.has_v) {
#pragma unroll
      for (int v = 0; v < 25; v++) P.r[v][q] = pv[v] + K.wv * rhs[v];
    }
    if (K.has_cp) {
#pragma unroll
      for (int v = 0; v < 25; v++) P.cp[v][q] = pv[v];
    }
    return;
  }
  if (K.addlinear) {
#pragma unroll
    for (int v = 0; v < 25; v++) rhs[v] = P.p[v][q] + K.c * rhs[v];
    double ng[6], nA[6];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      ng[c] = rhs[c];
      nA[c] = rhs[7 + c];
    }
    const double a0 = ng[3] * ng[5] - ng[4] * ng[4], a1 = ng[2] * ng[4] - ng[1] * ng[5];
    const double a2 = ng[1] * ng[4] - ng[2] * ng[3], a3 = ng[0] * ng[5] - ng[2] * ng[2];
    const double a4 = ng[1] * ng[2] - ng[0] * ng[4], a5 = ng[0] * ng[3] - ng[1] * ng[1];
    const double detnginv = 1.0 / (ng[0] * a0 + ng[1] * a1 + ng[2] * a2);
    if (K.subtractA) {
      const double trA =
          detnginv * (a0 * nA[0] + a3 * nA[3] + a5 * nA[5] + 2.0 * (a1 * nA[1] + a2 * nA[2] + a4 * nA[4]));
#pragma unroll
      for (int c = 0; c < 6; c++) rhs[7 + c] = nA[c] - (trA / 3.0) * ng[c];
    }
    if (K.normalizedetg) {
      const double aux = pow(detnginv, 1.0 / 3.0);
#pragma unroll
      for (int c = 0; c < 6; c++) rhs[c] = aux * ng[c];
    }
  }
#pragma unroll
  for (int v = 0; v < 25; v++) P.r[v][q] = rhs[v];
}

// ------------------------------------------------------------------ algebra
template <class D>
__device__ __forceinline__ bool z4c_point(D &dv, const Z4cPtrs &P, const Z4cK &K, long q) {
  const double *const *u = P.u;
  // point values (list order of z4.c:234-242)
  double g[6], A[6], Gm[3], beta[3], Bv[3];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    g[c] = u[c][q];
    A[c] = u[7 + c][q];
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    Gm[a] = u[14 + a][q];
    beta[a] = u[19 + a][q];
    Bv[a] = u[22 + a][q];
  }
  const double chi = u[6][q], Khat = u[13][q], Theta = u[17][q], alpha = u[18][q];

  // advection terms (lie*, adv*) of all 25 fields, weighted by their
  // coefficients in the right-hand sides
  dv.set_adv(beta, K);
  double rhs[25];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    rhs[c] = dv.adv(u[c]);
    rhs[7 + c] = dv.adv(u[7 + c]);
  }
  rhs[6] = dv.adv(u[6]);
  rhs[13] = dv.adv(u[13]);
  double advG[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    advG[a] = dv.adv(u[14 + a]);
    rhs[19 + a] = K.withShiftadv * dv.adv(u[19 + a]);
    rhs[22 + a] = K.withBadv * dv.adv(u[22 + a]);
  }
  rhs[17] = dv.adv(u[17]);
  rhs[18] = K.oplogwithshift * dv.adv(u[18]);

  BH_PHASE();
  // inverse conformal metric
  double ig[6];
  ig[0] = g[3] * g[5] - g[4] * g[4];
  ig[1] = g[2] * g[4] - g[1] * g[5];
  ig[2] = g[1] * g[4] - g[2] * g[3];
  ig[3] = g[0] * g[5] - g[2] * g[2];
  ig[4] = g[1] * g[2] - g[0] * g[4];
  ig[5] = g[0] * g[3] - g[1] * g[1];
  const double detginv = 1.0 / (g[0] * ig[0] + g[1] * ig[1] + g[2] * ig[2]);
#pragma unroll
  for (int c = 0; c < 6; c++) ig[c] *= detginv;

  BH_PHASE();
  // Christoffels: lower Gd[c][ab] = Gamma_{c ab}, upper Gu[c][ab] = Gamma^c_ab
  double Gd[3][6];
  {
    double dg[3][6];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int ab = 0; ab < 6; ab++) dg[c][ab] = dv.d1(u[ab], c);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++)
          Gd[c][sy(a, b)] = 0.5 * (dg[a][sy(b, c)] + dg[b][sy(a, c)] - dg[c][sy(a, b)]);
  }
  double Gu[3][6];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int ab = 0; ab < 6; ab++)
      Gu[c][ab] = ig[sy(c, 0)] * Gd[0][ab] + ig[sy(c, 1)] * Gd[1][ab] + ig[sy(c, 2)] * Gd[2][ab];
  double Gfg[3];
#pragma unroll
  for (int a = 0; a < 3; a++)
    Gfg[a] = ig[0] * Gu[a][0] + ig[3] * Gu[a][3] + ig[5] * Gu[a][5] +
             2.0 * (ig[1] * Gu[a][1] + ig[2] * Gu[a][2] + ig[4] * Gu[a][4]);

  BH_PHASE();
  // Ricci of the conformal metric
  double R[6];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = a; b < 3; b++) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
          double t = 0.0;
#pragma unroll
          for (int e = 0; e < 3; e++)
            t += Gu[e][sy(c, a)] * Gd[b][sy(e, d)] + Gu[e][sy(c, b)] * Gd[a][sy(e, d)] +
                 Gu[e][sy(a, d)] * Gd[e][sy(c, b)];
          s += ig[sy(c, d)] * t;
        }
      double t2 = 0.0;
#pragma unroll
      for (int c = 0; c < 3; c++) t2 += Gfg[c] * (Gd[a][sy(b, c)] + Gd[b][sy(a, c)]);
      R[sy(a, b)] = s + 0.5 * t2;
    }
  BH_PHASE();
  // - 1/2 g^cd d_c d_d g_ab, consumed component by component
#pragma unroll
  for (int ab = 0; ab < 6; ab++) {
    const double *f = u[ab];
    const double lap = ig[0] * dv.d2(f, 0) + ig[3] * dv.d2(f, 1) + ig[5] * dv.d2(f, 2) +
                       2.0 * (ig[1] * dv.dd(f, 0, 1) + ig[2] * dv.dd(f, 0, 2) + ig[4] * dv.dd(f, 1, 2));
    R[ab] -= 0.5 * lap;
    BH_PHASE();
  }
  BH_PHASE();
  // + 1/2 (g_ca d_b G^c + g_cb d_a G^c)
  {
    double dG[3][3];  // dG[b][c] = d_b G^c
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int c = 0; c < 3; c++) dG[b][c] = dv.d1(u[14 + c], b);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) {
        double t = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) t += g[sy(c, a)] * dG[b][c] + g[sy(c, b)] * dG[a][c];
        R[sy(a, b)] += 0.5 * t;
      }
  }

  BH_PHASE();
  // conformal factor
  const double chig = (chi < K.chiDivFloor) ? K.chiDivFloor : chi;
  const double oocp = 1.0 / K.chipsipower;
  const double psim4 = exp(-4.0 * oocp * log(chig));
  const double oochig = 1.0 / chig;
  double dchi[3], df[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    dchi[a] = dv.d1(u[6], a);
    df[a] = oocp * dchi[a] * oochig;
  }
  double cddf[6];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = a; b < 3; b++) {
      const int ab = sy(a, b);
      cddf[ab] = oocp * dv.dd(u[6], a, b) * oochig - K.chipsipower * df[a] * df[b] -
                 (Gu[0][ab] * df[0] + Gu[1][ab] * df[1] + Gu[2][ab] * df[2]);
    }
  auto trace = [&](const double *t6) {
    return ig[0] * t6[0] + ig[3] * t6[3] + ig[5] * t6[5] + 2.0 * (ig[1] * t6[1] + ig[2] * t6[2] + ig[4] * t6[4]);
  };
  auto contract = [&](const double *x, const double *y) {
    return ig[0] * x[0] * y[0] + ig[3] * x[1] * y[1] + ig[5] * x[2] * y[2] +
           ig[1] * (x[0] * y[1] + x[1] * y[0]) + ig[2] * (x[0] * y[2] + x[2] * y[0]) +
           ig[4] * (x[1] * y[2] + x[2] * y[1]);
  };
  const double trcddf = trace(cddf);
  const double dfdf = contract(df, df);
  // R + Rphi
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = a; b < 3; b++) {
      const int ab = sy(a, b);
      R[ab] += 4.0 * df[a] * df[b] - 2.0 * cddf[ab] - 2.0 * g[ab] * trcddf - 4.0 * g[ab] * dfdf;
    }

  BH_PHASE();
  // lapse derivatives
  double da[3];
#pragma unroll
  for (int a = 0; a < 3; a++) da[a] = dv.d1(u[18], a);
  const double dfda = contract(df, da);
  double cdda[6];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = a; b < 3; b++) {
      const int ab = sy(a, b);
      cdda[ab] = dv.dd(u[18], a, b) - (Gu[0][ab] * da[0] + Gu[1][ab] * da[1] + Gu[2][ab] * da[2]) -
                 2.0 * (df[a] * da[b] + df[b] * da[a] - g[ab] * dfda);
    }
  const double trcdda = psim4 * trace(cdda);

  BH_PHASE();
  // extrinsic-curvature terms
  double Ainv[6], AA[6];
  {
    double Au[3][3];  // Au[a][b] = g^ac A_cb (mixed)
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++)
        Au[a][b] = ig[sy(a, 0)] * A[sy(0, b)] + ig[sy(a, 1)] * A[sy(1, b)] + ig[sy(a, 2)] * A[sy(2, b)];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) {
        Ainv[sy(a, b)] = Au[a][0] * ig[sy(0, b)] + Au[a][1] * ig[sy(1, b)] + Au[a][2] * ig[sy(2, b)];
        AA[sy(a, b)] = A[sy(a, 0)] * Au[0][b] + A[sy(a, 1)] * Au[1][b] + A[sy(a, 2)] * Au[2][b];
      }
  }
  const double cAA = trace(AA);
  const double K_ = Khat + 2.0 * Theta;
  double dKT[3];  // 2 dKhat + dTheta
#pragma unroll
  for (int a = 0; a < 3; a++) dKT[a] = 2.0 * dv.d1(u[13], a) + dv.d1(u[17], a);
  double divAinv[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double s = 0.0;
#pragma unroll
    for (int bc = 0; bc < 6; bc++) s += ((bc == 0 || bc == 3 || bc == 5) ? 1.0 : 2.0) * Gu[a][bc] * Ainv[bc];
    s -= 1.5 * oochig * (Ainv[sy(a, 0)] * dchi[0] + Ainv[sy(a, 1)] * dchi[1] + Ainv[sy(a, 2)] * dchi[2]);
    s -= (1.0 / 3.0) * (ig[sy(a, 0)] * dKT[0] + ig[sy(a, 1)] * dKT[1] + ig[sy(a, 2)] * dKT[2]);
    divAinv[a] = s;
  }
  const double Rhat = psim4 * trace(R);   // R holds R + Rphi
  const double Hhat = Rhat + 2.0 * K_ * K_ / 3.0 - cAA;

  BH_PHASE();
  // shift terms
  double db[3][3];  // db[a][b] = d_a beta^b
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) db[a][b] = dv.d1(u[19 + b], a);
  const double divbeta = db[0][0] + db[1][1] + db[2][2];
  const double tdb = (2.0 / 3.0) * divbeta;
  double pG[3];
  {
    // g^bc d_b d_c beta^a + g^ab (1/3) d_b d_c beta^c
    double ddbc[3][6];
#pragma unroll
    for (int cc = 0; cc < 3; cc++)
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++) ddbc[cc][sy(a, b)] = dv.dd(u[19 + cc], a, b);
    double ootd[3];
#pragma unroll
    for (int a = 0; a < 3; a++) ootd[a] = (1.0 / 3.0) * (ddbc[0][sy(a, 0)] + ddbc[1][sy(a, 1)] + ddbc[2][sy(a, 2)]);
#pragma unroll
    for (int a = 0; a < 3; a++)
      pG[a] = trace(ddbc[a]) + ig[sy(a, 0)] * ootd[0] + ig[sy(a, 1)] * ootd[1] + ig[sy(a, 2)] * ootd[2] -
              (Gfg[0] * db[0][a] + Gfg[1] * db[1][a] + Gfg[2] * db[2][a]) + Gfg[a] * tdb + advG[a];
  }

  BH_PHASE();
  // right-hand sides
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = a; b < 3; b++) {
      const int ab = sy(a, b);
      double lg = -g[ab] * tdb, lA = -A[ab] * tdb;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        lg += db[a][c] * g[sy(b, c)] + db[b][c] * g[sy(a, c)];
        lA += db[a][c] * A[sy(b, c)] + db[b][c] * A[sy(a, c)];
      }
      rhs[ab] += -2.0 * alpha * A[ab] + lg;
      rhs[7 + ab] += psim4 * (-cdda[ab] + alpha * R[ab]) - (1.0 / 3.0) * g[ab] * (-trcdda + alpha * Rhat) +
                     alpha * (K_ * A[ab] - 2.0 * AA[ab]) + lA;
    }
  double rG[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    rG[a] = -2.0 * (Ainv[sy(a, 0)] * da[0] + Ainv[sy(a, 1)] * da[1] + Ainv[sy(a, 2)] * da[2]) +
            2.0 * alpha * divAinv[a] - 2.0 * alpha * K.kappa1 * (Gm[a] - Gfg[a]) + pG[a];
    rhs[14 + a] = rG[a];
  }
  rhs[13] += -trcdda + alpha * (cAA + K_ * K_ / 3.0) + K.kappa1 * (1.0 - K.kappa2) * alpha * Theta;
  rhs[6] += (K.chipsipower / 6.0) * chig * divbeta - (K.chipsipower / 6.0) * chig * alpha * K_;
  rhs[17] = (alpha * (0.5 * Hhat - (2.0 + K.kappa2) * K.kappa1 * Theta) + rhs[17]) * K.zeroThetaRHS;

  BH_PHASE();
  // gauge
  const double falpha =
      K.oploglapse * K.lapseharmonicf + K.oploglapse2 * (8.0 / 3.0) / (3.0 - alpha) + K.harmoniclapse * alpha;
  rhs[18] -= falpha * alpha * Khat;
  if (K.withB != 0.0) {
    const double betaF = K.shiftgammacoeff * pow(alpha, K.shiftalphapower);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      rhs[19 + a] += (K.gamma2factor + K.gamma0factor * betaF) * Bv[a];
      rhs[22 + a] += (K.gamma0factor + K.gamma2factor * betaF) * (rG[a] - K.withGadv * advG[a]) -
                     K.shiftdriver * Bv[a];
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      rhs[22 + a] = 0.0;
      rhs[19 + a] += Gm[a] - K.shiftdriver * beta[a];
    }
  }

  BH_PHASE();
  // Kreiss-Oliger dissipation
  if (dv.ko_on(K)) {
#pragma unroll
    for (int v = 0; v < 25; v++) rhs[v] += K.dissfactor * dv.ko(u[v]);
  }

  bool finite = true;
#pragma unroll
  for (int v = 0; v < 25; v++) finite = finite && isfinite(rhs[v]);

  store_point(P, K, q, rhs);
  return finite;
}

// ------------------------------------------------------- bulk kernel (LDS)
// Output tile of TX x TY x TZ points per block (256 threads, one point per
// thread).  Each of the 25 grid functions is staged in turn into an LDS tile
// with a halo of 3 points (the reach of the lop-sided advection stencil), so
// every stencil neighbour is an immediate-offset ds_read_b64 from one
// per-thread LDS address.  The algebra is ordered by staged field: each
// field's derivatives, advection and dissipation are consumed as soon as the
// field is in LDS.  The bulk uses 16 x 4 x 4 tiles (long x rows for
// coalesced loads); the box-edge shell, two points thick, uses slabs of
// 2 x 16 x 8, 16 x 2 x 8 and 16 x 8 x 2 on the x, y and z faces.
constexpr int HALO = 3;
constexpr int NTHR = 256;
// PX: unused LDS columns at the end of each tile row (the row stride SX is
// the staged width WX plus PX), chosen so that the 16 lanes of a read phase
// fall on distinct banks
// ORD: which point of the tile thread t of the block computes (64 threads
// form a wave): 0, x fastest, then y, then z; 1, y fastest, then z, then x
// (a wave keeps one x, as the layers of an x-face slab do); 2, x fastest,
// then z, then y (a wave keeps one y).
template <int TX_, int TY_, int TZ_, int PX = 0, int ORD = 0>
struct Lay {
  static constexpr int TX = TX_, TY = TY_, TZ = TZ_;
  static constexpr int WX = TX + 2 * HALO;
  static constexpr int SX = WX + PX, SY = TY + 2 * HALO, SZ = TZ + 2 * HALO;
  static constexpr int TILE = SX * SY * SZ;
  static constexpr int NLOAD = (TILE + NTHR - 1) / NTHR;
  static_assert(TX * TY * TZ == NTHR, "one point per thread");
  static_assert(ORD >= 0 && ORD <= 2, "thread order");
  __device__ static void coords(int t, int &x, int &y, int &z) {
    if constexpr (ORD == 1) {
      y = t % TY;
      z = (t / TY) % TZ;
      x = t / (TY * TZ);
    } else if constexpr (ORD == 2) {
      x = t % TX;
      z = (t / TX) % TZ;
      y = t / (TX * TZ);
    } else {
      x = t % TX;
      y = (t / TX) % TY;
      z = t / (TX * TY);
    }
  }
};
using LBulk = Lay<16, 4, 4>;
// x faces: rows of 2 points, so without padding (stride 8 doubles) the
// lanes (x, y) = (0..1, 0..7) of a read phase meet on 4 banks; a stride of 10
// spreads them over all 32
using LSlabX = Lay<2, 16, 8, 2>;
using LSlabY = Lay<16, 2, 8>;
using LSlabZ = Lay<16, 8, 2>;
using LLayX = Lay<1, 16, 16>;
using LLayY = Lay<16, 1, 16>;
using LLayZ = Lay<16, 16, 1>;

// a thread's tile corner in the LDS tile of layout L (its point is at +HALO)
template <class L>
struct TP {
  const double *p;
};

template <int DX, int DY, int DZ, class L>
__device__ __forceinline__ double at(TP<L> T) {
  return T.p[(DX + HALO) + L::SX * ((DY + HALO) + L::SY * (DZ + HALO))];
}
// neighbour at offset o along direction A
template <int A, int O, class L>
__device__ __forceinline__ double ax(TP<L> T) {
  return at<(A == 0 ? O : 0), (A == 1 ? O : 0), (A == 2 ? O : 0)>(T);
}
template <int A, class L>
__device__ __forceinline__ double t_d1(TP<L> T, const Z4cGeo &G) {
  return ((ax<A, -2>(T) - ax<A, 2>(T)) + 8.0 * (ax<A, 1>(T) - ax<A, -1>(T))) * G.i12h[A];
}
template <int A, class L>
__device__ __forceinline__ double t_d2(TP<L> T, const Z4cGeo &G) {
  return (16.0 * (ax<A, -1>(T) + ax<A, 1>(T)) - (ax<A, -2>(T) + ax<A, 2>(T)) - 30.0 * ax<A, 0>(T)) * G.i12hh[A];
}
// mixed derivative d_A d_B (A < B): tensor product of 4th-order first differences
template <int A, int B, int OA, class L>
__device__ __forceinline__ double t_row(TP<L> T) {
  constexpr int X0 = (A == 0 ? OA : 0), Y0 = (A == 1 ? OA : 0), Z0 = (A == 2 ? OA : 0);
  constexpr int BX = (B == 0), BY = (B == 1), BZ = (B == 2);
  return (at<X0 - 2 * BX, Y0 - 2 * BY, Z0 - 2 * BZ>(T) - at<X0 + 2 * BX, Y0 + 2 * BY, Z0 + 2 * BZ>(T)) +
         8.0 * (at<X0 + BX, Y0 + BY, Z0 + BZ>(T) - at<X0 - BX, Y0 - BY, Z0 - BZ>(T));
}
template <int A, int B, class L>
__device__ __forceinline__ double t_dd(TP<L> T, const Z4cGeo &G) {
  return ((t_row<A, B, -2>(T) - t_row<A, B, 2>(T)) + 8.0 * (t_row<A, B, 1>(T) - t_row<A, B, -1>(T))) *
         G.i144hh[A + B - 1];
}
// lop-sided advection: pw = max(beta/h,0)/12, mw = min(beta/h,0)/12 per direction
template <int A, class L>
__device__ __forceinline__ double t_adv1(TP<L> T, double pw, double mw) {
  return pw * ((18.0 * ax<A, 1>(T) - 3.0 * ax<A, -1>(T)) + (ax<A, 3>(T) - 6.0 * ax<A, 2>(T)) -
               10.0 * ax<A, 0>(T)) +
         mw * ((6.0 * ax<A, -2>(T) - ax<A, -3>(T)) + (3.0 * ax<A, 1>(T) - 18.0 * ax<A, -1>(T)) +
               10.0 * ax<A, 0>(T));
}
template <class L>
__device__ __forceinline__ double t_adv(TP<L> T, const double pw[3], const double mw[3]) {
  return t_adv1<0>(T, pw[0], mw[0]) + t_adv1<1>(T, pw[1], mw[1]) + t_adv1<2>(T, pw[2], mw[2]);
}
template <int A, class L>
__device__ __forceinline__ double t_ko1(TP<L> T) {
  return (ax<A, -2>(T) + ax<A, 2>(T)) - 4.0 * (ax<A, -1>(T) + ax<A, 1>(T)) + 6.0 * ax<A, 0>(T);
}
template <class L>
__device__ __forceinline__ double t_ko(TP<L> T, const Z4cGeo &G) {
  return -(t_ko1<0>(T) * G.ih[0] + t_ko1<1>(T) * G.ih[1] + t_ko1<2>(T) * G.ih[2]);
}
// all six second derivatives (xx xy xz yy yz zz)
template <class L>
__device__ __forceinline__ void t_hess(TP<L> T, const Z4cGeo &G, double d[6]) {
  d[0] = t_d2<0>(T, G);
  d[1] = t_dd<0, 1>(T, G);
  d[2] = t_dd<0, 2>(T, G);
  d[3] = t_d2<1>(T, G);
  d[4] = t_dd<1, 2>(T, G);
  d[5] = t_d2<2>(T, G);
}

// ---- stencils of the box-edge shell (bam_amr_loops_box.h:186-215,
// bam_amr_advection.h:90-146): a point one away from an edge uses order-2
// centred differences and order-2 advection rows, dissipation needs two or
// more points to every edge, and the advection row shifts away from an edge
// closer than rstencil.  The shell instantiation (SH = true) of the tile
// kernel selects these per point.  Every stencil entry with a nonzero
// coefficient lies inside the box, hence inside the tile's 3-point halo.
// The advection weights, first-derivative and dissipation coefficients are
// set once per point, so every point reads the same neighbours at fixed LDS
// offsets; a neighbour outside the box is read with coefficient 0 from a
// staging slot that holds box data.  Only the second derivatives branch on
// the point's order.
struct Sten {
  int o2;               // order-2 stencils (one point from an edge)
  int ko;               // dissipation on (two or more points from every edge)
  int offp[3], offm[3]; // advection row for beta > 0 / beta < 0 per direction
  // advection weights of the point's rows, as the reference forms them
  // (W = -sign(w) w stencil[offset][k] at offset sign(w) (k - offset),
  // bam_amr_advection.h:90-146): aw[a][o + 3] at offset o = -3..3 along a
  // (the centre entries unused), awc the sum of the three centre weights
  double aw[3][7];
  double awc;
  // first derivative along a: c1[a] (f(+1) - f(-1)) + c2[a] (f(-2) - f(+2)),
  // 1/(2h), 0 at order 2 and 8/(12h), 1/(12h) at order 4; kof: 1 where
  // dissipation is on, else 0
  double c1[3], c2[3];
  double kof;
};

// The point's advection weights along A (w = beta^A / h_A), set once per
// thread: the stages then read the 19 neighbours at fixed LDS offsets.
template <int A>
__device__ __forceinline__ void adv_weights(Sten &S, double w) {
  const int sg = w > 0.0 ? 1 : -1;
  const int off = w > 0.0 ? S.offp[A] : S.offm[A];
#pragma unroll
  for (int o = -3; o <= 3; o++) {
    const int k = off + sg * o;  // row entry k lands on offset sg (k - off) = o
    double c = 0.0;
    if (S.o2) {
      if (k >= 0 && k < 3) c = kAdvSt2[off][k];
    } else {
      if (k >= 0 && k < 5) c = kAdvSt4[off][k];
    }
    S.aw[A][o + 3] = (w == 0.0) ? 0.0 : -sg * w * c;
  }
}

template <bool SH, int A, class L>
__device__ __forceinline__ double s_d1(TP<L> T, const Z4cGeo &G, const Sten &S) {
  if constexpr (SH) return S.c2[A] * (ax<A, -2>(T) - ax<A, 2>(T)) + S.c1[A] * (ax<A, 1>(T) - ax<A, -1>(T));
  return t_d1<A>(T, G);
}
template <bool SH, class L>
__device__ __forceinline__ void s_grad(TP<L> T, const Z4cGeo &G, const Sten &S, double d[3]) {
  d[0] = s_d1<SH, 0>(T, G, S);
  d[1] = s_d1<SH, 1>(T, G, S);
  d[2] = s_d1<SH, 2>(T, G, S);
}
template <int A, int B, class L>
__device__ __forceinline__ double t_dd2(TP<L> T, const Z4cGeo &G) {
  constexpr int AX = (A == 0), AY = (A == 1), AZ = (A == 2), BX = (B == 0), BY = (B == 1), BZ = (B == 2);
  return ((at<AX + BX, AY + BY, AZ + BZ>(T) - at<AX - BX, AY - BY, AZ - BZ>(T)) -
          (at<-AX + BX, -AY + BY, -AZ + BZ>(T) - at<-AX - BX, -AY - BY, -AZ - BZ>(T))) *
         G.i4hh[A + B - 1];
}
template <bool SH, class L>
__device__ __forceinline__ void s_hess(TP<L> T, const Z4cGeo &G, const Sten &S, double d[6]) {
  if (SH && S.o2) {
    d[0] = (ax<0, -1>(T) + ax<0, 1>(T) - 2.0 * ax<0, 0>(T)) * G.ihh[0];
    d[1] = t_dd2<0, 1>(T, G);
    d[2] = t_dd2<0, 2>(T, G);
    d[3] = (ax<1, -1>(T) + ax<1, 1>(T) - 2.0 * ax<1, 0>(T)) * G.ihh[1];
    d[4] = t_dd2<1, 2>(T, G);
    d[5] = (ax<2, -1>(T) + ax<2, 1>(T) - 2.0 * ax<2, 0>(T)) * G.ihh[2];
    return;
  }
  t_hess(T, G, d);
}
// advection with the point's weights (shell) or the lop-sided rows (bulk)
template <int A, class L>
__device__ __forceinline__ double s_advw1(TP<L> T, const double w[7]) {
  return w[0] * ax<A, -3>(T) + w[1] * ax<A, -2>(T) + w[2] * ax<A, -1>(T) + w[4] * ax<A, 1>(T) +
         w[5] * ax<A, 2>(T) + w[6] * ax<A, 3>(T);
}
template <bool SH, class L>
__device__ __forceinline__ double s_adv(TP<L> T, const double pw[3], const double mw[3], const Sten &S) {
  if constexpr (SH)
    return S.awc * at<0, 0, 0>(T) + s_advw1<0>(T, S.aw[0]) + s_advw1<1>(T, S.aw[1]) + s_advw1<2>(T, S.aw[2]);
  return t_adv(T, pw, mw);
}
template <bool SH, class L>
__device__ __forceinline__ double s_ko(TP<L> T, const Z4cGeo &G, const Sten &S) {
  if constexpr (SH) return S.kof * t_ko(T, G);
  return t_ko(T, G);
}

// Staging of one field into the LDS tile, software-pipelined: the global
// loads of the next field are issued into registers (NLOAD per thread, all
// in flight together) before the current field is computed, and written to
// LDS between two barriers at the start of the next stage.
// staging order of the 25 fields: gt_ij, Gamma^i, chi, alpha, Khat, Theta,
// At_ij, beta^i, B^i
__device__ constexpr int STAGE_ORDER[25] = {0,  1,  2,  3,  4,  5,  14, 15, 16, 6,  18, 13, 17,
                                            7,  8,  9,  10, 11, 12, 19, 20, 21, 22, 23, 24};

template <class L>
struct Stager {
  static constexpr int NLOAD = L::NLOAD;
  int goff[NLOAD];
  unsigned ok;      // bit k: slot k lies inside the tile and the box
  long cbase;       // linear index of the tile corner (i0-3, j0-3, k0-3); may lie before the box
  double pf[NLOAD];
  __device__ Stager(const Z4cGeo &G, int i0, int j0, int k0, int tid) {
    cbase = G.noffset + (i0 - HALO) + G.sj * (j0 - HALO) + G.sk * (k0 - HALO);
    ok = 0;
#pragma unroll
    for (int k = 0; k < NLOAD; k++) {
      const int idx = tid + k * NTHR;
      const int x = idx % L::SX, y = (idx / L::SX) % L::SY, z = idx / (L::SX * L::SY);
      const int gi = i0 - HALO + x, gj = j0 - HALO + y, gk = k0 - HALO + z;
      const bool in = idx < L::TILE && x < L::WX && gi >= 0 && gj >= 0 && gk >= 0 && gi < G.m && gj < G.n && gk < G.o;
      goff[k] = in ? (int)(x + G.sj * y + G.sk * z) : 0;
      ok |= in ? (1u << k) : 0u;
    }
  }
  // slots outside the box are staged as 0; no stencil of a computed point reads them
  __device__ __forceinline__ void load(const double *__restrict__ f) {
#pragma unroll
    for (int k = 0; k < NLOAD; k++) pf[k] = (ok >> k & 1u) ? f[cbase + goff[k]] : 0.0;
  }
  __device__ __forceinline__ void commit(double *tile, int tid) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NLOAD; k++)
      if (tid + k * NTHR < L::TILE) tile[tid + k * NTHR] = pf[k];
    __syncthreads();
  }
};

// Staging through the LDS-DMA path: global_load_lds moves each field
// straight from HBM into one of NB LDS buffers (no registers, no ds_write), so
// NB - 1 fields are in flight while one is computed.  A wave's 64 lanes write
// 64 consecutive units of LDS (lane l -> unit l), so the tile is cut into
// chunks of 64 units, NPER per wave.  VEC = 1: a unit is one dword, the low
// (even unit) or high half of a tile element; VEC = 4: a unit is 16 bytes,
// the two tile elements (x, x + 1) with x even, which needs an even row
// length and 16-byte aligned pairs in HBM (checked by the host).  Slots of
// the tile outside the box load the nearest point of the box (a pair stays a
// pair), so that every staged value is box data: the shell's advection reads
// such slots with weight 0.
#define BH_LDS(p) ((__attribute__((address_space(3))) void *)(p))
template <class L, int NB, int VEC>
struct DmaStager {
  static_assert(VEC == 1 || VEC == 4, "dword or 16-byte units");
  static_assert(VEC == 1 || (L::SX % 2 == 0 && L::TX % 2 == 0), "16-byte units need even rows and tile steps");
  static constexpr int NW = NTHR / 64;
  static constexpr int NUNIT = VEC == 1 ? 2 * L::TILE : L::TILE / 2;
  static constexpr int NCH = (NUNIT + 63) / 64;
  static constexpr int NPER = (NCH + NW - 1) / NW;
  static constexpr int CHD = 64 * VEC / 2;      // doubles per chunk
  static constexpr int BUFD = NPER * NW * CHD;  // doubles per buffer, padded to whole chunks
  static_assert(NB >= 2 && (NB - 2) * NPER <= 63, "vmcnt range");
  unsigned boff[NPER];  // byte offset of each chunk's unit from the field's base
  int wave;
  __device__ DmaStager(const Z4cGeo &G, int i0, int j0, int k0, int tid) {
    const int lane = tid & 63;
    wave = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
    for (int c = 0; c < NPER; c++) {
      const int d = (wave * NPER + c) * 64 + lane;
      const int e = VEC == 1 ? d >> 1 : 2 * d;
      const int x = e % L::SX, y = (e / L::SX) % L::SY, z = e / (L::SX * L::SY);
      const int gi = i0 - HALO + x, gj = j0 - HALO + y, gk = k0 - HALO + z;
      const bool in = e < L::TILE && x < L::WX;
      // clamped into the box; for 16-byte units gi is even and so is the
      // box's row length, so the clamped pair is (0, 1) or (m - 2, m - 1)
      const int ci = min(max(gi, 0), G.m - (VEC == 1 ? 1 : 2)), cj = min(max(gj, 0), G.n - 1),
                ck = min(max(gk, 0), G.o - 1);
      const long el = in ? G.noffset + ci + G.sj * cj + G.sk * ck : G.noffset;
      boff[c] = (unsigned)(el * 8 + (VEC == 1 && in ? (d & 1) * 4 : 0));
    }
  }
  __device__ __forceinline__ void issue(const double *f, double *buf) const {
#pragma unroll
    for (int c = 0; c < NPER; c++) {
      const void *src = (const void *)((const char *)f + boff[c]);
      if constexpr (VEC == 1)
        __builtin_amdgcn_global_load_lds(src, BH_LDS(buf + (wave * NPER + c) * CHD), 4, 0, 0);
      else
        __builtin_amdgcn_global_load_lds(src, BH_LDS(buf + (wave * NPER + c) * CHD), 16, 0, 0);
    }
  }
};
// s_waitcnt vmcnt(n) lgkmcnt(0) (expcnt not waited), n a small constant
template <int N>
__device__ __forceinline__ void wait_vm_lgkm0() {
  static_assert(N >= 0 && N <= 63, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4));
}
template <int NPER>
__device__ __forceinline__ void wait_fields_after(int k) {
  switch (k) {
    case 0: wait_vm_lgkm0<0>(); break;
    case 1: wait_vm_lgkm0<NPER>(); break;
    case 2: wait_vm_lgkm0<(2 * NPER <= 63 ? 2 * NPER : 0)>(); break;
    default: wait_vm_lgkm0<(3 * NPER <= 63 ? 3 * NPER : 0)>(); break;
  }
}
// workgroup barrier without a memory fence (a fence would also wait for the
// LDS-DMA loads still in flight); the empty asm statements keep the compiler
// from moving memory accesses across it
__device__ __forceinline__ void bare_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Pin a value in registers at this point of the program: an empty volatile
// asm that reads and writes it.  Placed after each staged field, it keeps the
// compiler from sinking that field's arithmetic past the next stage's
// barrier, which would keep every staged tile's values live to the end of
// the kernel (and spill).
__device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }
template <int N>
__device__ __forceinline__ void pin(double (&a)[N]) {
#pragma unroll
  for (int i = 0; i < N; i++) pin(a[i]);
}

// Regions of one launch: index ranges [lo, hi] of boxes of a level (bulk
// regions, or slabs of the box-edge shell), each with its box geometry, and
// the first block of each region.  All boxes of a level are computed by one
// launch per tile layout.
// The grid spacing constants are common to the regions of one launch (one
// level), so a region carries only integer geometry.
constexpr int kMaxRegions = 32;
struct RegionGeo {
  int m, n, o;
  int lo[3], hi[3];
  long noffset, sj, sk;
};
struct Spacing {
  double h[3], ih[3], i12h[3], i12hh[3], i144hh[3], i2h[3], ihh[3], i4hh[3];
};
struct TileJobs {
  int n;
  int blk0[kMaxRegions + 1];  // first tile of each region; blk0[n] = number of tiles
  int xcd;                    // 1: XCD-aware numbering of the tiles (grid padded to a multiple of 8)
  RegionGeo reg[kMaxRegions];
  Spacing sp;
};
static_assert(sizeof(TileJobs) <= 3072, "kernel argument size");
static_assert(sizeof(Z4cPtrs) + sizeof(TileJobs) + sizeof(Z4cK) + sizeof(int *) <= 4096,
              "arguments of the tile kernel within 4 KiB");

// SH = false: every point of the regions is >= 3 from each edge (standard
// stencils).  SH = true: the regions are slabs of the box-edge shell; only
// points closer than 3 to an edge are computed, with the per-point rules of
// Sten.
template <class L, int NB, int VEC>
struct StageBuf {  // LDS doubles of the staging buffers
  static constexpr int n = NB * DmaStager<L, NB, VEC>::BUFD;
};
template <class L, int VEC>
struct StageBuf<L, 0, VEC> {
  static constexpr int n = L::TILE;
};

// NB = 0: one LDS tile, staged through registers (Stager); NB >= 2: NB LDS
// buffers filled by the LDS-DMA path (DmaStager).  OCC: blocks per CU the
// register allocation must allow (2 caps a thread at 256 registers, with
// spills, for two waves per SIMD).
template <class L, bool SH, int NB, int VEC, int OCC>
__global__ void __launch_bounds__(NTHR, OCC) z4c_rhs_tile(Z4cPtrs P, TileJobs J, Z4cK K, int *flag) {
  __shared__ double tile[StageBuf<L, NB, VEC>::n];
  // Tile of this block.  Workgroups are dispatched to the 8 XCDs round-robin
  // (workgroup w to XCD w mod 8), each XCD with its own L2.  With xcd set,
  // workgroup w takes tile (w mod 8) * per + w / 8, so every XCD works through
  // one contiguous range of tiles and neighbouring tiles, which stage the same
  // halo points, meet in the same L2.
  int t = (int)blockIdx.x;
  if (J.xcd) {
    const int per = (int)gridDim.x / 8;
    t = (t % 8) * per + t / 8;
    if (t >= J.blk0[J.n]) return;  // padding workgroup (whole block, before any barrier)
  }
  // region of this block (uniform): the regions' tiles are numbered in order
  int r = 0;
  while (r + 1 < J.n && t >= J.blk0[r + 1]) r++;
  // copies: indexing the argument array by r through a reference would stage it in scratch
  const RegionGeo Rg = J.reg[r];
  Z4cGeo G;
  G.m = Rg.m;
  G.n = Rg.n;
  G.o = Rg.o;
  G.noffset = Rg.noffset;
  G.sj = Rg.sj;
  G.sk = Rg.sk;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    G.lo[a] = Rg.lo[a];
    G.hi[a] = Rg.hi[a];
    G.h[a] = J.sp.h[a];
    G.ih[a] = J.sp.ih[a];
    G.i12h[a] = J.sp.i12h[a];
    G.i12hh[a] = J.sp.i12hh[a];
    G.i144hh[a] = J.sp.i144hh[a];
    G.i2h[a] = J.sp.i2h[a];
    G.ihh[a] = J.sp.ihh[a];
    G.i4hh[a] = J.sp.i4hh[a];
  }
  G.edge_only = 0;
  const int b = t - J.blk0[r];
  const int ntx = (G.hi[0] - G.lo[0] + L::TX) / L::TX, nty = (G.hi[1] - G.lo[1] + L::TY) / L::TY;
  const int tid = (int)threadIdx.x;
  int tx, ty, tz;
  L::coords(tid, tx, ty, tz);
  const int i0 = G.lo[0] + (b % ntx) * L::TX, j0 = G.lo[1] + ((b / ntx) % nty) * L::TY,
            k0 = G.lo[2] + (b / (ntx * nty)) * L::TZ;
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  Sten S{};
  bool active = (i <= G.hi[0] && j <= G.hi[1] && k <= G.hi[2]);
  if (SH) {
    // positions past hi (padding threads, never stored) are clamped to hi, so
    // that their rows and stencil reach stay within the tables and the halo
    const int pos[3] = {min(i, G.hi[0]), min(j, G.hi[1]), min(k, G.hi[2])};
    const int pmax[3] = {G.m - 1, G.n - 1, G.o - 1};
    int dmin = 1 << 30;
#pragma unroll
    for (int a = 0; a < 3; a++) dmin = min(dmin, min(pos[a], pmax[a] - pos[a]));
    active = active && dmin <= 2;
    S.o2 = dmin == 1;
    S.ko = dmin >= 2;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int dl = pos[a], dh = pmax[a] - pos[a];
      if (S.o2) {
        S.offm[a] = dl >= 2 ? 0 : 2 - dl;
        S.offp[a] = dh >= 2 ? 0 : 2 - dh;
      } else {
        S.offm[a] = (dl >= 3 ? 0 : 3 - dl) + 1;
        S.offp[a] = (dh >= 3 ? 0 : 3 - dh) + 1;
      }
    }
  }
  const long q = G.noffset + i + G.sj * j + G.sk * k;
  const int toff = tx + L::SX * (ty + L::SY * tz);
  TP<L> T{tile + toff};
  const double *const *u = P.u;

  // the first fields' transfers go out before the point values are loaded,
  // so that a block waits for one round trip to HBM at its start, not two
  using St = typename std::conditional<NB == 0, Stager<L>, DmaStager<L, (NB == 0 ? 2 : NB), VEC>>::type;
  St st(G, i0, j0, k0, tid);
  constexpr int NBUF = NB == 0 ? 1 : NB;
  constexpr int BUFD = StageBuf<L, NB, VEC>::n / NBUF;
  if constexpr (NB == 0) {
    st.load(u[STAGE_ORDER[0]]);
  } else {
#pragma unroll
    for (int p = 0; p < NB - 1; p++) st.issue(u[STAGE_ORDER[p]], tile + p * BUFD);
  }

  // point values; threads outside the box use flat data so that their
  // (discarded) arithmetic stays finite
  double g[6], A[6], Gm[3], beta[3], Bv[3];
  double chi = 1.0, Khat = 0.0, Theta = 0.0, alpha = 1.0;
  for (int c = 0; c < 6; c++) {
    g[c] = (c == 0 || c == 3 || c == 5) ? 1.0 : 0.0;
    A[c] = 0.0;
  }
  for (int a = 0; a < 3; a++) Gm[a] = beta[a] = Bv[a] = 0.0;
  if (active) {
#pragma unroll
    for (int c = 0; c < 6; c++) {
      g[c] = u[c][q];
      A[c] = u[7 + c][q];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      Gm[a] = u[14 + a][q];
      beta[a] = u[19 + a][q];
      Bv[a] = u[22 + a][q];
    }
    chi = u[6][q];
    Khat = u[13][q];
    Theta = u[17][q];
    alpha = u[18][q];
  }
  // stage s: make field s visible in LDS, start the transfer of a later field
  auto advance = [&](int sidx) {
    if constexpr (NB == 0) {
      st.commit(tile, tid);
      if (sidx + 1 < 25) st.load(u[STAGE_ORDER[sidx + 1]]);
    } else {
      // this wave's loads of field sidx are done once no more than the later
      // fields' loads are outstanding; the barrier makes every wave's done
      wait_fields_after<St::NPER>(min(NB - 2, 24 - sidx));
      bare_barrier();
      // the buffer of field sidx + NB - 1 held field sidx - 1, whose reads
      // every wave completed before the barrier
      if (sidx + NB - 1 < 25) st.issue(u[STAGE_ORDER[sidx + NB - 1]], tile + ((sidx + NB - 1) % NB) * BUFD);
      T.p = tile + (sidx % NB) * BUFD + toff;
    }
  };
  double pw[3], mw[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double w = beta[a] * G.ih[a];
    pw[a] = (w > 0.0 ? w : 0.0) * (1.0 / 12.0);
    mw[a] = (w < 0.0 ? w : 0.0) * (1.0 / 12.0);
  }
  if constexpr (SH) {
    adv_weights<0>(S, beta[0] * G.ih[0]);
    adv_weights<1>(S, beta[1] * G.ih[1]);
    adv_weights<2>(S, beta[2] * G.ih[2]);
    S.awc = S.aw[0][3] + S.aw[1][3] + S.aw[2][3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      S.c1[a] = S.o2 ? G.i2h[a] : 8.0 * G.i12h[a];
      S.c2[a] = S.o2 ? 0.0 : G.i12h[a];
    }
    S.kof = S.ko ? 1.0 : 0.0;
  }
  const bool ko = (K.order_dissipation == 4);
  const double eps = ko ? K.dissfactor : 0.0;

  // inverse conformal metric
  double ig[6];
  ig[0] = g[3] * g[5] - g[4] * g[4];
  ig[1] = g[2] * g[4] - g[1] * g[5];
  ig[2] = g[1] * g[4] - g[2] * g[3];
  ig[3] = g[0] * g[5] - g[2] * g[2];
  ig[4] = g[1] * g[2] - g[0] * g[4];
  ig[5] = g[0] * g[3] - g[1] * g[1];
  const double detginv = 1.0 / (g[0] * ig[0] + g[1] * ig[1] + g[2] * ig[2]);
#pragma unroll
  for (int c = 0; c < 6; c++) ig[c] *= detginv;
  auto trace = [&](const double *t6) {
    return ig[0] * t6[0] + ig[3] * t6[3] + ig[5] * t6[5] + 2.0 * (ig[1] * t6[1] + ig[2] * t6[2] + ig[4] * t6[4]);
  };
  auto contract = [&](const double *x, const double *y) {
    return ig[0] * x[0] * y[0] + ig[3] * x[1] * y[1] + ig[5] * x[2] * y[2] + ig[1] * (x[0] * y[1] + x[1] * y[0]) +
           ig[2] * (x[0] * y[2] + x[2] * y[0]) + ig[4] * (x[1] * y[2] + x[2] * y[1]);
  };

  double rhs[25];
  double R[6];  // Ricci accumulators (conformal + phi part)

  // ---- conformal metric: first derivatives, Laplacian part of Ricci
  double dg[3][6];
static_for<6>([&](auto I_) {
    constexpr int ab = decltype(I_)::value;
    advance(ab);
    dg[0][ab] = s_d1<SH, 0>(T, G, S);
    dg[1][ab] = s_d1<SH, 1>(T, G, S);
    dg[2][ab] = s_d1<SH, 2>(T, G, S);
    double h6[6];
    s_hess<SH>(T, G, S, h6);
    R[ab] = -0.5 * trace(h6);
    rhs[ab] = s_adv<SH>(T, pw, mw, S) + eps * s_ko<SH>(T, G, S);
    pin(dg[0][ab]);
    pin(dg[1][ab]);
    pin(dg[2][ab]);
    pin(R[ab]);
    pin(rhs[ab]);
  });
  // Christoffels
  double Gu[3][6], Gfg[3];
  {
    double Gd[3][6];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++)
          Gd[c][sy(a, b)] = 0.5 * (dg[a][sy(b, c)] + dg[b][sy(a, c)] - dg[c][sy(a, b)]);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int ab = 0; ab < 6; ab++)
        Gu[c][ab] = ig[sy(c, 0)] * Gd[0][ab] + ig[sy(c, 1)] * Gd[1][ab] + ig[sy(c, 2)] * Gd[2][ab];
#pragma unroll
    for (int a = 0; a < 3; a++)
      Gfg[a] = ig[0] * Gu[a][0] + ig[3] * Gu[a][3] + ig[5] * Gu[a][5] +
               2.0 * (ig[1] * Gu[a][1] + ig[2] * Gu[a][2] + ig[4] * Gu[a][4]);
    // quadratic Christoffel terms of R_ab
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int d = 0; d < 3; d++) {
            double t = 0.0;
#pragma unroll
            for (int e = 0; e < 3; e++)
              t += Gu[e][sy(c, a)] * Gd[b][sy(e, d)] + Gu[e][sy(c, b)] * Gd[a][sy(e, d)] +
                   Gu[e][sy(a, d)] * Gd[e][sy(c, b)];
            s += ig[sy(c, d)] * t;
          }
        double t2 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) t2 += Gfg[c] * (Gd[a][sy(b, c)] + Gd[b][sy(a, c)]);
        R[sy(a, b)] += s + 0.5 * t2;
      }
  }
  pin(R);
  pin(Gu[0]);
  pin(Gu[1]);
  pin(Gu[2]);
  pin(Gfg);

  // ---- Gamma^i: + 1/2 (g_ca d_b G^c + g_cb d_a G^c) in R_ab
  double advG[3], koG[3];
static_for<3>([&](auto I_) {
    constexpr int c = decltype(I_)::value;
    advance(6 + c);
    double dG[3];
    s_grad<SH>(T, G, S,dG);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) R[sy(a, b)] += 0.5 * (g[sy(c, a)] * dG[b] + g[sy(c, b)] * dG[a]);
    advG[c] = s_adv<SH>(T, pw, mw, S);
    koG[c] = eps * s_ko<SH>(T, G, S);
    pin(R);
    pin(advG[c]);
    pin(koG[c]);
  });

  // ---- conformal factor
  const double chig = (chi < K.chiDivFloor) ? K.chiDivFloor : chi;
  const double oocp = 1.0 / K.chipsipower;
  const double psim4 = exp(-4.0 * oocp * log(chig));
  const double oochig = 1.0 / chig;
  double dchi[3], df[3];
  {
    advance(9);
    s_grad<SH>(T, G, S,dchi);
    double h6[6];
    s_hess<SH>(T, G, S, h6);
    rhs[6] = s_adv<SH>(T, pw, mw, S) + eps * s_ko<SH>(T, G, S);
#pragma unroll
    for (int a = 0; a < 3; a++) df[a] = oocp * dchi[a] * oochig;
    double cddf[6];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) {
        const int ab = sy(a, b);
        cddf[ab] = oocp * h6[ab] * oochig - K.chipsipower * df[a] * df[b] -
                   (Gu[0][ab] * df[0] + Gu[1][ab] * df[1] + Gu[2][ab] * df[2]);
      }
    const double trcddf = trace(cddf);
    const double dfdf = contract(df, df);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) {
        const int ab = sy(a, b);
        R[ab] += 4.0 * df[a] * df[b] - 2.0 * cddf[ab] - 2.0 * g[ab] * trcddf - 4.0 * g[ab] * dfdf;
      }
    pin(R);
    pin(dchi);
    pin(df);
    pin(rhs[6]);
  }

  // ---- lapse
  double da[3], cdda[6];
  {
    advance(10);
    s_grad<SH>(T, G, S,da);
    double h6[6];
    s_hess<SH>(T, G, S, h6);
    rhs[18] = K.oplogwithshift * s_adv<SH>(T, pw, mw, S) + eps * s_ko<SH>(T, G, S);
    const double dfda = contract(df, da);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) {
        const int ab = sy(a, b);
        cdda[ab] = h6[ab] - (Gu[0][ab] * da[0] + Gu[1][ab] * da[1] + Gu[2][ab] * da[2]) -
                   2.0 * (df[a] * da[b] + df[b] * da[a] - g[ab] * dfda);
      }
    pin(da);
    pin(cdda);
    pin(rhs[18]);
  }
  const double trcdda = psim4 * trace(cdda);

  // ---- Khat and Theta: 2 dKhat + dTheta
  double dKT[3], koT;
  {
    advance(11);
    dKT[0] = 2.0 * s_d1<SH, 0>(T, G, S);
    dKT[1] = 2.0 * s_d1<SH, 1>(T, G, S);
    dKT[2] = 2.0 * s_d1<SH, 2>(T, G, S);
    rhs[13] = s_adv<SH>(T, pw, mw, S) + eps * s_ko<SH>(T, G, S);
    pin(dKT);
    pin(rhs[13]);
    advance(12);
    dKT[0] += s_d1<SH, 0>(T, G, S);
    dKT[1] += s_d1<SH, 1>(T, G, S);
    dKT[2] += s_d1<SH, 2>(T, G, S);
    rhs[17] = s_adv<SH>(T, pw, mw, S);  // KO of Theta is added after the zeroThetaRHS factor
    koT = eps * s_ko<SH>(T, G, S);
    pin(dKT);
    pin(rhs[17]);
    pin(koT);
  }

  // ---- trace-free extrinsic curvature: advection and dissipation only
static_for<6>([&](auto I_) {
    constexpr int ab = decltype(I_)::value;
    advance(13 + ab);
    rhs[7 + ab] = s_adv<SH>(T, pw, mw, S) + eps * s_ko<SH>(T, G, S);
    pin(rhs[7 + ab]);
  });

  // ---- shift: d_a beta^b, g^bc d_b d_c beta^a, (1/3) d_a d_c beta^c
  double db[3][3], pG[3], ootd[3];
static_for<3>([&](auto I_) {
    constexpr int c = decltype(I_)::value;
    advance(19 + c);
    db[0][c] = s_d1<SH, 0>(T, G, S);
    db[1][c] = s_d1<SH, 1>(T, G, S);
    db[2][c] = s_d1<SH, 2>(T, G, S);
    double h6[6];
    s_hess<SH>(T, G, S, h6);
    pG[c] = trace(h6);
#pragma unroll
    for (int a = 0; a < 3; a++) ootd[a] = (c == 0 ? 0.0 : ootd[a]) + (1.0 / 3.0) * h6[sy(a, c)];
    rhs[19 + c] = K.withShiftadv * s_adv<SH>(T, pw, mw, S) + eps * s_ko<SH>(T, G, S);
    pin(db[0][c]);
    pin(db[1][c]);
    pin(db[2][c]);
    pin(pG[c]);
    pin(ootd);
    pin(rhs[19 + c]);
  });
  // ---- B: advection and dissipation only
static_for<3>([&](auto I_) {
    constexpr int c = decltype(I_)::value;
    advance(22 + c);
    rhs[22 + c] = (K.withB != 0.0 ? K.withBadv * s_adv<SH>(T, pw, mw, S) : 0.0) + eps * s_ko<SH>(T, G, S);
    pin(rhs[22 + c]);
  });

  // ---- algebra
  double Ainv[6], AA[6];
  {
    double Au[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++)
        Au[a][b] = ig[sy(a, 0)] * A[sy(0, b)] + ig[sy(a, 1)] * A[sy(1, b)] + ig[sy(a, 2)] * A[sy(2, b)];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = a; b < 3; b++) {
        Ainv[sy(a, b)] = Au[a][0] * ig[sy(0, b)] + Au[a][1] * ig[sy(1, b)] + Au[a][2] * ig[sy(2, b)];
        AA[sy(a, b)] = A[sy(a, 0)] * Au[0][b] + A[sy(a, 1)] * Au[1][b] + A[sy(a, 2)] * Au[2][b];
      }
  }
  const double cAA = trace(AA);
  const double K_ = Khat + 2.0 * Theta;
  double divAinv[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double s = 0.0;
#pragma unroll
    for (int bc = 0; bc < 6; bc++) s += ((bc == 0 || bc == 3 || bc == 5) ? 1.0 : 2.0) * Gu[a][bc] * Ainv[bc];
    s -= 1.5 * oochig * (Ainv[sy(a, 0)] * dchi[0] + Ainv[sy(a, 1)] * dchi[1] + Ainv[sy(a, 2)] * dchi[2]);
    s -= (1.0 / 3.0) * (ig[sy(a, 0)] * dKT[0] + ig[sy(a, 1)] * dKT[1] + ig[sy(a, 2)] * dKT[2]);
    divAinv[a] = s;
  }
  const double Rhat = psim4 * trace(R);
  const double Hhat = Rhat + 2.0 * K_ * K_ / 3.0 - cAA;
  const double divbeta = db[0][0] + db[1][1] + db[2][2];
  const double tdb = (2.0 / 3.0) * divbeta;

#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = a; b < 3; b++) {
      const int ab = sy(a, b);
      double lg = -g[ab] * tdb, lA = -A[ab] * tdb;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        lg += db[a][c] * g[sy(b, c)] + db[b][c] * g[sy(a, c)];
        lA += db[a][c] * A[sy(b, c)] + db[b][c] * A[sy(a, c)];
      }
      rhs[ab] += -2.0 * alpha * A[ab] + lg;
      rhs[7 + ab] += psim4 * (-cdda[ab] + alpha * R[ab]) - (1.0 / 3.0) * g[ab] * (-trcdda + alpha * Rhat) +
                     alpha * (K_ * A[ab] - 2.0 * AA[ab]) + lA;
    }
  double rG[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double pseudolie = pG[a] + ig[sy(a, 0)] * ootd[0] + ig[sy(a, 1)] * ootd[1] + ig[sy(a, 2)] * ootd[2] -
                             (Gfg[0] * db[0][a] + Gfg[1] * db[1][a] + Gfg[2] * db[2][a]) + Gfg[a] * tdb + advG[a];
    rG[a] = -2.0 * (Ainv[sy(a, 0)] * da[0] + Ainv[sy(a, 1)] * da[1] + Ainv[sy(a, 2)] * da[2]) +
            2.0 * alpha * divAinv[a] - 2.0 * alpha * K.kappa1 * (Gm[a] - Gfg[a]) + pseudolie;
    rhs[14 + a] = rG[a] + koG[a];
  }
  rhs[13] += -trcdda + alpha * (cAA + K_ * K_ / 3.0) + K.kappa1 * (1.0 - K.kappa2) * alpha * Theta;
  rhs[6] += (K.chipsipower / 6.0) * chig * (divbeta - alpha * K_);
  rhs[17] = (alpha * (0.5 * Hhat - (2.0 + K.kappa2) * K.kappa1 * Theta) + rhs[17]) * K.zeroThetaRHS + koT;

  const double falpha =
      K.oploglapse * K.lapseharmonicf + K.oploglapse2 * (8.0 / 3.0) / (3.0 - alpha) + K.harmoniclapse * alpha;
  rhs[18] -= falpha * alpha * Khat;
  if (K.withB != 0.0) {
    const double betaF = K.shiftgammacoeff * pow(alpha, K.shiftalphapower);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      rhs[19 + a] += (K.gamma2factor + K.gamma0factor * betaF) * Bv[a];
      rhs[22 + a] += (K.gamma0factor + K.gamma2factor * betaF) * (rG[a] - K.withGadv * advG[a]) -
                     K.shiftdriver * Bv[a];
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; a++) rhs[19 + a] += Gm[a] - K.shiftdriver * beta[a];
  }
  if (!active) return;
  bool finite = true;
#pragma unroll
  for (int v = 0; v < 25; v++) finite = finite && isfinite(rhs[v]);
  if (!finite) *flag = 1;
  store_point(P, K, q, rhs);
}

// ------------------------------------------------------------------ kernels
// all computed points that are not in [blo, bhi] (the bulk region), or all
// computed points when edge_only == 0
__global__ void __launch_bounds__(256) z4c_rhs_edge(Z4cPtrs P, Z4cGeo G, Z4cK K, int blo0, int blo1, int blo2,
                                                   int bhi0, int bhi1, int bhi2, int *flag) {
  const int i = G.lo[0] + blockIdx.x * 64 + threadIdx.x;
  const int j = G.lo[1] + blockIdx.y * 4 + threadIdx.y;
  const int k = G.lo[2] + blockIdx.z;
  if (i > G.hi[0] || j > G.hi[1] || k > G.hi[2]) return;
  if (G.edge_only && i >= blo0 && i <= bhi0 && j >= blo1 && j <= bhi1 && k >= blo2 && k <= bhi2) return;
  const long q = G.noffset + i + G.sj * j + G.sk * k;
  Edge dv(G, q, i, j, k, K);
  if (!z4c_point(dv, P, K, q)) *flag = 1;
}

__global__ void set_int(int *p, int v) { *p = v; }

static Z4cK make_k(const bamhip_z4c_params *p, double c) {
  Z4cK K;
  K.order_centered = p->order_centered;
  K.order_advection = p->order_advection;
  K.lopsided = p->advection_lopsided;
  K.order_dissipation = p->order_dissipation;
  K.dissfactor = p->dissfactor;
  K.kappa1 = p->kappa1;
  K.kappa2 = p->kappa2;
  K.zeroThetaRHS = p->zeroThetaRHS;
  K.subtractA = p->subtractA;
  K.normalizedetg = p->normalizedetg;
  K.oploglapse = p->oploglapse;
  K.oploglapse2 = p->oploglapse2;
  K.oplogwithshift = p->oplogwithshift;
  K.harmoniclapse = p->harmoniclapse;
  K.lapseharmonicf = p->lapseharmonicf;
  K.gamma0factor = p->gamma0factor;
  K.gamma2factor = p->gamma2factor;
  K.withGadv = p->withGadv;
  K.withShiftadv = p->withShiftadv;
  K.withBadv = p->withBadv;
  K.withB = p->withB;
  K.shiftalphapower = p->shiftalphapower;
  K.shiftgammacoeff = p->shiftgammacoeff;
  K.shiftdriver = p->shiftdriver;
  K.chiDivFloor = p->chiDivFloor;
  K.chipsipower = p->chipsipower;
  K.c = c;
  K.addlinear = (c != 0.0);
  K.rk = 0;
  K.has_v = 0;
  K.has_cp = 0;
  K.wc = K.wv = 0.0;
  return K;
}

static void make_geo(const DevBox &B, Z4cGeo &G) {
  G.m = B.m;
  G.n = B.n;
  G.o = B.o;
  G.noffset = B.noffset;
  G.lo[0] = B.skip[0];
  G.hi[0] = B.m - 1 - B.skip[1];
  G.lo[1] = B.skip[2];
  G.hi[1] = B.n - 1 - B.skip[3];
  G.lo[2] = B.skip[4];
  G.hi[2] = B.o - 1 - B.skip[5];
  G.sj = B.m;
  G.sk = (long)B.m * B.n;
  const double h[3] = {B.dx, B.dy, B.dz};
  for (int a = 0; a < 3; a++) {
    G.h[a] = h[a];
    G.ih[a] = 1.0 / h[a];
    G.i12h[a] = 1.0 / (12.0 * h[a]);
    G.i12hh[a] = 1.0 / (12.0 * h[a] * h[a]);
    G.i2h[a] = 1.0 / (2.0 * h[a]);
    G.ihh[a] = 1.0 / (h[a] * h[a]);
  }
  G.i144hh[0] = 1.0 / (144.0 * h[0] * h[1]);
  G.i144hh[1] = 1.0 / (144.0 * h[0] * h[2]);
  G.i144hh[2] = 1.0 / (144.0 * h[1] * h[2]);
  G.i4hh[0] = 1.0 / (4.0 * h[0] * h[1]);
  G.i4hh[1] = 1.0 / (4.0 * h[0] * h[2]);
  G.i4hh[2] = 1.0 / (4.0 * h[1] * h[2]);
  G.edge_only = 0;
}

// tile regions of one level, by layout
struct LevelTiles {
  std::vector<Z4cGeo> bulk;
  std::vector<Z4cGeo> sx, sy, sz;  // shell as slabs (up to two points thick) normal to x, y, z
  std::vector<Z4cGeo> lx, ly, lz;  // the same slabs cut into single layers
};

static void add_region(std::vector<Z4cGeo> &v, const Z4cGeo &G0, const int lo[3], const int hi[3]) {
  if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) return;
  Z4cGeo G = G0;
  for (int a = 0; a < 3; a++) {
    G.lo[a] = lo[a];
    G.hi[a] = hi[a];
  }
  v.push_back(G);
}

template <class L>
static int ntiles(const Z4cGeo &G) {
  return ((G.hi[0] - G.lo[0] + L::TX) / L::TX) * ((G.hi[1] - G.lo[1] + L::TY) / L::TY) *
         ((G.hi[2] - G.lo[2] + L::TZ) / L::TZ);
}

// one launch per kMaxRegions regions of a layout
static int g_xcd_map = 1;  // XCD-aware tile numbering in the tile kernels

static bool same_spacing(const Z4cGeo &a, const Z4cGeo &b) {
  return a.h[0] == b.h[0] && a.h[1] == b.h[1] && a.h[2] == b.h[2];
}

// one launch per kMaxRegions regions of a layout (and per grid spacing)
template <class L, bool SH, int NB, int VEC, int OCC>
static int launch_regions(const Z4cPtrs &P, const std::vector<Z4cGeo> &regs, const Z4cK &K, int *flag,
                          hipStream_t st) {
  size_t r0 = 0;
  while (r0 < regs.size()) {
    TileJobs J{};
    const Z4cGeo &G0 = regs[r0];
    for (int a = 0; a < 3; a++) {
      J.sp.h[a] = G0.h[a];
      J.sp.ih[a] = G0.ih[a];
      J.sp.i12h[a] = G0.i12h[a];
      J.sp.i12hh[a] = G0.i12hh[a];
      J.sp.i144hh[a] = G0.i144hh[a];
      J.sp.i2h[a] = G0.i2h[a];
      J.sp.ihh[a] = G0.ihh[a];
      J.sp.i4hh[a] = G0.i4hh[a];
    }
    int blk = 0;
    J.n = 0;
    while (r0 < regs.size() && J.n < kMaxRegions && same_spacing(regs[r0], G0)) {
      const Z4cGeo &G = regs[r0++];
      RegionGeo &Rg = J.reg[J.n];
      Rg.m = G.m;
      Rg.n = G.n;
      Rg.o = G.o;
      Rg.noffset = G.noffset;
      Rg.sj = G.sj;
      Rg.sk = G.sk;
      for (int a = 0; a < 3; a++) {
        Rg.lo[a] = G.lo[a];
        Rg.hi[a] = G.hi[a];
      }
      J.blk0[J.n++] = blk;
      blk += ntiles<L>(G);
    }
    J.blk0[J.n] = blk;
    if (blk == 0) continue;
    J.xcd = g_xcd_map;
    const unsigned grid = g_xcd_map ? (unsigned)(8 * ((blk + 7) / 8)) : (unsigned)blk;
    z4c_rhs_tile<L, SH, NB, VEC, OCC><<<grid, NTHR, 0, st>>>(P, J, K, flag);
    BH_HIP(hipGetLastError());
  }
  return BAMHIP_OK;
}

// Split one box into its bulk region (>= 3 points from every edge) and the
// slabs of its edge shell (fast parameters), or launch the generic per-point
// kernel for the other parameter choices.
static int plan_box(const DevBox &B, const Z4cPtrs &P, const Z4cK &K, int *flag, bool fast, LevelTiles &T) {
  Z4cGeo G;
  make_geo(B, G);
  if (G.hi[0] < G.lo[0] || G.hi[1] < G.lo[1] || G.hi[2] < G.lo[2]) return BAMHIP_OK;
  int blo[3], bhi[3];
  const int dims[3] = {B.m, B.n, B.o};
  bool have_bulk = fast;
  for (int a = 0; a < 3; a++) {
    blo[a] = G.lo[a] > 3 ? G.lo[a] : 3;
    bhi[a] = G.hi[a] < dims[a] - 4 ? G.hi[a] : dims[a] - 4;
    if (bhi[a] < blo[a]) have_bulk = false;
  }
  if (have_bulk) add_region(T.bulk, G, blo, bhi);
  const bool need_edge = !have_bulk || blo[0] > G.lo[0] || bhi[0] < G.hi[0] || blo[1] > G.lo[1] ||
                         bhi[1] < G.hi[1] || blo[2] > G.lo[2] || bhi[2] < G.hi[2];
  if (need_edge && fast) {
    // The shell of points closer than 3 to an edge, as disjoint slabs: the
    // layers along x over all y, z; along y over the x interior; along z over
    // the x and y interior.  Layers of axis a: [lo, min(hi, 2)] and
    // [max(dim - 3, end of the first + 1), hi].
    int lay[3][2][2], ilo[3], ihi[3];
    for (int a = 0; a < 3; a++) {
      lay[a][0][0] = G.lo[a];
      lay[a][0][1] = std::min(G.hi[a], 2);
      lay[a][1][0] = std::max(dims[a] - 3, lay[a][0][1] + 1);
      lay[a][1][1] = G.hi[a];
      ilo[a] = std::max(G.lo[a], 3);
      ihi[a] = std::min(G.hi[a], dims[a] - 4);
    }
    for (int s = 0; s < 2; s++) {
      const int xl[3] = {lay[0][s][0], G.lo[1], G.lo[2]}, xh[3] = {lay[0][s][1], G.hi[1], G.hi[2]};
      add_region(T.sx, G, xl, xh);
      const int yl[3] = {ilo[0], lay[1][s][0], G.lo[2]}, yh[3] = {ihi[0], lay[1][s][1], G.hi[2]};
      add_region(T.sy, G, yl, yh);
      const int zl[3] = {ilo[0], ilo[1], lay[2][s][0]}, zh[3] = {ihi[0], ihi[1], lay[2][s][1]};
      add_region(T.sz, G, zl, zh);
      // single layers: the distance to the edge along the slab normal is then
      // uniform over a layer, and so are the stencil rules for most points
      const int *lo3[3] = {xl, yl, zl}, *hi3[3] = {xh, yh, zh};
      std::vector<Z4cGeo> *dst[3] = {&T.lx, &T.ly, &T.lz};
      for (int a = 0; a < 3; a++)
        for (int c = lo3[a][a]; c <= hi3[a][a]; c++) {
          int l1[3] = {lo3[a][0], lo3[a][1], lo3[a][2]}, h1[3] = {hi3[a][0], hi3[a][1], hi3[a][2]};
          l1[a] = h1[a] = c;
          add_region(*dst[a], G, l1, h1);
        }
    }
  } else if (need_edge) {
    G.edge_only = have_bulk ? 1 : 0;
    dim3 grid((G.hi[0] - G.lo[0] + 64) / 64, (G.hi[1] - G.lo[1] + 4) / 4, G.hi[2] - G.lo[2] + 1);
    {
      TimerScope t("z4_rhs_edge");
      z4c_rhs_edge<<<grid, dim3(64, 4, 1), 0, stream()>>>(P, G, K, blo[0], blo[1], blo[2], bhi[0], bhi[1], bhi[2],
                                                          flag);
    }
    BH_HIP(hipGetLastError());
  }
  return BAMHIP_OK;
}

// ---- autotuner of the tile kernels
// A variant is one instantiation of the tile kernel: tile shape (TX x TY x
// TZ, so the z depth of a block) and staging (registers, or LDS-DMA with NB
// buffers).  Every variant evaluates the same per-point arithmetic in the same
// order, so all give bitwise the same right-hand side.  The first time a set
// of regions (a level's boxes) is seen outside graph capture, each variant is
// timed on it (one warm-up launch, then two timed with HIP events) and the
// fastest is kept for that set.
using RegionLists = std::vector<const std::vector<Z4cGeo> *>;
using LaunchFn = int (*)(const Z4cPtrs &, const RegionLists &, const Z4cK &, int *, hipStream_t);
struct Variant {
  const char *name;
  bool dma;    // LDS-DMA staging: byte offsets within a field must fit 32 bits
  bool pairs;  // 16-byte LDS-DMA units: needs 16-byte aligned element pairs
  LaunchFn fn;
};
template <class L, int NB, int VEC, int OCC>
static int bulk_variant(const Z4cPtrs &P, const RegionLists &R, const Z4cK &K, int *flag, hipStream_t st) {
  return launch_regions<L, false, NB, VEC, OCC>(P, *R[0], K, flag, st);
}
template <int NB, int VEC, int OCC>
static int shell_variant(const Z4cPtrs &P, const RegionLists &R, const Z4cK &K, int *flag, hipStream_t st) {
  int e;
  if ((e = launch_regions<LSlabX, true, NB, VEC, OCC>(P, *R[0], K, flag, st))) return e;
  if ((e = launch_regions<LSlabY, true, NB, VEC, OCC>(P, *R[1], K, flag, st))) return e;
  return launch_regions<LSlabZ, true, NB, VEC, OCC>(P, *R[2], K, flag, st);
}
// the slabs with waves that keep one distance to the slab's face
using LSlabXo = Lay<2, 16, 8, 2, 1>;
using LSlabYo = Lay<16, 2, 8, 0, 2>;
template <int NB, int VEC, int OCC>
static int shell_ord_variant(const Z4cPtrs &P, const RegionLists &R, const Z4cK &K, int *flag, hipStream_t st) {
  int e;
  if ((e = launch_regions<LSlabXo, true, NB, VEC, OCC>(P, *R[0], K, flag, st))) return e;
  if ((e = launch_regions<LSlabYo, true, NB, VEC, OCC>(P, *R[1], K, flag, st))) return e;
  return launch_regions<LSlabZ, true, NB, VEC, OCC>(P, *R[2], K, flag, st);
}
template <int NB, int OCC>
static int layer_variant(const Z4cPtrs &P, const RegionLists &R, const Z4cK &K, int *flag, hipStream_t st) {
  int e;
  if ((e = launch_regions<LLayX, true, NB, 1, OCC>(P, *R[3], K, flag, st))) return e;
  if ((e = launch_regions<LLayY, true, NB, 1, OCC>(P, *R[4], K, flag, st))) return e;
  return launch_regions<LLayZ, true, NB, 1, OCC>(P, *R[5], K, flag, st);
}
constexpr int kMaxVariants = 10;
static const Variant kVariants[2][kMaxVariants] = {
    {{"bulk 16x4x4 dma3", true, false, bulk_variant<LBulk, 3, 1, 1>},
     {"bulk 16x4x4 reg", false, false, bulk_variant<LBulk, 0, 1, 1>},
     {"bulk 16x4x4 dma4", true, false, bulk_variant<LBulk, 4, 1, 1>},
     {"bulk 32x4x2 dma3", true, false, bulk_variant<Lay<32, 4, 2>, 3, 1, 1>},
     {"bulk 16x8x2 dma3", true, false, bulk_variant<Lay<16, 8, 2>, 3, 1, 1>},
     {"bulk 16x4x4 dma3 x4", true, true, bulk_variant<LBulk, 3, 4, 1>},
     {"bulk 16x4x4 dma4 x4", true, true, bulk_variant<LBulk, 4, 4, 1>},
     {"bulk 32x4x2 dma3 x4", true, true, bulk_variant<Lay<32, 4, 2>, 3, 4, 1>},
     {"bulk 16x4x4 dma4 occ2", true, false, bulk_variant<LBulk, 4, 1, 2>}},
    {{"shell layers dma3", true, false, layer_variant<3, 1>},
     {"shell slabs reg", false, false, shell_variant<0, 1, 1>},
     {"shell slabs dma3", true, false, shell_variant<3, 1, 1>},
     {"shell layers dma4", true, false, layer_variant<4, 1>},
     {"shell slabs dma3 x4", true, true, shell_variant<3, 4, 1>},
     {"shell slabs ord dma3 x4", true, true, shell_ord_variant<3, 4, 1>}}};
static const int kNumVariants[2] = {9, 6};
// variant 0 of each kind is the default; variant 1 needs no LDS-DMA

struct TuneEntry {
  int choice;
  double ms[kMaxVariants];
};
struct TuneState {
  int overlap = 1;  // shell on the side stream, concurrent with the bulk
  int mode = 1;  // 0: defaults only; 1: tune each new region set
  int forced[2] = {-1, -1};
  std::map<std::string, TuneEntry> chosen[2];
};
static TuneState g_tune;

static std::string region_signature(const RegionLists &R) {
  std::string s;
  char buf[96];
  for (const auto *v : R) {
    s += '|';
    for (const Z4cGeo &G : *v) {
      snprintf(buf, sizeof buf, "%d,%d,%d:%d-%d,%d-%d,%d-%d;", G.m, G.n, G.o, G.lo[0], G.hi[0], G.lo[1], G.hi[1],
               G.lo[2], G.hi[2]);
      s += buf;
    }
  }
  return s;
}

// Variant of kind for the region lists R: the pinned one, the one kept for
// this region set, or (mode 1, outside capture, repeatable launches) the
// fastest after timing all of them on the library stream.  In the last case
// the chosen variant has also been run once more, so the outputs are in place
// and *done is set.
static int choose_variant(int kind, const Z4cPtrs &P, const RegionLists &R, const Z4cK &K, int *flag,
                          bool repeatable, bool dma_ok, bool pairs_ok, int *idx, bool *done) {
  const Variant *V = kVariants[kind];
  const int nv = kNumVariants[kind];
  *done = false;
  *idx = g_tune.forced[kind];
  if (*idx >= 0) {
    if (V[*idx].dma && !dma_ok)
      return set_error(BAMHIP_EINVAL, "z4c tile variant '%s' needs a level below 2^29 points", V[*idx].name);
    if (V[*idx].pairs && !pairs_ok)
      return set_error(BAMHIP_EINVAL, "z4c tile variant '%s' needs even box sizes and offsets", V[*idx].name);
    return BAMHIP_OK;
  }
  const std::string sig = region_signature(R);
  auto it = g_tune.chosen[kind].find(sig);
  if (it != g_tune.chosen[kind].end()) {
    *idx = it->second.choice;
    return BAMHIP_OK;
  }
  const int dflt = dma_ok ? 0 : 1;
  *idx = dflt;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  BH_HIP(hipStreamIsCapturing(stream(), &cs));
  if (g_tune.mode == 0 || !repeatable || cs != hipStreamCaptureStatusNone) return BAMHIP_OK;
  TuneEntry te;
  te.choice = dflt;
  hipEvent_t a, b;
  BH_HIP(hipEventCreate(&a));
  BH_HIP(hipEventCreate(&b));
  double best = 1e300;
  int e = BAMHIP_OK;
  for (int v = 0; v < nv && e == BAMHIP_OK; v++) {
    te.ms[v] = -1.0;
    if ((V[v].dma && !dma_ok) || (V[v].pairs && !pairs_ok)) continue;
    if ((e = V[v].fn(P, R, K, flag, stream()))) break;
    BH_HIP(hipEventRecord(a, stream()));
    for (int rep = 0; rep < 2 && e == BAMHIP_OK; rep++) e = V[v].fn(P, R, K, flag, stream());
    BH_HIP(hipEventRecord(b, stream()));
    BH_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    BH_HIP(hipEventElapsedTime(&ms, a, b));
    te.ms[v] = ms / 2.0;
    if (te.ms[v] < best) {
      best = te.ms[v];
      te.choice = v;
    }
  }
  BH_HIP(hipEventDestroy(a));
  BH_HIP(hipEventDestroy(b));
  if (e) return e;
  g_tune.chosen[kind][sig] = te;
  *idx = te.choice;
  // the outputs come from the chosen variant, as in every later call
  *done = true;
  return V[te.choice].fn(P, R, K, flag, stream());
}

// The tile and edge kernels of one RHS evaluation (plain, addlinear or fused
// RK4 stage, as K says) over all boxes of a level.  A fused stage rewrites u_c,
// so its kernels cannot be timed on the real lists: the tuner then times the
// plain RHS into scratch (when given) and the stage runs the chosen variant.
static int run_rhs(bamhip_level *lev, const Z4cPtrs &P, const Z4cK &K, bool check_finite, double *const *scratch) {
  int e;
  if (check_finite) set_int<<<1, 1, 0, stream()>>>(lev->d_flag, 0);
  const bool fast = (K.order_centered == 4 && K.order_advection == 4 && K.lopsided == 1 &&
                     (K.order_dissipation == 4 || K.order_dissipation == 0));
  LevelTiles T;
  for (const DevBox &B : lev->boxes)
    if ((e = plan_box(B, P, K, lev->d_flag, fast, T))) return e;
  // repeated launches (tuning) give the same result unless the kernels write
  // what they read
  bool repeatable = !K.rk;
  for (int v = 0; v < 25 && repeatable; v++)
    for (int w = 0; w < 25; w++)
      if (P.r[v] == P.p[w]) repeatable = false;
  Z4cPtrs PT = P;  // lists the tuner runs on
  Z4cK KT = K;
  if (K.rk && scratch) {
    for (int v = 0; v < 25; v++) {
      PT.r[v] = scratch[v];
      PT.p[v] = P.u[v];
    }
    KT.rk = 0;
    KT.has_v = 0;
    KT.addlinear = 0;
    KT.c = 0.0;
    repeatable = true;
  }
  const bool dma_ok = (long)lev->npoints * 8 < (1L << 32);
  const RegionLists RB = {&T.bulk}, RS = {&T.sx, &T.sy, &T.sz, &T.lx, &T.ly, &T.lz};
  const bool have_shell = !T.sx.empty() || !T.sy.empty() || !T.sz.empty();
  int ib = -1, is = -1;
  bool db = false, ds = false;
  // 16-byte units hold element pairs (x, x + 1) with x even in every box:
  // even box rows, box offsets and tile rows (first staged x = i0 - 3 even)
  bool pairs_ok = true;
  for (const auto *v : {&T.bulk, &T.sx, &T.sy, &T.sz})
    for (const Z4cGeo &G : *v) pairs_ok = pairs_ok && G.m % 2 == 0 && G.noffset % 2 == 0 && (G.lo[0] - HALO) % 2 == 0;
  if (!T.bulk.empty() &&
      (e = choose_variant(0, PT, RB, KT, lev->d_flag, repeatable, dma_ok, pairs_ok, &ib, &db)))
    return e;
  if (have_shell && (e = choose_variant(1, PT, RS, KT, lev->d_flag, repeatable, dma_ok, pairs_ok, &is, &ds)))
    return e;
  if (K.rk) db = ds = false;  // the tuner ran the plain RHS; the stage itself is still to run
  if (ib >= 0 && is >= 0 && !db && !ds && g_tune.overlap) {
    // the shell on the side stream beside the bulk: the two write disjoint
    // points, and the shell's few blocks fill the CUs the bulk's last wave of
    // blocks leaves idle.  The bulk is enqueued first (1% faster on the FMR
    // hierarchy than the shell first).
    TimerScope t("z4_rhs");
    if ((e = fork_side())) return e;
    if ((e = kVariants[0][ib].fn(P, RB, K, lev->d_flag, stream()))) return e;
    if ((e = kVariants[1][is].fn(P, RS, K, lev->d_flag, side_stream()))) return e;
    if ((e = join_side())) return e;
  } else {
    if (ib >= 0 && !db) {
      TimerScope t("z4_rhs");
      if ((e = kVariants[0][ib].fn(P, RB, K, lev->d_flag, stream()))) return e;
    }
    if (is >= 0 && !ds) {
      TimerScope t("z4_rhs_edge");
      if ((e = kVariants[1][is].fn(P, RS, K, lev->d_flag, stream()))) return e;
    }
  }
  if (check_finite) {
    int h = 0;
    BH_HIP(hipMemcpyAsync(&h, lev->d_flag, sizeof(int), hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    if (h) return set_error(BAMHIP_ENONFINITE, "bamhip_z4c_rhs: non-finite right-hand side on level %d", lev->l);
  }
  return BAMHIP_OK;
}

int z4c_rhs_impl(bamhip_level *lev, const bamhip_z4c_params *par, const int *unew, const int *upre, double c,
                 const int *ucur, bool check_finite) {
  if (!lev || !par) return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs: NULL argument");
  if (par->order_centered != 2 && par->order_centered != 4)
    return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs: order_centered=%d not implemented (2, 4)", par->order_centered);
  if (par->order_advection != 2 && par->order_advection != 4)
    return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs: order_advection=%d not implemented (2, 4)",
                     par->order_advection);
  if (par->advection_lopsided < 0 || par->advection_lopsided > 2)
    return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs: advection_lopsided=%d", par->advection_lopsided);
  if (par->order_dissipation != 0 && par->order_dissipation != 4)
    return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs: order_dissipation=%d not implemented (0, 4)",
                     par->order_dissipation);
  Z4cPtrs P;
  int e;
  if ((e = gather_ptrs(lev, ucur, 25, const_cast<double **>(P.u), "bamhip_z4c_rhs(ucur)"))) return e;
  if ((e = gather_ptrs(lev, unew, 25, P.r, "bamhip_z4c_rhs(unew)"))) return e;
  if (c != 0.0) {
    if ((e = gather_ptrs(lev, upre, 25, const_cast<double **>(P.p), "bamhip_z4c_rhs(upre)"))) return e;
  } else {
    for (int v = 0; v < 25; v++) P.p[v] = P.u[v];
  }
  for (int v = 0; v < 25; v++) {
    P.c[v] = nullptr;
    P.a[v] = nullptr;
    P.cp[v] = nullptr;
  }
  for (int v = 0; v < 25; v++)
    for (int w = 0; w < 25; w++)
      if (P.r[v] == P.u[w])
        return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs: unew[%d] aliases ucur[%d] (stencils need the old data)", v, w);
  // the stencil reach (3 with lop-sided advection, 4 with one-sided) must stay
  // inside every box
  for (const DevBox &B : lev->boxes)
    for (int f = 0; f < 6; f++)
      if (B.skip[f] < 1) return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs: box face %d has no flagged layer", f);
  const Z4cK K = make_k(par, c);
  return run_rhs(lev, P, K, check_finite, nullptr);
}

static int check_params(const bamhip_z4c_params *par, const char *what) {
  if (par->order_centered != 2 && par->order_centered != 4)
    return set_error(BAMHIP_EINVAL, "%s: order_centered=%d not implemented (2, 4)", what, par->order_centered);
  if (par->order_advection != 2 && par->order_advection != 4)
    return set_error(BAMHIP_EINVAL, "%s: order_advection=%d not implemented (2, 4)", what, par->order_advection);
  if (par->advection_lopsided < 0 || par->advection_lopsided > 2)
    return set_error(BAMHIP_EINVAL, "%s: advection_lopsided=%d", what, par->advection_lopsided);
  if (par->order_dissipation != 0 && par->order_dissipation != 4)
    return set_error(BAMHIP_EINVAL, "%s: order_dissipation=%d not implemented (0, 4)", what, par->order_dissipation);
  return BAMHIP_OK;
}

// One RK4 stage with the update fused into the RHS kernels' store: at every
// computed point accout = acc + wc RHS(src), when vout != NULL vout = base +
// wv RHS(src), and when copy != NULL copy = base.  src is read by the
// stencils, so no output may alias it; acc and base are read at the point
// only, before any store, so an output may alias them (the accumulator read
// from the list the next state goes to, for one).  The outputs must differ
// from each other.  scratch (optional): 25 variables the tuner may use as a
// plain RHS output the first time a box set is seen; it must not alias an
// input (src, acc, base).
int z4c_rhs_rk4_impl(bamhip_level *lev, const bamhip_z4c_params *par, const int *src, const int *acc,
                     const int *accout, const int *base, const int *vout, const int *copy, double wc, double wv,
                     const int *scratch) {
  if (!lev || !par || !src || !acc || !accout || !base)
    return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs_rk4_stage: NULL argument");
  int e;
  if ((e = check_params(par, "bamhip_z4c_rhs_rk4_stage"))) return e;
  Z4cPtrs P;
  if ((e = gather_ptrs(lev, src, 25, const_cast<double **>(P.u), "bamhip_z4c_rhs_rk4_stage(src)"))) return e;
  if ((e = gather_ptrs(lev, acc, 25, const_cast<double **>(P.a), "bamhip_z4c_rhs_rk4_stage(acc)"))) return e;
  if ((e = gather_ptrs(lev, accout, 25, P.c, "bamhip_z4c_rhs_rk4_stage(accout)"))) return e;
  if ((e = gather_ptrs(lev, base, 25, const_cast<double **>(P.p), "bamhip_z4c_rhs_rk4_stage(base)"))) return e;
  if (vout) {
    if ((e = gather_ptrs(lev, vout, 25, P.r, "bamhip_z4c_rhs_rk4_stage(vout)"))) return e;
  } else {
    for (int v = 0; v < 25; v++) P.r[v] = nullptr;
  }
  if (copy) {
    if ((e = gather_ptrs(lev, copy, 25, P.cp, "bamhip_z4c_rhs_rk4_stage(copy)"))) return e;
  } else {
    for (int v = 0; v < 25; v++) P.cp[v] = nullptr;
  }
  for (int v = 0; v < 25; v++)
    for (int w = 0; w < 25; w++) {
      const double *outs[3] = {P.c[v], P.r[v], P.cp[v]}, *outw[3] = {P.c[w], P.r[w], P.cp[w]};
      for (int x = 0; x < 3; x++) {
        if (!outs[x]) continue;
        if (outs[x] == P.u[w])
          return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs_rk4_stage: an output list aliases src");
        for (int y = 0; y < 3; y++)
          if (y != x && outs[x] == outw[y])
            return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs_rk4_stage: the output lists alias each other");
      }
    }
  for (const DevBox &B : lev->boxes)
    for (int f = 0; f < 6; f++)
      if (B.skip[f] < 1)
        return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs_rk4_stage: box face %d has no flagged layer", f);
  double *scr[BAMHIP_Z4C_NVARS];
  if (scratch) {
    if ((e = gather_ptrs(lev, scratch, 25, scr, "bamhip_z4c_rhs_rk4_stage(scratch)"))) return e;
    for (int v = 0; v < 25; v++)
      for (int w = 0; w < 25; w++)
        if (scr[v] == P.u[w] || scr[v] == P.a[w] || scr[v] == P.p[w])
          return set_error(BAMHIP_EINVAL, "bamhip_z4c_rhs_rk4_stage: scratch aliases an input list");
  }
  Z4cK K = make_k(par, 0.0);
  K.rk = 1;
  K.has_v = vout ? 1 : 0;
  K.has_cp = copy ? 1 : 0;
  K.wc = wc;
  K.wv = wv;
  return run_rhs(lev, P, K, lev->check_mode != 0, scratch ? scr : nullptr);
}

}  // namespace bamhip

extern "C" int bamhip_z4c_rhs_rk4_stage(bamhip_level *lev, const bamhip_z4c_params *par, const int *src,
                                        const int *acc, const int *accout, const int *base, const int *vout,
                                        const int *copy, double wc, double wv, const int *scratch) {
  return bamhip::z4c_rhs_rk4_impl(lev, par, src, acc, accout, base, vout, copy, wc, wv, scratch);
}

extern "C" int bamhip_z4c_rhs(bamhip_level *lev, const bamhip_z4c_params *par, const int *unew, const int *upre,
                              double c, const int *ucur) {
  return bamhip::z4c_rhs_impl(lev, par, unew, upre, c, ucur, lev ? lev->check_mode != 0 : true);
}

extern "C" int bamhip_z4c_tune_mode(int mode) {
  if (mode != 0 && mode != 1) return bamhip::set_error(BAMHIP_EINVAL, "bamhip_z4c_tune_mode: mode %d (0, 1)", mode);
  const int old = bamhip::g_tune.mode;
  bamhip::g_tune.mode = mode;
  return old;
}

extern "C" int bamhip_z4c_set_overlap(int on) {
  const int old = bamhip::g_tune.overlap;
  bamhip::g_tune.overlap = on != 0;
  return old;
}

extern "C" int bamhip_z4c_set_xcd_map(int on) {
  const int old = bamhip::g_xcd_map;
  bamhip::g_xcd_map = on != 0;
  return old;
}

extern "C" int bamhip_z4c_tune_clear(void) {
  bamhip::g_tune.chosen[0].clear();
  bamhip::g_tune.chosen[1].clear();
  return BAMHIP_OK;
}

extern "C" int bamhip_z4c_variant_count(int kind) {
  if (kind != 0 && kind != 1) return bamhip::set_error(BAMHIP_EINVAL, "bamhip_z4c_variant_count: kind %d", kind);
  return bamhip::kNumVariants[kind];
}

extern "C" const char *bamhip_z4c_variant_name(int kind, int idx) {
  if (kind != 0 && kind != 1) return nullptr;
  if (idx < 0 || idx >= bamhip::kNumVariants[kind]) return nullptr;
  return bamhip::kVariants[kind][idx].name;
}

extern "C" int bamhip_z4c_force_variant(int kind, int idx) {
  if (kind != 0 && kind != 1) return bamhip::set_error(BAMHIP_EINVAL, "bamhip_z4c_force_variant: kind %d", kind);
  if (idx < -1 || idx >= bamhip::kNumVariants[kind])
    return bamhip::set_error(BAMHIP_EINVAL, "bamhip_z4c_force_variant: variant %d of %d", idx,
                             bamhip::kNumVariants[kind]);
  bamhip::g_tune.forced[kind] = idx;
  return BAMHIP_OK;
}

extern "C" int bamhip_z4c_tune_report(char *buf, int len) {
  std::string s;
  char line[160];
  for (int kind = 0; kind < 2; kind++)
    for (const auto &kv : bamhip::g_tune.chosen[kind]) {
      const bamhip::TuneEntry &te = kv.second;
      snprintf(line, sizeof line, "%s <- %zu regions:", bamhip::kVariants[kind][te.choice].name,
               (size_t)std::count(kv.first.begin(), kv.first.end(), ';'));
      s += line;
      for (int v = 0; v < bamhip::kNumVariants[kind]; v++) {
        snprintf(line, sizeof line, " %s=%.4fms", bamhip::kVariants[kind][v].name, te.ms[v]);
        s += line;
      }
      s += '\n';
    }
  if (buf && len > 0) {
    const int n = (int)std::min<size_t>(s.size(), (size_t)len - 1);
    memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return (int)s.size() + 1;
}
