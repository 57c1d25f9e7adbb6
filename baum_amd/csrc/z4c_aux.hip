// z4c_aux.hip -- the streaming and boundary kernels around the Z4c RHS:
// SoA vector ops and the fused RK4 stage update, the radiative and periodic
// boundaries, the algebraic constraints and the Hamiltonian-constraint norm.
#include <climits>
#include <cmath>
#include <vector>

#include "bamhip_internal.h"

namespace bamhip {

int z4c_rhs_impl(bamhip_level *lev, const bamhip_z4c_params *par, const int *unew, const int *upre, double c,
                 const int *ucur, bool check_finite);

struct Ptrs25 {
  double *p[BAMHIP_Z4C_NVARS];
};

// --------------------------------------------------------------- streaming
__global__ void fill_kernel(double *__restrict__ p, long n, double v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// r = ca*a + cb*b for nv variables (vladd, variables.c:650-719); 2D grid:
// blockIdx.y selects the variable, the x dimension strides over points.
// r may alias a or b (vladd(u, 1, u, c, w) is a common call), so the
// pointers are not restrict-qualified.
__global__ void vladd_kernel(Ptrs25 r, Ptrs25 a, Ptrs25 b, double ca, double cb, long n) {
  const int v = blockIdx.y;
  double *rv = r.p[v];
  const double *av = a.p[v];
  const double *bv = b.p[v];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    if (ca != 0.0) s += ca * av[i];
    if (cb != 0.0) s += cb * bv[i];
    rv[i] = s;
  }
}

// fused RK4 stage update (rk.c:231-252): uc += wc*r ; v = up + wv*r
// (wv == 0: last stage, no v)
__global__ void rk4_update_kernel(Ptrs25 uc, Ptrs25 up, Ptrs25 r, Ptrs25 v, double wc, double wv, long n) {
  const int k = blockIdx.y;
  double *__restrict__ c = uc.p[k];
  const double *__restrict__ p = up.p[k];
  const double *__restrict__ rr = r.p[k];
  double *__restrict__ vv = v.p[k];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double ri = rr[i];
    c[i] += wc * ri;
    if (wv != 0.0) vv[i] = p[i] + wv * ri;
  }
}

static dim3 stream_grid(long n, int nv) {
  long blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  return dim3((unsigned)blocks, nv, 1);
}

// ---------------------------------------------------------------- radiative
// per-variable far limit and speed (z4_startup, z4.c:200-228)
struct RadInfo {
  double var0[25], v[25];
};

// one thread per surface point of one box, all 25 variables
// rk = 0: unew = RHS (c == 0) or upre + c RHS at the boundary points; rk = 1
// (an RK4 stage with the update fused in): uc = ucin + wc RHS, when has_v
// unew = upre + wv RHS, when has_cp copy = upre
struct RkFuse {
  Ptrs25 uc, ucin, copy;
  double wc, wv;
  int rk, has_v, has_cp;
};
__global__ void radiative_kernel(Ptrs25 unew, Ptrs25 upre, Ptrs25 ucur, RadInfo ri, DevBox B, double c, RkFuse F) {
  const int M = B.m, N = B.n, O = B.o;
  const long nface_x = (long)N * O, nface_y = (long)(M - 2) * O, nface_z = (long)(M - 2) * (N - 2);
  const long total = 2 * (nface_x + nface_y + nface_z);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= total) return;
  int i, j, k, face;
  long r = t;
  if (r < 2 * nface_x) {
    face = (int)(r / nface_x);
    r %= nface_x;
    i = face ? M - 1 : 0;
    j = (int)(r % N);
    k = (int)(r / N);
  } else if ((r -= 2 * nface_x) < 2 * nface_y) {
    face = 2 + (int)(r / nface_y);
    r %= nface_y;
    j = (face == 3) ? N - 1 : 0;
    i = 1 + (int)(r % (M - 2));
    k = (int)(r / (M - 2));
  } else {
    r -= 2 * nface_y;
    face = 4 + (int)(r / nface_z);
    r %= nface_z;
    k = (face == 5) ? O - 1 : 0;
    i = 1 + (int)(r % (M - 2));
    j = 1 + (int)(r / (M - 2));
  }
  if (j < 0 || j >= N || k < 0 || k >= O || i < 0 || i >= M) return;
  // a surface point is PHYBOUND unless a face whose flagged layers contain it
  // carries another flag (periodic / ghost / symmetry / refinement override)
  const int dist[6] = {i, M - 1 - i, j, N - 1 - j, k, O - 1 - k};
  bool phy = false, other = false;
  for (int f = 0; f < 6; f++) {
    if (dist[f] == 0 && B.bflag[f] == BAMHIP_PHYBOUND) phy = true;
    if (dist[f] < B.skip[f] && B.bflag[f] != BAMHIP_PHYBOUND) other = true;
  }
  if (!phy || other) return;
  const long di = 1, dj = M, dk = (long)M * N;
  const long q = B.noffset + i + dj * j + dk * k;
  // neighbour indices, a missing neighbour equals q (bam_amr_loops_box.h:338-386)
  const long mcc = (i > 0) ? q - di : q, pcc = (i + 1 < M) ? q + di : q;
  const long cmc = (j > 0) ? q - dj : q, cpc = (j + 1 < N) ? q + dj : q;
  const long ccm = (k > 0) ? q - dk : q, ccp = (k + 1 < O) ? q + dk : q;
  const long Mcc = (i > 1) ? mcc - di : q, Pcc = (i + 2 < M) ? pcc + di : q;
  const long cMc = (j > 1) ? cmc - dj : q, cPc = (j + 2 < N) ? cpc + dj : q;
  const long ccM = (k > 1) ? ccm - dk : q, ccP = (k + 2 < O) ? ccp + dk : q;
  const double x = B.x0 + i * B.dx, y = B.y0 + j * B.dy, z = B.z0 + k * B.dz;
  const double rr = sqrt(x * x + y * y + z * z);
  const double nx = x / rr, ny = y / rr, nz = z / rr;
  // one-sided or centred 2nd-order difference times 2h (radiative.c:14-40)
  auto del = [&](const double *u, double vx, long mm, long m, long p, long pp) {
    const double cen = u[p] - u[m];
    const double rgt = 4 * u[p] - u[pp] - 3 * u[q];
    const double lft = 3 * u[q] + u[mm] - 4 * u[m];
    if (vx > 0) return (mm != q) ? lft : ((m != q) ? cen : rgt);
    if (vx < 0) return (pp != q) ? rgt : ((m != q) ? cen : lft);
    if (p != q && m != q) return cen;
    return (m == q) ? rgt : lft;
  };
  for (int v = 0; v < 25; v++) {
    const double *u = ucur.p[v];
    const double sp = ri.v[v];
    const double vx = sp * nx, vy = sp * ny, vz = sp * nz;
    double rhs = -sp * (u[q] - ri.var0[v]) / rr;
    rhs -= vx * del(u, vx, Mcc, mcc, pcc, Pcc) / (2 * B.dx) + vy * del(u, vy, cMc, cmc, cpc, cPc) / (2 * B.dy) +
           vz * del(u, vz, ccM, ccm, ccp, ccP) / (2 * B.dz);
    if (F.rk) {
      // both loads before the stores: an output may alias ucin or upre
      const double a = F.ucin.p[v][q], b = (F.has_v || F.has_cp) ? upre.p[v][q] : 0.0;
      F.uc.p[v][q] = a + F.wc * rhs;
      if (F.has_v) unew.p[v][q] = b + F.wv * rhs;
      if (F.has_cp) F.copy.p[v][q] = b;
    } else {
      unew.p[v][q] = (c != 0.0) ? upre.p[v][q] + c * rhs : rhs;
    }
  }
}

// dst[0 .. nd) = src (or the constant val when fill is set) for nv <= 25
// variables at the points of one box that no RHS kernel computes (the skip
// layers of width skip[f] at face f), enumerated as x slabs (full yz
// planes), y slabs (interior x) and z slabs (interior x and y), x fastest in
// the y and z slabs so that a wave's accesses are contiguous.  One thread
// per point and variable (blockIdx.y).
struct SkipDst {
  Ptrs25 d[3];
  int nd, nv, fill;
  double val;
};
// the boxes of one launch: box b takes blocks [blk0[b], blk0[b + 1])
constexpr int kSkipBoxes = 16;
struct SkipBoxes {
  int n;
  int blk0[kSkipBoxes + 1];
  DevBox box[kSkipBoxes];
};
static_assert(sizeof(SkipDst) + sizeof(Ptrs25) + sizeof(SkipBoxes) <= 4096, "kernel argument size");
__global__ void skip_copy_kernel(SkipDst D, Ptrs25 src, SkipBoxes SB) {
  int b = 0;
  while (b + 1 < SB.n && (int)blockIdx.x >= SB.blk0[b + 1]) b++;
  const DevBox B = SB.box[b];
  const int M = B.m, N = B.n, O = B.o;
  const int *s = B.skip;
  const int ix = M - s[0] - s[1], iy = N - s[2] - s[3];
  // 32-bit index arithmetic: the host keeps a box's skip points below 2^31
  const int cx = (s[0] + s[1]) * N * O, cy = ix * (s[2] + s[3]) * O, cz = ix * iy * (s[4] + s[5]);
  int r = ((int)blockIdx.x - SB.blk0[b]) * (int)blockDim.x + (int)threadIdx.x;
  int i, j, k;
  if (r < cx) {
    const int l = r % (s[0] + s[1]);
    r /= (s[0] + s[1]);
    i = l < s[0] ? l : M - s[1] + (l - s[0]);
    j = r % N;
    k = r / N;
  } else if ((r -= cx) < cy) {
    i = s[0] + r % ix;
    r /= ix;
    const int l = r % (s[2] + s[3]);
    j = l < s[2] ? l : N - s[3] + (l - s[2]);
    k = r / (s[2] + s[3]);
  } else if ((r -= cy) < cz) {
    i = s[0] + r % ix;
    r /= ix;
    j = s[2] + r % iy;
    const int l = r / iy;
    k = l < s[4] ? l : O - s[5] + (l - s[4]);
  } else {
    return;
  }
  const long q = B.noffset + i + (long)M * j + (long)M * N * k;
  const int v = blockIdx.y;
  const double x = D.fill ? D.val : src.p[v][q];
  for (int t = 0; t < D.nd; t++) D.d[t].p[v][q] = x;
}

// launch skip_copy_kernel over every box of the level, up to kSkipBoxes
// boxes per launch
static int skip_launch(bamhip_level *lev, const SkipDst &D, const Ptrs25 &src, const char *what) {
  SkipBoxes SB{};
  int blk = 0;
  auto flush = [&]() -> int {
    if (SB.n == 0) return BAMHIP_OK;
    SB.blk0[SB.n] = blk;
    skip_copy_kernel<<<dim3((unsigned)blk, (unsigned)D.nv, 1), 256, 0, stream()>>>(D, src, SB);
    BH_HIP(hipGetLastError());
    SB.n = 0;
    blk = 0;
    return BAMHIP_OK;
  };
  int e;
  for (const DevBox &B : lev->boxes) {
    const long ix = B.m - B.skip[0] - B.skip[1], iy = B.n - B.skip[2] - B.skip[3];
    if (ix < 1 || iy < 1 || B.o - B.skip[4] - B.skip[5] < 1)
      return set_error(BAMHIP_EINVAL, "%s: a box has no inner points", what);
    const long total = (long)(B.skip[0] + B.skip[1]) * B.n * B.o + ix * (B.skip[2] + B.skip[3]) * B.o +
                       ix * iy * (B.skip[4] + B.skip[5]);
    if (total <= 0) continue;
    if (total > (long)INT_MAX - 255)
      return set_error(BAMHIP_EINVAL, "%s: a box has %ld skip points (at most 2^31 - 256)", what, total);
    if (SB.n == kSkipBoxes && (e = flush())) return e;
    SB.box[SB.n] = B;
    SB.blk0[SB.n++] = blk;
    blk += (int)((total + 255) / 256);
  }
  return flush();
}

// periodic copy of g layers along direction d for nv variables
// (periodic.c:476-560 with g = 2): layers [0,g) <- [n-2g, n-g), [n-g, n) <- [g, 2g)
__global__ void periodic_kernel(Ptrs25 u, int nv, DevBox B, int d, int g) {
  const int dims[3] = {B.m, B.n, B.o};
  const long st[3] = {1, B.m, (long)B.m * B.n};
  const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
  const long plane = (long)dims[e1] * dims[e2];
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= plane * g) return;
  const int l = (int)(t / plane);
  const long r = t % plane;
  const int a = (int)(r % dims[e1]), b = (int)(r / dims[e1]);
  const long base = B.noffset + a * st[e1] + b * st[e2];
  const int nd = dims[d];
  const int v = blockIdx.y;
  if (v >= nv) return;
  double *p = u.p[v];
  p[base + l * st[d]] = p[base + (nd - 2 * g + l) * st[d]];
  p[base + (nd - g + l) * st[d]] = p[base + (g + l) * st[d]];
}

// All periodic ghost points of a box in one pass: a point within g layers
// of a periodic face takes the value of its image (i -> i +- (n - 2g)) in
// every periodic direction at once, which equals the reference's sequence
// of x, y, z slab copies (edges and corners included).  One thread per
// ghost point and variable; the ghost set is enumerated as 6 slabs without
// overlap: x slabs (full yz planes), y slabs (interior x), z slabs
// (interior x and y).
__global__ void periodic_all_kernel(Ptrs25 u, DevBox B, int g, int px, int py, int pz) {
  const int M = B.m, N = B.n, O = B.o;
  const int ix = px ? M - 2 * g : M, iy = py ? N - 2 * g : N;   // extent of the non-ghost part
  const long sx = px ? (long)g * N * O : 0;                     // points per x slab
  const long sy = py ? (long)g * ix * O : 0;
  const long sz = pz ? (long)g * ix * iy : 0;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long r = t;
  int i, j, k;
  if (r < 2 * sx) {
    const int side = (int)(r / sx);
    r %= sx;
    const int l = (int)(r % g);
    r /= g;
    j = (int)(r % N);
    k = (int)(r / N);
    i = side ? M - g + l : l;
  } else if ((r -= 2 * sx) < 2 * sy) {
    const int side = (int)(r / sy);
    r %= sy;
    const int l = (int)(r % g);
    r /= g;
    i = (px ? g : 0) + (int)(r % ix);
    k = (int)(r / ix);
    j = side ? N - g + l : l;
  } else if ((r -= 2 * sy) < 2 * sz) {
    const int side = (int)(r / sz);
    r %= sz;
    const int l = (int)(r % g);
    r /= g;
    i = (px ? g : 0) + (int)(r % ix);
    j = (py ? g : 0) + (int)(r / ix);
    k = side ? O - g + l : l;
  } else {
    return;
  }
  auto img = [g](int a, int n, int per) {
    if (!per) return a;
    if (a < g) return a + (n - 2 * g);
    if (a >= n - g) return a - (n - 2 * g);
    return a;
  };
  const long dst = B.noffset + i + (long)M * (j + (long)N * k);
  const long src = B.noffset + img(i, M, px) + (long)M * (img(j, N, py) + (long)N * img(k, O, pz));
  double *p = u.p[blockIdx.y];
  p[dst] = p[src];
}

// ---------------------------------------------------------------- algcon
// z4_algcon.m:14-48 at one point
__device__ __forceinline__ void algcon_point(double g[6], double A[6], int subtractA, int normalizedetg) {
  auto det = [&]() {
    return g[0] * (g[3] * g[5] - g[4] * g[4]) - g[1] * (g[1] * g[5] - g[4] * g[2]) + g[2] * (g[1] * g[4] - g[3] * g[2]);
  };
  double detg = det();
  if (detg <= 0.0) detg = 1.0;
  if (normalizedetg) {
    const double e = detg - 1.0;
    const double aux = (fabs(e) < 1e-12) ? 1.0 - e / 3.0 : pow(detg, -1.0 / 3.0);
#pragma unroll
    for (int c = 0; c < 6; c++) g[c] *= aux;
  }
  detg = det();
  if (detg <= 0.0) detg = 1.0;
  const double a0 = g[3] * g[5] - g[4] * g[4], a1 = g[2] * g[4] - g[1] * g[5], a2 = g[1] * g[4] - g[2] * g[3];
  const double a3 = g[0] * g[5] - g[2] * g[2], a4 = g[1] * g[2] - g[0] * g[4], a5 = g[0] * g[3] - g[1] * g[1];
  const double trA = (a0 * A[0] + a3 * A[3] + a5 * A[5] + 2.0 * (a1 * A[1] + a2 * A[2] + a4 * A[4])) / detg;
  if (subtractA) {
#pragma unroll
    for (int c = 0; c < 6; c++) A[c] -= trA / 3.0 * g[c];
  }
}

// z4_algcon.m:14-48 on all points
__global__ void algcon_kernel(Ptrs25 g6, Ptrs25 A6, long n, int subtractA, int normalizedetg) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    double g[6], A[6];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      g[c] = g6.p[c][q];
      A[c] = A6.p[c][q];
    }
    algcon_point(g, A, subtractA, normalizedetg);
#pragma unroll
    for (int c = 0; c < 6; c++) {
      g6.p[c][q] = g[c];
      A6.p[c][q] = A[c];
    }
  }
}

// First RK4 stage update of a step with the list operations before it folded
// in (evolve.c:298-340, rk.c:225-240): with bo, upp = ba*uc + bb*up (the
// time-interpolation begin, orders 2 and 3, as bamhip_vladd forms it); then
// up = uc; uc += wc*r; v = up + wv*r.  Each value is the same expression as
// in the separate vladd and rk4_update passes, so the results are bitwise
// theirs.
__global__ void rk4_first_kernel(Ptrs25 uc, Ptrs25 up, Ptrs25 r, Ptrs25 v, Ptrs25 upp, double wc, double wv, int bo,
                                 double ba, double bb, long n) {
  const int k = blockIdx.y;
  double *__restrict__ c = uc.p[k];
  double *__restrict__ p = up.p[k];
  const double *__restrict__ rr = r.p[k];
  double *__restrict__ vv = v.p[k];
  double *__restrict__ pp = upp.p[k];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double ci = c[i];
    if (bo) {
      double s = 0.0;
      if (ba != 0.0) s += ba * ci;
      if (bb != 0.0) s += bb * p[i];
      pp[i] = s;
    }
    p[i] = ci;
    const double ri = rr[i];
    double cn = ci;
    cn += wc * ri;
    c[i] = cn;
    vv[i] = ci + wv * ri;
  }
}

// Last RK4 stage update of a step with the list operations after it folded
// in (rk.c:250-252, evolve.c:367-380, z4.c:49-71): uc += wc*r; with bo,
// upp += bc*uc (the time-interpolation end, as bamhip_vladd forms it); with
// algcon, z4_algcon on the new point.  One thread per point, all 25 fields.
__global__ void rk4_last_kernel(Ptrs25 uc, Ptrs25 r, Ptrs25 upp, double wc, int bo, double bc, int algcon,
                                int subtractA, int normalizedetg, long n) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    double u[25];
#pragma unroll
    for (int k = 0; k < 25; k++) {
      u[k] = uc.p[k][q];
      u[k] += wc * r.p[k][q];
    }
    if (bo) {
#pragma unroll
      for (int k = 0; k < 25; k++) {
        double s = 0.0;
        s += 1.0 * upp.p[k][q];
        s += bc * u[k];
        upp.p[k][q] = s;
      }
    }
    if (algcon) {
      double g[6], A[6];
#pragma unroll
      for (int c = 0; c < 6; c++) {
        g[c] = u[c];
        A[c] = u[7 + c];
      }
      algcon_point(g, A, subtractA, normalizedetg);
#pragma unroll
      for (int c = 0; c < 6; c++) {
        u[c] = g[c];
        u[7 + c] = A[c];
      }
    }
#pragma unroll
    for (int k = 0; k < 25; k++) uc.p[k][q] = u[k];
  }
}

// -------------------------------------------------- Hamiltonian constraint
// z4_to_adm (z4.c:93-124) into 12 scratch fields, then the order-4 (order 2
// one point from the edge) Hamiltonian constraint of adm_constraints_N.m at
// inner points, then a per-block partial sum of H^2 and a count.
__global__ void z4_to_adm_kernel(Ptrs25 u, double *adm, long n, double chipsipower) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const double psi = pow(u.p[6][q], 1.0 / chipsipower);
    const double psi4 = pow(psi, 4.0);
    const double Kh = u.p[13][q];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      const double gg = psi4 * u.p[c][q];
      adm[c * n + q] = gg;
      adm[(6 + c) * n + q] = psi4 * u.p[7 + c][q] + Kh * gg / 3.0;
    }
  }
}

__device__ __forceinline__ int syi(int a, int b) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo * (5 - lo) / 2 + hi;
}

// Hamiltonian and momentum constraints (adm_constraints_N.m:14-116, usepsi
// off, no matter) at the inner points of one box; per block the sums of H^2,
// mom_x^2, mom_y^2, mom_z^2 and the point count go to partial[5 * block].
struct MomPtrs {
  double *p[3];
};
__global__ void __launch_bounds__(256) ham_kernel(const double *__restrict__ adm, long n, DevBox B, double *ham,
                                                  MomPtrs mom, double *partial) {
  __shared__ double red[5][256];
  const int M = B.m, N = B.n, O = B.o;
  const int i = B.skip[0] + blockIdx.x * 64 + threadIdx.x;
  const int j = B.skip[2] + blockIdx.y * 4 + threadIdx.y;
  const int k = B.skip[4] + blockIdx.z;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  double sums[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (i <= M - 1 - B.skip[1] && j <= N - 1 - B.skip[3] && k <= O - 1 - B.skip[5]) {
    const long st[3] = {1, M, (long)M * N};
    const double h[3] = {B.dx, B.dy, B.dz};
    const long q = B.noffset + i + st[1] * j + st[2] * k;
    const bool o2 = (i == 1 || j == 1 || k == 1 || i == M - 2 || j == N - 2 || k == O - 2);
    double g[6], Kd[6], dg[3][6], ddg[6][6], dK[3][6];
    for (int c6 = 0; c6 < 6; c6++) {
      const double *f = adm + c6 * n, *kf = adm + (6 + c6) * n;
      g[c6] = f[q];
      Kd[c6] = kf[q];
      for (int e = 0; e < 3; e++) {
        const long s = st[e];
        dg[e][c6] = o2 ? (f[q + s] - f[q - s]) / (2 * h[e])
                       : (f[q - 2 * s] - 8 * f[q - s] + 8 * f[q + s] - f[q + 2 * s]) / (12 * h[e]);
        dK[e][c6] = o2 ? (kf[q + s] - kf[q - s]) / (2 * h[e])
                       : (kf[q - 2 * s] - 8 * kf[q - s] + 8 * kf[q + s] - kf[q + 2 * s]) / (12 * h[e]);
      }
      for (int e = 0; e < 3; e++)
        for (int e2 = e; e2 < 3; e2++) {
          const long s = st[e], s2 = st[e2];
          double v;
          if (e == e2)
            v = o2 ? (f[q - s] - 2 * f[q] + f[q + s]) / (h[e] * h[e])
                   : (-f[q - 2 * s] + 16 * f[q - s] - 30 * f[q] + 16 * f[q + s] - f[q + 2 * s]) / (12 * h[e] * h[e]);
          else if (o2)
            v = (f[q + s + s2] - f[q + s - s2] - f[q - s + s2] + f[q - s - s2]) / (4 * h[e] * h[e2]);
          else {
            const int of[4] = {-2, -1, 1, 2};
            const double w[4] = {1, -8, 8, -1};
            double t = 0;
            for (int a = 0; a < 4; a++)
              for (int b2 = 0; b2 < 4; b2++) t += w[a] * w[b2] * f[q + of[a] * s + of[b2] * s2];
            v = t / (144 * h[e] * h[e2]);
          }
          ddg[syi(e, e2)][c6] = v;
        }
    }
    double gi[6];
    gi[0] = g[3] * g[5] - g[4] * g[4];
    gi[1] = g[2] * g[4] - g[1] * g[5];
    gi[2] = g[1] * g[4] - g[2] * g[3];
    gi[3] = g[0] * g[5] - g[2] * g[2];
    gi[4] = g[1] * g[2] - g[0] * g[4];
    gi[5] = g[0] * g[3] - g[1] * g[1];
    const double id = 1.0 / (g[0] * gi[0] + g[1] * gi[1] + g[2] * gi[2]);
    for (int c6 = 0; c6 < 6; c6++) gi[c6] *= id;
    double gdo[3][6], gam[3][6];
    for (int cc = 0; cc < 3; cc++)
      for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++)
          gdo[cc][syi(a, b)] = 0.5 * (dg[a][syi(b, cc)] + dg[b][syi(a, cc)] - dg[cc][syi(a, b)]);
    for (int cc = 0; cc < 3; cc++)
      for (int ab = 0; ab < 6; ab++)
        gam[cc][ab] = gi[syi(cc, 0)] * gdo[0][ab] + gi[syi(cc, 1)] * gdo[1][ab] + gi[syi(cc, 2)] * gdo[2][ab];
    double R = 0.0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double Rab = 0.0;
        for (int cc = 0; cc < 3; cc++)
          for (int d = 0; d < 3; d++) {
            double t = 0.5 * (-ddg[syi(cc, d)][syi(a, b)] - ddg[syi(a, b)][syi(cc, d)] + ddg[syi(a, cc)][syi(b, d)] +
                              ddg[syi(b, cc)][syi(a, d)]);
            for (int e = 0; e < 3; e++) t += gam[e][syi(a, cc)] * gdo[e][syi(b, d)] - gam[e][syi(a, b)] * gdo[e][syi(cc, d)];
            Rab += gi[syi(cc, d)] * t;
          }
        R += gi[syi(a, b)] * Rab;
      }
    double Kud[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        Kud[a][b] = gi[syi(a, 0)] * Kd[syi(0, b)] + gi[syi(a, 1)] * Kd[syi(1, b)] + gi[syi(a, 2)] * Kd[syi(2, b)];
    double trK = Kud[0][0] + Kud[1][1] + Kud[2][2], KK = 0.0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) KK += Kud[a][b] * Kud[b][a];
    const double H = R + trK * trK - KK;
    // D_a K_bc = d_a K_bc - Gamma^d_ab K_dc - Gamma^d_ac K_bd, raised index
    // D^a K_bc = g^ad D_d K_bc; mom^a = g^ab D^c K_bc - g^bc D^a K_bc
    double cdK[3][6];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        for (int cc = b; cc < 3; cc++) {
          double t = dK[a][syi(b, cc)];
          for (int d = 0; d < 3; d++) t -= gam[d][syi(a, b)] * Kd[syi(d, cc)] + gam[d][syi(a, cc)] * Kd[syi(b, d)];
          cdK[a][syi(b, cc)] = t;
        }
    double momu[3];
    for (int a = 0; a < 3; a++) {
      double t = 0.0;
      for (int b = 0; b < 3; b++)
        for (int cc = 0; cc < 3; cc++) {
          double up1 = 0.0, up2 = 0.0;  // D^c K_bc and D^a K_bc
          for (int d = 0; d < 3; d++) {
            up1 += gi[syi(cc, d)] * cdK[d][syi(b, cc)];
            up2 += gi[syi(a, d)] * cdK[d][syi(b, cc)];
          }
          t += gi[syi(a, b)] * up1 - gi[syi(b, cc)] * up2;
        }
      momu[a] = t;
    }
    if (ham) ham[q] = H;
    sums[0] = H * H;
    for (int a = 0; a < 3; a++) {
      const double Ma = g[syi(a, 0)] * momu[0] + g[syi(a, 1)] * momu[1] + g[syi(a, 2)] * momu[2];
      if (mom.p[a]) mom.p[a][q] = Ma;
      sums[1 + a] = Ma * Ma;
    }
    sums[4] = 1.0;
  }
  for (int v = 0; v < 5; v++) red[v][tid] = sums[v];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s)
      for (int v = 0; v < 5; v++) red[v][tid] += red[v][tid + s];
    __syncthreads();
  }
  if (tid == 0) {
    const long bid = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    for (int v = 0; v < 5; v++) partial[5 * bid + v] = red[v][0];
  }
}

static int get25(bamhip_level *lev, const int *idx, int n, Ptrs25 &P, const char *what) {
  for (int v = 0; v < 25; v++) P.p[v] = nullptr;
  return gather_ptrs(lev, idx, n, P.p, what);
}

}  // namespace bamhip

using namespace bamhip;

extern "C" {

int bamhip_level_fill(bamhip_level *lev, int var, double value) {
  double *p = bamhip_level_ptr(lev, var);
  if (!p) return set_error(BAMHIP_EINVAL, "bamhip_level_fill: variable %d not enabled", var);
  fill_kernel<<<stream_grid(lev->npoints, 1).x, 256, 0, stream()>>>(p, lev->npoints, value);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

int bamhip_vladd(bamhip_level *lev, int n, const int *r, double ca, const int *a, double cb, const int *b) {
  if (!lev || n < 1 || n > 25) return set_error(BAMHIP_EINVAL, "bamhip_vladd: n=%d (1..25)", n);
  Ptrs25 R, A, Bp;
  int e;
  if ((e = get25(lev, r, n, R, "bamhip_vladd(r)"))) return e;
  if (ca != 0.0 && (e = get25(lev, a, n, A, "bamhip_vladd(a)"))) return e;
  if (cb != 0.0 && (e = get25(lev, b, n, Bp, "bamhip_vladd(b)"))) return e;
  if (ca == 0.0) A = R;
  if (cb == 0.0) Bp = R;
  TimerScope t("vladd");
  vladd_kernel<<<stream_grid(lev->npoints, n), 256, 0, stream()>>>(R, A, Bp, ca, cb, lev->npoints);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

}  // extern "C"

// The boundary of one RHS evaluation or fused RK4 stage on every box:
// radiative condition on PHYBOUND faces, periodic copy of pgh layers of the
// lists in per_lists on PERBOUND face pairs.
static int boundary_impl(bamhip_level *lev, const bamhip_z4c_params *par, Ptrs25 N, Ptrs25 Pp, Ptrs25 U, double c,
                         const RkFuse &F, const Ptrs25 *per_lists, int nper, int pgh, const char *what) {
  RadInfo ri;
  double vg = 1.0;
  if (par->oploglapse) vg = sqrt(par->lapseharmonicf);
  else if (par->oploglapse2) vg = sqrt(4.0 / 3.0);
  for (int v = 0; v < 25; v++) {
    ri.var0[v] = (v == 0 || v == 3 || v == 5 || v == 6 || v == 18) ? 1.0 : 0.0;
    ri.v[v] = (v == 6 || v == 18 || v == 13) ? vg : 1.0;
  }
  TimerScope t("set_boundary");
  for (const DevBox &B : lev->boxes) {
    bool anyphy = false, anyper = false;
    for (int f = 0; f < 6; f++) {
      anyphy |= B.bflag[f] == BAMHIP_PHYBOUND;
      anyper |= B.bflag[f] == BAMHIP_PERBOUND;
    }
    if (anyphy) {
      if (B.m < 3 || B.n < 3 || B.o < 3) return set_error(BAMHIP_EINVAL, "%s: box too small for the radiative condition", what);
      const long total = 2L * ((long)B.n * B.o + (long)(B.m - 2) * B.o + (long)(B.m - 2) * (B.n - 2));
      radiative_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream()>>>(N, Pp, U, ri, B, c, F);
      BH_HIP(hipGetLastError());
    }
    if (anyper) {
      if (pgh < 1) return set_error(BAMHIP_EINVAL, "%s: periodic faces need pgh >= 1", what);
      const int dims[3] = {B.m, B.n, B.o};
      int per[3];
      for (int d = 0; d < 3; d++) {
        per[d] = (B.bflag[2 * d] == BAMHIP_PERBOUND && B.bflag[2 * d + 1] == BAMHIP_PERBOUND);
        if (per[d] && dims[d] < 3 * pgh)
          return set_error(BAMHIP_EINVAL, "%s: periodic direction %d has %d < 3*%d points", what, d, dims[d], pgh);
      }
      const long ix = per[0] ? B.m - 2 * pgh : B.m, iy = per[1] ? B.n - 2 * pgh : B.n;
      const long total = 2 * ((per[0] ? (long)pgh * B.n * B.o : 0) + (per[1] ? (long)pgh * ix * B.o : 0) +
                              (per[2] ? (long)pgh * ix * iy : 0));
      if (total > 0) {
        dim3 grid((unsigned)((total + 255) / 256), 25, 1);
        for (int l = 0; l < nper; l++) {
          periodic_all_kernel<<<grid, 256, 0, stream()>>>(per_lists[l], B, pgh, per[0], per[1], per[2]);
          BH_HIP(hipGetLastError());
        }
      }
    }
  }
  return BAMHIP_OK;
}

extern "C" {

int bamhip_z4c_set_boundary(bamhip_level *lev, const bamhip_z4c_params *par, const int *unew, const int *upre,
                            double c, const int *ucur, int pgh) {
  if (!lev || !par) return set_error(BAMHIP_EINVAL, "bamhip_z4c_set_boundary: NULL argument");
  Ptrs25 N, Pp, U;
  int e;
  if ((e = get25(lev, unew, 25, N, "bamhip_z4c_set_boundary(unew)"))) return e;
  if ((e = get25(lev, ucur, 25, U, "bamhip_z4c_set_boundary(ucur)"))) return e;
  if (c != 0.0) {
    if ((e = get25(lev, upre, 25, Pp, "bamhip_z4c_set_boundary(upre)"))) return e;
  } else {
    Pp = U;
  }
  RkFuse F{};
  return boundary_impl(lev, par, N, Pp, U, c, F, &N, 1, pgh, "bamhip_z4c_set_boundary");
}

int bamhip_z4c_bc_rk4_stage(bamhip_level *lev, const bamhip_z4c_params *par, const int *src, const int *acc,
                            const int *accout, const int *base, const int *vout, const int *copy, double wc,
                            double wv, int pgh, int fill_skip) {
  if (!lev || !par || !src || !acc || !accout || !base)
    return set_error(BAMHIP_EINVAL, "bamhip_z4c_bc_rk4_stage: NULL argument");
  Ptrs25 U, Pp, V;
  RkFuse F{};
  int e;
  if ((e = get25(lev, src, 25, U, "bamhip_z4c_bc_rk4_stage(src)"))) return e;
  if ((e = get25(lev, acc, 25, F.ucin, "bamhip_z4c_bc_rk4_stage(acc)"))) return e;
  if ((e = get25(lev, accout, 25, F.uc, "bamhip_z4c_bc_rk4_stage(accout)"))) return e;
  if ((e = get25(lev, base, 25, Pp, "bamhip_z4c_bc_rk4_stage(base)"))) return e;
  if (copy && (e = get25(lev, copy, 25, F.copy, "bamhip_z4c_bc_rk4_stage(copy)"))) return e;
  F.rk = 1;
  F.wc = wc;
  F.wv = wv;
  F.has_v = vout != nullptr;
  F.has_cp = copy != nullptr;
  if (vout) {
    if ((e = get25(lev, vout, 25, V, "bamhip_z4c_bc_rk4_stage(vout)"))) return e;
  } else {
    V = Pp;
  }
  // Points no kernel computes (RHS 0 there): the next state and the copy are
  // base, and an accumulator written to another list than it is read from is
  // acc + wc * 0 = base there, since acc equals base at those points in the
  // stage that starts the rotation.  Nothing else writes those points during
  // the stages, so the caller may fill them in the first stage only.
  SkipDst D{};
  D.nv = 25;
  if (vout) D.d[D.nd++] = V;
  if (copy) D.d[D.nd++] = F.copy;
  if (accout != acc) {
    bool same = true;
    for (int v = 0; v < 25; v++) same = same && F.uc.p[v] == F.ucin.p[v];
    if (!same) D.d[D.nd++] = F.uc;
  }
  if (fill_skip && D.nd > 0 && (e = skip_launch(lev, D, Pp, "bamhip_z4c_bc_rk4_stage"))) return e;
  const Ptrs25 lists[2] = {F.uc, V};
  return boundary_impl(lev, par, V, Pp, U, 0.0, F, lists, vout ? 2 : 1, pgh, "bamhip_z4c_bc_rk4_stage");
}

int bamhip_skip_fill(bamhip_level *lev, int n, const int *vars, double value) {
  if (!lev || !vars || n < 1 || n > 25) return set_error(BAMHIP_EINVAL, "bamhip_skip_fill: bad arguments");
  SkipDst D{};
  int e;
  if ((e = get25(lev, vars, n, D.d[0], "bamhip_skip_fill"))) return e;
  D.nd = 1;
  D.nv = n;
  D.fill = 1;
  D.val = value;
  return skip_launch(lev, D, D.d[0], "bamhip_skip_fill");
}

int bamhip_z4c_algcon(bamhip_level *lev, const int *g, const int *A, int subtractA, int normalizedetg) {
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_z4c_algcon: NULL level");
  Ptrs25 Gp, Ap;
  int e;
  if ((e = get25(lev, g, 6, Gp, "bamhip_z4c_algcon(g)"))) return e;
  if ((e = get25(lev, A, 6, Ap, "bamhip_z4c_algcon(A)"))) return e;
  TimerScope t("algcon");
  algcon_kernel<<<stream_grid(lev->npoints, 1).x, 256, 0, stream()>>>(Gp, Ap, lev->npoints, subtractA, normalizedetg);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

int bamhip_rk4_stage_update(bamhip_level *lev, int n, const int *uc, const int *up, const int *r, const int *v,
                            double wc, double wv) {
  Ptrs25 C, P, R, V;
  int e;
  if ((e = get25(lev, uc, n, C, "rk4_update(uc)"))) return e;
  if ((e = get25(lev, r, n, R, "rk4_update(r)"))) return e;
  if (wv != 0.0) {
    if ((e = get25(lev, up, n, P, "rk4_update(up)"))) return e;
    if ((e = get25(lev, v, n, V, "rk4_update(v)"))) return e;
  } else {
    P = C;
    V = C;
  }
  TimerScope t("rk4_update");
  rk4_update_kernel<<<stream_grid(lev->npoints, n), 256, 0, stream()>>>(C, P, R, V, wc, wv, lev->npoints);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

int bamhip_rk4_first_update(bamhip_level *lev, const int *uc, const int *up, const int *r, const int *v, double wc,
                            double wv, const int *upp, double ba, double bb) {
  Ptrs25 C, P, R, V, PP;
  int e;
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_rk4_first_update: NULL level");
  if ((e = get25(lev, uc, 25, C, "rk4_first_update(uc)"))) return e;
  if ((e = get25(lev, up, 25, P, "rk4_first_update(up)"))) return e;
  if ((e = get25(lev, r, 25, R, "rk4_first_update(r)"))) return e;
  if ((e = get25(lev, v, 25, V, "rk4_first_update(v)"))) return e;
  if (upp) {
    if ((e = get25(lev, upp, 25, PP, "rk4_first_update(upp)"))) return e;
  } else {
    PP = P;
  }
  TimerScope t("rk4_update");
  rk4_first_kernel<<<stream_grid(lev->npoints, 25), 256, 0, stream()>>>(C, P, R, V, PP, wc, wv, upp != nullptr, ba, bb,
                                                                         lev->npoints);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

int bamhip_rk4_last_update(bamhip_level *lev, const int *uc, const int *r, double wc, const int *upp, double bc,
                           int algcon, int subtractA, int normalizedetg) {
  Ptrs25 C, R, PP;
  int e;
  if (!lev) return set_error(BAMHIP_EINVAL, "bamhip_rk4_last_update: NULL level");
  if ((e = get25(lev, uc, 25, C, "rk4_last_update(uc)"))) return e;
  if ((e = get25(lev, r, 25, R, "rk4_last_update(r)"))) return e;
  if (upp) {
    if ((e = get25(lev, upp, 25, PP, "rk4_last_update(upp)"))) return e;
  } else {
    PP = C;
  }
  TimerScope t("rk4_update");
  rk4_last_kernel<<<stream_grid(lev->npoints, 1).x, 256, 0, stream()>>>(C, R, PP, wc, upp != nullptr, bc, algcon,
                                                                         subtractA, normalizedetg, lev->npoints);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

int bamhip_z4c_constraints_l2(bamhip_level *lev, const int *u, double chipsipower, int ham, const int *mom,
                              double *l2) {
  if (!lev || !l2) return set_error(BAMHIP_EINVAL, "bamhip_z4c_constraints_l2: NULL argument");
  Ptrs25 U;
  int e;
  if ((e = get25(lev, u, 25, U, "bamhip_z4c_constraints_l2(u)"))) return e;
  double *hp = nullptr;
  if (ham >= 0) {
    hp = bamhip_level_ptr(lev, ham);
    if (!hp) return set_error(BAMHIP_EINVAL, "bamhip_z4c_constraints_l2: ham variable %d not enabled", ham);
  }
  MomPtrs mp{{nullptr, nullptr, nullptr}};
  if (mom)
    for (int a = 0; a < 3; a++) {
      mp.p[a] = bamhip_level_ptr(lev, mom[a]);
      if (!mp.p[a])
        return set_error(BAMHIP_EINVAL, "bamhip_z4c_constraints_l2: mom variable %d not enabled", mom[a]);
    }
  double *adm = nullptr;
  BH_HIP(hipMallocAsync((void **)&adm, sizeof(double) * 12 * lev->npoints, stream()));
  z4_to_adm_kernel<<<stream_grid(lev->npoints, 1).x, 256, 0, stream()>>>(U, adm, lev->npoints, chipsipower);
  double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (const DevBox &B : lev->boxes) {
    const int ni = B.m - B.skip[0] - B.skip[1], nj = B.n - B.skip[2] - B.skip[3], nk = B.o - B.skip[4] - B.skip[5];
    if (ni <= 0 || nj <= 0 || nk <= 0) continue;
    dim3 grid((ni + 63) / 64, (nj + 3) / 4, nk);
    const long nb = (long)grid.x * grid.y * grid.z;
    double *partial = nullptr;
    BH_HIP(hipMallocAsync((void **)&partial, sizeof(double) * 5 * nb, stream()));
    ham_kernel<<<grid, dim3(64, 4, 1), 0, stream()>>>(adm, lev->npoints, B, hp, mp, partial);
    std::vector<double> hpart(5 * nb);
    BH_HIP(hipMemcpyAsync(hpart.data(), partial, sizeof(double) * 5 * nb, hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    BH_HIP(hipFreeAsync(partial, stream()));
    for (long b = 0; b < nb; b++)
      for (int v = 0; v < 5; v++) sum[v] += hpart[5 * b + v];
  }
  BH_HIP(hipFreeAsync(adm, stream()));
  BH_HIP(hipStreamSynchronize(stream()));
  for (int v = 0; v < 4; v++) l2[v] = sum[4] > 0 ? sqrt(sum[v] / sum[4]) : 0.0;
  return BAMHIP_OK;
}

int bamhip_z4c_ham_l2(bamhip_level *lev, const int *u, double chipsipower, int ham, double *l2) {
  if (!l2) return set_error(BAMHIP_EINVAL, "bamhip_z4c_ham_l2: NULL argument");
  double all[4];
  const int e = bamhip_z4c_constraints_l2(lev, u, chipsipower, ham, nullptr, all);
  if (e == BAMHIP_OK) *l2 = all[0];
  return e;
}

}  // extern "C"
