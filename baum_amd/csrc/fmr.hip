// fmr.hip -- fixed mesh refinement coupling on gfx950: R/P flags, transfer
// regions, Lagrange prolongation (coarse -> fine buffer shell) and
// restriction (fine -> coarse interior).
//
// Reference: src/main/amr/flag.c:122-262 (setRPflags),
// src/main/amr/box.c:284-309 (parent / point index snapping),
// src/utility/interpolate/interpolate_boxtobox_N.c:585-865 (prolong_boxtobox_N),
// :870-1050 (prolong_varlist_boxtobox_N), :1055-1220 (restrict_boxtobox_N),
// :1224-1370 (restrict_varlist_boxtobox_N), :92-279 (restrict_varlist,
// prolong_varlist), src/main/amr/inject.c:66-110 (restrict_prolong_varlist).
//
// The reference runs three directional sweeps through scratch arrays.  Here
// one thread produces a 2x2x2 octet of fine points (prolongation) or one
// coarse point (restriction) and evaluates the very same 1-D sums, in the
// same order, in registers: each fine value is a fixed function of the
// coarse data, so the result is bitwise the reference's, with no scratch
// arrays and one pass over HBM.  FP contraction is off so that every
// product and sum rounds as in the reference's C.
#pragma clang fp contract(off)

#include <algorithm>
#include <climits>
#include <cmath>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr int kMaxJobs = 8;   // boxes per launch (kernel-argument table)
constexpr int kMaxVars = 32;  // variables per launch
constexpr int kBlock = 256;

struct VarPtrs {
  double *dst[kMaxVars];
  const double *src[kMaxVars];
};

// one fine box of a prolongation launch
struct PJob {
  long foff, coff;          // box offsets in the level arrays
  int fm, fmn, cm, cmn;     // row and plane strides of fine and coarse box
  int ifb[6], ifb0[6], icb[6];
  int nox, noy, noz;        // octets per row, column and plane
  int ntx, nty, ntz;        // octet tiles of the LDS kernel per direction
  long noct;                // octets in the bounding box
  long blk0;                // first block of this box
};
struct PArgs {
  int njobs;
  PJob job[kMaxJobs];
  double c1[8], c2[8];
};

// one fine box of a restriction launch
struct RJob {
  long foff, coff;
  int fm, fmn, cm, cmn;
  int icb[6], ifb[6];
  int ncx, ncy, ncz;
  int ntx, nty, ntz;        // tiles of the LDS kernel per direction
  long npts, blk0;
};
struct RArgs {
  int njobs;
  RJob job[kMaxJobs];
  double c[8];
};

template <class J>
__device__ __forceinline__ int find_job(const J *job, int njobs) {
  int jb = 0;
  while (jb + 1 < njobs && (long)blockIdx.x >= job[jb + 1].blk0) jb++;
  return jb;
}

// Prolongation of one octet: fine points (fi+a, fj+r, fk+s), a,r,s in {0,1},
// from the (N+1)^3 coarse points starting at (cx, cy, cz).  z sweep, then y,
// then x, with weights c1 (left child, window start) and c2 (right child,
// window shifted by one): interpolate_boxtobox_N.c:652-672, 712-731, 770-794.
template <int N>
__global__ void __launch_bounds__(kBlock) prolong_kernel(PArgs a, VarPtrs vp) {
  const int jb = find_job(a.job, a.njobs);
  const PJob &J = a.job[jb];
  const long q = ((long)blockIdx.x - J.blk0) * kBlock + threadIdx.x;
  if (q >= J.noct) return;
  const int ox = (int)(q % J.nox);
  const long t = q / J.nox;
  const int oy = (int)(t % J.noy), oz = (int)(t / J.noy);
  const int fi = J.ifb[0] + 2 * ox, fj = J.ifb[2] + 2 * oy, fk = J.ifb[4] + 2 * oz;
  // an x pair is skipped when its first point lies in fbox0 (:845-850)
  bool w[2][2];
  bool any = false;
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int j = fj + r, k = fk + s;
      const bool valid = j <= J.ifb[3] && k <= J.ifb[5];
      const bool skip = fi >= J.ifb0[0] && fi <= J.ifb0[1] && j >= J.ifb0[2] && j <= J.ifb0[3] &&
                        k >= J.ifb0[4] && k <= J.ifb0[5];
      w[r][s] = valid && !skip;
      any |= w[r][s];
    }
  if (!any) return;
  const int v = blockIdx.y;
  const double *__restrict__ cv =
      vp.src[v] + J.coff + (J.icb[0] + ox) + (long)(J.icb[2] + oy) * J.cm + (long)(J.icb[4] + oz) * J.cmn;
  // The y and x sums are accumulated as their terms become available: the
  // left child's sum takes terms 0..N-1 and the right child's terms 1..N,
  // each in increasing order, which is the reference's order of additions.
  double o[2][2][2] = {};   // [r][s][left/right child in x]
#pragma unroll 1
  for (int ax = 0; ax <= N; ax++) {
    double e[2][2] = {};    // [r][s]
#pragma unroll
    for (int by = 0; by <= N; by++) {
      double u[N + 1];
#pragma unroll
      for (int cz = 0; cz <= N; cz++) u[cz] = cv[ax + (long)by * J.cm + (long)cz * J.cmn];
      double d[2] = {0, 0};
#pragma unroll
      for (int p = 0; p < N; p++) d[0] += a.c1[p] * u[p];
#pragma unroll
      for (int p = 0; p < N; p++) d[1] += a.c2[p] * u[p + 1];
#pragma unroll
      for (int s = 0; s < 2; s++) {
        if (by < N) e[0][s] += a.c1[by] * d[s];
        if (by > 0) e[1][s] += a.c2[by - 1] * d[s];
      }
    }
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
      for (int s = 0; s < 2; s++) {
        if (ax < N) o[r][s][0] += a.c1[ax] * e[r][s];
        if (ax > 0) o[r][s][1] += a.c2[ax - 1] * e[r][s];
      }
  }
  double *__restrict__ fv = vp.dst[v] + J.foff;
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int s = 0; s < 2; s++) {
      if (!w[r][s]) continue;
      const long qf = fi + (long)(fj + r) * J.fm + (long)(fk + s) * J.fmn;
      fv[qf] = o[r][s][0];
      fv[qf + 1] = o[r][s][1];
    }
}

// Stage the SX x SY x SZ box at src (row strides sj, sk) into the LDS array
// t, 0 outside [0, ex) x [0, ey) x [0, ez): every thread's loads are issued
// before its first LDS store, so a block waits for one memory round trip,
// not one per element it stages.
template <int SX, int SY, int SZ>
__device__ __forceinline__ void stage_tile(double *t, const double *__restrict__ src, long sj, long sk, int ex, int ey,
                                           int ez, int tid) {
  constexpr int NT = SX * SY * SZ, NL = (NT + kBlock - 1) / kBlock;
  double r[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const int idx = tid + l * kBlock;
    const int x = idx % SX, y = (idx / SX) % SY, z = idx / (SX * SY);
    r[l] = (idx < NT && x < ex && y < ey && z < ez) ? src[x + (long)y * sj + (long)z * sk] : 0.0;
  }
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const int idx = tid + l * kBlock;
    if (idx < NT) t[idx] = r[l];
  }
}

// Prolongation of a tile of PTX x PTY x PTZ octets through LDS: the coarse
// footprint is staged once, then the z sums of every coarse (x, y) column for
// both children in z, the y sums for both children in y and the x sums, as
// the reference's three sweeps (:652-672, 712-731, 770-794).  Every fine
// value is the same sequence of products and sums as in prolong_kernel (a
// left child sums terms 0..N-1 of its window with c1, a right child terms
// 1..N with c2, each from 0 in increasing order), so the results are bitwise
// equal; the coarse data is read once per tile instead of (N+1)^3 / 8 times
// per fine point.
constexpr int PTX = 8, PTY = 4, PTZ = 4;
template <int N>
__global__ void __launch_bounds__(kBlock) prolong_tile_kernel(PArgs a, VarPtrs vp) {
  constexpr int CX = PTX + N, CY = PTY + N, CZ = PTZ + N;  // coarse footprint
  __shared__ double cin[CZ * CY * CX];
  __shared__ double dz[2 * PTZ * CY * CX];   // [fine z][coarse y][coarse x]
  __shared__ double dy[2 * PTZ * 2 * PTY * CX];  // [fine z][fine y][coarse x]
  const int jb = find_job(a.job, a.njobs);
  const PJob &J = a.job[jb];
  const long b = (long)blockIdx.x - J.blk0;
  const int ox0 = (int)(b % J.ntx) * PTX, oy0 = (int)((b / J.ntx) % J.nty) * PTY,
            oz0 = (int)(b / ((long)J.ntx * J.nty)) * PTZ;
  const int vx = min(PTX, J.nox - ox0), vy = min(PTY, J.noy - oy0), vz = min(PTZ, J.noz - oz0);
  const int tid = threadIdx.x;
  // does any fine point of the tile get a value?  (fbox0 skip, :845-850)
  // The tile's valid points (j <= ifb[3], k <= ifb[5]) form a box, with the
  // x test on each octet's left child fi, and so does the skipped set, so the
  // tile has nothing to set exactly when its valid box is empty or lies in
  // the skipped box: a test on block-uniform values, before any barrier.
  {
    const int xlo = J.ifb[0] + 2 * ox0, xhi = J.ifb[0] + 2 * (ox0 + vx - 1);
    const int ylo = J.ifb[2] + 2 * oy0, yhi = min(J.ifb[2] + 2 * (oy0 + vy) - 1, J.ifb[3]);
    const int zlo = J.ifb[4] + 2 * oz0, zhi = min(J.ifb[4] + 2 * (oz0 + vz) - 1, J.ifb[5]);
    const bool none = yhi < ylo || zhi < zlo;
    const bool inside = xlo >= J.ifb0[0] && xhi <= J.ifb0[1] && ylo >= J.ifb0[2] && yhi <= J.ifb0[3] &&
                        zlo >= J.ifb0[4] && zhi <= J.ifb0[5];
    if (none || inside) return;
  }
  const int v = blockIdx.y;
  const double *__restrict__ cv =
      vp.src[v] + J.coff + (J.icb[0] + ox0) + (long)(J.icb[2] + oy0) * J.cm + (long)(J.icb[4] + oz0) * J.cmn;
  const int ex = vx + N, ey = vy + N, ez = vz + N;  // coarse points the valid octets read
  stage_tile<CX, CY, CZ>(cin, cv, J.cm, J.cmn, ex, ey, ez, tid);
  __syncthreads();
  for (int idx = tid; idx < 2 * PTZ * CY * CX; idx += kBlock) {  // z sums
    const int x = idx % CX, y = (idx / CX) % CY, fz = idx / (CX * CY);
    const int oz = fz / 2, sz = fz % 2;
    double d = 0;
    if (sz == 0) {
#pragma unroll
      for (int p = 0; p < N; p++) d += a.c1[p] * cin[x + CX * (y + CY * (oz + p))];
    } else {
#pragma unroll
      for (int p = 0; p < N; p++) d += a.c2[p] * cin[x + CX * (y + CY * (oz + p + 1))];
    }
    dz[idx] = d;
  }
  __syncthreads();
  for (int idx = tid; idx < 4 * PTZ * PTY * CX; idx += kBlock) {  // y sums
    const int x = idx % CX, fy = (idx / CX) % (2 * PTY), fz = idx / (CX * 2 * PTY);
    const int oy = fy / 2, r = fy % 2;
    double e = 0;
    if (r == 0) {
#pragma unroll
      for (int by = 0; by < N; by++) e += a.c1[by] * dz[x + CX * ((oy + by) + CY * fz)];
    } else {
#pragma unroll
      for (int by = 1; by <= N; by++) e += a.c2[by - 1] * dz[x + CX * ((oy + by) + CY * fz)];
    }
    dy[idx] = e;
  }
  __syncthreads();
  double *__restrict__ fv = vp.dst[v] + J.foff;
  for (int idx = tid; idx < 8 * PTX * PTY * PTZ; idx += kBlock) {  // x sums and stores
    const int fx = idx % (2 * PTX), fy = (idx / (2 * PTX)) % (2 * PTY), fz = idx / (4 * PTX * PTY);
    const int ox = fx / 2, oy = fy / 2, oz = fz / 2, ac = fx % 2;
    if (ox >= vx || oy >= vy || oz >= vz) continue;
    const int fi = J.ifb[0] + 2 * (ox0 + ox), j = J.ifb[2] + 2 * oy0 + fy, k = J.ifb[4] + 2 * oz0 + fz;
    const bool valid = j <= J.ifb[3] && k <= J.ifb[5];
    const bool skip = fi >= J.ifb0[0] && fi <= J.ifb0[1] && j >= J.ifb0[2] && j <= J.ifb0[3] && k >= J.ifb0[4] &&
                      k <= J.ifb0[5];
    if (!valid || skip) continue;
    const double *row = dy + CX * (fy + 2 * PTY * fz);
    double o = 0;
    if (ac == 0) {
#pragma unroll
      for (int ax = 0; ax < N; ax++) o += a.c1[ax] * row[ox + ax];
    } else {
#pragma unroll
      for (int ax = 1; ax <= N; ax++) o += a.c2[ax - 1] * row[ox + ax];
    }
    fv[fi + ac + (long)j * J.fm + (long)k * J.fmn] = o;
  }
}

// Restriction of one coarse point from the N^3 fine points starting at
// (ifb + 2*offset): z sums, then y, then x (:1108-1123, 1145-1158, 1181-1196),
// accumulated in the reference's order.
template <int N>
__global__ void __launch_bounds__(kBlock) restrict_kernel(RArgs a, VarPtrs vp) {
  const int jb = find_job(a.job, a.njobs);
  const RJob &J = a.job[jb];
  const long q = ((long)blockIdx.x - J.blk0) * kBlock + threadIdx.x;
  if (q >= J.npts) return;
  const int cx = (int)(q % J.ncx);
  const long t = q / J.ncx;
  const int cy = (int)(t % J.ncy), cz = (int)(t / J.ncy);
  const int v = blockIdx.y;
  const double *__restrict__ fv = vp.src[v] + J.foff + (J.ifb[0] + 2 * cx) + (long)(J.ifb[2] + 2 * cy) * J.fm +
                                  (long)(J.ifb[4] + 2 * cz) * J.fmn;
  double sum = 0;
#pragma unroll 1
  for (int ax = 0; ax < N; ax++) {
    double dv = 0;
#pragma unroll
    for (int by = 0; by < N; by++) {
      double e = 0;
#pragma unroll
      for (int p = 0; p < N; p++) e += a.c[p] * fv[ax + (long)by * J.fm + (long)p * J.fmn];
      dv += a.c[by] * e;
    }
    sum += a.c[ax] * dv;
  }
  vp.dst[v][J.coff + (J.icb[0] + cx) + (long)(J.icb[2] + cy) * J.cm + (long)(J.icb[4] + cz) * J.cmn] = sum;
}

// Restriction of a tile of RTX x RTY x RTZ coarse points through LDS: the
// fine footprint is staged once, then the z sums of every fine (x, y) column,
// the y sums and the x sums are formed as three sweeps, as the reference
// does (:1108-1196).  Each coarse value is the same sequence of products and
// sums as in restrict_kernel (a z sum e(ax, by) per fine column, then
// dv(ax) = 0 + c0 e(ax, 0) + ..., then 0 + c0 dv(0) + ...), so the results
// are bitwise equal; the fine data is read from HBM once instead of N^3 / 8
// times.
constexpr int RTX = 8, RTY = 8, RTZ = 4;
template <int N>
__global__ void __launch_bounds__(kBlock) restrict_tile_kernel(RArgs a, VarPtrs vp) {
  constexpr int FX = 2 * RTX + N - 2, FY = 2 * RTY + N - 2, FZ = 2 * RTZ + N - 2;
  __shared__ double fin[FZ * FY * FX];
  __shared__ double ez[RTZ * FY * FX];
  __shared__ double ey[RTZ * RTY * FX];
  const int jb = find_job(a.job, a.njobs);
  const RJob &J = a.job[jb];
  const long b = (long)blockIdx.x - J.blk0;
  const int tx0 = (int)(b % J.ntx) * RTX, ty0 = (int)((b / J.ntx) % J.nty) * RTY,
            tz0 = (int)(b / ((long)J.ntx * J.nty)) * RTZ;
  const int vx = min(RTX, J.ncx - tx0), vy = min(RTY, J.ncy - ty0), vz = min(RTZ, J.ncz - tz0);
  // fine points the valid coarse points of the tile read (inside the fine box)
  const int ex = 2 * (vx - 1) + N, eyy = 2 * (vy - 1) + N, ezz = 2 * (vz - 1) + N;
  const int tid = threadIdx.x;
  const int v = blockIdx.y;
  const double *__restrict__ src = vp.src[v] + J.foff + (J.ifb[0] + 2 * tx0) + (long)(J.ifb[2] + 2 * ty0) * J.fm +
                                   (long)(J.ifb[4] + 2 * tz0) * J.fmn;
  stage_tile<FX, FY, FZ>(fin, src, J.fm, J.fmn, ex, eyy, ezz, tid);
  __syncthreads();
  for (int idx = tid; idx < RTZ * FY * FX; idx += kBlock) {  // z sums
    const int x = idx % FX, y = (idx / FX) % FY, cz = idx / (FX * FY);
    double e = 0;
#pragma unroll
    for (int p = 0; p < N; p++) e += a.c[p] * fin[x + FX * (y + FY * (2 * cz + p))];
    ez[idx] = e;
  }
  __syncthreads();
  for (int idx = tid; idx < RTZ * RTY * FX; idx += kBlock) {  // y sums
    const int x = idx % FX, cy = (idx / FX) % RTY, cz = idx / (FX * RTY);
    double dv = 0;
#pragma unroll
    for (int by = 0; by < N; by++) dv += a.c[by] * ez[x + FX * ((2 * cy + by) + FY * cz)];
    ey[idx] = dv;
  }
  __syncthreads();
  const int cx = tid % RTX, cy = (tid / RTX) % RTY, cz = tid / (RTX * RTY);  // x sums, one coarse point each
  if (cx < vx && cy < vy && cz < vz) {
    double sum = 0;
#pragma unroll
    for (int ax = 0; ax < N; ax++) sum += a.c[ax] * ey[(2 * cx + ax) + FX * (cy + RTY * cz)];
    vp.dst[v][J.coff + (J.icb[0] + tx0 + cx) + (long)(J.icb[2] + ty0 + cy) * J.cm +
              (long)(J.icb[4] + tz0 + cz) * J.cmn] = sum;
  }
}
static_assert(RTX * RTY * RTZ == kBlock, "one coarse point per thread in the x sweep");

// flagprolong of one fine box (flag.c:143-206; no symmetry halves, no
// process ghosts): depth from the nearest face if at most nbuffer, else 0
__global__ void pflag_kernel(DevBox F, double *__restrict__ fp, int nbuffer) {
  const long np = (long)F.m * F.n * F.o;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int i = (int)(q % F.m), j = (int)((q / F.m) % F.n), k = (int)(q / ((long)F.m * F.n));
  const double bb0 = F.x0, bb1 = F.x0 + (F.m - 1) * F.dx;
  const double bb2 = F.y0, bb3 = F.y0 + (F.n - 1) * F.dy;
  const double bb4 = F.z0, bb5 = F.z0 + (F.o - 1) * F.dz;
  const double x = F.x0 + i * F.dx, y = F.y0 + j * F.dy, z = F.z0 + k * F.dz;
  const int npi = (int)np;
  int ii = npi, jj = npi, kk = npi, n;
  n = (int)(1.5 + (x - bb0) / F.dx);
  if (n <= nbuffer) ii = n;
  n = (int)(1.5 - (x - bb1) / F.dx);
  if (n <= nbuffer) ii = n;
  n = (int)(1.5 + (y - bb2) / F.dy);
  if (n <= nbuffer) jj = n;
  n = (int)(1.5 - (y - bb3) / F.dy);
  if (n <= nbuffer) jj = n;
  n = (int)(1.5 + (z - bb4) / F.dz);
  if (n <= nbuffer) kk = n;
  n = (int)(1.5 - (z - bb5) / F.dz);
  if (n <= nbuffer) kk = n;
  if (ii > jj) ii = jj;
  if (ii > kk) ii = kk;
  fp[q] = (ii < npi) ? ii : 0;
}

// box_parent_index (box.c:284-287): nearest coarse point of a fine point
__host__ __device__ __forceinline__ int parent_index(double cx0, double cdx, double fx0, double fdx, int fi) {
  return (int)(0.5 + (fx0 + fi * fdx - cx0) / cdx);
}

// box_point_index (box.c:304-307): snap down
__host__ __device__ __forceinline__ int point_index(double X0, double dX, double x0, double dx, int i) {
  return (int)((x0 + i * dx - X0) / dX);
}

// fine indices whose nearest parent is coarse index ic (at most 4 candidates)
__device__ __forceinline__ int children(double cx0, double cdx, double fx0, double fdx, int fm, int ic, int *out) {
  const int t = (int)floor((cx0 + ic * cdx - fx0) / fdx);
  int n = 0;
  for (int f = t - 1; f <= t + 2; f++)
    if (f >= 0 && f < fm && parent_index(cx0, cdx, fx0, fdx, f) == ic) out[n++] = f;
  return n;
}

// flagrestrict on the footprint of one fine box (flag.c:222-262) as a gather:
// a parent of a flag-0 child gets 3, then, visiting the children in the
// reference's loop order, flag >= 3 sets 2 and flag 2 sets 1 (last wins)
__global__ void rflag_kernel(DevBox C, DevBox F, const double *__restrict__ fp, double *__restrict__ fr, int lo0,
                             int lo1, int lo2, int nx, int ny, int nz) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= (long)nx * ny * nz) return;
  const int ic = lo0 + (int)(q % nx), jc = lo1 + (int)((q / nx) % ny), kc = lo2 + (int)(q / ((long)nx * ny));
  int ci[4], cj[4], ck[4];
  const int ni = children(C.x0, C.dx, F.x0, F.dx, F.m, ic, ci);
  const int nj = children(C.y0, C.dy, F.y0, F.dy, F.n, jc, cj);
  const int nk = children(C.z0, C.dz, F.z0, F.dz, F.o, kc, ck);
  if (!ni || !nj || !nk) return;
  double r = fr[ic + (long)C.m * (jc + (long)C.n * kc)];
  bool zero = false;
  double last = -1;
  for (int c = 0; c < nk; c++)
    for (int b = 0; b < nj; b++)
      for (int a = 0; a < ni; a++) {
        const double p = fp[ci[a] + (long)F.m * (cj[b] + (long)F.n * ck[c])];
        if (p == 0) zero = true;
        if (p >= 3) last = 2;
        else if (p == 2) last = 1;
      }
  if (zero) r = 3;
  if (last >= 0) r = last;
  fr[ic + (long)C.m * (jc + (long)C.n * kc)] = r;
}

// bounding boxes of selected / unselected points of an index range of one
// box: out[0..5] selected bbox, out[6..11] unselected bbox, out[12] count.
// mode 0: 0 < f <= zone, 1: f == 1, 2: every point, 3: f >= rflag
struct Scan {
  int m, n;
  int lo[3], cnt[3];
  int mode;
  double zone, rflag;
};

__global__ void scan_init_kernel(int *out, int nsets) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsets * 13) return;
  const int e = q % 13;
  out[q] = (e == 12) ? 0 : ((e % 2 == 0) ? INT_MAX : INT_MIN);
}

__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

__global__ void __launch_bounds__(kBlock) scan_kernel(Scan s, const double *__restrict__ flag, int *out) {
  int bb[12];
  for (int e = 0; e < 12; e++) bb[e] = (e % 2 == 0) ? INT_MAX : INT_MIN;
  int count = 0;
  const long total = (long)s.cnt[0] * s.cnt[1] * s.cnt[2];
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
    const int i = s.lo[0] + (int)(q % s.cnt[0]);
    const int j = s.lo[1] + (int)((q / s.cnt[0]) % s.cnt[1]);
    const int k = s.lo[2] + (int)(q / ((long)s.cnt[0] * s.cnt[1]));
    const double f = flag[i + (long)s.m * (j + (long)s.n * k)];
    bool sel;
    if (s.mode == 0) sel = f > 0 && s.zone >= f;
    else if (s.mode == 1) sel = f == 1;
    else if (s.mode == 2) sel = true;
    else sel = f >= s.rflag;
    int *b = sel ? bb : bb + 6;
    count += sel;
    b[0] = min(b[0], i);
    b[1] = max(b[1], i);
    b[2] = min(b[2], j);
    b[3] = max(b[3], j);
    b[4] = min(b[4], k);
    b[5] = max(b[5], k);
  }
  for (int e = 0; e < 12; e++) bb[e] = (e % 2 == 0) ? wave_min(bb[e]) : wave_max(bb[e]);
  for (int o = 32; o > 0; o >>= 1) count += __shfl_xor(count, o);
  if ((threadIdx.x & 63) == 0) {
    for (int e = 0; e < 12; e++) {
      if (e % 2 == 0) {
        if (bb[e] != INT_MAX) atomicMin(&out[e], bb[e]);
      } else if (bb[e] != INT_MIN) {
        atomicMax(&out[e], bb[e]);
      }
    }
    if (count) atomicAdd(&out[12], count);
  }
}

void lagrange(int n, double x, double *c) {
  for (int i = 0; i < n; i++) {
    double d = 1;
    for (int j = 0; j < n; j++)
      if (j != i) d *= (x - j) / (double)(i - j);
    c[i] = d;
  }
}

// coarse index range covered by fine box F (parents of its corner points)
void footprint(const DevBox &C, const DevBox &F, int *lo, int *hi) {
  lo[0] = parent_index(C.x0, C.dx, F.x0, F.dx, 0);
  hi[0] = parent_index(C.x0, C.dx, F.x0, F.dx, F.m - 1);
  lo[1] = parent_index(C.y0, C.dy, F.y0, F.dy, 0);
  hi[1] = parent_index(C.y0, C.dy, F.y0, F.dy, F.n - 1);
  lo[2] = parent_index(C.z0, C.dz, F.z0, F.dz, 0);
  hi[2] = parent_index(C.z0, C.dz, F.z0, F.dz, F.o - 1);
}

unsigned nblocks(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int gather_list(bamhip_level *lev, const int *idx, int n, double **out, const char *what) {
  for (int i = 0; i < n; i++) {
    const int k = idx[i];
    if (k < 0 || k >= lev->nvariables || !lev->v[k])
      return set_error(BAMHIP_EINVAL, "%s: variable %d (entry %d) is not enabled", what, k, i);
    out[i] = lev->v[k];
  }
  return BAMHIP_OK;
}

int check_order(int order, const char *what) {
  if (order != 2 && order != 4 && order != 6 && order != 8)
    return set_error(BAMHIP_EINVAL, "%s: order_RP=%d (2, 4, 6 or 8)", what, order);
  return BAMHIP_OK;
}

int check_link(bamhip_level *lc, bamhip_level *lf, const char *what) {
  if (!lc || !lf) return set_error(BAMHIP_EINVAL, "%s: NULL level", what);
  if (lf->fmr.lc != lc)
    return set_error(BAMHIP_EINVAL, "%s: levels %d/%d are not linked (call bamhip_set_rp_flags first)", what,
                     lc->l, lf->l);
  return BAMHIP_OK;
}

// Launch one kernel per chunk of kMaxJobs boxes and kMaxVars variables:
// grid.x runs over the boxes' blocks, grid.y over the variables.
template <class Job>
long assign_blocks(Job *dst, const Job *src, int nj, long (*units)(const Job &)) {
  long blk = 0;
  for (int j = 0; j < nj; j++) {
    dst[j] = src[j];
    dst[j].blk0 = blk;
    blk += nblocks(units(dst[j]));
  }
  return blk;
}

long prolong_units(const PJob &J) { return J.noct; }
long prolong_tiles(const PJob &J) { return (long)J.ntx * J.nty * J.ntz * kBlock; }
int g_prolong_lds = 1;  // prolongation through LDS tiles (prolong_tile_kernel)
long restrict_units(const RJob &J) { return J.npts; }
long restrict_tiles(const RJob &J) { return (long)J.ntx * J.nty * J.ntz * kBlock; }

int launch_prolong(PArgs &args, const std::vector<PJob> &jobs, double *const *dst, double *const *src, int n,
                   int order) {
  for (size_t j0 = 0; j0 < jobs.size(); j0 += kMaxJobs) {
    const int nj = (int)std::min<size_t>(kMaxJobs, jobs.size() - j0);
    args.njobs = nj;
    const long blk = assign_blocks(args.job, jobs.data() + j0, nj, g_prolong_lds ? prolong_tiles : prolong_units);
    if (blk == 0) continue;
    for (int v0 = 0; v0 < n; v0 += kMaxVars) {
      const int nv = std::min(kMaxVars, n - v0);
      VarPtrs vp{};
      for (int v = 0; v < nv; v++) {
        vp.dst[v] = dst[v0 + v];
        vp.src[v] = src[v0 + v];
      }
      const dim3 grid((unsigned)blk, (unsigned)nv);
      hipStream_t st = stream();
      if (g_prolong_lds) {
        switch (order) {
          case 2: prolong_tile_kernel<2><<<grid, kBlock, 0, st>>>(args, vp); break;
          case 4: prolong_tile_kernel<4><<<grid, kBlock, 0, st>>>(args, vp); break;
          case 6: prolong_tile_kernel<6><<<grid, kBlock, 0, st>>>(args, vp); break;
          default: prolong_tile_kernel<8><<<grid, kBlock, 0, st>>>(args, vp); break;
        }
      } else {
        switch (order) {
          case 2: prolong_kernel<2><<<grid, kBlock, 0, st>>>(args, vp); break;
          case 4: prolong_kernel<4><<<grid, kBlock, 0, st>>>(args, vp); break;
          case 6: prolong_kernel<6><<<grid, kBlock, 0, st>>>(args, vp); break;
          default: prolong_kernel<8><<<grid, kBlock, 0, st>>>(args, vp); break;
        }
      }
      BH_HIP(hipGetLastError());
    }
  }
  return BAMHIP_OK;
}

int g_restrict_lds = 1;  // restriction through LDS tiles (restrict_tile_kernel)

int launch_restrict(RArgs &args, const std::vector<RJob> &jobs, double *const *dst, double *const *src, int n,
                    int order) {
  for (size_t j0 = 0; j0 < jobs.size(); j0 += kMaxJobs) {
    const int nj = (int)std::min<size_t>(kMaxJobs, jobs.size() - j0);
    args.njobs = nj;
    const long blk = assign_blocks(args.job, jobs.data() + j0, nj, g_restrict_lds ? restrict_tiles : restrict_units);
    if (blk == 0) continue;
    for (int v0 = 0; v0 < n; v0 += kMaxVars) {
      const int nv = std::min(kMaxVars, n - v0);
      VarPtrs vp{};
      for (int v = 0; v < nv; v++) {
        vp.dst[v] = dst[v0 + v];
        vp.src[v] = src[v0 + v];
      }
      const dim3 grid((unsigned)blk, (unsigned)nv);
      hipStream_t st = stream();
      if (g_restrict_lds) {
        switch (order) {
          case 2: restrict_tile_kernel<2><<<grid, kBlock, 0, st>>>(args, vp); break;
          case 4: restrict_tile_kernel<4><<<grid, kBlock, 0, st>>>(args, vp); break;
          case 6: restrict_tile_kernel<6><<<grid, kBlock, 0, st>>>(args, vp); break;
          default: restrict_tile_kernel<8><<<grid, kBlock, 0, st>>>(args, vp); break;
        }
      } else {
        switch (order) {
          case 2: restrict_kernel<2><<<grid, kBlock, 0, st>>>(args, vp); break;
          case 4: restrict_kernel<4><<<grid, kBlock, 0, st>>>(args, vp); break;
          case 6: restrict_kernel<6><<<grid, kBlock, 0, st>>>(args, vp); break;
          default: restrict_kernel<8><<<grid, kBlock, 0, st>>>(args, vp); break;
        }
      }
      BH_HIP(hipGetLastError());
    }
  }
  return BAMHIP_OK;
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" {

int bamhip_set_rp_flags(bamhip_level *lc, bamhip_level *lf, const int *parent, int nbuffer, int fp, int fr) {
  const char *what = "bamhip_set_rp_flags";
  if (!lc || !lf || !parent) return set_error(BAMHIP_EINVAL, "%s: NULL argument", what);
  if (nbuffer < 1) return set_error(BAMHIP_EINVAL, "%s: amr_nbuffer=%d", what, nbuffer);
  if (fp < 0 || fp >= lf->nvariables || !lf->v[fp])
    return set_error(BAMHIP_EINVAL, "%s: flagprolong variable %d not enabled on the fine level", what, fp);
  if (fr < 0 || fr >= lc->nvariables || !lc->v[fr])
    return set_error(BAMHIP_EINVAL, "%s: flagrestrict variable %d not enabled on the coarse level", what, fr);
  const int nf = lf->nboxes;
  for (int b = 0; b < nf; b++) {
    const int pb = parent[b];
    if (pb < 0 || pb >= lc->nboxes) return set_error(BAMHIP_EINVAL, "%s: fine box %d has parent %d", what, b, pb);
    const DevBox &F = lf->boxes[b], &C = lc->boxes[pb];
    if (std::fabs(2 * F.dx - C.dx) > 1e-12 * C.dx || std::fabs(2 * F.dy - C.dy) > 1e-12 * C.dy ||
        std::fabs(2 * F.dz - C.dz) > 1e-12 * C.dz)
      return set_error(BAMHIP_EINVAL, "%s: fine box %d spacing is not half of its parent's", what, b);
    for (int f = 0; f < 6; f++)
      if (F.bflag[f] == BAMHIP_SYMBOUND || F.bflag[f] == BAMHIP_GHOBOUND)
        return set_error(BAMHIP_EINVAL, "%s: fine box %d face %d: symmetry and process ghosts are not supported",
                         what, b, f);
    int lo[3], hi[3];
    footprint(C, F, lo, hi);
    if (lo[0] < 0 || lo[1] < 0 || lo[2] < 0 || hi[0] >= C.m || hi[1] >= C.n || hi[2] >= C.o)
      return set_error(BAMHIP_EINVAL, "%s: fine box %d is not inside coarse box %d", what, b, pb);
  }
  FmrLink &L = lf->fmr;
  L.lc = lc;
  L.parent.assign(parent, parent + nf);
  L.nbuffer = nbuffer;
  L.fp = fp;
  L.fr = fr;
  hipStream_t st = stream();
  {
    TimerScope t("rp_flags");
    for (int b = 0; b < nf; b++) {
      const DevBox &F = lf->boxes[b];
      pflag_kernel<<<nblocks((long)F.m * F.n * F.o), kBlock, 0, st>>>(F, lf->v[fp] + F.noffset, nbuffer);
      BH_HIP(hipGetLastError());
    }
    if (int e = bamhip_level_fill(lc, fr, 0.0)) return e;
    for (int b = 0; b < nf; b++) {
      const DevBox &F = lf->boxes[b], &C = lc->boxes[parent[b]];
      int lo[3], hi[3];
      footprint(C, F, lo, hi);
      const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
      rflag_kernel<<<nblocks((long)nx * ny * nz), kBlock, 0, st>>>(C, F, lf->v[fp] + F.noffset,
                                                                    lc->v[fr] + C.noffset, lo[0], lo[1], lo[2],
                                                                    nx, ny, nz);
      BH_HIP(hipGetLastError());
    }
  }
  // regions: 3 prolongation modes and 2 restriction modes per fine box
  const int nsets = 5 * nf;
  int *d_out = nullptr;
  BH_HIP(hipMalloc(&d_out, sizeof(int) * 13 * nsets));
  scan_init_kernel<<<nblocks(13L * nsets), kBlock, 0, st>>>(d_out, nsets);
  for (int b = 0; b < nf; b++) {
    const DevBox &F = lf->boxes[b], &C = lc->boxes[parent[b]];
    for (int mode = 0; mode < 5; mode++) {
      Scan s;
      const double *flag;
      if (mode < 3) {
        s.m = F.m;
        s.n = F.n;
        s.lo[0] = s.lo[1] = s.lo[2] = 0;
        s.cnt[0] = F.m;
        s.cnt[1] = F.n;
        s.cnt[2] = F.o;
        s.mode = mode;
        flag = lf->v[fp] + F.noffset;
      } else {
        int lo[3], hi[3];
        footprint(C, F, lo, hi);
        s.m = C.m;
        s.n = C.n;
        for (int d = 0; d < 3; d++) {
          s.lo[d] = lo[d];
          s.cnt[d] = hi[d] - lo[d] + 1;
        }
        s.mode = 3;
        flag = lc->v[fr] + C.noffset;
      }
      s.zone = nbuffer;
      s.rflag = (mode == 3) ? 3 : 2;
      const long tot = (long)s.cnt[0] * s.cnt[1] * s.cnt[2];
      const unsigned g = (unsigned)std::min<long>(nblocks(tot), 1024);
      scan_kernel<<<g, kBlock, 0, st>>>(s, flag, d_out + 13 * (b * 5 + mode));
      BH_HIP(hipGetLastError());
    }
  }
  std::vector<int> h(13 * nsets);
  hipError_t e1 = hipMemcpyAsync(h.data(), d_out, sizeof(int) * h.size(), hipMemcpyDeviceToHost, st);
  hipError_t e2 = hipStreamSynchronize(st);
  (void)hipFree(d_out);
  BH_HIP(e1);
  BH_HIP(e2);
  for (int m = 0; m < 3; m++) L.pro[m].assign(nf, ProlongRegion{});
  for (int m = 0; m < 2; m++) L.res[m].assign(nf, RestrictRegion{});
  for (int b = 0; b < nf; b++) {
    for (int mode = 0; mode < 5; mode++) {
      const int *o = &h[13 * (b * 5 + mode)];
      if (mode < 3) {
        ProlongRegion &R = L.pro[mode][b];
        R.n = o[12];
        for (int e = 0; e < 6; e++) {
          R.ifbox[e] = o[e];
          R.ifbox0[e] = o[6 + e];
        }
      } else {
        RestrictRegion &R = L.res[mode - 3][b];
        R.n = o[12];
        for (int e = 0; e < 6; e++) R.icbox[e] = o[e];
      }
    }
  }
  return BAMHIP_OK;
}

int bamhip_prolong_varlist(bamhip_level *lc, bamhip_level *lf, int n, const int *uc, const int *uf, int nbuffer,
                           int order) {
  const char *what = "bamhip_prolong_varlist";
  if (int e = check_link(lc, lf, what)) return e;
  if (int e = check_order(order, what)) return e;
  if (n < 0 || (n > 0 && (!uc || !uf))) return set_error(BAMHIP_EINVAL, "%s: bad variable lists", what);
  if (n == 0) return BAMHIP_OK;
  int mode;
  if (nbuffer >= 3 || nbuffer == -3) mode = 0;
  else if (nbuffer == 1) mode = 1;
  else if (nbuffer == -1) mode = 2;
  else return set_error(BAMHIP_EINVAL, "%s: nbuffer = %d, P needs >= 3, or = 1, -1, -3", what, nbuffer);
  std::vector<double *> src(n), dst(n);
  if (int e = gather_list(lc, uc, n, src.data(), what)) return e;
  if (int e = gather_list(lf, uf, n, dst.data(), what)) return e;
  const FmrLink &L = lf->fmr;
  const int b2 = order / 2;
  std::vector<PJob> jobs;
  for (int b = 0; b < lf->nboxes; b++) {
    const ProlongRegion &R = L.pro[mode][b];
    if (!R.n) continue;
    const DevBox &F = lf->boxes[b], &C = lc->boxes[L.parent[b]];
    PJob J{};
    J.foff = F.noffset;
    J.coff = C.noffset;
    J.fm = F.m;
    J.fmn = F.m * F.n;
    J.cm = C.m;
    J.cmn = C.m * C.n;
    for (int e = 0; e < 6; e++) {
      J.ifb[e] = R.ifbox[e];
      J.ifb0[e] = R.ifbox0[e];
    }
    J.icb[0] = point_index(C.x0, C.dx, F.x0, F.dx, R.ifbox[0]) - b2 + 1;
    J.icb[2] = point_index(C.y0, C.dy, F.y0, F.dy, R.ifbox[2]) - b2 + 1;
    J.icb[4] = point_index(C.z0, C.dz, F.z0, F.dz, R.ifbox[4]) - b2 + 1;
    J.icb[1] = point_index(C.x0, C.dx, F.x0, F.dx, R.ifbox[1]) + b2;
    J.icb[3] = point_index(C.y0, C.dy, F.y0, F.dy, R.ifbox[3]) + b2;
    J.icb[5] = point_index(C.z0, C.dz, F.z0, F.dz, R.ifbox[5]) + b2;
    if (J.icb[0] < 0 || J.icb[2] < 0 || J.icb[4] < 0 || J.icb[1] >= C.m || J.icb[3] >= C.n || J.icb[5] >= C.o)
      return set_error(BAMHIP_EINVAL, "%s: prolongation: boxes are too small, enlarge nxyz (fine box %d)", what, b);
    const int ex = R.ifbox[1] - R.ifbox[0] + 1, ey = R.ifbox[3] - R.ifbox[2] + 1, ez = R.ifbox[5] - R.ifbox[4] + 1;
    if (ex % 2)
      return set_error(BAMHIP_EINVAL, "%s: fine box %d: prolongation region has an odd x extent %d", what, b, ex);
    // the three sweeps cover 2*(icb[1]-icb[0]+1-N) fine points per direction
    const int cover[3] = {2 * (J.icb[1] - J.icb[0] + 1 - order), 2 * (J.icb[3] - J.icb[2] + 1 - order),
                          2 * (J.icb[5] - J.icb[4] + 1 - order)};
    if (cover[0] < ex || cover[1] < ey || cover[2] < ez)
      return set_error(BAMHIP_EINVAL, "%s: fine box %d: region not aligned with its parent", what, b);
    J.nox = ex / 2;
    J.noy = (ey + 1) / 2;
    J.noz = (ez + 1) / 2;
    J.noct = (long)J.nox * J.noy * J.noz;
    J.ntx = (J.nox + PTX - 1) / PTX;
    J.nty = (J.noy + PTY - 1) / PTY;
    J.ntz = (J.noz + PTZ - 1) / PTZ;
    jobs.push_back(J);
  }
  PArgs args{};
  lagrange(order, order / 2.0 - 0.25, args.c1);
  lagrange(order, order / 2.0 - 0.75, args.c2);
  TimerScope t("prolong");
  return launch_prolong(args, jobs, dst.data(), src.data(), n, order);
}

int bamhip_restrict_varlist(bamhip_level *lf, bamhip_level *lc, int n, const int *uf, const int *uc, int nbuffer,
                            int order) {
  const char *what = "bamhip_restrict_varlist";
  if (int e = check_link(lc, lf, what)) return e;
  if (int e = check_order(order, what)) return e;
  if (n < 0 || (n > 0 && (!uc || !uf))) return set_error(BAMHIP_EINVAL, "%s: bad variable lists", what);
  int mode;
  if (nbuffer >= 3) mode = 0;
  else if (nbuffer == 1) mode = 1;
  else return set_error(BAMHIP_EINVAL, "%s: nbuffer = %d, R needs 1 or >= 3", what, nbuffer);
  if (n == 0) return BAMHIP_OK;
  std::vector<double *> src(n), dst(n);
  if (int e = gather_list(lf, uf, n, src.data(), what)) return e;
  if (int e = gather_list(lc, uc, n, dst.data(), what)) return e;
  const FmrLink &L = lf->fmr;
  const int b2 = order / 2;
  std::vector<RJob> jobs;
  for (int b = 0; b < lf->nboxes; b++) {
    const RestrictRegion &R = L.res[mode][b];
    if (!R.n) continue;
    const DevBox &F = lf->boxes[b], &C = lc->boxes[L.parent[b]];
    RJob J{};
    J.foff = F.noffset;
    J.coff = C.noffset;
    J.fm = F.m;
    J.fmn = F.m * F.n;
    J.cm = C.m;
    J.cmn = C.m * C.n;
    for (int e = 0; e < 6; e++) J.icb[e] = R.icbox[e];
    J.ifb[0] = point_index(F.x0, F.dx, C.x0, C.dx, R.icbox[0]) - b2 + 1;
    J.ifb[2] = point_index(F.y0, F.dy, C.y0, C.dy, R.icbox[2]) - b2 + 1;
    J.ifb[4] = point_index(F.z0, F.dz, C.z0, C.dz, R.icbox[4]) - b2 + 1;
    J.ifb[1] = point_index(F.x0, F.dx, C.x0, C.dx, R.icbox[1]) + b2;
    J.ifb[3] = point_index(F.y0, F.dy, C.y0, C.dy, R.icbox[3]) + b2;
    J.ifb[5] = point_index(F.z0, F.dz, C.z0, C.dz, R.icbox[5]) + b2;
    if (J.ifb[0] < 0 || J.ifb[2] < 0 || J.ifb[4] < 0 || J.ifb[1] >= F.m || J.ifb[3] >= F.n || J.ifb[5] >= F.o)
      return set_error(BAMHIP_EINVAL, "%s: restriction: boxes are too small, enlarge nxyz (fine box %d)", what, b);
    J.ncx = R.icbox[1] - R.icbox[0] + 1;
    J.ncy = R.icbox[3] - R.icbox[2] + 1;
    const int ncz = R.icbox[5] - R.icbox[4] + 1;
    // the sweeps read fine points up to ifb[0] + 2*(ncx-1) + N-1 (:1098-1196)
    if (2 * (J.ncx - 1) + order - 1 > J.ifb[1] - J.ifb[0] || 2 * (J.ncy - 1) + order - 1 > J.ifb[3] - J.ifb[2] ||
        2 * (ncz - 1) + order - 1 > J.ifb[5] - J.ifb[4])
      return set_error(BAMHIP_EINVAL, "%s: fine box %d: region not aligned with its parent", what, b);
    J.ncz = ncz;
    J.npts = (long)J.ncx * J.ncy * ncz;
    J.ntx = (J.ncx + RTX - 1) / RTX;
    J.nty = (J.ncy + RTY - 1) / RTY;
    J.ntz = (ncz + RTZ - 1) / RTZ;
    jobs.push_back(J);
  }
  RArgs args{};
  lagrange(order, order / 2.0 - 0.5, args.c);
  TimerScope t("restrict");
  return launch_restrict(args, jobs, dst.data(), src.data(), n, order);
}

int bamhip_restrict_prolong_varlist(bamhip_level *lc, bamhip_level *lf, int n, const int *uc, const int *uf,
                                    int nbuffer, int order) {
  if (int e = bamhip_restrict_varlist(lf, lc, n, uf, uc, nbuffer, order)) return e;
  return bamhip_prolong_varlist(lc, lf, n, uc, uf, nbuffer, order);
}

}  // extern "C"
