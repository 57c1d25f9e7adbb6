// move.hip -- fine levels that follow the punctures (src/main/amr/move.c).
//
// move_box regrids the levels below a coarse level: the nested boxes around
// the punctures are found on the parent (find_nested_bboxes), a level whose
// boxes changed is replaced by a new one (make_level_bboxes, box.c:465-560)
// that is filled by prolongation from the parent everywhere and then by a
// copy of the points it shares with the old level (move_overlap, move.c:
// 712-802).  The copy is one kernel per (old box, new box) pair over the
// shared index block, all variables of a chunk in one launch (blockIdx.y);
// the prolongation is the FMR prolongation with nbuffer = -1.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bamhip_internal.h"

namespace bamhip {
namespace {

constexpr double kEq = 1e-10;  // dequaleps, bam_main.h:19
constexpr int kCopyBlock = 256;
constexpr int kCopyVars = 32;

bool dless(double a, double b) { return a < b - kEq; }
bool dequal(double a, double b) { return !(dless(a, b) || dless(b, a)); }

struct CopyPtrs {
  const double *src[kCopyVars];
  double *dst[kCopyVars];
};
struct CopyArgs {
  long ooff, noff;       // noffset of the old and the new box
  int om, nm;            // x extents
  long omn, nmn;         // x*y extents
  int io[3], in[3];      // first shared index in the old and the new box
  int cnt[3];            // shared points per direction
};

__global__ void __launch_bounds__(kCopyBlock) copy_overlap_kernel(CopyArgs a, CopyPtrs p) {
  const long t = (long)blockIdx.x * kCopyBlock + threadIdx.x;
  const long nxy = (long)a.cnt[0] * a.cnt[1];
  if (t >= nxy * a.cnt[2]) return;
  const int i = (int)(t % a.cnt[0]), j = (int)((t / a.cnt[0]) % a.cnt[1]), k = (int)(t / nxy);
  const long so = a.ooff + (a.io[0] + i) + a.om * (long)(a.io[1] + j) + a.omn * (a.io[2] + k);
  const long dn = a.noff + (a.in[0] + i) + a.nm * (long)(a.in[1] + j) + a.nmn * (a.in[2] + k);
  p.dst[blockIdx.y][dn] = p.src[blockIdx.y][so];
}

// first and last point coordinates of a box per direction (tB::bbox)
void box_bbox(const DevBox &B, double *bb) {
  bb[0] = B.x0, bb[1] = B.x0 + (B.m - 1) * B.dx;
  bb[2] = B.y0, bb[3] = B.y0 + (B.n - 1) * B.dy;
  bb[4] = B.z0, bb[5] = B.z0 + (B.o - 1) * B.dz;
}

// box_overlap_range (move.c:487-533): index ranges of the shared block in
// box a (ia) and box b (ib); false if the boxes do not overlap
bool overlap_range(const DevBox &A, const DevBox &Bx, int *ia, int *ib) {
  double a[6], b[6], o[6];
  box_bbox(A, a);
  box_bbox(Bx, b);
  for (int d = 0; d < 3; d++) {
    o[2 * d] = std::fmax(a[2 * d], b[2 * d]);
    o[2 * d + 1] = std::fmin(a[2 * d + 1], b[2 * d + 1]);
    if (dless(o[2 * d + 1], o[2 * d])) return false;
  }
  const double da[3] = {A.dx, A.dy, A.dz}, db[3] = {Bx.dx, Bx.dy, Bx.dz};
  for (int k = 0; k < 6; k++) {
    ia[k] = (int)std::floor((o[k] - a[(k / 2) * 2] + kEq) / da[k / 2]);
    ib[k] = (int)std::floor((o[k] - b[(k / 2) * 2] + kEq) / db[k / 2]);
  }
  return true;
}

// the coarse bounding box a fine box was made from (inverse of the box
// construction in make_level_bboxes)
void coarse_bbox(const DevBox &F, double *bb) {
  const int ext[3] = {F.m, F.n, F.o};
  const double x0[3] = {F.x0, F.y0, F.z0}, h[3] = {F.dx, F.dy, F.dz};
  for (int d = 0; d < 3; d++) {
    bb[2 * d] = x0[d] + h[d] / 2;
    bb[2 * d + 1] = bb[2 * d] + (ext[d] / 2 - 1) * (2 * h[d]);
  }
}

// the first coarse box that holds fine box F (tB::pr of make_level_bboxes)
int find_parent(const bamhip_level *lc, const DevBox &F) {
  double f[6];
  box_bbox(F, f);
  for (int b = 0; b < lc->nboxes; b++) {
    double c[6];
    box_bbox(lc->boxes[b], c);
    bool in = true;
    for (int d = 0; d < 3; d++)
      if (dless(f[2 * d], c[2 * d]) || dless(c[2 * d + 1], f[2 * d + 1])) in = false;
    if (in) return b;
  }
  return -1;
}

__global__ void or_flag_kernel(int *dst, const int *src) {
  if (*src) *dst = 1;
}

int link_level(bamhip_level *lc, bamhip_level *lf, const bamhip_fmr_hierarchy *H, const char *what) {
  std::vector<int> parent(lf->nboxes);
  for (int b = 0; b < lf->nboxes; b++) {
    parent[b] = find_parent(lc, lf->boxes[b]);
    if (parent[b] < 0)
      return set_error(BAMHIP_EINVAL, "%s: level %d box %d is not inside a box of its parent", what, lf->l, b);
  }
  return bamhip_set_rp_flags(lc, lf, parent.data(), H->nbuffer, H->fp, H->fr);
}

bool list_enabled(const bamhip_level *lev, const int *list) {
  if (!list) return false;
  for (int i = 0; i < BAMHIP_Z4C_NVARS; i++)
    if (list[i] < 0 || list[i] >= lev->nvariables || !lev->v[list[i]]) return false;
  return true;
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" {

int bamhip_nested_box_size(const bamhip_nested_sizes *S, int lfine, int *n) {
  const char *what = "bamhip_nested_box_size";
  if (!S || !n || lfine < 1 || S->nlist < 0 || (S->nlist > 0 && !S->list) || S->nbuffer < 0)
    return set_error(BAMHIP_EINVAL, "%s: bad arguments", what);
  // the coarsest grid's size; ny takes nx, as the reference does
  n[0] = n[1] = S->nx;
  n[2] = S->nz;
  if (n[0] % 2 || n[1] % 2 || n[2] % 2) return set_error(BAMHIP_EINVAL, "%s: nested boxes need even nx, ny, nz", what);
  if (lfine > S->move_lcube && S->move_nxyz) n[0] = n[1] = n[2] = S->move_nxyz;
  if (S->nlist) {
    const int j = (S->nlist - 1 > lfine) ? lfine : S->nlist - 1;
    n[0] = n[1] = n[2] = S->list[j];
    if (n[0] % 4) return set_error(BAMHIP_EINVAL, "%s: box sizes in amr_nxyz have to be multiples of 4", what);
  }
  for (int i = 0; i < 3; i++) {
    if (S->addextra) n[i] += 2 * S->nbuffer;
    // factor two nesting, rounded up to an even number of coarse points
    n[i] = n[i] / 2 + (n[i] / 2) % 2;
    if (n[i] < 2) return set_error(BAMHIP_EINVAL, "%s: box of %d coarse points on level %d", what, n[i], lfine);
  }
  return BAMHIP_OK;
}

int bamhip_find_nested_bboxes(const bamhip_nested_sizes *S, int lfine, int np, const double *pos, const double *h,
                              int *nboxes, double *bbox, int *ibbox) {
  const char *what = "bamhip_find_nested_bboxes";
  if (np < 1 || np > 10 || !pos || !h || !nboxes || !bbox || !ibbox)
    return set_error(BAMHIP_EINVAL, "%s: bad arguments (np = %d)", what, np);
  for (int i = 0; i < 3; i++)
    if (!(h[i] > 0)) return set_error(BAMHIP_EINVAL, "%s: spacing %d is not positive", what, i);
  int n[3];
  if (int e = bamhip_nested_box_size(S, lfine, n)) return e;
  // levels up to amr_move_lcube: one centred cube, every point at the origin
  const bool cube = lfine <= S->move_lcube;
  double bb[10][6];
  for (int j = 0; j < np; j++)
    for (int i = 0; i < 3; i++) {
      const double p0 = cube ? 0.0 : pos[3 * j + i];
      const double p = ((int)(p0 / h[i])) * h[i];  // truncation: symmetric about the origin
      bb[j][2 * i] = p - (n[i] - 1) * h[i] / 2;
      bb[j][2 * i + 1] = p + (n[i] - 1) * h[i] / 2;
    }
  int nb = np;
  for (int i = 0; i < nb - 1; i++)
    for (int j = i + 1; j < nb; j++) {
      bool ov = true;
      for (int d = 0; d < 3; d++)
        if (dless(bb[i][2 * d + 1], bb[j][2 * d]) || dless(bb[j][2 * d + 1], bb[i][2 * d])) ov = false;
      if (!ov) continue;
      for (int d = 0; d < 3; d++) {
        bb[i][2 * d] = std::fmin(bb[i][2 * d], bb[j][2 * d]);
        bb[i][2 * d + 1] = std::fmax(bb[i][2 * d + 1], bb[j][2 * d + 1]);
      }
      for (int k = j; k < nb - 1; k++) std::memcpy(bb[k], bb[k + 1], sizeof bb[k]);
      nb--;
      j = i;  // compare the joined box with all others again
    }
  *nboxes = nb;
  for (int j = 0; j < nb; j++)
    for (int i = 0; i < 6; i++) {
      bbox[6 * j + i] = bb[j][i];
      ibbox[6 * j + i] = (i % 2) ? (int)((bb[j][i] - bb[j][i - 1]) / h[i / 2] + 0.5) : 0;
    }
  return BAMHIP_OK;
}

int bamhip_level_copy_overlap(bamhip_level *lold, bamhip_level *lnew, int n, const int *vars) {
  const char *what = "bamhip_level_copy_overlap";
  if (!lold || !lnew || n < 0 || (n > 0 && !vars)) return set_error(BAMHIP_EINVAL, "%s: bad arguments", what);
  std::vector<const double *> src;
  std::vector<double *> dst;
  for (int i = 0; i < n; i++) {
    const int k = vars[i];
    if (k < 0 || k >= lold->nvariables || k >= lnew->nvariables)
      return set_error(BAMHIP_EINVAL, "%s: variable index %d out of range", what, k);
    if (!lold->v[k] || !lnew->v[k]) continue;  // storage on one level only
    src.push_back(lold->v[k]);
    dst.push_back(lnew->v[k]);
  }
  struct Job {
    CopyArgs a;
    long pts;
  };
  std::vector<Job> jobs;
  for (int i = 0; i < lold->nboxes; i++)
    for (int j = 0; j < lnew->nboxes; j++) {
      const DevBox &A = lold->boxes[i], &Bx = lnew->boxes[j];
      if (std::fabs(A.dx - Bx.dx) > 1e-12 * A.dx || std::fabs(A.dy - Bx.dy) > 1e-12 * A.dy ||
          std::fabs(A.dz - Bx.dz) > 1e-12 * A.dz)
        return set_error(BAMHIP_EINVAL, "%s: old box %d and new box %d differ in spacing", what, i, j);
      const double s[3] = {(Bx.x0 - A.x0) / A.dx, (Bx.y0 - A.y0) / A.dy, (Bx.z0 - A.z0) / A.dz};
      for (int d = 0; d < 3; d++)
        if (std::fabs(s[d] - std::nearbyint(s[d])) > 1e-8)
          return set_error(BAMHIP_EINVAL, "%s: points of old box %d and new box %d are not aligned", what, i, j);
      int io[6], in[6];
      if (!overlap_range(A, Bx, io, in)) continue;
      Job J{};
      const int ea[3] = {A.m, A.n, A.o}, eb[3] = {Bx.m, Bx.n, Bx.o};
      for (int d = 0; d < 3; d++) {
        J.a.io[d] = io[2 * d];
        J.a.in[d] = in[2 * d];
        J.a.cnt[d] = io[2 * d + 1] - io[2 * d] + 1;
        // the kernel's reads and writes stay inside both boxes
        if (J.a.cnt[d] < 1 || io[2 * d] < 0 || in[2 * d] < 0 || io[2 * d] + J.a.cnt[d] > ea[d] ||
            in[2 * d] + J.a.cnt[d] > eb[d])
          return set_error(BAMHIP_EINVAL, "%s: overlap of old box %d and new box %d out of range", what, i, j);
      }
      J.a.ooff = A.noffset;
      J.a.noff = Bx.noffset;
      J.a.om = A.m;
      J.a.omn = (long)A.m * A.n;
      J.a.nm = Bx.m;
      J.a.nmn = (long)Bx.m * Bx.n;
      J.pts = (long)J.a.cnt[0] * J.a.cnt[1] * J.a.cnt[2];
      jobs.push_back(J);
    }
  const int nv = (int)src.size();
  if (!nv || jobs.empty()) return BAMHIP_OK;
  TimerScope t("move_box_copy");
  hipStream_t st = stream();
  // pairs in order on one stream: a later pair overwrites an earlier one
  for (const Job &J : jobs)
    for (int v0 = 0; v0 < nv; v0 += kCopyVars) {
      const int cv = std::min(kCopyVars, nv - v0);
      CopyPtrs p{};
      for (int v = 0; v < cv; v++) {
        p.src[v] = src[v0 + v];
        p.dst[v] = dst[v0 + v];
      }
      const dim3 grid((unsigned)((J.pts + kCopyBlock - 1) / kCopyBlock), (unsigned)cv);
      copy_overlap_kernel<<<grid, kCopyBlock, 0, st>>>(J.a, p);
      BH_HIP(hipGetLastError());
    }
  return BAMHIP_OK;
}

int bamhip_move_box(const bamhip_fmr_hierarchy *H, int l, int *moved) {
  const char *what = "bamhip_move_box";
  if (moved) *moved = 0;
  if (!H || !H->levels || !H->iteration) return set_error(BAMHIP_EINVAL, "%s: NULL hierarchy", what);
  if (H->move_nxyz <= 0) return BAMHIP_OK;  // amr != move
  if (l <= 0 || l > H->lmax) return BAMHIP_OK;
  // wait until level l-1 is the coarsest level of its time
  auto time_of = [H](int k) { return H->iteration[k] * (H->dt0 / (double)(1L << k)); };
  if (l - 2 >= 0 && !dless(time_of(l - 1), time_of(l - 2))) return BAMHIP_OK;
  const int np = H->npunc;
  if (np < 1 || np > 10 || !H->punc) return set_error(BAMHIP_EINVAL, "%s: needs 1 to 10 punctures", what);
  bamhip_nested_sizes S;
  S.nx = H->nx;
  S.nz = H->nz;
  S.move_nxyz = H->move_nxyz;
  S.move_lcube = H->move_lcube;
  S.nlist = H->n_amr_nxyz;
  S.list = H->amr_nxyz;
  S.nbuffer = H->nbuffer;
  S.addextra = H->addextra;
  std::vector<double> punc(3 * np);
  hipStream_t st = stream();
  BH_HIP(hipMemcpyAsync(punc.data(), H->punc, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, st));
  BH_HIP(hipStreamSynchronize(st));
  TimerScope t("move_box");
  bool parent_new = false;
  for (; l <= H->lmax; l++) {
    bamhip_level *lc = H->levels[l - 1], *lold = H->levels[l];
    const double h[3] = {lc->boxes[0].dx, lc->boxes[0].dy, lc->boxes[0].dz};
    int nb = 0;
    double bb[60];
    int ib[60];
    if (int e = bamhip_find_nested_bboxes(&S, l, np, punc.data(), h, &nb, bb, ib)) return e;
    bool same = nb == lold->nboxes;
    for (int b = 0; same && b < nb; b++) {
      double cb[6];
      coarse_bbox(lold->boxes[b], cb);
      for (int i = 0; i < 6; i++)
        if (!dequal(cb[i], bb[6 * b + i])) same = false;
    }
    if (same) {
      if (parent_new)
        if (int e = link_level(lc, lold, H, what)) return e;
      parent_new = false;
      continue;
    }
    // the new level: make_level_bboxes with REFBOUND faces
    std::vector<bamhip_box> boxes(nb);
    for (int b = 0; b < nb; b++) {
      bamhip_box &B = boxes[b];
      const DevBox &O = lold->boxes[0];
      B.m = 2 * (ib[6 * b + 1] + 1);
      B.n = 2 * (ib[6 * b + 3] + 1);
      B.o = 2 * (ib[6 * b + 5] + 1);
      B.dx = h[0] / 2;
      B.dy = h[1] / 2;
      B.dz = h[2] / 2;
      B.x0 = bb[6 * b + 0] - B.dx / 2;
      B.y0 = bb[6 * b + 2] - B.dy / 2;
      B.z0 = bb[6 * b + 4] - B.dz / 2;
      for (int f = 0; f < 6; f++) {
        B.skip[f] = O.bflag[f] == BAMHIP_REFBOUND ? O.skip[f] : 1;
        B.bflag[f] = BAMHIP_REFBOUND;
      }
    }
    bamhip_level *lnew = nullptr;
    if (int e = bamhip_level_create(&lnew, l, nb, boxes.data(), lold->nvariables)) return e;
    int e = BAMHIP_OK;
    for (int v = 0; v < lold->nvariables && !e; v++)
      if (lold->v[v] && !(e = bamhip_level_enable(lnew, v))) e = bamhip_level_fill(lnew, v, 0.0);
    lnew->check_mode = lold->check_mode;
    if (!e) e = link_level(lc, lnew, H, what);
    const int *lists[3] = {H->u, H->up, H->upp};
    std::vector<int> all;
    for (int k = 0; k < 3 && !e; k++) {
      if (!list_enabled(lc, lists[k]) || !list_enabled(lnew, lists[k])) continue;
      TimerScope tp("move_box_prol");
      e = bamhip_prolong_varlist(lc, lnew, BAMHIP_Z4C_NVARS, lists[k], lists[k], -1, H->order_RP);
      all.insert(all.end(), lists[k], lists[k] + BAMHIP_Z4C_NVARS);
    }
    if (!e) e = bamhip_level_copy_overlap(lold, lnew, (int)all.size(), all.data());
    if (e) {
      bamhip_level_free(lnew);
      return e;
    }
    // a non-finite value the old level's deferred check recorded stays
    // reported; the old level's storage goes once the copy has run
    or_flag_kernel<<<1, 1, 0, st>>>(lnew->d_flag, lold->d_flag);
    BH_HIP(hipGetLastError());
    BH_HIP(hipStreamSynchronize(st));
    bamhip_level_free(lold);
    H->levels[l] = lnew;
    parent_new = true;
    if (moved) *moved = 1;
  }
  return BAMHIP_OK;
}

}  // extern "C"
