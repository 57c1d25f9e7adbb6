/* evolve_rk4.c -- host driver of one RK4 step on a device-resident level.
 *
 * The C counterpart of evolve_rk_rk4 (src/utility/evolve/rk.c:220-257)
 * followed by z4_algcon_wrapper (src/projects/z4/z4.c:49-71): the same
 * 2-register scheme (u_p saved once, r and v per stage) with BAM's stage
 * coefficients, calling the device library through its C ABI only.  The
 * vladdto + vladd pair of every stage is one fused device pass, and the
 * non-finite check of the RHS is deferred to the end of the step so that the
 * four stages run without host synchronisation.
 */
#include "../../include/bamhip.h"

/* 0: separate update pass; 1: update fused into the RHS kernels; 2 (default):
 * fused on levels of at least 2^19 points.  The fused store adds its
 * operands' reads and writes to every RHS block's end, in place of the
 * separate passes.  Measured on one MI355X: fused is faster on the 192^3 box
 * (7.5 M points) and on the levels of the 7-level hierarchy (0.55 to 2.2 M
 * points), slower on a single 64^3 box (0.29 M points). */
static int g_fused = 2;
#define BAMHIP_FUSE_MIN_POINTS (1L << 19)

int bamhip_z4c_set_rk_fusion(int mode)
{
  const int old = g_fused;
  if (mode < 0 || mode > 2) return BAMHIP_EINVAL;
  g_fused = mode;
  return old;
}

int bamhip_z4c_rk4_stages(bamhip_level *lev, const bamhip_z4c_params *par, double dt, const int *uc, const int *up,
                          const int *r, const int *v, int pgh)
{
  /* stage weights of rk4 (rk.c:98-103): b = 1/6 1/3 1/3 1/6, c = 1/2 1/2 1 */
  const double wc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double wv[4] = {dt / 2, dt / 2, dt, 0.0};
  int e, s;
  long npoints = 0;
  if ((e = bamhip_level_npoints(lev, &npoints))) return e;
  if (g_fused == 1 || (g_fused == 2 && npoints >= BAMHIP_FUSE_MIN_POINTS)) {
    /* The update inside the RHS kernels' store, the four lists rotating so
     * that no stage writes a list its stencils read and no pass copies
     * u_p = u_c (U = uc, P = up):
     *   stage 1 reads U:   r = U + wc k, v = U + wv k, P = U
     *   stage 2 reads v:   U = r + wc k, r = P + wv k
     *   stage 3 reads r:   U = U + wc k, v = P + wv k
     *   stage 4 reads v:   U = U + wc k
     * The same sums as the separate passes, so bitwise the same state; P
     * ends as the step's start state.  Stage 1 also sets r, v and P to U at
     * the points no kernel computes, which no later stage writes. */
    const int *src[4] = {uc, v, r, v};
    const int *acc[4] = {uc, r, uc, uc};
    const int *accout[4] = {r, uc, uc, uc};
    const int *base[4] = {uc, up, up, up};
    const int *vout[4] = {v, r, v, 0};
    for (s = 0; s < 4; s++) {
      const int *copy = s == 0 ? up : 0;
      if ((e = bamhip_z4c_rhs_rk4_stage(lev, par, src[s], acc[s], accout[s], base[s], vout[s], copy, wc[s], wv[s],
                                        s == 0 ? r : 0)))
        return e;
      if ((e = bamhip_z4c_bc_rk4_stage(lev, par, src[s], acc[s], accout[s], base[s], vout[s], copy, wc[s], wv[s], pgh,
                                       s == 0)))
        return e;
    }
    /* r held stage states, and u_p at the points no kernel computes; the
     * separate passes of a later step take r = 0 there */
    return bamhip_skip_fill(lev, BAMHIP_Z4C_NVARS, r, 0.0);
  }
  /* u_p = u_c */
  if ((e = bamhip_vladd(lev, BAMHIP_Z4C_NVARS, up, 1.0, uc, 0.0, 0))) return e;
  for (s = 0; s < 4; s++) {
    const int *src = (s == 0) ? uc : v;
    if ((e = bamhip_z4c_rhs(lev, par, r, 0, 0.0, src))) return e;
    if ((e = bamhip_z4c_set_boundary(lev, par, r, 0, 0.0, src, pgh))) return e;
    if ((e = bamhip_rk4_stage_update(lev, BAMHIP_Z4C_NVARS, uc, up, r, v, wc[s], wv[s]))) return e;
  }
  return BAMHIP_OK;
}

/* One RK4 step of a level together with the list operations around it in
 * the Berger-Oliger evolution (evolve.c:298-380, z4.c:49-71): the time
 * interpolation begin of order o into upp (upp may be NULL), the four
 * stages, the time interpolation end and algcon.  When the stages run
 * separately from the RHS (small levels, see bamhip_z4c_set_rk_fusion), the
 * copy u_p = u_c and the interpolation begin are folded into the first stage
 * update, the interpolation end and algcon into the last; elsewhere the
 * operations run one by one.  Bitwise the same either way. */
int bamhip_z4c_rk4_step_bo(bamhip_level *lev, const bamhip_z4c_params *par, double dt, const int *uc, const int *up,
                           const int *r, const int *v, int pgh, const int *upp, int timeinterpolation, int iteration,
                           int algcon)
{
  static const double cend[4] = {0, 0, 1.0 / 2, 3. / 8};
  const double wc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double wv[4] = {dt / 2, dt / 2, dt, 0.0};
  const int o = (iteration + 2 > timeinterpolation) ? timeinterpolation : iteration + 2;
  long npoints = 0;
  int e, s;
  if (upp && (o < 2 || o > 3)) return BAMHIP_EINVAL;
  if ((e = bamhip_level_npoints(lev, &npoints))) return e;
  if (g_fused == 1 || (g_fused == 2 && npoints >= BAMHIP_FUSE_MIN_POINTS)) {
    if (upp && (e = bamhip_bo_interp_begin(lev, timeinterpolation, iteration, BAMHIP_Z4C_NVARS, upp, uc, up, 0, 0, 0)))
      return e;
    if ((e = bamhip_z4c_rk4_stages(lev, par, dt, uc, up, r, v, pgh))) return e;
    if (upp && (e = bamhip_bo_interp_end(lev, timeinterpolation, iteration, BAMHIP_Z4C_NVARS, upp, uc))) return e;
    return algcon ? bamhip_z4c_algcon_list(lev, par, uc) : BAMHIP_OK;
  }
  for (s = 0; s < 4; s++) {
    const int *src = (s == 0) ? uc : v;
    if ((e = bamhip_z4c_rhs(lev, par, r, 0, 0.0, src))) return e;
    if ((e = bamhip_z4c_set_boundary(lev, par, r, 0, 0.0, src, pgh))) return e;
    if (s == 0)
      e = bamhip_rk4_first_update(lev, uc, up, r, v, wc[0], wv[0], upp, o == 2 ? 0.5 : 0.75, o == 2 ? 0.0 : -0.125);
    else if (s < 3)
      e = bamhip_rk4_stage_update(lev, BAMHIP_Z4C_NVARS, uc, up, r, v, wc[s], wv[s]);
    else
      e = bamhip_rk4_last_update(lev, uc, r, wc[3], upp, cend[o], algcon, par->subtractA, par->normalizedetg);
    if (e) return e;
  }
  return BAMHIP_OK;
}

int bamhip_z4c_algcon_list(bamhip_level *lev, const bamhip_z4c_params *par, const int *uc)
{
  int gi[6], ai[6], s;
  for (s = 0; s < 6; s++) {
    gi[s] = uc[s];
    ai[s] = uc[7 + s];
  }
  return bamhip_z4c_algcon(lev, gi, ai, par->subtractA, par->normalizedetg);
}

int bamhip_z4c_rk4_step(bamhip_level *lev, const bamhip_z4c_params *par, double dt, const int *uc, const int *up,
                        const int *r, const int *v, int pgh, int algcon)
{
  int e, flag = 0, mode;
  if ((e = bamhip_level_get_check(lev, &mode))) return e;
  if ((e = bamhip_level_set_check(lev, 0))) return e;
  if ((e = bamhip_level_nonfinite(lev, 1, 0))) goto out;
  if ((e = bamhip_z4c_rk4_stages(lev, par, dt, uc, up, r, v, pgh))) goto out;
  if (algcon && (e = bamhip_z4c_algcon_list(lev, par, uc))) goto out;
  if ((e = bamhip_level_nonfinite(lev, 0, &flag))) goto out;
  if (flag) e = BAMHIP_ENONFINITE;
out:
  bamhip_level_set_check(lev, mode);
  return e;
}
