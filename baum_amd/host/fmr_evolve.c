/* fmr_evolve.c -- Berger-Oliger evolution of a device-resident FMR hierarchy.
 *
 * The C counterpart of advance_levelandsublevels / advance_level
 * (src/main/main/main.c:397-453,456-490) with the level step of evolve()
 * (src/utility/evolve/evolve.c:298-380: time interpolation of the parent
 * state around the RK4 step) and z4_algcon_wrapper after it (EVOLVE order of
 * bam_z4.c:41-44), and restrict_prolong_evolve (src/main/amr/inject.c:
 * 116-132) after each step of a level with a parent: the parent's time-
 * interpolated u_pp is the coarse list unless both levels are at the same
 * time (parentaligned, inject.c:10-19).  Subcycling on every level (amr =
 * bo, lbo = 0).  After the finest level's step and its restriction, the
 * puncture tracker of analyze_level (main.c:419-423; track_puncture,
 * moving_punctures.c:111-196), and move_box for the levels below (main.c:
 * 425-436, bamhip_move_box) when the boxes follow the punctures.
 *
 * The whole recursion issues device work only: the levels run with the
 * non-finite check deferred, so a coarse step can be captured as a graph.
 * The caller reads the flags (bamhip_level_nonfinite) when it wants the
 * reference's errorexit.
 */
#include <stddef.h>

#include "../../include/bamhip.h"

static int advance_level(const bamhip_fmr_hierarchy *H, int l)
{
  bamhip_level *lev = H->levels[l];
  const double dt = H->dt0 / (double)(1L << l);
  const int it = H->iteration[l];
  int e;
  if ((e = bamhip_z4c_rk4_step_bo(lev, &H->params[l], dt, H->u, H->up, H->r, H->v, l == 0 ? H->pgh : 0,
                                   l < H->lmax ? H->upp : NULL, H->timeinterpolation, it, 1)))
    return e;
  H->iteration[l] = it + 1;
  return BAMHIP_OK;
}

int bamhip_advance_levelandsublevels(const bamhip_fmr_hierarchy *H, int l)
{
  int e;
  if (!H || l < 0 || l > H->lmax) return BAMHIP_EINVAL;
  if (H->timeinterpolation < 2 || H->timeinterpolation > 3) return BAMHIP_EINVAL;
  if ((e = advance_level(H, l))) return e;
  if (l < H->lmax) {
    if ((e = bamhip_advance_levelandsublevels(H, l + 1))) return e;
    if ((e = bamhip_advance_levelandsublevels(H, l + 1))) return e;
  }
  if (l > 0) {
    /* parentaligned: level l has caught up with level l-1 */
    const int aligned = H->iteration[l] == 2 * H->iteration[l - 1];
    if ((e = bamhip_restrict_prolong_varlist(H->levels[l - 1], H->levels[l], BAMHIP_Z4C_NVARS,
                                             aligned ? H->u : H->upp, H->u, 3, H->order_RP)))
      return e;
  }
  /* analyze_level: the punctures move with the shift of the finest level */
  if (l == H->lmax && H->npunc > 0) {
    const double dt = H->dt0 / (double)(1L << l);
    if ((e = bamhip_puncture_track(H->levels[l], H->u + 19, H->up + 19, dt, H->order_RP, H->punc_flags, H->npunc,
                                   H->punc)))
      return e;
  }
  /* move boxes (main.c:425-436) */
  if (H->move_nxyz > 0 && (e = bamhip_move_box(H, l + 1, NULL))) return e;
  return BAMHIP_OK;
}

int bamhip_fmr_set_check(const bamhip_fmr_hierarchy *H, int mode)
{
  int l, e;
  if (!H) return BAMHIP_EINVAL;
  for (l = 0; l <= H->lmax; l++)
    if ((e = bamhip_level_set_check(H->levels[l], mode))) return e;
  return BAMHIP_OK;
}

int bamhip_fmr_nonfinite(const bamhip_fmr_hierarchy *H, int reset, int *flag)
{
  int l, e, f, any = 0;
  if (!H) return BAMHIP_EINVAL;
  for (l = 0; l <= H->lmax; l++) {
    f = 0;
    if ((e = bamhip_level_nonfinite(H->levels[l], reset, flag ? &f : NULL))) return e;
    any |= f;
  }
  if (flag) *flag = any;
  return BAMHIP_OK;
}
