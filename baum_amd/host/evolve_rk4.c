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

int bamhip_z4c_rk4_stages(bamhip_level *lev, const bamhip_z4c_params *par, double dt, const int *uc, const int *up,
                          const int *r, const int *v, int pgh)
{
  /* stage weights of rk4 (rk.c:98-103): b = 1/6 1/3 1/3 1/6, c = 1/2 1/2 1 */
  const double wc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double wv[4] = {dt / 2, dt / 2, dt, 0.0};
  int e, s;
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
  int e, flag = 0;
  if ((e = bamhip_level_set_check(lev, 0))) return e;
  if ((e = bamhip_level_nonfinite(lev, 1, 0))) goto out;
  if ((e = bamhip_z4c_rk4_stages(lev, par, dt, uc, up, r, v, pgh))) goto out;
  if (algcon && (e = bamhip_z4c_algcon_list(lev, par, uc))) goto out;
  if ((e = bamhip_level_nonfinite(lev, 0, &flag))) goto out;
  if (flag) e = BAMHIP_ENONFINITE;
out:
  bamhip_level_set_check(lev, 1);
  return e;
}
