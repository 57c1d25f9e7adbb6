/* z4c_grhd.c -- the right-hand side of the coupled Z4c + GRHD system.
 *
 * The C counterpart of evolve_rhs (src/utility/evolve/evolve.c:37-111) for
 * physics = z4 matter grhd, each step a device entry point of this library:
 *   rhs_geometry          z4_rhs_movpunc_N             bamhip_z4c_rhs
 *   rhs_geometry_adm      z4_to_adm_vl (z4.c:93-124)   bamhip_z4c_to_adm
 *   rhs_matter            matter_rhs (matter.c:373-482):
 *     MATTER.c2p            (hybrid p-root etc.)       bamhip_grhd_c2p(_learned)
 *     MATTER.set_atm_prerhs grhd_atm_set_prerhs        bamhip_grhd_atm_set_prerhs
 *     matter_rhs_prepare    (matter.c:323-352)         bamhip_vladd
 *     MATTER.sources        grhd_sources               bamhip_grhd_sources
 *     fluxes                matter_fluxes              bamhip_grhd_fluxes
 *   rhs_geometry_source   z4_rhs_add_source            bamhip_z4c_matter_source
 *   set_boundary          (Z4c list)                   bamhip_z4c_set_boundary
 * The matter variables carry no boundary information in the reference's
 * GRHD project, so the boundary is applied to the 25 Z4c fields.  Device
 * work only, apart from the Z4c RHS's non-finite check in check mode 1.
 */
#include <stddef.h>

#include "../../include/bamhip.h"

int bamhip_z4c_grhd_rhs(bamhip_level *lev, const bamhip_z4c_grhd *S, const int *unew, const int *upre, double c,
                        const int *ucur)
{
  int e;
  if (!lev || !S || !unew || !ucur || !S->z4 || !S->c2p || !S->hrsc || !S->adm_g || !S->adm_K || !S->prim ||
      !S->madm || !S->status || !S->iters)
    return BAMHIP_EINVAL;
  if (c != 0. && !upre) return BAMHIP_EINVAL;
  const int *qn = unew + BAMHIP_Z4C_NVARS, *qc = ucur + BAMHIP_Z4C_NVARS;
  const int *qp = upre ? upre + BAMHIP_Z4C_NVARS : NULL;
  int metric[10];
  metric[0] = ucur[18];
  for (int a = 0; a < 3; a++) metric[1 + a] = ucur[19 + a];
  for (int a = 0; a < 6; a++) metric[4 + a] = S->adm_g[a];
  /* geometry */
  if ((e = bamhip_z4c_rhs(lev, S->z4, unew, upre, c, ucur))) return e;
  if ((e = bamhip_z4c_to_adm(lev, ucur, S->z4->chipsipower, S->adm_g, S->adm_K))) return e;
  /* matter: c2p, prepare, sources, fluxes */
  if (S->c2p_method == BAMHIP_C2P_LEARNED)
    e = bamhip_grhd_c2p_learned(lev, S->c2p, S->mlp, S->polish_max, qc, S->adm_g, S->prim, S->status, S->iters,
                                S->fallbacks);
  else
    e = bamhip_grhd_c2p(lev, S->c2p, S->c2p_method, qc, S->adm_g, S->prim, S->status, S->iters);
  if (e) return e;
  /* (MATTER.USEATM && MATTER.USEATM_PRERHS): the atmosphere of u_p from the
   * new primitives, before matter_rhs_prepare copies u_p (c != 0) */
  if (S->atm_prerhs && qp && (e = bamhip_grhd_atm_set_prerhs(lev, S->c2p, qp, S->prim))) return e;
  if ((e = bamhip_vladd(lev, 5, qn, c != 0. ? 1. : 0., c != 0. ? qp : qc, 0., qc))) return e;
  if ((e = bamhip_grhd_sources(lev, S->hrsc, S->c2p->rhoatm, S->z4->order_centered, qn, c, metric, S->adm_K,
                               S->prim, S->madm)))
    return e;
  if ((e = bamhip_grhd_fluxes(lev, S->hrsc, qn, c, metric, S->prim))) return e;
  /* matter terms of the geometry, boundary */
  if ((e = bamhip_z4c_matter_source(lev, S->z4, unew, c, ucur, S->madm))) return e;
  return bamhip_z4c_set_boundary(lev, S->z4, unew, upre ? upre : unew, c, ucur, S->pgh);
}

/* One RK4 step of the coupled system: evolve_rk_rk4 (src/utility/evolve/
 * rk.c:220-257) over the 30 evolved fields with the RHS above, called as
 * evolve_rhs(r, u_p, 0.0, v) so that the atmosphere reset acts on u_p, then
 * z4_algcon_wrapper on the Z4c fields (z4.c:49-71).  The RHS's non-finite
 * check is deferred to the end of the step, as in bamhip_z4c_rk4_step, and
 * the caller's check mode is restored.  Lists hold 30 indices (25 Z4c, then
 * D, tau, S_i). */
int bamhip_z4c_grhd_rk4_step(bamhip_level *lev, const bamhip_z4c_grhd *S, double dt, const int *uc, const int *up,
                             const int *r, const int *v, int algcon)
{
  const int nz = BAMHIP_Z4C_NVARS;
  const double wc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double wv[4] = {dt / 2, dt / 2, dt, 0.0};
  int e, s, flag = 0, mode;
  if (!lev || !S || !uc || !up || !r || !v) return BAMHIP_EINVAL;
  if ((e = bamhip_level_get_check(lev, &mode))) return e;
  if ((e = bamhip_level_set_check(lev, 0))) return e;
  if ((e = bamhip_level_nonfinite(lev, 1, 0))) goto out;
  /* u_p = u_c */
  if ((e = bamhip_vladd(lev, nz, up, 1.0, uc, 0.0, 0))) goto out;
  if ((e = bamhip_vladd(lev, 5, up + nz, 1.0, uc + nz, 0.0, 0))) goto out;
  for (s = 0; s < 4; s++) {
    const int *src = (s == 0) ? uc : v;
    if ((e = bamhip_z4c_grhd_rhs(lev, S, r, up, 0.0, src))) goto out;
    if ((e = bamhip_rk4_stage_update(lev, nz, uc, up, r, v, wc[s], wv[s]))) goto out;
    if ((e = bamhip_rk4_stage_update(lev, 5, uc + nz, up + nz, r + nz, v + nz, wc[s], wv[s]))) goto out;
  }
  if (algcon && (e = bamhip_z4c_algcon_list(lev, S->z4, uc))) goto out;
  if ((e = bamhip_level_nonfinite(lev, 0, &flag))) goto out;
  if (flag) e = BAMHIP_ENONFINITE;
out:
  bamhip_level_set_check(lev, mode);
  return e;
}
