/* dissipation.c -- the Kreiss-Oliger factor of a level.
 *
 * get_dissipation_factor (src/utility/evolve/evolve.c:486-530): the factor
 * dissipation_factor / 2^order_dissipation, replaced, in this order of
 * precedence, by
 *   dissipation_factor_shells  on shell levels (when set, != -1),
 *   dissipation_factor_move    on levels l > amr_move_lcube (when set),
 *   0.2 dissipation_factor     on levels l > amr_move_lcube unless physics
 *                              contains Z4d,
 *   dissipation_factor_level0  on level 0 (when set),
 * each divided by 2^order_dissipation.  The parameter defaults are those of
 * bam_main.c:69-72 and bam_amr.c:85-90 (-1 = unset).
 */
#include <math.h>

#include "../../include/bamhip.h"

int bamhip_dissipation_factor(const bamhip_dissipation_rule *R, int l, int shells, double *factor)
{
  if (!R || !factor || l < 0) return BAMHIP_EINVAL;
  const double scale = pow(2.0, (double)R->order_dissipation);
  if (shells && R->factor_shells != -1.)
    *factor = R->factor_shells / scale;
  else if (l > R->amr_move_lcube && R->factor_move != -1.)
    *factor = R->factor_move / scale;
  else if (l > R->amr_move_lcube && !R->physics_z4d)
    *factor = 0.2 * R->dissipation_factor / scale;
  else if (l == 0 && R->factor_level0 != -1.)
    *factor = R->factor_level0 / scale;
  else
    *factor = R->dissipation_factor / scale;
  return BAMHIP_OK;
}
