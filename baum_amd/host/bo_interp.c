/* bo_interp.c -- Berger-Oliger time interpolation of a parent level.
 *
 * The C counterpart of the in-time interpolation that evolve() wraps around
 * the step of every level below the finest (src/utility/evolve/evolve.c:
 * 298-340 before the step, 367-380 after it): u_pp, the parent data at the
 * child's half step, is accumulated with the weights of the order-o
 * extrapolation, o = min(timeinterpolation, iteration + 2).  Every list
 * operation is a device vladd through the C ABI, in the reference's order,
 * so the results are those of the reference's vladd sequence.
 */
#include <stddef.h>

#include "../../include/bamhip.h"

/* vladd over lists longer than one device launch takes */
static int vladd_n(bamhip_level *lev, int n, const int *r, double ca, const int *a, double cb, const int *b)
{
  int e, k;
  for (k = 0; k < n; k += BAMHIP_Z4C_NVARS) {
    const int m = (n - k < BAMHIP_Z4C_NVARS) ? n - k : BAMHIP_Z4C_NVARS;
    if ((e = bamhip_vladd(lev, m, r + k, ca, a ? a + k : NULL, cb, b ? b + k : NULL))) return e;
  }
  return BAMHIP_OK;
}

static int effective_order(int timeinterpolation, int iteration)
{
  return (iteration + 2 > timeinterpolation) ? timeinterpolation : iteration + 2;
}

int bamhip_bo_interp_begin(bamhip_level *lev, int timeinterpolation, int iteration, int n, const int *upp,
                           const int *uc, const int *up, const int *up2, const int *up3, const int *up4)
{
  const int o = effective_order(timeinterpolation, iteration);
  int e;
  if (!lev || n < 1 || !upp || !uc) return BAMHIP_EINVAL;
  if (o < 2 || o > 6) return BAMHIP_EINVAL;
  if ((o >= 3 && !up) || (o >= 4 && !up2) || (o >= 5 && !up3) || (o >= 6 && !up4)) return BAMHIP_EINVAL;
  switch (o) {
    case 2:
      e = vladd_n(lev, n, upp, 0.5, uc, 0.0, uc);
      break;
    case 3:
      e = vladd_n(lev, n, upp, 0.75, uc, -0.125, up);
      break;
    case 4:
      if ((e = vladd_n(lev, n, upp, 15.0 / 16, uc, -5.0 / 16, up))) return e;
      e = vladd_n(lev, n, upp, 1.0, upp, 1.0 / 16, up2);
      break;
    case 5:
      if ((e = vladd_n(lev, n, upp, 140.0 / 128, uc, -70.0 / 128, up))) return e;
      if ((e = vladd_n(lev, n, upp, 1.0, upp, 28.0 / 128, up2))) return e;
      e = vladd_n(lev, n, upp, 1.0, upp, -5.0 / 128, up3);
      break;
    default:
      if ((e = vladd_n(lev, n, upp, 315.0 / 256, uc, -210.0 / 256, up))) return e;
      if ((e = vladd_n(lev, n, upp, 1.0, upp, 126.0 / 256, up2))) return e;
      if ((e = vladd_n(lev, n, upp, 1.0, upp, -45.0 / 256, up3))) return e;
      e = vladd_n(lev, n, upp, 1.0, upp, 7.0 / 256, up4);
      break;
  }
  if (e) return e;
  /* shift back time levels (vlcopy) */
  if (o >= 6 && (e = vladd_n(lev, n, up4, 1.0, up3, 0.0, NULL))) return e;
  if (o >= 5 && (e = vladd_n(lev, n, up3, 1.0, up2, 0.0, NULL))) return e;
  if (o >= 4 && (e = vladd_n(lev, n, up2, 1.0, up, 0.0, NULL))) return e;
  return BAMHIP_OK;
}

int bamhip_bo_interp_end(bamhip_level *lev, int timeinterpolation, int iteration, int n, const int *upp,
                         const int *uc)
{
  static const double c[7] = {0, 0, 1.0 / 2, 3. / 8, 5. / 16, 35. / 128, 63. / 256};
  const int o = effective_order(timeinterpolation, iteration);
  if (!lev || n < 1 || !upp || !uc || o < 2 || o > 6) return BAMHIP_EINVAL;
  return vladd_n(lev, n, upp, 1.0, upp, c[o], uc);
}
