/* interp_oracle.c -- TEST INFRASTRUCTURE: CPU restatement of BAM's Lagrange
 * prolongation weights (src/utility/interpolate/lagrange.c:13-27) and the
 * 1-D cell-centred 2:1 prolongation pattern of prolong_boxtobox_N
 * (src/utility/interpolate/interpolate_boxtobox_N.c:626-700).  Checker only.
 */
#include "oracle.h"

/* N-point Lagrange weights at x for nodes 0,1,...,n-1 (unit spacing) */
void ob_lagrange_coeffs(int n, double x, double *c)
{
  for (int i = 0; i < n; i++) {
    double d = 1;
    for (int j = 0; j < n; j++)
      if (j != i) d *= (x - j) / (double)(i - j);
    c[i] = d;
  }
}

/* 1-D prolongation of nc coarse values to 2*nc fine values; fine 2i sits at
 * coarse coordinate i+1/4 and 2i+1 at i+3/4.  Each coarse window of `order`
 * points starting at s yields fine values at s+order/2-1/4 (weights c1) and,
 * from the window shifted by one, at s+1+order/2-3/4 (weights c2).  Fine points
 * without a full window are left untouched. */
void ob_prolong_1d(int order, int nc, const double *uc, double *uf)
{
  double c1[16], c2[16];
  ob_lagrange_coeffs(order, order / 2.0 - 0.25, c1);
  ob_lagrange_coeffs(order, order / 2.0 - 0.75, c2);
  for (int s = 0; s + order <= nc; s++) {
    /* fine point at coarse coordinate s + order/2 - 1/4 */
    double a = 0;
    for (int p = 0; p < order; p++) a += c1[p] * uc[s + p];
    const int ia = 2 * (s + order / 2 - 1) + 1;
    if (ia >= 0 && ia < 2 * nc) uf[ia] = a;
    /* fine point at coarse coordinate s + order/2 + 1/4 (window s+1) */
    if (s + 1 + order <= nc) {
      double b = 0;
      for (int p = 0; p < order; p++) b += c2[p] * uc[s + 1 + p];
      const int ib = 2 * (s + order / 2);
      if (ib >= 0 && ib < 2 * nc) uf[ib] = b;
    }
  }
}
