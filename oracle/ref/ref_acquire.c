/* ORACLE (test infrastructure) — the reference's acquire.c compiled through this unit (oracle/ref/build_ref.sh) so
 * that its file-static interpolation can be called from the tests. Nothing of the reference is copied; this file only
 * includes the source where it lies and wraps three of its functions. */
#include "acquire.c"

/* anch_herm4 (acquire.c:2196-2210): cubic Hermite weights through knots t[0..n-1] at x. */
void ref_anch_herm4(const double *t, int n, double x, int *b, double *w) { anch_herm4(t, n, x, b, w); }

/* ob_anchor_m0 / ob_anchor_side (acquire.c:41-45) for a ring of this span. */
int ref_anchor_m0(int span, int i) {
  ob_layout_t L;
  memset(&L, 0, sizeof L);
  L.w = span + 2 * OB_MARGIN;
  return ob_anchor_m0(&L, i);
}
int ref_anchor_side(int span) {
  ob_layout_t L;
  memset(&L, 0, sizeof L);
  L.w = span + 2 * OB_MARGIN;
  return ob_anchor_side(&L);
}
