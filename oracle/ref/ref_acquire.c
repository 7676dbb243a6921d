/* ORACLE (test infrastructure) — the reference's acquire.c compiled through this unit (oracle/ref/build_ref.sh) so
 * that its file-static functions can be called from the tests. Nothing of the reference is copied; this file only
 * includes the source where it lies and wraps some of its functions. */
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

/* border_fill_axis (acquire.c:1531-1605) on the frame layout of a format made by ref_format_new: the border nodes'
   displacements in nodes (nx * ny pairs) with use[k] the marked nodes' trust (score 1 or 0; corners always used), the
   axis (NaN, NaN: the frame's centre); the interior written back. */
#include "lizard_phy.h"
void ref_border_fill_axis(const void *p, const float *H, int iw, int ih, float ax, float ay, float *nodes,
                          const unsigned char *use) {
  const lizard_t *f = p;
  const ob_layout_t *L = &f->frame;
  const int nn = L->nx * L->ny;
  node_t *nd = calloc((size_t)nn, sizeof(node_t));
  for (int k = 0; k < nn; k++) {
    nd[k].dx = nodes[2 * k];
    nd[k].dy = nodes[2 * k + 1];
    nd[k].done = 1;
    nd[k].score = use && !use[k] ? 0.0f : 1.0f;
  }
  homo_t Hh;
  memcpy(Hh.h, H, sizeof Hh.h);
  lizard_set_axis(ax, ay);
  border_fill_axis(L, &Hh, iw, ih, nd);
  lizard_set_axis(NAN, NAN);
  for (int k = 0; k < nn; k++) {
    nodes[2 * k] = nd[k].dx;
    nodes[2 * k + 1] = nd[k].dy;
  }
  free(nd);
}

/* lensm_at (acquire.c:1615-1620): a lens model (K radial terms, coefficients t, axis, R) at image point (u, v). */
void ref_lensm_at(int K, const double *t, double ax, double ay, double R, double u, double v, float *dx, float *dy) {
  lensm_t m;
  memset(&m, 0, sizeof m);
  m.K = K;
  for (int a = 0; a < 8 + K; a++) m.t[a] = t[a];
  m.ax = ax;
  m.ay = ay;
  m.R = R;
  lensm_at(&m, u, v, dx, dy);
}

/* border_fill_anchors (acquire.c:2219-2358): as ref_border_fill_axis, with the anchors' readings (49 x 5 floats: X, Y,
   dx, dy, found) and pixels a module mpx. */
void ref_border_fill_anchors(const void *p, const float *H, int iw, int ih, float ax, float ay, float *nodes,
                             const unsigned char *use, const float *anchors, float mpx) {
  const lizard_t *f = p;
  const ob_layout_t *L = &f->frame;
  const int nn = L->nx * L->ny;
  node_t *nd = calloc((size_t)nn, sizeof(node_t));
  for (int k = 0; k < nn; k++) {
    nd[k].dx = nodes[2 * k];
    nd[k].dy = nodes[2 * k + 1];
    nd[k].done = 1;
    nd[k].score = use && !use[k] ? 0.0f : 1.0f;
  }
  anchor_t A[OB_ANCH_K * OB_ANCH_K];
  memset(A, 0, sizeof A);
  for (int k = 0; k < OB_ANCH_K * OB_ANCH_K; k++) {
    A[k].X = anchors[5 * k];
    A[k].Y = anchors[5 * k + 1];
    A[k].dx = anchors[5 * k + 2];
    A[k].dy = anchors[5 * k + 3];
    A[k].ok = anchors[5 * k + 4] != 0;
  }
  homo_t Hh;
  memcpy(Hh.h, H, sizeof Hh.h);
  lizard_set_axis(ax, ay);
  border_fill_anchors(L, &Hh, iw, ih, nd, A, mpx);
  lizard_set_axis(NAN, NAN);
  for (int k = 0; k < nn; k++) {
    nodes[2 * k] = nd[k].dx;
    nodes[2 * k + 1] = nd[k].dy;
  }
  free(nd);
}
