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

/* An anchor record as 8 floats: X, Y, dx, dy, peak, u, v, found. */
static void anchors_out(const anchor_t *A, float *out) {
  for (int k = 0; k < OB_ANCH_K * OB_ANCH_K; k++) {
    float *o = out + 8 * k;
    o[0] = A[k].X; o[1] = A[k].Y; o[2] = A[k].dx; o[3] = A[k].dy;
    o[4] = A[k].peak; o[5] = A[k].u; o[6] = A[k].v; o[7] = (float)A[k].ok;
  }
}
/* A lens model as 15 doubles: K, ax, ay, R, the 11 coefficients. */
static void model_out(const lensm_t *m, double *out) {
  out[0] = m->K; out[1] = m->ax; out[2] = m->ay; out[3] = m->R;
  for (int a = 0; a < LENS_P; a++) out[4 + a] = a < 8 + m->K ? m->t[a] : 0;
}
static void model_in(const double *in, lensm_t *m) {
  memset(m, 0, sizeof *m);
  m->K = (int)in[0]; m->ax = in[1]; m->ay = in[2]; m->R = in[3];
  for (int a = 0; a < LENS_P; a++) m->t[a] = in[4 + a];
}

/* anchor_find (acquire.c:1876-2036) on a tight iw x ih grey capture at gamma 1, every anchor searched through H and
   the lens model (15 doubles as model_out writes them; NULL: H alone); records into anchors (49 x 8). */
int ref_anchor_find(const void *p, const unsigned char *img, int iw, int ih, const float *H, const double *model,
                    float *anchors) {
  const lizard_t *f = p;
  image_t im;
  ob_image_init(&im, img, iw, ih, 1.0f);
  homo_t Hh;
  memcpy(Hh.h, H, sizeof Hh.h);
  lensm_t lm;
  if (model) model_in(model, &lm);
  anchor_t A[OB_ANCH_K * OB_ANCH_K];
  memset(A, 0, sizeof A);
  const int found = anchor_find(&f->frame, &im, &Hh, model ? &lm : 0, A);
  anchors_out(A, anchors);
  return found;
}

/* anchor_register (acquire.c:2128-2191) as acquire_core calls it, with the capture unblurred for a corner's search
   again: H (in and out), the lattice's corners in pixels quad (in and out), pixels a module mpx; the anchors' records
   and the corners' and anchors' lens model out. Returns the reference's result. */
int ref_anchor_register(const void *p, const unsigned char *img, int iw, int ih, float *H, float mpx, float *quad,
                        float *anchors, double *model) {
  const lizard_t *f = p;
  const ob_layout_t *L = &f->frame;
  image_t im;
  ob_image_init(&im, img, iw, ih, 1.0f);
  homo_t Hh;
  memcpy(Hh.h, H, sizeof Hh.h);
  node_t *nodes = calloc((size_t)L->nx * L->ny, sizeof(node_t));
  for (int c = 0; c < 4; c++) nodes[((c >= 2) ? L->ny - 1 : 0) * L->nx + ((c == 1 || c == 2) ? L->nx - 1 : 0)].done = 1;
  anchor_t A[OB_ANCH_K * OB_ANCH_K];
  memset(A, 0, sizeof A);
  lensm_t m;
  memset(&m, 0, sizeof m);
  const int have = anchor_register(L, &im, &im, &Hh, iw, ih, mpx, quad, nodes, A, &m);
  free(nodes);
  memcpy(H, Hh.h, sizeof Hh.h);
  anchors_out(A, anchors);
  model_out(&m, model);
  return have;
}

/* acquire_core's border (acquire.c:5014-5058): the capture's border band blurred where mpx / 2 >= 1 (border_blur),
   then every marked node refined ring by ring outward from the corners, rows within a ring (refine_node), through H
   and the lens model (NULL: none), the corners held done at 0. nodes: nx * ny x 4 floats (dx, dy, score, done). */
int ref_refine_border(const void *p, const unsigned char *img, int iw, int ih, const float *H, float mpx,
                      const double *model, float *out) {
  const lizard_t *f = p;
  const ob_layout_t *L = &f->frame;
  image_t im;
  ob_image_init(&im, img, iw, ih, 1.0f);
  image_t rim = im;
  homo_t Hh;
  memcpy(Hh.h, H, sizeof Hh.h);
  uint8_t *blurred = NULL;
  if (mpx / 2 >= 1 && (blurred = border_blur(L, &Hh, img, iw, ih, (int)(mpx / 2), NULL))) rim.px = blurred;
  lensm_t lm;
  if (model) model_in(model, &lm);
  const int nx = L->nx, ny = L->ny;
  node_t *nodes = calloc((size_t)nx * ny, sizeof(node_t));
  for (int c = 0; c < 4; c++) nodes[((c >= 2) ? ny - 1 : 0) * nx + ((c == 1 || c == 2) ? nx - 1 : 0)].done = 1;
  for (int ring = 0; ring <= nx + ny; ring++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int ri = (i < nx - 1 - i ? i : nx - 1 - i) + (j < ny - 1 - j ? j : ny - 1 - j);
        if (ri != ring || !L->node_mark[j * nx + i] || nodes[j * nx + i].done) continue;
        refine_node(L, &rim, &Hh, i, j, nodes, model ? &lm : 0);
      }
  for (int k = 0; k < nx * ny; k++) {
    out[4 * k] = nodes[k].dx; out[4 * k + 1] = nodes[k].dy; out[4 * k + 2] = nodes[k].score; out[4 * k + 3] = (float)nodes[k].done;
  }
  free(nodes);
  free(blurred);
  return nx;
}
