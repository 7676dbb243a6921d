/* ORACLE (test infrastructure) — a driver over the reference decoder's own C sources, compiled from where they lie in
 * the reference checkout by oracle/ref/build_ref.sh into oracle/_ref/. Nothing of the reference is copied; this file
 * only calls its functions, the way the reference's own test programs do (liblizard/test/phy_x86.c includes
 * lizard_phy.c and defines ob_debug_hash the same way). Used to pin the oracle and as the CPU baseline. */
#include "lizard_phy.c"
#include "rs.h"
#include <stdio.h>

uint32_t ob_debug_hash;

uint32_t ref_crc32c(const uint8_t *p, int n) { return ob_crc32c(p, n); }

/* ---------------------------------------------------------------- LDPC (ldpc.h) */
static ldpc_t *code_for(int rate) {
  static _Thread_local ldpc_t codes[LDPC_RATES];
  static _Thread_local int ready[LDPC_RATES];
  if (rate < 0 || rate >= LDPC_RATES) return 0;
  if (!ready[rate]) { if (ldpc_init(&codes[rate], rate)) return 0; ready[rate] = 1; }
  return &codes[rate];
}
int ref_ldpc_dims(int rate, int *dims) {
  ldpc_t *c = code_for(rate);
  if (!c) return -1;
  dims[0] = c->n; dims[1] = c->k; dims[2] = c->m; dims[3] = c->z; dims[4] = c->mb; dims[5] = c->kb; dims[6] = c->np;
  dims[7] = c->nt; dims[8] = c->edges; dims[9] = c->slots_total; dims[10] = c->w4 != 0;
  return 0;
}
void ref_ldpc_encode(int rate, const uint8_t *data, uint8_t *cw) { ldpc_encode(code_for(rate), data, cw); }
int ref_ldpc_decode(int rate, const int8_t *llr, uint8_t *out, int max_iter, int stall_it, float ratio, int16_t *L_out, int8_t *R_out) {
  ldpc_t *c = code_for(rate);
  const int its = ldpc_decode_stall(c, llr, out, max_iter, stall_it, ratio);
  if (L_out) memcpy(L_out, c->wl, (size_t)c->n * sizeof(int16_t));
  if (R_out) for (int s = 0; s < c->slots_total; s++) memcpy(R_out + (size_t)s * c->z, c->wr + (size_t)s * c->zp, (size_t)c->z);
  return its;
}
/* Many codewords of one rate: what the CPU baseline times (one decoder per thread, as liblizard/core/cpu/pool.h). */
int ref_ldpc_decode_many(int rate, int count, const int8_t *llr, uint8_t *out, int8_t *its, int max_iter, int stall_it, float ratio) {
  ldpc_t *c = code_for(rate);
  int ok = 0;
  for (int i = 0; i < count; i++) {
    const int r = ldpc_decode_stall(c, llr + (size_t)i * c->n, out + (size_t)i * c->n, max_iter, stall_it, ratio);
    its[i] = (int8_t)r;
    ok += r > 0;
  }
  return ok;
}

/* ---------------------------------------------------------------- a frame's format (lizard_phy.h) */
void *ref_format_new(int n, int subch, int span) {
  lizard_t *f = calloc(1, sizeof *f);
  if (lizard_init(f, n, subch, span)) { free(f); return 0; }
  return f;
}
void ref_format_free(void *p) { if (p) { lizard_free(p); free(p); } }
/* The display rate the format's word states (lizard_fmt_fps): 0, or -1 where refused. */
int ref_set_fps(void *p, int fps) { return lizard_fmt_fps(p, fps); }
void ref_format_info(const void *p, int *info) {
  const lizard_t *f = p;
  info[0] = f->n; info[1] = f->subch; info[2] = f->blocks; info[3] = f->tiers; info[4] = f->vblocks; info[5] = f->used;
  info[6] = f->bw; info[7] = f->pxm; info[8] = f->px; info[9] = f->margin;
}
const int *ref_format_pos(const void *p) { return ((const lizard_t *)p)->pos; }
void ref_format_blocks(const void *p, int *rate, int *slot0, int *nslots) {
  const lizard_t *f = p;
  for (int b = 0; b < f->blocks; b++) { rate[b] = f->tier[f->block_tier[b]].rate; slot0[b] = f->block_slot[b]; nslots[b] = f->tier[f->block_tier[b]].slots; }
}
/* cx[n], qx[n], x1r[n], x1i[n], x2r[n], x2i[n], sx2, sq2 (lizard_tables) */
void ref_format_tables(const void *p, float *out) { lizard_tables(p, out); }
void ref_format_erase(const void *p, int *elo, int *ehi, double *hole_inv) {
  const lizard_t *f = p;
  for (int i = 0; i < OB_ANCH_K; i++) { elo[i] = f->ws->elo[i]; ehi[i] = f->ws->ehi[i]; }
  memcpy(hole_inv, f->ws->hole_inv, sizeof f->ws->hole_inv);
}

/* The finish path's seam (lizard_finish_ext) with nothing acquired: a held registration with no nodes stands for
   the acquisition, which neither the grid path nor the soft-value path reads. */
static void hold_nothing(const lizard_t *f) {
  struct lizard_ws *w = f->ws;
  memset(&w->hold_reg, 0, sizeof w->hold_reg);
  memset(&w->hold_im, 0, sizeof w->hold_im);
  w->hold_valid = 1;
}
static void block_out(const lizard_t *f, int8_t *its, float *est, float *pilot) {
  for (int b = 0; b < f->blocks; b++) {
    if (its) its[b] = f->ws->blk_its[b];
    if (est) est[b] = f->ws->blk_est[b];
    if (pilot) pilot[b] = f->ws->blk_pilot[b];
  }
}
int ref_finish_llr(const void *p, const int8_t *llr, const float *est, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est) {
  const lizard_t *f = p;
  ob_result_t res;
  hold_nothing(f);
  const int good = lizard_finish_ext(f, 0, llr, est, blocks, ok, &res);
  block_out(f, its, blk_est, 0);
  return good;
}
/* grid: n x n samples in strip layout, transformed in place. Out: the frame's soft values (subch * 640), each
   sub-channel's estimate and rms, the pilot shift, the spectrum at every used coefficient (re, im interleaved, pos
   order, before the pilots' turn), per-block verdicts, estimates and pilot readings. */
int ref_finish_grid(const void *p, float *grid, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est, float *blk_pilot,
                    int8_t *llr_out, double *sub_est, float *sub_rms, float *align, float *coef) {
  const lizard_t *f = p;
  struct lizard_ws *w = f->ws;
  ob_result_t res;
  if (coef) lizard_debug((lizard_t *)f, 1, 0, 0);
  hold_nothing(f);
  const int good = lizard_finish_ext(f, grid, 0, 0, blocks, ok, &res);
  block_out(f, its, blk_est, blk_pilot);
  const int S = f->subch * LIZARD_SUB;
  if (llr_out) memcpy(llr_out, w->llrf, (size_t)S * 2);
  if (sub_est) memcpy(sub_est, w->sub_est, (size_t)f->subch * sizeof(double));
  if (sub_rms) memcpy(sub_rms, w->sub_rms, (size_t)f->subch * sizeof(float));
  if (align) { align[0] = w->align[0]; align[1] = w->align[1]; }
  if (coef) memcpy(coef, w->dbg_coef, (size_t)S * 2 * sizeof(float));
  return good;
}

/* The encoder: blocks * 476 bytes, pilot count c, the drive (px * px floats) and the frame's slot bits (0/1). */
int ref_encode(void *p, const uint8_t *blocks, int c, float *drive, uint8_t *slots) {
  lizard_t *f = p;
  const int8_t *sym = 0;
  if (lizard_debug(f, 1, &sym, 0) || lizard_parity(f, c)) return -1;
  lizard_encode(f, blocks, drive);
  if (slots) for (int i = 0; i < f->subch * LIZARD_SUB * 2; i++) slots[i] = sym[i] < 0;
  return 0;
}
void ref_paint_grey(const void *p, const float *drive, uint8_t *grey) { lizard_paint_grey(p, drive, grey); }
/* The encoder's picture, after ref_encode: the n x n clipped picture in [0, 1], row-major, that lizard_encode
   resamples into its drive (its workspace's rs_pic, lizard_phy.c:1183-1200). */
int ref_encoded_picture(const void *p, float *pic) {
  const lizard_t *f = p;
  const struct lizard_ws *w = f->ws;
  if (!w || !w->rs_pic) return -1;
  memcpy(pic, w->rs_pic, (size_t)f->n * f->n * sizeof(float));
  return 0;
}

/* ---------------------------------------------------------------- the sampler (acquire.c) */
/* The picture grid through a registration given as H (9 floats, h8 = 1) and the node lattice's displacements
   (nx * ny pairs), read from a tight uint8 capture at gamma 1: ob_sample_grid (the float walk, the definition) when
   pic == 0, ob_sample_grid_pic (the decoder's form) when pic == 1. Output row-major n x n. */
int ref_sample_grid(const void *p, const uint8_t *img, int iw, int ih, const float *H, const float *nodes, int pic, float *out) {
  const lizard_t *f = p;
  image_t im;
  ob_image_init(&im, img, iw, ih, 1.0f);
  ob_reg_t reg;
  memcpy(reg.H.h, H, sizeof reg.H.h);
  const int nn = f->frame.nx * f->frame.ny;
  reg.nodes = calloc((size_t)nn, sizeof(node_t));
  for (int i = 0; i < nn; i++) { reg.nodes[i].dx = nodes[2 * i]; reg.nodes[i].dy = nodes[2 * i + 1]; reg.nodes[i].done = 1; }
  const float g0 = f->margin + 0.5f / f->scale, step = 1.0f / f->scale;
  if (pic) ob_sample_grid_pic(&f->frame, &im, &reg, g0, g0, step, 0, f->n, 0, f->n, out, f->n);
  else ob_sample_grid(&f->frame, &im, &reg, g0, g0, step, 0, f->n, 0, f->n, out, f->n);
  ob_reg_free(&reg);
  return 0;
}
/* ---------------------------------------------------------------- the format word (fmt.c, rs.c, acquire.c) */
void ref_rs_encode(uint8_t *cw, int k, int nroots) { rs_encode(cw, k, nroots); }
int ref_rs_decode(uint8_t *cw, int n, int nroots, const int *eras, int neras) {
  return neras ? rs_decode_er(cw, n, nroots, eras, neras) : rs_decode(cw, n, nroots);
}
int ref_fmt_encode(int version, int fps, uint8_t *cw, int bytes) { ob_fmt_t f = { version, fps }; return ob_fmt_encode(&f, cw, bytes); }
int ref_fmt_decode(const float *q, int cells, int bytes, int *version, int *fps) {
  ob_fmt_t f = { 0, 0 };
  const int r = ob_fmt_decode(q, cells, bytes, &f);
  *version = f.version; *fps = f.fps;
  return r;
}
int ref_fmt_geom(int side, int *cells, int *bytes) { *cells = ob_fmt_side_cells(side, OB_THIN_RESERVE); *bytes = ob_fmt_bytes(side, OB_THIN_RESERVE); return 0; }
/* The border's module map (layout.c ob_layout_init, then ob_layout_set_fmt with cw when given): side^2 kinds. */
int ref_layout_kinds(int side, const uint8_t *cw, uint8_t *kinds) {
  ob_layout_t L;
  if (ob_layout_init(&L, side)) return -1;
  if (cw) ob_layout_set_fmt(&L, cw);
  memcpy(kinds, L.kind, (size_t)side * side);
  ob_layout_free(&L);
  return 0;
}
/* The word read as lizard_acquire holds it (lizard_phy.c:1554-1556): the soft values (4 * cells, into q) and the
   decode through a registration given as H and the node lattice's displacements. dims out: cells, bytes. Returns 1
   decoded (version, fps set), 0 read but not decoded, -1 nothing read. */
int ref_read_word(const void *p, const uint8_t *img, int iw, int ih, const float *H, const float *nodes, float *q,
                  int *dims, int *version, int *fps) {
  const lizard_t *f = p;
  const struct lizard_ws *w = f->ws;
  image_t im;
  ob_image_init(&im, img, iw, ih, 1.0f);
  ob_reg_t reg;
  memcpy(reg.H.h, H, sizeof reg.H.h);
  const int nn = f->frame.nx * f->frame.ny;
  reg.nodes = calloc((size_t)nn, sizeof(node_t));
  for (int i = 0; i < nn; i++) { reg.nodes[i].dx = nodes[2 * i]; reg.nodes[i].dy = nodes[2 * i + 1]; reg.nodes[i].done = 1; }
  dims[0] = w->fmt_cells; dims[1] = w->fmt_bytes;
  const int got = w->fmt_cells && ob_thin_read_fmt(&f->frame, &im, &reg, q, w->fmt_cells);
  ob_fmt_t fm = { 0, 0 };
  const int ok = got && ob_fmt_decode(q, w->fmt_cells, w->fmt_bytes, &fm) == 0;
  *version = ok ? fm.version : 0; *fps = ok ? fm.fps : 0;
  ob_reg_free(&reg);
  return got ? ok : -1;
}

/* The node lattice: nx, ny, then node_x[nx], node_y[ny] in modules. */
int ref_node_lattice(const void *p, int *dims, float *nx_out, float *ny_out) {
  const lizard_t *f = p;
  dims[0] = f->frame.nx; dims[1] = f->frame.ny;
  if (nx_out) for (int i = 0; i < f->frame.nx; i++) nx_out[i] = f->frame.node_x[i];
  if (ny_out) for (int i = 0; i < f->frame.ny; i++) ny_out[i] = f->frame.node_y[i];
  return 0;
}

/* lizard_finish on a capture through a given registration (what lizard_acquire would have held): the reference's
   own sampler, erase, transform, soft values, LDPC and CRC. */
/* The reference's own finder (lizard_acquire: the thin frame's finder, the corner nodes, the anchors, the border mesh)
   on a tight uint8 capture at gamma 1. Returns 1 registered (H 9 floats, nodes nx * ny pairs, quad 8 floats, orient),
   else 0; the hold is let go either way. */
int ref_acquire(const void *p, const uint8_t *img, int iw, int ih, float *H, float *nodes, float *quad, int *orient) {
  const lizard_t *f = p;
  ob_result_t res;
  const int got = lizard_acquire(f, img, iw, ih, 1.0f, &res);
  if (got) {
    const ob_reg_t *r = lizard_held_reg(f);
    memcpy(H, r->H.h, 9 * sizeof(float));
    const int nn = f->frame.nx * f->frame.ny;
    for (int i = 0; i < nn; i++) { nodes[2 * i] = r->nodes[i].dx; nodes[2 * i + 1] = r->nodes[i].dy; }
    memcpy(quad, res.quad, sizeof res.quad);
    *orient = res.orient;
  }
  lizard_release(f);
  return got ? 1 : 0;
}
int ref_decode_reg(const void *p, const uint8_t *img, int iw, int ih, const float *H, const float *nodes, uint8_t *blocks,
                   uint8_t *ok, int8_t *its, float *blk_est, int8_t *llr_out, float *align) {
  const lizard_t *f = p;
  struct lizard_ws *w = f->ws;
  ob_image_init(&w->hold_im, img, iw, ih, 1.0f);
  memcpy(w->hold_reg.H.h, H, sizeof w->hold_reg.H.h);
  const int nn = f->frame.nx * f->frame.ny;
  w->hold_reg.nodes = calloc((size_t)nn, sizeof(node_t));
  for (int i = 0; i < nn; i++) { w->hold_reg.nodes[i].dx = nodes[2 * i]; w->hold_reg.nodes[i].dy = nodes[2 * i + 1]; w->hold_reg.nodes[i].done = 1; }
  w->hold_valid = 1;
  ob_result_t res;
  const int good = lizard_finish_ext(f, 0, 0, 0, blocks, ok, &res);
  block_out(f, its, blk_est, 0);
  if (llr_out) memcpy(llr_out, w->llrf, (size_t)f->subch * LIZARD_SUB * 2);
  if (align) { align[0] = w->align[0]; align[1] = w->align[1]; }
  return good;
}
