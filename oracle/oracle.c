/* ORACLE — test infrastructure only (see oracle.h). Plain C, scalar loops, written for clarity over speed.
 *
 * The LDPC decoder here is check-centric over explicit per-check variable lists (each check's data bits, then the
 * staircase bit shared with the previous check, then its own), which is the arithmetic of the reference's layered
 * decoder stated edge by edge: liblizard/src/ldpc.c:141-192 (scalar layer kernel), :589-617 (the slot order and check
 * 0's absent edge), :681-705 (the syndrome), :1080-1172 (iteration order, stall rule, outputs).
 */
#include "oracle.h"
#include "oracle_codes.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { Z = 128, SUB = 320, BLOCK = 476 };

/* ------------------------------------------------------------------ CRC-32C
 * Castagnoli, reflected polynomial 0x82F63B78, register 0xFFFFFFFF, result complemented (research/SPEC.md 7.1;
 * liblizard/src/layout.c:44-54). Bit at a time. */
uint32_t or_crc32c(const uint8_t *p, int n) {
  uint32_t c = 0xFFFFFFFFu;
  for (int i = 0; i < n; i++) {
    c ^= p[i];
    for (int b = 0; b < 8; b++) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : c >> 1;
  }
  return c ^ 0xFFFFFFFFu;
}

/* ------------------------------------------------------------------ whitening and bit map
 * PRBS-23, x^23 + x^18 + 1, all-ones start: w[i] = w[i-23] ^ w[i-18] with w[j<0] = 1 (SPEC 7.3;
 * liblizard/src/lizard_phy.c:98-101). Stated as the recurrence on the output sequence. */
void or_whiten(uint8_t *out, int n) {
  for (int i = 0; i < n; i++) {
    const int a = i - 23 >= 0 ? out[i - 23] : 1, b = i - 18 >= 0 ? out[i - 18] : 1;
    out[i] = (uint8_t)(a ^ b);
  }
}

static int gcd(int a, int b) { while (b) { int t = a % b; a = b; b = t; } return a; }
/* Stride: floor(0.3819660113 nt + 0.5), raised until prime to nt (SPEC 7.3; lizard_phy.c:112-116). */
int or_perm_stride(int nt) {
  int P = (int)floor(0.3819660113 * nt + 0.5);
  while (gcd(P, nt) != 1) P++;
  return P;
}
void or_perm(int *perm, int nt) {
  const long P = or_perm_stride(nt);
  for (int i = 0; i < nt; i++) perm[i] = (int)((P * i) % nt);
}

/* ------------------------------------------------------------------ codes */
typedef struct { int rate, mb, kb, np; const short *sh; } base_t;
static const base_t BASES[4] = {
  { 6, 5, 31, 1, &OR_SHIFT78[0][0] }, { 4, 10, 30, 0, &OR_SHIFT34[0][0] },
  { 3, 15, 30, 0, &OR_SHIFT23[0][0] }, { 2, 30, 30, 0, &OR_SHIFT12[0][0] },
};
static const base_t *base_for(int rate) {
  for (int i = 0; i < 4; i++) if (BASES[i].rate == rate) return &BASES[i];
  return 0;
}

/* Sent bits n - np (ldpc.c:44), plus 16 pilots at 7/8 and 32 at 3/4 (lizard_phy.c:583-586). */
int or_sent_bits(int rate) { const base_t *b = base_for(rate); return b ? (b->kb - b->np + b->mb) * Z : 0; }
int or_block_slots(int rate) {
  const int nt = or_sent_bits(rate);
  return nt ? nt + (rate == 6 ? 16 : rate == 4 ? 32 : 0) : 0;
}

struct or_code {
  int rate, mb, kb, np, n, k, m, nt, edges, slots_total;
  int *row_ptr, *col_idx;  /* H row-major, per check: data bits by column, then parity c-1 (c > 0), then c */
  int *lay_ptr;            /* per block row: first slot of the row's messages (reference message layout) */
  const short *sh;
};

static int modz(int a) { a %= Z; return a < 0 ? a + Z : a; }

or_code *or_code_new(int rate) {
  const base_t *b = base_for(rate);
  if (!b) return 0;
  or_code *c = calloc(1, sizeof *c);
  c->rate = rate; c->mb = b->mb; c->kb = b->kb; c->np = b->np * Z; c->sh = b->sh;
  c->n = (b->kb + b->mb) * Z; c->k = b->kb * Z; c->m = b->mb * Z; c->nt = c->n - c->np;
  /* check c = mb * i + r (block row r, position i): data bit j z + ((i - sh[r][j]) mod z) (SPEC 8.2; ldpc.c:67-74) */
  int e = 0;
  for (int r = 0; r < c->mb; r++) for (int j = 0; j < c->kb; j++) e += b->sh[r * c->kb + j] >= 0;
  c->edges = e * Z + 2 * c->m - 1;
  c->row_ptr = malloc((size_t)(c->m + 1) * sizeof(int));
  c->col_idx = malloc((size_t)c->edges * sizeof(int));
  c->lay_ptr = malloc((size_t)(c->mb + 1) * sizeof(int));
  int at = 0;
  for (int chk = 0; chk < c->m; chk++) {
    const int i = chk / c->mb, r = chk % c->mb;
    c->row_ptr[chk] = at;
    for (int j = 0; j < c->kb; j++) { const int s = b->sh[r * c->kb + j]; if (s >= 0) c->col_idx[at++] = j * Z + modz(i - s); }
    if (chk > 0) c->col_idx[at++] = c->k + chk - 1;
    c->col_idx[at++] = c->k + chk;
  }
  c->row_ptr[c->m] = at;
  /* messages: block row r holds (data circulants + 2) slots of z, rows one after another (ldpc.c:76-84, :1119) */
  int s = 0;
  for (int r = 0; r < c->mb; r++) {
    c->lay_ptr[r] = s;
    for (int j = 0; j < c->kb; j++) s += b->sh[r * c->kb + j] >= 0;
    s += 2;
  }
  c->lay_ptr[c->mb] = s;
  c->slots_total = s;
  return c;
}

void or_code_free(or_code *c) { if (!c) return; free(c->row_ptr); free(c->col_idx); free(c->lay_ptr); free(c); }

void or_code_dims(const or_code *c, int *d) {
  d[0] = c->n; d[1] = c->k; d[2] = c->m; d[3] = Z; d[4] = c->mb; d[5] = c->kb; d[6] = c->np; d[7] = c->nt;
  d[8] = c->edges; d[9] = c->slots_total;
}

/* FNV-1a over row pointers then column indices (liblizard/test/ldpc_table.c:12-18). */
uint32_t or_code_hash(const or_code *c) {
  uint32_t h = 2166136261u;
  for (int i = 0; i <= c->m; i++) h = (h ^ (uint32_t)c->row_ptr[i]) * 16777619u;
  for (int e = 0; e < c->edges; e++) h = (h ^ (uint32_t)c->col_idx[e]) * 16777619u;
  return h;
}

/* p_c = p_(c-1) ^ (xor of the data bits in check c), p_(-1) = 0 (SPEC 8.4; ldpc.c:117-126). */
void or_encode(const or_code *c, const uint8_t *data, uint8_t *cw) {
  memcpy(cw, data, (size_t)c->k);
  uint8_t p = 0;
  for (int chk = 0; chk < c->m; chk++) {
    for (int e = c->row_ptr[chk]; e < c->row_ptr[chk + 1]; e++) if (c->col_idx[e] < c->k) p ^= data[c->col_idx[e]];
    cw[c->k + chk] = p;
  }
}

/* Where a posterior sits in the reference decoder's workspace: data bits in code order, parity bit mb i + r at
   k + r z + i (ldpc.c:1109-1110). */
static int ws_index(const or_code *c, int v) {
  if (v < c->k) return v;
  const int chk = v - c->k;
  return c->k + (chk % c->mb) * Z + chk / c->mb;
}

static int check_unsat(const or_code *c, const int16_t *P) {
  int bad = 0;
  for (int chk = 0; chk < c->m; chk++) {
    int par = 0;
    for (int e = c->row_ptr[chk]; e < c->row_ptr[chk + 1]; e++) par ^= P[c->col_idx[e]] < 0;
    bad += par;
  }
  return bad;
}

/* Layered normalised min-sum with the reference's integer semantics (SPEC 8.5; ldpc.c:171-191, :1080-1172):
   posteriors int16 clamped to +-8191, messages int8, magnitudes (min * 13) >> 4 capped at 127, first least
   magnitude wins ties, zero counts positive. Layer = block row; the sweep runs rows 0..mb-1 on even iterations and
   mb-1..0 on odd ones. Check 0 has no previous parity bit: that slot reads as a certain 0 of value 8191 and its
   message is zeroed before the layer (ldpc.c:598-600, :1120). Syndrome after every iteration; with stall_it > 0 the
   unsatisfied counts after iteration 1 and iteration stall_it decide the stall (ldpc.c:1143-1148). */
int or_decode(const or_code *c, const int8_t *llr, uint8_t *out, int max_iter, int stall_it, float stall_ratio,
              int16_t *L_out, int8_t *R_out) {
  int16_t *P = malloc((size_t)c->n * sizeof(int16_t));
  int8_t *R = calloc((size_t)c->slots_total * Z, 1);
  for (int v = 0; v < c->n; v++) P[v] = llr[v];
  int q[64], vars[64];
  int used = -1, first = 0;
  for (int it = 0; it < max_iter; it++) {
    for (int step = 0; step < c->mb; step++) {
      const int r = (it & 1) ? c->mb - 1 - step : step;
      int8_t *Rl = R + (size_t)c->lay_ptr[r] * Z;
      const int ns = c->lay_ptr[r + 1] - c->lay_ptr[r], nd = ns - 2;
      if (r == 0) Rl[nd * Z + 0] = 0;
      for (int i = 0; i < Z; i++) {
        const int chk = c->mb * i + r;
        /* the check's slots: data bits by column, the previous parity bit (absent for check 0: -1), its own */
        int at = 0;
        for (int j = 0; j < c->kb; j++) { const int s = c->sh[r * c->kb + j]; if (s >= 0) vars[at++] = j * Z + modz(i - s); }
        vars[at++] = chk > 0 ? c->k + chk - 1 : -1;
        vars[at++] = c->k + chk;
        int min1 = 32767, min2 = 32767, arg = -1, sign = 0;
        for (int e = 0; e < ns; e++) {
          const int post = vars[e] >= 0 ? P[vars[e]] : 8191;
          q[e] = post - Rl[e * Z + i];
          const int a = q[e] < 0 ? -q[e] : q[e];
          sign ^= q[e] < 0;
          if (a < min1) { min2 = min1; min1 = a; arg = e; } else if (a < min2) min2 = a;
        }
        int m1 = (min1 * 13) >> 4, m2 = (min2 * 13) >> 4;
        if (m1 > 127) m1 = 127;
        if (m2 > 127) m2 = 127;
        for (int e = 0; e < ns; e++) {
          const int mag = e == arg ? m2 : m1, msg = (sign ^ (q[e] < 0)) ? -mag : mag;
          Rl[e * Z + i] = (int8_t)msg;
          int t = q[e] + msg;
          t = t > 8191 ? 8191 : t < -8191 ? -8191 : t;
          if (vars[e] >= 0) P[vars[e]] = (int16_t)t;
        }
      }
    }
    if (stall_it > 0 && (it == 0 || it == stall_it - 1)) {
      const int bad = check_unsat(c, P);
      if (!bad) { used = it + 1; break; }
      if (it == 0) first = bad;
      else if ((float)bad > stall_ratio * (float)first) { used = -3; break; }
    } else if (!check_unsat(c, P)) { used = it + 1; break; }
  }
  for (int v = 0; v < c->n; v++) out[v] = P[v] < 0;
  if (L_out) for (int v = 0; v < c->n; v++) L_out[ws_index(c, v)] = P[v];
  if (R_out) memcpy(R_out, R, (size_t)c->slots_total * Z);
  free(P); free(R);
  return used;
}

/* ------------------------------------------------------------------ the rate profile and the coefficient order
 * Full search over (a, t, c) in rising order with strict improvement, so the first of equals is kept; m fills
 * (SPEC 3.1; lizard_phy.c:594-621 searches the same space with pruning). */
int or_tiers_for(int subch, int *rate, int *blocks) {
  if (subch < 64 || subch > 1024 || subch % 16) return 0;
  static const int RATE[4] = { 6, 4, 3, 2 }, TGT[4] = { 390, 0, 150, 200 };
  long long s[4], S = (long long)subch * 640, best = -1;
  for (int q = 0; q < 4; q++) s[q] = or_block_slots(RATE[q]);
  int keep[4] = { 0 };
  for (long long a = 0; a * s[0] + s[1] <= S; a++)
    for (long long t = 0; a * s[0] + t * s[2] + s[1] <= S; t++)
      for (long long cc = 0; a * s[0] + t * s[2] + cc * s[3] + s[1] <= S; cc++) {
        const long long d0 = 1000 * s[0] * a - TGT[0] * S, d2 = 1000 * s[2] * t - TGT[2] * S, d3 = 1000 * s[3] * cc - TGT[3] * S;
        const long long cost = d0 * d0 + d2 * d2 + d3 * d3;
        if (best < 0 || cost < best) {
          best = cost; keep[0] = (int)a; keep[2] = (int)t; keep[3] = (int)cc;
          keep[1] = (int)((S - a * s[0] - t * s[2] - cc * s[3]) / s[1]);
        }
      }
  int nt = 0;
  for (int q = 0; q < 4; q++) if (keep[q]) { rate[nt] = RATE[q]; blocks[nt] = keep[q]; nt++; }
  return nt;
}

/* The first picture size with n >= 3 R, R = sqrt(640 subch / pi) (SPEC 3.2; lizard_phy.c:1491-1495). */
int or_n_for(int subch) {
  static const int PIC[7] = { 384, 512, 768, 1024, 1152, 1280, 1536 };
  const double r = sqrt(640.0 * subch / 3.14159265358979323846);
  for (int i = 0; i < 7; i++) if (PIC[i] >= 3 * r) return PIC[i];
  return 1536;
}

typedef struct { int r2, u, v; } cand_t;
static int by_freq(const void *pa, const void *pb) {
  const cand_t *a = pa, *b = pb;
  if (a->r2 != b->r2) return a->r2 < b->r2 ? -1 : 1;
  if (a->v != b->v) return a->v < b->v ? -1 : 1;
  return a->u < b->u ? -1 : a->u > b->u;
}
/* Upper half-plane without DC, Nyquist row and column, sorted by u^2 + v^2, then v, then u (SPEC 6.3); offsets into
   blocks of bw columns v, each n rows u (SPEC 6.12 P1; lizard_phy.c:532-535). */
int or_pos(int n, int subch, int *pos) {
  const int h = n / 2, need = subch * SUB;
  cand_t *c = malloc((size_t)h * (2 * h) * sizeof *c);
  int cnt = 0;
  for (int v = 0; v < h; v++) for (int u = -h + 1; u < h; u++) {
    if (v == 0 && u <= 0) continue;
    c[cnt].r2 = u * u + v * v; c[cnt].u = u; c[cnt].v = v; cnt++;
  }
  if (need > cnt) { free(c); return -1; }
  qsort(c, (size_t)cnt, sizeof *c, by_freq);
  const int bw = h < 64 ? h : 64;
  int vmax = 0;
  for (int i = 0; i < need; i++) {
    pos[i] = (c[i].v / bw) * n * bw + ((c[i].u + n) % n) * bw + c[i].v % bw;
    if (c[i].v > vmax) vmax = c[i].v;
  }
  free(c);
  return vmax / bw + 1;
}

/* ------------------------------------------------------------------ a frame's format */
struct or_format {
  int n, subch, blocks, tiers, vblocks, used;
  int tier_rate[4], tier_blocks[4];
  int *pos, *blk_rate, *blk_slot, *blk_nslots;
  uint8_t *white;
  or_code *code[4];
  int *perm[4];
};

or_format *or_format_new(int n, int subch) {
  or_format *f = calloc(1, sizeof *f);
  f->n = n; f->subch = subch;
  f->tiers = or_tiers_for(subch, f->tier_rate, f->tier_blocks);
  if (!f->tiers) { free(f); return 0; }
  f->pos = malloc((size_t)subch * SUB * sizeof(int));
  f->vblocks = or_pos(n, subch, f->pos);
  if (f->vblocks < 0) { free(f->pos); free(f); return 0; }
  for (int t = 0; t < f->tiers; t++) f->blocks += f->tier_blocks[t];
  f->blk_rate = malloc((size_t)f->blocks * 3 * sizeof(int));
  f->blk_slot = f->blk_rate + f->blocks; f->blk_nslots = f->blk_slot + f->blocks;
  int b = 0, at = 0;
  for (int t = 0; t < f->tiers; t++) {
    f->code[t] = or_code_new(f->tier_rate[t]);
    const int nt = f->code[t]->nt;
    f->perm[t] = malloc((size_t)nt * sizeof(int));
    or_perm(f->perm[t], nt);
    /* sent bit j is codeword bit j, or j + np past k - np (lizard_phy.c:111, :571-575) */
    for (int i = 0; i < nt; i++) { const int j = f->perm[t][i]; f->perm[t][i] = j < f->code[t]->k - f->code[t]->np ? j : j + f->code[t]->np; }
    for (int q = 0; q < f->tier_blocks[t]; q++, b++) {
      f->blk_rate[b] = t; f->blk_slot[b] = at; f->blk_nslots[b] = or_block_slots(f->tier_rate[t]); at += f->blk_nslots[b];
    }
  }
  f->used = at;
  f->white = malloc((size_t)subch * SUB * 2);
  or_whiten(f->white, subch * SUB * 2);
  return f;
}

void or_format_free(or_format *f) {
  if (!f) return;
  for (int t = 0; t < f->tiers; t++) { or_code_free(f->code[t]); free(f->perm[t]); }
  free(f->pos); free(f->blk_rate); free(f->white); free(f);
}

void or_format_info(const or_format *f, int *info) {
  info[0] = f->n; info[1] = f->subch; info[2] = f->blocks; info[3] = f->tiers; info[4] = f->vblocks; info[5] = f->used;
}
void or_format_blocks(const or_format *f, int *rate, int *slot0, int *nslots) {
  for (int b = 0; b < f->blocks; b++) { rate[b] = f->tier_rate[f->blk_rate[b]]; slot0[b] = f->blk_slot[b]; nslots[b] = f->blk_nslots[b]; }
}
const int *or_format_pos(const or_format *f) { return f->pos; }

/* ------------------------------------------------------------------ finish from soft values */
int or_finish_llr(const or_format *f, const int8_t *llr, const float *est, uint8_t *blocks, uint8_t *ok, int8_t *its,
                  float *blk_est) {
  int good = 0;
  int8_t *cw = malloc(8192);
  uint8_t *bits = malloc(8192);
  for (int b = 0; b < f->blocks; b++) {
    const int t = f->blk_rate[b], s0 = f->blk_slot[b], nb = f->blk_nslots[b];
    const or_code *c = f->code[t];
    ok[b] = 0; its[b] = 0;
    memset(blocks + (size_t)b * BLOCK, 0, BLOCK);
    /* the block's estimate: each sub-channel's estimate weighted by the block's slots in it (lizard_phy.c:1782-1787) */
    double info = 0;
    for (int j = s0 / 640; j * 640 < s0 + nb; j++) {
      const int lo = s0 > 640 * j ? s0 : 640 * j, hi = s0 + nb < 640 * (j + 1) ? s0 + nb : 640 * (j + 1);
      info += (double)est[j] * (hi - lo);
    }
    blk_est[b] = (float)(info / nb);
    /* decline gate: under 0.2 k / nt (lizard_phy.c:1818, SPEC 8.5) */
    if (blk_est[b] < 0.2f * (float)c->k / (float)c->nt) { its[b] = -2; continue; }
    /* slot order to codeword order, whitening off by sign; unsent bits 0, then 3840..k-1 known zeros at +127
       (lizard_phy.c:156-192, :1823) */
    for (int i = 0; i < c->n; i++) cw[i] = 0;
    for (int i = 0; i < c->nt; i++) {
      const int8_t v = llr[s0 + i];
      cw[f->perm[t][i]] = f->white[s0 + i] ? (int8_t)(-v) : v;
    }
    for (int i = 8 * (BLOCK + 4); i < c->k; i++) cw[i] = 127;
    const int it = or_decode(c, cw, bits, 30, 9, 0.95f, 0, 0);
    its[b] = (int8_t)it;
    if (it < 0) continue;
    uint8_t *dst = blocks + (size_t)b * BLOCK;
    for (int j = 0; j < BLOCK; j++) {
      int v = 0;
      for (int q = 0; q < 8; q++) v = (v << 1) | bits[8 * j + q];
      dst[j] = (uint8_t)v;
    }
    uint32_t crc = 0;
    for (int i = 0; i < 32; i++) crc = (crc << 1) | bits[8 * BLOCK + i];
    if (crc != or_crc32c(dst, BLOCK)) continue;
    ok[b] = 1; good++;
  }
  free(cw); free(bits);
  return good;
}

/* ------------------------------------------------------------------ frame slots (the encoder's bit side) */
void or_frame_slots(const or_format *f, const uint8_t *blocks, int cnt, uint8_t *slots) {
  const int S = f->subch * 640;
  uint8_t data[4096], cw[8192];
  memcpy(slots, f->white, (size_t)S);   /* padding: the whitening alone */
  for (int b = 0; b < f->blocks; b++) {
    const int t = f->blk_rate[b], s0 = f->blk_slot[b], nb = f->blk_nslots[b];
    const or_code *c = f->code[t];
    const uint8_t *src = blocks + (size_t)b * BLOCK;
    memset(data, 0, sizeof data);
    for (int i = 0; i < 8 * BLOCK; i++) data[i] = (src[i >> 3] >> (7 - (i & 7))) & 1;
    const uint32_t crc = or_crc32c(src, BLOCK);
    for (int i = 0; i < 32; i++) data[8 * BLOCK + i] = (crc >> (31 - i)) & 1;
    or_encode(c, data, cw);
    /* slot i: codeword bit perm[i], or past nt the pilot bit (bit 0 of the count on even blocks, bit 1 on odd),
       whitened (SPEC 7.3; lizard_phy.c:121-153) */
    const uint8_t q = (uint8_t)((cnt >> (b & 1)) & 1);
    for (int i = 0; i < nb; i++) slots[s0 + i] = (uint8_t)((i < c->nt ? cw[f->perm[t][i]] : q) ^ f->white[s0 + i]);
  }
}
