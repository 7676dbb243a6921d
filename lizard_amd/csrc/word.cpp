// The format word: Reed-Solomon over GF(256), the word's layout in the border band, and its soft read (see word.hpp).
#include "word.hpp"

#include <cstring>
#include <vector>

#include "common.hpp"
#include "spectrum.hpp"

namespace lz {
namespace {

struct Gf {
  uint8_t exp[512];
  uint8_t log[256];
  Gf() {
    int x = 1;
    for (int i = 0; i < 255; i++) {
      exp[i] = static_cast<uint8_t>(x);
      log[x] = static_cast<uint8_t>(i);
      x <<= 1;
      if (x & 0x100) x ^= 0x11d;
    }
    for (int i = 255; i < 512; i++) exp[i] = exp[i - 255];
    log[0] = 0;
  }
  uint8_t mul(uint8_t a, uint8_t b) const { return a && b ? exp[log[a] + log[b]] : 0; }
  uint8_t div(uint8_t a, uint8_t b) const { return a ? exp[log[a] + 255 - log[b]] : 0; }
  uint8_t pow_a(int e) const { return exp[((e % 255) + 255) % 255]; }
};

const Gf &gf() {
  static const Gf g;
  return g;
}

// p(x) at x (coefficients lowest degree first)
uint8_t eval_low(const uint8_t *p, int deg, uint8_t x) {
  const Gf &g = gf();
  uint8_t r = 0;
  for (int i = deg; i >= 0; i--) r = g.mul(r, x) ^ p[i];
  return r;
}

}  // namespace

void rs_encode(uint8_t *cw, int k, int nroots) {
  const Gf &g = gf();
  // generator, lowest degree first: prod (x + alpha^i)
  std::vector<uint8_t> gen(nroots + 1, 0);
  gen[0] = 1;
  for (int i = 0; i < nroots; i++) {
    const uint8_t a = g.pow_a(i);
    for (int j = i + 1; j >= 1; j--) gen[j] = gen[j - 1] ^ g.mul(gen[j], a);
    gen[0] = g.mul(gen[0], a);
  }
  // remainder of data(x) x^nroots by gen(x), by long division from the highest coefficient
  std::vector<uint8_t> rem(nroots, 0);  // rem[i]: coefficient of x^(nroots - 1 - i)
  for (int i = 0; i < k; i++) {
    const uint8_t fb = cw[i] ^ rem[0];
    for (int j = 0; j + 1 < nroots; j++) rem[j] = rem[j + 1] ^ g.mul(fb, gen[nroots - 1 - j]);
    rem[nroots - 1] = g.mul(fb, gen[0]);
  }
  std::memcpy(cw + k, rem.data(), nroots);
}

int rs_decode(uint8_t *cw, int n, int nroots, const int *eras, int neras) {
  const Gf &g = gf();
  if (n > 255 || nroots < 1 || nroots >= n || neras < 0 || neras > nroots) return -1;
  // syndromes S_i = c(alpha^i), c(x) = sum cw[j] x^(n - 1 - j)
  std::vector<uint8_t> S(nroots);
  bool any = false;
  for (int i = 0; i < nroots; i++) {
    uint8_t s = 0;
    const uint8_t a = g.pow_a(i);
    for (int j = 0; j < n; j++) s = g.mul(s, a) ^ cw[j];
    S[i] = s;
    any |= s != 0;
  }
  if (!any) return 0;
  // the erasures' locator, lowest degree first: prod (1 + X_e x), X_e = alpha^(n - 1 - position)
  std::vector<uint8_t> lam(nroots + 1, 0), prev(nroots + 1, 0), tmp(nroots + 1, 0);
  lam[0] = 1;
  for (int e = 0; e < neras; e++) {
    if (eras[e] < 0 || eras[e] >= n) return -1;
    const uint8_t X = g.pow_a(n - 1 - eras[e]);
    for (int j = e + 1; j >= 1; j--) lam[j] ^= g.mul(lam[j - 1], X);
  }
  // Berlekamp-Massey over the remaining syndromes, started from the erasure locator (errors and erasures at once)
  prev = lam;
  int L = neras, shift = 1;
  uint8_t bprev = 1;
  for (int r = neras; r < nroots; r++) {
    uint8_t d = 0;
    for (int i = 0; i <= L && i <= r; i++) d ^= g.mul(lam[i], S[r - i]);
    if (!d) {
      shift++;
      continue;
    }
    tmp = lam;
    const uint8_t c = g.div(d, bprev);
    for (int i = 0; i + shift <= nroots; i++) lam[i + shift] ^= g.mul(c, prev[i]);
    if (2 * L <= r + neras) {
      L = r + 1 + neras - L;
      prev = tmp;
      bprev = d;
      shift = 1;
    } else {
      shift++;
    }
  }
  int deg = nroots;
  while (deg > 0 && !lam[deg]) deg--;
  const int errors = deg - neras;
  if (deg != L || errors < 0 || 2 * errors + neras > nroots) return -1;
  // Chien search: position j is in error where lam(X_j^-1) = 0
  std::vector<int> pos;
  for (int j = 0; j < n; j++)
    if (!eval_low(lam.data(), deg, g.pow_a(-(n - 1 - j)))) pos.push_back(j);
  if (static_cast<int>(pos.size()) != deg) return -1;
  // Forney (first root alpha^0): e_j = X_j omega(X_j^-1) / lam'(X_j^-1), omega = S lam mod x^nroots
  std::vector<uint8_t> om(nroots, 0);
  for (int i = 0; i < nroots; i++)
    for (int k = 0; k <= i && k <= deg; k++) om[i] ^= g.mul(lam[k], S[i - k]);
  std::vector<uint8_t> dl(deg > 0 ? deg : 1, 0);  // formal derivative: odd terms
  for (int i = 1; i <= deg; i += 2) dl[i - 1] = lam[i];
  for (int j : pos) {
    const uint8_t X = g.pow_a(n - 1 - j), xi = g.pow_a(-(n - 1 - j));
    const uint8_t num = eval_low(om.data(), nroots - 1, xi), den = eval_low(dl.data(), deg - 1, xi);
    if (!den) return -1;
    cw[j] ^= g.mul(X, g.div(num, den));
  }
  // a codeword now, or the pattern was past the code's reach
  for (int i = 0; i < nroots; i++) {
    uint8_t s = 0;
    const uint8_t a = g.pow_a(i);
    for (int j = 0; j < n; j++) s = g.mul(s, a) ^ cw[j];
    if (s) return -1;
  }
  return deg;
}

// ---- the band (layout.h): cells OB_CELL = 2 modules along the border, runs of 4 track then 4 word cells, 15
// modules kept at each end of a side
namespace {
constexpr int CELL = 2, BAND_RUN = 4, RESERVE = 15, BORDER_SIDE = 30;
constexpr float DARK_DEPTH = 2.0f;  // the middle of the outer line (1 + line 2 / 2)
constexpr float GAP_MID = 4.0f, TRACK_MID = 6.0f, FMT_MID = 6.0f;  // band middles as depths
inline int band_pair(int k) { return (k / BAND_RUN) * 2 * BAND_RUN + k % BAND_RUN; }
inline int track_at(int j) { return band_pair(j) * CELL; }
inline int word_at(int i) { return (band_pair(i) + BAND_RUN) * CELL; }
inline int thin_cells(int side) { return side > 2 * RESERVE ? (side - 2 * RESERVE) / CELL : 0; }
inline int track_cells(int side) { return (thin_cells(side) + 1) / 2; }

// depths into the border (layout.h): the light module, the solid line 1..2, the light gap 3..4, the band 5..6, then
// light to the guard interval at 12 (the picture at 15 less its 3 modules of guard); corner marks fill [0, 12)^2
constexpr int LINE = 2, GAP = 1 + LINE, BAND = GAP + 2, THIN = BAND + 2, CORNER = 12, GUARD = 12;
}  // namespace

int thin_track(int side, int j) {
  // a per-side hash of the Manchester pair's index, flipped in the pair's second cell (layout.c:57-62, the format's)
  uint32_t v = static_cast<uint32_t>(side * 4099 + (j >> 1)) * 2654435761u + 40503u;
  v ^= v >> 15;
  v *= 2246822519u;
  v ^= v >> 13;
  return static_cast<int>((v ^ static_cast<uint32_t>(j)) & 1);
}

void border_kinds(int span, const uint8_t *cw, uint8_t *kinds) {
  const int L = span + BORDER_SIDE;
  for (int y = 0; y < L; y++)
    for (int x = 0; x < L; x++) {
      const int dx = x < L - 1 - x ? x : L - 1 - x, dy = y < L - 1 - y ? y : L - 1 - y, d = dx < dy ? dx : dy;
      int k;
      if (dx < CORNER && dy < CORNER) {
        // the corner mark: solid from depth 1 to CORNER - 2 on both axes, the gap ring left light through it
        const bool in = dx >= 1 && dx <= CORNER - 2 && dy >= 1 && dy <= CORNER - 2;
        k = in && !(d >= GAP && d < BAND) ? KIND_DARK : KIND_LIGHT;
      } else if (d >= GUARD) {
        k = KIND_DATA;
      } else if (d >= 1 && d <= LINE) {
        k = KIND_DARK;
      } else if (d >= BAND && d < THIN) {
        // the band: even runs of BAND_RUN cells the track, odd runs the word's (light until the word is set)
        const int along = dy <= dx ? x : y, side = dy <= dx ? (y < L / 2 ? 0 : 2) : (x < L / 2 ? 3 : 1);
        k = KIND_LIGHT;
        if (along >= RESERVE && along < L - RESERVE) {
          const int cell = (along - RESERVE) / CELL, run = cell / BAND_RUN;
          if (!(run & 1) && thin_track(side, run / 2 * BAND_RUN + cell % BAND_RUN)) k = KIND_DARK;
        }
      } else {
        k = KIND_LIGHT;
      }
      kinds[y * L + x] = static_cast<uint8_t>(k);
    }
  // the word's cells, two modules along by the band's two deep, flagged (layout.c:74-89); none without a word
  if (!cw) return;
  const int cells = fmt_side_cells(L), bytes = fmt_bytes(L);
  for (int side = 0; side < 4; side++)
    for (int i = 0; i < cells; i++) {
      const uint8_t v = static_cast<uint8_t>((fmt_bit(cw, bytes, side, i) ? KIND_DARK : KIND_LIGHT) | KIND_WORD);
      for (int j = 0; j < CELL; j++)
        for (int e = 0; e < 2; e++) {
          const int t = RESERVE + word_at(i) + j, dd = BAND + e;
          const int x = side == 0 || side == 2 ? t : side == 3 ? dd : L - 1 - dd;
          const int y = side == 0 ? dd : side == 2 ? L - 1 - dd : t;
          kinds[y * L + x] = v;
        }
    }
}

namespace {
// module coordinates on side 0..3 (top, right, bottom, left), t along it, d deep
inline void side_xy(int len, int side, float t, float d, float *x, float *y) {
  *x = side == 0 || side == 2 ? t : side == 3 ? d : static_cast<float>(len) - d;
  *y = side == 0 ? d : side == 2 ? static_cast<float>(len) - d : t;
}

struct Reader {
  const uint8_t *img;
  int w, h, stride;
  const float *H, *nodes;
  int nx;
  float node[MAX_NODES];
  float lut[256];
  // module point -> pixel: H plus the node lattice's bilinear displacement (acquire.c:5088-5103), in that order
  void map(float mx, float my, float *u, float *v) const {
    auto cell = [&](float m) {
      int lo = 0, hi = nx - 2;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (m >= node[mid]) lo = mid;
        else hi = mid - 1;
      }
      return lo;
    };
    const int j = cell(my), i = cell(mx);
    const float fv = (my - node[j]) / (node[j + 1] - node[j]), fu = (mx - node[i]) / (node[i + 1] - node[i]);
    const float *n00 = nodes + 2 * (j * nx + i), *n10 = n00 + 2, *n01 = n00 + 2 * nx, *n11 = n01 + 2;
    const float dx = (n00[0] * (1 - fu) + n10[0] * fu) * (1 - fv) + (n01[0] * (1 - fu) + n11[0] * fu) * fv;
    const float dy = (n00[1] * (1 - fu) + n10[1] * fu) * (1 - fv) + (n01[1] * (1 - fu) + n11[1] * fu) * fv;
    const float ww = H[6] * mx + H[7] * my + 1;
    *u = (H[0] * mx + H[1] * my + H[2]) / ww + dx;
    *v = (H[3] * mx + H[4] * my + H[5]) / ww + dy;
  }
  // bilinear grey read, pixel centres at half integers, clamped to the frame (internal.h:11-23)
  float sample(float x, float y) const {
    x -= 0.5f;
    y -= 0.5f;
    if (x < 0) x = 0;
    else if (x > w - 1.001f) x = w - 1.001f;
    if (y < 0) y = 0;
    else if (y > h - 1.001f) y = h - 1.001f;
    const int xi = x >= 0 ? static_cast<int>(x) : 0, yi = y >= 0 ? static_cast<int>(y) : 0;
    const float fx = x - xi, fy = y - yi;
    const uint8_t *p = img + static_cast<size_t>(yi) * stride + xi;
    const float a = lut[p[0]], b = lut[p[1]], c = lut[p[stride]], d = lut[p[stride + 1]];
    return (a + (b - a) * fx) * (1 - fy) + (c + (d - c) * fx) * fy;
  }
  float at(float mx, float my) const {
    float u, v;
    map(mx, my, &u, &v);
    return sample(u, v);
  }
};
}  // namespace

int fmt_side_cells(int side) {
  const int w = thin_cells(side) / (2 * BAND_RUN) * BAND_RUN;
  return w >= FMT_SIDE_CELLS_MIN ? w : 0;
}

int fmt_bytes(int side) {
  const int b = 4 * (fmt_side_cells(side) / 8);
  return b > FMT_BYTES_MAX ? FMT_BYTES_MAX : b;
}

int fmt_encode(const FmtWord &f, uint8_t *cw, int bytes) {
  if (bytes < FMT_BYTES_MIN || bytes > FMT_BYTES_MAX || f.version < FMT_VERSION_MIN || f.version > FMT_VERSION_MAX ||
      f.fps < 0 || f.fps > 255)
    return -1;
  std::memset(cw, 0, bytes);
  cw[0] = FMT_MAGIC;
  cw[1] = static_cast<uint8_t>(f.version);
  cw[2] = static_cast<uint8_t>(f.fps);
  rs_encode(cw, FMT_DATA, bytes - FMT_DATA);
  return 0;
}

int fmt_bit(const uint8_t *cw, int bytes, int edge, int cell) {
  const int k = 4 * (cell >> 3) + edge;
  if (edge < 0 || edge > 3 || cell < 0 || k >= bytes) return 0;
  return (cw[k] >> (7 - (cell & 7))) & 1;
}

int fmt_decode(const float *q, int cells, int bytes, FmtWord *f) {
  const int nroots = bytes - FMT_DATA;
  if (bytes < FMT_BYTES_MIN || bytes > FMT_BYTES_MAX || cells < 8 * ((bytes + 3) / 4)) return -1;
  uint8_t hard[FMT_BYTES_MAX], got[FMT_BYTES_MAX];
  float conf[FMT_BYTES_MAX];
  for (int k = 0; k < bytes; k++) {
    const float *qs = q + (k & 3) * cells + 8 * (k >> 2);
    uint8_t v = 0;
    float lo = 1;
    for (int b = 0; b < 8; b++) {
      v = static_cast<uint8_t>((v << 1) | (qs[b] > 0));
      const float a = qs[b] < 0 ? -qs[b] : qs[b];
      lo = a < lo ? a : lo;
    }
    hard[k] = v;
    conf[k] = lo;  // the byte's least confident bit
  }
  auto accept = [&](const uint8_t *cw) {
    if (cw[0] != FMT_MAGIC || cw[1] < FMT_VERSION_MIN || cw[1] > FMT_VERSION_MAX) return false;
    f->version = cw[1];
    f->fps = cw[2];
    return true;
  };
  std::memcpy(got, hard, bytes);
  if (rs_decode(got, bytes, nroots) >= 0 && accept(got)) return 0;
  // bytes by confidence, least first (the lower index on a tie), erased two more a step while three roots remain
  int order[FMT_BYTES_MAX];
  for (int k = 0; k < bytes; k++) order[k] = k;
  for (int a = 1; a < bytes; a++) {  // stable insertion by confidence
    const int k = order[a];
    int b = a;
    while (b > 0 && conf[order[b - 1]] > conf[k]) {
      order[b] = order[b - 1];
      b--;
    }
    order[b] = k;
  }
  for (int ne = 2; ne <= nroots - 3; ne += 2) {
    std::memcpy(got, hard, bytes);
    if (rs_decode(got, bytes, nroots, order, ne) >= 0 && accept(got)) return 0;
  }
  return -1;
}

bool read_word_soft(int span, const uint8_t *img, int w, int h, int stride, const float *H, const float *nodes,
                    float *q, int cells) {
  const int len = span + BORDER_SIDE;
  Reader R{img, w, h, stride, H, nodes, 0, {}, {}};
  R.nx = node_lattice(span, R.node);
  for (int g = 0; g < 256; g++) R.lut[g] = static_cast<float>(g) / 255.0f;
  if (cells < 1) return false;
  // the border's contrast: the outer line against the gap beside it, every 16 modules along every side
  float dark = 0, light = 0, x, y;
  int nref = 0;
  for (int side = 0; side < 4; side++)
    for (int t = RESERVE; t < len - RESERVE; t += 16, nref++) {
      side_xy(len, side, t + 0.5f, DARK_DEPTH, &x, &y);
      dark += R.at(x, y);
      side_xy(len, side, t + 0.5f, GAP_MID, &x, &y);
      light += R.at(x, y);
    }
  if (!nref) return false;
  const float con = (light - dark) / nref;
  if (con < 0.08f) return false;
  // each word cell against the mean level of the track's Manchester pairs within 3 pairs (12 modules) of it
  constexpr int REACH = 12 / (2 * CELL), PAIRMOD = 4 * CELL;
  std::vector<float> mid(len / PAIRMOD + 2);
  for (int side = 0; side < 4; side++) {
    const int npair = track_cells(len) / 2;
    if (npair < 1) return false;
    for (int pr = 0; pr < npair; pr++) {
      float s = 0;
      for (int k = 0; k < 2; k++) {
        side_xy(len, side, RESERVE + track_at(2 * pr + k) + CELL / 2.0f, TRACK_MID, &x, &y);
        s += R.at(x, y);
      }
      mid[pr] = 0.5f * s;
    }
    for (int i = 0; i < cells; i++) {
      const int off = word_at(i);
      side_xy(len, side, RESERVE + off + CELL / 2.0f, FMT_MID, &x, &y);
      const int at = off / PAIRMOD, lo = at - REACH < 0 ? 0 : at - REACH, hi = at + REACH >= npair ? npair - 1 : at + REACH;
      float m = 0;
      for (int k = lo; k <= hi; k++) m += mid[k];
      q[side * cells + i] = R.at(x, y) - m / (hi - lo + 1);
    }
  }
  // the word's own two levels (about half its cells dark): a two-means threshold from the track's, then dark
  // positive, in units of the contrast, clipped to +-0.5
  float ref = 0;
  for (int it = 0; it < 4; it++) {
    float lo = 0, hi = 0;
    int nlo = 0;
    for (int i = 0; i < 4 * cells; i++) {
      if (q[i] < ref) {
        lo += q[i];
        nlo++;
      } else {
        hi += q[i];
      }
    }
    if (nlo < cells / 2 || 4 * cells - nlo < cells / 2) break;
    ref = 0.5f * (lo / nlo + hi / (4 * cells - nlo));
  }
  for (int i = 0; i < 4 * cells; i++) {
    const float s = (ref - q[i]) / con;
    q[i] = s > 0.5f ? 0.5f : s < -0.5f ? -0.5f : s;
  }
  return true;
}

}  // namespace lz
