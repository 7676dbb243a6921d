// The format word (research/SPEC.md 5.3-5.4; liblizard/src/fmt.h, fmt.c, rs.c, acquire.c ob_thin_read_fmt): the
// border band's Reed-Solomon word that names a symbol's version and display rate, read from a capture through a
// registration. Host code: about 570 bilinear reads a frame.
#pragma once
#include <stdint.h>

namespace lz {

constexpr int FMT_MAGIC = 0x4c, FMT_DATA = 3, FMT_BYTES_MIN = 16, FMT_BYTES_MAX = 64, FMT_SIDE_CELLS_MIN = 32;
constexpr int FMT_VERSION_MIN = 4, FMT_VERSION_MAX = 64;

struct FmtWord {
  int version = 0, fps = 0;
};

// Reed-Solomon over GF(256) (x^8 + x^4 + x^3 + x^2 + 1), generator roots alpha^0 .. alpha^(nroots - 1), codeword
// k data bytes then nroots parity bytes, byte j the coefficient of x^(n - 1 - j).
void rs_encode(uint8_t *cw, int k, int nroots);
// Errors and erasures (positions eras[0 .. neras)), corrected in place when 2 errors + erasures <= nroots and the
// result is a codeword: returns errors + erasures corrected, or -1.
int rs_decode(uint8_t *cw, int n, int nroots, const int *eras = nullptr, int neras = 0);

// The word's geometry on a symbol `side` modules square (the band between reserve modules at each end).
int fmt_side_cells(int side);
int fmt_bytes(int side);
int fmt_encode(const FmtWord &f, uint8_t *cw, int bytes);
// Bit `cell` of frame side `edge` as the word is laid out (byte k on side k % 4, cells 8 (k / 4) ..).
int fmt_bit(const uint8_t *cw, int bytes, int edge, int cell);
// Soft values q[side * cells + i] in -0.5 .. 0.5 (positive dark) -> the word: hard decisions first, then the least
// confident bytes as erasures, two more a step. 0, or -1 where nothing decoded with the magic and a valid version.
int fmt_decode(const float *q, int cells, int bytes, FmtWord *f);

// The border's module map (liblizard/src/layout.c thin_init + ob_layout_set_fmt): (span + 30)^2 kinds, row-major,
// KIND_DATA where the picture and its guard interval are painted, else KIND_LIGHT or KIND_DARK, | KIND_WORD on the
// word's cells. cw: the word's bytes (fmt_bytes of the side), or null for none (its cells light, not flagged).
constexpr int KIND_DATA = 0, KIND_LIGHT = 1, KIND_DARK = 2, KIND_WORD = 8;
void border_kinds(int span, const uint8_t *cw, uint8_t *kinds);
// 1 where timing-track cell j of side 0..3 (top, right, bottom, left) is dark (layout.c ob_thin_track).
int thin_track(int side, int j);

// The soft word of a capture (tight uint8 luma w x h, row stride) through a registration: H (module -> pixel, 9
// floats) and the node lattice's displacements (nodes: nx * nx pairs, the layout of ring `span`). q: 4 * cells values.
// false where the border shows no contrast (no band to read).
bool read_word_soft(int span, const uint8_t *img, int w, int h, int stride, const float *H, const float *nodes,
                    float *q, int cells);

}  // namespace lz
