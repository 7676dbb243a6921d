// A LIZARD frame format's layout facts, as the receive path reads them: the rate profile's tiers, where each block's
// slots lie, the coefficient order (positions in the half-plane spectrum), the whitening, each tier's bit map and the
// format's LDPC codes. Host tables, built once per format; the device copies live with the decoder.
#pragma once
#include <stdint.h>

#include <vector>

namespace lz {

constexpr int Z = 128;           // circulant size of every code
constexpr int SUB = 320;         // coefficients per sub-channel (640 slots)
constexpr int BLOCK_BYTES = 476; // a block: 4-byte id + 472 payload
constexpr int INFO_BITS = 3840;  // block bytes + CRC-32C
constexpr int MAX_N = 7680;      // longest codeword (1/2)
constexpr int MAX_SLOTS_TOTAL = 210;  // most message slots of any code (1/2: 30 block rows, 7 slots)
constexpr int NCODES = 4;        // code index: 0 = 7/8, 1 = 3/4, 2 = 2/3, 3 = 1/2

// One code's shape and block-row (layer) view.
struct Code {
  int rate = 0;   // rate index (6, 4, 3, 2)
  int n = 0, k = 0, m = 0, mb = 0, kb = 0, np = 0, nt = 0, edges = 0, slots_total = 0;
  std::vector<int> lay_ptr;   // mb + 1: first data circulant of each block row
  std::vector<int> lay_col, lay_shift;
};

int code_index(int rate);            // rate index -> 0..3, or -1
const Code &code(int idx);           // the four codes, built on first use
int block_slots(int rate);           // a block's slots: nt + pilots (16 at 7/8, 32 at 3/4)

struct Tier { int rate, blocks, slots; };

struct Format {
  int n = 0, subch = 0, version = 0, blocks = 0, tiers = 0, bw = 0, vblocks = 0, used = 0;
  Tier tier[4];
  std::vector<int> blk_code;     // per block: code index 0..3
  std::vector<int> blk_slot;     // per block: first frame slot
  std::vector<int> blk_nslots;   // per block: slots
  std::vector<int> pos;          // subch * 320: offset of each coefficient in the half-plane block layout
  std::vector<uint8_t> white;    // subch * 640 whitening bits, frame slot order
  // per code index: slot i (< nt) -> posterior index in the decoder's workspace order (data bits in code order,
  // parity bit mb*i + r at k + r*128 + i); empty for codes the format does not use
  std::vector<int16_t> slot_to_post[NCODES];
};

int tiers_for(int subch, Tier out[4]);    // the rate profile; returns the tier count (4) or 0
int n_for(int subch);                     // the picture size a format takes
bool build_format(int version, Format &f);
const Format *format_for(int version);    // cached, thread-safe

// Index of codeword bit v in the decoder's workspace order.
inline int post_index(const Code &c, int v) {
  if (v < c.k) return v;
  const int chk = v - c.k;
  return c.k + (chk % c.mb) * Z + chk / c.mb;
}

// CRC-32C of a byte string (host), and the per-bit contribution table of a 476-byte block: crc(m) = crc0 ^ XOR of
// tab[i] over the set bits i of m (bit i = bit 7 - i % 8 of byte i / 8).
uint32_t crc32c(const uint8_t *p, int n);
const std::vector<uint32_t> &crc_bit_table(uint32_t *crc0);

}  // namespace lz
