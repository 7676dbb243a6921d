// SHAKE256 (FIPS 202) for the test stream's payloads: Keccak-f[1600] over a 1600-bit state, rate 136 bytes, the
// domain and pad bits 0x1F ... 0x80. The round constants come from FIPS 202's LFSR rc(t) (x^8 + x^6 + x^5 + x^4 + 1)
// and the rotation offsets from its walk (x, y) -> (y, 2x + 3y) with offsets (t + 1)(t + 2) / 2, built once.
#include <stdint.h>

#include <cstring>

#include "shake.hpp"

namespace lz {
namespace {

struct KeccakTables {
  uint64_t rc[24];
  int rho[25];
  KeccakTables() {
    // rc(t): bit t of the LFSR's output (FIPS 202 algorithm 5)
    auto rcbit = [](int t) {
      if (t % 255 == 0) return 1;
      unsigned r = 0x80;  // R = 10000000, bit 0 the leftmost
      for (int i = 1; i <= t % 255; i++) {
        // R = 0 || R, then R[0] ^= R[8], R[4] ^= R[8], R[5] ^= R[8], R[6] ^= R[8], then truncate to 8 bits
        unsigned r9 = r >> 1;  // shifted right as the leftmost bit moves one place on
        const unsigned b8 = r & 1;
        r9 ^= b8 << 7;
        r9 ^= b8 << 3;
        r9 ^= b8 << 2;
        r9 ^= b8 << 1;
        r = r9 & 0xFF;
      }
      return static_cast<int>((r >> 7) & 1);
    };
    for (int ir = 0; ir < 24; ir++) {
      uint64_t c = 0;
      for (int j = 0; j <= 6; j++)
        if (rcbit(j + 7 * ir)) c |= 1ull << ((1 << j) - 1);
      rc[ir] = c;
    }
    rho[0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; t++) {
      rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
      const int nx = y, ny = (2 * x + 3 * y) % 5;
      x = nx;
      y = ny;
    }
  }
};

const KeccakTables &tables() {
  static const KeccakTables t;
  return t;
}

inline uint64_t rotl(uint64_t v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }

void keccak_f(uint64_t A[25]) {
  const KeccakTables &T = tables();
  for (int ir = 0; ir < 24; ir++) {
    uint64_t C[5], D[5], B[25];
    for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];  // theta
    for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl(C[(x + 1) % 5], 1);
    for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
    for (int x = 0; x < 5; x++)  // rho and pi: B[y, 2x + 3y] = rot(A[x, y])
      for (int y = 0; y < 5; y++) B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(A[x + 5 * y], T.rho[x + 5 * y]);
    for (int y = 0; y < 5; y++)  // chi
      for (int x = 0; x < 5; x++) A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
    A[0] ^= T.rc[ir];  // iota
  }
}

}  // namespace

void shake256(const uint8_t *in, size_t in_len, uint8_t *out, size_t out_len) {
  constexpr size_t RATE = 136;
  uint64_t A[25] = {};
  uint8_t block[RATE];
  auto absorb = [&](const uint8_t *b) {
    for (size_t i = 0; i < RATE / 8; i++) {
      uint64_t w = 0;
      for (int k = 0; k < 8; k++) w |= static_cast<uint64_t>(b[8 * i + k]) << (8 * k);
      A[i] ^= w;
    }
    keccak_f(A);
  };
  while (in_len >= RATE) {
    absorb(in);
    in += RATE;
    in_len -= RATE;
  }
  std::memset(block, 0, RATE);
  if (in_len) std::memcpy(block, in, in_len);
  block[in_len] ^= 0x1F;
  block[RATE - 1] ^= 0x80;
  absorb(block);
  while (out_len) {
    const size_t n = out_len < RATE ? out_len : RATE;
    for (size_t i = 0; i < n; i++) out[i] = static_cast<uint8_t>(A[i / 8] >> (8 * (i % 8)));
    out += n;
    out_len -= n;
    if (out_len) keccak_f(A);
  }
}

void stream_fill(uint32_t id, uint8_t *out, uint32_t len) {
  const uint8_t b[4] = {static_cast<uint8_t>(id), static_cast<uint8_t>(id >> 8), static_cast<uint8_t>(id >> 16),
                        static_cast<uint8_t>(id >> 24)};
  shake256(b, 4, out, len);
}

}  // namespace lz
