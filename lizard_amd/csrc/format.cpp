// Format tables for the receive path. Each piece states the format's rule (research/SPEC.md sections cited) and the
// reference function it must agree with.
#include "format.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

#include "ldpc_codes.h"

namespace lz {

// ------------------------------------------------------------------------------------------------ codes
// H = [Hd | staircase], check mb*i + r holds data bit 128 j + ((i - sh[r][j]) mod 128) and parity bits c - 1, c
// (SPEC 8.2; liblizard/src/ldpc.c:50-113).
static Code make_code(const CodeBase &b) {
  Code c;
  c.rate = b.rate;
  c.mb = b.mb;
  c.kb = b.kb;
  c.np = b.np * Z;
  c.n = (b.kb + b.mb) * Z;
  c.k = b.kb * Z;
  c.m = b.mb * Z;
  c.nt = c.n - c.np;
  c.lay_ptr.assign(b.ptr, b.ptr + b.mb + 1);
  c.lay_col.assign(b.col, b.col + b.ptr[b.mb]);
  c.lay_shift.assign(b.shift, b.shift + b.ptr[b.mb]);
  c.edges = b.ptr[b.mb] * Z + 2 * c.m - 1;
  c.slots_total = b.ptr[b.mb] + 2 * b.mb;
  return c;
}

int code_index(int rate) {
  for (int i = 0; i < NCODES; i++)
    if (CODE_BASES[i].rate == rate) return i;
  return -1;
}

const Code &code(int idx) {
  static const Code codes[NCODES] = {make_code(CODE_BASES[0]), make_code(CODE_BASES[1]), make_code(CODE_BASES[2]),
                                     make_code(CODE_BASES[3])};
  return codes[idx];
}

int block_slots(int rate) {
  const int i = code_index(rate);
  if (i < 0) return 0;
  return code(i).nt + (rate == 6 ? 16 : rate == 4 ? 32 : 0);  // pilots: lizard_phy.c:583-586
}

// ------------------------------------------------------------------------------------------------ rate profile
// Inner first: a blocks at 7/8, m at 3/4 (filling), t at 2/3, c at 1/2, with (a, t, c) minimising
// (1000 a s7 - 390 S)^2 + (1000 t s2 - 150 S)^2 + (1000 c s1 - 200 S)^2, the first of equals in rising a, t, c
// (SPEC 3.1; lizard_phy.c:594-621). Each term only grows past its target, so each loop stops one past it.
int tiers_for(int subch, Tier out[4]) {
  if (subch < 64 || subch > 1024 || subch % 16) return 0;
  const int rate[4] = {6, 4, 3, 2};
  const long long S = 640LL * subch;
  long long s[4];
  for (int q = 0; q < 4; q++) s[q] = block_slots(rate[q]);
  auto term = [&](long long cnt, int q, long long tgt) {
    const long long d = 1000 * s[q] * cnt - tgt * S;
    return d * d;
  };
  long long best = -1;
  long long ka = 0, kt = 0, kc = 0;
  for (long long a = 0; a * s[0] + s[1] <= S; a++) {
    const long long ca = term(a, 0, 390);
    for (long long t = 0; a * s[0] + t * s[2] + s[1] <= S; t++) {
      const long long ct = ca + term(t, 2, 150);
      for (long long c = 0; a * s[0] + t * s[2] + c * s[3] + s[1] <= S; c++) {
        const long long cost = ct + term(c, 3, 200);
        if (best < 0 || cost < best) {
          best = cost;
          ka = a;
          kt = t;
          kc = c;
        }
        if (1000 * s[3] * c > 200 * S) break;
      }
      if (1000 * s[2] * t > 150 * S) break;
    }
    if (1000 * s[0] * a > 390 * S) break;
  }
  const long long km = (S - ka * s[0] - kt * s[2] - kc * s[3]) / s[1];
  const long long cnt[4] = {ka, km, kt, kc};
  int nt = 0;
  for (int q = 0; q < 4; q++)
    if (cnt[q]) out[nt++] = Tier{rate[q], static_cast<int>(cnt[q]), static_cast<int>(s[q])};
  return nt;
}

// The first of 384, 512, 768, 1024, 1152, 1280, 1536 with n >= 3 sqrt(640 subch / pi) (SPEC 3.2;
// lizard_phy.c:1491-1495).
int n_for(int subch) {
  static const int pics[7] = {384, 512, 768, 1024, 1152, 1280, 1536};
  const double r = std::sqrt(2.0 * SUB * subch / 3.14159265358979323846);
  for (int p : pics)
    if (p >= 3 * r) return p;
  return 1536;
}

// ------------------------------------------------------------------------------------------------ CRC-32C
static const uint32_t *crc_table() {
  static uint32_t t[256];
  static std::once_flag once;
  std::call_once(once, [] {
    for (uint32_t b = 0; b < 256; b++) {
      uint32_t c = b;
      for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
      t[b] = c;
    }
  });
  return t;
}

uint32_t crc32c(const uint8_t *p, int n) {
  const uint32_t *t = crc_table();
  uint32_t c = 0xFFFFFFFFu;
  for (int i = 0; i < n; i++) c = (c >> 8) ^ t[(c ^ p[i]) & 0xFF];
  return ~c;
}

// The CRC is affine in the message: crc(m) = crc(0) ^ L(m), L linear, so each set bit contributes a fixed word.
const std::vector<uint32_t> &crc_bit_table(uint32_t *crc0) {
  static std::vector<uint32_t> tab;
  static uint32_t zero_crc;
  static std::once_flag once;
  std::call_once(once, [] {
    std::vector<uint8_t> m(BLOCK_BYTES, 0);
    zero_crc = crc32c(m.data(), BLOCK_BYTES);
    tab.resize(8 * BLOCK_BYTES);
    for (int i = 0; i < 8 * BLOCK_BYTES; i++) {
      m[i >> 3] = static_cast<uint8_t>(0x80u >> (i & 7));
      tab[i] = crc32c(m.data(), BLOCK_BYTES) ^ zero_crc;
      m[i >> 3] = 0;
    }
  });
  if (crc0) *crc0 = zero_crc;
  return tab;
}

// ------------------------------------------------------------------------------------------------ a format
static int gcd(int a, int b) {
  while (b) {
    const int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

bool build_format(int version, Format &f) {
  const int subch = 16 * version;
  f = Format();
  f.version = version;
  f.subch = subch;
  f.tiers = tiers_for(subch, f.tier);
  if (!f.tiers) return false;
  f.n = n_for(subch);
  const int n = f.n, h = n / 2, need = subch * SUB;
  // Coefficient order: the upper half-plane without DC and the Nyquist row and column, by u^2 + v^2, then v, then u
  // (SPEC 6.3); a 64-bit key sorts on all three at once.
  std::vector<uint64_t> key;
  key.reserve(static_cast<size_t>(h) * (2 * h));
  for (int v = 0; v < h; v++)
    for (int u = -h + 1; u < h; u++) {
      if (v == 0 && u <= 0) continue;
      const uint64_t r2 = static_cast<uint64_t>(u * u + v * v);
      key.push_back(r2 << 32 | static_cast<uint64_t>(v) << 16 | static_cast<uint64_t>(u + h));
    }
  if (static_cast<int>(key.size()) < need) return false;
  std::partial_sort(key.begin(), key.begin() + need, key.end());
  f.bw = h < 64 ? h : 64;
  f.pos.resize(need);
  int vmax = 0;
  for (int i = 0; i < need; i++) {
    const int v = static_cast<int>((key[i] >> 16) & 0xFFFF), u = static_cast<int>(key[i] & 0xFFFF) - h;
    // half-plane blocks of bw columns v, each n rows u (SPEC 6.12 P1; lizard_phy.c:533)
    f.pos[i] = (v / f.bw) * n * f.bw + ((u + n) % n) * f.bw + v % f.bw;
    vmax = v > vmax ? v : vmax;
  }
  f.vblocks = vmax / f.bw + 1;
  // Blocks one after another from slot 0, inner tiers first (SPEC 6.4; lizard_phy.c:563-567).
  int at = 0;
  for (int t = 0; t < f.tiers; t++)
    for (int q = 0; q < f.tier[t].blocks; q++) {
      f.blk_code.push_back(code_index(f.tier[t].rate));
      f.blk_slot.push_back(at);
      f.blk_nslots.push_back(f.tier[t].slots);
      at += f.tier[t].slots;
    }
  f.blocks = static_cast<int>(f.blk_slot.size());
  f.used = at;
  if (at > 2 * need) return false;
  // Whitening: the 23-bit register x^23 + x^18 + 1 from all ones (SPEC 7.3; lizard_phy.c:98-101).
  f.white.resize(2 * need);
  uint32_t s = 0x7FFFFF;
  for (int i = 0; i < 2 * need; i++) {
    const uint32_t b = ((s >> 22) ^ (s >> 17)) & 1u;
    s = ((s << 1) | b) & 0x7FFFFFu;
    f.white[i] = static_cast<uint8_t>(b);
  }
  // Bit map: slot i carries sent bit i P mod nt, P = floor(0.3819660113 nt + 0.5) raised until prime to nt; sent bit
  // j is codeword bit j below k - np and j + np above (SPEC 7.3; lizard_phy.c:103-116). Stored as the posterior index
  // the decoder keeps that codeword bit at.
  for (int t = 0; t < f.tiers; t++) {
    const int ci = code_index(f.tier[t].rate);
    const Code &c = code(ci);
    int P = static_cast<int>(0.3819660113 * c.nt + 0.5);
    while (gcd(P, c.nt) != 1) P++;
    auto &m = f.slot_to_post[ci];
    m.resize(c.nt);
    for (int i = 0; i < c.nt; i++) {
      const int j = static_cast<int>(static_cast<long long>(i) * P % c.nt);
      const int v = j < c.k - c.np ? j : j + c.np;
      m[i] = static_cast<int16_t>(post_index(c, v));
    }
  }
  return true;
}

const Format *format_for(int version) {
  static Format cache[65];
  static bool ready[65];
  static std::mutex mu;
  if (version < 4 || version > 64) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!ready[version]) {
    if (!build_format(version, cache[version])) return nullptr;
    ready[version] = true;
  }
  return &cache[version];
}

}  // namespace lz
