// The layered min-sum decoder specialised per code at compile time: every circulant's column and shift is a constant,
// so a block row's posterior reads are issued back to back with immediate offsets, and a check's int8 messages live
// in VGPRs (four to a register, byte indices fixed per block row). A 128-lane workgroup (two waves) decodes one
// codeword: lane i owns check position i of every block row. Posteriors (int16) stay in LDS.
//
// The arithmetic is the reference decoder's scalar form (liblizard/src/ldpc.c:171-191, :589-617, :681-705,
// :1080-1172; research/SPEC.md 8.5). Shared by the
// gfx950 kernel and a host-side emulation (tests walk the 128 lanes on a CPU).
#pragma once
#include <stdint.h>

#include <type_traits>

#include "ldpc_codes.h"

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#endif
#define LZ_INLINE inline
#else
#define LZ_INLINE __forceinline__
#endif

namespace lz {
namespace tm {

template <int B, int E, class F>
__host__ __device__ LZ_INLINE void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Compile-time facts of code CI (0 = 7/8, 1 = 3/4, 2 = 2/3, 3 = 1/2).
template <int CI>
struct Shape {
  static constexpr const CodeBase &cb = CODE_BASES[CI];
  static constexpr int MB = cb.mb, KB = cb.kb, K = cb.kb * 128, N = (cb.kb + cb.mb) * 128;
  static constexpr int SLOTS = cb.ptr[cb.mb] + 2 * cb.mb;   // message slots of a check position (all block rows)
  static constexpr int WORDS = (SLOTS + 3) / 4;              // VGPRs a check position's messages take
  static constexpr int d0(int r) { return cb.ptr[r]; }
  static constexpr int nd(int r) { return cb.ptr[r + 1] - cb.ptr[r]; }
  static constexpr int rbase(int r) { return cb.ptr[r] + 2 * r; }
  static constexpr int col(int e) { return cb.col[e]; }
  static constexpr int shift(int e) { return cb.shift[e]; }
  static constexpr int max_col_degree() {
    int best = 0;
    for (int j = 0; j < KB; j++) {
      int d = 0;
      for (int e = 0; e < cb.ptr[MB]; e++) d += cb.col[e] == j;
      best = d > best ? d : best;
    }
    return best;
  }
};

// Two int16 lanes in one 32-bit register (packed 16-bit instructions on the GPU, element-wise on the host).
typedef int16_t v2i16 __attribute__((vector_size(4)));
typedef uint16_t v2u16 __attribute__((vector_size(4)));

// Messages of the check position a lane owns: slot s at byte s % 4 of word s / 4.
template <int CI>
struct Msgs {
  uint32_t m[Shape<CI>::WORDS];
};

// Keeps the compiler from hoisting every block row's addresses out of the iteration loop (they depend only on the
// lane, so all of them would stay live in VGPRs): the lane's offset 2 i is re-read, opaquely, once per block row.
__host__ __device__ LZ_INLINE int fresh(int i) {
#ifdef __HIP_DEVICE_COMPILE__
  asm volatile("" : "+v"(i));
#endif
  return i;
}

template <int S>
__host__ __device__ LZ_INLINE int get_msg(const uint32_t *w) {
  return static_cast<int>(static_cast<int8_t>(static_cast<uint8_t>(w[S >> 2] >> ((S & 3) * 8))));
}
// Byte permute of the 8-byte pool {hi, lo} (lo bytes 0-3, hi bytes 4-7): result byte i = pool byte sel[i]
// (v_perm_b32 on the GPU).
__host__ __device__ LZ_INLINE uint32_t perm_b32(uint32_t hi, uint32_t lo, uint32_t sel) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_perm(hi, lo, sel);
#else
  const uint64_t pool = (static_cast<uint64_t>(hi) << 32) | lo;
  uint32_t out = 0;
  for (int i = 0; i < 4; i++) out |= static_cast<uint32_t>((pool >> (8 * ((sel >> (8 * i)) & 7))) & 0xFF) << (8 * i);
  return out;
#endif
}

// Messages S and S + 1 from the two 16-bit halves of m (their low bytes), one byte permute a word.
template <int S>
__host__ __device__ LZ_INLINE void put_msg2(uint32_t *w, uint32_t m) {
  constexpr int p = S & 3;
  if constexpr (p < 3) {
    // byte p <- m byte 0 (pool 4), byte p + 1 <- m byte 2 (pool 6), the rest kept
    constexpr uint32_t sel = (p == 0 ? 0x03020604u : p == 1 ? 0x03060400u : 0x06040100u);
    w[S >> 2] = perm_b32(m, w[S >> 2], sel);
  } else {
    w[S >> 2] = perm_b32(m, w[S >> 2], 0x04020100u);
    w[(S >> 2) + 1] = perm_b32(m, w[(S >> 2) + 1], 0x03020106u);
  }
}
// Message S from the low byte of m.
template <int S>
__host__ __device__ LZ_INLINE void put_msg1(uint32_t *w, uint32_t m) {
  constexpr int p = S & 3;
  constexpr uint32_t sel = (p == 0 ? 0x03020104u : p == 1 ? 0x03020400u : p == 2 ? 0x03040100u : 0x04020100u);
  w[S >> 2] = perm_b32(m, w[S >> 2], sel);
}

template <int S>
__host__ __device__ LZ_INLINE void put_msg(uint32_t *w, int v) {
  constexpr int sh = (S & 3) * 8;
  w[S >> 2] = (w[S >> 2] & ~(0xFFu << sh)) | ((static_cast<uint32_t>(v) & 0xFFu) << sh);
}

// Byte offset of posterior (i - SH) mod 128 of a 128-entry block from 2 i: (2 i + 256 - 2 SH) mod 256, the block's
// own offset going into the LDS instruction's immediate. On the GPU one add whose result keeps only its low byte
// (a sub-dword write, dst_sel BYTE_0) in place of an add and a mask; its constant is set in a scalar register within
// the same statement, so no scalar register stays taken.
template <int SH>
__host__ __device__ LZ_INLINE int rot_off(int i2) {
  if constexpr (SH == 0) {
    return i2;
  } else {
#ifdef __HIP_DEVICE_COMPILE__
    int r, c;
    asm("s_movk_i32 %1, %2\n\tv_add_u32_sdwa %0, %1, %3 dst_sel:BYTE_0 dst_unused:UNUSED_PAD src0_sel:DWORD "
        "src1_sel:DWORD"
        : "=v"(r), "=&s"(c)
        : "n"(256 - 2 * SH), "v"(i2));
    return r;
#else
    return (i2 + (256 - 2 * SH)) & 0xFF;
#endif
  }
}

__host__ __device__ LZ_INLINE int16_t &at_byte(int16_t *base, int off) {
  return *reinterpret_cast<int16_t *>(reinterpret_cast<char *>(base) + off);
}
__host__ __device__ LZ_INLINE int16_t at_byte(const int16_t *base, int off) {
  return *reinterpret_cast<const int16_t *>(reinterpret_cast<const char *>(base) + off);
}

// Per 16-bit half of q: 16 |q| + J as max(16 q + J, J - 16 q) (|q| < 2048). On the GPU two packed multiply-adds
// whose constants are inline (16, -16 and J broadcast to both halves) and a packed max.
template <int J>
__host__ __device__ LZ_INLINE v2i16 key16(v2i16 q) {
#ifdef __HIP_DEVICE_COMPILE__
  v2i16 x, y;
  asm("v_pk_mad_i16 %0, %1, 16, %2 op_sel_hi:[1,0,0]" : "=v"(x) : "v"(q), "n"(J));
  asm("v_pk_mad_i16 %0, %1, -16, %2 op_sel_hi:[1,0,0]" : "=v"(y) : "v"(q), "n"(J));
  return x > y ? x : y;
#else
  v2i16 r;
  for (int h = 0; h < 2; h++) {
    const int a = q[h] < 0 ? -q[h] : q[h];
    r[h] = static_cast<int16_t>(16 * a + J);
  }
  return r;
#endif
}

// Block row R of code CI for the lane's check position i.
template <int CI, int R>
__host__ __device__ LZ_INLINE void layer(int16_t *L, Msgs<CI> &M, int lane) {
  using S = Shape<CI>;
  const int i2 = fresh(2 * lane);
  constexpr int D0 = S::d0(R), ND = S::nd(R), NS = ND + 2, RB = S::rbase(R), K = S::K;
  static_assert(NS <= 32, "edge indices fit the key's five bits");
  // the check's slots: its circulants' data bits, the parity bit before it, its own (ldpc.c:589-602); byte offsets
  // into L
  int adr[NS];
  static_for<0, ND>([&](auto e) {
    constexpr int col = S::col(D0 + e), sh = S::shift(D0 + e);
    adr[e] = col * 256 + rot_off<sh>(i2);
  });
  adr[ND] = R > 0 ? 2 * (K + (R - 1) * 128) + i2 : 2 * (K + (S::MB - 1) * 128 - 1) + i2;
  adr[ND + 1] = 2 * (K + R * 128) + i2;
  // check 0 (block row 0, position 0) has no parity bit before it: the reference reads it as a certain 0 of 8191 with
  // its message zeroed. Any value above every real |q| gives the same messages and posteriors (it is never a
  // minimum, adds no sign and is not written back); 2000 keeps its key within 16 bits.
  const bool absent = R == 0 && i2 == 0;
  // pass 1, two edges at a time in 16-bit halves (edge 2j low, 2j + 1 high; an odd row's last high half is a phantom
  // edge, q = 0, key above every real one): q = posterior - message, the sign product, and per half the key
  // 16 |q| + j (j < 16). Doubled, plus the half, that is |q| << 5 | e: the smallest is the first smallest magnitude
  // (the reference's strict "<" keeps the first of equal minima), the second smallest the next magnitude up or an
  // equal later one. A posterior is its channel value (|int8| <= 128) plus its column's messages (|m| <= 127), so it
  // never reaches the reference's +-8191 clamp and |q| <= 128 + 127 (column degree - 1) < 2047: 16 |q| + j fits a
  // signed 16-bit half, computed as max(16 q + j, j - 16 q) (two packed multiply-adds and a packed max).
  static_assert(128 + 127 * S::max_col_degree() < 8191, "the posterior clamp never acts");
  static_assert(16 * 2000 + 15 < 32767 && 128 + 127 * (S::max_col_degree() - 1) < 2000, "keys fit 16 bits");
  constexpr int NP = (NS + 1) / 2;
  static_assert(NP <= 16, "pair indices fit the key's four bits");
  v2i16 q2[NP];
  v2i16 k1 = {0x7FFF, 0x7FFF}, k2 = {0x7FFF, 0x7FFF};
  uint32_t sx = 0;
  static_for<0, NP>([&](auto j) {
    constexpr int e0 = 2 * j, e1 = 2 * j + 1;
    int p0 = at_byte(L, adr[e0]), m0 = get_msg<RB + e0>(M.m);
    if (e0 == ND && absent) {
      p0 = 2000;
      m0 = 0;
    }
    uint32_t w = static_cast<uint16_t>(p0 - m0);
    if constexpr (e1 < NS) {
      int p1 = at_byte(L, adr[e1]);
#ifdef __HIP_DEVICE_COMPILE__
      if constexpr (!(R == 0 && e1 == ND)) {
        // the high half written in place: posterior minus the sign-extended message byte, the low half kept
        asm("v_sub_u16_sdwa %0, %1, sext(%2) dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:WORD_0 "
            "src1_sel:BYTE_%3"
            : "+v"(w)
            : "v"(p1), "v"(M.m[(RB + e1) >> 2]), "n"((RB + e1) & 3));
      } else
#endif
      {
        int m1 = get_msg<RB + e1>(M.m);
        if (e1 == ND && absent) {
          p1 = 2000;
          m1 = 0;
        }
        w |= static_cast<uint32_t>(p1 - m1) << 16;
      }
    }
    sx ^= w;
    v2i16 q;
    __builtin_memcpy(&q, &w, 4);
    q2[j] = q;
    v2i16 key = key16<j>(q);
    if constexpr (e1 >= NS) key[1] = 0x7FFF;
    const v2i16 lo = k1 < key ? k1 : key, hi = k1 < key ? key : k1;
    k2 = hi < k2 ? hi : k2;  // the middle of k1 <= k2 and key
    k1 = lo;
  });
  // the halves merged: the first minimum is the smaller half key, the low half's on equal keys (edge 2j comes before
  // 2j + 1); the second the smaller of the other half's first and both halves' second (only its magnitude counts)
  const int a0 = k1[0], a1 = k1[1], b0 = k2[0], b1 = k2[1];
  const bool hi_first = a1 < a0;
  const int key1 = hi_first ? a1 : a0, kmax = hi_first ? a0 : a1, kb = b0 < b1 ? b0 : b1;
  const int key2 = kmax < kb ? kmax : kb;
  // (m1, m2) = min((|q| 13) >> 4, 127) in the two halves at once, |q| = key >> 4 (|q| 13 < 2^16)
  v2u16 mm = {static_cast<uint16_t>(key1), static_cast<uint16_t>(key2)};
  mm = ((mm >> 4) * static_cast<uint16_t>(13)) >> 4;
  const v2u16 cap = {127, 127};
  mm = mm < cap ? mm : cap;
  // the magnitudes with the check's sign product (XOR of every q's sign) applied once: an edge's message is that,
  // negated where its own q is negative (the product of the other edges' signs)
  const int16_t sg = static_cast<int16_t>(sx ^ (sx >> 16));  // the check's sign in bit 15
  const v2i16 cs = v2i16{sg, sg} >> 15;
  v2i16 m12;
  __builtin_memcpy(&m12, &mm, 4);
  m12 = (m12 ^ cs) - cs;
  const v2i16 m21 = {m12[1], m12[0]};
  // pass 2, the new messages and posteriors in the same halves: the pair holding the first minimum takes m2 in that
  // half, the others m1 in both
  const v2i16 m11 = {m12[0], m12[0]};
  const v2i16 m11d = hi_first ? m12 : m21;
  const int argpair = key1 & 15;
  static_for<0, NP>([&](auto j) {
    constexpr int e0 = 2 * j, e1 = 2 * j + 1;
    const v2i16 q = q2[j];
    const v2i16 mag = argpair == j ? m11d : m11;
    const v2i16 qs = q >> 15;  // -1 where the edge's own q is negative
    const v2i16 msg = (mag ^ qs) - qs;
    const v2i16 t = q + msg;  // within +-8191 (see above): the reference's clamp is the identity here
    uint32_t mw;
    __builtin_memcpy(&mw, &msg, 4);
    if constexpr (e1 < NS) put_msg2<RB + e0>(M.m, mw);
    else put_msg1<RB + e0>(M.m, mw);
    if (!(e0 == ND && absent)) at_byte(L, adr[e0]) = t[0];
    if constexpr (e1 < NS) {
      if (!(e1 == ND && absent)) at_byte(L, adr[e1]) = t[1];
    }
  });
}

// Block row r (runtime) dispatched to its compile-time form.
template <int CI, int R = 0>
__host__ __device__ LZ_INLINE void layer_at(int r, int16_t *L, Msgs<CI> &M, int lane) {
  if constexpr (R < Shape<CI>::MB) {
    if (r == R) layer<CI, R>(L, M, lane);
    else layer_at<CI, R + 1>(r, L, M, lane);
  }
}

// The lane's unsatisfied checks over every block row (the sign of the XOR of a check's posteriors).
template <int CI>
__host__ __device__ LZ_INLINE int unsat(const int16_t *L, int lane) {
  using S = Shape<CI>;
  int bad = 0;
  static_for<0, S::MB>([&](auto r) {
    constexpr int R = r, D0 = S::d0(R), ND = S::nd(R), K = S::K;
    const int i2 = fresh(2 * lane);
    int x = at_byte(L, 2 * (K + R * 128) + i2);
    static_for<0, ND>([&](auto e) {
      constexpr int col = S::col(D0 + e), sh = S::shift(D0 + e);
      x ^= at_byte(L, col * 256 + rot_off<sh>(i2));
    });
    if (R > 0) x ^= at_byte(L, 2 * (K + (R - 1) * 128) + i2);
    else x ^= i2 > 0 ? at_byte(L, 2 * (K + (S::MB - 1) * 128 - 1) + i2) : 0;
    bad += x < 0;
  });
  return bad;
}

// Whether any of the wave's checks is unsatisfied, block row by block row, stopping at the first row that has one
// (the reference's early-out test, ldpc.c layers_clear, for the iterations where only "all satisfied" decides).
// On the GPU the answer is the wave's (a ballot); on the host, the lane's.
template <int CI>
__host__ __device__ LZ_INLINE int unsat_any(const int16_t *L, int lane) {
  using S = Shape<CI>;
  int found = 0;
  static_for<0, S::MB>([&](auto r) {
    if (found) return;
    constexpr int R = r, D0 = S::d0(R), ND = S::nd(R), K = S::K;
    const int i2 = fresh(2 * lane);
    int x = at_byte(L, 2 * (K + R * 128) + i2);
    static_for<0, ND>([&](auto e) {
      constexpr int col = S::col(D0 + e), sh = S::shift(D0 + e);
      x ^= at_byte(L, col * 256 + rot_off<sh>(i2));
    });
    if (R > 0) x ^= at_byte(L, 2 * (K + (R - 1) * 128) + i2);
    else x ^= i2 > 0 ? at_byte(L, 2 * (K + (S::MB - 1) * 128 - 1) + i2) : 0;
#ifdef __HIP_DEVICE_COMPILE__
    found = __builtin_amdgcn_ballot_w64(x < 0) != 0;
#else
    found = x < 0;
#endif
  });
  return found;
}

template <int CI>
__host__ __device__ LZ_INLINE void clear_msgs(Msgs<CI> &M) {
#pragma unroll
  for (int w = 0; w < Shape<CI>::WORDS; w++) M.m[w] = 0;
}

// The messages out in the reference decoder's layout: slot s of check position i at s * 128 + i.
template <int CI>
__host__ __device__ LZ_INLINE void dump_msgs(const Msgs<CI> &M, int8_t *out, int i) {
  static_for<0, Shape<CI>::SLOTS>([&](auto s) { out[s * 128 + i] = static_cast<int8_t>(get_msg<s>(M.m)); });
}

}  // namespace tm
}  // namespace lz
