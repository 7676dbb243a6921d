// BLAKE3's compression function, for the host and the GPU, from the published specification (J. O'Connor, J.-P.
// Aumasson, S. Neves, Z. Wilcox-O'Hearn, "BLAKE3: one function, fast everywhere", 2020, sections 2.1-2.5): a 16-word
// state from the chaining value, the first four IV words, the 64-bit counter, the block's length and its flags; seven
// rounds of the G function over the columns then the diagonals, the message words permuted between rounds; the new
// chaining value the state's halves XORed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lz {

constexpr uint32_t B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8;
constexpr int B3_BLOCK = 64, B3_CHUNK = 1024;

__host__ __device__ constexpr uint32_t b3_iv(int i) {
  return i == 0 ? 0x6A09E667u : i == 1 ? 0xBB67AE85u : i == 2 ? 0x3C6EF372u : i == 3 ? 0xA54FF53Au
       : i == 4 ? 0x510E527Fu : i == 5 ? 0x9B05688Cu : i == 6 ? 0x1F83D9ABu : 0x5BE0CD19u;
}

// The message word G takes as its i-th input (0..15) in round r: the permutation applied r times.
__host__ __device__ constexpr int b3_sched(int r, int i) {
  constexpr int P[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  int k = i;
  for (int t = 0; t < r; t++) k = P[k];
  return k;
}

__host__ __device__ __forceinline__ uint32_t b3_rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

__host__ __device__ __forceinline__ void b3_g(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, uint32_t mx,
                                              uint32_t my) {
  a = a + b + mx;
  d = b3_rotr(d ^ a, 16);
  c = c + d;
  b = b3_rotr(b ^ c, 12);
  a = a + b + my;
  d = b3_rotr(d ^ a, 8);
  c = c + d;
  b = b3_rotr(b ^ c, 7);
}

// cv <- the chaining value of compressing block m (16 little-endian words) under cv.
__host__ __device__ __forceinline__ void b3_compress(uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                                                     uint32_t len, uint32_t flags) {
  uint32_t v0 = cv[0], v1 = cv[1], v2 = cv[2], v3 = cv[3], v4 = cv[4], v5 = cv[5], v6 = cv[6], v7 = cv[7];
  uint32_t v8 = b3_iv(0), v9 = b3_iv(1), v10 = b3_iv(2), v11 = b3_iv(3);
  uint32_t v12 = static_cast<uint32_t>(counter), v13 = static_cast<uint32_t>(counter >> 32), v14 = len, v15 = flags;
#pragma unroll
  for (int r = 0; r < 7; r++) {
    b3_g(v0, v4, v8, v12, m[b3_sched(r, 0)], m[b3_sched(r, 1)]);
    b3_g(v1, v5, v9, v13, m[b3_sched(r, 2)], m[b3_sched(r, 3)]);
    b3_g(v2, v6, v10, v14, m[b3_sched(r, 4)], m[b3_sched(r, 5)]);
    b3_g(v3, v7, v11, v15, m[b3_sched(r, 6)], m[b3_sched(r, 7)]);
    b3_g(v0, v5, v10, v15, m[b3_sched(r, 8)], m[b3_sched(r, 9)]);
    b3_g(v1, v6, v11, v12, m[b3_sched(r, 10)], m[b3_sched(r, 11)]);
    b3_g(v2, v7, v8, v13, m[b3_sched(r, 12)], m[b3_sched(r, 13)]);
    b3_g(v3, v4, v9, v14, m[b3_sched(r, 14)], m[b3_sched(r, 15)]);
  }
  cv[0] = v0 ^ v8;
  cv[1] = v1 ^ v9;
  cv[2] = v2 ^ v10;
  cv[3] = v3 ^ v11;
  cv[4] = v4 ^ v12;
  cv[5] = v5 ^ v13;
  cv[6] = v6 ^ v14;
  cv[7] = v7 ^ v15;
}

// A parent node: the two children's chaining values as the block, under the IV.
__host__ __device__ __forceinline__ void b3_parent(const uint32_t *left, const uint32_t *right, uint32_t flags,
                                                   uint32_t out[8]) {
  uint32_t m[16], cv[8];
  for (int i = 0; i < 8; i++) {
    m[i] = left[i];
    m[8 + i] = right[i];
    cv[i] = b3_iv(i);
  }
  b3_compress(cv, m, 0, B3_BLOCK, B3_PARENT | flags);
  for (int i = 0; i < 8; i++) out[i] = cv[i];
}

}  // namespace lz
