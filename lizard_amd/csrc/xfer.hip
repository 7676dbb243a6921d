// The transfer's BLAKE3 tree on the GPU (the receiver's check of a recovered chunk and of the file, liblizard/src/xfer.c
// :134-229): a file is cut into chunks of 2^k bytes, each a whole subtree of BLAKE3's tree (it starts at a multiple of
// its own size), so each chunk's chaining value comes from the chunk alone and the chunks' values combine to the
// file's root, which is what b3sum prints.
//   k_b3_leaves  one lane a 1 KiB BLAKE3 chunk: its 16 blocks compressed in turn (counter = the chunk's index in the
//                input), ROOT on the last block where the input is that one chunk
//   k_b3_level   one level of every subtree at once: node j of the next level is the parent of nodes 2j and 2j + 1,
//                or node 2j carried up where it ends an odd count (BLAKE3's left-full tree); ROOT on the input's top
//                node when the whole input is one tree
//   k_b3_gather  each subtree's top node into the chunks' chaining values
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <vector>

#include "b3.hpp"
#include "common.hpp"
#include "xfer.hpp"

namespace lz {
namespace {

__global__ __launch_bounds__(256) void k_b3_leaves(const uint8_t *data, uint64_t len, uint64_t nchunks, int root,
                                                   uint32_t *cvs) {
  const uint64_t c = blockIdx.x * 256ull + threadIdx.x;
  if (c >= nchunks) return;
  const uint64_t at = c * B3_CHUNK;
  const uint32_t clen = static_cast<uint32_t>(len - at < B3_CHUNK ? len - at : B3_CHUNK);
  uint32_t cv[8];
#pragma unroll
  for (int i = 0; i < 8; i++) cv[i] = b3_iv(i);
  if (clen == B3_CHUNK) {
    const uint4 *p = reinterpret_cast<const uint4 *>(data + at);
    for (int b = 0; b < 16; b++) {
      uint32_t m[16];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint4 v = p[4 * b + q];
        m[4 * q] = v.x;
        m[4 * q + 1] = v.y;
        m[4 * q + 2] = v.z;
        m[4 * q + 3] = v.w;
      }
      const uint32_t flags = (b == 0 ? B3_CHUNK_START : 0u) | (b == 15 ? B3_CHUNK_END | (root ? B3_ROOT : 0u) : 0u);
      b3_compress(cv, m, c, B3_BLOCK, flags);
    }
  } else {
    // the input's short last chunk (or an empty input): its blocks zero-filled past the end
    const int nb = clen ? static_cast<int>((clen + B3_BLOCK - 1) / B3_BLOCK) : 1;
    for (int b = 0; b < nb; b++) {
      const uint32_t base = static_cast<uint32_t>(b) * B3_BLOCK;
      const uint32_t bl = clen - base < B3_BLOCK ? clen - base : B3_BLOCK;
      uint32_t m[16];
#pragma unroll
      for (int w = 0; w < 16; w++) {
        uint32_t word = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const uint32_t i = 4 * w + k;
          if (i < bl) word |= static_cast<uint32_t>(data[at + base + i]) << (8 * k);
        }
        m[w] = word;
      }
      const uint32_t flags = (b == 0 ? B3_CHUNK_START : 0u) | (b == nb - 1 ? B3_CHUNK_END | (root ? B3_ROOT : 0u) : 0u);
      b3_compress(cv, m, c, bl, flags);
    }
  }
  uint4 *o = reinterpret_cast<uint4 *>(cvs + c * 8);
  o[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
  o[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

// subtree s keeps its nodes at [s * P, s * P + n): n = n_full, the last subtree's n_last
__global__ __launch_bounds__(256) void k_b3_level(const uint32_t *in, uint32_t *out, uint64_t S, uint64_t P,
                                                  uint64_t n_full, uint64_t n_last, int root) {
  const uint64_t half = (n_full + 1) / 2;
  const uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  const uint64_t s = t / half, j = t - s * half;
  if (s >= S) return;
  const uint64_t n = s + 1 == S ? n_last : n_full;
  if (j >= (n + 1) / 2) return;
  const uint32_t *a = in + (s * P + 2 * j) * 8;
  uint32_t o[8];
  if (2 * j + 1 < n) {
    b3_parent(a, a + 8, root && n == 2 ? B3_ROOT : 0u, o);
  } else {
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = a[i];
  }
  uint4 *d = reinterpret_cast<uint4 *>(out + (s * P + j) * 8);
  d[0] = make_uint4(o[0], o[1], o[2], o[3]);
  d[1] = make_uint4(o[4], o[5], o[6], o[7]);
}

__global__ __launch_bounds__(256) void k_b3_gather(const uint32_t *in, uint64_t S, uint64_t P, uint32_t *cvs) {
  const uint64_t s = blockIdx.x * 256ull + threadIdx.x;
  if (s >= S) return;
  const uint4 *a = reinterpret_cast<const uint4 *>(in + s * P * 8);
  uint4 *o = reinterpret_cast<uint4 *>(cvs + s * 8);
  o[0] = a[0];
  o[1] = a[1];
}

unsigned blocks_for(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// leaves into a, then levels between a and b; the buffer holding the subtrees' tops is returned
uint32_t *tree(const uint8_t *d_data, uint64_t len, uint64_t S, uint64_t P, uint64_t n_last, int root, uint32_t *a,
               uint32_t *b, hipStream_t s) {
  const uint64_t nchunks = len ? (len + B3_CHUNK - 1) / B3_CHUNK : 1;
  hipLaunchKernelGGL(k_b3_leaves, dim3(blocks_for(nchunks)), dim3(256), 0, s, d_data, len, nchunks,
                     root && nchunks == 1 ? 1 : 0, a);
  uint64_t nf = P, nl = n_last;
  while (nf > 1) {
    const uint64_t threads = S * ((nf + 1) / 2);
    hipLaunchKernelGGL(k_b3_level, dim3(blocks_for(threads)), dim3(256), 0, s, a, b, S, P, nf, nl, root);
    std::swap(a, b);
    nf = (nf + 1) / 2;
    nl = (nl + 1) / 2;
  }
  return a;
}

}  // namespace

uint64_t b3_chunk_cvs(const uint8_t *d_file, uint64_t length, int chunk_log2, uint32_t *d_a, uint32_t *d_b,
                      uint32_t *d_cvs, hipStream_t s) {
  const uint64_t csize = 1ull << chunk_log2, P = csize / B3_CHUNK;
  const uint64_t S = (length + csize - 1) / csize;
  const uint64_t last = length - (S - 1) * csize, n_last = (last + B3_CHUNK - 1) / B3_CHUNK;
  uint32_t *top = tree(d_file, length, S, P, n_last, 0, d_a, d_b, s);
  hipLaunchKernelGGL(k_b3_gather, dim3(blocks_for(S)), dim3(256), 0, s, top, S, P, d_cvs);
  hip_check(hipGetLastError(), "blake3 launches");
  return S;
}

void b3_hash_device(const uint8_t *d_data, uint64_t length, uint32_t *d_a, uint32_t *d_b, uint32_t *d_out,
                    hipStream_t s) {
  const uint64_t nchunks = length ? (length + B3_CHUNK - 1) / B3_CHUNK : 1;
  uint32_t *top = tree(d_data, length, 1, nchunks, nchunks, 1, d_a, d_b, s);
  hip_check(hipMemcpyAsync(d_out, top, 32, hipMemcpyDeviceToDevice, s), "blake3 root");
  hip_check(hipGetLastError(), "blake3 launches");
}

uint64_t b3_scratch_words(uint64_t length) {
  const uint64_t nchunks = length ? (length + B3_CHUNK - 1) / B3_CHUNK : 1;
  return nchunks * 8;
}

// ---- the host's side: the root from the chunks' values, and a whole hash for small inputs ----

namespace {
void merge(const uint32_t *cvs, uint32_t n, int root, uint32_t out[8]) {
  if (n == 1) {
    std::memcpy(out, cvs, 32);
    return;
  }
  uint32_t p = 1;  // the largest power of two below n: the left subtree's share
  while (2 * p < n) p *= 2;
  uint32_t l[8], r[8];
  merge(cvs, p, 0, l);
  merge(cvs + static_cast<size_t>(p) * 8, n - p, 0, r);
  b3_parent(l, r, root ? B3_ROOT : 0u, out);
}
}  // namespace

void b3_root(const uint32_t *cvs, uint32_t n, uint32_t out[8]) { merge(cvs, n, 1, out); }

void b3_hash_host(const uint8_t *data, uint64_t len, uint32_t out[8]) {
  const uint64_t nchunks = len ? (len + B3_CHUNK - 1) / B3_CHUNK : 1;
  std::vector<uint32_t> cvs(nchunks * 8);
  for (uint64_t c = 0; c < nchunks; c++) {
    const uint64_t at = c * B3_CHUNK;
    const uint32_t clen = static_cast<uint32_t>(len - at < B3_CHUNK ? len - at : B3_CHUNK);
    uint32_t cv[8];
    for (int i = 0; i < 8; i++) cv[i] = b3_iv(i);
    const int nb = clen ? static_cast<int>((clen + B3_BLOCK - 1) / B3_BLOCK) : 1;
    for (int b = 0; b < nb; b++) {
      const uint32_t base = static_cast<uint32_t>(b) * B3_BLOCK, bl = clen - base < B3_BLOCK ? clen - base : B3_BLOCK;
      uint32_t m[16] = {};
      for (uint32_t i = 0; i < bl; i++) m[i / 4] |= static_cast<uint32_t>(data[at + base + i]) << (8 * (i % 4));
      const uint32_t flags = (b == 0 ? B3_CHUNK_START : 0u) |
                             (b == nb - 1 ? B3_CHUNK_END | (nchunks == 1 ? B3_ROOT : 0u) : 0u);
      b3_compress(cv, m, c, bl, flags);
    }
    std::memcpy(&cvs[c * 8], cv, 32);
  }
  if (nchunks == 1) {
    std::memcpy(out, cvs.data(), 32);
    return;
  }
  merge(cvs.data(), static_cast<uint32_t>(nchunks), 1, out);
}

}  // namespace lz
