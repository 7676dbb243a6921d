// The transfer's BLAKE3 tree (liblizard/src/xfer.c:134-229): the chunks' chaining values and the whole input's hash
// on the GPU, the root from the chunks' values and a whole hash on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lz {

// Words of scratch each of the two work buffers needs for an input of `length` bytes (8 a 1 KiB BLAKE3 chunk).
uint64_t b3_scratch_words(uint64_t length);
// Every 2^chunk_log2-byte chunk's chaining value (8 words each, little-endian) of a device input (16-byte aligned) into
// d_cvs; returns the chunk count. d_a, d_b: work buffers of b3_scratch_words(length) words.
uint64_t b3_chunk_cvs(const uint8_t *d_file, uint64_t length, int chunk_log2, uint32_t *d_a, uint32_t *d_b,
                      uint32_t *d_cvs, hipStream_t s);
// BLAKE3 of a whole device input (16-byte aligned) into d_out (8 words).
void b3_hash_device(const uint8_t *d_data, uint64_t length, uint32_t *d_a, uint32_t *d_b, uint32_t *d_out,
                    hipStream_t s);
// The root (the file's hash) from its n >= 2 chunks' chaining values: BLAKE3's left-full merge.
void b3_root(const uint32_t *cvs, uint32_t n, uint32_t out[8]);
// BLAKE3 of a host input.
void b3_hash_host(const uint8_t *data, uint64_t len, uint32_t out[8]);

}  // namespace lz
