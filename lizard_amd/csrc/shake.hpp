// SHAKE256 (FIPS 202) and the test stream built on it (liblizard/src/shake.h): a block's payload a function of its id.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace lz {
void shake256(const uint8_t *in, size_t in_len, uint8_t *out, size_t out_len);
// SHAKE256 of the id as four little-endian bytes, len bytes of it.
void stream_fill(uint32_t id, uint8_t *out, uint32_t len);
}  // namespace lz
