// The transfer's control blocks on the host: the header (a file's length, chunk size, codec, BLAKE3 root, name and
// media type, and for a one-chunk file its bytes as sent and its fountain seed attempt) and the manifest (each
// chunk's chaining value, bytes as sent and seed attempt, twelve entries a block, tagged with the root's first 8
// bytes), as 472-byte block payloads. The layout is the transfer's wire format (liblizard/src/xfer.h:40-92, version
// 3); the rules every block is held to, on writing and on reading, are xfer.c's (xfer.c:11-132): a header or an entry
// that no file could have is refused, and every byte past a length is zero.
#include <cstring>
#include <vector>

#include "xfer.hpp"
#include "common.hpp"

namespace lz {
namespace {

constexpr int PAYLOAD = 472, CV = 32, ENTRY = CV + 4 + 1, PER_BLOCK = 12;
constexpr int VERSION = 3, HASH_BLAKE3 = 1, CODEC_NONE = 0, CODEC_ZSTD = 1, LOG2_MIN = 10, LOG2_MAX = 24;
constexpr uint32_t MAX_CHUNKS = (1u << 14) - 1, SYMBOL_BITS = 18, CONTROL = MAX_CHUNKS;
constexpr uint32_t MAX_MANIFEST = (MAX_CHUNKS + PER_BLOCK - 1) / PER_BLOCK;
constexpr int NAME_BYTES = 255, TYPE_BYTES = 158;
constexpr int ERR = -1, ERR_VERSION = -2, ERR_HASH = -3, ERR_CODEC = -4;

// the header's byte offsets
enum : int {
  H_VERSION = 0, H_HASH = 1, H_LOG2 = 2, H_CODEC = 3, H_SEED = 4, H_NAME_LEN = 5, H_TYPE_LEN = 6, H_RESERVED = 7,
  H_LENGTH = 8, H_CHUNKS = 16, H_ROOT = 20, H_NAME = 52, H_TYPE = H_NAME + NAME_BYTES, H_SENT = H_TYPE + TYPE_BYTES,
  H_TAIL = H_SENT + 4,
};
// the manifest's
enum : int { M_TAG = 0, M_TAG_BYTES = 8, M_ENTRIES = 8 };
static_assert(H_TAIL == 469 && M_ENTRIES + PER_BLOCK * ENTRY == 452, "the version 3 layout");

uint64_t read_le(const uint8_t *p, int n) {
  uint64_t v = 0;
  while (n--) v = (v << 8) | p[n];
  return v;
}
void write_le(uint8_t *p, uint64_t v, int n) {
  for (int i = 0; i < n; i++, v >>= 8) p[i] = static_cast<uint8_t>(v);
}
bool all_zero(const uint8_t *p, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (p[i]) return false;
  return true;
}

uint32_t chunk_count(uint64_t length, int log2) {
  if (log2 < LOG2_MIN || log2 > LOG2_MAX) return 0;
  const uint64_t n = (length >> log2) + ((length & ((uint64_t{1} << log2) - 1)) != 0);  // the ceiling, any length
  return n > MAX_CHUNKS ? 0 : static_cast<uint32_t>(n);
}

uint64_t chunk_len(const lizx_xfer_header &h, uint32_t index) {
  // a chunk size outside the format's has no chunks (and no shift past 63 bits)
  if (index >= h.chunks || h.chunk_log2 < LOG2_MIN || h.chunk_log2 > LOG2_MAX) return 0;
  const uint64_t size = uint64_t{1} << h.chunk_log2, at = static_cast<uint64_t>(index) << h.chunk_log2;
  if (at >= h.length) return 0;  // chunks beyond the length (an inconsistent header)
  const uint64_t left = h.length - at;
  return left < size ? left : size;
}

uint32_t source_blocks(uint32_t sent) { return (sent + PAYLOAD - 1) / PAYLOAD; }

// whether `sent` and `seed` can be chunk index's: 1 to its length, shorter only under a codec, a seed attempt only
// for a chunk that is fountained (two source blocks or more)
bool sent_ok(const lizx_xfer_header &h, uint32_t index, uint32_t sent, int seed) {
  const uint64_t len = chunk_len(h, index);
  if (!len || !sent || sent > len) return false;
  if (sent < len && h.codec == CODEC_NONE) return false;
  return seed >= 0 && seed <= 255 && !(seed && source_blocks(sent) < 2);
}

int manifest_blocks(uint32_t chunks) {
  return chunks < 2 || chunks > MAX_CHUNKS ? 0 : static_cast<int>((chunks + PER_BLOCK - 1) / PER_BLOCK);
}

// 0, or the code of the first rule the header breaks (the version, the hash, the codec, then the rest)
int check_header(const lizx_xfer_header &h) {
  if (h.version != VERSION) return ERR_VERSION;
  if (h.hash != HASH_BLAKE3) return ERR_HASH;
  if (h.codec != CODEC_NONE && h.codec != CODEC_ZSTD) return ERR_CODEC;
  if (h.chunk_log2 < LOG2_MIN || h.chunk_log2 > LOG2_MAX) return ERR;
  if (h.chunks != chunk_count(h.length, h.chunk_log2) || (h.length && !h.chunks)) return ERR;
  if (h.type_len > TYPE_BYTES) return ERR;
  if (!all_zero(h.name + h.name_len, NAME_BYTES - h.name_len) || !all_zero(h.type + h.type_len, TYPE_BYTES - h.type_len))
    return ERR;
  if (h.chunks == 1) return sent_ok(h, 0, h.sent, h.seed) ? 0 : ERR;
  return h.sent || h.seed ? ERR : 0;
}

uint32_t entries_of(uint32_t chunks, uint32_t m) {
  const uint32_t first = m * PER_BLOCK;
  return chunks - first < static_cast<uint32_t>(PER_BLOCK) ? chunks - first : PER_BLOCK;
}

}  // namespace
}  // namespace lz

using namespace lz;

extern "C" {

LIZ_API uint32_t lizx_xfer_chunk_count(uint64_t length, int chunk_log2) { return chunk_count(length, chunk_log2); }

LIZ_API uint64_t lizx_xfer_chunk_len(const lizx_xfer_header *h, uint32_t index) { return h ? chunk_len(*h, index) : 0; }

LIZ_API uint32_t lizx_xfer_blocks(uint32_t sent) { return source_blocks(sent); }

LIZ_API int lizx_xfer_sent_ok(const lizx_xfer_header *h, uint32_t index, uint32_t sent, int seed) {
  return h && sent_ok(*h, index, sent, seed) ? 0 : ERR;
}

LIZ_API int lizx_xfer_manifest_blocks(uint32_t chunks) { return manifest_blocks(chunks); }

LIZ_API int lizx_xfer_id_kind(uint32_t id) {
  const uint32_t chunk = id >> SYMBOL_BITS, symbol = id & ((1u << SYMBOL_BITS) - 1);
  if (chunk != CONTROL) return LIZX_XFER_DATA;
  return !symbol ? LIZX_XFER_HEADER : symbol <= MAX_MANIFEST ? LIZX_XFER_MANIFEST : LIZX_XFER_RESERVED;
}

LIZ_API int lizx_xfer_header_init(lizx_xfer_header *h, uint64_t length, int chunk_log2) {
  if (!h) return ERR;
  std::memset(h, 0, sizeof *h);
  const uint32_t chunks = chunk_count(length, chunk_log2);
  if (chunk_log2 < LOG2_MIN || chunk_log2 > LOG2_MAX || (length && !chunks)) return ERR;
  h->version = VERSION;
  h->hash = HASH_BLAKE3;
  h->chunk_log2 = static_cast<uint8_t>(chunk_log2);
  h->codec = CODEC_NONE;
  h->length = length;
  h->chunks = chunks;
  return 0;
}

LIZ_API int lizx_xfer_header_write(const lizx_xfer_header *h, uint8_t *payload) {
  if (!h || !payload || check_header(*h)) return ERR;
  std::memset(payload, 0, PAYLOAD);
  const uint8_t bytes[7] = {h->version, h->hash, h->chunk_log2, h->codec, h->seed, h->name_len, h->type_len};
  std::memcpy(payload + H_VERSION, bytes, sizeof bytes);
  write_le(payload + H_LENGTH, h->length, 8);
  write_le(payload + H_CHUNKS, h->chunks, 4);
  std::memcpy(payload + H_ROOT, h->root, CV);
  std::memcpy(payload + H_NAME, h->name, h->name_len);
  std::memcpy(payload + H_TYPE, h->type, h->type_len);
  write_le(payload + H_SENT, h->sent, 4);
  return 0;
}

LIZ_API int lizx_xfer_header_parse(const uint8_t *payload, lizx_xfer_header *h) {
  if (!payload || !h) return ERR;
  std::memset(h, 0, sizeof *h);
  h->version = payload[H_VERSION];
  if (h->version != VERSION) return ERR_VERSION;
  h->hash = payload[H_HASH];
  h->chunk_log2 = payload[H_LOG2];
  h->codec = payload[H_CODEC];
  h->seed = payload[H_SEED];
  h->name_len = payload[H_NAME_LEN];
  h->type_len = payload[H_TYPE_LEN];
  if (payload[H_RESERVED] || h->type_len > TYPE_BYTES || !all_zero(payload + H_TAIL, PAYLOAD - H_TAIL)) return ERR;
  h->length = read_le(payload + H_LENGTH, 8);
  h->chunks = static_cast<uint32_t>(read_le(payload + H_CHUNKS, 4));
  std::memcpy(h->root, payload + H_ROOT, CV);
  std::memcpy(h->name, payload + H_NAME, NAME_BYTES);
  std::memcpy(h->type, payload + H_TYPE, TYPE_BYTES);
  h->sent = static_cast<uint32_t>(read_le(payload + H_SENT, 4));
  return check_header(*h);
}

LIZ_API int lizx_xfer_manifest_write(const uint8_t *cvs, const uint32_t *sent, const uint8_t *seeds, uint32_t chunks,
                                     const uint8_t *root, uint32_t m, uint8_t *payload) {
  if (!cvs || !sent || !seeds || !root || !payload || m >= static_cast<uint32_t>(manifest_blocks(chunks))) return ERR;
  const uint32_t first = m * PER_BLOCK, n = entries_of(chunks, m);
  std::memset(payload, 0, PAYLOAD);
  std::memcpy(payload + M_TAG, root, M_TAG_BYTES);
  for (uint32_t i = 0; i < n; i++) {
    uint8_t *e = payload + M_ENTRIES + static_cast<size_t>(i) * ENTRY;
    std::memcpy(e, cvs + static_cast<size_t>(first + i) * CV, CV);
    write_le(e + CV, sent[first + i], 4);
    e[CV + 4] = seeds[first + i];
  }
  return 0;
}

LIZ_API int lizx_xfer_manifest_parse(const uint8_t *payload, const lizx_xfer_header *h, uint32_t m, uint8_t *cvs,
                                     uint32_t *sent, uint8_t *seeds) {
  if (!payload || !h || !cvs || !sent || !seeds) return ERR;
  const uint32_t chunks = h->chunks;
  if (m >= static_cast<uint32_t>(manifest_blocks(chunks)) || std::memcmp(payload + M_TAG, h->root, M_TAG_BYTES))
    return ERR;
  const uint32_t first = m * PER_BLOCK, n = entries_of(chunks, m);
  const size_t used = M_ENTRIES + static_cast<size_t>(n) * ENTRY;
  if (!all_zero(payload + used, PAYLOAD - used)) return ERR;
  // every entry checked before any is kept: a refused block writes nothing
  for (uint32_t i = 0; i < n; i++) {
    const uint8_t *e = payload + M_ENTRIES + static_cast<size_t>(i) * ENTRY;
    if (!sent_ok(*h, first + i, static_cast<uint32_t>(read_le(e + CV, 4)), e[CV + 4])) return ERR;
  }
  for (uint32_t i = 0; i < n; i++) {
    const uint8_t *e = payload + M_ENTRIES + static_cast<size_t>(i) * ENTRY;
    std::memcpy(cvs + static_cast<size_t>(first + i) * CV, e, CV);
    sent[first + i] = static_cast<uint32_t>(read_le(e + CV, 4));
    seeds[first + i] = e[CV + 4];
  }
  return 0;
}

LIZ_API int lizx_xfer_manifest_check(const uint8_t *cvs, uint32_t chunks, const uint8_t *root) {
  if (!cvs || !root || chunks < 2 || chunks > MAX_CHUNKS) return ERR;
  std::vector<uint32_t> w(static_cast<size_t>(chunks) * 8);
  std::memcpy(w.data(), cvs, w.size() * 4);  // little-endian words
  uint32_t r[8];
  b3_root(w.data(), chunks, r);
  return std::memcmp(r, root, CV) ? ERR : 0;
}

}  // extern "C"
