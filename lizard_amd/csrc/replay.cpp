// Replays: the .lizard file (research/SPEC.md 15; liblizard/core/rx/lizfile.h), a run of captures as a receiver's
// decoder was handed them, to run recorded captures through the GPU path.
//   the archive  a ZIP of stored entries (method 0, no data descriptor, no ZIP64; every entry's CRC-32 checked on
//                read, dates 1980-01-01 00:00 on write), the end record's comment "LIZARD replay 1"; entries NNNN.zst
//                a frame each, stats.jsonl where the run has rows, meta.json
//   meta.json    frames, w and h (or sizes[i]), ms, layout
//   a frame      one zstd frame (the system's libzstd, loaded at first use) of the w h residuals of the median edge
//                predictor: pred = min(a, b) where c >= max(a, b), max(a, b) where c <= min(a, b), else a + b - c, with
//                a, b, c the left, upper and upper-left neighbours, 0 outside the frame; r = (p - pred) mod 256
#include <dlfcn.h>
#include <stdint.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "replay.hpp"

namespace lz {
namespace {

// ---- zstd through the system's shared library ----
struct Zstd {
  size_t (*compressBound)(size_t) = nullptr;
  size_t (*compress)(void *, size_t, const void *, size_t, int) = nullptr;
  size_t (*decompress)(void *, size_t, const void *, size_t) = nullptr;
  unsigned (*isError)(size_t) = nullptr;
  size_t (*findFrameCompressedSize)(const void *, size_t) = nullptr;
  bool ok = false;
  Zstd() {
    void *h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("libzstd.so", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    compressBound = reinterpret_cast<size_t (*)(size_t)>(dlsym(h, "ZSTD_compressBound"));
    compress = reinterpret_cast<size_t (*)(void *, size_t, const void *, size_t, int)>(dlsym(h, "ZSTD_compress"));
    decompress = reinterpret_cast<size_t (*)(void *, size_t, const void *, size_t)>(dlsym(h, "ZSTD_decompress"));
    isError = reinterpret_cast<unsigned (*)(size_t)>(dlsym(h, "ZSTD_isError"));
    findFrameCompressedSize = reinterpret_cast<size_t (*)(const void *, size_t)>(dlsym(h, "ZSTD_findFrameCompressedSize"));
    ok = compressBound && compress && decompress && isError && findFrameCompressedSize;
  }
};

const Zstd &zstd() {
  static const Zstd z;
  if (!z.ok) throw Error(LIZ_E_UNSUPPORTED, "no zstd: the system's libzstd.so.1 could not be loaded");
  return z;
}

// ---- CRC-32 (ISO-HDLC: reflected 0xEDB88320, register and result inverted) ----
struct Crc32Table {
  uint32_t t[256];
  Crc32Table() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1)));
      t[i] = c;
    }
  }
};

uint32_t crc32_of(const uint8_t *p, size_t n) {
  static const Crc32Table T;  // built once, thread-safe
  const uint32_t *tab = T.t;
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; i++) c = tab[(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return ~c;
}

inline int med(int a, int b, int c) {
  const int mx = a > b ? a : b, mn = a < b ? a : b;
  return c >= mx ? mn : c <= mn ? mx : a + b - c;
}

// ---- a small JSON reader: what meta.json holds ----
struct Json {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  double num = 0;
  std::string str;
  std::vector<Json> items;
  std::vector<std::string> keys;
  const Json *get(const char *k) const {
    for (size_t i = 0; i < keys.size(); i++)
      if (keys[i] == k) return &items[i];
    return nullptr;
  }
};

struct JsonParser {
  const char *p, *e;
  void ws() {
    while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
  }
  void fail() { throw Error(LIZ_E_FORMAT, "meta.json is not JSON"); }
  std::string string() {
    if (p >= e || *p != '"') fail();
    p++;
    std::string s;
    while (p < e && *p != '"') {
      if (*p == '\\') {
        if (++p >= e) fail();
        const char c = *p;
        s.push_back(c == 'n' ? '\n' : c == 't' ? '\t' : c == 'r' ? '\r' : c == 'b' ? '\b' : c == 'f' ? '\f' : c);
        if (c == 'u') {  // kept as is: the fields read here are ASCII
          s.pop_back();
          s += "\\u";
        }
      } else {
        s.push_back(*p);
      }
      p++;
    }
    if (p >= e) fail();
    p++;
    return s;
  }
  Json value(int depth = 0) {
    if (depth > 64) fail();
    ws();
    if (p >= e) fail();
    Json j;
    if (*p == '{') {
      j.kind = Json::OBJ;
      p++;
      ws();
      if (p < e && *p == '}') {
        p++;
        return j;
      }
      for (;;) {
        ws();
        j.keys.push_back(string());
        ws();
        if (p >= e || *p != ':') fail();
        p++;
        j.items.push_back(value(depth + 1));
        ws();
        if (p < e && *p == ',') {
          p++;
          continue;
        }
        if (p < e && *p == '}') {
          p++;
          return j;
        }
        fail();
      }
    }
    if (*p == '[') {
      j.kind = Json::ARR;
      p++;
      ws();
      if (p < e && *p == ']') {
        p++;
        return j;
      }
      for (;;) {
        j.items.push_back(value(depth + 1));
        ws();
        if (p < e && *p == ',') {
          p++;
          continue;
        }
        if (p < e && *p == ']') {
          p++;
          return j;
        }
        fail();
      }
    }
    if (*p == '"') {
      j.kind = Json::STR;
      j.str = string();
      return j;
    }
    if (e - p >= 4 && !std::strncmp(p, "null", 4)) {
      p += 4;
      return j;
    }
    if (e - p >= 4 && !std::strncmp(p, "true", 4)) {
      p += 4;
      j.kind = Json::BOOL;
      j.num = 1;
      return j;
    }
    if (e - p >= 5 && !std::strncmp(p, "false", 5)) {
      p += 5;
      j.kind = Json::BOOL;
      return j;
    }
    char *end = nullptr;
    const std::string tmp(p, static_cast<size_t>(std::min<ptrdiff_t>(e - p, 64)));
    j.num = std::strtod(tmp.c_str(), &end);
    if (end == tmp.c_str()) fail();
    p += end - tmp.c_str();
    j.kind = Json::NUM;
    return j;
  }
};

uint32_t le32(const uint8_t *b) { return b[0] | b[1] << 8 | b[2] << 16 | static_cast<uint32_t>(b[3]) << 24; }
uint16_t le16(const uint8_t *b) { return static_cast<uint16_t>(b[0] | b[1] << 8); }
void put16(std::vector<uint8_t> &v, uint32_t x) {
  v.push_back(static_cast<uint8_t>(x));
  v.push_back(static_cast<uint8_t>(x >> 8));
}
void put32(std::vector<uint8_t> &v, uint32_t x) {
  put16(v, x & 0xFFFF);
  put16(v, x >> 16);
}

std::string frame_name(size_t i) {
  char b[32];
  std::snprintf(b, sizeof b, "%04zu.zst", i);
  return b;
}

}  // namespace

void loco_residuals(const uint8_t *px, uint32_t w, uint32_t h, uint32_t stride, uint8_t *r) {
  for (uint32_t y = 0; y < h; y++) {
    const uint8_t *row = px + static_cast<size_t>(y) * stride, *up = y ? row - stride : nullptr;
    for (uint32_t x = 0; x < w; x++) {
      const int a = x ? row[x - 1] : 0, b = up ? up[x] : 0, c = x && up ? up[x - 1] : 0;
      r[static_cast<size_t>(y) * w + x] = static_cast<uint8_t>(row[x] - med(a, b, c));
    }
  }
}

void loco_restore(const uint8_t *r, uint32_t w, uint32_t h, uint8_t *px) {
  for (uint32_t y = 0; y < h; y++) {
    uint8_t *row = px + static_cast<size_t>(y) * w;
    const uint8_t *up = y ? row - w : nullptr, *rr = r + static_cast<size_t>(y) * w;
    for (uint32_t x = 0; x < w; x++) {
      const int a = x ? row[x - 1] : 0, b = up ? up[x] : 0, c = x && up ? up[x - 1] : 0;
      row[x] = static_cast<uint8_t>(rr[x] + med(a, b, c));
    }
  }
}

// ---- reading ----

ReplayReader::ReplayReader(const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) throw Error(LIZ_E_IO, "cannot open " + path);
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (size < 22) {
    std::fclose(f);
    throw Error(LIZ_E_FORMAT, "not a ZIP archive");
  }
  file_.resize(static_cast<size_t>(size));
  const size_t got = std::fread(file_.data(), 1, file_.size(), f);
  std::fclose(f);
  if (got != file_.size()) throw Error(LIZ_E_IO, "a short read of " + path);
  const uint8_t *b = file_.data();
  const size_t n = file_.size();
  // the end of central directory record: the last signature whose comment runs to the file's end
  size_t eocd = SIZE_MAX;
  for (size_t at = n - 22 + 1; at-- > 0 && n - at <= 22 + 65535;)
    if (le32(b + at) == 0x06054b50u && at + 22 + le16(b + at + 20) == n) {
      eocd = at;
      break;
    }
  if (eocd == SIZE_MAX) throw Error(LIZ_E_FORMAT, "not a ZIP archive (no end record)");
  const std::string comment(reinterpret_cast<const char *>(b + eocd + 22), le16(b + eocd + 20));
  if (comment.compare(0, 14, "LIZARD replay ") != 0) throw Error(LIZ_E_FORMAT, "not a LIZARD replay (its comment)");
  if (comment != "LIZARD replay 1") throw Error(LIZ_E_FORMAT, "a LIZARD replay of a version this library does not read");
  const uint32_t count = le16(b + eocd + 10), cd_size = le32(b + eocd + 12), cd_at = le32(b + eocd + 16);
  if (static_cast<uint64_t>(cd_at) + cd_size > eocd) throw Error(LIZ_E_FORMAT, "a central directory past the end");
  size_t at = cd_at;
  for (uint32_t k = 0; k < count; k++) {
    if (at + 46 > eocd || le32(b + at) != 0x02014b50u) throw Error(LIZ_E_FORMAT, "a broken central directory");
    Entry e;
    const uint16_t method = le16(b + at + 10);
    e.crc = le32(b + at + 16);
    const uint32_t csize = le32(b + at + 20), usize = le32(b + at + 24);
    const uint16_t nl = le16(b + at + 28), xl = le16(b + at + 30), cl = le16(b + at + 32);
    const uint32_t local = le32(b + at + 42);
    if (at + 46 + nl + xl + cl > eocd) throw Error(LIZ_E_FORMAT, "a broken central directory");
    e.name.assign(reinterpret_cast<const char *>(b + at + 46), nl);
    if (method != 0 || csize != usize) throw Error(LIZ_E_FORMAT, "an entry not stored as it is: " + e.name);
    if (static_cast<uint64_t>(local) + 30 > n || le32(b + local) != 0x04034b50u)
      throw Error(LIZ_E_FORMAT, "a broken local header: " + e.name);
    e.at = local + 30 + le16(b + local + 26) + le16(b + local + 28);
    e.size = csize;
    if (e.at + e.size > n) throw Error(LIZ_E_FORMAT, "an entry past the end: " + e.name);
    entries_.push_back(e);
    at += 46 + nl + xl + cl;
  }
  const Entry *meta = find("meta.json");
  if (!meta) throw Error(LIZ_E_FORMAT, "no meta.json");
  std::vector<uint8_t> mb;
  read_entry(*meta, mb);
  meta_.assign(mb.begin(), mb.end());
  if (const Entry *st = find("stats.jsonl")) {
    std::vector<uint8_t> sb;
    read_entry(*st, sb);
    stats_.assign(sb.begin(), sb.end());
  }
  JsonParser jp{meta_.data(), meta_.data() + meta_.size()};
  const Json m = jp.value();
  if (m.kind != Json::OBJ) throw Error(LIZ_E_FORMAT, "meta.json is not an object");
  // the count and sizes as the reference's reader takes them (liblizard/core/rx/lizfile.cpp:200-212 framesOf): the
  // count from "frames", else the number of "sizes"; a frame's size from "sizes", else "w" and "h"; each a whole
  // number from 1 to 16384 (checked before any conversion)
  const auto whole = [](const Json *j, double lo, double hi, double *out) {
    if (!j || j->kind != Json::NUM || !(j->num >= lo && j->num <= hi) || j->num != std::floor(j->num)) return false;
    *out = j->num;
    return true;
  };
  const Json *fr = m.get("frames"), *W = m.get("w"), *H = m.get("h"), *sizes = m.get("sizes"), *ms = m.get("ms"),
             *lay = m.get("layout");
  const bool have_sizes = sizes && sizes->kind == Json::ARR;
  double nfd = 0;
  if (fr) {
    if (!whole(fr, 0, 65535, &nfd)) throw Error(LIZ_E_FORMAT, "meta.json's frame count is not a count of 0 to 65535");
  } else if (have_sizes) {
    nfd = static_cast<double>(sizes->items.size());
    if (nfd > 65535) throw Error(LIZ_E_FORMAT, "meta.json lists more than 65535 frames");
  } else {
    throw Error(LIZ_E_FORMAT, "meta.json has no frame count and no sizes");
  }
  const size_t nf = static_cast<size_t>(nfd);
  if (lay && lay->kind == Json::STR) layout_ = lay->str;
  for (size_t i = 0; i < nf; i++) {
    Frame fi;
    const Json *sz = have_sizes && i < sizes->items.size() ? &sizes->items[i] : nullptr;
    double fw = 0, fh = 0;
    bool got = false;
    if (sz && sz->kind == Json::ARR && sz->items.size() == 2)
      got = whole(&sz->items[0], 1, 16384, &fw) && whole(&sz->items[1], 1, 16384, &fh);
    else
      got = whole(W, 1, 16384, &fw) && whole(H, 1, 16384, &fh);
    if (!got) throw Error(LIZ_E_FORMAT, "meta.json gives frame " + std::to_string(i) + " no size of 1 to 16384");
    fi.w = static_cast<uint32_t>(fw);
    fi.h = static_cast<uint32_t>(fh);
    fi.ms = ms && ms->kind == Json::ARR && i < ms->items.size() && ms->items[i].kind == Json::NUM ? ms->items[i].num
                                                                                                   : std::nan("");
    const Entry *e = find(frame_name(i));
    if (!e) throw Error(LIZ_E_FORMAT, "no entry for frame " + std::to_string(i));
    fi.entry = static_cast<int>(e - entries_.data());
    frames_.push_back(fi);
  }
}

const ReplayReader::Entry *ReplayReader::find(const std::string &name) const {
  for (const Entry &e : entries_)
    if (e.name == name) return &e;
  return nullptr;
}

void ReplayReader::read_entry(const Entry &e, std::vector<uint8_t> &out) const {
  out.assign(file_.begin() + static_cast<ptrdiff_t>(e.at), file_.begin() + static_cast<ptrdiff_t>(e.at + e.size));
  if (crc32_of(out.data(), out.size()) != e.crc) throw Error(LIZ_E_FORMAT, "a CRC that does not hold: " + e.name);
}

void ReplayReader::frame(size_t i, uint8_t *luma) {
  if (i >= frames_.size()) throw Error(LIZ_E_ARG, "no such frame");
  const Frame &fi = frames_[i];
  std::vector<uint8_t> z;
  read_entry(entries_[fi.entry], z);
  const Zstd &zs = zstd();
  const size_t want = static_cast<size_t>(fi.w) * fi.h;
  const size_t one = zs.findFrameCompressedSize(z.data(), z.size());
  if (zs.isError(one) || one != z.size()) throw Error(LIZ_E_FORMAT, "a frame's entry is not one zstd frame");
  res_.resize(want);
  const size_t got = zs.decompress(res_.data(), want, z.data(), z.size());
  if (zs.isError(got) || got != want) throw Error(LIZ_E_FORMAT, "a frame's entry does not hold its w h bytes");
  loco_restore(res_.data(), fi.w, fi.h, luma);
}

// ---- writing ----

ReplayWriter::ReplayWriter(const std::string &path) : path_(path) {
  zstd();
  f_ = std::fopen((path + ".part").c_str(), "wb");
  if (!f_) throw Error(LIZ_E_IO, "cannot write " + path + ".part");
}

ReplayWriter::~ReplayWriter() {
  if (f_) {
    std::fclose(f_);
    std::remove((path_ + ".part").c_str());
  }
}

void ReplayWriter::entry(const std::string &name, const uint8_t *data, size_t n) {
  if (n > 0xFFFFFFF0u || offset_ + n + 30 + name.size() > 0xFFFFFFF0u) throw Error(LIZ_E_TOOBIG, "past ZIP's 4 GB");
  if (central_count_ >= 65535) throw Error(LIZ_E_TOOBIG, "past ZIP's 65,535 entries");
  const uint32_t crc = crc32_of(data, n);
  std::vector<uint8_t> h;
  put32(h, 0x04034b50u);
  put16(h, 10);      // version needed
  put16(h, 0);       // flags
  put16(h, 0);       // stored
  put16(h, 0);       // 00:00
  put16(h, 0x0021);  // 1980-01-01
  put32(h, crc);
  put32(h, static_cast<uint32_t>(n));
  put32(h, static_cast<uint32_t>(n));
  put16(h, static_cast<uint32_t>(name.size()));
  put16(h, 0);
  h.insert(h.end(), name.begin(), name.end());
  if (std::fwrite(h.data(), 1, h.size(), f_) != h.size() || (n && std::fwrite(data, 1, n, f_) != n))
    throw Error(LIZ_E_IO, "a write the file system refused");
  put32(central_, 0x02014b50u);
  put16(central_, 20);  // made by
  put16(central_, 10);
  put16(central_, 0);
  put16(central_, 0);
  put16(central_, 0);
  put16(central_, 0x0021);
  put32(central_, crc);
  put32(central_, static_cast<uint32_t>(n));
  put32(central_, static_cast<uint32_t>(n));
  put16(central_, static_cast<uint32_t>(name.size()));
  put16(central_, 0);
  put16(central_, 0);
  put16(central_, 0);
  put16(central_, 0);
  put32(central_, 0);
  put32(central_, static_cast<uint32_t>(offset_));
  central_.insert(central_.end(), name.begin(), name.end());
  central_count_++;
  offset_ += h.size() + n;
}

void ReplayWriter::frame(const uint8_t *px, uint32_t w, uint32_t h, uint32_t stride) {
  if (!f_) throw Error(LIZ_E_STATE, "the replay is finished");
  const Zstd &zs = zstd();
  res_.resize(static_cast<size_t>(w) * h);
  loco_residuals(px, w, h, stride, res_.data());
  buf_.resize(zs.compressBound(res_.size()));
  const size_t n = zs.compress(buf_.data(), buf_.size(), res_.data(), res_.size(), 1);
  if (zs.isError(n)) throw Error(LIZ_E_INTERNAL, "zstd refused a frame");
  entry(frame_name(frames_), buf_.data(), n);
  frames_++;
}

int64_t ReplayWriter::finish(const std::string &meta, const std::string &stats) {
  if (!f_) throw Error(LIZ_E_STATE, "the replay is finished");
  if (!stats.empty()) entry("stats.jsonl", reinterpret_cast<const uint8_t *>(stats.data()), stats.size());
  entry("meta.json", reinterpret_cast<const uint8_t *>(meta.data()), meta.size());
  const std::string comment = "LIZARD replay 1";
  std::vector<uint8_t> end;
  put32(end, 0x06054b50u);
  put16(end, 0);
  put16(end, 0);
  put16(end, central_count_);
  put16(end, central_count_);
  put32(end, static_cast<uint32_t>(central_.size()));
  put32(end, static_cast<uint32_t>(offset_));
  put16(end, static_cast<uint32_t>(comment.size()));
  end.insert(end.end(), comment.begin(), comment.end());
  if (std::fwrite(central_.data(), 1, central_.size(), f_) != central_.size() ||
      std::fwrite(end.data(), 1, end.size(), f_) != end.size())
    throw Error(LIZ_E_IO, "a write the file system refused");
  const int64_t total = static_cast<int64_t>(offset_ + central_.size() + end.size());
  std::fclose(f_);
  f_ = nullptr;
  if (std::rename((path_ + ".part").c_str(), path_.c_str()) != 0) throw Error(LIZ_E_IO, "cannot rename to " + path_);
  return total;
}

}  // namespace lz
