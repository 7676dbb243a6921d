// Replays: the .lizard file (research/SPEC.md 15): reading one a frame at a time, and writing one.
#pragma once
#include <stdint.h>

#include <cstdio>
#include <string>
#include <vector>

namespace lz {

// The median edge predictor's residuals of a w x h luma (rows stride apart), and the luma back from them.
void loco_residuals(const uint8_t *px, uint32_t w, uint32_t h, uint32_t stride, uint8_t *r);
void loco_restore(const uint8_t *r, uint32_t w, uint32_t h, uint8_t *px);

class ReplayReader {
 public:
  struct Frame {
    uint32_t w = 0, h = 0;
    double ms = 0;  // NaN where meta.json has none
    int entry = -1;
  };
  explicit ReplayReader(const std::string &path);  // throws Error
  const std::vector<Frame> &frames() const { return frames_; }
  const std::string &meta() const { return meta_; }
  const std::string &stats() const { return stats_; }
  const std::string &layout() const { return layout_; }
  void frame(size_t i, uint8_t *luma);  // w h bytes, rows w apart; throws Error

 private:
  struct Entry {
    std::string name;
    uint64_t at = 0, size = 0;
    uint32_t crc = 0;
  };
  std::vector<uint8_t> file_, res_;
  std::vector<Entry> entries_;
  std::vector<Frame> frames_;
  std::string meta_, stats_, layout_ = "1:1";
  const Entry *find(const std::string &name) const;
  void read_entry(const Entry &e, std::vector<uint8_t> &out) const;
};

class ReplayWriter {
 public:
  explicit ReplayWriter(const std::string &path);  // writes path.part, renamed over path at finish
  ~ReplayWriter();
  void frame(const uint8_t *px, uint32_t w, uint32_t h, uint32_t stride);
  int64_t finish(const std::string &meta, const std::string &stats);  // the file's bytes
  size_t frames() const { return frames_; }

 private:
  std::string path_;
  std::FILE *f_ = nullptr;
  std::vector<uint8_t> central_, res_, buf_;
  uint64_t offset_ = 0;
  uint32_t central_count_ = 0;
  size_t frames_ = 0;
  void entry(const std::string &name, const uint8_t *data, size_t n);
};

}  // namespace lz
