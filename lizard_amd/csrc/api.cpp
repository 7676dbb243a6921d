// The C ABI (include/lizard_amd.h): liblizard's decode API and the lizx_* seams, over the HIP stages.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "format.hpp"
#include "ldpc.hpp"
#include "channel.hpp"
#include "word.hpp"
#include "register.hpp"
#include "border.hpp"
#include "finder.hpp"
#include "xfer.hpp"
#include "shake.hpp"
#include "replay.hpp"
#include "spectrum.hpp"

#define LIZARD_AMD_VERSION "lizard-amd 0.1 (gfx950)"

namespace lz {

// The calling thread's last failure (liblizard/api/api.cpp:16-30): a fixed array, so no thread_local destructor.
static thread_local char last_error[1024];
static thread_local int last_code = LIZ_OK;

int fail(int code, const std::string &message) {
  const size_t n = message.size() < sizeof last_error ? message.size() : sizeof last_error - 1;
  std::memcpy(last_error, message.data(), n);
  last_error[n] = '\0';
  last_code = code;
  return code;
}
void clear_error() {
  last_error[0] = '\0';
  last_code = LIZ_OK;
}

static std::atomic<double> g_hang_ms{-1.0};

double hang_bound_ms() {
  double ms = g_hang_ms.load();
  if (ms < 0) {
    const char *e = std::getenv("LIZ_HANG_MS");
    ms = e ? std::max(0.0, std::atof(e)) : 4000.0;
    g_hang_ms.store(ms);
  }
  return ms;
}

void set_hang_bound_ms(double ms) { g_hang_ms.store(ms); }

void wait_bounded(hipStream_t s, const char *what) {
  const double bound = hang_bound_ms();
  if (bound <= 0) {
    hip_check(hipStreamSynchronize(s), what);
    return;
  }
  const auto t0 = std::chrono::steady_clock::now();
  for (int spins = 0;; spins++) {
    const hipError_t q = hipStreamQuery(s);
    if (q == hipSuccess) return;
    if (q != hipErrorNotReady) hip_check(q, what);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms > bound)
      throw Error(LIZ_E_STATE, std::string(what) + ": not done within the hang bound of " + std::to_string(bound) +
                                   " ms, taken as hung; the handle refuses further work");
    if (spins < 2000) std::this_thread::yield();
    else std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

hipStream_t aux_stream() {
  static std::mutex m;
  static hipStream_t streams[64] = {};
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  need(dev >= 0 && dev < 64, "a device index under 64");
  std::lock_guard<std::mutex> lk(m);
  if (!streams[dev]) hip_check(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking), "hipStreamCreate");
  return streams[dev];
}

int current_device(int want) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) throw Error(LIZ_E_UNSUPPORTED, "no HIP device");
  int dev = want;
  if (dev < 0) hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev >= count) throw Error(LIZ_E_ARG, "no such device");
  hip_check(hipSetDevice(dev), "hipSetDevice");
  return dev;
}

// The most blocks a frame of a picture up to nmax carries (liblizard/api/codec.cpp:108-111 through cpu_dec_top).
static int top_blocks(int nmax) {
  int top = 0;
  for (int v = 4; v <= 64; v++) {
    const Format *f = format_for(v);
    if (f && f->n <= nmax && f->blocks > top) top = f->blocks;
  }
  return top;
}

}  // namespace lz

using namespace lz;

struct lizx_ldpc {
  std::unique_ptr<LdpcEngine> eng;
  DevBuf tasks, rate_dev, llr, bits, its, bytes, crc_ok, post, msg, off;
  std::vector<LdpcTask> host_tasks;
  int per_cu = 0;  // the persistent grid's workgroups a CU (0: as many as fit); a launch shape the tuner picks
};

// A frame's per-block results in one device area (and one pinned host mirror), so they come back in one copy:
// bytes (B x 476), iterations (B), verdicts (B), estimates (B floats), pilot readings (B floats), the pilots' grid
// shift (2 floats).
struct Results {
  int B = 0;
  size_t at_its = 0, at_ok = 0, at_est = 0, at_pilot = 0, at_align = 0, total = 0;
  void layout(int blocks) {
    B = blocks;
    at_its = static_cast<size_t>(B) * BLOCK_BYTES;
    at_ok = at_its + B;
    at_est = (at_ok + B + 3) & ~static_cast<size_t>(3);
    at_pilot = at_est + 4 * static_cast<size_t>(B);
    at_align = at_pilot + 4 * static_cast<size_t>(B);
    total = at_align + 2 * sizeof(float);
  }
};

struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (p) hip_check(hipHostFree(p), "hipHostFree");
    p = nullptr;
    bytes = 0;
    hip_check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc");
    bytes = n;
  }
};

struct liz_decoder {
  int nmax = 0, top = 0;
  std::unique_ptr<LdpcEngine> eng;
  DevBuf tasks, llr, est, grid, res;
  PinnedBuf res_host;
  Results lay;
  int tasks_version = 0;  // the format whose frame-mode task list sits in tasks
  std::vector<LdpcTask> host_tasks;
  std::vector<uint8_t> blocks, ok;
  std::vector<std::unique_ptr<FramePlan>> plans;  // one per (version, span) met
  FrameBufs fb;
  DevBuf img, nodes, wframe;
  PinnedBuf img_host;  // the capture's box staged in pinned memory, then one contiguous copy
  // a cropped capture's upload as one copy: its WarpFrame, nodes and box side by side (CAP_* offsets), staged in
  // cap_host and copied to cap (the WarpFrame at cap's start, a fixed address the frame graph reads)
  DevBuf cap;
  PinnedBuf cap_host;
  WarpFrame wframe_host{};
  std::vector<uint8_t> luma;
  // registration refinement
  DevBuf reg_ref, reg_part, reg_sum, reg_state, reg_trace, reg_lens;
  DevBuf spec_w, spec_X, spec_C, spec_G, spec_dl, spec_part, spec_mask;  // the spectral loss's work areas
  std::vector<double> spec_part_host;
  std::vector<float> reg_nodes, reg_part_host;
  RegState reg_state_host{};
  // the held registration for liz_decode
  int held_version = 0, held_span = 0;
  lizx_registration held_reg{};
  std::vector<float> held_nodes;
  // the sender's pictures and the camera channel
  DevBuf mod_blocks, mod_bits, mod_G, chan_a, chan_b, chan_kinds;
  FinderBufs finder;
  DevBuf b3_a, b3_b, b3_out;
  std::vector<float> found_nodes;
  // the single-frame decode's kernels as one HIP graph, kept for the format and work areas it was captured with
  bool use_graphs = true;
  // the single-frame picture stages' waves a workgroup (the same bits at every shape): one wave a workgroup for the
  // row pass and the demap spreads a lone frame over the most CUs (measured fastest on n = 384 frames)
  LaunchShape shape = [] {
    LaunchShape s;
    s.rows = 1, s.cols = 2, s.demap = 1;
    return s;
  }();
  int anchor_mode = 1;  // the finder registers from the anchors outward (lizx_finder_anchors)
  struct FrameGraph {
    int version = 0, span = 0;
    bool dbg = false;
    const void *keys[9] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } fg;
  void drop_frame_graph() {
    if (fg.exec) (void)hipGraphExecDestroy(fg.exec);
    if (fg.graph) (void)hipGraphDestroy(fg.graph);
    fg = FrameGraph{};
  }
  ~liz_decoder() { drop_frame_graph(); }
  const FramePlan &plan(int version, int span) {
    for (auto &p : plans)
      if (p->version == version && p->span == span) return *p;
    plans.push_back(make_plan(version, span));
    return *plans.back();
  }
};

// The decline gate, LDPC and CRC of one frame's blocks from soft values and estimates already on the device
// (lizard_phy.c:1775-1870): the tables and the frame's task list made ready (host copies), and the launch's
// parameters.
static LdpcParams finish_params(liz_decoder *d, const Format *f, const int8_t *d_llr, const double *d_est) {
  LdpcEngine &e = *d->eng;
  e.ensure_frame_tables(static_cast<int>(f->white.size()));
  const int B = f->blocks;
  d->lay.layout(B);
  d->res.reserve(d->lay.total);
  if (d->tasks_version != f->version) {
    d->host_tasks.resize(B);
    for (int b = 0; b < B; b++) d->host_tasks[b] = LdpcTask{f->blk_slot[b], -1, f->blk_code[b], f->blk_slot[b], 0, b};
    std::stable_sort(d->host_tasks.begin(), d->host_tasks.end(),
                     [](const LdpcTask &a, const LdpcTask &b) { return a.code > b.code; });
    d->tasks.reserve(B * sizeof(LdpcTask));
    copy_wait(d->tasks.p, d->host_tasks.data(), B * sizeof(LdpcTask), hipMemcpyHostToDevice, e.stream, "tasks");
    d->tasks_version = f->version;
  }
  uint8_t *r = d->res.as<uint8_t>();
  LdpcParams p{};
  p.tasks = d->tasks.as<LdpcTask>();
  p.count = B;
  p.llr = d_llr;
  p.its = reinterpret_cast<int8_t *>(r + d->lay.at_its);
  p.blk_est = reinterpret_cast<float *>(r + d->lay.at_est);
  p.bytes = r;
  p.crc_ok = r + d->lay.at_ok;
  p.white = e.white.as<uint8_t>();
  for (int ci = 0; ci < NCODES; ci++) p.s2p[ci] = e.s2p[ci].as<int16_t>();
  p.est = d_est;
  p.crc_tab = e.crc_tab.as<uint32_t>();
  p.crc0 = e.crc0;
  p.max_iter = 30;  // lizard_phy.c:1824 with LIZARD_STALL_IT, LIZARD_STALL_RATIO (:81-82)
  p.stall_it = 9;
  p.stall_ratio = 0.95f;
  p.mode = MODE_FRAME;
  p.flags = 1;
  return p;
}

// The same, enqueued on the decoder's stream.
static void finish_blocks(liz_decoder *d, const Format *f, const int8_t *d_llr, const double *d_est) {
  d->eng->launch(finish_params(d, f, d_llr, d_est), false, true);
}

// The frame's results area's pilot readings (where the picture stages write them).
static float *pilot_area(liz_decoder *d, const Format *f) {
  d->lay.layout(f->blocks);
  d->res.reserve(d->lay.total);
  return reinterpret_cast<float *>(d->res.as<uint8_t>() + d->lay.at_pilot);
}

// The frame's results area's grid shift (where k_align writes it, in place of the work area's).
static float *align_area(liz_decoder *d, const Format *f) {
  d->lay.layout(f->blocks);
  d->res.reserve(d->lay.total);
  return reinterpret_cast<float *>(d->res.as<uint8_t>() + d->lay.at_align);
}

// The per-block results of finish_blocks to the host in one copy (after the stream's work); returns the verified
// count. pilot (optional) receives the pilot readings, align (optional) the grid shift.
static int collect_blocks(liz_decoder *d, const Format *f, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est,
                          float *pilot = nullptr, float *align = nullptr) {
  LdpcEngine &e = *d->eng;
  const int B = f->blocks;
  d->res_host.reserve(d->lay.total);
  hip_check(hipMemcpyAsync(d->res_host.p, d->res.p, d->lay.total, hipMemcpyDeviceToHost, e.stream), "results");
  e.sync();
  const uint8_t *h = static_cast<const uint8_t *>(d->res_host.p);
  std::memcpy(its, h + d->lay.at_its, B);
  std::memcpy(ok, h + d->lay.at_ok, B);
  std::memcpy(blk_est, h + d->lay.at_est, B * sizeof(float));
  if (pilot) std::memcpy(pilot, h + d->lay.at_pilot, B * sizeof(float));
  if (align) std::memcpy(align, h + d->lay.at_align, 2 * sizeof(float));
  int good = 0;
  for (int b = 0; b < B; b++) {
    if (its[b] == -2) ok[b] = 0;
    if (ok[b]) {
      std::memcpy(blocks + static_cast<size_t>(b) * BLOCK_BYTES, h + static_cast<size_t>(b) * BLOCK_BYTES, BLOCK_BYTES);
      good++;
    }
  }
  return good;
}

// Builds the task list for count codewords in codeword mode.
static void codeword_tasks(int count, const uint8_t *rate, const int64_t *off, std::vector<LdpcTask> &t) {
  t.resize(count);
  for (int i = 0; i < count; i++) {
    const int ci = code_index(rate[i]);
    if (ci < 0) throw Error(LIZ_E_ARG, "a rate index with no code (6, 4, 3 or 2)");
    need(off[i] >= 0, "a negative offset");
    t[i] = LdpcTask{off[i], off[i], ci, 0, 0, i};
  }
  // the longest codewords (1/2) first, the shortest last: the persistent queue's tail stays short
  std::stable_sort(t.begin(), t.end(), [](const LdpcTask &a, const LdpcTask &b) { return a.code > b.code; });
}

// The systematic QC-IRA encoder (liblizard/src/ldpc.c:117): data bits as they are, then parity bit m (check i of block row
// r, m = i * mb + r) as the running XOR of every check's data sum in that order (the staircase's accumulator).
static void encode_word(const Code &c, const uint8_t *d, uint8_t *o) {
  std::memcpy(o, d, c.k);
  uint8_t p = 0;
  for (int chk = 0; chk < c.m; chk++) {
    const int i = chk / c.mb, r = chk % c.mb;
    for (int e = c.lay_ptr[r]; e < c.lay_ptr[r + 1]; e++) p ^= d[c.lay_col[e] * Z + ((i - c.lay_shift[e]) & (Z - 1))];
    o[c.k + chk] = p;
  }
}

static void check_params(int max_iter, int stall_it, float ratio) {
  need(max_iter >= 1 && max_iter <= 127, "max_iter is 1 to 127");
  need(stall_it >= 0 && stall_it <= 127, "stall_it is 0 to 127");
  need(ratio == ratio, "stall_ratio is a number");
}

extern "C" {

static int decode_capture(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                          const lizx_registration *reg, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est,
                          lizx_soft *soft);


LIZ_API const char *liz_version(void) { return LIZARD_AMD_VERSION; }
LIZ_API const char *liz_last_error(void) { return last_error; }
LIZ_API int liz_last_error_code(void) { return last_code; }

LIZ_API int lizx_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

LIZ_API void *lizx_dev_alloc(size_t bytes) {
  void *p = nullptr;
  guard([&] {
    hip_check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
    return 0;
  });
  return p;
}
LIZ_API void lizx_dev_free(void *p) {
  if (p) (void)hipFree(p);
}
LIZ_API int lizx_dev_upload(void *dst, const void *src, size_t bytes) {
  return guard([&] {
    copy_wait(dst, src, bytes, hipMemcpyHostToDevice, aux_stream(), "upload");
    return 0;
  });
}
LIZ_API int lizx_dev_download(void *dst, const void *src, size_t bytes) {
  return guard([&] {
    copy_wait(dst, src, bytes, hipMemcpyDeviceToHost, aux_stream(), "download");
    return 0;
  });
}

LIZ_API int lizx_format_info(int version, int *info) {
  return guard([&] {
    need(info, "a place for the info");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    info[0] = f->n;
    info[1] = f->subch;
    info[2] = f->blocks;
    info[3] = f->tiers;
    info[4] = f->vblocks;
    info[5] = f->used;
    return 0;
  });
}

// The sender's bit side (lizard_phy.c:1087-1150): every block's CRC-32C, LDPC codeword, bit map and pilot tail, the
// padding past the last block, whitened; one byte (0/1) a frame slot.
static void frame_slots_of(const Format *f, const uint8_t *blocks, int cnt, uint8_t *slots) {
  std::memcpy(slots, f->white.data(), f->white.size());  // the padding: the whitening alone
  std::vector<uint8_t> data(4096), cw(MAX_N);
  for (int b = 0; b < f->blocks; b++) {
    const Code &c = code(f->blk_code[b]);
    const uint8_t *src = blocks + static_cast<size_t>(b) * BLOCK_BYTES;
    std::fill(data.begin(), data.end(), 0);
    for (int i = 0; i < 8 * BLOCK_BYTES; i++) data[i] = (src[i >> 3] >> (7 - (i & 7))) & 1;
    const uint32_t crc = crc32c(src, BLOCK_BYTES);
    for (int i = 0; i < 32; i++) data[8 * BLOCK_BYTES + i] = (crc >> (31 - i)) & 1;
    encode_word(c, data.data(), cw.data());
    // slot i: codeword bit at the bit map's position, or past nt the pilot bit (bit 0 of the count on even blocks,
    // bit 1 on odd ones), whitened (SPEC 7.3)
    const auto &s2p = f->slot_to_post[f->blk_code[b]];
    const uint8_t q = static_cast<uint8_t>((cnt >> (b & 1)) & 1);
    const int s0 = f->blk_slot[b], nb = f->blk_nslots[b];
    for (int i = 0; i < nb; i++) {
      uint8_t bit = q;
      if (i < c.nt) {
        // the map holds the posterior index of the codeword bit; back to the codeword index
        const int pi = s2p[i];
        const int v = pi < c.k ? pi : c.k + ((pi - c.k) % Z) * c.mb + (pi - c.k) / Z;
        bit = cw[v];
      }
      slots[s0 + i] = static_cast<uint8_t>(bit ^ f->white[s0 + i]);
    }
  }
}

LIZ_API int lizx_frame_slots(int version, const uint8_t *blocks, int cnt, uint8_t *slots) {
  return guard([&] {
    need(blocks && slots && cnt >= 0 && cnt <= 3, "blocks, slots and a count 0 to 3");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    frame_slots_of(f, blocks, cnt, slots);
    return 0;
  });
}

LIZ_API int lizx_format_blocks(int version, int *rate, int *slot0, int *nslots) {
  return guard([&] {
    need(rate && slot0 && nslots, "outputs");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    for (int b = 0; b < f->blocks; b++) {
      rate[b] = code(f->blk_code[b]).rate;
      slot0[b] = f->blk_slot[b];
      nslots[b] = f->blk_nslots[b];
    }
    return f->blocks;
  });
}

LIZ_API int lizx_format_coefs(int version, int16_t *u, int16_t *v) {
  return guard([&] {
    need(u && v, "outputs");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    const int n = f->n, bw = f->bw;
    for (size_t i = 0; i < f->pos.size(); i++) {
      const int pp = f->pos[i], vv = (pp / (n * bw)) * bw + pp % bw, uw = (pp / bw) % n;
      u[i] = static_cast<int16_t>(uw < n / 2 ? uw : uw - n);
      v[i] = static_cast<int16_t>(vv);
    }
    return 0;
  });
}

LIZ_API int lizx_ldpc_dims(int rate, int *dims) {
  return guard([&] {
    need(dims, "a place for the dims");
    const int ci = code_index(rate);
    if (ci < 0) throw Error(LIZ_E_ARG, "a rate index with no code (6, 4, 3 or 2)");
    const Code &c = code(ci);
    const int d[12] = {c.n, c.k, c.m, Z, c.mb, c.kb, c.np, c.nt, c.edges, c.slots_total, c.slots_total * Z,
                       c.n * 2};
    std::memcpy(dims, d, sizeof d);
    return 0;
  });
}

// p_c = p_(c-1) ^ (xor of the data bits in check c), p_(-1) = 0 (research/SPEC.md 8.4): check mb*i + r holds data bit
// 128 j + ((i - sh[r][j]) mod 128) for each circulant (j, sh) of block row r.
LIZ_API int lizx_ldpc_encode(int rate, int count, const uint8_t *data, uint8_t *cw) {
  return guard([&] {
    const int ci = code_index(rate);
    if (ci < 0) throw Error(LIZ_E_ARG, "a rate index with no code (6, 4, 3 or 2)");
    need(count >= 0 && (count == 0 || (data && cw)), "data and codeword buffers");
    const Code &c = code(ci);
    for (int w = 0; w < count; w++) {
      const uint8_t *d = data + static_cast<size_t>(w) * c.k;
      uint8_t *o = cw + static_cast<size_t>(w) * c.n;
      encode_word(c, d, o);
    }
    return 0;
  });
}

LIZ_API uint32_t lizx_crc32c(const uint8_t *p, size_t n) { return p ? crc32c(p, static_cast<int>(n)) : 0u; }

LIZ_API int lizx_copy_probe(void *d_dst, const void *d_src, size_t bytes, float *ms) {
  return guard([&] {
    need(d_dst && d_src && ms && bytes % 16 == 0, "device buffers and a size in 16-byte units");
    *ms = copy_probe(d_dst, d_src, bytes);
    return 0;
  });
}

LIZ_API lizx_ldpc *lizx_ldpc_new(int device) {
  return guard_new<lizx_ldpc>([&] {
    const int dev = current_device(device);
    auto h = std::make_unique<lizx_ldpc>();
    h->eng = std::make_unique<LdpcEngine>(dev);
    return h.release();
  });
}
LIZ_API void lizx_ldpc_free(lizx_ldpc *h) { delete h; }

LIZ_API int lizx_ldpc_decode_device(lizx_ldpc *h, int count, const uint8_t *d_rate, const int8_t *d_llr,
                                    const int64_t *d_off, uint8_t *d_bits, int8_t *d_its, int max_iter, int stall_it,
                                    float stall_ratio, int flags, uint8_t *d_bytes, uint8_t *d_crc_ok) {
  return guard([&] {
    need(h && count >= 0 && d_rate && d_llr && d_off && d_its, "a decoder and device arrays");
    check_params(max_iter, stall_it, stall_ratio);
    need(!(flags & LIZX_LDPC_CRC) || (d_bytes && d_crc_ok), "CRC outputs");
    need(!(flags & LIZX_LDPC_TRACE), "traces are a host-call option");
    if (!count) return 0;
    // the task list is made on the device from the rates and offsets (which also zeroes the queue counter)
    hip_check(hipSetDevice(h->eng->device), "hipSetDevice");
    h->tasks.reserve(count * sizeof(LdpcTask));
    h->eng->make_tasks(h->tasks.as<LdpcTask>(), d_rate, d_off, count);
    LdpcParams p{};
    p.tasks = h->tasks.as<LdpcTask>();
    p.count = count;
    p.llr = d_llr;
    p.bits = d_bits;
    p.its = d_its;
    p.bytes = d_bytes;
    p.crc_ok = d_crc_ok;
    p.crc_tab = h->eng->crc_tab.as<uint32_t>();
    p.crc0 = h->eng->crc0;
    p.max_iter = max_iter;
    p.stall_it = stall_it;
    p.stall_ratio = stall_ratio;
    p.mode = MODE_CODEWORD;
    p.flags = flags & LIZX_LDPC_CRC;
    h->eng->launch(p, (flags & LIZX_LDPC_PLAIN) != 0, true, true, h->per_cu);
    return 0;
  });
}

// lizx_ldpc_decode_device with its task list and queue counter in the library's fixed per-device work area: nothing
// allocated, freed or grown. The kernels are the same; only where their scratch lives differs.
LIZ_API int lizx_ldpc_decode_device_noalloc(lizx_ldpc *h, int count, const uint8_t *d_rate, const int8_t *d_llr,
                                            const int64_t *d_off, uint8_t *d_bits, int8_t *d_its, int max_iter,
                                            int stall_it, float stall_ratio, int flags, uint8_t *d_bytes,
                                            uint8_t *d_crc_ok) {
  return guard([&] {
    need(h && count >= 0 && d_rate && d_llr && d_off && d_its, "a decoder and device arrays");
    if (count > SCRATCH_TASKS)
      throw Error(LIZ_E_ARG, "more codewords than the fixed work area holds (16384 a call)");
    check_params(max_iter, stall_it, stall_ratio);
    need(!(flags & LIZX_LDPC_CRC) || (d_bytes && d_crc_ok), "CRC outputs");
    need(!(flags & LIZX_LDPC_TRACE), "traces are a host-call option");
    if (!count) return 0;
    hip_check(hipSetDevice(h->eng->device), "hipSetDevice");
    const LdpcScratch sc = ldpc_scratch(h->eng->device);  // looked up when the handle was made: no call here
    h->eng->make_tasks(sc.tasks, d_rate, d_off, count, sc.queue);
    LdpcParams p{};
    p.tasks = sc.tasks;
    p.count = count;
    p.queue = sc.queue;
    p.llr = d_llr;
    p.bits = d_bits;
    p.its = d_its;
    p.bytes = d_bytes;
    p.crc_ok = d_crc_ok;
    p.crc_tab = h->eng->crc_tab.as<uint32_t>();
    p.crc0 = h->eng->crc0;
    p.max_iter = max_iter;
    p.stall_it = stall_it;
    p.stall_ratio = stall_ratio;
    p.mode = MODE_CODEWORD;
    p.flags = flags & LIZX_LDPC_CRC;
    h->eng->launch(p, (flags & LIZX_LDPC_PLAIN) != 0, true, true, h->per_cu);
    return 0;
  });
}

LIZ_API int lizx_ldpc_tune(lizx_ldpc *h, int count, const uint8_t *d_rate, const int8_t *d_llr, const int64_t *d_off,
                           int8_t *d_its, int max_iter, int stall_it, float stall_ratio, int reps, int *per_cu) {
  return guard([&] {
    need(h && count > 0 && reps > 0 && reps <= 100, "a decoder, codewords and 1 to 100 repeats");
    LdpcEngine &e = *h->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    int best = 0;
    float best_ms = 0;
    for (int c : {0, 2, 3, 4, 6, 8, 10, 12, 16}) {
      if (c > e.wg_per_cu) continue;
      h->per_cu = c;
      for (int r = 0; r <= reps; r++) {
        if (r == 1) e.sync();  // the first run settles; the next reps are timed
        if (lizx_ldpc_decode_device(h, count, d_rate, d_llr, d_off, nullptr, d_its, max_iter, stall_it, stall_ratio, 0,
                                    nullptr, nullptr))
          throw Error(liz_last_error_code(), liz_last_error());
      }
      e.sync();
      if (c == 0 || e.last_ms < best_ms) {
        best_ms = e.last_ms;
        best = c;
      }
    }
    h->per_cu = best;
    if (per_cu) *per_cu = best;
    return 0;
  });
}

LIZ_API int lizx_ldpc_shape(lizx_ldpc *h, int per_cu) {
  return guard([&] {
    need(h && per_cu >= 0, "a decoder and a count of workgroups a CU");
    h->per_cu = per_cu >= h->eng->wg_per_cu ? 0 : per_cu;
    return h->per_cu ? h->per_cu : h->eng->wg_per_cu;
  });
}

LIZ_API int lizx_ldpc_sync(lizx_ldpc *h) {
  return guard([&] {
    need(h, "a decoder");
    h->eng->sync();
    return 0;
  });
}

LIZ_API float lizx_ldpc_last_ms(lizx_ldpc *h) { return h ? h->eng->last_ms : -1.0f; }

LIZ_API int lizx_ldpc_decode(lizx_ldpc *h, int count, const uint8_t *rate, const int8_t *llr, const int64_t *off,
                             uint8_t *bits, int8_t *its, int max_iter, int stall_it, float stall_ratio, int flags,
                             uint8_t *bytes, uint8_t *crc_ok, int16_t *post, int8_t *msg) {
  return guard([&] {
    need(h && count >= 0 && rate && llr && off && its, "a decoder, rates, LLRs, offsets and iteration outputs");
    check_params(max_iter, stall_it, stall_ratio);
    const bool crc = flags & LIZX_LDPC_CRC, trace = flags & LIZX_LDPC_TRACE;
    need(!crc || (bytes && crc_ok), "CRC outputs");
    need(!trace || (post && msg), "trace outputs");
    if (!count) return 0;
    codeword_tasks(count, rate, off, h->host_tasks);
    // per codeword (input order): its length; the device buffers cover every codeword's range
    std::vector<int> len(count);
    int64_t span = 0;
    for (int i = 0; i < count; i++) {
      len[i] = code(code_index(rate[i])).n;
      span = std::max<int64_t>(span, off[i] + len[i]);
    }
    LdpcEngine &e = *h->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    h->tasks.reserve(count * sizeof(LdpcTask));
    h->llr.reserve(span);
    h->its.reserve(count);
    if (bits) h->bits.reserve(span);
    if (crc) {
      h->bytes.reserve(static_cast<size_t>(count) * BLOCK_BYTES);
      h->crc_ok.reserve(count);
    }
    if (trace) {
      h->post.reserve(span * sizeof(int16_t));
      h->msg.reserve(static_cast<size_t>(count) * MSG_TRACE_STRIDE);
    }
    hip_check(hipMemcpyAsync(h->tasks.p, h->host_tasks.data(), count * sizeof(LdpcTask), hipMemcpyHostToDevice,
                             e.stream),
              "tasks");
    hip_check(hipMemcpyAsync(h->llr.p, llr, span, hipMemcpyHostToDevice, e.stream), "LLRs");
    LdpcParams p{};
    p.tasks = h->tasks.as<LdpcTask>();
    p.count = count;
    p.llr = h->llr.as<int8_t>();
    p.bits = bits ? h->bits.as<uint8_t>() : nullptr;
    p.its = h->its.as<int8_t>();
    p.bytes = crc ? h->bytes.as<uint8_t>() : nullptr;
    p.crc_ok = crc ? h->crc_ok.as<uint8_t>() : nullptr;
    p.post = trace ? h->post.as<int16_t>() : nullptr;
    p.msg = trace ? h->msg.as<int8_t>() : nullptr;
    p.crc_tab = e.crc_tab.as<uint32_t>();
    p.crc0 = e.crc0;
    p.max_iter = max_iter;
    p.stall_it = stall_it;
    p.stall_ratio = stall_ratio;
    p.mode = MODE_CODEWORD;
    p.flags = (crc ? 1 : 0) | (trace ? 2 : 0);
    e.launch(p, (flags & LIZX_LDPC_PLAIN) != 0, true);
    e.sync();
    copy_wait(its, h->its.p, count, hipMemcpyDeviceToHost, h->eng->stream, "iterations");
    if (bits) {
      // only the codewords' own ranges: the caller's buffer between them is left as it was
      std::vector<uint8_t> all(span);
      copy_wait(all.data(), h->bits.p, span, hipMemcpyDeviceToHost, h->eng->stream, "bits");
      for (int i = 0; i < count; i++) std::memcpy(bits + off[i], all.data() + off[i], len[i]);
    }
    if (crc) {
      copy_wait(bytes, h->bytes.p, static_cast<size_t>(count) * BLOCK_BYTES, hipMemcpyDeviceToHost, h->eng->stream, "bytes");
      copy_wait(crc_ok, h->crc_ok.p, count, hipMemcpyDeviceToHost, h->eng->stream, "CRC verdicts");
    }
    if (trace) {
      std::vector<int16_t> all(span);
      copy_wait(all.data(), h->post.p, span * sizeof(int16_t), hipMemcpyDeviceToHost, h->eng->stream, "posteriors");
      for (int i = 0; i < count; i++)
        std::memcpy(post + off[i], all.data() + off[i], len[i] * sizeof(int16_t));
      copy_wait(msg, h->msg.p, static_cast<size_t>(count) * MSG_TRACE_STRIDE, hipMemcpyDeviceToHost, h->eng->stream,
                "messages");
    }
    int cleared = 0;
    for (int i = 0; i < count; i++) cleared += its[i] > 0;
    return cleared;
  });
}

LIZ_API liz_decoder *liz_decoder_new(int nmax) {
  return guard_new<liz_decoder>([&] {
    if (!nmax) nmax = 1536;
    need(nmax >= 384 && nmax <= 1536, "nmax is 384 to 1536 (or 0)");
    const int dev = current_device(-1);
    auto d = std::make_unique<liz_decoder>();
    d->nmax = nmax;
    d->top = top_blocks(nmax);
    d->eng = std::make_unique<LdpcEngine>(dev);
    return d.release();
  });
}

LIZ_API void liz_decoder_free(liz_decoder *d) { delete d; }

LIZ_API int lizx_decoder_graphs(liz_decoder *d, int on) {
  return guard([&] {
    need(d, "a decoder");
    d->use_graphs = on != 0;
    if (!on) d->drop_frame_graph();
    return 0;
  });
}

LIZ_API int lizx_decoder_shape(liz_decoder *d, int rows, int cols, int demap) {
  return guard([&] {
    need(d, "a decoder");
    const auto ok = [](int w) { return w == 1 || w == 2 || w == 4; };
    need(ok(rows) && ok(cols) && ok(demap), "waves a workgroup: 1, 2 or 4");
    d->shape.rows = rows;
    d->shape.cols = cols;
    d->shape.demap = demap;
    d->drop_frame_graph();
    return 0;
  });
}

LIZ_API int lizx_decoder_otf(liz_decoder *d, float sigma) {
  return guard([&] {
    need(d, "a decoder");
    hip_check(hipSetDevice(d->eng->device), "hipSetDevice");
    d->fb.set_otf(&sigma, 1, d->eng->stream);
    return 0;
  });
}

static int lane_anchor_find(FinderLane &ln, int span, const uint8_t *d_img, int w, int h, const float *H,
                            const LensModel *lm, AnchorHit *A);
static void border_refine(FinderLane &ln, int span, const uint8_t *d_img, int w, int h, const float *H, float mpx,
                          const LensModel *lm, std::vector<NodeState> &st);

LIZ_API int lizx_hang_bound(double ms) {
  return guard([&] {
    need(ms >= 0 && ms <= 3.6e6, "a hang bound of 0 (none) to 3600000 ms");
    set_hang_bound_ms(ms);
    return 0;
  });
}

LIZ_API int lizx_finder_anchors(liz_decoder *d, int on) {
  return guard([&] {
    need(d, "a decoder");
    need(on >= 0 && on <= 2, "0 (off), 1 (anchors over the finder's border) or 2 (anchors and the border refined)");
    d->anchor_mode = on;
    return 0;
  });
}

// A capture's upload into the finder's device image (rows w bytes apart).
static const uint8_t *finder_upload(liz_decoder *d, const uint8_t *img, int w, int h, int stride) {
  need(img && w >= 16 && h >= 16 && w <= 16384 && h <= 16384 && stride >= w,
       "a capture of 16 to 16384 pixels a side, rows stride >= w bytes apart");
  LdpcEngine &e = *d->eng;
  hip_check(hipSetDevice(e.device), "hipSetDevice");
  d->finder.img.reserve(static_cast<size_t>(w) * h);
  hip_check(hipMemcpy2DAsync(d->finder.img.p, w, img, stride, w, h, hipMemcpyHostToDevice, e.stream), "capture");
  d->finder.main.s = e.stream;
  d->finder.main.anch.reserve(ANCH_AREA);
  d->finder.main.h_anch.reserve(ANCH_AREA);
  return d->finder.img.as<uint8_t>();
}

static void model_of(const double *model, LensModel &m) {
  m.K = static_cast<int>(model[0]);
  need(m.K >= 0 && m.K <= LENS_TERMS, "a lens model of 0 to 3 radial terms");
  m.ax = model[1];
  m.ay = model[2];
  m.R = model[3];
  for (int a = 0; a < LENS_P; a++) m.t[a] = model[4 + a];
}

static void model_to(const LensModel &m, double *model) {
  model[0] = m.K;
  model[1] = m.ax;
  model[2] = m.ay;
  model[3] = m.R;
  for (int a = 0; a < LENS_P; a++) model[4 + a] = m.t[a];
}

static void records_to(const AnchorHit *A, float *out) {
  for (int k = 0; k < ANCH_N; k++) {
    float *o = out + 8 * k;
    o[0] = A[k].X, o[1] = A[k].Y, o[2] = A[k].dx, o[3] = A[k].dy;
    o[4] = A[k].peak, o[5] = A[k].u, o[6] = A[k].v, o[7] = static_cast<float>(A[k].ok);
  }
}

LIZ_API int lizx_anchor_find(liz_decoder *d, int span, const uint8_t *img, int w, int h, int stride, const float *H,
                             const double *model, float *anchors) {
  return guard([&] {
    need(d && H && anchors, "a decoder, a homography and room for the records");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    const uint8_t *d_img = finder_upload(d, img, w, h, stride);
    LensModel m;
    if (model) model_of(model, m);
    AnchorHit A[ANCH_N] = {};
    const int found = lane_anchor_find(d->finder.main, span, d_img, w, h, H, model ? &m : nullptr, A);
    records_to(A, anchors);
    return found;
  });
}

LIZ_API int lizx_anchor_register(liz_decoder *d, int span, const uint8_t *img, int w, int h, int stride, float *H,
                                 float mpx, float *quad, float *anchors, double *model, int *moved) {
  return guard([&] {
    need(d && H && quad && anchors && model, "a decoder, a homography, the corners and room for the results");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    need(mpx > 0, "a positive module size in pixels");
    const uint8_t *d_img = finder_upload(d, img, w, h, stride);
    AnchorHit A[ANCH_N] = {};
    LensModel m;
    m.K = 0;
    const AnchorFind find = [&](const float *Hs, const LensModel *lm, AnchorHit *an) {
      return lane_anchor_find(d->finder.main, span, d_img, w, h, Hs, lm, an);
    };
    int mv = 0;
    const int have = anchor_register(span, H, w, h, mpx, quad, find, A, &m, &mv);
    records_to(A, anchors);
    model_to(m, model);
    if (moved) *moved = mv;
    return have;
  });
}

LIZ_API int lizx_border_refine(liz_decoder *d, int span, const uint8_t *img, int w, int h, int stride, const float *H,
                               float mpx, const double *model, float *nodes) {
  return guard([&] {
    need(d && H && nodes, "a decoder, a homography and room for the nodes");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    need(mpx > 0, "a positive module size in pixels");
    const uint8_t *d_img = finder_upload(d, img, w, h, stride);
    LensModel m;
    if (model) model_of(model, m);
    std::vector<NodeState> st;
    border_refine(d->finder.main, span, d_img, w, h, H, mpx, model ? &m : nullptr, st);
    for (size_t k = 0; k < st.size(); k++) {
      nodes[4 * k] = st[k].dx;
      nodes[4 * k + 1] = st[k].dy;
      nodes[4 * k + 2] = st[k].score;
      nodes[4 * k + 3] = static_cast<float>(st[k].done);
    }
    float pos[MAX_NODES];
    return node_lattice(span, pos);
  });
}

LIZ_API int liz_decoder_max_blocks(const liz_decoder *d) {
  return guard([&] {
    need(d, "a decoder");
    return d->top;
  });
}

static bool find_registration(liz_decoder *d, const uint8_t *img, int w, int h, int stride, FindResult &fr,
                              lizx_registration &reg, std::vector<float> &nodes);

LIZ_API int liz_decode(liz_decoder *d, const uint8_t *px, int w, int h, int stride, liz_pixfmt fmt, int *held,
                       uint8_t *verified, liz_decoded *out) {
  return guard([&]() -> int {
    need(d && px && held && verified, "a decoder, an image, the held word and a place for the blocks");
    need(w > 0 && h > 0, "an image of at least a pixel");
    need(fmt == LIZ_GREY8 || fmt == LIZ_RGBX8 || fmt == LIZ_RGBA8 || fmt == LIZ_BGRA8, "a pixel format of lizard.h's");
    const int bpp = fmt == LIZ_GREY8 ? 1 : 4;
    need(stride >= w * bpp, "a stride of at least the image's width in bytes");
    if (out) *out = liz_decoded{};
    // tight luma as the reference's codec reads it (api/codec.cpp:128-140): BGRA/RGBA/RGBX weights 77/150/29
    const uint8_t *img = px;
    int istride = stride;
    if (fmt != LIZ_GREY8) {
      d->luma.resize(static_cast<size_t>(w) * h);
      for (int y = 0; y < h; y++) {
        const uint8_t *row = px + static_cast<size_t>(y) * stride;
        uint8_t *o = d->luma.data() + static_cast<size_t>(y) * w;
        for (int x = 0; x < w; x++) {
          const int r = fmt == LIZ_BGRA8 ? row[4 * x + 2] : row[4 * x], g = row[4 * x + 1],
                    b = fmt == LIZ_BGRA8 ? row[4 * x] : row[4 * x + 2];
          o[x] = static_cast<uint8_t>((77 * r + 150 * g + 29 * b + 128) >> 8);
        }
      }
      img = d->luma.data();
      istride = w;
    }
    // the registration: the held one, else the finder's on this frame (lizard_acquire's place in the reference)
    liz_decoded o{};
    o.ring = -1;
    int span = d->held_span;
    lizx_registration found{};
    const lizx_registration *reg = &d->held_reg;
    if (!span) {
      FindResult fr;
      if (w < 16 || h < 16 || !find_registration(d, img, w, h, istride, fr, found, d->found_nodes)) {
        if (out) *out = o;
        return 0;
      }
      span = fr.span;
      reg = &found;
    }
    o.found = 1;
    o.ring = span == 128 ? 0 : span == 256 ? 1 : 2;
    // the corner nodes in the image (acquire.c:656-660 corner_coords, 5026-5033): H at the lattice's first and last
    // node positions plus those nodes' displacements; top left, top right, bottom right, bottom left
    {
      float pos[MAX_NODES];
      const int nx = node_lattice(span, pos);
      const float *H = reg->H;
      for (int c = 0; c < 4; c++) {
        const int i = (c == 1 || c == 2) ? nx - 1 : 0, j = c >= 2 ? nx - 1 : 0;
        const float mx = pos[i], my = pos[j], ww = H[6] * mx + H[7] * my + H[8];
        o.quad[2 * c] = (H[0] * mx + H[1] * my + H[2]) / ww + reg->nodes[2 * (j * nx + i)];
        o.quad[2 * c + 1] = (H[3] * mx + H[4] * my + H[5]) / ww + reg->nodes[2 * (j * nx + i) + 1];
      }
    }
    // the frame's own word, read through the held registration (ob_thin_read_fmt acquire.c:4503, ob_fmt_decode
    // fmt.h:75); the version as any.c:48-51 takes it: the word's, else the caller's held one, else (an extension) the
    // version the hold states for frames that carry no word; with none of them the frame is not decoded
    const int side = span + 30, cells = fmt_side_cells(side);
    std::vector<float> q(4 * cells);
    FmtWord fw{};
    o.word = read_word_soft(span, img, w, h, istride, reg->H, reg->nodes, q.data(), cells) &&
             fmt_decode(q.data(), cells, fmt_bytes(side), &fw) == 0;
    int version = 0;
    if (o.word) {
      version = fw.version;
      o.fps = fw.fps;
      *held = version;  // api/codec.cpp:143: updated whenever a frame's own word reads
    } else if (*held >= FMT_VERSION_MIN && *held <= FMT_VERSION_MAX) {
      version = *held;
      o.held_used = 1;
    } else if (reg == &d->held_reg) {
      version = d->held_version;
    }
    const Format *f = version ? format_for(version) : nullptr;
    if (!f || f->n > d->nmax) {  // not finished: no version, or a picture beyond this decoder's (any.c:52-53)
      if (out) *out = o;
      return 0;
    }
    const int B = f->blocks;
    d->ok.resize(B);
    std::vector<int8_t> its(B);
    std::vector<float> est(B);
    std::vector<uint8_t> all(static_cast<size_t>(B) * BLOCK_BYTES);
    std::vector<float> pil(B);
    lizx_soft sp{};
    sp.blk_pilot = pil.data();
    decode_capture(d, version, span, img, w, h, istride, reg, all.data(), d->ok.data(), its.data(), est.data(), &sp);
    int k = 0;
    const int total = std::min(B, d->top);
    for (int b = 0; b < total; b++)
      if (d->ok[b]) std::memcpy(verified + static_cast<size_t>(k++) * BLOCK_BYTES, all.data() + static_cast<size_t>(b) * BLOCK_BYTES, BLOCK_BYTES);
    if (out) {
      o.n = f->n;
      o.blocks = version;
      o.total = B;
      o.verified = k;
      // the pilots over the even blocks and the odd (bits 0 and 1 of the picture's count): each parity's mean reading
      // and its standard error, NaN where no block of it read (lizard_phy.c:451-472 lizard_pilot over every block)
      o.pilot_blocks = 0;
      for (int g = 0; g < 2; g++) {
        double s1 = 0, s2 = 0;
        int c = 0;
        for (int b = g; b < B; b += 2)
          if (pil[b] == pil[b]) {
            s1 += pil[b];
            s2 += static_cast<double>(pil[b]) * pil[b];
            c++;
          }
        if (!c) {
          o.pilot_r[g] = o.pilot_sd[g] = __builtin_nanf("");
          continue;
        }
        const double mean = s1 / c, var = c > 1 ? (s2 - s1 * mean) / (c - 1) : 0;
        o.pilot_r[g] = static_cast<float>(mean);
        o.pilot_sd[g] = static_cast<float>(std::sqrt((var > 0 ? var : 0) / c));
        o.pilot_blocks += c;
      }
      *out = o;
    }
    return k;
  });
}

// The pixels of a w x h capture the warp through this registration can read, as [x0, x1) x [y0, y1): the picture's
// sample square widened by a module each way, through H (its corners bound a projective image while the denominator
// stays positive at all four), widened by the largest node displacement and by the bilinear read's next pixel and two
// more, clamped to the frame. The whole frame where the denominator does not stay positive.
static void read_box(const FramePlan &p, const lizx_registration *reg, int w, int h, int box[4]) {
  box[0] = 0, box[1] = 0, box[2] = w, box[3] = h;
  const double lo = 14.0, hi = 16.0 + p.span;  // module coordinates: the sampled square is 15 .. 15 + span
  const float *H = reg->H;
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
  for (int c = 0; c < 4; c++) {
    const double mx = (c & 1) ? hi : lo, my = (c & 2) ? hi : lo;
    const double ww = static_cast<double>(H[6]) * mx + static_cast<double>(H[7]) * my + H[8];
    if (!(ww > 1e-6)) return;
    const double x = (static_cast<double>(H[0]) * mx + static_cast<double>(H[1]) * my + H[2]) / ww;
    const double y = (static_cast<double>(H[3]) * mx + static_cast<double>(H[4]) * my + H[5]) / ww;
    if (!(std::fabs(x) < 1e7 && std::fabs(y) < 1e7)) return;
    xmin = std::min(xmin, x), xmax = std::max(xmax, x), ymin = std::min(ymin, y), ymax = std::max(ymax, y);
  }
  double dmx = 0, dmy = 0;
  const int nn = p.dev.nx * p.dev.ny;
  for (int i = 0; i < nn; i++) {
    const double ax = std::fabs(static_cast<double>(reg->nodes[2 * i])), ay = std::fabs(static_cast<double>(reg->nodes[2 * i + 1]));
    if (!(ax < 1e6 && ay < 1e6)) return;  // NaN or wild nodes: the whole frame
    dmx = std::max(dmx, ax), dmy = std::max(dmy, ay);
  }
  const auto clampi = [](double v, int a, int b) { return v < a ? a : v > b ? b : static_cast<int>(v); };
  box[0] = clampi(std::floor(xmin - dmx - 0.5) - 2, 0, w);
  box[1] = clampi(std::floor(ymin - dmy - 0.5) - 2, 0, h);
  box[2] = clampi(std::floor(xmax + dmx - 0.5) + 4, 0, w);
  box[3] = clampi(std::floor(ymax + dmy - 0.5) + 4, 0, h);
  if (box[2] - box[0] < 2 || box[3] - box[1] < 2) box[0] = 0, box[1] = 0, box[2] = w, box[3] = h;
}

// The cropped upload's layout: the WarpFrame at 0, the nodes at CAP_NODES, the capture's box at CAP_IMG.
constexpr size_t CAP_NODES = 256, CAP_IMG = CAP_NODES + ((MAX_NODES * MAX_NODES * sizeof(float2) + 255) & ~size_t{255});
static_assert(sizeof(WarpFrame) <= CAP_NODES, "the WarpFrame fits before the nodes");

// A capture and its registration onto the device as a WarpFrame (one frame), on the decoder's stream: only the
// pixels the warp can reach through this registration when crop (read_box), else the whole frame (the refinement
// moves the registration on the device). Every entry that stages a capture synchronises the stream before it
// returns, so the caller's buffers are read in time.
static const WarpFrame *stage_capture(liz_decoder *d, const FramePlan &p, const uint8_t *img, int w, int h, int stride,
                                      const lizx_registration *reg, bool crop = false) {
  need(img && reg && reg->nodes, "a capture and a registration with nodes");
  need(w >= 2 && h >= 2 && stride >= w, "a capture of at least 2 x 2 pixels, rows stride >= w bytes apart");
  need(warp_frame_fits(h, w), "a capture under 2^31 pixels with rows under 2^24 pixels");
  LdpcEngine &e = *d->eng;
  const size_t nn = static_cast<size_t>(p.dev.nx) * p.dev.ny;
  int box[4] = {0, 0, w, h};
  if (crop) read_box(p, reg, w, h, box);
  const int bw = box[2] - box[0], bh = box[3] - box[1];
  if (crop) {
    // the WarpFrame, the nodes and the box's rows gathered into pinned memory by the host, then one copy
    need(p.dev.nx <= MAX_NODES && p.dev.ny <= MAX_NODES, "a node lattice of at most 40 x 40");
    const size_t room = CAP_IMG + static_cast<size_t>(w) * h;  // the whole frame's, so a larger box never reallocates
    d->cap.reserve(room);
    d->cap_host.reserve(room);
    uint8_t *dev = d->cap.as<uint8_t>(), *stage = static_cast<uint8_t *>(d->cap_host.p);
    WarpFrame wf{};
    for (int k = 0; k < 9; k++) wf.H[k] = reg->H[k];
    wf.nodes = reinterpret_cast<const float2 *>(dev + CAP_NODES);
    // pixel (x, y) at img + y bw + x: the stored box's first pixel is (box[0], box[1])
    wf.img = dev + CAP_IMG - (static_cast<ptrdiff_t>(box[1]) * bw + box[0]);
    wf.stride = bw;
    wf.w = w;
    wf.h = h;
    std::memcpy(stage, &wf, sizeof wf);
    std::memcpy(stage + CAP_NODES, reg->nodes, nn * sizeof(float2));
    const uint8_t *src = img + static_cast<size_t>(box[1]) * stride + box[0];
    for (int y = 0; y < bh; y++)
      std::memcpy(stage + CAP_IMG + static_cast<size_t>(y) * bw, src + static_cast<size_t>(y) * stride, bw);
    hip_check(hipMemcpyAsync(dev, stage, CAP_IMG + static_cast<size_t>(bw) * bh, hipMemcpyHostToDevice, e.stream),
              "capture");
    return d->cap.as<WarpFrame>();
  }
  d->img.reserve(static_cast<size_t>(w) * h);  // the whole frame's room, so a later larger box never reallocates
  const uint8_t *src = img + static_cast<size_t>(box[1]) * stride + box[0];
  const size_t bytes = static_cast<size_t>(bw) * bh;
  if (bw == stride) {
    hip_check(hipMemcpyAsync(d->img.p, src, bytes, hipMemcpyHostToDevice, e.stream), "capture");
  } else {
    // the box's rows gathered into pinned memory by the host (a strided copy from pageable memory is slow), then one
    // copy; the pinned area is not reused before the stream has read it (every entry syncs before it returns)
    d->img_host.reserve(static_cast<size_t>(w) * h);
    uint8_t *stage = static_cast<uint8_t *>(d->img_host.p);
    for (int y = 0; y < bh; y++) std::memcpy(stage + static_cast<size_t>(y) * bw, src + static_cast<size_t>(y) * stride, bw);
    hip_check(hipMemcpyAsync(d->img.p, stage, bytes, hipMemcpyHostToDevice, e.stream), "capture");
  }
  d->nodes.reserve(nn * sizeof(float2));
  hip_check(hipMemcpyAsync(d->nodes.p, reg->nodes, nn * sizeof(float2), hipMemcpyHostToDevice, e.stream), "nodes");
  WarpFrame wf{};
  for (int k = 0; k < 9; k++) wf.H[k] = reg->H[k];
  wf.nodes = d->nodes.as<float2>();
  // pixel (x, y) at img + y bw + x: the stored box's first pixel is (box[0], box[1])
  wf.img = d->img.as<uint8_t>() - (static_cast<ptrdiff_t>(box[1]) * bw + box[0]);
  wf.stride = bw;
  wf.w = w;
  wf.h = h;
  d->wframe.reserve(sizeof(WarpFrame));
  d->wframe_host = wf;
  hip_check(hipMemcpyAsync(d->wframe.p, &d->wframe_host, sizeof wf, hipMemcpyHostToDevice, e.stream), "registration");
  return d->wframe.as<WarpFrame>();
}

// The soft values, estimates, rms and coefficients of the last frame to the caller's lizx_soft (the pilot shift comes
// back with the results).
static void copy_soft(liz_decoder *d, const FramePlan &p, lizx_soft *soft, bool dbg) {
  const size_t S = static_cast<size_t>(p.subch) * SUB;
  if (soft->llr) copy_wait(soft->llr, d->fb.llr.p, 2 * S, hipMemcpyDeviceToHost, d->eng->stream, "soft values");
  if (soft->sub_est)
    copy_wait(soft->sub_est, d->fb.est.p, p.subch * sizeof(double), hipMemcpyDeviceToHost, d->eng->stream, "estimates");
  if (soft->sub_rms)
    copy_wait(soft->sub_rms, d->fb.rms.p, p.subch * sizeof(float), hipMemcpyDeviceToHost, d->eng->stream, "rms");
  if (soft->coef && dbg)
    copy_wait(soft->coef, d->fb.coef_dbg.p, S * 2 * sizeof(float), hipMemcpyDeviceToHost, d->eng->stream, "coefficients");
  if (soft->soft && dbg)
    copy_wait(soft->soft, d->fb.coef_dbg.as<float>() + 2 * S, S * 2 * sizeof(float), hipMemcpyDeviceToHost, d->eng->stream,
              "soft floats");
}

static int decode_capture(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                          const lizx_registration *reg, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est,
                          lizx_soft *soft) {
  need(blocks && ok && its && blk_est, "outputs");
  const Format *f = format_for(version);
  if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
  if (f->n > d->nmax) throw Error(LIZ_E_FORMAT, "the format's picture is larger than the decoder's nmax");
  LdpcEngine &e = *d->eng;
  hip_check(hipSetDevice(e.device), "hipSetDevice");
  const FramePlan &p = d->plan(version, span);
  const WarpFrame *wf = stage_capture(d, p, img, w, h, stride, reg, true);
  const bool dbg = soft && (soft->coef || soft->soft);
  d->fb.reserve(p, 1, dbg);
  SoftBufs v = d->fb.view(dbg);
  v.blk_pilot = pilot_area(d, f);
  v.align = align_area(d, f);
  v.zero_int = e.queue.as<int>();  // the LDPC queue counter, reset by the first picture kernel (no launch of its own)
  const LdpcParams lp = finish_params(d, f, v.llr, v.est);
  if (!d->use_graphs) {
    launch_soft_warp(p, v, wf, 1, e.stream, d->shape);
    e.launch(lp, false, true, true);
  } else {
    // the frame's kernels (picture stages, LDPC) as one graph: captured once for this format and these work areas
    // (the capture and registration sit at fixed device addresses the kernels read), then replayed; the results'
    // copy stays outside it (with that copy captured as a graph node the process crashed on the first launch under
    // rocprofv3's kernel trace)
    auto &g = d->fg;
    const void *keys[9] = {v.X, v.C, v.llr, v.coef_dbg, d->res.p, d->tasks.p, wf, e.white.p, v.otf};
    bool same = g.exec && g.version == version && g.span == span && g.dbg == dbg;
    for (int k = 0; same && k < 9; k++) same = g.keys[k] == keys[k];
    if (!same) {
      d->drop_frame_graph();
      hip_check(hipStreamBeginCapture(e.stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
      try {
        launch_soft_warp(p, v, wf, 1, e.stream, d->shape);
        e.launch(lp, false, false, true);
      } catch (...) {
        hipGraph_t gg = nullptr;
        (void)hipStreamEndCapture(e.stream, &gg);
        if (gg) (void)hipGraphDestroy(gg);
        throw;
      }
      hip_check(hipStreamEndCapture(e.stream, &g.graph), "hipStreamEndCapture");
      hip_check(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      g.version = version;
      g.span = span;
      g.dbg = dbg;
      for (int k = 0; k < 9; k++) g.keys[k] = keys[k];
    }
    hip_check(hipGraphLaunch(g.exec, e.stream), "hipGraphLaunch");
  }
  const int good = collect_blocks(d, f, blocks, ok, its, blk_est, soft ? soft->blk_pilot : nullptr,
                                  soft ? soft->align : nullptr);
  if (soft) copy_soft(d, p, soft, dbg);
  return good;
}

// The device bit side writes the padding after the last block: the format's blocks must tile its slots from 0.
static void need_tiled(const Format *f) {
  for (int b = 0; b < f->blocks; b++)
    if (f->blk_slot[b] != (b ? f->blk_slot[b - 1] + f->blk_nslots[b - 1] : 0))
      throw Error(LIZ_E_INTERNAL, "the format's blocks do not tile its slots");
}

LIZ_API int lizx_frame_slots_device(liz_decoder *d, int version, int frames, const uint8_t *d_blocks, int cnt,
                                    uint8_t *d_slots) {
  return guard([&] {
    need(d && d_blocks && d_slots && frames > 0 && frames <= 4096 && cnt >= 0 && cnt <= 3,
         "a decoder, 1 to 4096 frames of blocks in device memory, a pilot count 0 to 3 and device memory for the slots");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    need_tiled(f);
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, f->n <= 384 ? 128 : f->n <= 768 ? 256 : 512);
    e.frame_slots(p.dev.blk, f->blocks, static_cast<int>(f->white.size()), d_blocks, cnt, d_slots, frames);
    e.sync();
    return 0;
  });
}

LIZ_API int lizx_modulate(liz_decoder *d, int version, int frames, const uint8_t *blocks, int cnt, float *d_pictures,
                          float *kernel_ms) {
  return guard([&] {
    need(d && blocks && d_pictures && frames > 0 && frames <= 4096 && cnt >= 0 && cnt <= 3,
         "a decoder, 1 to 4096 frames of blocks, a pilot count 0 to 3 and device memory for the pictures");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    need_tiled(f);
    // any ring's plan of the format serves: the picture depends on n and the coefficient order only
    const FramePlan &p = d->plan(version, f->n <= 384 ? 128 : f->n <= 768 ? 256 : 512);
    const size_t S2 = f->white.size(), fb = static_cast<size_t>(f->blocks) * BLOCK_BYTES;
    d->mod_blocks.reserve(fb * frames);
    d->mod_bits.reserve(S2 * frames);
    d->mod_G.reserve(static_cast<size_t>(frames) * p.E * p.n * sizeof(float2));
    hip_check(hipMemcpyAsync(d->mod_blocks.p, blocks, fb * frames, hipMemcpyHostToDevice, e.stream), "blocks");
    hip_check(hipEventRecord(e.ev0, e.stream), "hipEventRecord");
    e.frame_slots(p.dev.blk, f->blocks, static_cast<int>(S2), d->mod_blocks.as<uint8_t>(), cnt,
                  d->mod_bits.as<uint8_t>(), frames);
    launch_modulate(p, d->mod_bits.as<uint8_t>(), d->mod_G.as<float2>(), d_pictures, frames, e.stream);
    hip_check(hipEventRecord(e.ev1, e.stream), "hipEventRecord");
    e.sync();
    if (kernel_ms) hip_check(hipEventElapsedTime(kernel_ms, e.ev0, e.ev1), "hipEventElapsedTime");
    return 0;
  });
}

// The channel from a device picture through a camera; with a word's fps (>= 0) the whole symbol is painted round it
// (border, word, anchors, guard interval, white quiet zone), else the picture alone in grey 128.
static int capture_impl(liz_decoder *d, int version, int span, int fps, const float *d_picture, const lizx_camera *cam,
                        uint8_t *d_img) {
  {
    need(d && d_picture && cam && d_img && cam->w > 0 && cam->h > 0 && cam->w <= 16384 && cam->h <= 16384,
         "a decoder, a device picture, a camera of 1 to 16384 pixels a side and device memory for the capture");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    need(cam->blur >= 0 && cam->blur <= CHAN_MAX_R / 3.0f && cam->noise >= 0, "blur 0 to 5.33 px, noise >= 0");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    ChannelParams c{};
    c.W = cam->w;
    c.Hh = cam->h;
    c.n = f->n;
    c.span = span;
    const double *H = cam->H;
    const double det = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                       H[2] * (H[3] * H[7] - H[4] * H[6]);
    need(std::fabs(det) > 0, "an invertible homography");
    c.Hi[0] = (H[4] * H[8] - H[5] * H[7]) / det;
    c.Hi[1] = (H[2] * H[7] - H[1] * H[8]) / det;
    c.Hi[2] = (H[1] * H[5] - H[2] * H[4]) / det;
    c.Hi[3] = (H[5] * H[6] - H[3] * H[8]) / det;
    c.Hi[4] = (H[0] * H[8] - H[2] * H[6]) / det;
    c.Hi[5] = (H[2] * H[3] - H[0] * H[5]) / det;
    c.Hi[6] = (H[3] * H[7] - H[4] * H[6]) / det;
    c.Hi[7] = (H[1] * H[6] - H[0] * H[7]) / det;
    c.Hi[8] = (H[0] * H[4] - H[1] * H[3]) / det;
    c.k1 = cam->k1;
    c.k2 = cam->k2;
    c.ax = cam->w / 2.0;
    c.ay = cam->h / 2.0;
    c.D = std::hypot(c.ax, c.ay);
    c.r = cam->blur > 0 ? std::max(1, static_cast<int>(std::ceil(3.0 * cam->blur))) : 0;
    if (c.r > 0) {
      float sum = 0.0f;
      for (int k = -c.r; k <= c.r; k++) {
        const float t = static_cast<float>(k) / cam->blur;
        c.taps[k + c.r] = std::exp(-0.5f * t * t);
        sum += c.taps[k + c.r];
      }
      for (int k = 0; k <= 2 * c.r; k++) c.taps[k] /= sum;
    }
    c.noise = cam->noise;
    c.seed = cam->seed;
    if (fps >= 0) {
      const int L = span + 30;
      std::vector<uint8_t> kinds(static_cast<size_t>(L) * L);
      uint8_t cw[FMT_BYTES_MAX];
      FmtWord fw;
      fw.version = version;
      fw.fps = fps;
      fmt_encode(fw, cw, fmt_bytes(L));
      border_kinds(span, cw, kinds.data());
      d->chan_kinds.reserve(kinds.size());
      hip_check(hipMemcpyAsync(d->chan_kinds.p, kinds.data(), kinds.size(), hipMemcpyHostToDevice, e.stream),
                "border map");
      c.kinds = d->chan_kinds.as<uint8_t>();
      c.anch_a = span / 64;
      for (int i = 0; i < 7; i++) c.anch_m0[i] = anchor_m0(span, i);
    }
    const size_t px = static_cast<size_t>(c.W) * c.Hh;
    d->chan_a.reserve(px * sizeof(float));
    d->chan_b.reserve(px * sizeof(float));
    launch_channel(c, d_picture, d->chan_a.as<float>(), d->chan_b.as<float>(), d_img, e.stream);
    e.sync();
    return 0;
  }
}

LIZ_API int lizx_capture(liz_decoder *d, int version, int span, const float *d_picture, const lizx_camera *cam,
                         uint8_t *d_img) {
  return guard([&] { return capture_impl(d, version, span, -1, d_picture, cam, d_img); });
}

LIZ_API int lizx_capture_symbol(liz_decoder *d, int version, int span, int fps, const float *d_picture,
                                const lizx_camera *cam, uint8_t *d_img) {
  return guard([&] {
    need(fps >= 0 && fps <= 255, "fps 0..255");
    return capture_impl(d, version, span, fps, d_picture, cam, d_img);
  });
}

LIZ_API int lizx_b3_chunk_cvs(liz_decoder *d, const uint8_t *d_file, uint64_t length, int chunk_log2, uint8_t *d_cvs,
                              float *kernel_ms) {
  return guard([&] {
    need(d && d_file && d_cvs, "a decoder, a device file and device room for the chaining values");
    need(chunk_log2 >= 10 && chunk_log2 <= 24 && length >= 1, "chunk_log2 10..24 and a file of a byte or more");
    need((reinterpret_cast<uintptr_t>(d_file) & 15) == 0 && (reinterpret_cast<uintptr_t>(d_cvs) & 15) == 0,
         "16-byte aligned device buffers");
    const uint64_t chunks = (length + (1ull << chunk_log2) - 1) >> chunk_log2;
    need(chunks <= 16383, "at most 16,383 chunks");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const uint64_t words = b3_scratch_words(length);
    d->b3_a.reserve(words * 4);
    d->b3_b.reserve(words * 4);
    hip_check(hipEventRecord(e.ev0, e.stream), "hipEventRecord");
    b3_chunk_cvs(d_file, length, chunk_log2, d->b3_a.as<uint32_t>(), d->b3_b.as<uint32_t>(),
                 reinterpret_cast<uint32_t *>(d_cvs), e.stream);
    hip_check(hipEventRecord(e.ev1, e.stream), "hipEventRecord");
    e.sync();
    if (kernel_ms) hip_check(hipEventElapsedTime(kernel_ms, e.ev0, e.ev1), "hipEventElapsedTime");
    return static_cast<int>(chunks);
  });
}

LIZ_API int lizx_b3_hash_device(liz_decoder *d, const uint8_t *d_data, uint64_t length, uint8_t *out,
                                float *kernel_ms) {
  return guard([&] {
    need(d && d_data && out, "a decoder, device data and room for the hash");
    need((reinterpret_cast<uintptr_t>(d_data) & 15) == 0, "16-byte aligned device data");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const uint64_t words = b3_scratch_words(length);
    d->b3_a.reserve(words * 4);
    d->b3_b.reserve(words * 4);
    d->b3_out.reserve(32);
    hip_check(hipEventRecord(e.ev0, e.stream), "hipEventRecord");
    b3_hash_device(d_data, length, d->b3_a.as<uint32_t>(), d->b3_b.as<uint32_t>(), d->b3_out.as<uint32_t>(),
                   e.stream);
    hip_check(hipEventRecord(e.ev1, e.stream), "hipEventRecord");
    e.sync();
    if (kernel_ms) hip_check(hipEventElapsedTime(kernel_ms, e.ev0, e.ev1), "hipEventElapsedTime");
    copy_wait(out, d->b3_out.p, 32, hipMemcpyDeviceToHost, e.stream, "hash");
    return 0;
  });
}

LIZ_API int lizx_b3_root(const uint8_t *cvs, uint32_t chunks, uint8_t *root) {
  return guard([&] {
    need(cvs && root && chunks >= 2 && chunks <= 16383, "2 to 16,383 chaining values and room for the root");
    std::vector<uint32_t> w(static_cast<size_t>(chunks) * 8);
    std::memcpy(w.data(), cvs, w.size() * 4);  // little-endian words
    uint32_t r[8];
    b3_root(w.data(), chunks, r);
    std::memcpy(root, r, 32);
    return 0;
  });
}

LIZ_API int lizx_b3_hash(const uint8_t *data, uint64_t length, uint8_t *out) {
  return guard([&] {
    need((data || !length) && out, "data and room for the hash");
    uint32_t r[8];
    b3_hash_host(data, length, r);
    std::memcpy(out, r, 32);
    return 0;
  });
}

// ---- the format's arithmetic and the per-frame encoder (lizard.h sections 1 and 2) ----

namespace {
constexpr int RING_CELLS[3] = {64, 128, 256};  // liblizard/src/lizard_phy.h LIZARD_RING
constexpr int RING_DEFAULT = 1;

int ring_index(int ring) {
  if (ring == LIZ_RING_DEFAULT) return RING_DEFAULT;
  need(ring >= 0 && ring < LIZ_RINGS, "ring is 0 to 2 or LIZ_RING_DEFAULT");
  return ring;
}

int pxm_for(int n, int span) { return static_cast<int>(std::ceil(static_cast<float>(n) / span - 1e-4f)); }

liz_geometry geometry_of(const liz_format &fm) {
  need(fm.blocks >= LIZ_MIN_BLOCKS && fm.blocks <= LIZ_MAX_BLOCKS, "blocks is 4 to 64 (LIZARD-64 to LIZARD-1024)");
  need(fm.fps >= 1 && fm.fps <= 255, "fps is 1 to 255");
  need(fm.codes == 1 || fm.codes == 2, "codes is 1 or 2");
  const int r = ring_index(fm.ring);
  const Format *f = format_for(fm.blocks);
  if (!f) throw Error(LIZ_E_FORMAT, "no format of that size");
  liz_geometry g{};
  g.n = f->n;
  g.span = 2 * RING_CELLS[r];
  g.pxm = pxm_for(g.n, g.span);
  g.side = (g.span + 30 + 4) * g.pxm;
  g.gap = fm.codes > 1 ? LIZ_GAP_MODULES * g.pxm : 0;
  g.width = fm.codes * g.side + (fm.codes - 1) * g.gap;
  g.height = g.side;
  g.frame_blocks = fm.codes * f->blocks;
  return g;
}

// The room a symbol of this size needs on a display for its outer coefficients to keep 2.7 pixels a cycle: its width
// over the picture's times the top ring's radius in cycles.
double room_for(int blocks, int r) {
  const double radius = std::sqrt(2.0 * 320 * 16 * blocks / 3.14159265358979323846);
  return static_cast<double>(2 * RING_CELLS[r] + 30 + 4) / (2 * RING_CELLS[r]) * 2.7 * radius;
}

double lanczos3(double x) {
  x = std::fabs(x);
  if (x >= 3) return 0;
  const double px = 3.14159265358979323846 * x;
  return 3 * std::sin(px) * std::sin(px / 3) / (px * px);
}

// The resample's taps (lizard_phy.c:975-1000): output pixel x of the picture and its 3-module guard at K pixels a
// module reads picture samples i0 .. i0 + 5 (taken as periodic), the Lanczos-3 weights normalised to sum to one,
// rounded to float; a pixel that falls on a sample takes it alone.
void lanczos_taps(int n, int span, int K, std::vector<int32_t> &i0, std::vector<float> &w) {
  const int cp = 3, q = (span + 2 * cp) * K;
  const double sc = static_cast<double>(n) / span;
  i0.resize(q);
  w.assign(static_cast<size_t>(6) * q, 0.0f);
  for (int x = 0; x < q; x++) {
    const double u = (x + 0.5) * sc / K - cp * sc - 0.5, r = std::floor(u + 0.5);
    if (std::fabs(u - r) < 1e-6) {
      i0[x] = static_cast<int32_t>(r) - 2;
      w[6 * x + 2] = 1.0f;
      continue;
    }
    const int first = static_cast<int>(std::floor(u)) - 2;
    double tw[6], sum = 0;
    for (int t = 0; t < 6; t++) {
      tw[t] = lanczos3(u - (first + t));
      sum += tw[t];
    }
    i0[x] = first;
    for (int t = 0; t < 6; t++) w[6 * x + t] = static_cast<float>(tw[t] / sum);
  }
}
}  // namespace

struct liz_encoder {
  liz_format fmt{};
  liz_geometry g{};
  liz_decoder *d = nullptr;
  DevBuf pic, page, kinds, i0, w;
  std::vector<uint8_t> host;
  PageParams pp{};
  ~liz_encoder() {
    if (d) liz_decoder_free(d);
  }
};

LIZ_API int liz_ring_cells(int ring) {
  return guard([&] { return RING_CELLS[ring_index(ring)]; });
}

LIZ_API int liz_geometry_of(const liz_format *format, liz_geometry *out) {
  return guard([&] {
    need(format && out, "a format and a place for its geometry");
    *out = geometry_of(*format);
    return 0;
  });
}

LIZ_API double liz_room_for(int blocks, int ring) {
  double v = 0;
  const int rc = guard([&] {
    need(blocks >= LIZ_MIN_BLOCKS && blocks <= LIZ_MAX_BLOCKS, "blocks is 4 to 64");
    v = room_for(blocks, ring_index(ring));
    return 0;
  });
  return rc < 0 ? rc : v;
}

LIZ_API int liz_blocks_for(int blocks) {
  if (blocks < LIZ_MIN_BLOCKS || blocks > LIZ_MAX_BLOCKS) return 0;
  const Format *f = format_for(blocks);
  return f ? f->blocks : 0;
}

LIZ_API int liz_pick(double w, double h, int codes, int ring, int top_blocks) {
  return guard([&] {
    need(codes == 1 || codes == 2, "codes is 1 or 2");
    need(top_blocks >= LIZ_MIN_BLOCKS && top_blocks <= LIZ_MAX_BLOCKS, "top_blocks is 4 to 64");
    const int r = ring_index(ring);
    const double modules = 2 * RING_CELLS[r] + 30 + 4;
    const double gap = codes > 1 ? (codes - 1) * static_cast<double>(LIZ_GAP_MODULES) / modules : 0;
    const double room = std::max(64.0, std::min(w / (codes + gap), h));
    int best = LIZ_MIN_BLOCKS;
    for (int b = LIZ_MIN_BLOCKS; b <= top_blocks; b++)
      if (room_for(b, r) <= room) best = b;
    return best;
  });
}

LIZ_API int liz_simd(void) { return 1; }  // the codec's vector path, here the GPU's, is always the one that runs

LIZ_API void liz_stream_fill(uint32_t id, uint8_t *payload) {
  if (payload) stream_fill(id, payload, LIZ_PAYLOAD);
}

LIZ_API int liz_layout_rects(int w, int h, int layout, int rect[2][4]) {
  return guard([&] {
    need(rect && w > 0 && h > 0, "a frame and a place for its regions");
    need(layout == 1 || layout == 2, "layout is 1 (the centre square) or 2 (2:1)");
    // the centre square, or two squares side by side along the long side (core/rx/crops.h cropsOf)
    const bool two = layout == 2, wide = w >= h;
    const int lo = std::min(w, h), hi = std::max(w, h), count = two ? 2 : 1;
    const int side = two ? std::min(lo, hi / 2) : lo;
    for (int k = 0; k < count; k++) {
      const int along = (hi - count * side) / 2 + k * side, across = (lo - side) / 2;
      rect[k][0] = wide ? along : across;
      rect[k][1] = wide ? across : along;
      rect[k][2] = rect[k][3] = side;
    }
    return count;
  });
}

LIZ_API liz_encoder *liz_encoder_new(const liz_format *format) {
  return guard_new<liz_encoder>([&] {
    need(format, "a format");
    auto e = std::make_unique<liz_encoder>();
    e->fmt = *format;
    e->g = geometry_of(*format);
    e->d = liz_decoder_new(0);
    if (!e->d) throw Error(LIZ_E_UNSUPPORTED, liz_last_error());
    const liz_geometry &g = e->g;
    const int L = g.span + 30;
    std::vector<uint8_t> kinds(static_cast<size_t>(L) * L);
    uint8_t cw[FMT_BYTES_MAX];
    FmtWord fw;
    fw.version = format->blocks;
    fw.fps = format->fps;
    fmt_encode(fw, cw, fmt_bytes(L));
    border_kinds(g.span, cw, kinds.data());
    std::vector<int32_t> i0;
    std::vector<float> w;
    lanczos_taps(g.n, g.span, g.pxm, i0, w);
    hip_check(hipSetDevice(e->d->eng->device), "hipSetDevice");
    e->kinds.reserve(kinds.size());
    e->i0.reserve(i0.size() * sizeof(int32_t));
    e->w.reserve(w.size() * sizeof(float));
    e->pic.reserve(static_cast<size_t>(g.n) * g.n * sizeof(float));
    e->page.reserve(static_cast<size_t>(g.side) * g.side);
    copy_wait(e->kinds.p, kinds.data(), kinds.size(), hipMemcpyHostToDevice, e->d->eng->stream, "border map");
    copy_wait(e->i0.p, i0.data(), i0.size() * sizeof(int32_t), hipMemcpyHostToDevice, e->d->eng->stream, "taps");
    copy_wait(e->w.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, e->d->eng->stream, "weights");
    PageParams &pp = e->pp;
    pp.n = g.n;
    pp.span = g.span;
    pp.pxm = g.pxm;
    pp.q = static_cast<int>(i0.size());
    pp.kinds = e->kinds.as<uint8_t>();
    pp.i0 = e->i0.as<int32_t>();
    pp.w = e->w.as<float>();
    pp.anch_a = g.span / 64;
    for (int i = 0; i < 7; i++) pp.anch_m0[i] = anchor_m0(g.span, i);
    e->host.resize(static_cast<size_t>(g.side) * g.side);
    return e.release();
  });
}

LIZ_API void liz_encoder_free(liz_encoder *e) { delete e; }

LIZ_API int liz_encoder_geometry(const liz_encoder *e, liz_geometry *out) {
  return guard([&] {
    need(e && out, "an encoder and a place for its geometry");
    *out = e->g;
    return 0;
  });
}

LIZ_API int liz_encoder_paint(liz_encoder *e, const uint8_t *blocks, uint32_t picture, uint8_t *out, int stride,
                              liz_pixfmt fmt) {
  return guard([&] {
    need(e && blocks && out, "an encoder, blocks and a frame to paint");
    need(fmt == LIZ_GREY8 || fmt == LIZ_RGBX8 || fmt == LIZ_RGBA8 || fmt == LIZ_BGRA8, "a pixel format of lizard.h's");
    const int bpp = fmt == LIZ_GREY8 ? 1 : 4;
    const liz_geometry &g = e->g;
    need(stride >= g.width * bpp, "a stride of at least the frame's width in bytes");
    LdpcEngine &eng = *e->d->eng;
    const size_t per = static_cast<size_t>(g.frame_blocks / e->fmt.codes) * BLOCK_BYTES;
    for (int k = 0; k < e->fmt.codes; k++) {
      // the picture on the GPU (bit side, QPSK, inverse 2D DFT, clip), then the page round it
      const int rc = lizx_modulate(e->d, e->fmt.blocks, 1, blocks + k * per, static_cast<int>(picture & 3),
                                   e->pic.as<float>(), nullptr);
      if (rc < 0) throw Error(rc, liz_last_error());
      hip_check(hipSetDevice(eng.device), "hipSetDevice");
      launch_paint_page(e->pp, e->pic.as<float>(), e->page.as<uint8_t>(), eng.stream);
      eng.sync();
      copy_wait(e->host.data(), e->page.p, e->host.size(), hipMemcpyDeviceToHost, eng.stream, "page");
      // each symbol into its place, the gap after it its row's last pixel (the margin's colour)
      const int x0 = k * (g.side + g.gap), fill = k + 1 < e->fmt.codes ? g.gap : 0;
      for (int y = 0; y < g.side; y++) {
        const uint8_t *src = e->host.data() + static_cast<size_t>(y) * g.side;
        uint8_t *o = out + static_cast<size_t>(y) * stride + static_cast<size_t>(x0) * bpp;
        if (bpp == 1) {
          std::memcpy(o, src, g.side);
          std::memset(o + g.side, src[g.side - 1], fill);
        } else {
          for (int x = 0; x < g.side + fill; x++) {
            const uint8_t v = src[x < g.side ? x : g.side - 1];
            o[4 * x] = o[4 * x + 1] = o[4 * x + 2] = v;
            o[4 * x + 3] = 255;
          }
        }
      }
    }
    return 0;
  });
}

// ---- replays ----

struct lizx_replay {
  std::unique_ptr<ReplayReader> r;
};
struct lizx_replay_writer {
  std::unique_ptr<ReplayWriter> w;
};

LIZ_API lizx_replay *lizx_replay_open(const char *path) {
  return guard_new<lizx_replay>([&] {
    need(path, "a path");
    auto r = std::make_unique<lizx_replay>();
    r->r = std::make_unique<ReplayReader>(path);
    return r.release();
  });
}

LIZ_API void lizx_replay_close(lizx_replay *r) { delete r; }

LIZ_API int lizx_replay_frames(const lizx_replay *r) {
  return guard([&] {
    need(r, "a replay");
    return static_cast<int>(r->r->frames().size());
  });
}

LIZ_API int lizx_replay_frame_info(const lizx_replay *r, int i, int *w, int *h, double *ms) {
  return guard([&] {
    need(r && i >= 0 && static_cast<size_t>(i) < r->r->frames().size(), "a replay and one of its frames");
    const ReplayReader::Frame &f = r->r->frames()[i];
    if (w) *w = static_cast<int>(f.w);
    if (h) *h = static_cast<int>(f.h);
    if (ms) *ms = f.ms;
    return 0;
  });
}

LIZ_API int lizx_replay_frame(lizx_replay *r, int i, uint8_t *luma) {
  return guard([&] {
    need(r && luma && i >= 0, "a replay, a frame index and room for its luma");
    r->r->frame(static_cast<size_t>(i), luma);
    return 0;
  });
}

LIZ_API const char *lizx_replay_meta(const lizx_replay *r, int which) {
  if (!r) return nullptr;
  return which == 0 ? r->r->meta().c_str() : which == 1 ? r->r->stats().c_str() : r->r->layout().c_str();
}

LIZ_API lizx_replay_writer *lizx_replay_create(const char *path) {
  return guard_new<lizx_replay_writer>([&] {
    need(path, "a path");
    auto w = std::make_unique<lizx_replay_writer>();
    w->w = std::make_unique<ReplayWriter>(path);
    return w.release();
  });
}

LIZ_API int lizx_replay_add(lizx_replay_writer *w, const uint8_t *luma, int width, int height, int stride) {
  return guard([&] {
    need(w && luma && width > 0 && height > 0 && width <= 16384 && height <= 16384 && stride >= width,
         "a writer and a luma frame of 1 to 16384 pixels a side, rows stride >= w bytes apart");
    w->w->frame(luma, static_cast<uint32_t>(width), static_cast<uint32_t>(height), static_cast<uint32_t>(stride));
    return static_cast<int>(w->w->frames());
  });
}

LIZ_API int64_t lizx_replay_finish(lizx_replay_writer *w, const char *meta, const char *stats) {
  int64_t bytes = -1;
  const int rc = guard([&] {
    need(w && meta, "a writer and its meta.json");
    bytes = w->w->finish(meta, stats ? stats : "");
    return 0;
  });
  return rc < 0 ? rc : bytes;
}

LIZ_API void lizx_replay_writer_free(lizx_replay_writer *w) { delete w; }

LIZ_API int lizx_word_geometry(int span, int *cells, int *bytes) {
  return guard([&] {
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    if (cells) *cells = fmt_side_cells(span + 30);
    if (bytes) *bytes = fmt_bytes(span + 30);
    return 0;
  });
}

LIZ_API int lizx_border_kinds(int span, int version, int fps, uint8_t *kinds) {
  return guard([&] {
    need(kinds && (span == 128 || span == 256 || span == 512), "a ring's span and a place for the map");
    need(version == 0 || (version >= FMT_VERSION_MIN && version <= FMT_VERSION_MAX), "version 4..64, or 0 for no word");
    need(fps >= 0 && fps <= 255, "fps 0..255");
    uint8_t cw[FMT_BYTES_MAX];
    if (version) {
      FmtWord fw;
      fw.version = version;
      fw.fps = fps;
      fmt_encode(fw, cw, fmt_bytes(span + 30));
    }
    border_kinds(span, version ? cw : nullptr, kinds);
    return (span + 30) * (span + 30);
  });
}

LIZ_API int lizx_word_encode(int span, int version, int fps, uint8_t *cw) {
  return guard([&] {
    need(cw && (span == 128 || span == 256 || span == 512), "a ring's span and a place for the codeword");
    const int bytes = fmt_bytes(span + 30);
    FmtWord fw;
    fw.version = version;
    fw.fps = fps;
    need(fmt_encode(fw, cw, bytes) == 0, "a version 4 to 64 and a rate 0 to 255");
    return bytes;
  });
}

LIZ_API int lizx_word_decode(int span, const float *q, int *version, int *fps) {
  return guard([&] {
    need(q && (span == 128 || span == 256 || span == 512), "a ring's span and the soft values");
    const int side = span + 30;
    FmtWord fw;
    if (fmt_decode(q, fmt_side_cells(side), fmt_bytes(side), &fw) != 0) return 0;
    if (version) *version = fw.version;
    if (fps) *fps = fw.fps;
    return 1;
  });
}

LIZ_API int lizx_read_word(int span, const uint8_t *img, int w, int h, int stride, const lizx_registration *reg,
                           float *q, int *version, int *fps) {
  return guard([&] {
    need(img && reg && reg->nodes && w >= 2 && h >= 2 && stride >= w, "an image of 2 x 2 pixels or more and a registration");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    const int side = span + 30, cells = fmt_side_cells(side), bytes = fmt_bytes(side);
    if (!cells) return 0;
    std::vector<float> qv(4 * cells);
    if (!read_word_soft(span, img, w, h, stride, reg->H, reg->nodes, qv.data(), cells)) return 0;
    if (q) std::memcpy(q, qv.data(), qv.size() * sizeof(float));
    FmtWord fw;
    if (fmt_decode(qv.data(), cells, bytes, &fw) != 0) return 0;
    if (version) *version = fw.version;
    if (fps) *fps = fw.fps;
    return 1;
  });
}

LIZ_API int lizx_node_lattice(int span, float *pos) {
  return guard([&] {
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    float tmp[MAX_NODES];
    const int n = node_lattice(span, tmp);
    if (pos) std::memcpy(pos, tmp, n * sizeof(float));
    return n;
  });
}

LIZ_API int lizx_warp(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                      const lizx_registration *reg, float *grid) {
  return guard([&] {
    need(d && grid, "a decoder and a place for the grid");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, span);
    const WarpFrame *wf = stage_capture(d, p, img, w, h, stride, reg, true);
    const size_t n = p.n;
    d->grid.reserve(n * n * sizeof(float));
    launch_warp(p, wf, 1, d->grid.as<float>(), e.stream);
    e.sync();
    copy_wait(grid, d->grid.p, n * n * sizeof(float), hipMemcpyDeviceToHost, e.stream, "grid");
    return 0;
  });
}

LIZ_API int lizx_decode_reg(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                            const lizx_registration *reg, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est,
                            lizx_soft *soft) {
  return guard([&] {
    need(d, "a decoder");
    return decode_capture(d, version, span, img, w, h, stride, reg, blocks, ok, its, blk_est, soft);
  });
}

LIZ_API int lizx_decoder_hold(liz_decoder *d, int version, int span, const lizx_registration *reg) {
  return guard([&] {
    need(d, "a decoder");
    if (!reg) {  // let the hold go: liz_decode finds each frame's symbol itself
      d->held_span = d->held_version = 0;
      d->held_nodes.clear();
      return 0;
    }
    need(reg->nodes, "a registration with nodes");
    if (version && !format_for(version)) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64, or 0)");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    float pos[MAX_NODES];
    const int nx = node_lattice(span, pos);
    d->held_nodes.assign(reg->nodes, reg->nodes + 2 * nx * nx);
    d->held_reg = *reg;
    d->held_reg.nodes = d->held_nodes.data();
    d->held_version = version;
    d->held_span = span;
    return 0;
  });
}

// The anchors' search on a lane's stream for anchor_register: every anchor of A not yet found searched through H (and
// the lens model, when given) on the device capture; returns how many are found in all.
static int lane_anchor_find(FinderLane &ln, int span, const uint8_t *d_img, int w, int h, const float *H,
                            const LensModel *lm, AnchorHit *A) {
  AnchorJob &J = *ln.h_anch.as<AnchorJob>();
  AnchorHit *hh = reinterpret_cast<AnchorHit *>(ln.h_anch.as<char>() + ANCH_JOB_BYTES);
  J = AnchorJob{};
  J.img = d_img;
  J.w = w;
  J.h = h;
  J.span = span;
  anchor_pattern(span, J);
  for (int i = 0; i < ANCH_K; i++) J.m0[i] = static_cast<float>(anchor_m0(span, i));
  std::memcpy(J.H, H, sizeof J.H);
  J.lens_on = lm ? 1 : 0;
  if (lm) {
    J.K = lm->K;
    for (int a = 0; a < LENS_P; a++) J.t[a] = lm->t[a];
    J.ax = lm->ax;
    J.ay = lm->ay;
    J.R = lm->R;
  }
  for (int k = 0; k < ANCH_N; k++)
    if (A[k].ok) J.skip |= 1ull << k;
  AnchorJob *dj = ln.anch.as<AnchorJob>();
  AnchorHit *dh = reinterpret_cast<AnchorHit *>(ln.anch.as<char>() + ANCH_JOB_BYTES);
  hip_check(hipMemcpyAsync(dj, &J, sizeof J, hipMemcpyHostToDevice, ln.s), "anchor job");
  launch_anchor_find(dj, 1, dh, ln.s);
  hip_check(hipGetLastError(), "anchor search");
  hip_check(hipMemcpyAsync(hh, dh, ANCH_N * sizeof(AnchorHit), hipMemcpyDeviceToHost, ln.s), "anchor records");
  hip_check(hipStreamSynchronize(ln.s), "anchor search");
  int found = 0;
  for (int k = 0; k < ANCH_N; k++) {
    if (!((J.skip >> k) & 1ull)) A[k] = hh[k];
    found += A[k].ok ? 1 : 0;
  }
  return found;
}

// The lattice's corners in pixels through H and the nodes, top left and clockwise, and pixels a module along the
// quad's sides over the lattice's (liblizard/src/acquire.c:5009-5013).
static float lattice_quad(int span, const float *H, const float *nodes, float *quad) {
  float pos[MAX_NODES], src[8];
  const int nx = node_lattice(span, pos);
  for (int c = 0; c < 4; c++) {
    const int i = (c == 1 || c == 2) ? nx - 1 : 0, j = c >= 2 ? nx - 1 : 0;
    src[2 * c] = pos[i];
    src[2 * c + 1] = pos[j];
    float u, v;
    project_f(H, pos[i], pos[j], &u, &v);
    quad[2 * c] = u + nodes[2 * (j * nx + i)];
    quad[2 * c + 1] = v + nodes[2 * (j * nx + i) + 1];
  }
  float mpx = 0;
  for (int k = 0; k < 4; k++) {
    const int k1 = (k + 1) & 3;
    mpx += std::hypot(quad[2 * k1] - quad[2 * k], quad[2 * k1 + 1] - quad[2 * k + 1]) /
           std::hypot(src[2 * k1] - src[2 * k], src[2 * k1 + 1] - src[2 * k + 1]) / 4;
  }
  return mpx;
}

// The border's marked nodes refined on the GPU through H and, when given, the anchors' lens model, in the reference's
// order (border.hpp); st: the lattice's nx * nx nodes (the corners done at 0, as the reference starts them).
static void border_refine(FinderLane &ln, int span, const uint8_t *d_img, int w, int h, const float *H, float mpx,
                          const LensModel *lm, std::vector<NodeState> &st) {
  BorderBufs &B = ln.border;
  float pos[MAX_NODES];
  const int nx = node_lattice(span, pos), nn = nx * nx, Lw = span + 2 * BORDER;
  if (B.span != span) {
    std::vector<short> order, level;
    B.levels = border_schedule(span, order, level);
    B.count = static_cast<int>(order.size());
    std::vector<uint8_t> kinds(static_cast<size_t>(Lw) * Lw), cw(static_cast<size_t>(std::max(fmt_bytes(Lw), 1)), 0);
    border_kinds(span, cw.data(), kinds.data());  // the word's cells flagged (their bits are no template's)
    B.kinds.reserve(kinds.size());
    B.sched.reserve((order.size() + level.size()) * sizeof(short));
    copy_wait(B.kinds.p, kinds.data(), kinds.size(), hipMemcpyHostToDevice, ln.s, "border layout");
    copy_wait(B.sched.p, order.data(), order.size() * sizeof(short), hipMemcpyHostToDevice, ln.s, "border order");
    copy_wait(B.sched.as<short>() + order.size(), level.data(), level.size() * sizeof(short),
                        hipMemcpyHostToDevice, ln.s, "border levels");
    B.span = span;
  }
  const uint8_t *rim = d_img;
  if (mpx / 2 >= 1) {  // the templates read the border band under a box a module wide (acquire.c:5014-5019)
    const int r = static_cast<int>(mpx / 2);
    int rects[16];
    border_band(span, H, w, h, r, rects);
    B.rim.reserve(static_cast<size_t>(w) * h);
    launch_border_blur(d_img, B.rim.as<uint8_t>(), w, h, r, rects, ln.s);
    hip_check(hipGetLastError(), "border blur");
    rim = B.rim.as<uint8_t>();
  }
  const size_t job_bytes = (sizeof(BorderJob) + 255) / 256 * 256, node_bytes = nn * sizeof(NodeState);
  B.job.reserve(job_bytes);
  B.nodes.reserve(node_bytes);
  B.h.reserve(job_bytes + node_bytes);
  BorderJob &J = *B.h.as<BorderJob>();
  NodeState *hn = reinterpret_cast<NodeState *>(B.h.as<char>() + job_bytes);
  J = BorderJob{};
  J.img = rim;
  J.kinds = B.kinds.as<uint8_t>();
  J.order = B.sched.as<short>();
  J.level = B.sched.as<short>() + B.count;
  J.nodes = B.nodes.as<NodeState>();
  J.w = w;
  J.h = h;
  J.span = span;
  J.nx = nx;
  J.Lw = Lw;
  J.levels = B.levels;
  std::memcpy(J.H, H, sizeof J.H);
  std::memcpy(J.pos, pos, sizeof J.pos);
  J.lens_on = lm ? 1 : 0;
  if (lm) {
    J.K = lm->K;
    for (int a = 0; a < LENS_P; a++) J.t[a] = lm->t[a];
    J.ax = lm->ax;
    J.ay = lm->ay;
    J.R = lm->R;
  }
  for (int k = 0; k < nn; k++) hn[k] = NodeState{0, 0, 0, 0};
  for (int c = 0; c < 4; c++) hn[(c >= 2 ? nx - 1 : 0) * nx + ((c == 1 || c == 2) ? nx - 1 : 0)].done = 1;
  hip_check(hipMemcpyAsync(B.job.p, &J, sizeof J, hipMemcpyHostToDevice, ln.s), "border job");
  hip_check(hipMemcpyAsync(B.nodes.p, hn, node_bytes, hipMemcpyHostToDevice, ln.s), "border nodes");
  launch_border_nodes(B.job.as<BorderJob>(), 1, ln.s);
  hip_check(hipGetLastError(), "border nodes");
  hip_check(hipMemcpyAsync(hn, B.nodes.p, node_bytes, hipMemcpyDeviceToHost, ln.s), "border nodes");
  hip_check(hipStreamSynchronize(ln.s), "border nodes");
  st.assign(hn, hn + nn);
}

// The finder's registration (H, and the nodes of its radial lens) taken on from the anchors: anchor_register through
// H; then (mode 2, the reference's chain) the border's marked nodes refined from the anchors' lens model and the
// interior made from the trusted ones (score 0.35 or more) and the anchors, or (mode 1) the interior made from the
// finder's own border nodes and the anchors. H and nodes are replaced where 12 or more anchors are found and their lens
// model fits; returns the anchors found then (else 0, the registration unchanged).
static int anchor_stage(FinderLane &ln, int mode, int span, const uint8_t *d_img, int w, int h, float *H,
                        float *nodes) {
  float quad[8], Hn[9];
  const float mpx = lattice_quad(span, H, nodes, quad);
  if (!(mpx > 0)) return 0;
  std::memcpy(Hn, H, sizeof Hn);
  float pos[MAX_NODES];
  const int nx = node_lattice(span, pos), nn = nx * nx;
  if (mode == 2) {  // H through the lattice's corners, as the reference's chain holds them (displacement 0 there)
    float src[8];
    for (int c = 0; c < 4; c++) {
      src[2 * c] = pos[(c == 1 || c == 2) ? nx - 1 : 0];
      src[2 * c + 1] = pos[c >= 2 ? nx - 1 : 0];
    }
    if (!homography4(src, quad, Hn)) return 0;
  }
  AnchorHit A[ANCH_N];
  LensModel lm;
  int moved = 0;
  const AnchorFind find = [&](const float *Hs, const LensModel *m, AnchorHit *an) {
    return lane_anchor_find(ln, span, d_img, w, h, Hs, m, an);
  };
  if (!anchor_register(span, Hn, w, h, mpx, quad, find, A, &lm, &moved)) return 0;
  std::vector<uint8_t> use;
  if (mode == 2) {
    std::vector<NodeState> st;
    border_refine(ln, span, d_img, w, h, Hn, mpx, &lm, st);
    use.assign(nn, 0);
    for (int k = 0; k < nn; k++) {
      nodes[2 * k] = st[k].dx;
      nodes[2 * k + 1] = st[k].dy;
      use[k] = st[k].done && st[k].score >= 0.35f;
    }
  } else if (moved) {  // the border nodes about the new H: the same pixels
    for (int j = 0; j < nx; j++)
      for (int i = 0; i < nx; i++) {
        if (i && j && i < nx - 1 && j < nx - 1) continue;
        float u0, v0, u1, v1;
        project_f(H, pos[i], pos[j], &u0, &v0);
        project_f(Hn, pos[i], pos[j], &u1, &v1);
        nodes[2 * (j * nx + i)] += u0 - u1;
        nodes[2 * (j * nx + i) + 1] += v0 - v1;
      }
  }
  AnchorReading R[ANCH_N];
  int found = 0;
  for (int k = 0; k < ANCH_N; k++) {
    R[k] = AnchorReading{A[k].X, A[k].Y, A[k].dx, A[k].dy, A[k].ok};
    found += A[k].ok ? 1 : 0;
  }
  const float nan = __builtin_nanf("");
  anchor_fill(span, Hn, w, h, nan, nan, nodes, use.empty() ? nullptr : use.data(), R, mpx, nullptr);
  std::memcpy(H, Hn, sizeof Hn);
  return found;
}

// The finder on a capture: uploads it, searches, and writes the registration (nodes into `nodes`, nx * nx pairs).
static bool find_registration(liz_decoder *d, const uint8_t *img, int w, int h, int stride, FindResult &fr,
                              lizx_registration &reg, std::vector<float> &nodes) {
  need(img && w >= 16 && h >= 16 && stride >= w, "a capture of 16 x 16 pixels or more, rows stride >= w bytes apart");
  LdpcEngine &e = *d->eng;
  hip_check(hipSetDevice(e.device), "hipSetDevice");
  const auto t0 = std::chrono::steady_clock::now();
  d->finder.img.reserve(static_cast<size_t>(w) * h);
  if (stride == w) {
    hip_check(hipMemcpyAsync(d->finder.img.p, img, static_cast<size_t>(w) * h, hipMemcpyHostToDevice, e.stream),
              "capture");
  } else {
    hip_check(hipMemcpy2DAsync(d->finder.img.p, w, img, stride, w, h, hipMemcpyHostToDevice, e.stream), "capture");
  }
  e.wait("capture upload");
  fr.ms[0] = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  find_symbol(d->finder, d->finder.img.as<uint8_t>(), img, w, h, stride, e.stream, fr);
  if (!fr.found) return false;
  for (int i = 0; i < 9; i++) reg.H[i] = static_cast<float>(fr.H[i]);
  RegModel m;
  reg_model(fr.span, reg.H, w, h, m);
  float params[REG_PARAMS] = {};
  for (int i = 0; i < 8; i++) params[i] = reg.H[i];
  params[8] = static_cast<float>(fr.k1);
  reg_nodes(m, params, nodes);
  if (d->anchor_mode)
    anchor_stage(d->finder.main, d->anchor_mode, fr.span, d->finder.img.as<uint8_t>(), w, h, reg.H, nodes.data());
  reg.nodes = nodes.data();
  return true;
}

LIZ_API int lizx_find(liz_decoder *d, const uint8_t *img, int w, int h, int stride, int *span, float *H, float *nodes,
                      float *quad, float *info) {
  return guard([&] {
    need(d && H && nodes, "a decoder and room for H and the nodes");
    FindResult fr;
    lizx_registration got{};
    std::vector<float> nv;
    if (!find_registration(d, img, w, h, stride, fr, got, nv)) return 0;
    std::memcpy(H, got.H, sizeof got.H);
    std::memcpy(nodes, nv.data(), nv.size() * sizeof(float));
    if (span) *span = fr.span;
    if (quad) std::memcpy(quad, fr.quad, sizeof fr.quad);
    if (info) {
      info[0] = static_cast<float>(fr.k1);
      info[1] = fr.track;
      info[2] = fr.contrast;
      info[3] = fr.rms;
      info[4] = static_cast<float>(fr.points);
      for (int k = 0; k < 6; k++) info[5 + k] = fr.ms[k];
    }
    return 1;
  });
}

LIZ_API int lizx_find_batch(liz_decoder *d, int frames, const uint8_t *const *imgs, int w, int h, int stride,
                            int *spans, float *H, float *nodes, int *found, int *versions) {
  return guard([&] {
    need(d && imgs && spans && H && nodes && found, "a decoder, the frames and room for what is found");
    need(frames >= 1 && frames <= 256, "1 to 256 frames");
    need(w >= 16 && h >= 16 && w <= 16384 && h <= 16384 && stride >= w, "frames of 16 to 16384 pixels a side");
    for (int f = 0; f < frames; f++) need(imgs[f], "every frame's pixels");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const size_t fb = static_cast<size_t>(w) * h;
    d->finder.img.reserve(fb * finder_slots(d->finder, frames));
    std::vector<FindResult> res(frames);
    // each frame's registration (and word) on the thread that found it, while the other frames are still searched
    auto done = [&](int f, FinderLane &ln) {
      const FindResult &fr = res[f];
      found[f] = fr.found;
      spans[f] = fr.found ? fr.span : 0;
      if (versions) versions[f] = -1;
      if (!fr.found) return;
      float *Hf = H + 9 * f;
      for (int i = 0; i < 9; i++) Hf[i] = static_cast<float>(fr.H[i]);
      RegModel m;
      reg_model(fr.span, Hf, w, h, m);
      float params[REG_PARAMS] = {};
      for (int i = 0; i < 8; i++) params[i] = Hf[i];
      params[8] = static_cast<float>(fr.k1);
      std::vector<float> nv;
      reg_nodes(m, params, nv);
      if (d->anchor_mode)
        anchor_stage(ln, d->anchor_mode, fr.span, ln.d_img, w, h, Hf, nv.data());
      float *nf = nodes + static_cast<size_t>(f) * MAX_NODES * MAX_NODES * 2;
      std::memcpy(nf, nv.data(), nv.size() * sizeof(float));
      if (!versions) return;
      const int side = fr.span + 30, cells = fmt_side_cells(side), bytes = fmt_bytes(side);
      std::vector<float> qv(4 * static_cast<size_t>(cells));
      FmtWord fw;
      if (cells && read_word_soft(fr.span, imgs[f], w, h, stride, Hf, nf, qv.data(), cells) &&
          fmt_decode(qv.data(), cells, bytes, &fw) == 0)
        versions[f] = fw.version;
    };
    find_symbols(d->finder, frames, d->finder.img.as<uint8_t>(), imgs, w, h, stride, e.stream, res.data(), done);
    int hits = 0;
    for (int f = 0; f < frames; f++) hits += res[f].found ? 1 : 0;
    return hits;
  });
}

// The registration (H from params with h8 = 1, nodes from k1 and the anchors) of a refinement call, staged with its
// capture and reference picture.
static const WarpFrame *stage_refinement(liz_decoder *d, const FramePlan &p, const uint8_t *img, int w, int h, int stride,
                                         const float *H0, const float *params, const float *ref, RegModel &model) {
  need(H0 && params && ref, "a reference homography, parameters and a reference picture");
  reg_model(p.span, H0, w, h, model);
  reg_nodes(model, params, d->reg_nodes);
  lizx_registration reg{};
  for (int k = 0; k < 8; k++) reg.H[k] = params[k];
  reg.H[8] = 1.0f;
  reg.nodes = d->reg_nodes.data();
  const WarpFrame *wf = stage_capture(d, p, img, w, h, stride, &reg);
  LdpcEngine &e = *d->eng;
  const size_t nn = static_cast<size_t>(p.n) * p.n;
  const size_t cells = static_cast<size_t>(p.dev.nx - 1) * (p.dev.nx - 1) * REG_SPLIT;
  d->reg_ref.reserve(nn * sizeof(float));
  hip_check(hipMemcpyAsync(d->reg_ref.p, ref, nn * sizeof(float), hipMemcpyHostToDevice, e.stream), "reference picture");
  d->reg_part.reserve(cells * REG_PART * sizeof(float));
  d->reg_sum.reserve(REG_PART * sizeof(double));
  d->reg_state.reserve(sizeof(RegState));
  return wf;
}

LIZ_API int lizx_reg_border_fill(int span, const float *H, int w, int h, const float *axis, float *nodes,
                                 const uint8_t *use, double *model) {
  return guard([&] {
    need(H && nodes && w > 0 && h > 0, "a homography, nodes and a frame size");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    LensModel m;
    const float nan = __builtin_nanf("");
    const int K = border_fill(span, H, w, h, axis ? axis[0] : nan, axis ? axis[1] : nan, nodes, use, &m);
    if (model) {
      model[0] = m.K;
      model[1] = m.ax;
      model[2] = m.ay;
      model[3] = m.R;
      for (int a = 0; a < LENS_P; a++) model[4 + a] = m.t[a];
    }
    return K;
  });
}

LIZ_API int lizx_reg_anchor_fill(int span, const float *H, int w, int h, const float *axis, float *nodes,
                                 const uint8_t *use, const float *anchors, float mpx, double *model) {
  return guard([&] {
    need(H && nodes && anchors && w > 0 && h > 0, "a homography, nodes, anchors and a frame size");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    need(mpx > 0, "a positive module size in pixels");
    AnchorReading A[REG_ANCH * REG_ANCH];
    for (int k = 0; k < REG_ANCH * REG_ANCH; k++)
      A[k] = AnchorReading{anchors[5 * k], anchors[5 * k + 1], anchors[5 * k + 2], anchors[5 * k + 3],
                           anchors[5 * k + 4] != 0};
    LensModel m;
    const float nan = __builtin_nanf("");
    const int K = anchor_fill(span, H, w, h, axis ? axis[0] : nan, axis ? axis[1] : nan, nodes, use, A, mpx, &m);
    if (model) {
      model[0] = m.K;
      model[1] = m.ax;
      model[2] = m.ay;
      model[3] = m.R;
      for (int a = 0; a < LENS_P; a++) model[4 + a] = m.t[a];
    }
    return K;
  });
}

LIZ_API int lizx_reg_lens_field(int span, const float *H, const double *model, float *nodes) {
  return guard([&] {
    need(H && model && nodes, "a homography, a lens model and a place for the nodes");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    LensModel m;
    m.K = static_cast<int>(model[0]);
    need(m.K >= -1 && m.K <= LENS_TERMS, "a lens model of 0 to 3 radial terms (or -1: none)");
    m.ax = model[1];
    m.ay = model[2];
    m.R = model[3];
    need(m.R > 0, "a positive radius");
    for (int a = 0; a < LENS_P; a++) m.t[a] = model[4 + a];
    float pos[MAX_NODES];
    const int nx = node_lattice(span, pos);
    for (int j = 0; j < nx; j++)
      for (int i = 0; i < nx; i++) {
        float *o = nodes + 2 * (j * nx + i);
        if (m.K < 0) {
          o[0] = o[1] = 0.f;
          continue;
        }
        const float mx = pos[i], my = pos[j], ww = H[6] * mx + H[7] * my + 1;
        const float u = (H[0] * mx + H[1] * my + H[2]) / ww, v = (H[3] * mx + H[4] * my + H[5]) / ww;
        lens_at(m, u, v, &o[0], &o[1]);
      }
    return nx;
  });
}

LIZ_API int lizx_reg_nodes(int span, const float *H0, int w, int h, const float *params, float *nodes) {
  return guard([&] {
    need(H0 && params && nodes && w > 0 && h > 0, "a reference homography, parameters, a frame size and an output");
    need(span == 128 || span == 256 || span == 512, "span is a ring's: 128, 256 or 512 modules");
    RegModel m;
    reg_model(span, H0, w, h, m);
    std::vector<float> out;
    reg_nodes(m, params, out);
    std::copy(out.begin(), out.end(), nodes);
    return m.nx;
  });
}

// One pass of the warp's adjoint on a staged refinement (with d_dl: e = d_dl / 2, a loss's derivative with respect
// to the warped picture): the pieces' sums in FP64, in piece order: the sum of e^2, the gradient over the 107
// parameters (the nodes' gradient from the corners of the cells about them, then on to k1 and the anchors), J^T J
// over h0..h7 and k1 (k1's column 0 without d_lens) and the intensity statistics.
struct AdjointSums {
  double sq = 0;
  double grad[REG_PARAMS] = {};
  double jtj[REG_GN * REG_GN] = {};
  double stats[5] = {};
  std::vector<double> met;  // each node's share of the warp's metric (gx^2, gx gy, gy^2 weighted), nx * nx * 3
};
static void adjoint_pass(liz_decoder *d, const FramePlan &p, const WarpFrame *wf, const RegModel &model,
                         const float *params, float alpha, float beta, const float *d_dl, const float2 *d_lens,
                         AdjointSums &out) {
  LdpcEngine &e = *d->eng;
  RegState &st = d->reg_state_host;
  for (int k = 0; k < 8; k++) st.H[k] = params[k];
  st.H[8] = 1.0f;
  st.alpha = alpha;
  st.beta = beta;
  st.dk1 = 0.0f;
  hip_check(hipMemcpyAsync(d->reg_state.p, &st, sizeof st, hipMemcpyHostToDevice, e.stream), "state");
  launch_reg_accum(p, wf, d->reg_state.as<RegState>(), d->reg_ref.as<float>(), d_lens, d->reg_part.as<float>(),
                   e.stream, d_dl);
  const int nx = p.dev.nx, c = nx - 1;
  d->reg_part_host.resize(static_cast<size_t>(c) * c * REG_SPLIT * REG_PART);
  hip_check(hipMemcpyAsync(d->reg_part_host.data(), d->reg_part.p, d->reg_part_host.size() * sizeof(float),
                           hipMemcpyDeviceToHost, e.stream), "partial sums");
  e.wait("refinement pass");
  constexpr int JTJ0 = 1 + REG_GN, STAT0 = JTJ0 + REG_GN * (REG_GN + 1) / 2;
  double L = 0, gH[8] = {0, 0, 0, 0, 0, 0, 0, 0}, up[REG_GN * (REG_GN + 1) / 2] = {}, stats[5] = {};
  std::vector<double> G(static_cast<size_t>(nx) * nx * 2, 0.0);
  out.met.assign(static_cast<size_t>(nx) * nx * 3, 0.0);
  for (int row = 0; row < c * REG_SPLIT; row++)
    for (int ci = 0; ci < c; ci++) {
      const int cj = row / REG_SPLIT;
      const float *q = &d->reg_part_host[(static_cast<size_t>(row) * c + ci) * REG_PART];
      L += q[0];
      for (int k = 0; k < 8; k++) gH[k] += q[1 + k];
      for (int k = 0; k < REG_GN * (REG_GN + 1) / 2; k++) up[k] += q[JTJ0 + k];
      for (int k = 0; k < 5; k++) stats[k] += q[STAT0 + k];
      const int corner[4] = {cj * nx + ci, cj * nx + ci + 1, (cj + 1) * nx + ci, (cj + 1) * nx + ci + 1};
      for (int k = 0; k < 4; k++) {
        G[2 * corner[k]] += 2.0 * q[REG_NODE0 + 2 * k];
        G[2 * corner[k] + 1] += 2.0 * q[REG_NODE0 + 1 + 2 * k];
        for (int m = 0; m < 3; m++) out.met[3 * corner[k] + m] += q[REG_MET0 + 3 * k + m];
      }
    }
  out.sq = L;
  for (int k = 0; k < 8; k++) out.grad[k] = 2.0 * gH[k];
  reg_chain(model, G.data(), out.grad);
  for (int k = 0, q = 0; k < REG_GN; k++)
    for (int l = k; l < REG_GN; l++, q++) out.jtj[k * REG_GN + l] = out.jtj[l * REG_GN + k] = up[q];
  for (int k = 0; k < 5; k++) out.stats[k] = stats[k];
}

LIZ_API int lizx_reg_loss_grad(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                               const float *H0, const float *params, const float *ref, float alpha, float beta,
                               double *loss, double *grad) {
  return guard([&] {
    need(d && loss && grad, "a decoder and outputs");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, span);
    RegModel model;
    const WarpFrame *wf = stage_refinement(d, p, img, w, h, stride, H0, params, ref, model);
    AdjointSums a;
    adjoint_pass(d, p, wf, model, params, alpha, beta, nullptr, nullptr, a);
    *loss = a.sq;
    std::copy(a.grad, a.grad + REG_PARAMS, grad);
    return 0;
  });
}

// The spectral loss on a staged refinement: the warped picture, its residual's masked coefficients and their
// inverse transform (launch_spectral_adjoint), then the warp's adjoint of that derivative; the mask already on the
// device (d->spec_mask).
static double spectral_pass(liz_decoder *d, const FramePlan &p, const WarpFrame *wf, const RegModel &model,
                            const float *params, float alpha, float beta, const float2 *d_lens, AdjointSums &out) {
  LdpcEngine &e = *d->eng;
  const size_t nn = static_cast<size_t>(p.n) * p.n, S = static_cast<size_t>(p.subch) * SUB,
               X = static_cast<size_t>(p.E) * p.n;
  d->spec_w.reserve(nn * sizeof(float));
  d->spec_dl.reserve(nn * sizeof(float));
  d->spec_X.reserve(X * sizeof(float2));
  d->spec_G.reserve(X * sizeof(float2));
  d->spec_C.reserve(S * sizeof(float2));
  d->spec_part.reserve(((S + 255) / 256) * sizeof(double));
  launch_warp(p, wf, 1, d->spec_w.as<float>(), e.stream);
  const int parts = launch_spectral_adjoint(p, d->spec_w.as<float>(), d->reg_ref.as<float>(), alpha, beta,
                                            d->spec_mask.as<uint8_t>(), d->spec_X.as<float2>(),
                                            d->spec_C.as<float2>(), d->spec_part.as<double>(),
                                            d->spec_G.as<float2>(), d->spec_dl.as<float>(), e.stream);
  d->spec_part_host.resize(parts);
  hip_check(hipMemcpyAsync(d->spec_part_host.data(), d->spec_part.p, parts * sizeof(double), hipMemcpyDeviceToHost,
                           e.stream), "loss parts");
  adjoint_pass(d, p, wf, model, params, alpha, beta, d->spec_dl.as<float>(), d_lens, out);  // waits for the stream
  double L = 0;
  for (int k = 0; k < parts; k++) L += d->spec_part_host[k];
  return 2.0 * L / (static_cast<double>(p.n) * p.n);
}

static void upload_mask(liz_decoder *d, const FramePlan &p, const uint8_t *mask) {
  const size_t S = static_cast<size_t>(p.subch) * SUB;
  d->spec_mask.reserve(S);
  hip_check(hipMemcpyAsync(d->spec_mask.p, mask, S, hipMemcpyHostToDevice, d->eng->stream), "coefficient mask");
}

LIZ_API int lizx_reg_loss_grad_spectral(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h,
                                        int stride, const float *H0, const float *params, const float *ref,
                                        const uint8_t *mask, float alpha, float beta, double *loss, double *grad) {
  return guard([&] {
    need(d && loss && grad && mask, "a decoder, a coefficient mask and outputs");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, span);
    RegModel model;
    const WarpFrame *wf = stage_refinement(d, p, img, w, h, stride, H0, params, ref, model);
    upload_mask(d, p, mask);
    AdjointSums a;
    *loss = spectral_pass(d, p, wf, model, params, alpha, beta, nullptr, a);
    std::copy(a.grad, a.grad + REG_PARAMS, grad);
    return 0;
  });
}

// The staged registration moved to new parameters: the model's nodes for them and H (the capture stays staged).
static void restage(liz_decoder *d, const RegModel &model, const float *params) {
  LdpcEngine &e = *d->eng;
  reg_nodes(model, params, d->reg_nodes);
  hip_check(hipMemcpyAsync(d->nodes.p, d->reg_nodes.data(), d->reg_nodes.size() * sizeof(float), hipMemcpyHostToDevice,
                           e.stream), "nodes");
  for (int k = 0; k < 8; k++) d->wframe_host.H[k] = params[k];
  d->wframe_host.H[8] = 1.0f;
  hip_check(hipMemcpyAsync(d->wframe.p, &d->wframe_host, sizeof(WarpFrame), hipMemcpyHostToDevice, e.stream),
            "registration");
}

// (A + damping diag A) x = b for the n x n symmetric positive A (row-major), Jacobi-scaled, by Cholesky in FP64
// (host); false where A is not positive.
static bool solve_spd(const double *A0, const double *b0, int n, double damping, double *x) {
  std::vector<double> A(static_cast<size_t>(n) * n), b(n), sc(n);
  for (int k = 0; k < n; k++) {
    const double dg = A0[static_cast<size_t>(k) * n + k];
    sc[k] = dg > 0 ? 1.0 / std::sqrt(dg * (1.0 + damping)) : 1.0;
  }
  for (int k = 0; k < n; k++) {
    for (int l = 0; l < n; l++)
      A[static_cast<size_t>(k) * n + l] = A0[static_cast<size_t>(k) * n + l] * (k == l ? 1.0 + damping : 1.0) * sc[k] * sc[l];
    b[k] = b0[k] * sc[k];
  }
  auto at = [&](int i, int j) -> double & { return A[static_cast<size_t>(i) * n + j]; };
  for (int j = 0; j < n; j++) {
    double dd = at(j, j);
    for (int k = 0; k < j; k++) dd -= at(j, k) * at(j, k);
    if (!(dd > 0)) return false;
    at(j, j) = std::sqrt(dd);
    for (int i = j + 1; i < n; i++) {
      double v = at(i, j);
      for (int k = 0; k < j; k++) v -= at(i, k) * at(j, k);
      at(i, j) = v / at(j, j);
    }
  }
  for (int i = 0; i < n; i++) {
    double v = b[i];
    for (int k = 0; k < i; k++) v -= at(i, k) * b[k];
    b[i] = v / at(i, i);
  }
  for (int i = n - 1; i >= 0; i--) {
    double v = b[i];
    for (int k = i + 1; k < n; k++) v -= at(k, i) * x[k];
    x[i] = v / at(i, i);
  }
  for (int k = 0; k < n; k++) x[k] *= sc[k];
  return true;
}

// The anchors' metric from the nodes' (each node's 2 x 2 block; the blocks between neighbouring nodes left out):
// A[(a, c), (b, d)] = sum over nodes of phi_a phi_b M_node[c][d], 98 x 98.
static void anchor_metric(const RegModel &m, const std::vector<double> &met, std::vector<double> &A) {
  constexpr int NA = REG_ANCH * REG_ANCH, NP = 2 * NA;
  A.assign(static_cast<size_t>(NP) * NP, 0.0);
  for (int k = 0; k < m.nx * m.nx; k++) {
    const double *ph = &m.phi[static_cast<size_t>(k) * NA];
    const double mxx = met[3 * k], mxy = met[3 * k + 1], myy = met[3 * k + 2];
    for (int a = 0; a < NA; a++) {
      if (ph[a] == 0) continue;
      for (int b = 0; b < NA; b++) {
        const double w = ph[a] * ph[b];
        if (w == 0) continue;
        A[(2 * a) * NP + 2 * b] += w * mxx;
        A[(2 * a) * NP + 2 * b + 1] += w * mxy;
        A[(2 * a + 1) * NP + 2 * b] += w * mxy;
        A[(2 * a + 1) * NP + 2 * b + 1] += w * myy;
      }
    }
  }
}

LIZ_API int lizx_reg_refine_spectral(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h,
                                     int stride, const float *H0, float *params, const float *ref,
                                     const uint8_t *mask, int steps, int anchors, float *trace) {
  return guard([&] {
    need(d && params && mask && steps >= 0 && steps <= 100, "a decoder, parameters, a mask and 0 to 100 steps");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, span);
    RegModel model;
    const WarpFrame *wf = stage_refinement(d, p, img, w, h, stride, H0, params, ref, model);
    upload_mask(d, p, mask);
    std::vector<float> lf(model.lens.size());
    for (size_t k = 0; k < lf.size(); k++) lf[k] = static_cast<float>(model.lens[k]);
    d->reg_lens.reserve(lf.size() * sizeof(float));
    hip_check(hipMemcpyAsync(d->reg_lens.p, lf.data(), lf.size() * sizeof(float), hipMemcpyHostToDevice, e.stream),
              "lens field");
    const float2 *d_lens = d->reg_lens.as<float2>();
    const size_t S = static_cast<size_t>(p.subch) * SUB;
    size_t on = 0;
    for (size_t k = 0; k < S; k++) on += mask[k] != 0;
    // the loss's curvature is about the masked share of the picture's: the first trial step is the spatial Gauss-
    // Newton step over that share, halved until the loss falls
    const double share = std::max(0.05, static_cast<double>(on) / static_cast<double>(S));
    std::vector<float> cur(params, params + REG_PARAMS), trial(REG_PARAMS);
    AdjointSums a, b;
    // the intensity fit from a first pass (the statistics do not depend on alpha and beta), held after
    (void)spectral_pass(d, p, wf, model, cur.data(), 1.0f, 0.0f, d_lens, a);
    const double sw = a.stats[0], sr = a.stats[1], swr = a.stats[2], srr = a.stats[3], cnt = a.stats[4];
    const double var = srr - sr * sr / std::max(cnt, 1.0);
    const float alpha = var > 0 ? static_cast<float>((swr - sw * sr / cnt) / var) : 1.0f;
    const float beta = cnt > 0 ? static_cast<float>((sw - alpha * sr) / cnt) : 0.0f;
    double L = spectral_pass(d, p, wf, model, cur.data(), alpha, beta, d_lens, a);
    if (trace) trace[0] = static_cast<float>(L);
    // a step along dx over parameters [first, first + n): the step halved from the masked share's full one until
    // the loss falls (at most 6 halvings)
    auto try_step = [&](int first, int n, const double *dx) {
      for (double t = 1.0 / share; t > 1.0 / (share * 64.0); t *= 0.5) {
        trial = cur;
        for (int k = 0; k < n; k++) trial[first + k] = static_cast<float>(cur[first + k] + t * dx[k]);
        restage(d, model, trial.data());
        const double Lt = spectral_pass(d, p, wf, model, trial.data(), alpha, beta, d_lens, b);
        if (Lt < L) {
          cur = trial;
          a = b;
          L = Lt;
          return true;
        }
      }
      restage(d, model, cur.data());
      return false;
    };
    constexpr int NA2 = 2 * REG_ANCH * REG_ANCH;
    std::vector<double> AA;
    int s = 0;
    for (; s < steps; s++) {
      bool moved = false;
      double rhs[REG_GN], dx[REG_GN];
      for (int k = 0; k < REG_GN; k++) rhs[k] = -0.5 * a.grad[k];
      if (solve_spd(a.jtj, rhs, REG_GN, 1e-6, dx)) moved |= try_step(0, REG_GN, dx);
      if (anchors) {
        // the anchors by the nodes' metric through the mesh, H and k1 held
        double ra[NA2], da[NA2];
        for (int k = 0; k < NA2; k++) ra[k] = -0.5 * a.grad[REG_GN + k];
        anchor_metric(model, a.met, AA);
        if (solve_spd(AA.data(), ra, NA2, 1e-3, da)) moved |= try_step(REG_GN, NA2, da);
      }
      if (trace) trace[s + 1] = static_cast<float>(L);
      if (!moved) break;
    }
    for (int k = s + 1; trace && k <= steps; k++) trace[k] = static_cast<float>(L);
    std::copy(cur.begin(), cur.end(), params);
    return 0;
  });
}

// steps Gauss-Newton steps over H (h0..h7 in and out) and, with a lens field (d_lens, the unit field at the nodes),
// k1 too (its change from the staged nodes' into *dk1) on a staged refinement's warp; trace as lizx_reg_refine's.
static void gauss_newton(liz_decoder *d, const FramePlan &p, const WarpFrame *wf, float *H8, int steps, float *trace,
                         const float2 *d_lens = nullptr, float *dk1 = nullptr) {
  LdpcEngine &e = *d->eng;
  RegState &st = d->reg_state_host;
  for (int k = 0; k < 8; k++) st.H[k] = H8[k];
  st.H[8] = 1.0f;
  st.alpha = 1.0f;
  st.beta = 0.0f;
  st.dk1 = 0.0f;
  hip_check(hipMemcpyAsync(d->reg_state.p, &st, sizeof st, hipMemcpyHostToDevice, e.stream), "state");
  d->reg_trace.reserve((steps + 2) * sizeof(float));
  const int nparam = d_lens ? 9 : 8;
  // pass 0 fits the intensity model only; passes 1 .. steps move H (and k1); the last pass measures the result
  for (int s = 0; s <= steps + 1; s++) {
    launch_reg_accum(p, wf, d->reg_state.as<RegState>(), d->reg_ref.as<float>(), d_lens, d->reg_part.as<float>(),
                     e.stream);
    launch_gn_step(p, d->reg_part.as<float>(), d->reg_sum.as<double>(), d->reg_state.as<RegState>(), 1e-6f,
                   s >= 1 && s <= steps, nparam, d->reg_trace.as<float>(), s, e.stream);
  }
  hip_check(hipMemcpyAsync(&st, d->reg_state.p, sizeof st, hipMemcpyDeviceToHost, e.stream), "state");
  std::vector<float> tr(steps + 2);
  hip_check(hipMemcpyAsync(tr.data(), d->reg_trace.p, tr.size() * sizeof(float), hipMemcpyDeviceToHost, e.stream),
            "trace");
  e.wait("refinement");
  for (int k = 0; k < 8; k++) H8[k] = st.H[k];
  if (dk1) *dk1 = st.dk1;
  if (trace) std::copy(tr.begin() + 1, tr.end(), trace);  // the loss before each step and after the last
}

LIZ_API int lizx_reg_refine(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                            const float *H0, float *params, const float *ref, int steps, float *trace) {
  return guard([&] {
    need(d && steps >= 0 && steps <= 100, "a decoder and 0 to 100 steps");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, span);
    RegModel model;
    const WarpFrame *wf = stage_refinement(d, p, img, w, h, stride, H0, params, ref, model);
    gauss_newton(d, p, wf, params, steps, trace);
    return 0;
  });
}

LIZ_API int lizx_reg_refine_k1(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                               const float *H0, float *params, const float *ref, int steps, float *trace) {
  return guard([&] {
    need(d && steps >= 0 && steps <= 100, "a decoder and 0 to 100 steps");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, span);
    RegModel model;
    const WarpFrame *wf = stage_refinement(d, p, img, w, h, stride, H0, params, ref, model);
    // the unit lens field at the nodes (the nodes' derivative in k1), as the warp's float pairs
    std::vector<float> lf(model.lens.size());
    for (size_t k = 0; k < lf.size(); k++) lf[k] = static_cast<float>(model.lens[k]);
    d->reg_lens.reserve(lf.size() * sizeof(float));
    hip_check(hipMemcpyAsync(d->reg_lens.p, lf.data(), lf.size() * sizeof(float), hipMemcpyHostToDevice, e.stream),
              "lens field");
    float dk1 = 0;
    gauss_newton(d, p, wf, params, steps, trace, d->reg_lens.as<float2>(), &dk1);
    params[8] += dk1;
    return 0;
  });
}

LIZ_API int lizx_reg_refine_nodes(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                                  float *H, const float *nodes, const float *ref, int steps, float *trace) {
  return guard([&] {
    need(d && H && nodes && ref && steps >= 0 && steps <= 100,
         "a decoder, a homography, nodes, a reference picture and 0 to 100 steps");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, span);
    need(std::fabs(H[8] - 1.0f) < 1e-6f, "a homography with h8 = 1");
    lizx_registration reg{};
    std::memcpy(reg.H, H, sizeof reg.H);
    reg.nodes = nodes;
    const WarpFrame *wf = stage_capture(d, p, img, w, h, stride, &reg);
    const size_t nn = static_cast<size_t>(p.n) * p.n;
    const size_t cells = static_cast<size_t>(p.dev.nx - 1) * (p.dev.nx - 1) * REG_SPLIT;
    d->reg_ref.reserve(nn * sizeof(float));
    hip_check(hipMemcpyAsync(d->reg_ref.p, ref, nn * sizeof(float), hipMemcpyHostToDevice, e.stream),
              "reference picture");
    d->reg_part.reserve(cells * REG_PART * sizeof(float));
    d->reg_sum.reserve(REG_PART * sizeof(double));
    d->reg_state.reserve(sizeof(RegState));
    gauss_newton(d, p, wf, H, steps, trace);
    return 0;
  });
}

LIZ_API int lizx_finish_llr(liz_decoder *d, int version, const int8_t *llr, const float *est, uint8_t *blocks,
                            uint8_t *ok, int8_t *its, float *blk_est) {
  return guard([&] {
    need(d && llr && est && blocks && ok && its && blk_est, "a decoder, soft values, estimates and outputs");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    if (f->n > d->nmax) throw Error(LIZ_E_FORMAT, "the format's picture is larger than the decoder's nmax");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const int S = f->subch * 640;
    std::vector<double> est_d(f->subch);
    for (int j = 0; j < f->subch; j++) est_d[j] = est[j];
    d->llr.reserve(S);
    d->est.reserve(f->subch * sizeof(double));
    hip_check(hipMemcpyAsync(d->llr.p, llr, S, hipMemcpyHostToDevice, e.stream), "soft values");
    hip_check(hipMemcpyAsync(d->est.p, est_d.data(), f->subch * sizeof(double), hipMemcpyHostToDevice, e.stream),
              "estimates");
    finish_blocks(d, f, d->llr.as<int8_t>(), d->est.as<double>());
    return collect_blocks(d, f, blocks, ok, its, blk_est);
  });
}

LIZ_API int lizx_finish_grid(liz_decoder *d, int version, int span, const float *grid, uint8_t *blocks, uint8_t *ok,
                             int8_t *its, float *blk_est, lizx_soft *soft) {
  return guard([&] {
    need(d && grid && blocks && ok && its && blk_est, "a decoder, a grid and outputs");
    const Format *f = format_for(version);
    if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
    if (f->n > d->nmax) throw Error(LIZ_E_FORMAT, "the format's picture is larger than the decoder's nmax");
    LdpcEngine &e = *d->eng;
    hip_check(hipSetDevice(e.device), "hipSetDevice");
    const FramePlan &p = d->plan(version, span);
    const bool dbg = soft && (soft->coef || soft->soft);
    const size_t n = p.n;
    d->fb.reserve(p, 1, dbg);
    d->grid.reserve(n * n * sizeof(float));
    hip_check(hipMemcpyAsync(d->grid.p, grid, n * n * sizeof(float), hipMemcpyHostToDevice, e.stream), "grid");
    SoftBufs v = d->fb.view(dbg);
    v.blk_pilot = pilot_area(d, f);
    v.align = align_area(d, f);
    launch_soft(p, v, d->grid.as<float>(), 1, e.stream);
    finish_blocks(d, f, v.llr, v.est);
    const int good = collect_blocks(d, f, blocks, ok, its, blk_est, soft ? soft->blk_pilot : nullptr,
                                  soft ? soft->align : nullptr);
    if (soft) copy_soft(d, p, soft, dbg);
    return good;
  });
}

}  // extern "C"
