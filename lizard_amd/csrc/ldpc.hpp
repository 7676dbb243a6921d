// The LDPC stage on the GPU: task descriptors, the per-device engine that owns the code tables, and the launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "format.hpp"
#include "ldpc_dev.hpp"

namespace lz {

// One codeword to decode. Codeword mode: llr is the offset of its n LLRs in codeword order. Frame mode: llr is the
// offset of its block's first slot in the frame's soft values (slot order, whitened), slot0 that slot's index in the
// frame (the whitening), est the index of the frame's first sub-channel estimate (the decline gate).
struct LdpcTask {
  int64_t llr;
  int64_t bits;   // offset of n bytes of hard decisions (and, traced, of n int16 posteriors), or -1
  int32_t code;   // code index 0..3 (7/8, 3/4, 2/3, 1/2)
  int32_t slot0;
  int32_t est;
  int32_t out;    // output record: its[out], blk_est[out], crc_ok[out], bytes[out * 476]
};

enum { MODE_CODEWORD = 0, MODE_FRAME = 1 };

struct LdpcParams {
  const LdpcTask *tasks;
  int count;
  int *queue;              // persistent queue counter (zero before launch), or null: one task per workgroup
  const int8_t *llr;
  uint8_t *bits;
  int8_t *its;
  float *blk_est;
  uint8_t *bytes;
  uint8_t *crc_ok;
  int16_t *post;
  int8_t *msg;
  const uint8_t *white;    // frame mode: whitening bit per frame slot
  const int16_t *s2p[NCODES];  // frame mode: slot -> posterior index per code
  const double *est;       // frame mode: sub-channel estimates
  const uint32_t *crc_tab; // 3808 per-bit CRC words
  uint32_t crc0;
  int max_iter, stall_it;
  float stall_ratio;
  int mode, flags;
};

constexpr int LDPC_THREADS = 128;

// The sender's bit side on the device: per-code bit maps, the whitening, the CRC table.
struct SlotTables {
  const int16_t *s2p[NCODES];
  const uint8_t *white;
  const uint32_t *crc_tab;
  uint32_t crc0;
};
constexpr size_t MSG_TRACE_STRIDE = MAX_SLOTS_TOTAL * Z;  // 26880 bytes a codeword

// Per-device state: the code tables in constant memory, the CRC table, frame-mode bit maps, the queue counter.
struct LdpcEngine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf queue, crc_tab, s2p[NCODES], white;
  uint32_t crc0 = 0;
  int cus = 0, wg_per_cu = 0;
  float last_ms = 0;
  bool armed = false;
  bool lost = false;  // a wait on the stream passed the hang bound: the handle refuses further work
  explicit LdpcEngine(int device);
  ~LdpcEngine();
  // Enqueue a decode of count tasks already on the device. Plain: one launch per task. Timed: HIP events bracket the
  // launches from the first since the last sync() to the last.
  // per_cu > 0 caps the persistent grid at that many workgroups a CU (a launch shape: the bits do not change).
  // A non-null p.queue is the queue counter to use (zeroed by make_tasks); null: the engine's own.
  void launch(const LdpcParams &p, bool plain, bool timed, bool queue_zeroed = false, int per_cu = 0);
  // Writes the task list and zeroes the queue counter (d_queue, or null for the engine's own).
  void make_tasks(LdpcTask *d_tasks, const uint8_t *d_rate, const int64_t *d_off, int count, int *d_queue = nullptr);
  void sync();
  // Waits for the stream within the hang bound (wait_bounded); past it the engine is lost and every later wait throws.
  void wait(const char *what);
  void ensure_frame_tables(int max_slots);
  // The frame slots (0/1 bytes, total a frame) of frames x blocks x 476 bytes on the device, enqueued on the stream;
  // d_blk: per block first slot, nt, slots, code index (contiguous from slot 0).
  void frame_slots(const int4 *d_blk, int blocks, int total, const uint8_t *d_blocks, int cnt, uint8_t *d_slots,
                   int frames);
};

float copy_probe(void *dst, const void *src, size_t bytes);

// The device-resident work area of the decode that allocates nothing (lizx_ldpc_decode_device_noalloc): one per GPU
// for the process's life, shared by every call on that GPU. SCRATCH_TASKS task descriptors, then the queue counter.
constexpr int SCRATCH_TASKS = 16384;
struct LdpcScratch {
  LdpcTask *tasks;
  int *queue;
};
// This device's work area (its address looked up once per device, with that device current); throws if the lookup
// fails.
LdpcScratch ldpc_scratch(int device);

}  // namespace lz
