// Layered min-sum LDPC + CRC-32C on gfx950: one codeword per 128-lane workgroup (two waves; lane i owns check position
// i of every block row, a workgroup barrier between block rows), the codeword's int16 posteriors in LDS and its int8
// messages in VGPRs, the code's tables compiled in (ldpc_tmpl.hpp). A persistent grid pulls codewords off an atomic queue ordered by code, so codewords of
// four rates with uneven iteration counts share one launch; the plain path launches one workgroup per codeword.
//
// Follows liblizard/src/ldpc.c:1080-1172 (ldpc_decode_stall) and, for frame mode, liblizard/src/lizard_phy.c:1775-1870
// (decline gate, bit map, known zeros, bit pack, CRC compare).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "ldpc.hpp"
#include "ldpc_tmpl.hpp"

namespace lz {

__global__ void k_zero_int(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__device__ inline uint32_t wave_xor(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v ^= static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, 64));
  return v;
}
__device__ inline int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over the workgroup's two waves; slot is a 2-int LDS area no other lane uses until the next barrier.
__device__ inline int group_sum(int v, int *slot, int lane) {
  v = wave_sum(v);
  if ((lane & 63) == 0) slot[lane >> 6] = v;
  __syncthreads();
  return slot[0] + slot[1];
}
__device__ inline uint32_t group_xor(uint32_t v, uint32_t *slot, int lane) {
  v = wave_xor(v);
  if ((lane & 63) == 0) slot[lane >> 6] = v;
  __syncthreads();
  return slot[0] ^ slot[1];
}

// Codeword-mode tasks from device arrays of rate indices and offsets, ordered by code (1/2 first: the longest
// codewords lead the queue, the shortest fill its tail) by a counting sort in one workgroup; a rate with no code gets
// code -1 (not decoded, its 0) at the end. Also zeroes the queue counter for the decode that follows on the stream.
__global__ __launch_bounds__(1024) void k_make_tasks(LdpcTask *t, const uint8_t *rate, const int64_t *off, int count,
                                                     int *queue) {
  __shared__ int cnt[5], base[5];
  const int tid = threadIdx.x;
  if (tid < 5) cnt[tid] = 0;
  if (tid == 0 && queue) *queue = 0;
  __syncthreads();
  auto slot_of = [](int r) { return r == 2 ? 0 : r == 3 ? 1 : r == 4 ? 2 : r == 6 ? 3 : 4; };
  for (int i = tid; i < count; i += blockDim.x) atomicAdd(&cnt[slot_of(rate[i])], 1);
  __syncthreads();
  if (tid == 0) {
    int at = 0;
    for (int q = 0; q < 5; q++) {
      base[q] = at;
      at += cnt[q];
    }
  }
  __syncthreads();
  for (int i = tid; i < count; i += blockDim.x) {
    const int q = slot_of(rate[i]);
    const int at = atomicAdd(&base[q], 1);
    t[at] = LdpcTask{off[i], off[i], q < 4 ? 3 - q : -1, 0, 0, i};
  }
}

// One codeword of code CI by one workgroup (LDPC_THREADS = 128 lanes, lane = check position), L the workgroup's
// posteriors, red a 4-int LDS scratch for the reductions.
template <int CI>
__device__ __forceinline__ void decode_one(const LdpcParams &p, const LdpcTask &tk, int16_t *L, int *red, int lane) {
  using S = tm::Shape<CI>;
  constexpr int N = S::N, K = S::K, NP = S::cb.np * Z, NT = N - NP, MB = S::MB;
  // ---- posteriors from the channel (ldpc.c:1107-1110; frame mode: map_in and the known zeros, lizard_phy.c:156-192,
  // :1823)
  if (p.mode == MODE_CODEWORD) {
    const int8_t *src = p.llr + tk.llr;
    if ((reinterpret_cast<uintptr_t>(src) & 3) == 0) {
      // four channel values a load, every load of the codeword in flight before the first store; K is a multiple of
      // 4, so a word is all data bits or all parity bits
      static_assert(K % 4 == 0 && N % 4 == 0, "whole words");
      constexpr int NW = N / 4, R = (NW + LDPC_THREADS - 1) / LDPC_THREADS;
      const uint32_t *s4 = reinterpret_cast<const uint32_t *>(src);
      const int ln = tm::fresh(lane);  // lane-only addresses stay here, not hoisted out of the task loop
      uint32_t w[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = ln + LDPC_THREADS * r;
        w[r] = (NW % LDPC_THREADS == 0 || i < NW) ? s4[i] : 0u;
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = ln + LDPC_THREADS * r;
        if (NW % LDPC_THREADS != 0 && i >= NW) break;
        const int16_t x0 = static_cast<int8_t>(w[r]), x1 = static_cast<int8_t>(w[r] >> 8),
                      x2 = static_cast<int8_t>(w[r] >> 16), x3 = static_cast<int8_t>(w[r] >> 24);
        if (4 * i < K) {
          L[4 * i] = x0;
          L[4 * i + 1] = x1;
          L[4 * i + 2] = x2;
          L[4 * i + 3] = x3;
        } else {
          const int t = 4 * i - K;
          L[K + (t % MB) * Z + t / MB] = x0;
          L[K + ((t + 1) % MB) * Z + (t + 1) / MB] = x1;
          L[K + ((t + 2) % MB) * Z + (t + 2) / MB] = x2;
          L[K + ((t + 3) % MB) * Z + (t + 3) / MB] = x3;
        }
      }
    } else {
      for (int v = lane; v < K; v += LDPC_THREADS) L[v] = src[v];
      for (int v = lane; v < N - K; v += LDPC_THREADS) L[K + (v % MB) * Z + v / MB] = src[K + v];
    }
  } else {
    const int8_t *src = p.llr + tk.llr;
    const uint8_t *wh = p.white + tk.slot0;
    const int16_t *s2p = p.s2p[CI];
    for (int v = K - NP + lane; v < K; v += LDPC_THREADS) L[v] = v >= INFO_BITS ? 127 : 0;
    if (NT % 4 == 0 && ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(wh)) & 3) == 0 &&
        (reinterpret_cast<uintptr_t>(s2p) & 7) == 0) {
      // four slots a load from each of the soft values, the whitening and the bit map, all in flight at once
      constexpr int NW = NT / 4, R = (NW + LDPC_THREADS - 1) / LDPC_THREADS;
      const int ln = tm::fresh(lane);
      uint32_t xs[R], ws[R];
      uint2 ps[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = ln + LDPC_THREADS * r;
        const bool in = NW % LDPC_THREADS == 0 || i < NW;
        xs[r] = in ? reinterpret_cast<const uint32_t *>(src)[i] : 0u;
        ws[r] = in ? reinterpret_cast<const uint32_t *>(wh)[i] : 0u;
        ps[r] = in ? reinterpret_cast<const uint2 *>(s2p)[i] : make_uint2(0, 0);
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = ln + LDPC_THREADS * r;
        if (NW % LDPC_THREADS != 0 && i >= NW) break;
#pragma unroll
        for (int b = 0; b < 4; b++) {
          const int x = static_cast<int8_t>(xs[r] >> (8 * b));
          const uint32_t pw = b < 2 ? ps[r].x : ps[r].y;
          const int at = static_cast<int16_t>(pw >> (16 * (b & 1)));
          L[at] = static_cast<int16_t>(static_cast<int8_t>(((ws[r] >> (8 * b)) & 0xFF) ? -x : x));
        }
      }
    } else {
      for (int s = lane; s < NT; s += LDPC_THREADS) {
        const int x = src[s];
        L[s2p[s]] = static_cast<int16_t>(static_cast<int8_t>(wh[s] ? -x : x));
      }
    }
    for (int v = INFO_BITS + lane; v < K; v += LDPC_THREADS) L[v] = 127;
  }
  tm::Msgs<CI> M;
  tm::clear_msgs(M);
  __syncthreads();
  // ---- iterations: block rows in order on even iterations, reversed on odd; syndrome after each (ldpc.c:1116-1149)
  int used = -1, first = 0;
  for (int it = 0; it < p.max_iter; it++) {
    for (int step = 0; step < MB; step++) {
      const int r = (it & 1) ? MB - 1 - step : step;
      tm::layer_at<CI>(r, L, M, lane);
      __syncthreads();
    }
    // two alternating reduction slots: a slot is rewritten only after a later barrier has passed every read of it
    // the count where the stall rule reads it, else only whether every check holds (early out per block row)
    const bool count = p.stall_it > 0 && (it == 0 || it == p.stall_it - 1);
    const int bad = group_sum(count ? tm::unsat<CI>(L, lane) : tm::unsat_any<CI>(L, lane), red + 2 * (it & 1), lane);
    const int v = stop_rule(it, bad, p.stall_it, p.stall_ratio, &first);
    if (v) {
      used = v;
      break;
    }
  }
  // ---- outputs: hard decisions, CRC (lizard_phy.c:1828-1868), traces
  if (p.bits && tk.bits >= 0) {
    for (int v = lane; v < K; v += LDPC_THREADS) p.bits[tk.bits + v] = L[v] < 0;
    for (int v = lane; v < N - K; v += LDPC_THREADS) p.bits[tk.bits + K + v] = L[K + (v % MB) * Z + v / MB] < 0;
  }
  if (p.flags & 2) {
    for (int v = lane; v < N; v += LDPC_THREADS) p.post[tk.bits + v] = L[v];
    tm::dump_msgs<CI>(M, p.msg + static_cast<int64_t>(tk.out) * (MAX_SLOTS_TOTAL * Z), lane);
  }
  if (p.flags & 1) {
    // the CRC as the XOR of one table word per set bit: every table read of the lane in flight at once
    constexpr int RB = (BLOCK_BYTES + LDPC_THREADS - 1) / LDPC_THREADS;
    const int ln = tm::fresh(lane);
    uint32_t tab[RB][8];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int j = ln + LDPC_THREADS * r < BLOCK_BYTES ? ln + LDPC_THREADS * r : BLOCK_BYTES - 1;
#pragma unroll
      for (int q = 0; q < 8; q++) tab[r][q] = p.crc_tab[8 * j + q];
    }
    uint32_t x = 0;
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int j = ln + LDPC_THREADS * r;
      const bool in = j < BLOCK_BYTES;
      const int jj = in ? j : BLOCK_BYTES - 1;
      int b = 0;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int bit = L[8 * jj + q] < 0;
        b = (b << 1) | bit;
        x ^= (in && bit) ? tab[r][q] : 0u;
      }
      if (in) p.bytes[static_cast<int64_t>(tk.out) * BLOCK_BYTES + j] = static_cast<uint8_t>(b);
    }
    x = group_xor(x, reinterpret_cast<uint32_t *>(red) + 4, lane);
    if (lane == 0) {
      uint32_t rx = 0;
      for (int q = 0; q < 32; q++) rx = (rx << 1) | (L[8 * BLOCK_BYTES + q] < 0);
      p.crc_ok[tk.out] = used > 0 && (p.crc0 ^ x) == rx;
    }
  }
  if (lane == 0) p.its[tk.out] = static_cast<int8_t>(used);
}

// The frame-mode decline gate (lizard_phy.c:1782-1787, :1818): the block's estimate from its sub-channels', true
// where the block is not to be decoded.
template <int CI>
__device__ __forceinline__ bool declined(const LdpcParams &p, const LdpcTask &tk, int *red, int lane) {
  using S = tm::Shape<CI>;
  constexpr int K = S::K, NT = S::N - S::cb.np * Z, NB = NT + (CI == 0 ? 16 : CI == 1 ? 32 : 0);
  int d = 0;
  if (lane == 0) {
    const int s0 = tk.slot0;
    double info = 0;
    for (int j = s0 / 640; j * 640 < s0 + NB; j++) {
      const int lo = s0 > 640 * j ? s0 : 640 * j, hi = s0 + NB < 640 * (j + 1) ? s0 + NB : 640 * (j + 1);
      info += p.est[tk.est + j] * (hi - lo);
    }
    const float be = static_cast<float>(info / NB);
    p.blk_est[tk.out] = be;
    d = be < 0.2f * static_cast<float>(K) / static_cast<float>(NT);
    if (d) {
      p.its[tk.out] = -2;
      if (p.crc_ok) p.crc_ok[tk.out] = 0;
    }
  }
  if (lane == 0) red[6] = d;
  __syncthreads();
  return red[6] != 0;
}

template <int CI>
__device__ __forceinline__ void run_task(const LdpcParams &p, const LdpcTask &tk, int16_t *L, int *red, int lane) {
  if (p.mode == MODE_FRAME && declined<CI>(p, tk, red, lane)) return;
  decode_one<CI>(p, tk, L, red, lane);
}

__global__ __launch_bounds__(LDPC_THREADS) void k_ldpc(LdpcParams p) {
  __shared__ int16_t L[MAX_N];
  __shared__ int red[8];
  __shared__ int s_task;
  const int lane = threadIdx.x;
  for (int round = 0;; round++) {
    if (lane == 0) s_task = p.queue ? atomicAdd(p.queue, 1) : static_cast<int>(blockIdx.x);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_task);
    if (t >= p.count || (!p.queue && round > 0)) break;
    const LdpcTask tk = p.tasks[t];
    switch (__builtin_amdgcn_readfirstlane(tk.code)) {
      case 0: run_task<0>(p, tk, L, red, lane); break;
      case 1: run_task<1>(p, tk, L, red, lane); break;
      case 2: run_task<2>(p, tk, L, red, lane); break;
      case 3: run_task<3>(p, tk, L, red, lane); break;
      default:
        if (lane == 0) p.its[tk.out] = 0;
    }
    // every lane is done with this codeword (and has read s_task and red) before lane 0 takes the next
    __syncthreads();
    if (!p.queue) break;
  }
}

// ---- the sender's bit side (lizard_phy.c:1087-1150): one workgroup a block
// The block's 3808 bits MSB first and its CRC-32C's 32 (the XOR of one table word per set bit), zeros to k (the 7/8
// code's known zeros); parity bit c (check c = MB i + r, lane i) = the XOR of every syndrome up to c (the staircase,
// ldpc.c ldpc_encode): per lane the running XOR over its rows, then an exclusive prefix XOR across the lanes; the
// block's slots through the bit map, the pilot bit past nt, whitened.
template <int CI>
__device__ __forceinline__ void encode_slots(const SlotTables &t, const uint8_t *src, uint8_t *slots, int s0, int nb,
                                             int pilot, uint8_t *cw, uint32_t *red, int lane) {
  using S = tm::Shape<CI>;
  constexpr int K = S::K, MB = S::MB, NT = S::N - S::cb.np * Z;
  static_assert(MB <= 32, "a lane's running XOR in one word");
  uint32_t x = 0;
  for (int j = lane; j < BLOCK_BYTES; j += LDPC_THREADS) {
    const int b = src[j];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int bit = (b >> (7 - q)) & 1;
      cw[8 * j + q] = static_cast<uint8_t>(bit);
      x ^= bit ? t.crc_tab[8 * j + q] : 0u;
    }
  }
  for (int v = INFO_BITS + lane; v < K; v += LDPC_THREADS) cw[v] = 0;
  const uint32_t crc = t.crc0 ^ group_xor(x, red, lane);
  if (lane < 32) cw[8 * BLOCK_BYTES + lane] = static_cast<uint8_t>((crc >> (31 - lane)) & 1);
  __syncthreads();
  int run = 0;
  uint32_t loc = 0;
  tm::static_for<0, MB>([&](auto r) {
    constexpr int R = r, D0 = S::d0(R), ND = S::nd(R);
    int syn = 0;
    tm::static_for<0, ND>([&](auto e) { syn ^= cw[S::col(D0 + e) * Z + ((lane - S::shift(D0 + e)) & (Z - 1))]; });
    run ^= syn;
    loc |= static_cast<uint32_t>(run) << R;
  });
  int incl = run;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if ((lane & 63) >= o) incl ^= y;
  }
  if ((lane & 63) == 63) red[2 + (lane >> 6)] = static_cast<uint32_t>(incl);
  __syncthreads();
  const int pre = (incl ^ run) ^ (lane >= 64 ? static_cast<int>(red[2]) : 0);
#pragma unroll
  for (int r = 0; r < MB; r++) cw[K + MB * lane + r] = static_cast<uint8_t>(pre ^ ((loc >> r) & 1));
  __syncthreads();
  const int16_t *s2p = t.s2p[CI];
  for (int u = lane; u < nb; u += LDPC_THREADS) {
    int bit = pilot;
    if (u < NT) {
      const int pi = s2p[u];
      bit = cw[pi < K ? pi : K + ((pi - K) % Z) * MB + (pi - K) / Z];
    }
    slots[s0 + u] = static_cast<uint8_t>(bit ^ t.white[s0 + u]);
  }
}

// Grid (blocks + 1, frames): workgroup b < blocks encodes block b of frame f; workgroup blocks writes the padding
// past the last block (the whitening alone). blk: first slot, nt, slots, code index per block.
__global__ __launch_bounds__(LDPC_THREADS) void k_frame_slots(SlotTables t, const int4 *blk, int blocks, int total,
                                                              const uint8_t *blocks_all, int cnt, uint8_t *slots_all) {
  __shared__ uint8_t cw[MAX_N];
  __shared__ uint32_t red[4];
  const int b = blockIdx.x, f = blockIdx.y, lane = threadIdx.x;
  uint8_t *slots = slots_all + static_cast<size_t>(f) * total;
  if (b == blocks) {
    const int4 last = blk[blocks - 1];
    for (int u = last.x + last.z + lane; u < total; u += LDPC_THREADS) slots[u] = t.white[u];
    return;
  }
  const int4 bk = blk[b];
  const uint8_t *src = blocks_all + (static_cast<size_t>(f) * blocks + b) * BLOCK_BYTES;
  const int pilot = (cnt >> (b & 1)) & 1;
  switch (bk.w) {
    case 0: encode_slots<0>(t, src, slots, bk.x, bk.z, pilot, cw, red, lane); break;
    case 1: encode_slots<1>(t, src, slots, bk.x, bk.z, pilot, cw, red, lane); break;
    case 2: encode_slots<2>(t, src, slots, bk.x, bk.z, pilot, cw, red, lane); break;
    default: encode_slots<3>(t, src, slots, bk.x, bk.z, pilot, cw, red, lane); break;
  }
}

void LdpcEngine::frame_slots(const int4 *d_blk, int blocks, int total, const uint8_t *d_blocks, int cnt,
                             uint8_t *d_slots, int frames) {
  ensure_frame_tables(total);
  SlotTables t{};
  for (int ci = 0; ci < NCODES; ci++) t.s2p[ci] = s2p[ci].as<int16_t>();
  t.white = white.as<uint8_t>();
  t.crc_tab = crc_tab.as<uint32_t>();
  t.crc0 = crc0;
  hipLaunchKernelGGL(k_frame_slots, dim3(blocks + 1, frames), dim3(LDPC_THREADS), 0, stream, t, d_blk, blocks, total,
                     d_blocks, cnt, d_slots);
  hip_check(hipGetLastError(), "frame slots launch");
}

// The fixed work area: task descriptors then the queue counter, in device memory that the code object holds for the
// process's life (one copy per GPU, at its own address on each).
constexpr size_t SCRATCH_BYTES = SCRATCH_TASKS * sizeof(LdpcTask) + 256;
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

namespace {
std::mutex scratch_mu;
void *scratch_at[64];  // per device: the array's device address, looked up with that device current
}  // namespace

LdpcScratch ldpc_scratch(int dev) {
  if (dev < 0 || dev >= 64) throw Error(LIZ_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_at[dev]) {
    int cur = 0;
    hip_check(hipGetDevice(&cur), "hipGetDevice");
    if (cur != dev) throw Error(LIZ_E_INTERNAL, "the scratch area is looked up with its device current");
    void *ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (LDPC work area)");
    scratch_at[dev] = ptr;
  }
  uint8_t *base = static_cast<uint8_t *>(scratch_at[dev]);
  return LdpcScratch{reinterpret_cast<LdpcTask *>(base),
                     reinterpret_cast<int *>(base + SCRATCH_TASKS * sizeof(LdpcTask))};
}

LdpcEngine::LdpcEngine(int dev) : device(dev) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  (void)ldpc_scratch(device);  // the fixed work area's address, once per device, with the device current
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreate(&ev0), "hipEventCreate");
  hip_check(hipEventCreate(&ev1), "hipEventCreate");
  queue.reserve(sizeof(int) * 4);
  const auto &tab = crc_bit_table(&crc0);
  crc_tab.reserve(tab.size() * sizeof(uint32_t));
  copy_wait(crc_tab.p, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream, "CRC table");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  cus = prop.multiProcessorCount;
  hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&wg_per_cu, k_ldpc, LDPC_THREADS, 0), "occupancy");
  if (wg_per_cu < 1) wg_per_cu = 1;
}

LdpcEngine::~LdpcEngine() {
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

// The frame mode's bit maps (slot -> posterior index per code) and the whitening of the largest frame.
void LdpcEngine::ensure_frame_tables(int max_slots) {
  if (white.bytes >= static_cast<size_t>(max_slots)) return;
  const Format *top = format_for(64);
  if (!top) throw Error(LIZ_E_INTERNAL, "the top format could not be built");
  white.reserve(top->white.size());
  copy_wait(white.p, top->white.data(), top->white.size(), hipMemcpyHostToDevice, stream, "whitening");
  for (int ci = 0; ci < NCODES; ci++) {
    const auto &m = top->slot_to_post[ci];
    s2p[ci].reserve(m.size() * sizeof(int16_t));
    copy_wait(s2p[ci].p, m.data(), m.size() * sizeof(int16_t), hipMemcpyHostToDevice, stream, "bit map");
  }
}

void LdpcEngine::make_tasks(LdpcTask *d_tasks, const uint8_t *d_rate, const int64_t *d_off, int count, int *d_queue) {
  hipLaunchKernelGGL(k_make_tasks, dim3(1), dim3(1024), 0, stream, d_tasks, d_rate, d_off, count,
                     d_queue ? d_queue : queue.as<int>());
  hip_check(hipGetLastError(), "task launch");
}

void LdpcEngine::launch(const LdpcParams &p0, bool plain, bool timed, bool queue_zeroed, int per_cu) {
  if (p0.count <= 0) return;
  LdpcParams p = p0;
  if (timed && !armed) {
    hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
    armed = true;
  }
  if (plain) {
    p.queue = nullptr;
    // one workgroup per codeword, one launch each: the per-codeword path the batched one is measured against
    for (int t = 0; t < p0.count; t++) {
      LdpcParams q = p;
      q.tasks = p0.tasks + t;
      q.count = 1;
      hipLaunchKernelGGL(k_ldpc, dim3(1), dim3(LDPC_THREADS), 0, stream, q);
    }
  } else {
    p.queue = p0.queue ? p0.queue : queue.as<int>();
    if (!queue_zeroed) hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, stream, p.queue, 1);
    const int grid = std::min(p.count, cus * (per_cu > 0 ? std::min(per_cu, wg_per_cu) : wg_per_cu));
    hipLaunchKernelGGL(k_ldpc, dim3(grid), dim3(LDPC_THREADS), 0, stream, p);
  }
  hip_check(hipGetLastError(), "LDPC launch");
  if (timed) hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
}

// Waits for the stream; the time from the first timed launch since the last sync to the last one.
void LdpcEngine::wait(const char *what) {
  if (lost) throw Error(LIZ_E_STATE, "this handle's earlier work was taken as hung; free it and make another");
  try {
    wait_bounded(stream, what);
  } catch (const Error &e) {
    if (e.code == LIZ_E_STATE) lost = true;
    throw;
  }
}

void LdpcEngine::sync() {
  wait("decode");
  if (armed) {
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
    last_ms = ms;
    armed = false;
  }
}

}  // namespace lz

namespace lz {

__global__ __launch_bounds__(256) void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

float copy_probe(void *dst, const void *src, size_t bytes) {
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "hipEventCreate");
  hip_check(hipEventCreate(&b), "hipEventCreate");
  const size_t n = bytes / 16;
  hipDeviceProp_t prop;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  const unsigned grid = static_cast<unsigned>(prop.multiProcessorCount) * 8;
  hip_check(hipEventRecord(a, nullptr), "hipEventRecord");
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, nullptr, static_cast<uint4 *>(dst),
                     static_cast<const uint4 *>(src), n);
  hip_check(hipGetLastError(), "copy launch");
  hip_check(hipEventRecord(b, nullptr), "hipEventRecord");
  hip_check(hipEventSynchronize(b), "copy");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

}  // namespace lz
