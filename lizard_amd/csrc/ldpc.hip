// Layered min-sum LDPC + CRC-32C on gfx950: one codeword per 128-lane workgroup, a lane per check position of the
// current block row, the codeword's posteriors (int16) and messages (int8) in LDS, a block-row barrier between
// layers. A persistent grid pulls codewords off an atomic queue, so codewords of four rates with uneven iteration
// counts share one launch; the plain path launches one workgroup per codeword instead.
//
// Follows liblizard/src/ldpc.c:1080-1172 (ldpc_decode_stall) and, for frame mode, liblizard/src/lizard_phy.c:1775-1870
// (decline gate, bit map, known zeros, bit pack, CRC compare).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ldpc.hpp"

namespace lz {

__constant__ DevCode c_codes[NCODES];

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

__global__ __launch_bounds__(LDPC_THREADS) void k_ldpc(LdpcParams p) {
  __shared__ int16_t L[MAX_N];
  __shared__ int8_t R[MAX_SLOTS_TOTAL * Z];
  __shared__ int s_task, s_bad[2], s_decl;
  __shared__ uint32_t s_x[2];
  const int i = threadIdx.x, lane = i & 63, wave = i >> 6;
  if (i == 0) {
    s_bad[0] = 0;
    s_bad[1] = 0;
  }
  for (int round = 0;; round++) {
    if (i == 0) s_task = p.queue ? atomicAdd(p.queue, 1) : static_cast<int>(blockIdx.x);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_task);
    if (t >= p.count || (!p.queue && round > 0)) break;
    const LdpcTask tk = p.tasks[t];
    const int ci = __builtin_amdgcn_readfirstlane(tk.code);
    const DevCode &c = c_codes[ci];
    const int n = c.n, k = c.k;

    // ---- frame mode: the block's estimate and the decline gate (lizard_phy.c:1782-1787, :1818)
    if (p.mode == MODE_FRAME) {
      if (i == 0) {
        const int nb = c.nt + (ci == 0 ? 16 : ci == 1 ? 32 : 0);
        const int s0 = tk.slot0;
        double info = 0;
        for (int j = s0 / 640; j * 640 < s0 + nb; j++) {
          const int lo = s0 > 640 * j ? s0 : 640 * j, hi = s0 + nb < 640 * (j + 1) ? s0 + nb : 640 * (j + 1);
          info += p.est[tk.est + j] * (hi - lo);
        }
        const float be = static_cast<float>(info / nb);
        p.blk_est[tk.out] = be;
        s_decl = be < 0.2f * static_cast<float>(k) / static_cast<float>(c.nt);
      }
      __syncthreads();
      if (s_decl) {
        if (i == 0) {
          p.its[tk.out] = -2;
          if (p.crc_ok) p.crc_ok[tk.out] = 0;
        }
        __syncthreads();
        if (!p.queue) break;
        continue;
      }
    }

    // ---- posteriors from the channel, messages zero (ldpc.c:1107-1115; frame mode: map_in and the known zeros,
    // lizard_phy.c:156-192, :1823)
    if (p.mode == MODE_CODEWORD) {
      const int8_t *src = p.llr + tk.llr;
      for (int v = i; v < n; v += LDPC_THREADS) L[dev_post_index(c, v)] = src[v];
    } else {
      const int8_t *src = p.llr + tk.llr;
      const uint8_t *wh = p.white + tk.slot0;
      const int16_t *s2p = p.s2p[ci];
      for (int v = k - c.np + i; v < k; v += LDPC_THREADS) L[v] = 0;
      __syncthreads();
      for (int s = i; s < c.nt; s += LDPC_THREADS) {
        const int x = src[s];
        L[s2p[s]] = static_cast<int16_t>(static_cast<int8_t>(wh[s] ? -x : x));
      }
      __syncthreads();
      for (int v = INFO_BITS + i; v < k; v += LDPC_THREADS) L[v] = 127;
    }
    for (int s = i; s < c.slots_total * Z; s += LDPC_THREADS) R[s] = 0;
    __syncthreads();

    // ---- iterations: block rows in order on even iterations, reversed on odd; syndrome after each
    int used = -1, first = 0;
    for (int it = 0; it < p.max_iter; it++) {
      for (int step = 0; step < c.mb; step++) {
        const int r = (it & 1) ? c.mb - 1 - step : step;
        layer_update(c, L, R, r, i);
        __syncthreads();
      }
      const int part = wave_sum(lane_unsat(c, L, i));
      if (lane == 0) atomicAdd(&s_bad[it & 1], part);
      __syncthreads();
      const int bad = s_bad[it & 1];
      if (i == 0) s_bad[(it + 1) & 1] = 0;
      const int v = stop_rule(it, bad, p.stall_it, p.stall_ratio, &first);
      if (v) {
        used = v;
        break;
      }
    }

    // ---- outputs: hard decisions, CRC, trace
    if (p.bits && tk.bits >= 0)
      for (int v = i; v < n; v += LDPC_THREADS) p.bits[tk.bits + v] = L[dev_post_index(c, v)] < 0;
    if (p.flags & 1) {
      uint32_t x = 0;
      for (int j = i; j < BLOCK_BYTES; j += LDPC_THREADS) {
        int b = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int bit = L[8 * j + q] < 0;
          b = (b << 1) | bit;
          x ^= bit ? p.crc_tab[8 * j + q] : 0u;
        }
        p.bytes[static_cast<int64_t>(tk.out) * BLOCK_BYTES + j] = static_cast<uint8_t>(b);
      }
      x = wave_xor(x);
      if (lane == 0) s_x[wave] = x;
      __syncthreads();
      if (i == 0) {
        uint32_t rx = 0;
        for (int q = 0; q < 32; q++) rx = (rx << 1) | (L[8 * BLOCK_BYTES + q] < 0);
        p.crc_ok[tk.out] = used > 0 && (p.crc0 ^ s_x[0] ^ s_x[1]) == rx;
      }
    }
    if (p.flags & 2) {
      for (int v = i; v < n; v += LDPC_THREADS) p.post[tk.bits + v] = L[v];
      int8_t *m = p.msg + static_cast<int64_t>(tk.out) * (MAX_SLOTS_TOTAL * Z);
      for (int s = i; s < c.slots_total * Z; s += LDPC_THREADS) m[s] = R[s];
    }
    if (i == 0) p.its[tk.out] = static_cast<int8_t>(used);
    __syncthreads();
    // every lane has read the last count: the two count slots start the next codeword at zero
    if (i == 0) {
      s_bad[0] = 0;
      s_bad[1] = 0;
    }
    if (!p.queue) break;
  }
}

void fill_dev_code(const Code &c, DevCode &d) {
  d = DevCode{};
  d.n = c.n;
  d.k = c.k;
  d.m = c.m;
  d.mb = c.mb;
  d.kb = c.kb;
  d.np = c.np;
  d.nt = c.nt;
  d.slots_total = c.slots_total;
  for (int r = 0; r <= c.mb; r++) d.lay_ptr[r] = c.lay_ptr[r];
  for (int r = 0; r < c.mb; r++) d.rbase[r] = c.lay_ptr[r] + 2 * r;
  for (size_t e = 0; e < c.lay_col.size(); e++) d.ent[e] = static_cast<uint16_t>(c.lay_col[e] << 7 | c.lay_shift[e]);
}

LdpcEngine::LdpcEngine(int dev) : device(dev) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  DevCode dc[NCODES];
  for (int ci = 0; ci < NCODES; ci++) fill_dev_code(code(ci), dc[ci]);
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_codes), dc, sizeof dc), "code tables to the device");
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreate(&ev0), "hipEventCreate");
  hip_check(hipEventCreate(&ev1), "hipEventCreate");
  queue.reserve(sizeof(int) * 4);
  const auto &tab = crc_bit_table(&crc0);
  crc_tab.reserve(tab.size() * sizeof(uint32_t));
  hip_check(hipMemcpy(crc_tab.p, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "CRC table");
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
  hip_check(hipMemcpy(white.p, top->white.data(), top->white.size(), hipMemcpyHostToDevice), "whitening");
  for (int ci = 0; ci < NCODES; ci++) {
    const auto &m = top->slot_to_post[ci];
    s2p[ci].reserve(m.size() * sizeof(int16_t));
    hip_check(hipMemcpy(s2p[ci].p, m.data(), m.size() * sizeof(int16_t), hipMemcpyHostToDevice), "bit map");
  }
}

void LdpcEngine::launch(const LdpcParams &p0, bool plain, bool timed) {
  if (p0.count <= 0) return;
  LdpcParams p = p0;
  if (timed) hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
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
    p.queue = queue.as<int>();
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, stream, p.queue, 1);
    const int grid = std::min(p.count, cus * wg_per_cu);
    hipLaunchKernelGGL(k_ldpc, dim3(grid), dim3(LDPC_THREADS), 0, stream, p);
  }
  hip_check(hipGetLastError(), "LDPC launch");
  if (timed) hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
}

}  // namespace lz
