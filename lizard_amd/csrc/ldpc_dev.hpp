// The layered min-sum decoder's per-lane arithmetic, shared by the gfx950 kernel and the host-side lane emulation
// (tests build the emulation to walk a workgroup lane by lane on a CPU). A workgroup decodes one codeword; lane i of
// its 128 owns check position i of the current layer (block row), so the 128 checks of a layer run side by side and
// touch disjoint posteriors (a circulant puts each data bit in exactly one check of a row, and each staircase bit is
// shared only with the check before or after it, in another row).
//
// Integer semantics are the reference decoder's scalar form (liblizard/src/ldpc.c:171-191, :589-617, :1080-1172;
// research/SPEC.md 8.5): int16 posteriors clamped to +-8191, int8 messages, magnitude (min * 13) >> 4 capped at 127,
// the first least magnitude wins ties, zero counts positive, check 0's absent parity edge read as 8191 with its
// message zeroed before the layer.
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace lz {

constexpr int DZ = 128;
constexpr int DMAX_SLOTS = 23;   // most slots in a block row of any code (7/8: 21 circulants + 2)
constexpr int DMAX_ROWS = 30;
constexpr int DMAX_DATA = 150;   // most data circulants of any code (1/2)

// One code as the kernel reads it (constant memory).
struct DevCode {
  int32_t n, k, m, mb, kb, np, nt, slots_total;
  int32_t lay_ptr[DMAX_ROWS + 1];  // first data circulant of each block row
  int32_t rbase[DMAX_ROWS];        // first message slot of each block row: lay_ptr[r] + 2 r
  uint16_t ent[DMAX_DATA];         // per data circulant: column << 7 | shift
};

__host__ __device__ inline int dev_post_index(const DevCode &c, int v) {
  if (v < c.k) return v;
  const int chk = v - c.k;
  return c.k + (chk % c.mb) * DZ + chk / c.mb;
}

// Lane i's check in block row r: read its slots' posteriors less its old messages, take the check's two least
// magnitudes and its sign, write the new messages and posteriors.
__host__ __device__ inline void layer_update(const DevCode &c, int16_t *L, int8_t *R, int r, int i) {
  const int d0 = c.lay_ptr[r], nd = c.lay_ptr[r + 1] - d0;
  int8_t *Rl = R + c.rbase[r] * DZ;
  const int16_t *Lp = L + c.k;
  int idx[DMAX_SLOTS], q[DMAX_SLOTS];
  int min1 = 32767, min2 = 32767, arg = -1, sgn = 0;
  const bool absent = r == 0 && i == 0;  // check 0 has no parity bit before it
#pragma unroll
  for (int e = 0; e < DMAX_SLOTS; e++) {
    if (e < nd + 2) {
      int post, msg;
      if (e < nd) {
        const int ent = c.ent[d0 + e];
        idx[e] = (ent >> 7) * DZ + ((i - (ent & 127)) & (DZ - 1));
        post = L[idx[e]];
        msg = Rl[e * DZ + i];
      } else if (e == nd) {
        idx[e] = r > 0 ? c.k + (r - 1) * DZ + i : c.k + (c.mb - 1) * DZ + i - 1;
        post = absent ? 8191 : L[idx[e]];
        msg = absent ? 0 : Rl[e * DZ + i];
      } else {
        idx[e] = c.k + r * DZ + i;
        post = Lp[r * DZ + i];
        msg = Rl[e * DZ + i];
      }
      const int v = post - msg;
      q[e] = v;
      const int a = v < 0 ? -v : v;
      sgn ^= v < 0;
      const bool lt = a < min1;
      min2 = lt ? min1 : (a < min2 ? a : min2);
      min1 = lt ? a : min1;
      arg = lt ? e : arg;
    }
  }
  int m1 = (min1 * 13) >> 4, m2 = (min2 * 13) >> 4;
  m1 = m1 > 127 ? 127 : m1;
  m2 = m2 > 127 ? 127 : m2;
#pragma unroll
  for (int e = 0; e < DMAX_SLOTS; e++) {
    if (e < nd + 2) {
      const int mag = e == arg ? m2 : m1;
      const int msg = (sgn ^ (q[e] < 0)) ? -mag : mag;
      Rl[e * DZ + i] = static_cast<int8_t>(msg);
      int t = q[e] + msg;
      t = t > 8191 ? 8191 : (t < -8191 ? -8191 : t);
      if (!(e == nd && absent)) L[idx[e]] = static_cast<int16_t>(t);
    }
  }
}

// Lane i's unsatisfied checks over every block row: the parity of the signs of each check's posteriors (the absent
// edge of check 0 counts as a 0).
__host__ __device__ inline int lane_unsat(const DevCode &c, const int16_t *L, int i) {
  int bad = 0;
  for (int r = 0; r < c.mb; r++) {
    const int d0 = c.lay_ptr[r], nd = c.lay_ptr[r + 1] - d0;
    int par = 0;
    for (int e = 0; e < nd; e++) {
      const int ent = c.ent[d0 + e];
      par ^= L[(ent >> 7) * DZ + ((i - (ent & 127)) & (DZ - 1))] < 0;
    }
    if (r > 0) par ^= L[c.k + (r - 1) * DZ + i] < 0;
    else if (i > 0) par ^= L[c.k + (c.mb - 1) * DZ + i - 1] < 0;
    par ^= L[c.k + r * DZ + i] < 0;
    bad += par;
  }
  return bad;
}

// The stop rule after an iteration (ldpc.c:1143-1148): returns the final verdict (iterations, or -3) or 0 to go on.
__host__ __device__ inline int stop_rule(int it, int bad, int stall_it, float stall_ratio, int *first) {
  if (stall_it > 0 && (it == 0 || it == stall_it - 1)) {
    if (!bad) return it + 1;
    if (it == 0) *first = bad;
    else if (static_cast<float>(bad) > stall_ratio * static_cast<float>(*first)) return -3;
    return 0;
  }
  return bad ? 0 : it + 1;
}

}  // namespace lz
