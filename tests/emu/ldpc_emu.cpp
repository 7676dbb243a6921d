// Test harness: the LDPC kernel's per-lane code (lizard_amd/csrc/ldpc_dev.hpp) walked on a CPU, lane by lane between
// the kernel's barriers, in codeword mode. It checks the kernel's indexing and stop logic without a GPU; the tests
// hold it to the oracle. Built with g++ by __graft_entry__.build().
#include <cstring>
#include <vector>

#include "../../lizard_amd/csrc/format.hpp"
#include "../../lizard_amd/csrc/ldpc_dev.hpp"

namespace lz {
void fill_dev_code_host(const Code &c, DevCode &d) {
  d = DevCode{};
  d.n = c.n; d.k = c.k; d.m = c.m; d.mb = c.mb; d.kb = c.kb; d.np = c.np; d.nt = c.nt; d.slots_total = c.slots_total;
  for (int r = 0; r <= c.mb; r++) d.lay_ptr[r] = c.lay_ptr[r];
  for (int r = 0; r < c.mb; r++) d.rbase[r] = c.lay_ptr[r] + 2 * r;
  for (size_t e = 0; e < c.lay_col.size(); e++) d.ent[e] = static_cast<uint16_t>(c.lay_col[e] << 7 | c.lay_shift[e]);
}
}  // namespace lz

extern "C" int emu_ldpc_decode(int rate, const int8_t *llr, uint8_t *out, int max_iter, int stall_it, float ratio,
                               int16_t *post, int8_t *msg) {
  using namespace lz;
  const int ci = code_index(rate);
  if (ci < 0) return -100;
  DevCode c;
  fill_dev_code_host(code(ci), c);
  std::vector<int16_t> L(MAX_N);
  std::vector<int8_t> R(MAX_SLOTS_TOTAL * Z, 0);
  for (int i = 0; i < 128; i++)
    for (int v = i; v < c.n; v += 128) L[dev_post_index(c, v)] = llr[v];
  int used = -1, first = 0;
  for (int it = 0; it < max_iter; it++) {
    for (int step = 0; step < c.mb; step++) {
      const int r = (it & 1) ? c.mb - 1 - step : step;
      for (int i = 0; i < 128; i++) layer_update(c, L.data(), R.data(), r, i);
    }
    int bad = 0;
    for (int i = 0; i < 128; i++) bad += lane_unsat(c, L.data(), i);
    const int v = stop_rule(it, bad, stall_it, ratio, &first);
    if (v) {
      used = v;
      break;
    }
  }
  for (int v = 0; v < c.n; v++) out[v] = L[dev_post_index(c, v)] < 0;
  if (post) std::memcpy(post, L.data(), c.n * sizeof(int16_t));
  if (msg) std::memcpy(msg, R.data(), static_cast<size_t>(c.slots_total) * Z);
  return used;
}
