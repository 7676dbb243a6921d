// Test harness: the LDPC kernel's per-code lane code (lizard_amd/csrc/ldpc_tmpl.hpp) walked on a CPU, the 128 lanes of
// the workgroup one after another between the kernel's barriers, in codeword mode. It checks the kernel's indexing, message
// registers and stop logic without a GPU; the tests hold it to the oracle. Built with g++ by __graft_entry__.build().
#include <cstring>
#include <vector>

#include "../../lizard_amd/csrc/ldpc_dev.hpp"
#include "../../lizard_amd/csrc/ldpc_tmpl.hpp"

template <int CI>
static int emu(const int8_t *llr, uint8_t *out, int max_iter, int stall_it, float ratio, int16_t *post, int8_t *msg) {
  using S = lz::tm::Shape<CI>;
  std::vector<int16_t> L(S::N);
  std::vector<lz::tm::Msgs<CI>> M(128);
  for (int v = 0; v < S::K; v++) L[v] = llr[v];
  for (int v = 0; v < S::N - S::K; v++) L[S::K + (v % S::MB) * 128 + v / S::MB] = llr[S::K + v];
  for (auto &m : M) lz::tm::clear_msgs(m);
  int used = -1, first = 0;
  for (int it = 0; it < max_iter; it++) {
    for (int step = 0; step < S::MB; step++) {
      const int r = (it & 1) ? S::MB - 1 - step : step;
      for (int lane = 0; lane < 128; lane++) lz::tm::layer_at<CI>(r, L.data(), M[lane], lane);
    }
    // as the kernel: the count where the stall rule reads it, else only whether every check holds
    const bool count = stall_it > 0 && (it == 0 || it == stall_it - 1);
    int bad = 0;
    for (int lane = 0; lane < 128; lane++)
      bad += count ? lz::tm::unsat<CI>(L.data(), lane) : lz::tm::unsat_any<CI>(L.data(), lane);
    const int v = lz::stop_rule(it, bad, stall_it, ratio, &first);
    if (v) {
      used = v;
      break;
    }
  }
  for (int v = 0; v < S::K; v++) out[v] = L[v] < 0;
  for (int v = 0; v < S::N - S::K; v++) out[S::K + v] = L[S::K + (v % S::MB) * 128 + v / S::MB] < 0;
  if (post) std::memcpy(post, L.data(), S::N * sizeof(int16_t));
  if (msg)
    for (int lane = 0; lane < 128; lane++) lz::tm::dump_msgs<CI>(M[lane], msg, lane);
  return used;
}

extern "C" int emu_ldpc_decode(int rate, const int8_t *llr, uint8_t *out, int max_iter, int stall_it, float ratio,
                               int16_t *post, int8_t *msg) {
  switch (rate) {
    case 6: return emu<0>(llr, out, max_iter, stall_it, ratio, post, msg);
    case 4: return emu<1>(llr, out, max_iter, stall_it, ratio, post, msg);
    case 3: return emu<2>(llr, out, max_iter, stall_it, ratio, post, msg);
    case 2: return emu<3>(llr, out, max_iter, stall_it, ratio, post, msg);
    default: return -100;
  }
}
