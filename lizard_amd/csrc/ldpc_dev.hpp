// The LDPC stop rule, shared by the gfx950 kernel and the host-side lane emulation.
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#endif
#endif

namespace lz {

// After an iteration (liblizard/src/ldpc.c:1143-1148): the syndrome's unsatisfied count decides. With stall_it > 0
// the counts after the first iteration and after iteration stall_it are compared: still above stall_ratio of the
// first, the codeword is given up (-3). Returns the verdict (iterations used, or -3), or 0 to go on.
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
