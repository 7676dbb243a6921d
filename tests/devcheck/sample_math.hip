// Device check of the warp sampler's cheap float forms (lizard_amd/csrc/sample_math.hpp) against the plain operators,
// over every input the check names: grey over the 256 levels, recip_rn over all 2^32 float bit patterns, frac_pos over
// every float in [0, 2^31). Built beside the tests (test infrastructure, not the product); run by
// tests/test_warp_gpu.py.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "../../lizard_amd/csrc/sample_math.hpp"

namespace {

__global__ void k_check(uint32_t base, unsigned long long *cnt, uint32_t *first) {
  const uint32_t u = base + blockIdx.x * 256u + threadIdx.x;
  float x;
  memcpy(&x, &u, 4);
  const float q = __fdiv_rn(1.0f, x), r = lz::recip_rn(x);
  // NaN quotients (x NaN) compare as patterns too: both sides divide there
  if (__float_as_uint(r) != __float_as_uint(q)) {
    const unsigned long long k = atomicAdd(&cnt[0], 1ull);
    if (k < 4) first[k] = u;
  }
  if (u <= 0x4effffffu) {
    const float a = lz::frac_pos(x), b = __fsub_rn(x, static_cast<float>(static_cast<int>(x)));
    if (__float_as_uint(a) != __float_as_uint(b)) {
      const unsigned long long k = atomicAdd(&cnt[1], 1ull);
      if (k < 4) first[4 + k] = u;
    }
  }
  if (u < 256u && __float_as_uint(lz::grey(u)) != __float_as_uint(__fdiv_rn(static_cast<float>(u), 255.0f))) {
    const unsigned long long k = atomicAdd(&cnt[2], 1ull);
    if (k < 4) first[8 + k] = u;
  }
}

}  // namespace

// counts[3]: mismatches of recip_rn, frac_pos, grey; first[12]: up to four inputs of each. 0, or -1 on a HIP error.
extern "C" int lz_check_sample_math(unsigned long long *counts, uint32_t *first) {
  unsigned long long *d_cnt = nullptr;
  uint32_t *d_first = nullptr;
  if (hipMalloc(&d_cnt, 3 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMalloc(&d_first, 12 * sizeof(uint32_t)) != hipSuccess) return -1;
  int rc = 0;
  if (hipMemset(d_cnt, 0, 3 * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(d_first, 0, 12 * sizeof(uint32_t)) != hipSuccess)
    rc = -1;
  const uint32_t chunk = 1u << 28;  // 2^20 workgroups of 256 a launch, 16 launches
  for (uint64_t b = 0; rc == 0 && b < (1ull << 32); b += chunk) {
    hipLaunchKernelGGL(k_check, dim3(chunk / 256), dim3(256), 0, 0, static_cast<uint32_t>(b), d_cnt, d_first);
    if (hipDeviceSynchronize() != hipSuccess) rc = -1;
  }
  if (rc == 0 && (hipMemcpy(counts, d_cnt, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess ||
                  hipMemcpy(first, d_first, 12 * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess))
    rc = -1;
  (void)hipFree(d_cnt);
  (void)hipFree(d_first);
  return rc;
}
