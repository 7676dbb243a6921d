// Host-side plumbing shared by the library's translation units: the error model of the C ABI (a thread's last
// failure, no exception across the boundary; liblizard/api/api.hpp:15-44 and api.cpp:16-36 state the same contract)
// and HIP status checks.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/lizard_amd.h"

namespace lz {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

int fail(int code, const std::string &message);
void clear_error();
// Makes device `want` (or, when negative, the current one) current; throws LIZ_E_UNSUPPORTED without a HIP device.
int current_device(int want);

inline void need(bool ok, const char *what) {
  if (!ok) throw Error(LIZ_E_ARG, what);
}

// The hang bound (liblizard/core/wg/wg.h:48-55 HANG_MS): work not done this many ms after a wait began is taken as
// hung. LIZ_HANG_MS in the environment sets it (the reference's variable), lizx_hang_bound afterwards; 0: none.
double hang_bound_ms();
void set_hang_bound_ms(double ms);
// Waits for stream s as hipStreamSynchronize does, within the hang bound: LIZ_E_STATE "... taken as hung" past it.
void wait_bounded(hipStream_t s, const char *what);

inline void hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  const int code = e == hipErrorOutOfMemory ? LIZ_E_NOMEM
                   : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? LIZ_E_UNSUPPORTED
                                                                             : LIZ_E_INTERNAL;
  throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guard(F &&f) {
  try {
    clear_error();
    return f();
  } catch (const Error &e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc &) {
    return fail(LIZ_E_NOMEM, "out of host memory");
  } catch (const std::exception &e) {
    return fail(LIZ_E_INTERNAL, e.what());
  } catch (...) {
    return fail(LIZ_E_INTERNAL, "an unknown failure");
  }
}

template <class T, class F>
T *guard_new(F &&f) {
  T *out = nullptr;
  guard([&] {
    out = f();
    return 0;
  });
  return out;
}

// A copy ordered on stream s and waited for. The library makes no legacy-stream operation (hipMemcpy and its like):
// the HIP runtime fails those, and invalidates the capture, while any stream of the process captures a graph, and a
// decoder on another thread may be capturing its frame graph.
inline void copy_wait(void *dst, const void *src, size_t n, hipMemcpyKind kind, hipStream_t s, const char *what) {
  hip_check(hipMemcpyAsync(dst, src, n, kind, s), what);
  hip_check(hipStreamSynchronize(s), what);
}
// A non-blocking stream of the current device for the copies that belong to no handle's stream (plan tables).
hipStream_t aux_stream();

// A device buffer owned by RAII; grows, never shrinks.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    bytes = 0;
    hip_check(hipMalloc(&p, n), "hipMalloc");
    bytes = n;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

}  // namespace lz
