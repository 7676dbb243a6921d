// The finder's pixel work on the GPU (the kernels finder.hpp lists).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "finder.hpp"

namespace lz {
namespace {

constexpr float TWO_PI = 6.283185307179586f;

// Sobel gradient at (X, Y) in grey levels a pixel (the 3 x 3 kernels over 8), 0 on the frame's outermost pixels.
__device__ __forceinline__ void sobel(const uint8_t *img, int w, int h, int stride, int X, int Y, float *gx, float *gy) {
  if (X < 1 || Y < 1 || X > w - 2 || Y > h - 2) {
    *gx = *gy = 0.0f;
    return;
  }
  const uint8_t *p = img + static_cast<size_t>(Y) * stride + X;
  const float a = p[-stride - 1], b = p[-stride], c = p[-stride + 1];
  const float d = p[-1], f = p[1];
  const float g = p[stride - 1], hh = p[stride], k = p[stride + 1];
  *gx = ((c + 2.0f * f + k) - (a + 2.0f * d + g)) * 0.125f;
  *gy = ((g + 2.0f * hh + k) - (a + 2.0f * b + c)) * 0.125f;
}

// One 16 x 16 tile a workgroup, its gradients with a one-pixel halo in LDS for the thinning.
__global__ __launch_bounds__(256) void k_find_vote(const uint8_t *img, int w, int h, int stride, float thr,
                                                   const float2 *trig, int nr, float *acc) {
  __shared__ float sm[18][18], sgx[18][18], sgy[18][18];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int x0 = blockIdx.x * 16 - 1, y0 = blockIdx.y * 16 - 1;
  for (int i = threadIdx.x; i < 18 * 18; i += 256) {
    const int ly = i / 18, lx = i - ly * 18;
    float gx, gy;
    sobel(img, w, h, stride, x0 + lx, y0 + ly, &gx, &gy);
    sgx[ly][lx] = gx;
    sgy[ly][lx] = gy;
    sm[ly][lx] = sqrtf(gx * gx + gy * gy);
  }
  __syncthreads();
  const int X = x0 + 1 + tx, Y = y0 + 1 + ty;
  if (X >= w || Y >= h) return;
  const float m = sm[ty + 1][tx + 1];
  if (m < thr) return;
  const float gx = sgx[ty + 1][tx + 1], gy = sgy[ty + 1][tx + 1];
  // the neighbour along the gradient, to the nearest of the eight directions
  const float ax = fabsf(gx), ay = fabsf(gy);
  int dx, dy;
  if (ay <= 0.41421356f * ax) {
    dx = 1;
    dy = 0;
  } else if (ax <= 0.41421356f * ay) {
    dx = 0;
    dy = 1;
  } else {
    dx = 1;
    dy = (gx > 0) == (gy > 0) ? 1 : -1;
  }
  if (!(m > sm[ty + 1 + dy][tx + 1 + dx] && m >= sm[ty + 1 - dy][tx + 1 - dx])) return;
  const float b = atan2f(gy, gx) * (FIND_THETA / TWO_PI);
  const int bc = static_cast<int>(floorf(b + 0.5f));
  const float px = X + 0.5f - 0.5f * w, py = Y + 0.5f - 0.5f * h, R = 0.5f * (nr - 1);
  for (int db = -1; db <= 1; db++) {
    int t = bc + db;
    t = t < 0 ? t + FIND_THETA : t >= FIND_THETA ? t - FIND_THETA : t;
    const float2 cs = trig[t];
    const int r = static_cast<int>(floorf(cs.x * px + cs.y * py + R + 0.5f));
    if (r >= 0 && r < nr) atomicAdd(&acc[static_cast<size_t>(t) * nr + r], db ? 0.5f * m : m);
  }
}

__global__ __launch_bounds__(256) void k_find_zero(float *acc, size_t count, unsigned *d_max, int *d_count) {
  const size_t stride = static_cast<size_t>(gridDim.x) * 256;
  for (size_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += stride) acc[i] = 0.0f;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *d_max = 0u;
    *d_count = 0;
  }
}

__global__ __launch_bounds__(256) void k_find_max(const float *acc, size_t count, unsigned *d_max) {
  __shared__ float wmax[4];
  const size_t stride = static_cast<size_t>(gridDim.x) * 256;
  float v = 0.0f;
  for (size_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += stride) v = fmaxf(v, acc[i]);
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  // one atomic a workgroup; non-negative floats order as their bit patterns
  if (threadIdx.x == 0) atomicMax(d_max, __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
}

__global__ __launch_bounds__(256) void k_find_peaks(const float *acc, int nr, float frac, const unsigned *d_max,
                                                    FindPeak *peaks, int *d_count) {
  const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= static_cast<size_t>(FIND_THETA) * nr) return;
  const float v = acc[i], floor_v = frac * __uint_as_float(*d_max);
  if (v <= 0.0f || v < floor_v) return;
  const int t = static_cast<int>(i / nr), r = static_cast<int>(i - static_cast<size_t>(t) * nr);
  float score = 0.0f;
  for (int dt = -2; dt <= 2; dt++) {
    int tt = t + dt;
    tt = tt < 0 ? tt + FIND_THETA : tt >= FIND_THETA ? tt - FIND_THETA : tt;
    for (int dr = -3; dr <= 3; dr++) {
      const int rr = r + dr;
      if (rr < 0 || rr >= nr || (dt == 0 && dr == 0)) continue;
      const size_t j = static_cast<size_t>(tt) * nr + rr;
      const float u = acc[j];
      if (u > v || (u == v && j < i)) return;  // not the window's maximum (ties to the lowest index)
      if (dt >= -1 && dt <= 1 && dr >= -2 && dr <= 2) score += u;
    }
  }
  const int k = atomicAdd(d_count, 1);
  if (k >= FIND_MAX_PEAKS) return;
  FindPeak p;
  p.theta = t * (TWO_PI / FIND_THETA);
  p.rho = static_cast<float>(r) - 0.5f * (nr - 1);
  p.votes = v;
  p.score = score + v;
  peaks[k] = p;
}

// bilinear grey read, pixel centres at half integers, clamped to the frame
__device__ __forceinline__ float grey_at(const uint8_t *img, int w, int h, int stride, float x, float y) {
  x -= 0.5f;
  y -= 0.5f;
  x = fminf(fmaxf(x, 0.0f), w - 1.001f);
  y = fminf(fmaxf(y, 0.0f), h - 1.001f);
  const int xi = static_cast<int>(x), yi = static_cast<int>(y);
  const float fx = x - xi, fy = y - yi;
  const uint8_t *p = img + static_cast<size_t>(yi) * stride + xi;
  const float a = p[0], b = p[1], c = p[stride], d = p[stride + 1];
  return (a + (b - a) * fx) * (1 - fy) + (c + (d - c) * fx) * fy;
}

// Per probe: the profile along the normal every quarter pixel over +-reach, its central differences, and the rising
// edge (a local maximum of the difference at 0.3 of the window's largest or more) nearest the probe's point.
__global__ __launch_bounds__(256) void k_find_probe(const uint8_t *const *imgs, int w, int h, int stride,
                                                    const FindProbe *probes, int n, FindEdge *out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const FindProbe P = probes[k];
  const uint8_t *img = imgs[P.frame];
  const float reach = fminf(fmaxf(P.reach, 0.25f), 12.0f);
  const float step = 0.25f;
  const int ns = 2 * static_cast<int>(reach / step) + 1;
  auto val = [&](int i) {
    const float s = -reach + step * i;
    return grey_at(img, w, h, stride, P.x + s * P.nx, P.y + s * P.ny);
  };
  float dmax = 0.0f;
  {
    float a = val(0), b = val(1);
    for (int i = 1; i + 1 < ns; i++) {
      const float c = val(i + 1);
      dmax = fmaxf(dmax, c - a);
      a = b;
      b = c;
    }
  }
  FindEdge e{P.x, P.y, 0.0f, 0.0f};
  if (dmax > 1.0f) {
    // rolling window: v[i-2 .. i+2], differences d(i-1), d(i), d(i+1)
    float v0 = val(0), v1 = val(1), v2 = val(2), v3 = val(3);
    float best = 1e30f;
    for (int i = 2; i + 2 < ns; i++) {
      const float v4 = val(i + 2);
      const float dm = v2 - v0, d0 = v3 - v1, dp = v4 - v2;
      if (d0 > dm && d0 >= dp && d0 >= 0.3f * dmax) {
        const float den = dm - 2.0f * d0 + dp;
        float delta = den < 0.0f ? 0.5f * (dm - dp) / den : 0.0f;
        delta = fminf(fmaxf(delta, -0.5f), 0.5f);
        const float off = -reach + step * (i + delta);
        if (fabsf(off) < best) {
          best = fabsf(off);
          e.x = P.x + off * P.nx;
          e.y = P.y + off * P.ny;
          e.w = d0 / (2.0f * step);
          e.off = off;
        }
      }
      v0 = v1;
      v1 = v2;
      v2 = v3;
      v3 = v4;
    }
  }
  out[k] = e;
}

}  // namespace

void launch_find_zero(float *acc, size_t count, unsigned *d_max, int *d_count, hipStream_t s) {
  hipLaunchKernelGGL(k_find_zero, dim3(1024), dim3(256), 0, s, acc, count, d_max, d_count);
}

void launch_find_vote(const uint8_t *img, int w, int h, int stride, float thr, const float2 *trig, int nr, float *acc,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_find_vote, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, s, img, w, h, stride, thr, trig,
                     nr, acc);
}

void launch_find_peaks(const float *acc, int nr, float frac, unsigned *d_max, FindPeak *peaks, int *count,
                       hipStream_t s) {
  const size_t cells = static_cast<size_t>(FIND_THETA) * nr;
  hipLaunchKernelGGL(k_find_max, dim3(256), dim3(256), 0, s, acc, cells, d_max);
  hipLaunchKernelGGL(k_find_peaks, dim3(static_cast<unsigned>((cells + 255) / 256)), dim3(256), 0, s, acc, nr, frac,
                     d_max, peaks, count);
}

void launch_find_probe(const uint8_t *const *imgs, int w, int h, int stride, const FindProbe *probes, int n,
                       FindEdge *out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_find_probe, dim3((n + 255) / 256), dim3(256), 0, s, imgs, w, h, stride, probes, n, out);
}

}  // namespace lz
