// The camera channel on the GPU, for synthetic captures: a picture rendered into a W x H uint8 frame through a radial
// lens (two radial terms) and a homography, blurred and noised (the model of lizard_amd/sim.py's capture; the lens form of
// liblizard/src/acquire.c:1470-1512). Not part of the receive path: it makes that path's measurement inputs on the
// device.
//   k_chan_render  per capture pixel: the lens undone by fixed-point iteration, the inverse homography to module
//                  coordinates, the picture read bilinearly (grey 128 outside it), times 255; or, given the border's
//                  module map, the whole symbol as a page shows it (liblizard/src/lizard_phy.c:1150-1236, :700-716):
//                  3 paint pixels a module read bilinearly, each white outside the symbol (the quiet zone and the
//                  page), black or white on the border's cells, the anchors' squares, else the picture taken as
//                  periodic (so the 3-module guard interval round it is the picture wrapped round its own edge)
//   k_chan_blur_x  the Gaussian along x (edge pixels repeated)
//   k_chan_blur_y  the Gaussian along y, AWGN from a counter-based generator, rounded half-even and clamped to uint8
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "channel.hpp"
#include "common.hpp"

namespace lz {
namespace {

// One paint pixel (ix, iy) of the symbol, 3 a module from its outer corner, in [0, 1].
__device__ float paint_px(const ChannelParams &c, const float *pic, int ix, int iy) {
  const int L = c.span + 30;
  if (ix < 0 || iy < 0 || ix >= 3 * L || iy >= 3 * L) return 1.0f;
  const int mx = ix / 3, my = iy / 3, k = c.kinds[my * L + mx] & 3;
  if (k) return k == 2 ? 0.0f : 1.0f;
  // the anchors (lizard_phy.c:1218-1234): a light middle to a quarter of the half-side, a dark ring to 0.6 of it, a
  // light rim, by the Chebyshev distance of the pixel's middle from the square's
  int ai = -1, aj = -1;
  for (int i = 0; i < 7; i++) {
    if (mx >= c.anch_m0[i] && mx < c.anch_m0[i] + c.anch_a) ai = i;
    if (my >= c.anch_m0[i] && my < c.anch_m0[i] + c.anch_a) aj = i;
  }
  if (ai >= 0 && aj >= 0) {
    const float half = 1.5f * c.anch_a;
    const float dx = fabsf(static_cast<float>(ix - 3 * c.anch_m0[ai]) + 0.5f - half);
    const float dy = fabsf(static_cast<float>(iy - 3 * c.anch_m0[aj]) + 0.5f - half);
    const float d = fmaxf(dx, dy) / half;
    return d < 0.25f || d >= 0.6f ? 1.0f : 0.0f;
  }
  // the picture, periodic: sample coordinate of the pixel's middle (exact where n = 3 span)
  const int n = c.n;
  const double sc = static_cast<double>(n) / (3.0 * c.span), off = 15.0 * n / c.span + 0.5;
  double sx = (ix + 0.5) * sc - off, sy = (iy + 0.5) * sc - off;
  sx -= floor(sx / n) * n;
  sy -= floor(sy / n) * n;
  int x0 = static_cast<int>(sx), y0 = static_cast<int>(sy);
  x0 = x0 >= n ? n - 1 : x0;
  y0 = y0 >= n ? n - 1 : y0;
  const int x1 = x0 + 1 == n ? 0 : x0 + 1, y1 = y0 + 1 == n ? 0 : y0 + 1;
  const float fx = static_cast<float>(sx - x0), fy = static_cast<float>(sy - y0);
  const float top = pic[y0 * n + x0] * (1 - fx) + pic[y0 * n + x1] * fx;
  const float bot = pic[y1 * n + x0] * (1 - fx) + pic[y1 * n + x1] * fx;
  return top * (1 - fy) + bot * fy;
}

__global__ __launch_bounds__(256) void k_chan_render(ChannelParams c, const float *pic, float *out) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= c.W) return;
  const double X = x + 0.5, Y = y + 0.5;
  double px = X, py = Y;
  for (int it = 0; it < 6; it++) {
    const double qx = (px - c.ax) / c.D, qy = (py - c.ay) / c.D, r2 = qx * qx + qy * qy;
    const double f = c.k1 * r2 + c.k2 * r2 * r2;
    px = X - f * qx * c.D;
    py = Y - f * qy * c.D;
  }
  const double w = c.Hi[6] * px + c.Hi[7] * py + c.Hi[8];
  const double mx = (c.Hi[0] * px + c.Hi[1] * py + c.Hi[2]) / w, my = (c.Hi[3] * px + c.Hi[4] * py + c.Hi[5]) / w;
  if (c.kinds) {
    // paint pixels' middles at k + 1/2: bilinear over the four round the point
    const double u = 3.0 * mx - 0.5, t = 3.0 * my - 0.5;
    const double uf = floor(u), tf = floor(t);
    // far outside the symbol (or no finite point) every paint pixel is white; keep the indices in int range
    const double lim = 3.0 * (c.span + 30) + 8.0;
    if (!(uf >= -8.0 && tf >= -8.0 && uf <= lim && tf <= lim)) {
      out[static_cast<size_t>(y) * c.W + x] = 255.0f;
      return;
    }
    const int i0 = static_cast<int>(uf), j0 = static_cast<int>(tf);
    const float fx = static_cast<float>(u - uf), fy = static_cast<float>(t - tf);
    const float top = paint_px(c, pic, i0, j0) * (1 - fx) + paint_px(c, pic, i0 + 1, j0) * fx;
    const float bot = paint_px(c, pic, i0, j0 + 1) * (1 - fx) + paint_px(c, pic, i0 + 1, j0 + 1) * fx;
    out[static_cast<size_t>(y) * c.W + x] = (top * (1 - fy) + bot * fy) * 255.0f;
    return;
  }
  const int n = c.n;
  double tx = (mx - 15.0) * (static_cast<double>(n) / c.span) - 0.5, ty = (my - 15.0) * (static_cast<double>(n) / c.span) - 0.5;
  float v = 128.0f / 255.0f;  // outside the picture: grey 128
  if (tx > -0.5 && tx < n - 0.5 && ty > -0.5 && ty < n - 0.5) {
    const double lim = n - 1.000001;
    tx = tx < 0 ? 0 : tx > lim ? lim : tx;
    ty = ty < 0 ? 0 : ty > lim ? lim : ty;
    const int x0 = static_cast<int>(tx), y0 = static_cast<int>(ty);
    const int x1 = x0 + 1 < n ? x0 + 1 : n - 1, y1 = y0 + 1 < n ? y0 + 1 : n - 1;
    const float fx = static_cast<float>(tx - x0), fy = static_cast<float>(ty - y0);
    const float top = pic[y0 * n + x0] * (1 - fx) + pic[y0 * n + x1] * fx;
    const float bot = pic[y1 * n + x0] * (1 - fx) + pic[y1 * n + x1] * fx;
    v = top * (1 - fy) + bot * fy;
  }
  out[static_cast<size_t>(y) * c.W + x] = v * 255.0f;
}

__global__ __launch_bounds__(256) void k_chan_blur_x(ChannelParams c, const float *in, float *out) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= c.W) return;
  const float *row = in + static_cast<size_t>(y) * c.W;
  float s = 0.0f;
  for (int k = -c.r; k <= c.r; k++) {
    const int xx = x + k < 0 ? 0 : x + k >= c.W ? c.W - 1 : x + k;
    s += c.taps[k + c.r] * row[xx];
  }
  out[static_cast<size_t>(y) * c.W + x] = s;
}

// Two standard normals from the frame's seed and a pixel's index (splitmix64 twice, Box-Muller).
__device__ __forceinline__ float normal_at(uint64_t seed, uint64_t i) {
  auto mix = [](uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const uint64_t a = mix(seed + 0x9E3779B97F4A7C15ull * (2 * i + 1)), b = mix(seed + 0x9E3779B97F4A7C15ull * (2 * i + 2));
  const double u1 = (static_cast<double>(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);  // (0, 1]
  const double u2 = static_cast<double>(b >> 11) * (1.0 / 9007199254740992.0);
  return static_cast<float>(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

__global__ __launch_bounds__(256) void k_chan_blur_y(ChannelParams c, const float *in, uint8_t *img) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= c.W) return;
  float s;
  if (c.r > 0) {
    s = 0.0f;
    for (int k = -c.r; k <= c.r; k++) {
      const int yy = y + k < 0 ? 0 : y + k >= c.Hh ? c.Hh - 1 : y + k;
      s += c.taps[k + c.r] * in[static_cast<size_t>(yy) * c.W + x];
    }
  } else {
    s = in[static_cast<size_t>(y) * c.W + x];
  }
  const uint64_t i = static_cast<uint64_t>(y) * c.W + x;
  if (c.noise > 0) s += c.noise * normal_at(c.seed, i);
  s = rintf(s);
  img[i] = static_cast<uint8_t>(s < 0 ? 0 : s > 255 ? 255 : s);
}

// One page pixel a lane: the quiet zone white; the anchors' squares (light middle to a quarter of the half-side, dark
// to 0.6, light rim); over modules 12 .. L - 12 the resample, down then across in the reference's order of float
// operations (each a sum of six products taken in turn, clamped to [0, 1]); elsewhere the border's kind; then
// (int)(v 255 + 1/2) clamped.
__global__ __launch_bounds__(256) void k_paint_page(PageParams p, const float *pic, uint8_t *page) {
  const int L = p.span + 30, C = p.pxm, side = (L + 4) * C;
  const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
  if (X >= side) return;
  const int dx = X - 2 * C, dy = Y - 2 * C;
  float v = 1.0f;
  if (dx >= 0 && dy >= 0 && dx < L * C && dy < L * C) {
    const int mx = dx / C, my = dy / C;
    int ai = -1, aj = -1;
    for (int i = 0; i < 7; i++) {
      if (mx >= p.anch_m0[i] && mx < p.anch_m0[i] + p.anch_a) ai = i;
      if (my >= p.anch_m0[i] && my < p.anch_m0[i] + p.anch_a) aj = i;
    }
    if (ai >= 0 && aj >= 0) {
      const float half = 0.5f * static_cast<float>(p.anch_a * C);
      const float ex = fabsf(static_cast<float>(dx - p.anch_m0[ai] * C) + 0.5f - half);
      const float ey = fabsf(static_cast<float>(dy - p.anch_m0[aj] * C) + 0.5f - half);
      const float d = fmaxf(ex, ey) / half;
      v = d < 0.25f || d >= 0.6f ? 1.0f : 0.0f;
    } else if (mx >= 12 && my >= 12 && mx < L - 12 && my < L - 12) {
      const int n = p.n, x = dx - 12 * C, y = dy - 12 * C;
      const int ix = p.i0[x], iy = p.i0[y];
      const float *wx = p.w + 6 * x, *wy = p.w + 6 * y;
      int rows[6];
#pragma unroll
      for (int t = 0; t < 6; t++) {
        const int r = (iy + t) % n;
        rows[t] = r < 0 ? r + n : r;
      }
      float acc = 0.0f;
#pragma unroll
      for (int s = 0; s < 6; s++) {
        int k = (ix + s) % n;
        k = k < 0 ? k + n : k;
        float col = wy[0] * pic[rows[0] * n + k];
#pragma unroll
        for (int t = 1; t < 6; t++) col = col + wy[t] * pic[rows[t] * n + k];
        acc = s == 0 ? wx[0] * col : acc + wx[s] * col;
      }
      v = acc < 0.0f ? 0.0f : acc > 1.0f ? 1.0f : acc;
    } else {
      v = (p.kinds[my * L + mx] & 3) == 2 ? 0.0f : 1.0f;
    }
  }
  const float t = v * 255.0f + 0.5f;
  const int g = static_cast<int>(t);
  page[static_cast<size_t>(Y) * side + X] = static_cast<uint8_t>(g < 0 ? 0 : g > 255 ? 255 : g);
}

}  // namespace

void launch_paint_page(const PageParams &p, const float *d_pic, uint8_t *d_page, hipStream_t s) {
  const int side = (p.span + 34) * p.pxm;
  hipLaunchKernelGGL(k_paint_page, dim3((side + 255) / 256, side), dim3(256), 0, s, p, d_pic, d_page);
  hip_check(hipGetLastError(), "page launch");
}

void launch_channel(const ChannelParams &c, const float *d_pic, float *d_a, float *d_b, uint8_t *d_img,
                    hipStream_t s) {
  const dim3 g((c.W + 255) / 256, c.Hh);
  hipLaunchKernelGGL(k_chan_render, g, dim3(256), 0, s, c, d_pic, d_a);
  if (c.r > 0) {
    hipLaunchKernelGGL(k_chan_blur_x, g, dim3(256), 0, s, c, d_a, d_b);
    hipLaunchKernelGGL(k_chan_blur_y, g, dim3(256), 0, s, c, d_b, d_img);
  } else {
    hipLaunchKernelGGL(k_chan_blur_y, g, dim3(256), 0, s, c, d_a, d_img);
  }
  hip_check(hipGetLastError(), "channel launches");
}

}  // namespace lz
