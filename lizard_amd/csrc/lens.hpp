// The projective map, the lens model's rows and the capture's bilinear read, for host and device alike, in the
// reference's order of float and double operations (liblizard/src/internal.h:107-112 project, :11-23 sample,
// liblizard/src/acquire.c:1506-1512 lens_rows, :1615-1620 lensm_at), so both sides give the reference's values to the
// bit (every unit is built without contraction).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lz {

constexpr int LENS_TERMS = 3, LENS_P = 8 + LENS_TERMS;

// H applied to a module point in float.
__host__ __device__ inline void project_f(const float *h, float x, float y, float *u, float *v) {
  const float w = h[6] * x + h[7] * y + 1;
  *u = (h[0] * x + h[1] * y + h[2]) / w;
  *v = (h[3] * x + h[4] * y + h[5]) / w;
}

// The model's rows at q = (x, y): x, y, the shift and the two perspective terms on each axis, then K radial terms.
__host__ __device__ inline void lens_rows(double x, double y, int K, double *rx, double *ry) {
  rx[0] = x, rx[1] = y, rx[2] = 0, rx[3] = 0, rx[4] = 1, rx[5] = 0, rx[6] = -x * x, rx[7] = -x * y;
  ry[0] = 0, ry[1] = 0, ry[2] = x, ry[3] = y, ry[4] = 0, ry[5] = 1, ry[6] = -x * y, ry[7] = -y * y;
  const double s = x * x + y * y;
  double sk = s;
  for (int k = 0; k < K; k++, sk *= s) {
    rx[8 + k] = sk * x;
    ry[8 + k] = sk * y;
  }
}

// A model's displacement at image point (u, v): K radial terms, coefficients t, axis (ax, ay), R.
__host__ __device__ inline void lens_disp(int K, const double *t, double ax, double ay, double R, double u, double v,
                                          float *dx, float *dy) {
  double rx[LENS_P], ry[LENS_P], fx = 0, fy = 0;
  lens_rows((u - ax) / R, (v - ay) / R, K, rx, ry);
  for (int a = 0; a < 8 + K; a++) {
    fx += t[a] * rx[a];
    fy += t[a] * ry[a];
  }
  *dx = static_cast<float>(fx);
  *dy = static_cast<float>(fy);
}

// A point of a grey capture (rows w bytes apart) read bilinearly through lut (gamma 1: i / 255): the half pixel, the
// clamp to the frame, the blend.
__host__ __device__ inline float sample_px(const uint8_t *px, int w, int h, const float *lut, float x, float y) {
  x -= 0.5f;
  y -= 0.5f;
  if (x < 0) x = 0;
  else if (x > w - 1.001f) x = w - 1.001f;
  if (y < 0) y = 0;
  else if (y > h - 1.001f) y = h - 1.001f;
  const int xi = x >= 0 ? static_cast<int>(x) : 0, yi = y >= 0 ? static_cast<int>(y) : 0;
  const float fx = x - xi, fy = y - yi;
  const uint8_t *p = px + static_cast<size_t>(yi) * w + xi;
  const float a = lut[p[0]], b = lut[p[1]], c = lut[p[w]], d = lut[p[w + 1]];
  return (a + (b - a) * fx) * (1 - fy) + (c + (d - c) * fx) * fy;
}

}  // namespace lz
