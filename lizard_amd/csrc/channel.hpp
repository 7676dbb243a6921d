// The camera channel on the GPU (synthetic captures for measurement inputs; not part of the receive path).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lz {

constexpr int CHAN_MAX_R = 16;  // the Gaussian's half-width, at most

struct ChannelParams {
  int W, Hh, n, span;
  double Hi[9];     // capture pixel -> module coordinates (the homography's inverse)
  double k1, ax, ay, D;  // radial lens (k1 |q|^2 + k2 |q|^4) q D about (ax, ay), q = (p - a) / D
  double k2;
  int r;            // Gaussian half-width (0: no blur)
  float taps[2 * CHAN_MAX_R + 1];
  float noise;      // AWGN sigma, grey levels
  uint64_t seed;
  // the whole symbol (null: the picture alone, grey 128 round it): the border's module map ((span + 30)^2 kinds,
  // device), the anchors' side and first modules; painted at 3 pixels a module, white round it
  const uint8_t *kinds;
  int anch_a;
  int anch_m0[7];
};

// The sender's page (liblizard/src/lizard_phy.c:1150-1236 and lizard_paint_grey :700-716): the symbol at pxm pixels a
// module with its quiet zone, (span + 34) pxm pixels a side. The picture and its guard interval (modules 12 ..
// span + 18) are the picture resampled as periodic through Lanczos-3 (i0: a pixel's first tap, w: its 6 weights, the
// same table on both axes, q = (span + 6) pxm entries), the border's modules black or white by their kinds, the
// anchors' squares over the picture.
struct PageParams {
  int n, span, pxm, q;
  const uint8_t *kinds;
  const int32_t *i0;
  const float *w;
  int anch_a;
  int anch_m0[7];
};
void launch_paint_page(const PageParams &p, const float *d_pic, uint8_t *d_page, hipStream_t s);

// Picture (n x n floats in [0, 1], device) -> capture (W x Hh uint8, device); d_a and d_b W x Hh float work areas.
void launch_channel(const ChannelParams &c, const float *d_pic, float *d_a, float *d_b, uint8_t *d_img,
                    hipStream_t s);

}  // namespace lz
