// The anchors' search (anchor.hpp): one 256-lane workgroup an anchor, in the reference's order of float and double
// operations (liblizard/src/acquire.c:1876-2036, its scalar form), so its records are the reference's to the bit.
//   window     W x W points a / 16 modules apart about the anchor's square (W = 16 + 2 r, r = ceil(3 / h) steps of
//              h = a / 16 modules), each through H and the lens model and read bilinearly at gamma 1 (internal.h:11-23)
//   integrals  the window's integral images of the greys and their squares in FP64: a lane a row for the running row
//              sums, then a lane a column adding them down (the reference's sums, in its order)
//   correlate  every step (ox, oy) within r: the pattern's sum against the window as box sums of its two squares,
//              2 (S25 - S6) + S (1 - tm), over the window's spread; the first best step in row order, by lanes then a
//              tree that keeps the lower step on a tie
//   peak       a parabola each way where the best step is inside, and the anchor's place through H and the lens model
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "anchor.hpp"
#include "common.hpp"

namespace lz {
namespace {

constexpr int NT = 256;
constexpr int W1MAX = ANCH_WMAX + 1;
constexpr float ANCH_PEAK = 0.6f;  // acquire.c:1717

// A module point through H and, where the job has one, the lens model (acquire.c:1699-1702).
__device__ void anchor_map(const AnchorJob &j, float X, float Y, float *u, float *v) {
  project_f(j.H, X, Y, u, v);
  if (j.lens_on) {
    float dx, dy;
    lens_disp(j.K, j.t, j.ax, j.ay, j.R, *u, *v, &dx, &dy);
    *u += dx;
    *v += dy;
  }
}

// The box sum of side s at (x, y) of an integral image W1 entries a row.
__device__ double box(const double *I, int W1, int x, int y, int s) {
  return I[(y + s) * W1 + x + s] - I[y * W1 + x + s] - I[(y + s) * W1 + x] + I[y * W1 + x];
}

__global__ __launch_bounds__(NT) void k_anchor_find(const AnchorJob *__restrict__ jobs, AnchorHit *__restrict__ hits) {
  const AnchorJob &J = jobs[blockIdx.y];
  const int k = blockIdx.x, tid = threadIdx.x;
  if ((J.skip >> k) & 1ull) return;
  __shared__ double I1[W1MAX * W1MAX], I2[W1MAX * W1MAX];
  __shared__ float win[ANCH_WMAX * ANCH_WMAX];  // the window, then the correlations (O x O <= W x W)
  __shared__ float lut[256];
  __shared__ float rb[NT];
  __shared__ int ri[NT];
  const int a = J.span / 64;
  const float h = static_cast<float>(a) / ANCH_TS;
  const int r = static_cast<int>(ceilf(static_cast<float>(ANCH_REACH) / h)), W = ANCH_TS + 2 * r, O = 2 * r + 1;
  const int W1 = W + 1;
  const float x0 = J.m0[k % ANCH_K], y0 = J.m0[k / ANCH_K];
  lut[tid] = static_cast<float>(tid) / 255.0f;
  __syncthreads();
  for (int idx = tid; idx < W * W; idx += NT) {
    const int wy = idx / W, wx = idx - wy * W;
    float u, v;
    anchor_map(J, x0 + (wx - r + 0.5f) * h, y0 + (wy - r + 0.5f) * h, &u, &v);
    win[idx] = sample_px(J.img, J.w, J.h, lut, u, v);
  }
  for (int x = tid; x < W1; x += NT) I1[x] = I2[x] = 0;
  __syncthreads();
  if (tid < W) {  // row tid's running sums, at row tid + 1 of the integral images
    double s = 0, q = 0;
    double *o1 = I1 + (tid + 1) * W1, *o2 = I2 + (tid + 1) * W1;
    o1[0] = o2[0] = 0;
    for (int wx = 0; wx < W; wx++) {
      const double v = win[tid * W + wx];
      s += v;
      q += v * v;
      o1[wx + 1] = s;
      o2[wx + 1] = q;
    }
  }
  __syncthreads();
  if (tid >= 1 && tid < W1) {  // column tid: each row's entry is the one above plus the row's running sum
    for (int y = 2; y <= W; y++) {
      I1[y * W1 + tid] = I1[(y - 1) * W1 + tid] + I1[y * W1 + tid];
      I2[y * W1 + tid] = I2[(y - 1) * W1 + tid] + I2[y * W1 + tid];
    }
  }
  __syncthreads();
  float *C = win;
  const double tm1 = 1 - static_cast<double>(J.tm), tn = J.tn;
  const int l25 = J.j25 - J.i25 + 1, l6 = J.j6 - J.i6 + 1;
  float best = -2;
  int bi = r * O + r;
  for (int idx = tid; idx < O * O; idx += NT) {
    const int py = idx / O, px = idx - py * O;
    const double s1 = box(I1, W1, px, py, ANCH_TS), s2 = box(I2, W1, px, py, ANCH_TS);
    const double var = s2 - s1 * s1 / (ANCH_TS * ANCH_TS);
    const double s25 = J.j25 >= J.i25 ? box(I1, W1, px + J.i25, py + J.i25, l25) : 0;
    const double s6 = J.j6 >= J.i6 ? box(I1, W1, px + J.i6, py + J.i6, l6) : 0;
    const double st = 2 * (s25 - s6) + s1 * tm1;
    const float c = var > 1e-12 ? static_cast<float>(st / (sqrt(var) * tn)) : 0;
    C[idx] = c;
    if (c > best) {
      best = c;
      bi = idx;
    }
  }
  rb[tid] = best;
  ri[tid] = bi;
  __syncthreads();
  for (int m = NT / 2; m > 0; m >>= 1) {
    if (tid < m) {
      const float vb = rb[tid + m];
      const int ib = ri[tid + m];
      if (vb > rb[tid] || (vb == rb[tid] && ib < ri[tid])) {
        rb[tid] = vb;
        ri[tid] = ib;
      }
    }
    __syncthreads();
  }
  if (tid) return;
  best = rb[0];
  const int by = ri[0] / O - r, bx = ri[0] % O - r;
  float sx = 0, sy = 0;
  if (bx > -r && bx < r) {
    const float l = C[(by + r) * O + bx + r - 1], c = best, rr = C[(by + r) * O + bx + r + 1], d = l - 2 * c + rr;
    if (d < 0) sx = 0.5f * (l - rr) / d;
  }
  if (by > -r && by < r) {
    const float l = C[(by + r - 1) * O + bx + r], c = best, rr = C[(by + r + 1) * O + bx + r], d = l - 2 * c + rr;
    if (d < 0) sy = 0.5f * (l - rr) / d;
  }
  AnchorHit o;
  o.X = x0 + a * 0.5f;
  o.Y = y0 + a * 0.5f;
  o.peak = best;
  float u, v, hu, hv;
  anchor_map(J, o.X + (bx + sx) * h, o.Y + (by + sy) * h, &u, &v);
  project_f(J.H, o.X, o.Y, &hu, &hv);
  o.dx = u - hu;
  o.dy = v - hv;
  o.u = u;
  o.v = v;
  o.ok = best >= ANCH_PEAK && bx > -r && bx < r && by > -r && by < r;
  hits[blockIdx.y * ANCH_N + k] = o;
}

}  // namespace

void anchor_pattern(int span, AnchorJob &j) {
  const int a = span / 64, TS = ANCH_TS;
  need(a >= 2, "anchors of 2 modules or more (rings of 128 modules or more)");
  const float h = static_cast<float>(a) / TS;
  float T[ANCH_TS * ANCH_TS], tm = 0, tn = 0;
  for (int y = 0; y < TS; y++)
    for (int x = 0; x < TS; x++) {
      const float u = (x + 0.5f) * h - a * 0.5f, v = (y + 0.5f) * h - a * 0.5f;
      const float dc = std::fmax(std::fabs(u), std::fabs(v)) / (a * 0.5f);
      T[y * TS + x] = dc < 0.25f || dc >= 0.6f ? 1.0f : -1.0f;
      tm += T[y * TS + x];
    }
  tm /= static_cast<float>(TS * TS);
  for (int k = 0; k < TS * TS; k++) {
    T[k] -= tm;
    tn += T[k] * T[k];
  }
  j.tm = tm;
  j.tn = std::sqrt(tn);
  j.i25 = TS, j.j25 = -1, j.i6 = TS, j.j6 = -1;
  for (int x = 0; x < TS; x++) {
    const float d = std::fabs((x + 0.5f) * h - a * 0.5f) / (a * 0.5f);
    if (d < 0.25f) {
      if (x < j.i25) j.i25 = x;
      j.j25 = x;
    }
    if (d < 0.6f) {
      if (x < j.i6) j.i6 = x;
      j.j6 = x;
    }
  }
}

void launch_anchor_find(const AnchorJob *d_jobs, int frames, AnchorHit *d_hits, hipStream_t s) {
  hipLaunchKernelGGL(k_anchor_find, dim3(ANCH_N, frames), dim3(NT), 0, s, d_jobs, d_hits);
}

}  // namespace lz
