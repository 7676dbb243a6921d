// Picture samples to soft values on gfx950 (liblizard/src/lizard_phy.c:1566-1772, research/SPEC.md 6.9-6.13, 7.3,
// 8.5):
//   k_erase_fit  n / 32 workgroups a frame: the shading's least-squares sums on every eighth sample outside the
//                anchors, four fit rows a workgroup
//   k_rows       2 W rows a workgroup, two a wave packed as one complex sequence: the warp (or a supplied grid) read
//                on load with the anchors' samples erased, the shading's beta from the fit's parts, the detrend row
//                sums, the wave FFT along x, the two rows' transforms separated and pruned to u < E, written
//                column-major
//   k_cols       one column u a wave: the wave FFT along y, the column's coefficients scattered into coefficient order
//   k_align      one workgroup a frame: the detrend projections and the pilots' two-pass least-squares grid shift
//   k_demap      one wave a sub-channel: 320 coefficients gathered, detrended, turned, their moments, int8 LLRs
//   k_est        one lane a sub-channel: the estimate J from the argument the demap left (single-frame paths: in
//                k_pilot_read, one wave a block, which also reads the block's pilots)
// The wave FFT (WaveFft): n = 64 M points, M in registers a lane, an M-point DFT in registers, then the 64-point stage
// as R1 x 64 / R1 (a radix plan the autotuner picks) through the wave's own LDS; every size compiled in. Floating-point steps the reference rounds in a fixed order (the
// warp, the detrend terms, the turn, the LLR scale) use explicitly rounded operations, so no contraction into fused
// multiply-adds changes them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <type_traits>

#include "sample_math.hpp"
#include "spectrum.hpp"

namespace lz {

// ------------------------------------------------------------------------------------------------ FFT in LDS
// The transforms' products and sums may fuse (fused multiply-adds): the FFT is held to FP64 within a tolerance, and
// every path uses the same transform, so this changes no comparison between the build's own paths. Everywhere else
// (the warp, the detrend, the turn, the soft scale) every operation rounds on its own, in the reference's order.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(fast)
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int R>
__device__ __forceinline__ void dft(float2 *v) {
#pragma clang fp contract(fast)
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const float2 t0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y), t1 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 t2 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y), t3 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    v[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
    v[2] = make_float2(t0.x - t2.x, t0.y - t2.y);
    v[1] = make_float2(t1.x + t3.y, t1.y - t3.x);  // t1 - i t3
    v[3] = make_float2(t1.x - t3.y, t1.y + t3.x);  // t1 + i t3
  } else if constexpr (R == 3) {
    const float k = 0.866025403784438647f;  // sin(2 pi / 3)
    const float2 t1 = make_float2(v[1].x + v[2].x, v[1].y + v[2].y), t2 = make_float2(v[1].x - v[2].x, v[1].y - v[2].y);
    const float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    v[0] = make_float2(v[0].x + t1.x, v[0].y + t1.y);
    v[1] = make_float2(m.x + k * t2.y, m.y - k * t2.x);  // m - i k t2
    v[2] = make_float2(m.x - k * t2.y, m.y + k * t2.x);  // m + i k t2
  } else {
    static_assert(R == 5, "radix");
    const float c1 = 0.309016994374947424f, c2 = -0.809016994374947424f;   // cos(2 pi / 5), cos(4 pi / 5)
    const float s1 = 0.951056516295153572f, s2 = 0.587785252292473129f;   // sin(2 pi / 5), sin(4 pi / 5)
    const float2 a = v[0];
    const float2 t1 = make_float2(v[1].x + v[4].x, v[1].y + v[4].y), t2 = make_float2(v[2].x + v[3].x, v[2].y + v[3].y);
    const float2 t3 = make_float2(v[1].x - v[4].x, v[1].y - v[4].y), t4 = make_float2(v[2].x - v[3].x, v[2].y - v[3].y);
    const float2 m1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const float2 m2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    const float2 n1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const float2 n2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    v[0] = make_float2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
    v[1] = make_float2(m1.x + n1.y, m1.y - n1.x);  // m1 - i n1
    v[4] = make_float2(m1.x - n1.y, m1.y + n1.x);
    v[2] = make_float2(m2.x + n2.y, m2.y - n2.x);  // m2 - i n2
    v[3] = make_float2(m2.x - n2.y, m2.y + n2.x);
  }
}

// f(integral_constant<int, i>) for i in [B, E): loops whose indices must be constants (register arrays, twiddles)
template <int B, int E, class F>
__device__ __forceinline__ void sfor(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// ---- compile-time twiddles: cos and sin of 2 pi e / m by their Taylor series in double (|angle| <= pi after the
// reduction below, where 30 terms are exact to double precision), rounded to float once
constexpr double kPi = 3.14159265358979323846;
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 30; k++) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1, sum = 1;
  for (int k = 1; k < 30; k++) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// exp(-2 pi i e / m) as (cos, sin) pairs, the angle reduced to (-pi, pi] exactly in integers first
constexpr double angle_of(int e, int m) {
  const int r = ((e % m) + m) % m;
  const int s = 2 * r > m ? r - m : r;  // r or r - m, whichever lies in (-m/2, m/2]
  return -2.0 * kPi * s / m;
}

template <int M, int E>
__device__ __forceinline__ float2 twiddle(float2 x) {
#pragma clang fp contract(fast)
  constexpr int r = ((E % M) + M) % M;
  if constexpr (r == 0) {
    return x;
  } else if constexpr (4 * r == M) {  // -i
    return make_float2(x.y, -x.x);
  } else if constexpr (2 * r == M) {  // -1
    return make_float2(-x.x, -x.y);
  } else if constexpr (4 * r == 3 * M) {  // +i
    return make_float2(-x.y, x.x);
  } else {
    constexpr float c = static_cast<float>(taylor_cos(angle_of(E, M)));
    constexpr float sn = static_cast<float>(taylor_sin(angle_of(E, M)));
    return make_float2(x.x * c - x.y * sn, x.x * sn + x.y * c);
  }
}

constexpr int first_factor(int m) { return m % 4 == 0 ? 4 : m % 2 == 0 ? 2 : m % 3 == 0 ? 3 : 5; }

// In-register DFT of M points, natural order in and out (Cooley-Tukey over M = P Q, P the first factor; every
// index and twiddle a compile-time constant).
template <int M>
__device__ __forceinline__ void dft_small(float2 *v) {
  if constexpr (M == 1) {
    return;
  } else if constexpr (M <= 5) {
    dft<M>(v);
  } else {
    constexpr int P = first_factor(M), Q = M / P;
    float2 t[M];
    sfor<0, Q>([&](auto q) {
      float2 a[P];
      sfor<0, P>([&](auto p) { a[p] = v[q + Q * p]; });
      dft_small<P>(a);
      sfor<0, P>([&](auto k1) { t[k1 * Q + q] = twiddle<M, q * k1>(a[k1]); });
    });
    sfor<0, P>([&](auto k1) {
      dft_small<Q>(t + k1 * Q);
      sfor<0, Q>([&](auto k2) { v[k1 + P * k2] = t[k1 * Q + k2]; });
    });
  }
}

// A wave's LDS operations are ordered: wait for this lane's writes, then let no access move across.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One 64-lane wave transforms one complex sequence of N = 64 M points (unscaled forward DFT): lane l holds
// x[l + 64 m], m < M, in registers; an M-point DFT in registers, the twiddles w_N^(l k1), then the 64-point DFTs
// across lanes as two passes through the wave's own LDS area S (64 MP points, MP odd so that a lane stride of MP
// spreads over the banks): 64 = R1 x R2, R2 = 64 / R1 (the radix plan: 8 x 8, 4 x 16 or 16 x 4). Z(k) ends at
// S[pos(k)]. No workgroup barrier: waves run independently.
template <int N, int R1 = 8>
struct WaveFft {
  static constexpr int M = N / 64;
  static constexpr int R2 = 64 / R1;
  static constexpr int MP = M % 2 == 0 ? M + 1 : M;
  static constexpr int LDS = 64 * MP;               // float2 a wave
  static constexpr int TASKS1 = R1 * M, TASKS2 = R2 * M;  // R2-point DFTs in pass one, R1-point in pass two
  static constexpr int ROUNDS1 = (TASKS1 + 63) / 64, ROUNDS2 = (TASKS2 + 63) / 64;
  static_assert(N % 64 == 0 && M >= 1 && M <= 24, "picture sizes 64 M, M <= 24");
  static_assert(R1 == 4 || R1 == 8 || R1 == 16, "radix plans 4 x 16, 8 x 8, 16 x 4");

  // k = k1 + M (c + R2 d) sits at position (d + R1 c) MP + k1
  __device__ static __forceinline__ int pos(int k) {
    const int k1 = k % M, k2 = k / M;
    return ((k2 / R2) + R1 * (k2 % R2)) * MP + k1;
  }

  __device__ static __forceinline__ void run(float2 (&v)[M], float2 *S, const float2 *tw, int lane) {
    dft_small<M>(v);
#pragma unroll
    for (int k1 = 0; k1 < M; k1++) S[lane * MP + k1] = k1 ? cmul(v[k1], tw[lane * k1]) : v[k1];
    wave_sync();
    // 64 = R1 x R2 on each k1: n2 = a + R1 b. Pass one: R2 points over b, twiddle w_64^(a c), in place at a + R1 c.
#pragma unroll
    for (int r = 0; r < ROUNDS1; r++) {
      const int t = lane + 64 * r;
      if (TASKS1 % 64 == 0 || t < TASKS1) {
        const int k1 = t % M, a = t / M;
        float2 u[R2];
#pragma unroll
        for (int b = 0; b < R2; b++) u[b] = S[(a + R1 * b) * MP + k1];
        dft_small<R2>(u);
#pragma unroll
        for (int c = 0; c < R2; c++) S[(a + R1 * c) * MP + k1] = c ? cmul(u[c], tw[a * c * M]) : u[c];
      }
    }
    wave_sync();
    // pass two: R1 points over a, output d in place at d + R1 c
#pragma unroll
    for (int r = 0; r < ROUNDS2; r++) {
      const int t = lane + 64 * r;
      if (TASKS2 % 64 == 0 || t < TASKS2) {
        const int k1 = t % M, c = t / M;
        float2 u[R1];
#pragma unroll
        for (int a = 0; a < R1; a++) u[a] = S[(a + R1 * c) * MP + k1];
        dft_small<R1>(u);
#pragma unroll
        for (int d = 0; d < R1; d++) S[(d + R1 * c) * MP + k1] = u[d];
      }
    }
    wave_sync();
  }
};

// Calls f with std::integral_constant<int, n> for every picture size LIZARD uses (64 M, M in 6, 8, 12, 16, 18, 20, 24).
template <class F>
void with_size(int n, F &&f) {
  switch (n) {
    case 384: f(std::integral_constant<int, 384>{}); break;
    case 512: f(std::integral_constant<int, 512>{}); break;
    case 768: f(std::integral_constant<int, 768>{}); break;
    case 1024: f(std::integral_constant<int, 1024>{}); break;
    case 1152: f(std::integral_constant<int, 1152>{}); break;
    case 1280: f(std::integral_constant<int, 1280>{}); break;
    case 1536: f(std::integral_constant<int, 1536>{}); break;
    default: throw Error(LIZ_E_FORMAT, "a picture size outside 384, 512, 768, 1024, 1152, 1280, 1536");
  }
}

// ------------------------------------------------------------------------------------------------ anchors, shading
__device__ __forceinline__ bool in_square(const PlanDev &P, int t) {
  bool in = false;
#pragma unroll
  for (int i = 0; i < ANCH_K; i++) in |= t >= P.elo[i] && t <= P.ehi[i];
  return in;
}

// The shading's basis at sample (x, y): 1, c(x), q(x), c(y), q(y), c(x) c(y) (lizard_phy.c:337-340).
__device__ __forceinline__ void hole_basis(const PlanDev &P, int x, int y, double *b) {
  const double cx = P.cx[x], cy = P.cx[y];
  b[0] = 1;
  b[1] = cx;
  b[2] = P.qx[x];
  b[3] = cy;
  b[4] = P.qx[y];
  b[5] = __dmul_rn(cx, cy);
}

// The fill of an erased sample (lizard_phy.c:341-346): the shading there, summed in the basis' order.
__device__ __forceinline__ float hole_value(const PlanDev &P, const double *beta, int x, int y) {
  double b[6], v = 0;
  hole_basis(P, x, y, b);
#pragma unroll
  for (int r = 0; r < 6; r++) v = __dadd_rn(v, __dmul_rn(beta[r], b[r]));
  return static_cast<float>(v);
}

// K sums over the workgroup at once, each by the same fixed tree (level o = NT/2, ..., 1 adds element t + o to element
// t, t < o), so the sums are run-to-run identical: the levels across waves through LDS, the last six within wave 0 by
// shuffles. Every thread gets the sums. red holds K * NT doubles.
template <int NT, int K>
__device__ void block_sums(double (&v)[K], double *red, int tid) {
  static_assert(NT >= 64 && (NT & (NT - 1)) == 0, "a power-of-two workgroup of whole waves");
#pragma unroll
  for (int k = 0; k < K; k++) red[k * NT + tid] = v[k];
  __syncthreads();
  for (int o = NT / 2; o >= 64; o >>= 1) {
    if (tid < o) {
#pragma unroll
      for (int k = 0; k < K; k++) red[k * NT + tid] += red[k * NT + tid + o];
    }
    __syncthreads();
  }
  if (tid < 64) {
    double x[K];
#pragma unroll
    for (int k = 0; k < K; k++) x[k] = red[k * NT + tid];
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
      for (int k = 0; k < K; k++) {
        const double y = __shfl_down(x[k], o, 64);
        if (tid < o) x[k] += y;
      }
    }
    if (tid == 0) {
#pragma unroll
      for (int k = 0; k < K; k++) red[k * NT] = x[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = red[k * NT];
  __syncthreads();
}

// ------------------------------------------------------------------------------------------------ the warp

// A bilinear read of the uint8 capture at pixel coordinates (x, y) (pixel centres at half integers), clamped to the
// frame, grey level i / 255 (liblizard/src/internal.h:11-23 at gamma 1). In bounds for any input: NaN reads pixel 0
// (the reference leaves NaN to C's undefined float-to-int conversion). Branch-free: the clamp as two selects, the
// fraction x - (int) x as v_fract (exact for x >= 0, checked on the device for every float in [0, 2^31)), the
// pixel's offset in 32 bits (the frame's rows stride < 2^24 bytes apart and stride h < 2^31, which every host entry
// that builds a WarpFrame checks) from the frame's base held in scalar registers, read through global loads.
__device__ __forceinline__ float sample(const WarpFrame &F, float x, float y) {
  x = __fsub_rn(x, 0.5f);
  y = __fsub_rn(y, 0.5f);
  const float xm = __fsub_rn(static_cast<float>(F.w), 1.001f), ym = __fsub_rn(static_cast<float>(F.h), 1.001f);
  x = fminf(x > 0.0f ? x : 0.0f, xm);
  y = fminf(y > 0.0f ? y : 0.0f, ym);
  const int xi = static_cast<int>(x), yi = static_cast<int>(y);
  const float fx = frac_pos(x), fy = frac_pos(y);
  // a pixel pair (x, x + 1) as one 16-bit load at any byte address (the device's unaligned global access)
  struct __attribute__((packed)) Pair { uint16_t v; };
  const auto *g = (const __attribute__((address_space(1))) uint8_t *)F.img;
  const uint32_t o = __umul24(static_cast<uint32_t>(yi), static_cast<uint32_t>(F.stride)) + static_cast<uint32_t>(xi);
  const uint32_t t = ((const __attribute__((address_space(1))) Pair *)(g + o))->v;
  const uint32_t u = ((const __attribute__((address_space(1))) Pair *)(g + o + static_cast<uint32_t>(F.stride)))->v;
  const float a = grey(t & 0xff), b = grey(t >> 8), c = grey(u & 0xff), d = grey(u >> 8);
  const float top = __fadd_rn(a, __fmul_rn(__fsub_rn(b, a), fx)), bot = __fadd_rn(c, __fmul_rn(__fsub_rn(d, c), fx));
  return __fadd_rn(__fmul_rn(top, __fsub_rn(1.0f, fy)), __fmul_rn(bot, fy));
}

// The row part of a warp sample (everything that depends on y alone).
struct WarpRow {
  float my, v, omv, kw, ku, kv;
  int j;
};
__device__ __forceinline__ WarpRow warp_row(const PlanDev &P, const WarpFrame &F, int y) {
  WarpRow r;
  r.my = P.mcoord[y];
  r.j = P.mcell[y];
  r.v = __fdiv_rn(__fsub_rn(r.my, P.node_y[r.j]), __fsub_rn(P.node_y[r.j + 1], P.node_y[r.j]));
  r.omv = __fsub_rn(1.0f, r.v);
  r.kw = __fadd_rn(__fmul_rn(F.H[7], r.my), 1.0f);
  r.ku = __fadd_rn(__fmul_rn(F.H[1], r.my), F.H[2]);
  r.kv = __fadd_rn(__fmul_rn(F.H[4], r.my), F.H[5]);
  return r;
}

// Picture sample (x, y) of frame F: module coordinates (15 + (t + 1/2) span / n on either axis, SPEC 6.9), the
// homography H plus the node lattice's bilinear displacement (acquire.c:5095-5103), read from the capture. The float
// operations are the reference's scalar grid walk's, in its order (acquire.c:5121-5196 without OB_SIMD), so the
// samples are that walk's to the bit.
// The part of a warp sample that depends on its row and its node cell i alone: the node displacement's slope and
// intercept in mx on either axis, (ex, cx, ey, cy), in the walk's float order.
__device__ __forceinline__ float4 cell_terms(const PlanDev &P, const WarpFrame &F, const WarpRow &r, int i) {
  const int nx = P.nx, j = r.j;
  const float v = r.v, omv = r.omv;
  const float2 n00 = F.nodes[j * nx + i], n10 = F.nodes[j * nx + i + 1];
  const float2 n01 = F.nodes[(j + 1) * nx + i], n11 = F.nodes[(j + 1) * nx + i + 1];
  const float inv = P.node_inv[i];
  const float ex = __fmul_rn(__fadd_rn(__fmul_rn(__fsub_rn(n10.x, n00.x), omv), __fmul_rn(__fsub_rn(n11.x, n01.x), v)), inv);
  const float cx = __fsub_rn(__fadd_rn(__fmul_rn(n00.x, omv), __fmul_rn(n01.x, v)), __fmul_rn(ex, P.node_x[i]));
  const float ey = __fmul_rn(__fadd_rn(__fmul_rn(__fsub_rn(n10.y, n00.y), omv), __fmul_rn(__fsub_rn(n11.y, n01.y), v)), inv);
  const float cy = __fsub_rn(__fadd_rn(__fmul_rn(n00.y, omv), __fmul_rn(n01.y, v)), __fmul_rn(ey, P.node_x[i]));
  return make_float4(ex, cx, ey, cy);
}

// The capture's point for sample x of a row, its cell terms given.
__device__ __forceinline__ float2 warp_point(const PlanDev &P, const WarpFrame &F, const WarpRow &r, int x, float4 ct) {
  const float mx = P.mcoord[x];
  const float *H = F.H;
  const float rw = recip_rn(__fadd_rn(__fmul_rn(H[6], mx), r.kw));
  const float pu = __fmul_rn(__fadd_rn(__fmul_rn(H[0], mx), r.ku), rw), pv = __fmul_rn(__fadd_rn(__fmul_rn(H[3], mx), r.kv), rw);
  const float dx = __fadd_rn(ct.y, __fmul_rn(ct.x, mx)), dy = __fadd_rn(ct.w, __fmul_rn(ct.z, mx));
  return make_float2(__fadd_rn(pu, dx), __fadd_rn(pv, dy));
}

__device__ __forceinline__ float warp_at(const PlanDev &P, const WarpFrame &F, const WarpRow &r, int x) {
  const float2 q = warp_point(P, F, r, x, cell_terms(P, F, r, P.mcell[x]));
  return sample(F, q.x, q.y);
}

__device__ __forceinline__ float warp_sample(const PlanDev &P, const WarpFrame &F, int x, int y) {
  return warp_at(P, F, warp_row(P, F, y), x);
}

// Where the path's picture samples come from: a grid the caller supplied, or captures through their registrations.
// operator() reads one sample; row() and at() read a row's samples with the row's part done once.
struct GridSource {
  const float *grid;
  __device__ __forceinline__ float operator()(const PlanDev &P, int f, int x, int y) const {
    return grid[(static_cast<size_t>(f) * P.n + y) * P.n + x];
  }
  using Row = const float *;
  __device__ __forceinline__ Row row(const PlanDev &P, int f, int y) const {
    return grid + (static_cast<size_t>(f) * P.n + y) * P.n;
  }
  __device__ __forceinline__ float at(const PlanDev &, Row r, int x) const { return r[x]; }
};
struct GridListSource {
  const float *const *grids;
  __device__ __forceinline__ float operator()(const PlanDev &P, int f, int x, int y) const {
    return grids[f][static_cast<size_t>(y) * P.n + x];
  }
  using Row = const float *;
  __device__ __forceinline__ Row row(const PlanDev &P, int f, int y) const {
    return grids[f] + static_cast<size_t>(y) * P.n;
  }
  __device__ __forceinline__ float at(const PlanDev &, Row r, int x) const { return r[x]; }
};
struct WarpSource {
  const WarpFrame *frames;
  __device__ __forceinline__ float operator()(const PlanDev &P, int f, int x, int y) const {
    return warp_sample(P, frames[f], x, y);
  }
  struct Row {
    WarpFrame F;
    WarpRow r;
  };
  __device__ __forceinline__ Row row(const PlanDev &P, int f, int y) const {
    Row out;
    out.F = frames[f];
    out.r = warp_row(P, out.F, y);
    return out;
  }
  __device__ __forceinline__ float at(const PlanDev &P, const Row &r, int x) const { return warp_at(P, r.F, r.r, x); }
};

__global__ __launch_bounds__(256) void k_warp(PlanDev P, const WarpFrame *frames, float *grid) {
  const int f = blockIdx.y, n = P.n;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * n) return;
  grid[static_cast<size_t>(f) * n * n + t] = warp_sample(P, frames[f], t % n, t / n);
}

// The shading fit (lizard_phy.c:1605-1624): rhs over every eighth row and column outside the squares, beta =
// hole_inv rhs.
template <class Src>
__global__ __launch_bounds__(256) void k_erase_fit(PlanDev P, Src src, double *part_out, int *zero_int) {
  if (zero_int && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *zero_int = 0;
  __shared__ double red[4][6];
  const int tid = threadIdx.x, b = blockIdx.x, f = blockIdx.y, n = P.n, c = n / 8;
  double rhs[6] = {0, 0, 0, 0, 0, 0};
  // this part's FIT_ROWS rows of the fit lattice
  for (int k = tid; k < FIT_ROWS * c; k += 256) {
    const int y = 8 * (b * FIT_ROWS + k / c), x = 8 * (k % c);
    if (in_square(P, y) && in_square(P, x)) continue;
    const double v = src(P, f, x, y);
    double bs[6];
    hole_basis(P, x, y, bs);
#pragma unroll
    for (int r = 0; r < 6; r++) rhs[r] += bs[r] * v;
  }
  // a fixed tree: the wave's lanes by shuffles, then the four waves in order
#pragma unroll
  for (int r = 0; r < 6; r++) {
    for (int o = 32; o > 0; o >>= 1) rhs[r] += __shfl_xor(rhs[r], o, 64);
  }
  if ((tid & 63) == 0) {
#pragma unroll
    for (int r = 0; r < 6; r++) red[tid >> 6][r] = rhs[r];
  }
  __syncthreads();
  if (tid < 6) part_out[static_cast<size_t>(f) * FIT_STRIDE + b * 6 + tid] =
      (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

// ------------------------------------------------------------------------------------------------ the two passes
// Workgroups are dealt to the 8 XCDs round-robin (linear id mod 8), each XCD with its own L2. Where a grid row holds
// a multiple of 8 workgroups, block bx takes logical block (bx mod 8) (nb / 8) + bx / 8: each XCD gets a contiguous
// range, so neighbouring blocks' writes into the same cache lines meet in one L2 instead of leaving each line half
// written in two. Only the assignment changes, never a result.
__device__ __forceinline__ int xcd_block(int bx, int nb) { return nb % 8 ? bx : (bx & 7) * (nb >> 3) + (bx >> 3); }

// The same for a grid of (blocks, frames) whose blocks of one frame write into the same lines of that frame's output
// (the column pass's coefficients in coefficient order): where the frames are a multiple of 8, workgroup linear id L
// (on XCD L mod 8) takes frame (L mod 8) + 8 (L / 8 / nb) and block (L / 8) mod nb, so each frame's blocks all run on
// one XCD and its lines fill in that XCD's L2. A bijection of (block, frame); only the assignment changes.
__device__ __forceinline__ void xcd_frame(int &bx, int &f) {
  const int nb = gridDim.x, nf = gridDim.y;
  if (nf % 8) return;
  const int l = blockIdx.x + blockIdx.y * nb, k = l >> 3;
  f = (l & 7) + 8 * (k / nb);
  bx = k % nb;
}

// k_rows: W waves a workgroup, two picture rows a wave packed as one complex sequence (2 W rows a workgroup)

// Rows y and y + 1 of wave w: load with the anchors' samples erased, the detrend's row sums (s, s c(x), s q(x)) in
// double, one wave FFT of row y + i row y + 1, the two real rows' transforms separated (Z(u) and conj Z(n - u)) for
// u < E and written column-major, X[frame][u][y].
template <class Src, int N, int W, int R1>
__global__ __launch_bounds__(64 * W) void k_rows(PlanDev P, Src src, const double *fitpart, float2 *X, double *rowsum) {
  using F = WaveFft<N, R1>;
  constexpr int M = F::M;
  extern __shared__ float2 zs[];  // W waves x F::LDS
  __shared__ double bsh[6];
  constexpr bool WARP = std::is_same<Src, WarpSource>::value;
  // per wave, its two rows' node-cell terms (a cell's share of every sample in it)
  __shared__ float4 cterm[WARP ? W : 1][2][WARP ? MAX_NODES : 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, f = blockIdx.y;
  const int y = xcd_block(blockIdx.x, gridDim.x) * (2 * W) + 2 * w;
  // the shading beta = hole_inv rhs (lizard_phy.c:1618-1624) from the fit's parts summed in part order, by each
  // workgroup (a few dozen doubles; one launch fewer on the frame's path)
  {
    __shared__ double rhs_sh[6];
    if (tid < 6) {
      const double *Fp = fitpart + static_cast<size_t>(f) * FIT_STRIDE;
      double s = 0;
      for (int b = 0; b < N / 8 / FIT_ROWS; b++) s += Fp[b * 6 + tid];
      rhs_sh[tid] = s;
    }
    __syncthreads();
    if (tid < 6) {
      double s = 0;
      for (int k = 0; k < 6; k++) s += P.hole_inv[tid * 6 + k] * rhs_sh[k];
      bsh[tid] = s;
    }
    __syncthreads();
  }
  const bool yin0 = in_square(P, y), yin1 = in_square(P, y + 1);
  const auto row0 = src.row(P, f, y), row1 = src.row(P, f, y + 1);
  if constexpr (WARP) {
    for (int i = lane; i < P.nx - 1; i += 64) {
      cterm[w][0][i] = cell_terms(P, row0.F, row0.r, i);
      cterm[w][1][i] = cell_terms(P, row1.F, row1.r, i);
    }
    wave_sync();
  }
  float2 v[M];
  // each lane's share of the row sums in float (as the reference's float chains), the lanes' in double
  float s0f = 0, x0f = 0, q0f = 0, s1f = 0, x1f = 0, q1f = 0;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int x = lane + 64 * m;
    float a, b;
    if constexpr (WARP) {
      const int ci = P.mcell[x];
      const float2 q0 = warp_point(P, row0.F, row0.r, x, cterm[w][0][ci]);
      const float2 q1 = warp_point(P, row1.F, row1.r, x, cterm[w][1][ci]);
      a = sample(row0.F, q0.x, q0.y);
      b = sample(row1.F, q1.x, q1.y);
    } else {
      a = src.at(P, row0, x);
      b = src.at(P, row1, x);
    }
    if ((yin0 || yin1) && in_square(P, x)) {
      if (yin0) a = hole_value(P, bsh, x, y);
      if (yin1) b = hole_value(P, bsh, x, y + 1);
    }
    v[m] = make_float2(a, b);
    const float cx = P.cx[x], qx = P.qx[x];
    s0f += a;
    x0f += a * cx;
    q0f += a * qx;
    s1f += b;
    x1f += b * cx;
    q1f += b * qx;
  }
  double s0 = s0f, x0 = x0f, q0 = q0f, s1 = s1f, x1 = x1f, q1 = q1f;
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o, 64);
    x0 += __shfl_xor(x0, o, 64);
    q0 += __shfl_xor(q0, o, 64);
    s1 += __shfl_xor(s1, o, 64);
    x1 += __shfl_xor(x1, o, 64);
    q1 += __shfl_xor(q1, o, 64);
  }
  if (lane == 0) {
    double *o = rowsum + (static_cast<size_t>(f) * N + y) * 3;
    o[0] = s0;
    o[1] = x0;
    o[2] = q0;
    o[3] = s1;
    o[4] = x1;
    o[5] = q1;
  }
  float2 *S = zs + w * F::LDS;
  F::run(v, S, P.tw, lane);
  const int E = P.E;
  for (int u = lane; u < E; u += 64) {
    const float2 zu = S[F::pos(u)], zm = S[F::pos(u ? N - u : 0)];
    const float4 o = make_float4(0.5f * (zu.x + zm.x), 0.5f * (zu.y - zm.y), 0.5f * (zu.y + zm.y), -0.5f * (zu.x - zm.x));
    *reinterpret_cast<float4 *>(X + (static_cast<size_t>(f) * E + u) * N + y) = o;
  }
}

// Column u (one wave): the FFT along y, then every coefficient on this column straight into coefficient order:
// coefficient (u, v) is Y(u, v), coefficient (-u, v) is conj Y(u, n - v) (the picture is real).
template <int N, int WV, int R1>
__global__ __launch_bounds__(64 * WV) void k_cols(PlanDev P, const float2 *X, float2 *C) {
  using F = WaveFft<N, R1>;
  constexpr int M = F::M;
  extern __shared__ float2 zs[];  // WV waves x F::LDS
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, E = P.E;
  int bx = blockIdx.x, f = blockIdx.y;
  xcd_frame(bx, f);
  const int u = bx * WV + w;
  if (u >= E) return;  // the whole wave; nothing below waits on the workgroup
  const float2 *col = X + (static_cast<size_t>(f) * E + u) * N;
  float2 v[M];
#pragma unroll
  for (int m = 0; m < M; m++) v[m] = col[lane + 64 * m];
  float2 *S = zs + w * F::LDS;
  F::run(v, S, P.tw, lane);
  float2 *out = C + static_cast<size_t>(f) * P.subch * SUB;
  for (int e = P.col_off[u] + lane; e < P.col_off[u + 1]; e += 64) {
    const uint32_t q = static_cast<uint32_t>(P.col_ent[e]);
    float2 z = S[F::pos(static_cast<int>((q >> 19) & 0x7FF))];
    if (q >> 30) z.y = -z.y;
    out[q & 0x7FFFF] = z;
  }
}

// ------------------------------------------------------------------------------------------------ the coefficients
struct Proj { double px, py, pxy, pxx, pyy; };

// Coefficient ci = (u, v) after the detrend (lizard_phy.c:1697-1716; SPEC 6.10), in the reference's float order; C
// holds the frame's raw coefficients in coefficient order.
__device__ __forceinline__ float2 coef_at(const PlanDev &P, const float2 *C, const Proj &pr, int ci, int u, int v) {
  const int n = P.n;
  float2 y = C[ci];
  const int uw = u < 0 ? u + n : u;
  const float ar = static_cast<float>(__dmul_rn(pr.pxy, P.x1r[uw])), ai = static_cast<float>(__dmul_rn(pr.pxy, P.x1i[uw]));
  const float r = P.x1r[v], i = P.x1i[v];
  y.x = __fsub_rn(y.x, __fsub_rn(__fmul_rn(ar, r), __fmul_rn(ai, i)));
  y.y = __fsub_rn(y.y, __fadd_rn(__fmul_rn(ar, i), __fmul_rn(ai, r)));
  const double dn = n;
  if (v == 0) {
    y.x = __fsub_rn(y.x, static_cast<float>(__dmul_rn(dn, __dadd_rn(__dmul_rn(pr.px, P.x1r[uw]), __dmul_rn(pr.pxx, P.x2r[uw])))));
    y.y = __fsub_rn(y.y, static_cast<float>(__dmul_rn(dn, __dadd_rn(__dmul_rn(pr.px, P.x1i[uw]), __dmul_rn(pr.pxx, P.x2i[uw])))));
  }
  if (uw == 0) {
    y.x = __fsub_rn(y.x, static_cast<float>(__dmul_rn(dn, __dadd_rn(__dmul_rn(pr.py, P.x1r[v]), __dmul_rn(pr.pyy, P.x2r[v])))));
    y.y = __fsub_rn(y.y, static_cast<float>(__dmul_rn(dn, __dadd_rn(__dmul_rn(pr.py, P.x1i[v]), __dmul_rn(pr.pyy, P.x2i[v])))));
  }
  return y;
}

// Coefficient (u, v) turned back by the grid shift (dx, dy) through the 4096-step table, index rounded half-even
// (lizard_phy.c:214-218).
__device__ __forceinline__ float2 turn(const PlanDev &P, float2 y, int u, int v, float dx, float dy) {
  const float a = __fmul_rn(__fadd_rn(__fmul_rn(static_cast<float>(u), dx), __fmul_rn(static_cast<float>(v), dy)),
                            static_cast<float>(ALIGN_TURNS) / static_cast<float>(P.n));
  const unsigned k = static_cast<unsigned>(static_cast<int>(rintf(a))) & (ALIGN_TURNS - 1);
  const float2 cs = P.rot[k];
  return make_float2(__fadd_rn(__fmul_rn(y.x, cs.x), __fmul_rn(y.y, cs.y)),
                     __fsub_rn(__fmul_rn(y.y, cs.x), __fmul_rn(y.x, cs.y)));
}

__device__ __forceinline__ void coef_uv(const PlanDev &P, int i, int *u, int *v) {
  const int32_t c = P.coef[i];
  *u = static_cast<int16_t>(c & 0xFFFF);
  *v = c >> 16;
}

// The projections from the row sums, then the pilots' grid shift (lizard_phy.c:280-329, :1697, :1739).
__global__ __launch_bounds__(256) void k_align(PlanDev P, const double *rowsum, const float2 *C_all, double *proj_out,
                                               float *align_out) {
  __shared__ double red[6 * 256];
  __shared__ double s_d[2];
  __shared__ int s_go;
  const int tid = threadIdx.x, f = blockIdx.x, n = P.n;
  const double *rsum = rowsum + static_cast<size_t>(f) * n * 3;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
  for (int y = tid; y < n; y += 256) {
    const double rs = rsum[3 * y], rx = rsum[3 * y + 1], rq = rsum[3 * y + 2];
    a0 += rx;
    a1 += rs * P.cx[y];
    a2 += rx * P.cx[y];
    a3 += rq;
    a4 += rs * P.qx[y];
  }
  Proj pr;
  double sums[5] = {a0, a1, a2, a3, a4};
  block_sums<256, 5>(sums, red, tid);
  pr.px = sums[0] / (n * P.sx2);
  pr.py = sums[1] / (n * P.sx2);
  pr.pxy = sums[2] / (P.sx2 * P.sx2);
  pr.pxx = sums[3] / (n * P.sq2);
  pr.pyy = sums[4] / (n * P.sq2);
  if (tid == 0) {
    double *o = proj_out + f * 5;
    o[0] = pr.px;
    o[1] = pr.py;
    o[2] = pr.pxy;
    o[3] = pr.pxx;
    o[4] = pr.pyy;
    s_d[0] = s_d[1] = 0;
    s_go = 1;
  }
  __syncthreads();
  const float2 *Cf = C_all + static_cast<size_t>(f) * P.subch * SUB;
  const double k = 2.0 * 3.14159265358979323846 / n;
  double A = 0, B = 0, C = 0, Pq = 0, Q = 0, YY = 0;
  for (int it = 0; it < 2; it++) {
    if (!s_go) break;
    const float dx = static_cast<float>(s_d[0]), dy = static_cast<float>(s_d[1]);
    double a = 0, b = 0, c = 0, p = 0, q = 0, yy = 0;
    for (int t = tid; t < P.npilots; t += 256) {
      const int ci = P.pilot[t];
      int u, v;
      coef_uv(P, ci, &u, &v);
      const float2 y = turn(P, coef_at(P, Cf, pr, ci, u, v), u, v, dx, dy);
      const double xr = P.white[2 * ci] ? -1.0 : 1.0, xi = P.white[2 * ci + 1] ? -1.0 : 1.0;
      const double zr = y.x * xr + y.y * xi, zi = y.y * xr - y.x * xi;
      const double cu = k * zr * u, cv = k * zr * v;
      a += cu * cu;
      b += cu * cv;
      c += cv * cv;
      p += cu * zi;
      q += cv * zi;
      yy += zi * zi;
    }
    double six[6] = {a, b, c, p, q, yy};
    block_sums<256, 6>(six, red, tid);
    A = six[0];
    B = six[1];
    C = six[2];
    Pq = six[3];
    Q = six[4];
    YY = six[5];
    if (tid == 0) {
      const double det = A * C - B * B;
      if (P.npilots < 8 || !(det > 0)) {
        s_go = 0;
        s_d[0] = s_d[1] = 0;
      } else {
        s_d[0] += (Pq * C - Q * B) / det;
        s_d[1] += (A * Q - B * Pq) / det;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    float d0 = 0, d1 = 0;
    if (s_go) {
      const double det = A * C - B * B, ex = (Pq * C - Q * B) / det, ey = (A * Q - B * Pq) / det;
      const double s2 = (YY - ex * Pq - ey * Q) / (P.npilots - 2);
      const double dx = s_d[0], dy = s_d[1];
      const double chi = s2 > 0 ? (dx * dx * A + 2 * dx * dy * B + dy * dy * C) / s2 : 0;
      if (chi > 13.8 && dx == dx && dy == dy) {
        d0 = static_cast<float>(dx);
        d1 = static_cast<float>(dy);
      }
    }
    align_out[2 * f] = d0;
    align_out[2 * f + 1] = d1;
  }
}

// J at sigma = 2 sqrt(g) (ten Brink): the information a soft QPSK axis carries (lizard_phy.c:75).
__device__ __forceinline__ double axis_info(double g) {
  return g > 0 ? pow(1 - pow(2.0, -0.3073 * pow(2 * sqrt(g), 1.7870)), 1.1064) : 0;
}

// Sub-channel j of frame f (lizard_phy.c:1756-1772, :1322-1332; SPEC 8.5), one wave a sub-channel (WV a
// workgroup): each lane's five coefficients gathered, detrended and turned, the moments m2, m4 summed in double
// across the wave by shuffles (every lane ends with the totals, so the scale needs no broadcast and no barrier).
// Equalisation by a known Gaussian blur of sigma picture samples (an extension; the reference does not equalise, its
// per-sub-channel scale absorbing the blur, research/SPEC.md 6.10): coefficient (u, v) divided by the blur's OTF
// exp(-2 pi^2 sigma^2 (u^2 + v^2) / n^2), after the pilot turn. otf_rate is 2 pi^2 sigma^2 / n^2 for frame f (0: none).
__device__ __forceinline__ float otf_rate(const PlanDev &P, const float *otf, int f) {
  const float sg = otf ? otf[f] : 0.0f;
  return sg > 0 ? static_cast<float>(2.0 * kPi * kPi) * sg * sg / (static_cast<float>(P.n) * static_cast<float>(P.n))
                : 0.0f;
}
__device__ __forceinline__ float2 equalise(float2 c, float gk, int u, int v) {
  const float g = expf(gk * static_cast<float>(u * u + v * v));
  return make_float2(c.x * g, c.y * g);
}

template <int WV>
__global__ __launch_bounds__(64 * WV) void k_demap(PlanDev P, const float2 *C_all, const double *proj,
                                                   const float *align, int8_t *llr_all, double *est_all, float *rms_all,
                                                   float *dbg_all, const float *otf) {
  const int lane = threadIdx.x & 63, j = blockIdx.x * WV + (threadIdx.x >> 6), f = blockIdx.y;
  if (j >= P.subch) return;  // the whole wave
  const float gk = otf_rate(P, otf, f);
  const float2 *C = C_all + static_cast<size_t>(f) * P.subch * SUB;
  const Proj pr{proj[f * 5], proj[f * 5 + 1], proj[f * 5 + 2], proj[f * 5 + 3], proj[f * 5 + 4]};
  const float dx = align[2 * f], dy = align[2 * f + 1];
  constexpr int PER = SUB / 64;
  float2 y[PER];
  double s2 = 0, s4 = 0;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int ci = j * SUB + lane + 64 * q;
    int u, v;
    coef_uv(P, ci, &u, &v);
    float2 c = coef_at(P, C, pr, ci, u, v);
    if (dbg_all) reinterpret_cast<float2 *>(dbg_all)[static_cast<size_t>(f) * 2 * P.subch * SUB + ci] = c;
    if (dx != 0 || dy != 0) c = turn(P, c, u, v, dx, dy);
    if (gk > 0) c = equalise(c, gk, u, v);
    y[q] = c;
    const double e = static_cast<double>(c.x) * c.x + static_cast<double>(c.y) * c.y;
    s2 += e;
    s4 += e * e;
  }
  for (int o = 32; o > 0; o >>= 1) {
    s2 += __shfl_xor(s2, o, 64);
    s4 += __shfl_xor(s4, o, 64);
  }
  const double m2 = s2 / SUB, m4 = s4 / SUB;
  double a2 = 2 * m2 * m2 - m4;
  a2 = a2 > 0 ? sqrt(a2) : 0;
  double nv = m2 - a2;
  if (nv < 0.02 * m2) nv = 0.02 * m2;
  const float k = static_cast<float>(0.7 * 2 * sqrt(2 * a2) / nv);
  if (lane == 0) {
    rms_all[f * P.subch + j] = m2 > 0 ? static_cast<float>(sqrt(m2 / 2)) : 0.f;
    // the estimate's argument a2 / nv; est_in_place turns it into J (axis_info), many sub-channels a wave, so this
    // wave does not spend its issue on one lane's double pow chain
    est_all[f * P.subch + j] = a2 / nv;
  }
  int8_t *o = llr_all + static_cast<size_t>(f) * P.subch * 2 * SUB + 2 * j * SUB;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    float lr = __fmul_rn(k, y[q].x), li = __fmul_rn(k, y[q].y);
    if (dbg_all)
      reinterpret_cast<float2 *>(dbg_all)[static_cast<size_t>(f) * 2 * P.subch * SUB + P.subch * SUB + j * SUB + lane +
                                          64 * q] = make_float2(lr, li);
    lr = lr > 10.f ? 10.f : (lr < -10.f ? -10.f : lr);
    li = li > 10.f ? 10.f : (li < -10.f ? -10.f : li);
    const int at = 2 * (lane + 64 * q);
    o[at] = lr == lr ? static_cast<int8_t>(rintf(__fmul_rn(lr, 8.f))) : 0;
    o[at + 1] = li == li ? static_cast<int8_t>(rintf(__fmul_rn(li, 8.f))) : 0;
  }
}

// Block b's pilot reading (lizard_phy.c:1795-1809; SPEC 7.3): over the slots past its codeword, the axis value
// against the whitening's sign in units of its sub-channel's rms axis value, averaged over the tail; NaN where the
// block has no tail or no sub-channel of it read anything. One wave a block (tails are 16 or 32 slots).
// est[j] = axis_info(est[j]) in place: the sub-channel estimates from the arguments k_demap left (lizard_phy.c:75,
// :1771), one lane a sub-channel.
__device__ __forceinline__ void est_in_place(double *est, int j) { est[j] = axis_info(est[j]); }

__global__ __launch_bounds__(256) void k_est(double *est, int count) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < count) est_in_place(est, j);
}

// The pilot readings of block b of frame f; its wave also turns the frame's sub-channels b, b + blocks, ... into
// their estimates (the single-frame path's est step, without a launch of its own).
__global__ __launch_bounds__(64) void k_pilot_read(PlanDev P, const float2 *C_all, const double *proj,
                                                   const float *align, const float *rms_all, float *out,
                                                   const float *otf, double *est_all) {
  const int b = blockIdx.x, f = blockIdx.y, lane = threadIdx.x;
  for (int j = b + P.blocks * lane; j < P.subch; j += P.blocks * 64) est_in_place(est_all + f * P.subch, j);
  const float gk = otf_rate(P, otf, f);
  const int4 bk = P.blk[b];
  const int s0 = bk.x, nt = bk.y, nb = bk.z;
  const float2 *C = C_all + static_cast<size_t>(f) * P.subch * SUB;
  const Proj pr{proj[f * 5], proj[f * 5 + 1], proj[f * 5 + 2], proj[f * 5 + 3], proj[f * 5 + 4]};
  const float dx = align[2 * f], dy = align[2 * f + 1];
  double cs = 0;
  int got = 0;
  for (int i = nt + lane; i < nb; i += 64) {
    const int gs = s0 + i;
    const float rms = rms_all[f * P.subch + gs / (2 * SUB)];
    if (!(rms > 0)) continue;
    int u, v;
    coef_uv(P, gs >> 1, &u, &v);
    float2 y = coef_at(P, C, pr, gs >> 1, u, v);
    if (dx != 0 || dy != 0) y = turn(P, y, u, v, dx, dy);
    if (gk > 0) y = equalise(y, gk, u, v);
    const float a = (i & 1) ? y.y : y.x;
    cs += (P.white[gs] ? -1.0 : 1.0) * a / rms;
    got++;
  }
  for (int o = 32; o > 0; o >>= 1) {
    cs += __shfl_xor(cs, o, 64);
    got += __shfl_xor(got, o, 64);
  }
  if (lane == 0) out[f * P.blocks + b] = got ? static_cast<float>(cs / (nb - nt)) : __builtin_nanf("");
}

// ------------------------------------------------------------------------------------------------ host: plans
// 6 x 6 inverse by Gauss-Jordan with partial pivots.
static void invert6(double g[6][12]) {
  for (int c = 0; c < 6; c++) {
    int p = c;
    for (int r = c + 1; r < 6; r++)
      if (std::fabs(g[r][c]) > std::fabs(g[p][c])) p = r;
    if (std::fabs(g[p][c]) < 1e-12) throw Error(LIZ_E_INTERNAL, "singular shading fit");
    for (int k = 0; k < 12; k++) std::swap(g[c][k], g[p][k]);
    const double d = g[c][c];
    for (int k = 0; k < 12; k++) g[c][k] /= d;
    for (int r = 0; r < 6; r++)
      if (r != c) {
        const double m = g[r][c];
        for (int k = 0; k < 12; k++) g[r][k] -= m * g[c][k];
      }
  }
}

std::unique_ptr<FramePlan> make_plan(int version, int span) {
  auto p = std::make_unique<FramePlan>();
  const Format *f = format_for(version);
  if (!f) throw Error(LIZ_E_FORMAT, "no format at that version (4 to 64)");
  if (span != 128 && span != 256 && span != 512) throw Error(LIZ_E_ARG, "span is a ring's: 128, 256 or 512 modules");
  const int n = f->n, S = f->subch * SUB;
  p->version = version;
  p->span = span;
  p->n = n;
  p->subch = f->subch;
  p->fmt = f;
  PlanDev &d = p->dev;
  d.n = n;
  d.subch = f->subch;
  d.blocks = f->blocks;
  // the detrend basis and its transforms (lizard_phy.c:388-399; SPEC 6.10)
  std::vector<float> tab(6 * n);
  float *cx = tab.data(), *qx = cx + n, *x1r = qx + n, *x1i = x1r + n, *x2r = x1i + n, *x2i = x2r + n;
  double m2 = 0;
  std::vector<double> ec(n), es(n);
  for (int i = 0; i < n; i++) {
    const double c = (2.0 * i + 1 - n) / n;
    m2 += c * c;
    ec[i] = std::cos(-2 * 3.14159265358979323846 * i / n);
    es[i] = std::sin(-2 * 3.14159265358979323846 * i / n);
  }
  d.sx2 = m2;
  m2 /= n;
  d.sq2 = 0;
  for (int i = 0; i < n; i++) {
    const double c = (2.0 * i + 1 - n) / n, q = c * c - m2;
    cx[i] = static_cast<float>(c);
    qx[i] = static_cast<float>(q);
    d.sq2 += q * q;
  }
  for (int u = 0; u < n; u++) {
    double ar = 0, ai = 0, br = 0, bi = 0;
    for (int i = 0; i < n; i++) {
      const double c = (2.0 * i + 1 - n) / n, q = c * c - m2;
      const int k = static_cast<int>((static_cast<long long>(u) * i) % n);
      ar += c * ec[k];
      ai += c * es[k];
      br += q * ec[k];
      bi += q * es[k];
    }
    x1r[u] = static_cast<float>(ar);
    x1i[u] = static_cast<float>(ai);
    x2r[u] = static_cast<float>(br);
    x2i[u] = static_cast<float>(bi);
  }
  // anchors as erased (SPEC 6.13; lizard_phy.c:402-407, acquire.c:41-45)
  const float scale = static_cast<float>(n) / span;
  const int a = span / 64;
  for (int i = 0; i < ANCH_K; i++) {
    const int m0 = static_cast<int>(std::floor(BORDER + static_cast<double>(span) * (i + 0.5) / ANCH_K - a / 2.0 + 0.5));
    const int lo = static_cast<int>(std::floor(static_cast<float>(m0 - BORDER) * scale)) - ANCH_MARGIN;
    const int hi = static_cast<int>(std::ceil(static_cast<float>(m0 + a - BORDER) * scale)) - 1 + ANCH_MARGIN;
    d.elo[i] = lo < 0 ? 0 : lo;
    d.ehi[i] = hi > n - 1 ? n - 1 : hi;
  }
  auto in_sq = [&](int t) {
    for (int i = 0; i < ANCH_K; i++)
      if (t >= d.elo[i] && t <= d.ehi[i]) return true;
    return false;
  };
  double g[6][12] = {{0}};
  for (int y = 0; y < n; y += 8)
    for (int x = 0; x < n; x += 8) {
      if (in_sq(x) && in_sq(y)) continue;
      const double b[6] = {1, cx[x], qx[x], cx[y], qx[y], static_cast<double>(cx[x]) * cx[y]};
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) g[r][c] += b[r] * b[c];
    }
  for (int r = 0; r < 6; r++) g[r][6 + r] = 1;
  invert6(g);
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) d.hole_inv[r * 6 + c] = g[r][6 + c];
  // the coefficients' (u, v) and the half-plane's reach
  std::vector<int32_t> coef(S);
  const int bw = f->bw;
  int E = 0;
  for (int i = 0; i < S; i++) {
    const int pp = f->pos[i], v = (pp / (n * bw)) * bw + pp % bw, uw = (pp / bw) % n, u = uw < n / 2 ? uw : uw - n;
    coef[i] = static_cast<int32_t>((static_cast<uint32_t>(v) << 16) | (static_cast<uint32_t>(u) & 0xFFFFu));
    E = std::max(E, std::max(std::abs(u), v) + 1);
  }
  p->E = d.E = E;
  // each column's coefficients for the column pass: coefficient (u, v) from column |u|, at v (u >= 0) or conjugated
  // at n - v (u < 0); entry = ci | source index << 19 | conjugate << 30
  {
    std::vector<int32_t> off(E + 1, 0), ent(S);
    for (int i = 0; i < S; i++) {
      const int u = static_cast<int16_t>(coef[i] & 0xFFFF);
      off[std::abs(u) + 1]++;
    }
    for (int c = 0; c < E; c++) off[c + 1] += off[c];
    std::vector<int32_t> at(off.begin(), off.end() - 1);
    for (int i = 0; i < S; i++) {
      const int u = static_cast<int16_t>(coef[i] & 0xFFFF), v = coef[i] >> 16;
      const int k = u >= 0 ? v : (n - v) % n;
      ent[at[std::abs(u)]++] = static_cast<int32_t>(static_cast<uint32_t>(i) | (static_cast<uint32_t>(k) << 19) |
                                                    (static_cast<uint32_t>(u < 0) << 30));
    }
    if (S >= (1 << 19) || n > 2048) throw Error(LIZ_E_INTERNAL, "coefficient table packing");
    p->cols.reserve((off.size() + ent.size()) * sizeof(int32_t));
    copy_wait(p->cols.p, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, aux_stream(), "columns");
    copy_wait(static_cast<int32_t *>(p->cols.p) + off.size(), ent.data(), ent.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice, aux_stream(), "columns");
    d.col_off = p->cols.as<int32_t>();
    d.col_ent = p->cols.as<int32_t>() + off.size();
  }
  // the pilots: the tail coefficients of every 7/8 and 3/4 block (SPEC 7.3; lizard_phy.c:293-317)
  for (int b = 0; b < f->blocks; b++) {
    const Code &c = code(f->blk_code[b]);
    const int nb = f->blk_nslots[b], first = f->blk_slot[b] / 2;
    if (c.nt >= nb) continue;
    for (int s = c.nt / 2; s < nb / 2; s++) {
      p->pilot_host.push_back(first + s);
      p->pilot_block.push_back(b);
    }
  }
  d.npilots = static_cast<int>(p->pilot_host.size());
  std::vector<int4> blk(f->blocks);
  for (int b = 0; b < f->blocks; b++)
    blk[b] = make_int4(f->blk_slot[b], code(f->blk_code[b]).nt, f->blk_nslots[b], f->blk_code[b]);
  p->blk.reserve(blk.size() * sizeof(int4));
  copy_wait(p->blk.p, blk.data(), blk.size() * sizeof(int4), hipMemcpyHostToDevice, aux_stream(), "blocks");
  d.blk = p->blk.as<int4>();
  // the warp's sample grid: sample t at module g0 + t step (lizard_phy.c:1592; SPEC 6.9), its node cell found as the
  // reference's walk finds it (acquire.c:5111-5119)
  d.nx = d.ny = node_lattice(span, d.node_x);
  for (int k = 0; k < d.ny; k++) d.node_y[k] = d.node_x[k];
  for (int k = 0; k + 1 < d.nx; k++) {
    const volatile float span_k = d.node_x[k + 1] - d.node_x[k];
    d.node_inv[k] = 1.0f / span_k;
  }
  std::vector<float> mc(n);
  std::vector<int32_t> cell(n);
  {
    const volatile float g0 = static_cast<float>(BORDER) + 0.5f / scale, step = 1.0f / scale;
    for (int t = 0; t < n; t++) {
      const volatile float prod = static_cast<float>(t) * step;
      const float m = g0 + prod;
      int i = 0;
      while (i < d.nx - 2 && m >= d.node_x[i + 1]) i++;
      mc[t] = m;
      cell[t] = i;
    }
    for (int i = 0; i < d.nx; i++) {
      int t = 0;
      while (t < n && cell[t] < i) t++;
      d.cell0[i] = i == d.nx - 1 ? n : t;
    }
  }
  with_size(n, [](auto) {});  // a size the transforms are built for, or LIZ_E_FORMAT
  std::vector<float2> tw(n), rot(ALIGN_TURNS);
  for (int k = 0; k < n; k++) {
    const double th = -2 * 3.14159265358979323846 * k / n;
    tw[k] = make_float2(static_cast<float>(std::cos(th)), static_cast<float>(std::sin(th)));
  }
  for (int i = 0; i < ALIGN_TURNS; i++) {
    const double th = 2.0 * 3.14159265358979323846 * i / ALIGN_TURNS;
    rot[i] = make_float2(static_cast<float>(std::cos(th)), static_cast<float>(std::sin(th)));
  }
  // upload
  p->tables.reserve(tab.size() * sizeof(float));
  copy_wait(p->tables.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, aux_stream(), "plan tables");
  p->coef.reserve(coef.size() * sizeof(int32_t));
  copy_wait(p->coef.p, coef.data(), coef.size() * sizeof(int32_t), hipMemcpyHostToDevice, aux_stream(), "coefficients");
  p->pilot.reserve(std::max<size_t>(1, p->pilot_host.size()) * sizeof(int32_t));
  if (!p->pilot_host.empty())
    copy_wait(p->pilot.p, p->pilot_host.data(), p->pilot_host.size() * sizeof(int32_t), hipMemcpyHostToDevice, aux_stream(),
              "pilots");
  p->tw.reserve(n * sizeof(float2));
  copy_wait(p->tw.p, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice, aux_stream(), "twiddles");
  p->rot.reserve(ALIGN_TURNS * sizeof(float2));
  copy_wait(p->rot.p, rot.data(), ALIGN_TURNS * sizeof(float2), hipMemcpyHostToDevice, aux_stream(), "turn table");
  p->mtab.reserve(n * (sizeof(float) + sizeof(int32_t)));
  copy_wait(p->mtab.p, mc.data(), n * sizeof(float), hipMemcpyHostToDevice, aux_stream(), "module coordinates");
  copy_wait(static_cast<char *>(p->mtab.p) + n * sizeof(float), cell.data(), n * sizeof(int32_t),
                      hipMemcpyHostToDevice, aux_stream(),
            "node cells");
  d.mcoord = p->mtab.as<float>();
  d.mcell = reinterpret_cast<const int32_t *>(static_cast<char *>(p->mtab.p) + n * sizeof(float));
  p->white.reserve(f->white.size());
  copy_wait(p->white.p, f->white.data(), f->white.size(), hipMemcpyHostToDevice, aux_stream(), "whitening");
  const float *tb = p->tables.as<float>();
  d.cx = tb;
  d.qx = tb + n;
  d.x1r = tb + 2 * n;
  d.x1i = tb + 3 * n;
  d.x2r = tb + 4 * n;
  d.x2i = tb + 5 * n;
  d.tw = p->tw.as<float2>();
  d.rot = p->rot.as<float2>();
  d.coef = p->coef.as<int32_t>();
  d.pilot = p->pilot.as<int32_t>();
  d.white = p->white.as<uint8_t>();
  return p;
}

void FrameBufs::reserve(const FramePlan &p, int frames, bool debug) {
  const size_t n = p.n, E = p.E, S = static_cast<size_t>(p.subch) * SUB;
  X.reserve(frames * E * n * sizeof(float2));
  C.reserve(frames * S * sizeof(float2));
  rowsum.reserve(frames * n * 3 * sizeof(double));
  beta.reserve(frames * FIT_STRIDE * sizeof(double));
  proj.reserve(frames * 5 * sizeof(double));
  align.reserve(frames * 2 * sizeof(float));
  llr.reserve(frames * 2 * S);
  est.reserve(frames * p.subch * sizeof(double));
  rms.reserve(frames * p.subch * sizeof(float));
  if (debug) coef_dbg.reserve(frames * 2 * S * sizeof(float2));  // the coefficients, then the soft floats
  pilot_r.reserve(frames * p.dev.blocks * sizeof(float));
}

SoftBufs FrameBufs::view(bool debug) {
  SoftBufs v;
  v.beta = beta.as<double>();
  v.rowsum = rowsum.as<double>();
  v.proj = proj.as<double>();
  v.X = X.as<float2>();
  v.C = C.as<float2>();
  v.align = align.as<float>();
  v.llr = llr.as<int8_t>();
  v.est = est.as<double>();
  v.rms = rms.as<float>();
  v.coef_dbg = debug ? coef_dbg.as<float>() : nullptr;
  v.blk_pilot = pilot_r.p ? pilot_r.as<float>() : nullptr;
  v.otf = has_otf ? otf.as<float>() : nullptr;
  return v;
}

void FrameBufs::set_otf(const float *sigma, int frames, hipStream_t s) {
  bool any = false;
  for (int f = 0; sigma && f < frames; f++) {
    if (!(sigma[f] >= 0.0f && sigma[f] <= 8.0f)) throw Error(LIZ_E_ARG, "a blur sigma of 0 to 8 picture samples");
    any |= sigma[f] > 0.0f;
  }
  has_otf = any;
  if (!any) return;
  otf.reserve(static_cast<size_t>(frames) * sizeof(float));
  copy_wait(otf.p, sigma, static_cast<size_t>(frames) * sizeof(float), hipMemcpyHostToDevice, s, "blur sigma");
}

SoftBufs SoftBufs::at(const FramePlan &p, int f) const {
  const size_t n = p.n, E = p.E, S = static_cast<size_t>(p.subch) * SUB, F = f;
  SoftBufs v;
  v.beta = beta + F * FIT_STRIDE;
  v.rowsum = rowsum + F * n * 3;
  v.proj = proj + F * 5;
  v.X = X + F * E * n;
  v.C = C + F * S;
  v.align = align + F * 2;
  v.llr = llr + F * 2 * S;
  v.est = est + F * p.subch;
  v.rms = rms + F * p.subch;
  v.coef_dbg = coef_dbg ? coef_dbg + F * 4 * S : nullptr;
  v.blk_pilot = blk_pilot ? blk_pilot + F * p.dev.blocks : nullptr;
  v.otf = otf ? otf + F : nullptr;
  return v;
}

// Calls f with std::integral_constant<int, w> for the waves-per-workgroup choices the launches are built for.
template <class F>
void with_waves(int w, F &&f) {
  switch (w) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    default: throw Error(LIZ_E_ARG, "waves per workgroup: 1, 2 or 4");
  }
}

// The wave FFT's radix plan for its 64-point stage (R1 x 64 / R1).
template <class F>
void with_radix(int r, F &&f) {
  switch (r) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    default: throw Error(LIZ_E_ARG, "radix plan: 4, 8 or 16");
  }
}

template <class Src>
static void launch_soft_src(const FramePlan &p, const SoftBufs &b, Src src, int frames, hipStream_t s,
                            const LaunchShape &sh) {
  const int n = p.n, E = p.E;
  hipLaunchKernelGGL(k_erase_fit<Src>, dim3(n / 8 / FIT_ROWS, frames), dim3(256), 0, s, p.dev, src, b.beta,
                     b.zero_int);
  with_size(n, [&](auto N) {
    constexpr int NN = decltype(N)::value;
    with_waves(sh.rows, [&](auto Wr) {
      constexpr int W = decltype(Wr)::value;
      with_radix(sh.rows_radix, [&](auto Rr) {
        constexpr int R = decltype(Rr)::value;
        hipLaunchKernelGGL((k_rows<Src, NN, W, R>), dim3(NN / (2 * W), frames), dim3(64 * W),
                           W * WaveFft<NN>::LDS * sizeof(float2), s, p.dev, src, b.beta, b.X, b.rowsum);
      });
    });
    with_waves(sh.cols, [&](auto Wc) {
      constexpr int W = decltype(Wc)::value;
      with_radix(sh.cols_radix, [&](auto Rc) {
        constexpr int R = decltype(Rc)::value;
        hipLaunchKernelGGL((k_cols<NN, W, R>), dim3((E + W - 1) / W, frames), dim3(64 * W),
                           W * WaveFft<NN>::LDS * sizeof(float2), s, p.dev, b.X, b.C);
      });
    });
  });
  hipLaunchKernelGGL(k_align, dim3(frames), dim3(256), 0, s, p.dev, b.rowsum, b.C, b.proj, b.align);
  with_waves(sh.demap, [&](auto Wd) {
    constexpr int W = decltype(Wd)::value;
    hipLaunchKernelGGL(k_demap<W>, dim3((p.subch + W - 1) / W, frames), dim3(64 * W), 0, s, p.dev, b.C, b.proj,
                       b.align, b.llr, b.est, b.rms, b.coef_dbg, b.otf);
  });
  if (b.blk_pilot)
    hipLaunchKernelGGL(k_pilot_read, dim3(p.dev.blocks, frames), dim3(64), 0, s, p.dev, b.C, b.proj, b.align, b.rms,
                       b.blk_pilot, b.otf, b.est);
  else
    hipLaunchKernelGGL(k_est, dim3((frames * p.subch + 255) / 256), dim3(256), 0, s, b.est, frames * p.subch);
  hip_check(hipGetLastError(), "soft-value launches");
}

void launch_soft(const FramePlan &p, const SoftBufs &b, const float *d_grid, int frames, hipStream_t s) {
  launch_soft_src(p, b, GridSource{d_grid}, frames, s, LaunchShape{});
}

void launch_soft_grids(const FramePlan &p, const SoftBufs &b, const float *const *d_grids, int frames, hipStream_t s,
                       const LaunchShape &sh) {
  launch_soft_src(p, b, GridListSource{d_grids}, frames, s, sh);
}

void launch_soft_warp(const FramePlan &p, const SoftBufs &b, const WarpFrame *d_frames, int frames, hipStream_t s,
                      const LaunchShape &sh) {
  launch_soft_src(p, b, WarpSource{d_frames}, frames, s, sh);
}

void launch_warp(const FramePlan &p, const WarpFrame *d_frames, int frames, float *d_grid, hipStream_t s) {
  const int n = p.n;
  hipLaunchKernelGGL(k_warp, dim3((n * n + 255) / 256, frames), dim3(256), 0, s, p.dev, d_frames, d_grid);
  hip_check(hipGetLastError(), "warp launch");
}

// ------------------------------------------------------------------------------------------------ the sender's picture
// Frame slots to the n x n picture (research/SPEC.md 6.5-6.7; the picture liblizard/src/lizard_phy.c:1087-1200
// resamples into its drive): QPSK (+-1 +- i) / sqrt 2 at each coefficient (u, v), the Hermitian spectrum, the
// unscaled inverse DFT exp(+2 pi i (u x + v y) / n) as conj(DFT(conj)) through the same wave FFT, clipped at
// L = 2 sqrt(640 subch) into [0, 1]. Two passes, the receive path's in reverse order:
//   k_mod_cols  one column u < E a wave: the column's coefficients (and, for u = 0, their Hermitian partners) set in
//               the wave's LDS, the inverse DFT along v, G[frame][u][y] column-major
//   k_mod_rows  two rows a wave: each row's spectrum G[u][y] (u < E) with conj G[u][y] at n - u, the two rows as one
//               complex sequence, the inverse DFT along u, row y the real part and row y + 1 the imaginary

// Frame slot pair (2 ci, 2 ci + 1) as its QPSK point.
__device__ __forceinline__ float2 qpsk(const uint8_t *bits, int ci) {
  constexpr float a = 0.70710678118654752f;
  return make_float2(bits[2 * ci] ? -a : a, bits[2 * ci + 1] ? -a : a);
}

// COEF: the column's values are the coefficients coef[ci] where mask[ci] (else 0), one frame, instead of QPSK points
// from frame slots (the spectral loss's adjoint).
template <int N, int WV, bool COEF = false>
__global__ __launch_bounds__(64 * WV) void k_mod_cols(PlanDev P, const uint8_t *bits_all, float2 *G,
                                                      const float2 *coef = nullptr, const uint8_t *mask = nullptr) {
  using F = WaveFft<N>;
  constexpr int M = F::M;
  extern __shared__ float2 zs[];  // WV waves x F::LDS (at least N points each)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, f = blockIdx.y, E = P.E;
  const int u = blockIdx.x * WV + w;
  if (u >= E) return;  // the whole wave; nothing below waits on the workgroup
  float2 *T = zs + w * F::LDS;
  const uint8_t *bits = bits_all + static_cast<size_t>(f) * P.subch * SUB * 2;
#pragma unroll
  for (int m = 0; m < M; m++) T[lane + 64 * m] = make_float2(0.0f, 0.0f);
  wave_sync();
  // coefficient (u, v) at row v, coefficient (-u, v) as conj at row n - v (the Hermitian partner of (-u, v) lies in
  // column u); in column 0 every coefficient's partner (0, -v) too
  for (int e = P.col_off[u] + lane; e < P.col_off[u + 1]; e += 64) {
    const uint32_t q = static_cast<uint32_t>(P.col_ent[e]);
    float2 h;
    if constexpr (COEF) {
      const int ci = static_cast<int>(q & 0x7FFFF);
      h = mask[ci] ? coef[ci] : make_float2(0.0f, 0.0f);
    } else {
      h = qpsk(bits, static_cast<int>(q & 0x7FFFF));
    }
    const int k = static_cast<int>((q >> 19) & 0x7FF);
    T[k] = (q >> 30) ? make_float2(h.x, -h.y) : h;
    if (u == 0 && k) T[N - k] = make_float2(h.x, -h.y);
  }
  wave_sync();
  float2 v[M];
#pragma unroll
  for (int m = 0; m < M; m++) {
    const float2 t = T[lane + 64 * m];
    v[m] = make_float2(t.x, -t.y);
  }
  wave_sync();
  F::run(v, T, P.tw, lane);
  float2 *out = G + (static_cast<size_t>(f) * E + u) * N;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int y = lane + 64 * m;
    const float2 z = T[F::pos(y)];
    out[y] = make_float2(z.x, -z.y);
  }
}

// LIN: the rows are the transform times L, unclipped (the spectral loss's adjoint), instead of the clipped picture.
template <int N, int W, bool LIN = false>
__global__ __launch_bounds__(64 * W) void k_mod_rows(PlanDev P, const float2 *G, float *pic, float L) {
  using F = WaveFft<N>;
  constexpr int M = F::M;
  extern __shared__ float2 zs[];  // W waves x F::LDS
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, f = blockIdx.y, E = P.E;
  const int y = (blockIdx.x * W + w) * 2;
  float2 *T = zs + w * F::LDS;
#pragma unroll
  for (int m = 0; m < M; m++) T[lane + 64 * m] = make_float2(0.0f, 0.0f);
  wave_sync();
  // W = Z_y + i Z_y+1: at u, a + i b; at n - u, conj a + i conj b
  const float2 *col = G + static_cast<size_t>(f) * E * N + y;
  for (int u = lane; u < E; u += 64) {
    const float4 ab = *reinterpret_cast<const float4 *>(col + static_cast<size_t>(u) * N);
    T[u] = make_float2(ab.x - ab.w, ab.y + ab.z);
    if (u) T[N - u] = make_float2(ab.x + ab.w, ab.z - ab.y);
  }
  wave_sync();
  float2 v[M];
#pragma unroll
  for (int m = 0; m < M; m++) {
    const float2 t = T[lane + 64 * m];
    v[m] = make_float2(t.x, -t.y);
  }
  wave_sync();
  F::run(v, T, P.tw, lane);
  const float inv2L = 0.5f / L;
  float *row0 = pic + (static_cast<size_t>(f) * N + y) * N, *row1 = row0 + N;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int x = lane + 64 * m;
    const float2 z = T[F::pos(x)];  // conj of the DFT: (p_y, p_y+1) = (z.x, -z.y)
    if constexpr (LIN) {
      row0[x] = L * z.x;
      row1[x] = -L * z.y;
    } else {
      row0[x] = 0.5f + fminf(fmaxf(z.x, -L), L) * inv2L;
      row1[x] = 0.5f + fminf(fmaxf(-z.y, -L), L) * inv2L;
    }
  }
}

// ------------------------------------------------------------------------------------------------ the spectral loss
// The refinement's loss on the format's coefficients (an extension; the reference has no refinement): with e = w -
// alpha r - beta the warped picture's residual against the reference picture and E its unscaled 2D DFT,
// L = (2 / n^2) sum over the masked coefficients of |E(u, v)|^2 (each listed coefficient stands for itself and its
// Hermitian partner; over every coefficient of the spectrum this is sum e^2, by Parseval). Its derivative with respect
// to the picture is dL/dw(x, y) = (4 / n^2) Re sum_masked E(u, v) exp(+2 pi i (u x + v y) / n) = (2 / n^2) times the
// real picture whose Hermitian spectrum is the masked E: the sender's inverse transform (k_mod_cols, k_mod_rows) on
// those coefficients. The forward transform is the row and column passes without the erase and the detrend.

// Rows y, y + 1 of e (one frame) as one complex sequence, transformed along x, the two real rows' transforms separated
// for u < E into X[u][y] (the column pass's input, as k_rows writes it).
template <int N, int W>
__global__ __launch_bounds__(64 * W) void k_res_rows(PlanDev P, const float *wpic, const float *ref, float alpha,
                                                     float beta, float2 *X) {
  using F = WaveFft<N>;
  constexpr int M = F::M;
  extern __shared__ float2 zs[];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int y = (blockIdx.x * W + w) * 2;
  const float *a0 = wpic + static_cast<size_t>(y) * N, *a1 = a0 + N;
  const float *r0 = ref + static_cast<size_t>(y) * N, *r1 = r0 + N;
  float2 v[M];
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int x = lane + 64 * m;
    v[m] = make_float2(a0[x] - alpha * r0[x] - beta, a1[x] - alpha * r1[x] - beta);
  }
  float2 *S = zs + w * F::LDS;
  F::run(v, S, P.tw, lane);
  for (int u = lane; u < P.E; u += 64) {
    const float2 zu = S[F::pos(u)], zm = S[F::pos(u ? N - u : 0)];
    const float4 o = make_float4(0.5f * (zu.x + zm.x), 0.5f * (zu.y - zm.y), 0.5f * (zu.y + zm.y), -0.5f * (zu.x - zm.x));
    *reinterpret_cast<float4 *>(X + static_cast<size_t>(u) * N + y) = o;
  }
}

// Per 256 coefficients: sum of mask |C|^2 in FP64 by a fixed tree.
__global__ __launch_bounds__(256) void k_mask_loss(int S, const float2 *C, const uint8_t *mask, double *part) {
  __shared__ double red[256];
  const int tid = threadIdx.x, i = blockIdx.x * 256 + tid;
  double a = 0;
  if (i < S && mask[i]) a = static_cast<double>(C[i].x) * C[i].x + static_cast<double>(C[i].y) * C[i].y;
  red[tid] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  if (tid == 0) part[blockIdx.x] = red[0];
}

int launch_spectral_adjoint(const FramePlan &p, const float *d_w, const float *d_ref, float alpha, float beta,
                            const uint8_t *d_mask, float2 *d_X, float2 *d_C, double *d_part, float2 *d_G, float *d_dl,
                            hipStream_t s) {
  const int S = p.subch * SUB, parts = (S + 255) / 256;
  with_size(p.n, [&](auto Nc) {
    constexpr int NN = decltype(Nc)::value, W = 4;
    const size_t lds = W * WaveFft<NN>::LDS * sizeof(float2);
    hipLaunchKernelGGL((k_res_rows<NN, W>), dim3(NN / (2 * W)), dim3(64 * W), lds, s, p.dev, d_w, d_ref, alpha, beta,
                       d_X);
    hipLaunchKernelGGL((k_cols<NN, W, 8>), dim3((p.E + W - 1) / W, 1), dim3(64 * W), lds, s, p.dev, d_X, d_C);
    hipLaunchKernelGGL(k_mask_loss, dim3(parts), dim3(256), 0, s, S, d_C, d_mask, d_part);
    hipLaunchKernelGGL((k_mod_cols<NN, W, true>), dim3((p.E + W - 1) / W, 1), dim3(64 * W), lds, s, p.dev, nullptr,
                       d_G, d_C, d_mask);
    hipLaunchKernelGGL((k_mod_rows<NN, W, true>), dim3(NN / (2 * W), 1), dim3(64 * W), lds, s, p.dev, d_G, d_dl,
                       static_cast<float>(2.0 / (static_cast<double>(NN) * NN)));
  });
  hip_check(hipGetLastError(), "spectral loss launches");
  return parts;
}

void launch_modulate(const FramePlan &p, const uint8_t *d_bits, float2 *d_G, float *d_pic, int frames, hipStream_t s) {
  const float L = static_cast<float>(2.0 * std::sqrt(640.0 * p.subch));
  with_size(p.n, [&](auto Nc) {
    constexpr int NN = decltype(Nc)::value, W = 4;
    hipLaunchKernelGGL((k_mod_cols<NN, W>), dim3((p.E + W - 1) / W, frames), dim3(64 * W),
                       W * WaveFft<NN>::LDS * sizeof(float2), s, p.dev, d_bits, d_G);
    hipLaunchKernelGGL((k_mod_rows<NN, W>), dim3(NN / (2 * W), frames), dim3(64 * W),
                       W * WaveFft<NN>::LDS * sizeof(float2), s, p.dev, d_G, d_pic, L);
  });
  hip_check(hipGetLastError(), "modulator launches");
}

// Node centres every 16 modules from 3.5 (the middle of the line-and-track band) to side - 3.5 (layout.c:110-118,
// layout.h OB_NODE_LINE, OB_THIN_PITCH, OB_MAX_NODES).
int node_lattice(int span, float *pos) {
  const int len = span + 2 * BORDER;
  int n = 0;
  for (float c = 3.5f; c < static_cast<float>(len) - 3.5f - 8.0f && n < MAX_NODES - 1; c += 16.0f) pos[n++] = c;
  pos[n++] = static_cast<float>(len) - 3.5f;
  return n;
}

}  // namespace lz
