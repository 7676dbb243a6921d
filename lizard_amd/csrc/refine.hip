// Registration refinement on gfx950: the warp's adjoint and Gauss-Newton over the homography.
//   k_reg_accum  one workgroup a node cell (or a share of its rows): every picture sample of the cell warped through H and the nodes' bilinear
//                blend, read bilinearly from the capture with its gradient, the residual against the reference
//                picture, and the cell's sums: e^2, J^T e and J^T J over the 8 homography parameters and k1 (the
//                lens field's blend at the sample), the intensity statistics, and e * (gx, gy) spread to the cell's
//                four nodes (the gradient with respect to the nodes, which the host carries on to k1 and the anchors)
//   k_gn_sum     one workgroup a sum: the cells' partial sums in FP64 in a fixed order
//   k_gn_solve   one lane: the least-squares intensity fit, the damped 8 x 8 (or 9 x 9, with k1) normal equations
//                solved in FP64 (Jacobi-scaled Cholesky, in registers), H (and k1) moved on the device
// The loop of passes runs on the device with no host round trip; partial sums are fixed trees (no atomics), so a
// pass is bit-identical run to run.
#include <hip/hip_runtime.h>

#include "register.hpp"

namespace lz {

namespace {

constexpr int NG = REG_GN, JTJ0 = 1 + NG, STAT0 = JTJ0 + NG * (NG + 1) / 2, NODE0 = STAT0 + 5, MET0 = NODE0 + 8;
static_assert(NODE0 == REG_NODE0 && MET0 == REG_MET0 && MET0 + 12 == REG_PART, "the partial sums' layout");

__global__ __launch_bounds__(256) void k_reg_accum(PlanDev P, const WarpFrame *frame, const RegState *st,
                                                   const float *ref, const float2 *lens, float *part,
                                                   const float *gsrc) {
  __shared__ float red[4][REG_PART];
  const int ci = blockIdx.x, cj = blockIdx.y / REG_SPLIT, piece = blockIdx.y % REG_SPLIT, tid = threadIdx.x;
  const int n = P.n, nx = P.nx;
  const int x0 = P.cell0[ci], x1 = P.cell0[ci + 1], ya = P.cell0[cj], yb = P.cell0[cj + 1];
  const int y0 = ya + (yb - ya) * piece / REG_SPLIT, y1 = ya + (yb - ya) * (piece + 1) / REG_SPLIT;
  const WarpFrame F = *frame;
  float H[9];
#pragma unroll
  for (int k = 0; k < 9; k++) H[k] = st->H[k];
  const float alpha = st->alpha, beta = st->beta, dk1 = lens ? st->dk1 : 0.0f;
  const int q00 = cj * nx + ci, q10 = q00 + 1, q01 = q00 + nx, q11 = q01 + 1;
  const float2 z2 = make_float2(0.f, 0.f);
  const float2 l00 = lens ? lens[q00] : z2, l10 = lens ? lens[q10] : z2, l01 = lens ? lens[q01] : z2,
               l11 = lens ? lens[q11] : z2;
  const float2 n00 = make_float2(F.nodes[q00].x + dk1 * l00.x, F.nodes[q00].y + dk1 * l00.y);
  const float2 n10 = make_float2(F.nodes[q10].x + dk1 * l10.x, F.nodes[q10].y + dk1 * l10.y);
  const float2 n01 = make_float2(F.nodes[q01].x + dk1 * l01.x, F.nodes[q01].y + dk1 * l01.y);
  const float2 n11 = make_float2(F.nodes[q11].x + dk1 * l11.x, F.nodes[q11].y + dk1 * l11.y);
  const float xa = P.node_x[ci], xr = 1.0f / (P.node_x[ci + 1] - xa);
  const float yn = P.node_y[cj], yr = 1.0f / (P.node_y[cj + 1] - yn);
  const float xm = static_cast<float>(F.w) - 1.001f, ym = static_cast<float>(F.h) - 1.001f;
  float acc[REG_PART];
#pragma unroll
  for (int k = 0; k < REG_PART; k++) acc[k] = 0.f;
  // the cell's samples in row order, four a thread at a time so that their capture reads are in flight together;
  // t / wx by the rounded-up reciprocal (exact while t wx < 2^32)
  const unsigned wx = static_cast<unsigned>(x1 - x0), cnt = wx * static_cast<unsigned>(y1 - y0);
  const unsigned magic = wx ? static_cast<unsigned>(0xFFFFFFFFull / wx) + 1u : 0u;
  constexpr int UNR = 4;
  constexpr float k255 = 1.0f / 255.0f;
  for (unsigned t0 = tid; t0 < cnt; t0 += 256 * UNR) {
    float mxs[UNR], mys[UNR], As[UNR], Bs[UNR], rWs[UNR], fxs[UNR], fys[UNR], rs[UNR], gs[UNR];
    uint8_t pa[UNR], pb[UNR], pc[UNR], pd[UNR];
    bool ok[UNR], inx[UNR], iny[UNR];
    float w00[UNR], w10[UNR], w01[UNR], w11[UNR];
#pragma unroll
    for (int j = 0; j < UNR; j++) {
      const unsigned t = t0 + 256u * j;
      ok[j] = t < cnt;
      const unsigned tt = ok[j] ? t : 0u;
      const unsigned row = __umulhi(tt, magic);
      const int y = y0 + static_cast<int>(row), x = x0 + static_cast<int>(tt - row * wx);
      const float mx = P.mcoord[x], my = P.mcoord[y];
      const float u = (mx - xa) * xr, v = (my - yn) * yr;
      w00[j] = (1 - u) * (1 - v);
      w10[j] = u * (1 - v);
      w01[j] = (1 - u) * v;
      w11[j] = u * v;
      const float dx = w00[j] * n00.x + w10[j] * n10.x + w01[j] * n01.x + w11[j] * n11.x;
      const float dy = w00[j] * n00.y + w10[j] * n10.y + w01[j] * n01.y + w11[j] * n11.y;
      const float A = H[0] * mx + H[1] * my + H[2], B = H[3] * mx + H[4] * my + H[5];
      const float rW = 1.0f / (H[6] * mx + H[7] * my + H[8]);
      float xs = A * rW + dx - 0.5f, ys = B * rW + dy - 0.5f;
      inx[j] = xs > 0.f && xs < xm;
      iny[j] = ys > 0.f && ys < ym;
      xs = fminf(fmaxf(xs, 0.f), xm);  // a NaN reads pixel 0
      ys = fminf(fmaxf(ys, 0.f), ym);
      const int xi = static_cast<int>(xs), yi = static_cast<int>(ys);
      fxs[j] = xs - xi;
      fys[j] = ys - yi;
      const uint8_t *p = F.img + static_cast<ptrdiff_t>(yi) * F.stride + xi;
      pa[j] = p[0];
      pb[j] = p[1];
      pc[j] = p[F.stride];
      pd[j] = p[F.stride + 1];
      rs[j] = ref[static_cast<size_t>(y) * n + x];
      gs[j] = gsrc ? gsrc[static_cast<size_t>(y) * n + x] : 0.0f;
      mxs[j] = mx;
      mys[j] = my;
      As[j] = A;
      Bs[j] = B;
      rWs[j] = rW;
    }
#pragma unroll
    for (int j = 0; j < UNR; j++) {
      if (!ok[j]) continue;
      const float fx = fxs[j], fy = fys[j], mx = mxs[j], my = mys[j], A = As[j], B = Bs[j], rW = rWs[j], r = rs[j];
      const float a = pa[j] * k255, b = pb[j] * k255, c = pc[j] * k255, d = pd[j] * k255;
      const float top = a + (b - a) * fx, bot = c + (d - c) * fx, val = top + (bot - top) * fy;
      const float gx = inx[j] ? (1 - fy) * (b - a) + fy * (d - c) : 0.f, gy = iny[j] ? bot - top : 0.f;
      // the residual, or half the loss's derivative with respect to the sample where one is given (the spectral
      // loss's adjoint through the FFT), so that 2 e J is that loss's gradient either way
      const float e = gsrc ? 0.5f * gs[j] : val - alpha * r - beta;
      // d(pixel)/d(h): x = A / W, y = B / W with W = h6 mx + h7 my + 1
      const float gp = -(gx * A + gy * B) * rW * rW;
      float J[NG];
      J[0] = gx * mx * rW;
      J[1] = gx * my * rW;
      J[2] = gx * rW;
      J[3] = gy * mx * rW;
      J[4] = gy * my * rW;
      J[5] = gy * rW;
      J[6] = gp * mx;
      J[7] = gp * my;
      // d(pixel)/d(k1): the unit lens field's blend at the sample
      J[8] = gx * (w00[j] * l00.x + w10[j] * l10.x + w01[j] * l01.x + w11[j] * l11.x) +
             gy * (w00[j] * l00.y + w10[j] * l10.y + w01[j] * l01.y + w11[j] * l11.y);
      acc[0] += e * e;
#pragma unroll
      for (int k = 0; k < NG; k++) acc[1 + k] += e * J[k];
      {
        int q = JTJ0;
#pragma unroll
        for (int k = 0; k < NG; k++)
#pragma unroll
          for (int l = k; l < NG; l++) acc[q++] += J[k] * J[l];
      }
      acc[STAT0] += val;
      acc[STAT0 + 1] += r;
      acc[STAT0 + 2] += val * r;
      acc[STAT0 + 3] += r * r;
      acc[STAT0 + 4] += 1.f;
      const float ex = e * gx, ey = e * gy;
      acc[NODE0] += ex * w00[j];
      acc[NODE0 + 1] += ey * w00[j];
      acc[NODE0 + 2] += ex * w10[j];
      acc[NODE0 + 3] += ey * w10[j];
      acc[NODE0 + 4] += ex * w01[j];
      acc[NODE0 + 5] += ey * w01[j];
      acc[NODE0 + 6] += ex * w11[j];
      acc[NODE0 + 7] += ey * w11[j];
      // the corners' shares of the warp's metric (the diagonal blocks of J^T J over the node displacements)
      const float gxx = gx * gx, gxy = gx * gy, gyy = gy * gy;
      const float wq[4] = {w00[j] * w00[j], w10[j] * w10[j], w01[j] * w01[j], w11[j] * w11[j]};
#pragma unroll
      for (int a = 0; a < 4; a++) {
        acc[MET0 + 3 * a] += wq[a] * gxx;
        acc[MET0 + 3 * a + 1] += wq[a] * gxy;
        acc[MET0 + 3 * a + 2] += wq[a] * gyy;
      }
    }
  }
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int k = 0; k < REG_PART; k++) {
    float s = acc[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) red[w][k] = s;
  }
  __syncthreads();
  if (tid < REG_PART)
    part[(static_cast<size_t>(blockIdx.y) * (nx - 1) + ci) * REG_PART + tid] =
        (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

// Component k of the cells' partial sums, in double, by one workgroup: each thread's cells in order, then a fixed
// tree over the threads.
__global__ __launch_bounds__(256) void k_gn_sum(const float *part, int cells, double *sum) {
  __shared__ double red[256];
  const int k = blockIdx.x, tid = threadIdx.x;
  double s = 0;
  for (int c = tid; c < cells; c += 256) s += part[static_cast<size_t>(c) * REG_PART + k];
  red[tid] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  if (tid == 0) sum[k] = red[0];
}

// The intensity fit and, when update, the Gauss-Newton step: the damped normal equations, Jacobi-scaled (the
// parameters' units span nine orders of magnitude, pixels against perspective terms), solved by Cholesky in FP64 with
// every index a constant (the matrix stays in registers). One lane does it; the matrix is NP x NP (8: H; 9: H, k1).
template <int NP>
__device__ void gn_update(const double *tot, RegState *st, float damping) {
  double A[NP][NP], b[NP], sc[NP];
  constexpr int ROW = NG;  // J^T J's stored rows hold NG entries less the row index
#pragma unroll
  for (int k = 0; k < NP; k++) {
    const double d = tot[JTJ0 + k * ROW - k * (k - 1) / 2];  // J^T J (k, k) in the upper triangle's row order
    sc[k] = d > 0 ? 1.0 / sqrt(d * (1.0 + damping)) : 1.0;
  }
#pragma unroll
  for (int k = 0; k < NP; k++) {
#pragma unroll
    for (int l = k; l < NP; l++) {
      const double v = tot[JTJ0 + k * ROW - k * (k - 1) / 2 + (l - k)] * (k == l ? 1.0 + damping : 1.0);
      A[k][l] = A[l][k] = v * sc[k] * sc[l];
    }
    b[k] = -tot[1 + k] * sc[k];
  }
  bool ok = true;
  // A = L L^T, L in the lower triangle of A
#pragma unroll
  for (int j = 0; j < NP; j++) {
    double d = A[j][j];
#pragma unroll
    for (int k = 0; k < j; k++) d -= A[j][k] * A[j][k];
    ok = ok && d > 1e-300;
    const double ljj = sqrt(d > 1e-300 ? d : 1e-300);
    A[j][j] = ljj;
#pragma unroll
    for (int i = j + 1; i < NP; i++) {
      double t = A[i][j];
#pragma unroll
      for (int k = 0; k < j; k++) t -= A[i][k] * A[j][k];
      A[i][j] = t / ljj;
    }
  }
  double x[NP];
#pragma unroll
  for (int i = 0; i < NP; i++) {
    double t = b[i];
#pragma unroll
    for (int k = 0; k < i; k++) t -= A[i][k] * x[k];
    x[i] = t / A[i][i];
  }
#pragma unroll
  for (int i = NP - 1; i >= 0; i--) {
    double t = x[i];
#pragma unroll
    for (int k = i + 1; k < NP; k++) t -= A[k][i] * x[k];
    x[i] = t / A[i][i];
  }
#pragma unroll
  for (int k = 0; k < NP; k++) ok = ok && x[k] == x[k];
  if (ok) {
#pragma unroll
    for (int k = 0; k < 8; k++) st->H[k] = static_cast<float>(st->H[k] + x[k] * sc[k]);
    if constexpr (NP > 8) st->dk1 = static_cast<float>(st->dk1 + x[8] * sc[8]);
  }
}

__global__ __launch_bounds__(64) void k_gn_solve(const double *tot, RegState *st, float damping, int update,
                                                 int nparam, float *trace, int step) {
  if (threadIdx.x != 0) return;
  if (trace) trace[step] = static_cast<float>(tot[0]);
  const double sw = tot[STAT0], sr = tot[STAT0 + 1], swr = tot[STAT0 + 2], srr = tot[STAT0 + 3], cnt = tot[STAT0 + 4];
  const double den = cnt * srr - sr * sr;
  const double al = den > 0 ? (cnt * swr - sw * sr) / den : 1.0;
  const double be = cnt > 0 ? (sw - al * sr) / cnt : 0.0;
  if (update) {
    if (nparam > 8) gn_update<9>(tot, st, damping);
    else gn_update<8>(tot, st, damping);
  }
  st->alpha = static_cast<float>(al);
  st->beta = static_cast<float>(be);
}

}  // namespace

void launch_reg_accum(const FramePlan &p, const WarpFrame *d_frame, const RegState *d_state, const float *d_ref,
                      const float2 *d_lens, float *d_part, hipStream_t s, const float *d_dl) {
  const int c = p.dev.nx - 1;
  hipLaunchKernelGGL(k_reg_accum, dim3(c, c * REG_SPLIT), dim3(256), 0, s, p.dev, d_frame, d_state, d_ref, d_lens,
                     d_part, d_dl);
  hip_check(hipGetLastError(), "refinement launch");
}

void launch_gn_step(const FramePlan &p, const float *d_part, double *d_sum, RegState *d_state, float damping, int update,
                    int nparam, float *d_trace, int step, hipStream_t s) {
  const int c = p.dev.nx - 1;
  hipLaunchKernelGGL(k_gn_sum, dim3(REG_PART), dim3(256), 0, s, d_part, c * c * REG_SPLIT, d_sum);
  hipLaunchKernelGGL(k_gn_solve, dim3(1), dim3(64), 0, s, d_sum, d_state, damping, update, nparam, d_trace, step);
  hip_check(hipGetLastError(), "Gauss-Newton launch");
}

}  // namespace lz
