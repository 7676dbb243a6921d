// Registration refinement on gfx950: the warp's adjoint and Gauss-Newton over the homography.
//   k_reg_accum  one workgroup a node cell: every picture sample of the cell warped through H and the nodes' bilinear
//                blend, read bilinearly from the capture with its gradient, the residual against the reference
//                picture, and the cell's sums: e^2, J^T e and J^T J over the 8 homography parameters, the intensity
//                statistics, and e * (gx, gy) spread to the cell's four nodes (the gradient with respect to the nodes,
//                which the host carries on to k1 and the anchors)
//   k_gn_step    one wave: the cells' sums in FP64, the least-squares intensity fit, the damped 8 x 8 normal equations
//                solved in FP64 (Jacobi-scaled Gauss-Jordan with partial pivots), H moved on the device
// The loop of passes runs on the device with no host round trip; partial sums are fixed trees (no atomics), so a
// pass is bit-identical run to run.
#include <hip/hip_runtime.h>

#include "register.hpp"

namespace lz {

namespace {

constexpr int JTJ0 = 9, STAT0 = 45, NODE0 = 50;

__global__ __launch_bounds__(256) void k_reg_accum(PlanDev P, const WarpFrame *frame, const RegState *st,
                                                   const float *ref, float *part) {
  __shared__ float red[4][REG_PART];
  const int ci = blockIdx.x, cj = blockIdx.y, tid = threadIdx.x, n = P.n, nx = P.nx;
  const int x0 = P.cell0[ci], x1 = P.cell0[ci + 1], y0 = P.cell0[cj], y1 = P.cell0[cj + 1];
  const int wx = x1 - x0, cnt = wx * (y1 - y0);
  const WarpFrame F = *frame;
  float H[9];
#pragma unroll
  for (int k = 0; k < 9; k++) H[k] = st->H[k];
  const float alpha = st->alpha, beta = st->beta;
  const float2 n00 = F.nodes[cj * nx + ci], n10 = F.nodes[cj * nx + ci + 1];
  const float2 n01 = F.nodes[(cj + 1) * nx + ci], n11 = F.nodes[(cj + 1) * nx + ci + 1];
  const float xa = P.node_x[ci], xr = 1.0f / (P.node_x[ci + 1] - xa);
  const float ya = P.node_y[cj], yr = 1.0f / (P.node_y[cj + 1] - ya);
  const float xm = static_cast<float>(F.w) - 1.001f, ym = static_cast<float>(F.h) - 1.001f;
  float acc[REG_PART];
#pragma unroll
  for (int k = 0; k < REG_PART; k++) acc[k] = 0.f;
  for (int t = tid; t < cnt; t += 256) {
    const int x = x0 + t % wx, y = y0 + t / wx;
    const float mx = P.mcoord[x], my = P.mcoord[y];
    const float u = (mx - xa) * xr, v = (my - ya) * yr;
    const float w00 = (1 - u) * (1 - v), w10 = u * (1 - v), w01 = (1 - u) * v, w11 = u * v;
    const float dx = w00 * n00.x + w10 * n10.x + w01 * n01.x + w11 * n11.x;
    const float dy = w00 * n00.y + w10 * n10.y + w01 * n01.y + w11 * n11.y;
    const float A = H[0] * mx + H[1] * my + H[2], B = H[3] * mx + H[4] * my + H[5];
    const float rW = 1.0f / (H[6] * mx + H[7] * my + H[8]);
    float xs = A * rW + dx - 0.5f, ys = B * rW + dy - 0.5f;
    const bool inx = xs > 0.f && xs < xm, iny = ys > 0.f && ys < ym;
    xs = fminf(fmaxf(xs, 0.f), xm);  // a NaN reads pixel 0
    ys = fminf(fmaxf(ys, 0.f), ym);
    const int xi = static_cast<int>(xs), yi = static_cast<int>(ys);
    const float fx = xs - xi, fy = ys - yi;
    const uint8_t *p = F.img + static_cast<size_t>(yi) * F.w + xi;
    constexpr float k255 = 1.0f / 255.0f;
    const float a = p[0] * k255, b = p[1] * k255, c = p[F.w] * k255, d = p[F.w + 1] * k255;
    const float top = a + (b - a) * fx, bot = c + (d - c) * fx, val = top + (bot - top) * fy;
    const float gx = inx ? (1 - fy) * (b - a) + fy * (d - c) : 0.f, gy = iny ? bot - top : 0.f;
    const float r = ref[static_cast<size_t>(y) * n + x];
    const float e = val - alpha * r - beta;
    // d(pixel)/d(h): x = A / W, y = B / W with W = h6 mx + h7 my + 1
    const float gp = -(gx * A + gy * B) * rW * rW;
    float J[8];
    J[0] = gx * mx * rW;
    J[1] = gx * my * rW;
    J[2] = gx * rW;
    J[3] = gy * mx * rW;
    J[4] = gy * my * rW;
    J[5] = gy * rW;
    J[6] = gp * mx;
    J[7] = gp * my;
    acc[0] += e * e;
#pragma unroll
    for (int k = 0; k < 8; k++) acc[1 + k] += e * J[k];
    {
      int q = JTJ0;
#pragma unroll
      for (int k = 0; k < 8; k++)
#pragma unroll
        for (int l = k; l < 8; l++) acc[q++] += J[k] * J[l];
    }
    acc[STAT0] += val;
    acc[STAT0 + 1] += r;
    acc[STAT0 + 2] += val * r;
    acc[STAT0 + 3] += r * r;
    acc[STAT0 + 4] += 1.f;
    const float ex = e * gx, ey = e * gy;
    acc[NODE0] += ex * w00;
    acc[NODE0 + 1] += ey * w00;
    acc[NODE0 + 2] += ex * w10;
    acc[NODE0 + 3] += ey * w10;
    acc[NODE0 + 4] += ex * w01;
    acc[NODE0 + 5] += ey * w01;
    acc[NODE0 + 6] += ex * w11;
    acc[NODE0 + 7] += ey * w11;
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
    part[(static_cast<size_t>(cj) * (nx - 1) + ci) * REG_PART + tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

__global__ __launch_bounds__(64) void k_gn_step(const float *part, int cells, RegState *st, float damping, int update,
                                                float *trace, int step) {
  __shared__ double tot[REG_PART];
  const int tid = threadIdx.x;
  if (tid < REG_PART) {
    double s = 0;
    for (int c = 0; c < cells; c++) s += part[static_cast<size_t>(c) * REG_PART + tid];
    tot[tid] = s;
  }
  __syncthreads();
  if (tid != 0) return;
  if (trace) trace[step] = static_cast<float>(tot[0]);
  const double sw = tot[STAT0], sr = tot[STAT0 + 1], swr = tot[STAT0 + 2], srr = tot[STAT0 + 3], cnt = tot[STAT0 + 4];
  const double den = cnt * srr - sr * sr;
  const double al = den > 0 ? (cnt * swr - sw * sr) / den : 1.0;
  const double be = cnt > 0 ? (sw - al * sr) / cnt : 0.0;
  if (update) {
    double M[8][9];
    int q = JTJ0;
    for (int k = 0; k < 8; k++)
      for (int l = k; l < 8; l++) M[k][l] = M[l][k] = tot[q++];
    for (int k = 0; k < 8; k++) {
      M[k][k] *= 1.0 + damping;
      M[k][8] = -tot[1 + k];
    }
    // Jacobi scaling: the parameters' units span nine orders of magnitude (pixels against perspective terms)
    double sc[8];
    for (int k = 0; k < 8; k++) sc[k] = M[k][k] > 0 ? 1.0 / sqrt(M[k][k]) : 1.0;
    for (int k = 0; k < 8; k++) {
      for (int l = 0; l < 8; l++) M[k][l] *= sc[k] * sc[l];
      M[k][8] *= sc[k];
    }
    bool ok = true;
    for (int c = 0; c < 8 && ok; c++) {
      int p = c;
      for (int r = c + 1; r < 8; r++)
        if (fabs(M[r][c]) > fabs(M[p][c])) p = r;
      if (!(fabs(M[p][c]) > 1e-12)) {
        ok = false;
        break;
      }
      for (int k = 0; k < 9; k++) {
        const double t = M[c][k];
        M[c][k] = M[p][k];
        M[p][k] = t;
      }
      for (int r = 0; r < 8; r++) {
        if (r == c) continue;
        const double f = M[r][c] / M[c][c];
        for (int k = c; k < 9; k++) M[r][k] -= f * M[c][k];
      }
    }
    if (ok) {
      double dl[8];
      for (int k = 0; k < 8; k++) {
        dl[k] = M[k][8] / M[k][k] * sc[k];
        ok = ok && dl[k] == dl[k];
      }
      if (ok)
        for (int k = 0; k < 8; k++) st->H[k] = static_cast<float>(st->H[k] + dl[k]);
    }
  }
  st->alpha = static_cast<float>(al);
  st->beta = static_cast<float>(be);
}

}  // namespace

void launch_reg_accum(const FramePlan &p, const WarpFrame *d_frame, const RegState *d_state, const float *d_ref,
                      float *d_part, hipStream_t s) {
  const int c = p.dev.nx - 1;
  hipLaunchKernelGGL(k_reg_accum, dim3(c, c), dim3(256), 0, s, p.dev, d_frame, d_state, d_ref, d_part);
  hip_check(hipGetLastError(), "refinement launch");
}

void launch_gn_step(const FramePlan &p, const float *d_part, RegState *d_state, float damping, int update, float *d_trace,
                    int step, hipStream_t s) {
  const int c = p.dev.nx - 1;
  hipLaunchKernelGGL(k_gn_step, dim3(1), dim3(64), 0, s, d_part, c * c, d_state, damping, update, d_trace, step);
  hip_check(hipGetLastError(), "Gauss-Newton launch");
}

}  // namespace lz
