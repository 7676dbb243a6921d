// The border's nodes (border.hpp): one 512-lane workgroup a frame, a wave a node, the nodes of a level at once, in the
// reference's order of float operations (liblizard/src/acquire.c:1242-1468 refine_node, :1205-1236 node_ncc, :848-859
// node_inside, their scalar forms), so each node is the reference's to the bit.
//   start     the mean displacement of the node's measured 8-neighbours less the lens model there, plus the model at the
//             node (acquire.c:1259-1272)
//   grid      the capture read on a half-module grid over the template's cells and the coarse search's reach
//   coarse    7 x 7 offsets of half a module, a lane each: the normalised correlation of the template (the border's
//             own light and dark cells, the picture and the format word left out) with the grid's samples; the first
//             best in row order (lanes, then a butterfly that keeps the lower offset on a tie)
//   climb     a quarter module then an eighth, a move at a time: the correlations a move lacks a lane each, then a
//             parabola each way through the peak
//   place     the offset turned into pixels through H's local Jacobian, or the start where the template scores
//             under 0.35
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "border.hpp"
#include "common.hpp"

namespace lz {
namespace {

constexpr int NT = 512, WAVES = NT / 64;
constexpr int GMAX = 480;  // the grid of a side node's template: (2 x 13 + 5) x (2 x 5 + 5) samples
constexpr int CMAX = 72;   // a side node's template cells (13 x 5)
constexpr float NODE_NONE = -2.0f;
constexpr int KD_DATA = 0, KD_DARK = 2, KD_WORD = 8;

__device__ bool in_capture(int w, int h, float u, float v) {
  return u >= 0.5f && u <= w - 0.5f && v >= 0.5f && v <= h - 0.5f;
}

// Whether the template's cell middles (rect r about (cx0, cy0)) moved by offsets in [ax, bx] x [ay, by] modules lie
// inside the capture with a pixel to spare, through H and the start (acquire.c:848-859).
__device__ bool node_inside(const BorderJob &J, int cx0, int cy0, const int *r, float pdx, float pdy, float ax, float bx,
                            float ay, float by) {
  const float x0 = cx0 + r[0] + 0.5f + ax, x1 = cx0 + r[1] - 0.5f + bx, y0 = cy0 + r[2] + 0.5f + ay,
              y1 = cy0 + r[3] - 0.5f + by;
  for (int c = 0; c < 4; c++) {
    const float x = c & 1 ? x1 : x0, y = c & 2 ? y1 : y0;
    if (!(J.H[6] * x + J.H[7] * y + 1 > 0)) return false;
    float u, v;
    project_f(J.H, x, y, &u, &v);
    if (!(u + pdx >= 1.5f && u + pdx <= J.w - 1.5f && v + pdy >= 1.5f && v + pdy <= J.h - 1.5f)) return false;
  }
  return true;
}

// The template's correlation at offset (ox, oy) modules (acquire.c:1205-1236): cd[k] the cell's module offset from
// (cx0, cy0), tv[k] its level (+1 light, -1 dark), in the reference's cell order.
__device__ float node_ncc(const BorderJob &J, const float *lut, int cx0, int cy0, const float *tv, const short2 *cd,
                          int cells, float pdx, float pdy, float ox, float oy, bool edge, int nmin) {
  float st = 0, ss = 0, stt = 0, sss = 0, sts = 0;
  int n = 0, nin = 0;
  for (int k = 0; k < cells; k++) {
    const float mx = cx0 + cd[k].x + 0.5f + ox, my = cy0 + cd[k].y + 0.5f + oy;
    float u, v;
    project_f(J.H, mx, my, &u, &v);
    const float s = sample_px(J.img, J.w, J.h, lut, u + pdx, v + pdy);
    if (edge) nin += in_capture(J.w, J.h, u + pdx, v + pdy) ? 1 : 0;
    const float t = tv[k];
    st += t;
    ss += s;
    stt += t * t;
    sss += s * s;
    sts += t * s;
    n++;
  }
  if (edge && nin < nmin) return NODE_NONE;
  const float vt = stt - st * st / n, vs = sss - ss * ss / n;
  if (vt <= 0 || vs <= 1e-9f) return 0;
  return (sts - st * ss / n) / sqrtf(vt * vs);
}

__global__ __launch_bounds__(NT) void k_border_nodes(const BorderJob *__restrict__ jobs) {
  const BorderJob &J = jobs[blockIdx.x];
  __shared__ NodeState nd[MAX_NODES * MAX_NODES];
  __shared__ float G[WAVES][GMAX], GM[WAVES][GMAX];
  __shared__ float tvs[WAVES][CMAX];
  __shared__ short2 cds[WAVES][CMAX];
  __shared__ short ats[WAVES][CMAX];
  __shared__ float lut[256];
  __shared__ float wst[WAVES], wvt[WAVES];
  __shared__ int wcells[WAVES];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, nx = J.nx, Lw = J.Lw;
  for (int k = tid; k < nx * nx; k += NT) nd[k] = J.nodes[k];
  if (tid < 256) lut[tid] = static_cast<float>(tid) / 255.0f;
  __syncthreads();
  for (int l = 0; l < J.levels; l++) {
    const int b0 = J.level[l], cnt = J.level[l + 1] - b0;
    for (int c0 = 0; c0 < cnt; c0 += WAVES) {
      const bool has = c0 + wave < cnt;
      const int node = has ? J.order[b0 + c0 + wave] : 0, i = node % nx, j = node / nx;
      // the template's rect, the start, the edge rule and the Jacobian, alike in every lane of the wave
      const int cx0 = static_cast<int>(J.pos[i]), cy0 = static_cast<int>(J.pos[j]);
      const float mcx = cx0 + 0.5f, mcy = cy0 + 0.5f;
      const int ex = (j == 0 || j == nx - 1) ? 6 : 2, ey = (i == 0 || i == nx - 1) ? 6 : 2;
      int rect[4];
      rect[0] = -ex < -cx0 ? -cx0 : -ex;
      rect[1] = ex + 1 > Lw - cx0 ? Lw - cx0 : ex + 1;
      rect[2] = -ey < -cy0 ? -cy0 : -ey;
      rect[3] = ey + 1 > Lw - cy0 ? Lw - cy0 : ey + 1;
      float pdx = 0, pdy = 0;
      int pn = 0;
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          const int ii = i + di, jj = j + dj;
          if (ii < 0 || jj < 0 || ii >= nx || jj >= nx || !nd[jj * nx + ii].done) continue;
          float rx = nd[jj * nx + ii].dx, ry = nd[jj * nx + ii].dy;
          if (J.lens_on) {
            float u, v, mx, my;
            project_f(J.H, J.pos[ii], J.pos[jj], &u, &v);
            lens_disp(J.K, J.t, J.ax, J.ay, J.R, u, v, &mx, &my);
            rx -= mx;
            ry -= my;
          }
          pdx += rx;
          pdy += ry;
          pn++;
        }
      if (pn) {
        pdx /= pn;
        pdy /= pn;
      }
      if (J.lens_on) {
        float u, v, mx, my;
        project_f(J.H, J.pos[i], J.pos[j], &u, &v);
        lens_disp(J.K, J.t, J.ax, J.ay, J.R, u, v, &mx, &my);
        pdx += mx;
        pdy += my;
      }
      const bool edge = !node_inside(J, cx0, cy0, rect, pdx, pdy, -1.5f, 1.5f, -1.5f, 1.5f);
      float px, py, ax, ay, bx, by;
      project_f(J.H, mcx, mcy, &px, &py);
      project_f(J.H, mcx + 1, mcy, &ax, &ay);
      project_f(J.H, mcx, mcy + 1, &bx, &by);
      const float jxx = ax - px, jxy = ay - py, jyx = bx - px, jyy = by - py;
      const int lo = rect[0], ylo = rect[2], side = rect[1] - rect[0], tall = rect[3] - rect[2];
      const int GW = 2 * side + 5, GH = 2 * tall + 5;
      if (has) {
        for (int idx = lane; idx < GW * GH; idx += 64) {
          const int gy = idx / GW, gx = idx - gy * GW;
          const float my = static_cast<float>(2 * (cy0 + ylo) + gy - 2) * 0.5f;
          const float mx = static_cast<float>(2 * (cx0 + lo) + gx - 2) * 0.5f;
          float u, v;
          project_f(J.H, mx, my, &u, &v);
          G[wave][idx] = sample_px(J.img, J.w, J.h, lut, u + pdx, v + pdy);
          GM[wave][idx] = !edge || in_capture(J.w, J.h, u + pdx, v + pdy) ? 1.0f : 0.0f;
        }
        if (lane == 0) {  // the template: the border's own cells in row order
          float st = 0, stt = 0;
          int cells = 0;
          for (int dy = 0; dy < tall; dy++)
            for (int dx = 0; dx < side; dx++) {
              const int kind = J.kinds[(cy0 + ylo + dy) * Lw + cx0 + lo + dx];
              if (kind == KD_DATA || (kind & KD_WORD)) continue;
              const float t = (kind & 3) == KD_DARK ? -1.0f : 1.0f;
              tvs[wave][cells] = t;
              ats[wave][cells] = static_cast<short>((2 * dy + 3) * GW + 2 * dx + 3);
              cds[wave][cells] = make_short2(static_cast<short>(lo + dx), static_cast<short>(ylo + dy));
              st += t;
              stt += t * t;
              cells++;
            }
          wst[wave] = st;
          wvt[wave] = stt - st * st / cells;
          wcells[wave] = cells;
        }
      }
      __syncthreads();
      if (has) {
        const int cells = wcells[wave], nmin = (cells + 1) / 2;
        const float st = wst[wave], vt = wvt[wave];
        const float *tv = tvs[wave];
        const short2 *cd = cds[wave];
        // the coarse search: offset (a, b) half-modules, a lane each
        float sc = NODE_NONE;
        if (lane < 49) {
          const int a = lane % 7 - 3, b = lane / 7 - 3, base = b * GW + a;
          float ss = 0, sss = 0, sts = 0, sn = 0;
          for (int k = 0; k < cells; k++) {
            const float v = G[wave][ats[wave][k] + base];
            ss += v;
            sss += v * v;
            sts += tv[k] * v;
          }
          if (edge)
            for (int k = 0; k < cells; k++) sn += GM[wave][ats[wave][k] + base];
          const float vs = sss - ss * ss / cells;
          sc = edge && sn < nmin ? NODE_NONE : vt <= 0 || vs <= 1e-9f ? 0 : (sts - st * ss / cells) / sqrtf(vt * vs);
          if (!(sc > NODE_NONE)) sc = NODE_NONE;
        }
        float bv = sc;
        int bi = lane;
        for (int m = 32; m > 0; m >>= 1) {
          const float ov = __shfl_xor(bv, m);
          const int oi = __shfl_xor(bi, m);
          if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
          }
        }
        float best = NODE_NONE, box = 0, boy = 0;
        if (bv > NODE_NONE) {
          best = bv;
          box = (bi % 7 - 3) * 0.5f;
          boy = (bi / 7 - 3) * 0.5f;
        }
        // the climb, its control alike in every lane; the correlations a move lacks a lane each
        float g[3][3] = {};
        int hv[3][3] = {};
        float step = 0.25f;
        int ma = 1, mb = 1;
        g[1][1] = best;
        hv[1][1] = 1;
        for (int pass = 0; pass < 2; pass++, step *= 0.5f) {
          for (int move = 0; move < 8; move++) {
            if (ma != 1 || mb != 1) {
              float ng[3][3];
              int nh[3][3];
              for (int b2 = 0; b2 < 3; b2++)
                for (int a2 = 0; a2 < 3; a2++) {
                  const int sb = b2 + mb - 1, sa = a2 + ma - 1;
                  const bool in = sb >= 0 && sb < 3 && sa >= 0 && sa < 3;
                  nh[b2][a2] = in && hv[sb][sa];
                  ng[b2][a2] = in ? g[sb][sa] : 0;
                }
              for (int b2 = 0; b2 < 3; b2++)
                for (int a2 = 0; a2 < 3; a2++) {
                  g[b2][a2] = ng[b2][a2];
                  hv[b2][a2] = nh[b2][a2];
                }
              ma = mb = 1;
            }
            const bool ce = !node_inside(J, cx0, cy0, rect, pdx, pdy, box - step, box + step, boy - step, boy + step);
            int m = 0, mine = -1;
            for (int c = 0; c < 9; c++)
              if (!hv[c / 3][c % 3]) {
                if (m == lane) mine = c;
                m++;
              }
            float val = 0;
            if (mine >= 0)
              val = node_ncc(J, lut, cx0, cy0, tv, cd, cells, pdx, pdy, box + (mine % 3 - 1) * step,
                             boy + (mine / 3 - 1) * step, ce, nmin);
            for (int c = 0, q = 0; c < 9; c++)
              if (!hv[c / 3][c % 3]) {
                g[c / 3][c % 3] = __shfl(val, q++);
                hv[c / 3][c % 3] = 1;
              }
            int ba = 1, bb = 1;
            for (int b2 = 0; b2 < 3; b2++)
              for (int a2 = 0; a2 < 3; a2++)
                if (g[b2][a2] > g[bb][ba]) {
                  ba = a2;
                  bb = b2;
                }
            if (ba == 1 && bb == 1) break;
            box += (ba - 1) * step;
            boy += (bb - 1) * step;
            ma = ba;
            mb = bb;
          }
          if (pass == 0) {  // the finer pass starts from the centre alone (the cell last moved to)
            const float c = g[mb][ma];
            for (int b2 = 0; b2 < 3; b2++)
              for (int a2 = 0; a2 < 3; a2++) hv[b2][a2] = 0;
            g[1][1] = c;
            hv[1][1] = 1;
            ma = mb = 1;
          }
        }
        best = g[1][1];
        const float dxx = g[1][0] - 2 * g[1][1] + g[1][2], dyy = g[0][1] - 2 * g[1][1] + g[2][1];
        if (dxx < -1e-6f && g[1][0] > NODE_NONE && g[1][2] > NODE_NONE) box += 0.125f * 0.5f * (g[1][0] - g[1][2]) / dxx;
        if (dyy < -1e-6f && g[0][1] > NODE_NONE && g[2][1] > NODE_NONE) boy += 0.125f * 0.5f * (g[0][1] - g[2][1]) / dyy;
        NodeState o;
        o.done = 1;
        o.score = best;
        if (best < 0.35f) {
          o.dx = pdx;
          o.dy = pdy;
        } else {
          o.dx = pdx + box * jxx + boy * jyx;
          o.dy = pdy + box * jxy + boy * jyy;
        }
        if (lane == 0) {
          nd[node] = o;
          J.nodes[node] = o;
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256) void k_border_blur(const uint8_t *__restrict__ img, uint8_t *__restrict__ out, int w,
                                                     int h, int r, int4 b0, int4 b1, int4 b2, int4 b3) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const int4 bs[4] = {b0, b1, b2, b3};
  bool in = false;
  for (int k = 0; k < 4; k++) in = in || (x >= bs[k].x && x < bs[k].z && y >= bs[k].y && y < bs[k].w);
  const size_t o = static_cast<size_t>(y) * w + x;
  if (!in) {
    out[o] = img[o];
    return;
  }
  int a = 0;
  for (int dy = -r; dy <= r; dy++) {
    const int sy = min(max(y + dy, 0), h - 1);
    const uint8_t *row = img + static_cast<size_t>(sy) * w;
    for (int dx = -r; dx <= r; dx++) a += row[min(max(x + dx, 0), w - 1)];
  }
  const int kk = (2 * r + 1) * (2 * r + 1);
  out[o] = static_cast<uint8_t>((a + kk / 2) / kk);
}

}  // namespace

int border_schedule(int span, std::vector<short> &order, std::vector<short> &level) {
  float pos[MAX_NODES];
  const int nx = node_lattice(span, pos), nn = nx * nx;
  const auto ring = [&](int i, int j) { return std::min(i, nx - 1 - i) + std::min(j, nx - 1 - j); };
  const auto marked = [&](int i, int j) {
    const bool ex = i == 0 || i == nx - 1, ey = j == 0 || j == nx - 1;
    return (ex || ey) && !(ex && ey);
  };
  // the reference's order: a counting sort on the ring, rows within a ring (acquire.c:5046-5055)
  order.clear();
  for (int rg = 0; rg <= nx; rg++)
    for (int j = 0; j < nx; j++)
      for (int i = 0; i < nx; i++)
        if (marked(i, j) && ring(i, j) == rg) order.push_back(static_cast<short>(j * nx + i));
  // a node's level: one past the highest of its neighbours that precede it (the corners: none)
  std::vector<int> lev(nn, -1), at(nn, -1);
  for (size_t k = 0; k < order.size(); k++) at[order[k]] = static_cast<int>(k);
  int levels = 0;
  for (size_t k = 0; k < order.size(); k++) {
    const int i = order[k] % nx, j = order[k] / nx;
    int lv = 0;
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        const int ii = i + di, jj = j + dj;
        if ((!di && !dj) || ii < 0 || jj < 0 || ii >= nx || jj >= nx) continue;
        const int q = jj * nx + ii;
        if (at[q] >= 0 && at[q] < static_cast<int>(k)) lv = std::max(lv, lev[q] + 1);
      }
    lev[order[k]] = lv;
    levels = std::max(levels, lv + 1);
  }
  std::vector<short> by;
  level.assign(1, 0);
  for (int l = 0; l < levels; l++) {
    for (short q : order)
      if (lev[q] == l) by.push_back(q);
    level.push_back(static_cast<short>(by.size()));
  }
  order.swap(by);
  return levels;
}

void border_band(int span, const float *H, int w, int h, int r, int rects[16]) {
  const float D = 16, Mo = 3, W = static_cast<float>(span + 2 * BORDER), Hh = W;
  const float band[4][4] = {{-Mo, -Mo, W + Mo, D}, {-Mo, Hh - D, W + Mo, Hh + Mo}, {-Mo, -Mo, D, Hh + Mo},
                            {W - D, -Mo, W + Mo, Hh + Mo}};
  for (int k = 0; k < 4; k++) {
    float lo[2] = {1e9f, 1e9f}, hi[2] = {-1e9f, -1e9f};
    for (int c = 0; c < 4; c++) {
      float u, v;
      project_f(H, band[k][c & 1 ? 2 : 0], band[k][c & 2 ? 3 : 1], &u, &v);
      lo[0] = std::fmin(lo[0], u);
      hi[0] = std::fmax(hi[0], u);
      lo[1] = std::fmin(lo[1], v);
      hi[1] = std::fmax(hi[1], v);
    }
    const int pad = r + 4;
    int *q = rects + 4 * k;
    q[0] = static_cast<int>(std::floor(lo[0])) - pad;
    q[1] = static_cast<int>(std::floor(lo[1])) - pad;
    q[2] = static_cast<int>(std::ceil(hi[0])) + pad;
    q[3] = static_cast<int>(std::ceil(hi[1])) + pad;
    q[0] = std::max(q[0], 0);
    q[1] = std::max(q[1], 0);
    q[2] = std::min(q[2], w);
    q[3] = std::min(q[3], h);
    if (q[2] <= q[0] || q[3] <= q[1]) q[0] = q[1] = q[2] = q[3] = 0;
  }
}

void launch_border_blur(const uint8_t *img, uint8_t *out, int w, int h, int r, const int rects[16], hipStream_t s) {
  const int4 b0 = make_int4(rects[0], rects[1], rects[2], rects[3]), b1 = make_int4(rects[4], rects[5], rects[6], rects[7]);
  const int4 b2 = make_int4(rects[8], rects[9], rects[10], rects[11]);
  const int4 b3 = make_int4(rects[12], rects[13], rects[14], rects[15]);
  hipLaunchKernelGGL(k_border_blur, dim3((w + 255) / 256, h), dim3(256), 0, s, img, out, w, h, r, b0, b1, b2, b3);
}

void launch_border_nodes(const BorderJob *d_jobs, int frames, hipStream_t s) {
  hipLaunchKernelGGL(k_border_nodes, dim3(frames), dim3(NT), 0, s, d_jobs);
}

}  // namespace lz
