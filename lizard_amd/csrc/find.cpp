// The finder's geometry on the host: from the GPU's Hough peaks and edge probes to the registration (finder.hpp).
//
//   sides     each side of the border is its solid line: an outer edge (light outside, so its gradient points out) and
//             an inner edge two modules in (the light gap inside, gradient in), parallel, so a side is a pair of
//             peaks half a turn apart whose lines are a positive thickness apart; four sides a quarter turn apart
//             round a convex quad, best by votes
//   edges     probes along both edges of each side (the GPU), a robust quadratic through each edge's points
//   fit       every edge point lies on a known module line (depth 1 or 3 from its side), so G (pixel -> module, the
//             homography's inverse) is linear least squares in its 8 entries (one equation a point, reweighted to
//             module units); the radial lens by a golden-section search on k1 round it; each ring's module count tried
//   track     the timing track's cells (a per-side hash, liblizard/src/layout.c:57-62) read under each of the four
//             rotations: the one that agrees names the orientation, and the ring
//   refine    probes from the fitted model at every half module along both edges, and the fit again
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "finder.hpp"
#include "word.hpp"

namespace lz {
namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int RINGS[3] = {128, 256, 512};
constexpr int BAND_CELL = 2, RUN = 4, RESV = 15;

struct Side {
  double th, nx, ny;  // outward normal
  double ro, ri;      // outer edge n.(p - c) = ro, inner n.(p - c) = ri (ri < ro)
  double score;
};

struct Pt {
  double x, y, w;
  int side, depth;  // side label 0..3 (top, right, bottom, left), edge depth 1 or 3
};

double wrap_pi(double a) {
  while (a > PI) a -= 2 * PI;
  while (a < -PI) a += 2 * PI;
  return a;
}

// lens: p_d = p_u + k1 |q|^2 q D, q = (p_u - a) / D (undone by fixed-point iteration where the fit needs it)
struct Lens {
  double ax, ay, D, k1;
  void distort(double ux, double uy, double *dx, double *dy) const {
    const double qx = (ux - ax) / D, qy = (uy - ay) / D, r2 = qx * qx + qy * qy;
    *dx = ux + k1 * r2 * qx * D;
    *dy = uy + k1 * r2 * qy * D;
  }
};

// the module line a side's edge lies on: axis 0 (mx) or 1 (my) and its value
void edge_line(int L, int side, int depth, int *axis, double *val) {
  switch (side) {
    case 0: *axis = 1; *val = depth; break;
    case 1: *axis = 0; *val = L - depth; break;
    case 2: *axis = 1; *val = L - depth; break;
    default: *axis = 0; *val = depth; break;
  }
}

void side_xy(int L, int side, double t, double d, double *x, double *y) {
  *x = side == 0 || side == 2 ? t : side == 3 ? d : L - d;
  *y = side == 0 ? d : side == 2 ? L - d : t;
}

bool solve8(double A[8][8], double b[8], double x[8]) {
  for (int c = 0; c < 8; c++) {
    int p = c;
    for (int r = c + 1; r < 8; r++)
      if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
    if (std::fabs(A[p][c]) < 1e-300) return false;
    if (p != c) {
      for (int k = 0; k < 8; k++) std::swap(A[c][k], A[p][k]);
      std::swap(b[c], b[p]);
    }
    for (int r = c + 1; r < 8; r++) {
      const double f = A[r][c] / A[c][c];
      for (int k = c; k < 8; k++) A[r][k] -= f * A[c][k];
      b[r] -= f * b[c];
    }
  }
  for (int c = 7; c >= 0; c--) {
    double s = b[c];
    for (int k = c + 1; k < 8; k++) s -= A[c][k] * x[k];
    x[c] = s / A[c][c];
  }
  return true;
}

bool inv3(const double *M, double *R) {
  const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                     M[2] * (M[3] * M[7] - M[4] * M[6]);
  if (!(std::fabs(det) > 1e-300)) return false;
  R[0] = (M[4] * M[8] - M[5] * M[7]) / det;
  R[1] = (M[2] * M[7] - M[1] * M[8]) / det;
  R[2] = (M[1] * M[5] - M[2] * M[4]) / det;
  R[3] = (M[5] * M[6] - M[3] * M[8]) / det;
  R[4] = (M[0] * M[8] - M[2] * M[6]) / det;
  R[5] = (M[2] * M[3] - M[0] * M[5]) / det;
  R[6] = (M[3] * M[7] - M[4] * M[6]) / det;
  R[7] = (M[1] * M[6] - M[0] * M[7]) / det;
  R[8] = (M[0] * M[4] - M[1] * M[3]) / det;
  return true;
}

void mul3(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

bool solve9(double A[9][9], double b[9], double x[9]) {
  for (int c = 0; c < 9; c++) {
    int p = c;
    for (int r = c + 1; r < 9; r++)
      if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
    if (std::fabs(A[p][c]) < 1e-300) return false;
    if (p != c) {
      for (int k = 0; k < 9; k++) std::swap(A[c][k], A[p][k]);
      std::swap(b[c], b[p]);
    }
    for (int r = c + 1; r < 9; r++) {
      const double f = A[r][c] / A[c][c];
      for (int k = c; k < 9; k++) A[r][k] -= f * A[c][k];
      b[r] -= f * b[c];
    }
  }
  for (int c = 8; c >= 0; c--) {
    double t = b[c];
    for (int k = c + 1; k < 9; k++) t -= A[c][k] * x[k];
    x[c] = t / A[c][c];
  }
  return true;
}

// The homography's inverse G (normalised undistorted pixel -> module) and the lens k1 through the edge points. First
// G alone at the given k1, linear least squares in its 8 entries (one equation a point: the point's module coordinate
// across its edge is the edge's); then Gauss-Newton in all 9 (G and k1, the residuals in modules, a point's
// undistorted place moving by -|q|^2 q D per unit k1). Points further than `cut` modules from their line are dropped
// from the third round on. Returns the rms residual in modules (or a large value); H (module -> pixel, h8 = 1) and k1
// out.
struct Fit {
  double cx, cy, s;  // pixel normalisation
  double solve(const std::vector<Pt> &pts, Lens lens, int L, double cut, int rounds, double *H, double *k1_out,
               int *used) const {
    const int n = static_cast<int>(pts.size());
    std::vector<double> ux(n), uy(n), fx(n), fy(n), res(n, 0.0);
    std::vector<char> keep(n, 1);
    std::vector<int> axis(n);
    std::vector<double> val(n);
    for (int k = 0; k < n; k++) edge_line(L, pts[k].side, pts[k].depth, &axis[k], &val[k]);
    std::vector<double> px(n), py(n);  // undistorted places, each round starting from the last
    for (int k = 0; k < n; k++) {
      px[k] = pts[k].x;
      py[k] = pts[k].y;
    }
    int iters = 8;
    auto place = [&]() {
      for (int k = 0; k < n; k++) {
        double x = px[k], y = py[k];
        for (int it = 0; it < iters; it++) {
          const double qx = (x - lens.ax) / lens.D, qy = (y - lens.ay) / lens.D, r2 = qx * qx + qy * qy;
          x = pts[k].x - lens.k1 * r2 * qx * lens.D;
          y = pts[k].y - lens.k1 * r2 * qy * lens.D;
        }
        px[k] = x;
        py[k] = y;
        const double qx = (x - lens.ax) / lens.D, qy = (y - lens.ay) / lens.D, r2 = qx * qx + qy * qy;
        ux[k] = (x - cx) / s;
        uy[k] = (y - cy) / s;
        fx[k] = -r2 * qx * lens.D / s;
        fy[k] = -r2 * qy * lens.D / s;
      }
    };
    double g[8] = {};
    place();
    {
      double A[8][8] = {}, b[8] = {};
      int m = 0;
      for (int k = 0; k < n; k++) {
        const double c = val[k], wk = pts[k].w;
        const int o = axis[k] ? 3 : 0;
        const int id[5] = {o, o + 1, o + 2, 6, 7};
        const double v[5] = {ux[k], uy[k], 1.0, -c * ux[k], -c * uy[k]};
        for (int a = 0; a < 5; a++) {
          b[id[a]] += wk * v[a] * c;
          for (int e = 0; e < 5; e++) A[id[a]][id[e]] += wk * v[a] * v[e];
        }
        m++;
      }
      if (m < 16 || !solve8(A, b, g)) return 1e9;
    }
    double rms = 1e9;
    iters = 3;
    for (int round = 0; round < rounds; round++) {
      if (round) place();
      double A[9][9] = {}, b[9] = {};
      double s2 = 0, sw = 0;
      int m = 0;
      for (int k = 0; k < n; k++) {
        const double d = g[6] * ux[k] + g[7] * uy[k] + 1.0, den = std::fabs(d) > 1e-9 ? d : 1e-9;
        const int o = axis[k] ? 3 : 0;
        const double *ga = g + o;
        const double mod = (ga[0] * ux[k] + ga[1] * uy[k] + ga[2]) / den;
        res[k] = mod - val[k];
        if (round >= 2) keep[k] = std::fabs(res[k]) <= cut;
        if (!keep[k]) continue;
        // d residual / d (g[o..o+2], g6, g7, k1)
        const double jx = (ga[0] - mod * g[6]) / den, jy = (ga[1] - mod * g[7]) / den;
        const int id[6] = {o, o + 1, o + 2, 6, 7, 8};
        const double J[6] = {ux[k] / den, uy[k] / den, 1.0 / den, -mod * ux[k] / den, -mod * uy[k] / den,
                             jx * fx[k] + jy * fy[k]};
        const double w = pts[k].w;
        for (int a = 0; a < 6; a++) {
          b[id[a]] -= w * J[a] * res[k];
          for (int e = 0; e < 6; e++) A[id[a]][id[e]] += w * J[a] * J[e];
        }
        s2 += w * res[k] * res[k];
        sw += w;
        m++;
      }
      rms = sw > 0 ? std::sqrt(s2 / sw) : 1e9;
      if (m < 16) return 1e9;
      double dx[9];
      if (!solve9(A, b, dx)) break;
      for (int i = 0; i < 8; i++) g[i] += dx[i];
      lens.k1 = std::max(-0.25, std::min(0.25, lens.k1 + std::max(-0.05, std::min(0.05, dx[8]))));
    }
    int u = 0;
    for (int k = 0; k < n; k++) u += keep[k];
    if (used) *used = u;
    if (k1_out) *k1_out = lens.k1;
    const double G[9] = {g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], 1.0};
    double Hn[9];
    if (!inv3(G, Hn)) return 1e9;
    const double S[9] = {s, 0, cx, 0, s, cy, 0, 0, 1};
    double Hp[9];
    mul3(S, Hn, Hp);
    for (int i = 0; i < 9; i++) H[i] = Hp[i] / Hp[8];
    return rms;
  }
};

// bilinear grey read, pixel centres at half integers, clamped
struct Img {
  const uint8_t *p;
  int w, h, stride;
  double at(double x, double y) const {
    x -= 0.5;
    y -= 0.5;
    x = x < 0 ? 0 : x > w - 1.001 ? w - 1.001 : x;
    y = y < 0 ? 0 : y > h - 1.001 ? h - 1.001 : y;
    const int xi = static_cast<int>(x), yi = static_cast<int>(y);
    const double fx = x - xi, fy = y - yi;
    const uint8_t *q = p + static_cast<size_t>(yi) * stride + xi;
    return (q[0] + (q[1] - q[0]) * fx) * (1 - fy) + (q[stride] + (q[stride + 1] - q[stride]) * fx) * fy;
  }
};

void map_pt(const double *H, const Lens &lens, double mx, double my, double *x, double *y) {
  const double W = H[6] * mx + H[7] * my + H[8];
  lens.distort((H[0] * mx + H[1] * my + H[2]) / W, (H[3] * mx + H[4] * my + H[5]) / W, x, y);
}

// The timing track's agreement under H: the mean of +-(cell - the track's mean level) by the cell's expected colour,
// over the border's contrast (dark line against gap); about 1 where H is right, about 0 where it is turned.
double track_score(const Img &im, const double *H, const Lens &lens, int L, float *contrast) {
  double dark = 0, light = 0;
  int nref = 0;
  for (int side = 0; side < 4; side++)
    for (int t = RESV; t < L - RESV; t += 16, nref++) {
      double mx, my, x, y;
      side_xy(L, side, t + 0.5, 2.0, &mx, &my);
      map_pt(H, lens, mx, my, &x, &y);
      dark += im.at(x, y);
      side_xy(L, side, t + 0.5, 4.0, &mx, &my);
      map_pt(H, lens, mx, my, &x, &y);
      light += im.at(x, y);
    }
  const double con = (light - dark) / (nref ? nref : 1);
  if (contrast) *contrast = static_cast<float>(con / 255.0);
  if (!(con > 8.0)) return 0.0;
  const int cells = (L > 2 * RESV ? (L - 2 * RESV) / BAND_CELL : 0), ntrack = (cells + 1) / 2;
  std::vector<double> v;
  std::vector<int> dk;
  for (int side = 0; side < 4; side++)
    for (int j = 0; j < ntrack; j++) {
      const int pair = (j / RUN) * 2 * RUN + j % RUN;
      double mx, my, x, y;
      side_xy(L, side, RESV + pair * BAND_CELL + 1.0, 6.0, &mx, &my);
      map_pt(H, lens, mx, my, &x, &y);
      v.push_back(im.at(x, y));
      dk.push_back(thin_track(side, j));
    }
  if (v.empty()) return 0.0;
  double mean = 0;
  for (double a : v) mean += a;
  mean /= v.size();
  double s = 0;
  for (size_t k = 0; k < v.size(); k++) s += dk[k] ? mean - v[k] : v[k] - mean;
  return s / v.size() / (0.5 * con);
}

// Robust quadratic through an edge's points in its side's frame (t along, o across): inliers kept in place.
void robust_edge(std::vector<Pt> &pts, double ox, double oy, double tx, double ty) {
  if (pts.size() < 8) {
    pts.clear();
    return;
  }
  std::vector<char> keep(pts.size(), 1);
  double cut = 1e9;
  for (int round = 0; round < 4; round++) {
    double A[3][3] = {}, b[3] = {};
    for (size_t k = 0; k < pts.size(); k++) {
      if (!keep[k]) continue;
      const double t = (pts[k].x - ox) * tx + (pts[k].y - oy) * ty, o = (pts[k].x - ox) * -ty + (pts[k].y - oy) * tx;
      const double ts = t / 512.0, r[3] = {1.0, ts, ts * ts}, w = pts[k].w;
      for (int i = 0; i < 3; i++) {
        b[i] += w * r[i] * o;
        for (int j = 0; j < 3; j++) A[i][j] += w * r[i] * r[j];
      }
    }
    // 3 x 3 solve
    double M[3][4];
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) M[i][j] = A[i][j];
      M[i][3] = b[i];
    }
    bool ok = true;
    for (int c = 0; c < 3 && ok; c++) {
      int p = c;
      for (int r = c + 1; r < 3; r++)
        if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
      if (std::fabs(M[p][c]) < 1e-12) ok = false;
      else {
        for (int k = 0; k < 4; k++) std::swap(M[c][k], M[p][k]);
        for (int r = 0; r < 3; r++)
          if (r != c) {
            const double f = M[r][c] / M[c][c];
            for (int k = c; k < 4; k++) M[r][k] -= f * M[c][k];
          }
      }
    }
    if (!ok) {
      pts.clear();
      return;
    }
    const double a0 = M[0][3] / M[0][0], a1 = M[1][3] / M[1][1], a2 = M[2][3] / M[2][2];
    std::vector<double> ad;
    for (size_t k = 0; k < pts.size(); k++) {
      const double t = (pts[k].x - ox) * tx + (pts[k].y - oy) * ty, o = (pts[k].x - ox) * -ty + (pts[k].y - oy) * tx;
      const double ts = t / 512.0, r = o - (a0 + a1 * ts + a2 * ts * ts);
      keep[k] = std::fabs(r) <= cut;
      ad.push_back(std::fabs(r));
    }
    std::vector<double> srt = ad;
    std::nth_element(srt.begin(), srt.begin() + srt.size() / 2, srt.end());
    cut = std::max(0.35, 3.0 * 1.4826 * srt[srt.size() / 2]);
    for (size_t k = 0; k < pts.size(); k++) keep[k] = ad[k] <= cut;
  }
  std::vector<Pt> out;
  for (size_t k = 0; k < pts.size(); k++)
    if (keep[k]) out.push_back(pts[k]);
  pts.swap(out);
}

}  // namespace

void find_symbol(FinderBufs &b, const uint8_t *d_img, const uint8_t *img, int w, int h, int stride, hipStream_t s,
                 FindResult &r) {
  const int d_stride = w;  // the device copy's rows are tight
  const float upload = r.ms[0];
  r = FindResult{};
  r.ms[0] = upload;
  auto t0 = std::chrono::steady_clock::now();
  auto lap = [&](int k) {
    const auto t1 = std::chrono::steady_clock::now();
    r.ms[k] = std::chrono::duration<float, std::milli>(t1 - t0).count();
    t0 = t1;
  };
  const int R = static_cast<int>(std::ceil(0.5 * std::hypot(w, h))) + 2, nr = 2 * R + 1;
  const size_t cells = static_cast<size_t>(FIND_THETA) * nr;
  b.acc.reserve(cells * sizeof(float));
  b.max.reserve(256);
  b.count.reserve(256);
  b.peaks.reserve(FIND_MAX_PEAKS * sizeof(FindPeak));
  if (b.trig_nr != 1) {
    std::vector<float> tr(2 * FIND_THETA);
    for (int t = 0; t < FIND_THETA; t++) {
      tr[2 * t] = static_cast<float>(std::cos(2 * PI * t / FIND_THETA));
      tr[2 * t + 1] = static_cast<float>(std::sin(2 * PI * t / FIND_THETA));
    }
    b.trig.reserve(tr.size() * sizeof(float));
    hip_check(hipStreamSynchronize(s), "finder stream");
    hip_check(hipMemcpy(b.trig.p, tr.data(), tr.size() * sizeof(float), hipMemcpyHostToDevice), "trig");
    b.trig_nr = 1;
  }
  launch_find_zero(b.acc.as<float>(), cells, b.max.as<unsigned>(), b.count.as<int>(), s);
  launch_find_vote(d_img, w, h, d_stride, 12.0f, b.trig.as<float2>(), nr, b.acc.as<float>(), s);
  launch_find_peaks(b.acc.as<float>(), nr, 0.08f, b.max.as<unsigned>(), b.peaks.as<FindPeak>(), b.count.as<int>(), s);
  hip_check(hipGetLastError(), "finder launches");
  int np = 0;
  hip_check(hipMemcpyAsync(&np, b.count.p, sizeof(int), hipMemcpyDeviceToHost, s), "peak count");
  hip_check(hipStreamSynchronize(s), "finder peaks");
  np = std::min(np, FIND_MAX_PEAKS);
  std::vector<FindPeak> pk(np);
  if (np) hip_check(hipMemcpy(pk.data(), b.peaks.p, np * sizeof(FindPeak), hipMemcpyDeviceToHost), "peaks");
  lap(1);
  std::sort(pk.begin(), pk.end(), [](const FindPeak &a, const FindPeak &c) { return a.score > c.score; });
  if (pk.size() > 96) pk.resize(96);

  // sides: an outer edge (gradient out) and an inner edge (gradient in) half a turn apart, a positive thickness apart
  const double cx = 0.5 * w, cy = 0.5 * h, dth = 2.5 * 2 * PI / FIND_THETA;
  std::vector<Side> sides;
  for (size_t a = 0; a < pk.size(); a++)
    for (size_t c = 0; c < pk.size(); c++) {
      if (a == c) continue;
      if (std::fabs(wrap_pi(pk[c].theta - pk[a].theta - PI)) > dth) continue;
      const double t = pk[a].rho + pk[c].rho;
      if (t < 1.2 || t > 80.0) continue;
      Side sd;
      sd.th = pk[a].theta;
      sd.nx = std::cos(sd.th);
      sd.ny = std::sin(sd.th);
      sd.ro = pk[a].rho;
      sd.ri = -pk[c].rho;
      sd.score = pk[a].score + pk[c].score;
      sides.push_back(sd);
    }
  std::sort(sides.begin(), sides.end(), [](const Side &a, const Side &c) { return a.score > c.score; });
  if (sides.size() > 40) sides.resize(40);
  if (sides.size() < 4) return;
  // four sides a quarter turn apart (within 30 degrees) round a convex quad that holds the others' midlines' corners
  auto mid = [](const Side &sd) { return 0.5 * (sd.ro + sd.ri); };
  auto meet = [&](const Side &p, const Side &q, double *x, double *y) {
    const double det = p.nx * q.ny - p.ny * q.nx;
    if (std::fabs(det) < 1e-6) return false;
    const double a = mid(p), c = mid(q);
    *x = cx + (a * q.ny - c * p.ny) / det;
    *y = cy + (p.nx * c - q.nx * a) / det;
    return true;
  };
  double best = -1;
  int pick[4] = {0, 0, 0, 0};
  const int m = static_cast<int>(sides.size());
  for (int i0 = 0; i0 < m; i0++)
    for (int i1 = i0 + 1; i1 < m; i1++)
      for (int i2 = i1 + 1; i2 < m; i2++)
        for (int i3 = i2 + 1; i3 < m; i3++) {
          const double sc = sides[i0].score + sides[i1].score + sides[i2].score + sides[i3].score;
          if (sc <= best) continue;
          int id[4] = {i0, i1, i2, i3};
          std::sort(id, id + 4, [&](int a, int c) { return sides[a].th < sides[c].th; });
          bool ok = true;
          double tmin = 1e9, tmax = 0;
          for (int k = 0; k < 4 && ok; k++) {
            const double d = wrap_pi(sides[id[(k + 1) & 3]].th - sides[id[k]].th);
            if (std::fabs(d - PI / 2) > PI / 6) ok = false;
            const double t = sides[id[k]].ro - sides[id[k]].ri;
            tmin = std::min(tmin, t);
            tmax = std::max(tmax, t);
          }
          if (!ok || tmax > 3.0 * tmin) continue;
          double qx[4], qy[4];
          for (int k = 0; k < 4 && ok; k++) ok = meet(sides[id[k]], sides[id[(k + 1) & 3]], &qx[k], &qy[k]);
          if (!ok) continue;
          const double mx = 0.25 * (qx[0] + qx[1] + qx[2] + qx[3]), my = 0.25 * (qy[0] + qy[1] + qy[2] + qy[3]);
          double smin = 1e9;
          for (int k = 0; k < 4 && ok; k++) {
            const Side &sd = sides[id[k]];
            if (sd.nx * (mx - cx) + sd.ny * (my - cy) >= mid(sd)) ok = false;  // the centre inside every side
            smin = std::min(smin, std::hypot(qx[(k + 1) & 3] - qx[k], qy[(k + 1) & 3] - qy[k]));
          }
          if (!ok || smin < 30.0 || smin < 20.0 * tmin) continue;
          best = sc;
          std::memcpy(pick, id, sizeof pick);
        }
  lap(2);
  if (best < 0) return;
  // labels: the side whose outward normal is nearest straight up is side 0 (top), then clockwise
  Side S4[4];
  {
    int top = 0;
    double bd = 1e9;
    for (int k = 0; k < 4; k++) {
      const double d = std::fabs(wrap_pi(sides[pick[k]].th + PI / 2));
      if (d < bd) {
        bd = d;
        top = k;
      }
    }
    for (int k = 0; k < 4; k++) S4[k] = sides[pick[(top + k) & 3]];
  }
  // corners of the midlines, TL = top & left, TR = top & right, BR, BL
  double Cx[4], Cy[4];
  meet(S4[0], S4[3], &Cx[0], &Cy[0]);
  meet(S4[0], S4[1], &Cx[1], &Cy[1]);
  meet(S4[2], S4[1], &Cx[2], &Cy[2]);
  meet(S4[2], S4[3], &Cx[3], &Cy[3]);

  b.probes.reserve(FIND_MAX_PROBES * sizeof(FindProbe));
  b.edges.reserve(FIND_MAX_PROBES * sizeof(FindEdge));
  std::vector<FindProbe> pr;
  std::vector<FindEdge> ed;
  auto run_probes = [&](float reach) {
    const int n = std::min(static_cast<int>(pr.size()), FIND_MAX_PROBES);
    if (n <= 0) return;
    hip_check(hipMemcpyAsync(b.probes.p, pr.data(), n * sizeof(FindProbe), hipMemcpyHostToDevice, s), "probes");
    launch_find_probe(d_img, w, h, d_stride, b.probes.as<FindProbe>(), n, reach, b.edges.as<FindEdge>(), s);
    hip_check(hipGetLastError(), "probe launch");
    ed.resize(n);
    hip_check(hipMemcpyAsync(ed.data(), b.edges.p, n * sizeof(FindEdge), hipMemcpyDeviceToHost, s), "edges");
    hip_check(hipStreamSynchronize(s), "probe sync");
  };

  // pass 1: probes along each side's straight edges between the corners, a robust quadratic through each edge
  std::vector<Pt> pts;
  const int ends[4][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}};  // the side's corners (top TL-TR, right TR-BR, ...)
  double thick = 0;
  for (int sd = 0; sd < 4; sd++) thick += 0.25 * (S4[sd].ro - S4[sd].ri);
  for (int pass = 0; pass < 2; pass++) {
    pr.clear();
    std::vector<int> lab;
    for (int sd = 0; sd < 4; sd++) {
      const Side &q = S4[sd];
      const double x0 = Cx[ends[sd][0]], y0 = Cy[ends[sd][0]], x1 = Cx[ends[sd][1]], y1 = Cy[ends[sd][1]];
      const double len = std::hypot(x1 - x0, y1 - y0), tx = (x1 - x0) / len, ty = (y1 - y0) / len;
      const double margin = std::max(4.0 * thick, 8.0);
      const int cnt = std::max(0, std::min(4000, static_cast<int>((len - 2 * margin) / 1.5)));
      for (int e = 0; e < 2; e++) {
        const double half = 0.5 * (q.ro - q.ri) * (e == 0 ? 1 : -1);
        for (int k = 0; k < cnt; k++) {
          const double t = margin + (len - 2 * margin) * (k + 0.5) / cnt;
          FindProbe p;
          p.x = static_cast<float>(x0 + t * tx + half * q.nx);
          p.y = static_cast<float>(y0 + t * ty + half * q.ny);
          p.nx = static_cast<float>(e == 0 ? q.nx : -q.nx);
          p.ny = static_cast<float>(e == 0 ? q.ny : -q.ny);
          pr.push_back(p);
          lab.push_back(sd * 2 + e);
        }
      }
    }
    run_probes(static_cast<float>(std::min(6.0, std::max(1.0, 0.75 * thick))));
    // per edge, the robust quadratic; the side's straight line moved to the edges' mean for the next pass
    pts.clear();
    for (int sd = 0; sd < 4; sd++)
      for (int e = 0; e < 2; e++) {
        std::vector<Pt> ep;
        for (size_t k = 0; k < ed.size(); k++)
          if (lab[k] == sd * 2 + e && ed[k].w > 0) ep.push_back(Pt{ed[k].x, ed[k].y, ed[k].w, sd, e ? 3 : 1});
        const double x0 = Cx[ends[sd][0]], y0 = Cy[ends[sd][0]], x1 = Cx[ends[sd][1]], y1 = Cy[ends[sd][1]];
        const double len = std::hypot(x1 - x0, y1 - y0);
        robust_edge(ep, x0, y0, (x1 - x0) / len, (y1 - y0) / len);
        if (pass == 0 && !ep.empty()) {
          double so = 0, sw = 0;
          for (const Pt &p : ep) {
            so += p.w * (S4[sd].nx * (p.x - cx) + S4[sd].ny * (p.y - cy));
            sw += p.w;
          }
          if (e == 0) S4[sd].ro = so / sw;
          else S4[sd].ri = so / sw;
        }
        pts.insert(pts.end(), ep.begin(), ep.end());
      }
    if (pass == 0) {
      meet(S4[0], S4[3], &Cx[0], &Cy[0]);
      meet(S4[0], S4[1], &Cx[1], &Cy[1]);
      meet(S4[2], S4[1], &Cx[2], &Cy[2]);
      meet(S4[2], S4[3], &Cx[3], &Cy[3]);
    }
  }
  lap(3);
  if (pts.size() < 64) return;

  // the fit for each ring, the lens by golden section; then the rotation by the track
  const Img im{img, w, h, stride};
  Fit fit{cx, cy, 0.5 * std::max(w, h)};
  Lens lens{cx, cy, std::hypot(cx, cy), 0.0};
  double bestH[9] = {}, bestk1 = 0, bestTrack = -1;
  int bestL = 0;
  std::vector<Pt> sub;  // every fourth point for the search
  for (size_t k = 0; k < pts.size(); k += 4) sub.push_back(pts[k]);
  // rings in the order the line's thickness (2 modules) against the side's length (L - 4 modules between the
  // midlines) suggests
  int order[3] = {0, 1, 2};
  {
    const double side = 0.25 * (std::hypot(Cx[1] - Cx[0], Cy[1] - Cy[0]) + std::hypot(Cx[2] - Cx[1], Cy[2] - Cy[1]) +
                                std::hypot(Cx[3] - Cx[2], Cy[3] - Cy[2]) + std::hypot(Cx[0] - Cx[3], Cy[0] - Cy[3]));
    const double modules = thick > 0 ? 2.0 * side / thick + 4.0 : 0.0;
    std::sort(order, order + 3, [&](int a, int c) {
      return std::fabs(std::log((RINGS[a] + 30) / modules)) < std::fabs(std::log((RINGS[c] + 30) / modules));
    });
  }
  for (int oi = 0; oi < 3; oi++) {
    if (bestTrack >= 0.7) break;  // a clear track under the likelier ring
    const int L = RINGS[order[oi]] + 30;
    Lens ln = lens;
    double H[9];
    const double rms = fit.solve(sub, ln, L, 1.0, 6, H, &ln.k1, nullptr);
    if (!(rms < 0.5)) continue;
    // the four rotations about the symbol's centre: module (x, y) -> (L - y, x) a quarter turn
    double Rq[9] = {0, -1, static_cast<double>(L), 1, 0, 0, 0, 0, 1}, Rk[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int rot = 0; rot < 4; rot++) {
      double Hr[9];
      mul3(H, Rk, Hr);
      for (int i = 0; i < 9; i++) Hr[i] /= Hr[8];
      float con = 0;
      const double sc = track_score(im, Hr, ln, L, &con);
      if (sc > bestTrack) {
        bestTrack = sc;
        std::memcpy(bestH, Hr, sizeof bestH);
        bestk1 = ln.k1;
        bestL = L;
      }
      double T[9];
      mul3(Rk, Rq, T);
      std::memcpy(Rk, T, sizeof T);
    }
  }
  lap(4);
  if (bestTrack < 0.4) return;
  const int L = bestL;

  // refine: probes from the model at every module along both edges of every side, and the fit again, twice
  lens.k1 = bestk1;
  std::memcpy(r.H, bestH, sizeof bestH);
  for (int pass = 0; pass < 2; pass++) {
    pr.clear();
    std::vector<Pt> lab;
    const double mod = std::hypot(r.H[0], r.H[3]);  // pixels a module near the origin, for the reach
    for (int sd = 0; sd < 4; sd++)
      for (int e = 0; e < 2; e++) {
        const int depth = e ? 3 : 1;
        for (double t = 3.0; t <= L - 3.0; t += 1.0) {
          double mx, my, x, y, ax, ay, bx, by;
          side_xy(L, sd, t, depth, &mx, &my);
          map_pt(r.H, lens, mx, my, &x, &y);
          // rising intensity: outward at depth 1 (light module outside), inward at depth 3 (the gap)
          side_xy(L, sd, t, depth - 0.25, &mx, &my);
          map_pt(r.H, lens, mx, my, &ax, &ay);
          side_xy(L, sd, t, depth + 0.25, &mx, &my);
          map_pt(r.H, lens, mx, my, &bx, &by);
          const double nx = e ? bx - ax : ax - bx, ny = e ? by - ay : ay - by, nn = std::hypot(nx, ny);
          if (!(nn > 0)) continue;
          pr.push_back(FindProbe{static_cast<float>(x), static_cast<float>(y), static_cast<float>(nx / nn),
                                 static_cast<float>(ny / nn)});
          lab.push_back(Pt{0, 0, 0, sd, depth});
        }
      }
    run_probes(static_cast<float>(std::min(3.0, std::max(0.75, (pass ? 0.6 : 0.9) * mod))));
    pts.clear();
    for (size_t k = 0; k < ed.size(); k++)
      if (ed[k].w > 0) pts.push_back(Pt{ed[k].x, ed[k].y, ed[k].w, lab[k].side, lab[k].depth});
    double H[9], k1 = lens.k1;
    int used = 0;
    const double rms = fit.solve(pts, lens, L, 0.3, 5, H, &k1, &used);
    if (!(rms < 0.3)) return;
    std::memcpy(r.H, H, sizeof H);
    lens.k1 = k1;
    r.k1 = k1;
    r.rms = static_cast<float>(rms);
    r.points = used;
  }
  lap(5);
  r.span = L - 30;
  r.found = 1;
  r.track = static_cast<float>(track_score(im, r.H, lens, L, &r.contrast));
  const double cm[4][2] = {{1, 1}, {L - 1.0, 1}, {L - 1.0, L - 1.0}, {1, L - 1.0}};
  for (int c = 0; c < 4; c++) {
    double x, y;
    map_pt(r.H, lens, cm[c][0], cm[c][1], &x, &y);
    r.quad[2 * c] = static_cast<float>(x);
    r.quad[2 * c + 1] = static_cast<float>(y);
  }
}

}  // namespace lz
