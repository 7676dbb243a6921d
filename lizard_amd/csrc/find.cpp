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
#include <condition_variable>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "anchor.hpp"
#include "finder.hpp"
#include "word.hpp"

namespace lz {
namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int RINGS[3] = {128, 256, 512};
constexpr int BAND_CELL = 2, RUN = 4, RESV = 15;
constexpr int EDGE_PROBES = 4000;  // the most probes along one edge of a side

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

// ---- the search as stages over one frame's state, so that many frames can share each stage's launches ----

struct FinderPool {
  std::vector<std::thread> threads;
  std::mutex m;
  std::condition_variable cv, done_cv;
  std::function<void(int, int)> task;  // (item, thread)
  int count = 0, next = 0, done = 0;
  uint64_t gen = 0;
  bool stop = false;
  explicit FinderPool(int n) {
    for (int t = 0; t < n; t++)
      threads.emplace_back([this, t] {
        uint64_t seen = 0;
        for (;;) {
          std::unique_lock<std::mutex> lk(m);
          cv.wait(lk, [&] { return stop || gen != seen; });
          if (stop) return;
          seen = gen;
          while (next < count) {
            const int i = next++;
            lk.unlock();
            task(i, t);
            lk.lock();
            if (++done == count) done_cv.notify_all();
          }
        }
      });
  }
  ~FinderPool() {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    cv.notify_all();
    for (auto &t : threads) t.join();
  }
  // fn(0 .. n - 1, thread) on the pool's threads; returns when every call has
  void run(int n, const std::function<void(int, int)> &fn) {
    if (n <= 0) return;
    std::unique_lock<std::mutex> lk(m);
    task = fn;
    count = n;
    next = 0;
    done = 0;
    gen++;
    cv.notify_all();
    done_cv.wait(lk, [&] { return done == count; });
  }
};

namespace {

constexpr int ENDS[4][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}};  // a side's corners (top TL-TR, right TR-BR, ...)

struct FindJob {
  const uint8_t *img = nullptr;  // host copy
  int w = 0, h = 0, stride = 0, frame = 0;
  double cx = 0, cy = 0;
  std::vector<FindPeak> pk;
  Side S4[4];
  double Cx[4] = {}, Cy[4] = {}, thick = 0;
  std::vector<FindProbe> pr;  // this stage's probes
  std::vector<int> lab1;      // straight-edge probes' labels: side * 2 + edge
  std::vector<Pt> lab2;       // model probes' labels: side and depth
  std::vector<Pt> pts;
  Fit fit{};
  Lens lens{};
  int L = 0;
  bool dead = false;
  FindResult *r = nullptr;
};

bool meet_sides(const Side &p, const Side &q, double cx, double cy, double *x, double *y) {
  const double det = p.nx * q.ny - p.ny * q.nx;
  if (std::fabs(det) < 1e-6) return false;
  const double a = 0.5 * (p.ro + p.ri), c = 0.5 * (q.ro + q.ri);
  *x = cx + (a * q.ny - c * p.ny) / det;
  *y = cy + (p.nx * c - q.nx * a) / det;
  return true;
}

void corners(FindJob &j) {
  meet_sides(j.S4[0], j.S4[3], j.cx, j.cy, &j.Cx[0], &j.Cy[0]);
  meet_sides(j.S4[0], j.S4[1], j.cx, j.cy, &j.Cx[1], &j.Cy[1]);
  meet_sides(j.S4[2], j.S4[1], j.cx, j.cy, &j.Cx[2], &j.Cy[2]);
  meet_sides(j.S4[2], j.S4[3], j.cx, j.cy, &j.Cx[3], &j.Cy[3]);
}

// sides: an outer edge (gradient out) and an inner edge (gradient in) half a turn apart, a positive thickness apart;
// four of them a quarter turn apart (within 30 degrees) round a convex quad that holds its centre, best by votes
void stage_sides(FindJob &j) {
  std::vector<FindPeak> &pk = j.pk;
  // by score, ties by votes, direction, distance: an order that does not depend on the peaks' arrival
  std::sort(pk.begin(), pk.end(), [](const FindPeak &a, const FindPeak &c) {
    if (a.score != c.score) return a.score > c.score;
    if (a.votes != c.votes) return a.votes > c.votes;
    if (a.theta != c.theta) return a.theta < c.theta;
    return a.rho < c.rho;
  });
  if (pk.size() > 96) pk.resize(96);
  const double cx = j.cx, cy = j.cy, dth = 2.5 * 2 * PI / FIND_THETA;
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
  std::sort(sides.begin(), sides.end(), [](const Side &a, const Side &c) {
    if (a.score != c.score) return a.score > c.score;
    if (a.th != c.th) return a.th < c.th;
    if (a.ro != c.ro) return a.ro < c.ro;
    return a.ri < c.ri;
  });
  if (sides.size() > 40) sides.resize(40);
  if (sides.size() < 4) {
    j.dead = true;
    return;
  }
  auto mid = [](const Side &sd) { return 0.5 * (sd.ro + sd.ri); };
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
          for (int k = 0; k < 4 && ok; k++) ok = meet_sides(sides[id[k]], sides[id[(k + 1) & 3]], cx, cy, &qx[k], &qy[k]);
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
  if (best < 0) {
    j.dead = true;
    return;
  }
  // labels: the side whose outward normal is nearest straight up is side 0 (top), then clockwise
  int top = 0;
  double bd = 1e9;
  for (int k = 0; k < 4; k++) {
    const double d = std::fabs(wrap_pi(sides[pick[k]].th + PI / 2));
    if (d < bd) {
      bd = d;
      top = k;
    }
  }
  for (int k = 0; k < 4; k++) j.S4[k] = sides[pick[(top + k) & 3]];
  corners(j);  // the midlines' corners: TL = top & left, TR = top & right, BR, BL
  j.thick = 0;
  for (int sd = 0; sd < 4; sd++) j.thick += 0.25 * (j.S4[sd].ro - j.S4[sd].ri);
}

// probes along each side's straight edges between the corners
void edge_probes(FindJob &j) {
  j.pr.clear();
  j.lab1.clear();
  const float reach = static_cast<float>(std::min(6.0, std::max(1.0, 0.75 * j.thick)));
  for (int sd = 0; sd < 4; sd++) {
    const Side &q = j.S4[sd];
    const double x0 = j.Cx[ENDS[sd][0]], y0 = j.Cy[ENDS[sd][0]], x1 = j.Cx[ENDS[sd][1]], y1 = j.Cy[ENDS[sd][1]];
    const double len = std::hypot(x1 - x0, y1 - y0), tx = (x1 - x0) / len, ty = (y1 - y0) / len;
    const double margin = std::max(4.0 * j.thick, 8.0);
    const int cnt = std::max(0, std::min(EDGE_PROBES, static_cast<int>((len - 2 * margin) / 1.5)));
    for (int e = 0; e < 2; e++) {
      const double half = 0.5 * (q.ro - q.ri) * (e == 0 ? 1 : -1);
      for (int k = 0; k < cnt; k++) {
        const double t = margin + (len - 2 * margin) * (k + 0.5) / cnt;
        FindProbe p;
        p.x = static_cast<float>(x0 + t * tx + half * q.nx);
        p.y = static_cast<float>(y0 + t * ty + half * q.ny);
        p.nx = static_cast<float>(e == 0 ? q.nx : -q.nx);
        p.ny = static_cast<float>(e == 0 ? q.ny : -q.ny);
        p.reach = reach;
        p.frame = j.frame;
        j.pr.push_back(p);
        j.lab1.push_back(sd * 2 + e);
      }
    }
  }
}

// per edge, a robust quadratic through its points; after the first pass the sides' straight lines move to their
// edges' means and the corners with them
void edge_fits(FindJob &j, const FindEdge *ed, int pass) {
  j.pts.clear();
  const int n = static_cast<int>(j.lab1.size());
  for (int sd = 0; sd < 4; sd++)
    for (int e = 0; e < 2; e++) {
      std::vector<Pt> ep;
      for (int k = 0; k < n; k++)
        if (j.lab1[k] == sd * 2 + e && ed[k].w > 0) ep.push_back(Pt{ed[k].x, ed[k].y, ed[k].w, sd, e ? 3 : 1});
      const double x0 = j.Cx[ENDS[sd][0]], y0 = j.Cy[ENDS[sd][0]], x1 = j.Cx[ENDS[sd][1]], y1 = j.Cy[ENDS[sd][1]];
      const double len = std::hypot(x1 - x0, y1 - y0);
      robust_edge(ep, x0, y0, (x1 - x0) / len, (y1 - y0) / len);
      if (pass == 0 && !ep.empty()) {
        double so = 0, sw = 0;
        for (const Pt &p : ep) {
          so += p.w * (j.S4[sd].nx * (p.x - j.cx) + j.S4[sd].ny * (p.y - j.cy));
          sw += p.w;
        }
        if (e == 0) j.S4[sd].ro = so / sw;
        else j.S4[sd].ri = so / sw;
      }
      j.pts.insert(j.pts.end(), ep.begin(), ep.end());
    }
  if (pass == 0) corners(j);
}

// the fit for each ring (in the order the line's thickness, 2 modules, against the side's length, L - 4 modules
// between the midlines, suggests), the lens within it; the rotation by the track
void stage_rings(FindJob &j) {
  if (j.pts.size() < 64) {
    j.dead = true;
    return;
  }
  const Img im{j.img, j.w, j.h, j.stride};
  j.fit = Fit{j.cx, j.cy, 0.5 * std::max(j.w, j.h)};
  j.lens = Lens{j.cx, j.cy, std::hypot(j.cx, j.cy), 0.0};
  double bestH[9] = {}, bestk1 = 0, bestTrack = -1;
  int bestL = 0;
  std::vector<Pt> sub;  // every fourth point for the search
  for (size_t k = 0; k < j.pts.size(); k += 4) sub.push_back(j.pts[k]);
  int order[3] = {0, 1, 2};
  {
    const double *Cx = j.Cx, *Cy = j.Cy;
    const double side = 0.25 * (std::hypot(Cx[1] - Cx[0], Cy[1] - Cy[0]) + std::hypot(Cx[2] - Cx[1], Cy[2] - Cy[1]) +
                                std::hypot(Cx[3] - Cx[2], Cy[3] - Cy[2]) + std::hypot(Cx[0] - Cx[3], Cy[0] - Cy[3]));
    const double modules = j.thick > 0 ? 2.0 * side / j.thick + 4.0 : 0.0;
    std::sort(order, order + 3, [&](int a, int c) {
      return std::fabs(std::log((RINGS[a] + 30) / modules)) < std::fabs(std::log((RINGS[c] + 30) / modules));
    });
  }
  for (int oi = 0; oi < 3; oi++) {
    if (bestTrack >= 0.7) break;  // a clear track under the likelier ring
    const int L = RINGS[order[oi]] + 30;
    Lens ln = j.lens;
    double H[9];
    const double rms = j.fit.solve(sub, ln, L, 1.0, 6, H, &ln.k1, nullptr);
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
  if (bestTrack < 0.4) {
    j.dead = true;
    return;
  }
  j.L = bestL;
  j.lens.k1 = bestk1;
  std::memcpy(j.r->H, bestH, sizeof bestH);
}

// probes from the model at every module along both edges of every side
void model_probes(FindJob &j, int pass) {
  j.pr.clear();
  j.lab2.clear();
  const double *H = j.r->H;
  const int L = j.L;
  const double mod = std::hypot(H[0], H[3]);  // pixels a module near the origin, for the reach
  const float reach = static_cast<float>(std::min(3.0, std::max(0.75, (pass ? 0.6 : 0.9) * mod)));
  for (int sd = 0; sd < 4; sd++)
    for (int e = 0; e < 2; e++) {
      const int depth = e ? 3 : 1;
      for (double t = 3.0; t <= L - 3.0; t += 1.0) {
        double mx, my, x, y, ax, ay, bx, by;
        side_xy(L, sd, t, depth, &mx, &my);
        map_pt(H, j.lens, mx, my, &x, &y);
        // rising intensity: outward at depth 1 (light module outside), inward at depth 3 (the gap)
        side_xy(L, sd, t, depth - 0.25, &mx, &my);
        map_pt(H, j.lens, mx, my, &ax, &ay);
        side_xy(L, sd, t, depth + 0.25, &mx, &my);
        map_pt(H, j.lens, mx, my, &bx, &by);
        const double nx = e ? bx - ax : ax - bx, ny = e ? by - ay : ay - by, nn = std::hypot(nx, ny);
        if (!(nn > 0)) continue;
        j.pr.push_back(FindProbe{static_cast<float>(x), static_cast<float>(y), static_cast<float>(nx / nn),
                                 static_cast<float>(ny / nn), reach, j.frame});
        j.lab2.push_back(Pt{0, 0, 0, sd, depth});
      }
    }
}

// the homography and lens again through the model's probes
void model_fit(FindJob &j, const FindEdge *ed) {
  j.pts.clear();
  const int n = static_cast<int>(j.lab2.size());
  for (int k = 0; k < n; k++)
    if (ed[k].w > 0) j.pts.push_back(Pt{ed[k].x, ed[k].y, ed[k].w, j.lab2[k].side, j.lab2[k].depth});
  double H[9], k1 = j.lens.k1;
  int used = 0;
  const double rms = j.fit.solve(j.pts, j.lens, j.L, 0.3, 5, H, &k1, &used);
  if (!(rms < 0.3)) {
    j.dead = true;
    return;
  }
  FindResult &r = *j.r;
  std::memcpy(r.H, H, sizeof H);
  j.lens.k1 = k1;
  r.k1 = k1;
  r.rms = static_cast<float>(rms);
  r.points = used;
}

void stage_finish(FindJob &j) {
  FindResult &r = *j.r;
  const int L = j.L;
  r.span = L - 30;
  r.found = 1;
  r.track = static_cast<float>(track_score(Img{j.img, j.w, j.h, j.stride}, r.H, j.lens, L, &r.contrast));
  const double cm[4][2] = {{1, 1}, {L - 1.0, 1}, {L - 1.0, L - 1.0}, {1, L - 1.0}};
  for (int c = 0; c < 4; c++) {
    double x, y;
    map_pt(r.H, j.lens, cm[c][0], cm[c][1], &x, &y);
    r.quad[2 * c] = static_cast<float>(x);
    r.quad[2 * c + 1] = static_cast<float>(y);
  }
}

// The Hough stage's buffers for F frames and the trig table.
int hough_setup(FinderBufs &b, int F, int w, int h, hipStream_t s, size_t *cells_out) {
  const int R = static_cast<int>(std::ceil(0.5 * std::hypot(w, h))) + 2, nr = 2 * R + 1;
  const size_t cells = static_cast<size_t>(FIND_THETA) * nr;
  b.acc.reserve(static_cast<size_t>(F) * cells * sizeof(float));
  b.max.reserve(static_cast<size_t>(F) * sizeof(unsigned) + 256);
  b.count.reserve(static_cast<size_t>(F) * sizeof(int) + 256);
  b.peaks.reserve(static_cast<size_t>(F) * FIND_MAX_PEAKS * sizeof(FindPeak));
  if (b.trig_nr != 1) {
    std::vector<float> tr(2 * FIND_THETA);
    for (int t = 0; t < FIND_THETA; t++) {
      tr[2 * t] = static_cast<float>(std::cos(2 * PI * t / FIND_THETA));
      tr[2 * t + 1] = static_cast<float>(std::sin(2 * PI * t / FIND_THETA));
    }
    b.trig.reserve(tr.size() * sizeof(float));
    hip_check(hipStreamSynchronize(s), "finder stream");
    copy_wait(b.trig.p, tr.data(), tr.size() * sizeof(float), hipMemcpyHostToDevice, s, "trig");
    b.trig_nr = 1;
  }
  *cells_out = cells;
  return nr;
}

// A frame's probes on its lane: one launch, its edges back into `ed`.
void run_probes(FinderLane &ln, const uint8_t *const *d_tab, int w, int h, const FindJob &j, std::vector<FindEdge> &ed) {
  const size_t n = j.pr.size();
  ed.resize(n);
  if (!n) return;
  ln.h_probes.reserve(n * sizeof(FindProbe));
  ln.h_edges.reserve(n * sizeof(FindEdge));
  ln.probes.reserve(n * sizeof(FindProbe));
  ln.edges.reserve(n * sizeof(FindEdge));
  std::copy(j.pr.begin(), j.pr.end(), ln.h_probes.as<FindProbe>());
  hip_check(hipMemcpyAsync(ln.probes.p, ln.h_probes.p, n * sizeof(FindProbe), hipMemcpyHostToDevice, ln.s), "probes");
  launch_find_probe(d_tab, w, h, w, ln.probes.as<FindProbe>(), static_cast<int>(n), ln.edges.as<FindEdge>(), ln.s);
  hip_check(hipGetLastError(), "probe launch");
  hip_check(hipMemcpyAsync(ln.h_edges.p, ln.edges.p, n * sizeof(FindEdge), hipMemcpyDeviceToHost, ln.s), "edges");
  hip_check(hipStreamSynchronize(ln.s), "probe sync");
  std::memcpy(ed.data(), ln.h_edges.p, n * sizeof(FindEdge));
}

// A frame from its peaks to its registration; the stage times (host clock) into its result.
void run_frame(FindJob &j, FinderLane &ln, const uint8_t *const *d_tab) {
  auto t0 = std::chrono::steady_clock::now();
  auto lap = [&](int k) {
    const auto t1 = std::chrono::steady_clock::now();
    j.r->ms[k] = std::chrono::duration<float, std::milli>(t1 - t0).count();
    t0 = t1;
  };
  std::vector<FindEdge> ed;
  stage_sides(j);
  lap(2);
  for (int pass = 0; pass < 2 && !j.dead; pass++) {
    edge_probes(j);
    run_probes(ln, d_tab, j.w, j.h, j, ed);
    edge_fits(j, ed.data(), pass);
  }
  lap(3);
  if (!j.dead) stage_rings(j);
  lap(4);
  for (int pass = 0; pass < 2 && !j.dead; pass++) {
    model_probes(j, pass);
    run_probes(ln, d_tab, j.w, j.h, j, ed);
    model_fit(j, ed.data());
  }
  lap(5);
  if (!j.dead) stage_finish(j);
}

}  // namespace

namespace {

// The most probes one frame's stage makes: both edges of four sides (the model's probes are fewer).
constexpr size_t MAX_FRAME_PROBES = static_cast<size_t>(EDGE_PROBES) * 2 * 4;

// A lane's buffers at their largest at once, so that no later call reallocates (and so synchronises) them.
void lane_reserve(FinderLane &ln) {
  ln.probes.reserve(MAX_FRAME_PROBES * sizeof(FindProbe));
  ln.edges.reserve(MAX_FRAME_PROBES * sizeof(FindEdge));
  ln.h_probes.reserve(MAX_FRAME_PROBES * sizeof(FindProbe));
  ln.h_edges.reserve(MAX_FRAME_PROBES * sizeof(FindEdge));
  ln.h_peaks.reserve(FIND_MAX_PEAKS * sizeof(FindPeak) + sizeof(int));
  ln.anch.reserve(ANCH_AREA);
  ln.h_anch.reserve(ANCH_AREA);
}

// A frame's vote and peaks on its lane, then the rest of its search.
void search_frame(FinderBufs &b, FinderLane &ln, FindJob &j, const uint8_t *d_img, const uint8_t *const *d_tab, int nr,
                  size_t cells) {
  const auto t0 = std::chrono::steady_clock::now();
  const int f = j.frame;
  float *acc = b.acc.as<float>() + f * cells;
  launch_find_zero(acc, cells, b.max.as<unsigned>() + f, b.count.as<int>() + f, ln.s);
  launch_find_vote(d_img, j.w, j.h, j.w, 12.0f, b.trig.as<float2>(), nr, acc, ln.s);
  launch_find_peaks(acc, nr, 0.08f, b.max.as<unsigned>() + f, b.peaks.as<FindPeak>() + f * FIND_MAX_PEAKS,
                    b.count.as<int>() + f, ln.s);
  hip_check(hipGetLastError(), "finder launches");
  ln.h_peaks.reserve(FIND_MAX_PEAKS * sizeof(FindPeak) + sizeof(int));
  FindPeak *hp = ln.h_peaks.as<FindPeak>();
  int *hn = reinterpret_cast<int *>(hp + FIND_MAX_PEAKS);
  hip_check(hipMemcpyAsync(hn, b.count.as<int>() + f, sizeof(int), hipMemcpyDeviceToHost, ln.s), "peak count");
  hip_check(hipMemcpyAsync(hp, b.peaks.as<FindPeak>() + f * FIND_MAX_PEAKS, FIND_MAX_PEAKS * sizeof(FindPeak),
                           hipMemcpyDeviceToHost, ln.s), "peaks");
  hip_check(hipStreamSynchronize(ln.s), "finder peaks");
  j.pk.assign(hp, hp + std::min(*hn, FIND_MAX_PEAKS));
  j.r->ms[1] = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  run_frame(j, ln, d_tab);
}

// The pool's threads and their lanes, made once.
void ensure_pool(FinderBufs &b) {
  if (b.pool) return;
  const int nt = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  b.pool = std::make_shared<FinderPool>(nt);
  for (int t = 0; t < nt; t++) {
    auto ln = std::make_unique<FinderLane>();
    hip_check(hipStreamCreateWithFlags(&ln->s, hipStreamNonBlocking), "finder lane stream");
    ln->own = true;
    lane_reserve(*ln);
    b.lanes.push_back(std::move(ln));
  }
}

// The search of frames whose device images are in place (upload false: one frame, at d_imgs) or are uploaded here,
// each into its lane's slot of d_imgs, on their lanes. The Hough stage's buffers are a lane's too: the frames a lane
// searches one after another reuse them.
void search(FinderBufs &b, int F, uint8_t *d_imgs, const uint8_t *const *imgs, int w, int h, int stride,
            hipStream_t s, FindResult *res, bool upload, const std::function<void(int, FinderLane &)> &done) {
  const float up = res[0].ms[0];
  for (int f = 0; f < F; f++) res[f] = FindResult{};
  res[0].ms[0] = up;
  const int slots = finder_slots(b, F);
  size_t cells = 0;
  const int nr = hough_setup(b, slots, w, h, s, &cells);
  const size_t frame_bytes = static_cast<size_t>(w) * h;
  std::vector<const uint8_t *> tab(slots);
  for (int t = 0; t < slots; t++) tab[t] = d_imgs + t * frame_bytes;
  b.imgtab.reserve(slots * sizeof(void *));
  hip_check(hipMemcpyAsync(b.imgtab.p, tab.data(), slots * sizeof(void *), hipMemcpyHostToDevice, s), "image table");
  hip_check(hipStreamSynchronize(s), "image table");
  const uint8_t *const *d_tab = b.imgtab.as<const uint8_t *>();
  std::vector<FindJob> jobs(F);
  for (int f = 0; f < F; f++) {
    FindJob &j = jobs[f];
    j.img = imgs[f];
    j.w = w;
    j.h = h;
    j.stride = stride;
    j.cx = 0.5 * w;
    j.cy = 0.5 * h;
    j.r = &res[f];
  }
  auto one = [&](int f, FinderLane &ln, int slot) {
    uint8_t *dst = d_imgs + slot * frame_bytes;
    jobs[f].frame = slot;  // the lane's Hough buffers, image and probes' image index
    ln.d_img = dst;
    if (upload) {
      const auto t0 = std::chrono::steady_clock::now();
      if (stride == w) {
        hip_check(hipMemcpyAsync(dst, imgs[f], frame_bytes, hipMemcpyHostToDevice, ln.s), "capture");
      } else {
        hip_check(hipMemcpy2DAsync(dst, w, imgs[f], stride, w, h, hipMemcpyHostToDevice, ln.s), "capture");
      }
      hip_check(hipStreamSynchronize(ln.s), "capture");
      res[f].ms[0] = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    search_frame(b, ln, jobs[f], dst, d_tab, nr, cells);
    if (done) done(f, ln);
  };
  if (F == 1) {
    b.main.s = s;
    lane_reserve(b.main);
    one(0, b.main, 0);
    return;
  }
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::mutex em;
  std::exception_ptr err;  // the first frame's failure, raised here on the caller's thread
  b.pool->run(F, [&](int f, int t) {
    try {
      hip_check(hipSetDevice(dev), "hipSetDevice");
      one(f, *b.lanes[t], t);
    } catch (...) {
      std::lock_guard<std::mutex> lk(em);
      if (!err) err = std::current_exception();
    }
  });
  if (err) std::rethrow_exception(err);
}

}  // namespace

int finder_slots(FinderBufs &b, int F) {
  if (F <= 1) return 1;
  ensure_pool(b);
  return static_cast<int>(b.lanes.size());
}

void find_symbols(FinderBufs &b, int F, uint8_t *d_imgs, const uint8_t *const *imgs, int w, int h, int stride,
                  hipStream_t s, FindResult *res, const std::function<void(int, FinderLane &)> &done) {
  search(b, F, d_imgs, imgs, w, h, stride, s, res, true, done);
}

void find_symbol(FinderBufs &b, const uint8_t *d_img, const uint8_t *img, int w, int h, int stride, hipStream_t s,
                 FindResult &r) {
  const uint8_t *imgs[1] = {img};
  search(b, 1, const_cast<uint8_t *>(d_img), imgs, w, h, stride, s, &r, false, {});
}

}  // namespace lz
