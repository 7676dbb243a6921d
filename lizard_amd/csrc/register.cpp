// The registration model on the host: the lens field at the lattice's nodes and the anchors' mesh weights, both in
// double, computed once per (ring, frame size, reference homography).
#include "register.hpp"

#include <functional>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <utility>

namespace lz {

// The anchors' squares (liblizard/src/acquire.c:41-45): side span / 64 modules, square i's first module.
int anchor_m0(int span, int i) {
  const int a = span / 64;
  return static_cast<int>(std::floor(BORDER + static_cast<double>(span) * (i + 0.5) / REG_ANCH - a / 2.0 + 0.5));
}

namespace {

// Cubic Hermite weights through knots t[0..n-1] at x: the value is sum w[a] v[b + a], a = 0..3; slopes from the
// neighbouring knots, one-sided at the ends (the interpolation of acquire.c:2196-2210).
void herm4(const double *t, int n, double x, int *b, double w[4]) {
  int i = 0;
  while (i < n - 2 && x >= t[i + 1]) i++;
  const double hc = t[i + 1] - t[i], s = (x - t[i]) / hc, s2 = s * s, s3 = s2 * s;
  const double h00 = 2 * s3 - 3 * s2 + 1, h01 = -2 * s3 + 3 * s2, A = (s3 - 2 * s2 + s) * hc, Bc = (s3 - s2) * hc;
  w[0] = w[3] = 0;
  w[1] = h00;
  w[2] = h01;
  if (i > 0) {
    const double d = t[i + 1] - t[i - 1];
    w[0] -= A / d;
    w[2] += A / d;
  } else {
    w[1] -= A / hc;
    w[2] += A / hc;
  }
  if (i + 2 < n) {
    const double d = t[i + 2] - t[i];
    w[1] -= Bc / d;
    w[3] += Bc / d;
  } else {
    w[1] -= Bc / hc;
    w[2] += Bc / hc;
  }
  *b = i - 1;
}

}  // namespace

// ---- the border's lens model and the Coons fill (acquire.c:1470-1605), in the reference's order of float and double
// operations so the nodes are its to the bit

namespace {

// Gauss-Jordan with partial pivots on the first n rows and columns of [A | b] (b in column LENS_P).
bool solve_lens(double A[LENS_P][LENS_P + 1], int n) {
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
    if (std::fabs(A[p][c]) < 1e-14) return false;
    for (int k = 0; k <= LENS_P; k++) std::swap(A[c][k], A[p][k]);
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const double f = A[r][c] / A[c][c];
      for (int k = c; k <= LENS_P; k++) A[r][k] -= f * A[c][k];
    }
  }
  return true;
}

// The least-squares model with K radial terms over the nodes in use; coefficients in t.
bool lens_fit(int K, int nn, const std::vector<int> &use, const std::vector<float> &qx, const std::vector<float> &qy,
              const float *nodes, double *t) {
  const int p = 8 + K;
  double A[LENS_P][LENS_P + 1] = {}, rx[LENS_P], ry[LENS_P];
  int n = 0;
  for (int m = 0; m < nn; m++) {
    if (!use[m]) continue;
    n++;
    lens_rows(qx[m], qy[m], K, rx, ry);
    for (int a = 0; a < p; a++) {
      for (int b = 0; b < p; b++) A[a][b] += rx[a] * rx[b] + ry[a] * ry[b];
      A[a][LENS_P] += rx[a] * nodes[2 * m] + ry[a] * nodes[2 * m + 1];
    }
  }
  for (int a = 0; a < p; a++) A[a][a] += 1e-12;
  if (2 * n <= p || !solve_lens(A, p)) return false;
  for (int a = 0; a < p; a++) t[a] = A[a][LENS_P] / A[a][a];
  return true;
}

}  // namespace

void lens_axis(int w, int h, float axis_x, float axis_y, LensModel &m) {
  float ax = axis_x, ay = axis_y;
  if (!(ax == ax) || !(ay == ay)) {
    ax = w * 0.5f;
    ay = h * 0.5f;
  }
  m.ax = ax;
  m.ay = ay;
  m.R = 0;
  for (int c = 0; c < 4; c++) {
    const double d = std::hypot((c & 1 ? w : 0) - static_cast<double>(ax), (c & 2 ? h : 0) - static_cast<double>(ay));
    if (d > m.R) m.R = d;
  }
  if (!(m.R > 0)) m.R = 1;
}

void lens_at(const LensModel &m, double u, double v, float *dx, float *dy) {
  lens_disp(m.K, m.t, m.ax, m.ay, m.R, u, v, dx, dy);
}

int border_fill(int span, const float *H, int w, int h, float axis_x, float axis_y, float *nodes, const uint8_t *use_in,
                LensModel *out) {
  float pos[MAX_NODES];
  const int nx = node_lattice(span, pos), ny = nx, nn = nx * ny;
  LensModel lm;
  lens_axis(w, h, axis_x, axis_y, lm);
  const float ax = static_cast<float>(lm.ax), ay = static_cast<float>(lm.ay);
  const double R = lm.R;
  std::vector<float> qx(nn), qy(nn);
  std::vector<int> use(nn);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      float u, v;
      project_f(H, pos[i], pos[j], &u, &v);
      const int k = j * nx + i;
      qx[k] = static_cast<float>((u - ax) / R);
      qy[k] = static_cast<float>((v - ay) / R);
      const bool ex = i == 0 || i == nx - 1, ey = j == 0 || j == ny - 1;
      const bool corner = ex && ey, mark = (ex || ey) && !corner;  // the marked nodes: the border's, not its corners
      use[k] = corner || (mark && (!use_in || use_in[k]));
    }
  // the axis inside the lattice's quad (on one side of all four edges, corners in ring order): one radial term at most
  const int cn[4] = {0, nx - 1, nn - 1, nn - nx};
  int npos = 0, nneg = 0;
  for (int c = 0; c < 4; c++) {
    const int a = cn[c], b = cn[(c + 1) & 3];
    const double cr = static_cast<double>(qx[a]) * qy[b] - static_cast<double>(qy[a]) * qx[b];
    npos += cr > 0;
    nneg += cr < 0;
  }
  const int kmax = (npos == 4 || nneg == 4) ? 1 : LENS_TERMS;
  // K by AICc over the used nodes' 2n residuals
  double rx[LENS_P], ry[LENS_P];
  int K = -1;
  double best = INFINITY;
  for (int k = 0; k <= kmax; k++) {
    const int p = 8 + k;
    double t[LENS_P] = {}, rss = 0;
    int n = 0;
    if (!lens_fit(k, nn, use, qx, qy, nodes, t)) continue;
    for (int m = 0; m < nn; m++) {
      if (!use[m]) continue;
      n += 2;
      lens_rows(qx[m], qy[m], k, rx, ry);
      double fx = 0, fy = 0;
      for (int a = 0; a < p; a++) {
        fx += t[a] * rx[a];
        fy += t[a] * ry[a];
      }
      rss += (nodes[2 * m] - fx) * (nodes[2 * m] - fx) + (nodes[2 * m + 1] - fy) * (nodes[2 * m + 1] - fy);
    }
    if (n - p - 1 <= 0) continue;
    const double aicc = n * std::log(rss / n + 1e-30) + 2.0 * p + 2.0 * p * (p + 1) / (n - p - 1);
    if (aicc < best) {
      best = aicc;
      K = k;
    }
  }
  double th[LENS_P] = {};
  if (K >= 0 && !lens_fit(K, nn, use, qx, qy, nodes, th)) K = -1;
  std::vector<float> mx(nn), my(nn);
  for (int k = 0; k < nn; k++) {
    double X = 0, Y = 0;
    if (K >= 0) {
      lens_rows(qx[k], qy[k], K, rx, ry);
      for (int a = 0; a < 8 + K; a++) {
        X += th[a] * rx[a];
        Y += th[a] * ry[a];
      }
    }
    mx[k] = static_cast<float>(X);
    my[k] = static_cast<float>(Y);
  }
  // the Coons patch of the border's residuals from the model, carried to every interior node, in float
  const auto ex = [&](int i, int j) { return nodes[2 * (j * nx + i)] - mx[j * nx + i]; };
  const auto ey = [&](int i, int j) { return nodes[2 * (j * nx + i) + 1] - my[j * nx + i]; };
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const float u = (pos[i] - pos[0]) / (pos[nx - 1] - pos[0]), v = (pos[j] - pos[0]) / (pos[ny - 1] - pos[0]);
      const float fx = (1 - u) * ex(0, j) + u * ex(nx - 1, j) + (1 - v) * ex(i, 0) + v * ex(i, ny - 1) -
                       ((1 - u) * (1 - v) * ex(0, 0) + u * (1 - v) * ex(nx - 1, 0) + (1 - u) * v * ex(0, ny - 1) +
                        u * v * ex(nx - 1, ny - 1));
      const float fy = (1 - u) * ey(0, j) + u * ey(nx - 1, j) + (1 - v) * ey(i, 0) + v * ey(i, ny - 1) -
                       ((1 - u) * (1 - v) * ey(0, 0) + u * (1 - v) * ey(nx - 1, 0) + (1 - u) * v * ey(0, ny - 1) +
                        u * v * ey(nx - 1, ny - 1));
      nodes[2 * (j * nx + i)] = mx[j * nx + i] + fx;
      nodes[2 * (j * nx + i) + 1] = my[j * nx + i] + fy;
    }
  if (out) {
    lm.K = K;
    for (int a = 0; a < LENS_P; a++) lm.t[a] = K >= 0 && a < 8 + K ? th[a] : 0.0;
    *out = lm;
  }
  return K;
}

namespace {

// The mesh cell a module coordinate falls in along one axis: the last k of 0 .. nn - 2 with k = 0 or m >= node[k].
int node_cell(const float *node, int nn, float m) {
  int lo = 0, hi = nn - 2;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m >= node[mid]) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// A module point through H and the node mesh (acquire.c:5095-5103), float, the reference's order.
void map_point(const float *pos, int nx, const float *H, const float *nodes, float mx, float my, float *pu, float *pv) {
  const int j = node_cell(pos, nx, my), i = node_cell(pos, nx, mx);
  const float v = (my - pos[j]) / (pos[j + 1] - pos[j]), u = (mx - pos[i]) / (pos[i + 1] - pos[i]);
  const float *n00 = nodes + 2 * (j * nx + i), *n10 = n00 + 2, *n01 = n00 + 2 * nx, *n11 = n01 + 2;
  const float dx = (n00[0] * (1 - u) + n10[0] * u) * (1 - v) + (n01[0] * (1 - u) + n11[0] * u) * v;
  const float dy = (n00[1] * (1 - u) + n10[1] * u) * (1 - v) + (n01[1] * (1 - u) + n11[1] * u) * v;
  project_f(H, mx, my, pu, pv);
  *pu += dx;
  *pv += dy;
}

// The lens model over points, K = 0..3 by AICc; robust where isb is given (border points further than the larger of
// tau_floor and three robust spreads of the border's residuals dropped and the model refitted, three rounds).
bool lensm_fit(int np, const std::vector<double> &qx, const std::vector<double> &qy, const std::vector<double> &dx,
               const std::vector<double> &dy, const std::vector<uint8_t> *isb, std::vector<uint8_t> &drop,
               double tau_floor, LensModel &m) {
  std::vector<double> res(isb ? np : 0);
  bool ok = false;
  for (int round = 0; round < 4; round++) {
    double best = INFINITY, tb[LENS_P] = {};
    int K = -1;
    // the normal equations at every term at once (a model of fewer terms takes their leading block), the upper
    // triangle summed and copied down
    double A3[LENS_P][LENS_P + 1] = {}, rx[LENS_P], ry[LENS_P];
    int n = 0;
    for (int i = 0; i < np; i++) {
      if (drop[i]) continue;
      n++;
      lens_rows(qx[i], qy[i], LENS_TERMS, rx, ry);
      for (int a = 0; a < LENS_P; a++) {
        for (int b = a; b < LENS_P; b++) A3[a][b] += rx[a] * rx[b] + ry[a] * ry[b];
        A3[a][LENS_P] += rx[a] * dx[i] + ry[a] * dy[i];
      }
    }
    for (int a = 1; a < LENS_P; a++)
      for (int b = 0; b < a; b++) A3[a][b] = A3[b][a];
    for (int k = 0; k <= LENS_TERMS; k++) {
      const int p = 8 + k;
      double A[LENS_P][LENS_P + 1];
      for (int a = 0; a < LENS_P; a++)
        for (int b = 0; b <= LENS_P; b++) A[a][b] = A3[a][b];
      for (int a = 0; a < p; a++) A[a][a] += 1e-12;
      if (2 * n - p - 1 <= 0 || !solve_lens(A, p)) continue;
      double t[LENS_P] = {}, rss = 0;
      for (int a = 0; a < p; a++) t[a] = A[a][LENS_P] / A[a][a];
      for (int i = 0; i < np; i++) {
        if (drop[i]) continue;
        lens_rows(qx[i], qy[i], k, rx, ry);
        double fx = 0, fy = 0;
        for (int a = 0; a < p; a++) {
          fx += t[a] * rx[a];
          fy += t[a] * ry[a];
        }
        rss += (dx[i] - fx) * (dx[i] - fx) + (dy[i] - fy) * (dy[i] - fy);
      }
      const double nn = 2.0 * n, aicc = nn * std::log(rss / nn + 1e-30) + 2.0 * p + 2.0 * p * (p + 1) / (nn - p - 1);
      if (aicc < best) {
        best = aicc;
        K = k;
        for (int a = 0; a < LENS_P; a++) tb[a] = t[a];
      }
    }
    if (K < 0) break;
    m.K = K;
    for (int a = 0; a < LENS_P; a++) m.t[a] = tb[a];
    ok = true;
    if (!isb || round == 3) break;
    int nb = 0, changed = 0;
    for (int i = 0; i < np; i++) {
      if (!(*isb)[i]) continue;
      float fx, fy;
      lens_at(m, qx[i] * m.R + m.ax, qy[i] * m.R + m.ay, &fx, &fy);
      res[nb++] = std::hypot(dx[i] - fx, dy[i] - fy);
    }
    if (!nb) break;
    std::vector<double> srt(res.begin(), res.begin() + nb);
    std::sort(srt.begin(), srt.end());
    const double tau = std::fmax(tau_floor, 3 * 1.4826 * srt[nb / 2]);
    for (int i = 0, b = 0; i < np; i++) {
      if (!(*isb)[i]) continue;
      const uint8_t bad = res[b++] > tau;
      if (bad != drop[i]) {
        drop[i] = bad;
        changed++;
      }
    }
    if (!changed) break;
  }
  return ok;
}

}  // namespace

int anchor_fill(int span, const float *H, int w, int h, float axis_x, float axis_y, float *nodes, const uint8_t *use_in,
                const AnchorReading *A, float mpx, LensModel *out) {
  constexpr int NA = REG_ANCH * REG_ANCH;
  float pos[MAX_NODES];
  const int nx = node_lattice(span, pos), ny = nx, nn = nx * ny;
  std::vector<double> qx, qy, dx, dy;
  std::vector<uint8_t> isb, drop;
  std::vector<int> at;
  LensModel lm;
  lm.K = 0;  // the zero model (H alone) until a fit replaces it
  lens_axis(w, h, axis_x, axis_y, lm);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      if (i && j && i < nx - 1 && j < ny - 1) continue;
      const int k = j * nx + i;
      const bool corner = (i == 0 || i == nx - 1) && (j == 0 || j == ny - 1);
      if (!(corner || !use_in || use_in[k])) continue;
      float u, v;
      project_f(H, pos[i], pos[j], &u, &v);
      qx.push_back((u - lm.ax) / lm.R);
      qy.push_back((v - lm.ay) / lm.R);
      dx.push_back(nodes[2 * k]);
      dy.push_back(nodes[2 * k + 1]);
      isb.push_back(1);
      at.push_back(k);
    }
  for (int k = 0; k < NA; k++) {
    if (!A[k].ok) continue;
    float u, v;
    project_f(H, A[k].X, A[k].Y, &u, &v);
    qx.push_back((u - lm.ax) / lm.R);
    qy.push_back((v - lm.ay) / lm.R);
    dx.push_back(A[k].dx);
    dy.push_back(A[k].dy);
    isb.push_back(0);
    at.push_back(-1 - k);
  }
  const int np = static_cast<int>(qx.size());
  drop.assign(np, 0);
  lensm_fit(np, qx, qy, dx, dy, &isb, drop, 0.5 * mpx, lm);
  std::vector<float> mx(nn), my(nn);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      float u, v;
      project_f(H, pos[i], pos[j], &u, &v);
      lens_at(lm, u, v, &mx[j * nx + i], &my[j * nx + i]);
    }
  // a border node the anchors outvote reads the model's displacement, so the patch leaves nothing there
  for (int p = 0; p < np; p++)
    if (isb[p] && drop[p]) {
      nodes[2 * at[p]] = mx[at[p]];
      nodes[2 * at[p] + 1] = my[at[p]];
    }
  const auto ex = [&](int i, int j) { return nodes[2 * (j * nx + i)] - mx[j * nx + i]; };
  const auto ey = [&](int i, int j) { return nodes[2 * (j * nx + i) + 1] - my[j * nx + i]; };
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const float u = (pos[i] - pos[0]) / (pos[nx - 1] - pos[0]), v = (pos[j] - pos[0]) / (pos[ny - 1] - pos[0]);
      const float fx = (1 - u) * ex(0, j) + u * ex(nx - 1, j) + (1 - v) * ex(i, 0) + v * ex(i, ny - 1) -
                       ((1 - u) * (1 - v) * ex(0, 0) + u * (1 - v) * ex(nx - 1, 0) + (1 - u) * v * ex(0, ny - 1) +
                        u * v * ex(nx - 1, ny - 1));
      const float fy = (1 - u) * ey(0, j) + u * ey(nx - 1, j) + (1 - v) * ey(i, 0) + v * ey(i, ny - 1) -
                       ((1 - u) * (1 - v) * ey(0, 0) + u * (1 - v) * ey(nx - 1, 0) + (1 - u) * v * ey(0, ny - 1) +
                        u * v * ey(nx - 1, ny - 1));
      nodes[2 * (j * nx + i)] = mx[j * nx + i] + fx;
      nodes[2 * (j * nx + i) + 1] = my[j * nx + i] + fy;
    }
  // what the anchors still see through that mesh, as the sampler blends it
  const double X0 = pos[0], SX = pos[nx - 1] - X0;
  std::vector<std::array<double, 2>> tv;
  std::vector<int> gi;
  for (int p = 0; p < np; p++) {
    if (isb[p]) continue;
    const AnchorReading &an = A[-1 - at[p]];
    float u, v, hu, hv;
    map_point(pos, nx, H, nodes, an.X, an.Y, &u, &v);
    project_f(H, an.X, an.Y, &hu, &hv);
    tv.push_back({static_cast<double>(an.dx - (u - hu)), static_cast<double>(an.dy - (v - hv))});
    gi.push_back(-1 - at[p]);
  }
  const int m = static_cast<int>(tv.size());
  // each reading shrunk by the share of it that is signal: the noise from each reading against the mean of its found
  // grid neighbours (robust: 1.4826 MAD over both components), the readings' power per component, max(0, 1 - s2 / P)
  {
    double g[NA][2] = {}, dd[2 * NA], pw = 0;
    uint8_t has[NA] = {};
    int na2 = 0, nd = 0;
    for (int q = 0; q < m; q++)
      if (gi[q] >= 0) {
        has[gi[q]] = 1;
        g[gi[q]][0] = tv[q][0];
        g[gi[q]][1] = tv[q][1];
        pw += tv[q][0] * tv[q][0] + tv[q][1] * tv[q][1];
        na2++;
      }
    for (int k = 0; k < NA; k++) {
      if (!has[k]) continue;
      const int i = k % REG_ANCH, j = k / REG_ANCH;
      const int nb[4] = {i > 0 ? k - 1 : -1, i < REG_ANCH - 1 ? k + 1 : -1, j > 0 ? k - REG_ANCH : -1,
                         j < REG_ANCH - 1 ? k + REG_ANCH : -1};
      double sx = 0, sy = 0;
      int c = 0;
      for (int e = 0; e < 4; e++)
        if (nb[e] >= 0 && has[nb[e]]) {
          sx += g[nb[e]][0];
          sy += g[nb[e]][1];
          c++;
        }
      if (c < 2) continue;
      const double f = 1 / std::sqrt(1 + 1.0 / c);
      dd[nd++] = std::fabs(g[k][0] - sx / c) * f;
      dd[nd++] = std::fabs(g[k][1] - sy / c) * f;
    }
    if (nd >= 16 && na2 > 0) {
      std::sort(dd, dd + nd);
      const double sg = 1.4826 * dd[nd / 2], s2 = sg * sg, P = pw / (2.0 * na2),
                   kf = P > 0 ? std::fmax(0, 1 - s2 / P) : 0;
      for (int q = 0; q < m; q++)
        if (gi[q] >= 0) {
          tv[q][0] *= kf;
          tv[q][1] *= kf;
        }
    }
  }
  // the knots on both axes: the lattice's ends (0) and the anchors' centres; an inner knot its anchor's shrunk
  // reading, a missing one the mean of its found 4-neighbours (rounds until none is left, at most 6; none near: 0)
  constexpr int KN = REG_ANCH + 2;
  double kx[KN], kv[KN][KN][2] = {};
  uint8_t kn[KN][KN] = {};
  const double ha = 0.5 * (span / 64);
  kx[0] = 0;
  kx[KN - 1] = 1;
  for (int i = 0; i < REG_ANCH; i++) kx[i + 1] = (anchor_m0(span, i) + ha - X0) / SX;
  for (int q = 0; q < m; q++)
    if (gi[q] >= 0) {
      const int i = gi[q] % REG_ANCH + 1, j = gi[q] / REG_ANCH + 1;
      kv[j][i][0] = tv[q][0];
      kv[j][i][1] = tv[q][1];
      kn[j][i] = 1;
    }
  for (int round = 0; round < 6; round++) {
    uint8_t now[KN][KN];
    std::memcpy(now, kn, sizeof now);
    int left = 0;
    for (int j = 1; j < KN - 1; j++)
      for (int i = 1; i < KN - 1; i++) {
        if (kn[j][i]) continue;
        const int nb[4][2] = {{i - 1, j}, {i + 1, j}, {i, j - 1}, {i, j + 1}};
        double s0 = 0, s1 = 0;
        int c = 0;
        for (int e = 0; e < 4; e++) {
          const int a = nb[e][0], b = nb[e][1];
          if (a < 1 || a > KN - 2 || b < 1 || b > KN - 2 || !kn[b][a]) continue;
          s0 += kv[b][a][0];
          s1 += kv[b][a][1];
          c++;
        }
        if (c) {
          kv[j][i][0] = s0 / c;
          kv[j][i][1] = s1 / c;
          now[j][i] = 1;
        } else {
          left++;
        }
      }
    std::memcpy(kn, now, sizeof kn);
    if (!left) break;
  }
  // each axis a cubic Hermite through the knots: a node is a 4 x 4 blend of knot values
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      int bx, by;
      double wx[4], wy[4];
      herm4(kx, KN, (pos[i] - X0) / SX, &bx, wx);
      herm4(kx, KN, (pos[j] - X0) / SX, &by, wy);
      double d0 = 0, d1 = 0;
      for (int b = 0; b < 4; b++)
        for (int a = 0; a < 4; a++) {
          const int qxi = bx + a, qyi = by + b;
          if (qxi < 0 || qxi >= KN || qyi < 0 || qyi >= KN || (wx[a] == 0 && wy[b] == 0)) continue;
          d0 += wy[b] * wx[a] * kv[qyi][qxi][0];
          d1 += wy[b] * wx[a] * kv[qyi][qxi][1];
        }
      nodes[2 * (j * nx + i)] += static_cast<float>(d0);
      nodes[2 * (j * nx + i) + 1] += static_cast<float>(d1);
    }
  if (out) *out = lm;
  return lm.K;
}

// ---- the registration from the anchors outward (acquire.c:2038-2191), the same float and double operations

namespace {

constexpr int ANCH_ROUNDS = 6, ANCH_CORNER = 3, ANCH_MIN = 12;  // acquire.c:1716, internal.h:186

// Gauss-Jordan with partial pivots on [A | b], 8 unknowns (acquire.c:610-640, its scalar form).
bool solve8(double A[8][9]) {
  for (int c = 0; c < 8; c++) {
    int p = c;
    for (int r = c + 1; r < 8; r++)
      if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
    if (std::fabs(A[p][c]) < 1e-12) return false;
    for (int k = 0; k < 9; k++) std::swap(A[c][k], A[p][k]);
    for (int r = 0; r < 8; r++) {
      if (r == c) continue;
      const double f = A[r][c] / A[c][c];
      for (int k = c; k < 9; k++) A[r][k] -= f * A[c][k];
    }
  }
  return true;
}

}  // namespace

bool homography4(const float *src, const float *dst, float *H) {
  double A[8][9];
  for (int i = 0; i < 4; i++) {
    const double x = src[2 * i], y = src[2 * i + 1], u = dst[2 * i], v = dst[2 * i + 1];
    const double r0[9] = {x, y, 1, 0, 0, 0, -u * x, -u * y, u}, r1[9] = {0, 0, 0, x, y, 1, -v * x, -v * y, v};
    std::memcpy(A[2 * i], r0, sizeof r0);
    std::memcpy(A[2 * i + 1], r1, sizeof r1);
  }
  if (!solve8(A)) return false;
  for (int i = 0; i < 8; i++) H[i] = static_cast<float>(A[i][8] / A[i][i]);
  H[8] = 1;
  return true;
}

namespace {

// A homography through the anchors found, least squares with h8 = 1 on coordinates centred and scaled; false where
// fewer than eight are found or they span fewer than three rows or columns (acquire.c:2061-2096).
bool anchor_homo(const AnchorHit *A, float *H) {
  int n = 0, rows = 0, cols = 0, rm = 0, cm = 0;
  double mx = 0, my = 0, mu = 0, mv = 0;
  for (int k = 0; k < ANCH_N; k++)
    if (A[k].ok) {
      n++;
      mx += A[k].X;
      my += A[k].Y;
      mu += A[k].u;
      mv += A[k].v;
      rm |= 1 << (k / ANCH_K);
      cm |= 1 << (k % ANCH_K);
    }
  for (int b = 0; b < ANCH_K; b++) {
    rows += (rm >> b) & 1;
    cols += (cm >> b) & 1;
  }
  if (n < 8 || rows < 3 || cols < 3) return false;
  mx /= n;
  my /= n;
  mu /= n;
  mv /= n;
  double sx = 0, su = 0;
  for (int k = 0; k < ANCH_N; k++)
    if (A[k].ok) {
      sx += std::hypot(A[k].X - mx, A[k].Y - my);
      su += std::hypot(A[k].u - mu, A[k].v - mv);
    }
  sx = sx / n / 1.4142135623730951;
  su = su / n / 1.4142135623730951;
  if (!(sx > 0) || !(su > 0)) return false;
  double M[LENS_P][LENS_P + 1] = {};
  for (int k = 0; k < ANCH_N; k++) {
    if (!A[k].ok) continue;
    const double x = (A[k].X - mx) / sx, y = (A[k].Y - my) / sx, u = (A[k].u - mu) / su, v = (A[k].v - mv) / su;
    const double r0[8] = {x, y, 1, 0, 0, 0, -u * x, -u * y}, r1[8] = {0, 0, 0, x, y, 1, -v * x, -v * y};
    for (int a = 0; a < 8; a++) {
      for (int b = 0; b < 8; b++) M[a][b] += r0[a] * r0[b] + r1[a] * r1[b];
      M[a][LENS_P] += r0[a] * u + r1[a] * v;
    }
  }
  if (!solve_lens(M, 8)) return false;
  double g[9];
  for (int a = 0; a < 8; a++) g[a] = M[a][LENS_P] / M[a][a];
  g[8] = 1;
  // the scalings undone: H = Tu^-1 G Tx
  const double tx[9] = {1 / sx, 0, -mx / sx, 0, 1 / sx, -my / sx, 0, 0, 1}, ti[9] = {su, 0, mu, 0, su, mv, 0, 0, 1};
  double gt[9], hh[9];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      gt[3 * r + c] = 0;
      for (int k = 0; k < 3; k++) gt[3 * r + c] += g[3 * r + k] * tx[3 * k + c];
    }
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      hh[3 * r + c] = 0;
      for (int k = 0; k < 3; k++) hh[3 * r + c] += ti[3 * r + k] * gt[3 * k + c];
    }
  if (!(std::fabs(hh[8]) > 1e-12)) return false;
  for (int a = 0; a < 9; a++) H[a] = static_cast<float>(hh[a] / hh[8]);
  return true;
}

// The lens model the anchors alone say about H, held to the border's robust rule (a reading further from it than the
// larger of mpx / 2 and three robust spreads dropped, and no longer found); false where it does not fit
// (acquire.c:2099-2117).
bool anchor_model_free(const float *H, int w, int h, AnchorHit *A, LensModel &m, float mpx) {
  std::vector<double> qx, qy, dx, dy;
  std::vector<int> at;
  lens_axis(w, h, NAN, NAN, m);
  for (int k = 0; k < ANCH_N; k++) {
    if (!A[k].ok) continue;
    float hu, hv;
    project_f(H, A[k].X, A[k].Y, &hu, &hv);
    A[k].dx = A[k].u - hu;
    A[k].dy = A[k].v - hv;
    qx.push_back((hu - m.ax) / m.R);
    qy.push_back((hv - m.ay) / m.R);
    dx.push_back(A[k].dx);
    dy.push_back(A[k].dy);
    at.push_back(k);
  }
  const int np = static_cast<int>(qx.size());
  std::vector<uint8_t> isb(np, 1), drop(np, 0);
  if (np < 4 || !lensm_fit(np, qx, qy, dx, dy, &isb, drop, 0.5 * mpx, m)) return false;
  for (int i = 0; i < np; i++)
    if (drop[i]) A[at[i]].ok = 0;
  return true;
}

// The lattice's corner c (0 top left, then clockwise) in modules.
void corner_node(const float *pos, int nx, int c, float *x, float *y) {
  *x = pos[(c == 1 || c == 2) ? nx - 1 : 0];
  *y = pos[c >= 2 ? nx - 1 : 0];
}

// The lens model the anchors and the four corners (displacement 0: H passes through them) say about H
// (acquire.c:2038-2057).
bool anchor_model(const float *pos, int nx, const float *H, int w, int h, const AnchorHit *A, LensModel &m) {
  std::vector<double> qx, qy, dx, dy;
  lens_axis(w, h, NAN, NAN, m);
  for (int c = 0; c < 4; c++) {
    float x, y, u, v;
    corner_node(pos, nx, c, &x, &y);
    project_f(H, x, y, &u, &v);
    qx.push_back((u - m.ax) / m.R);
    qy.push_back((v - m.ay) / m.R);
    dx.push_back(0);
    dy.push_back(0);
  }
  for (int k = 0; k < ANCH_N; k++) {
    if (!A[k].ok) continue;
    float u, v;
    project_f(H, A[k].X, A[k].Y, &u, &v);
    qx.push_back((u - m.ax) / m.R);
    qy.push_back((v - m.ay) / m.R);
    dx.push_back(A[k].dx);
    dy.push_back(A[k].dy);
  }
  const int np = static_cast<int>(qx.size());
  std::vector<uint8_t> drop(np, 0);
  return lensm_fit(np, qx, qy, dx, dy, nullptr, drop, 0, m);
}

}  // namespace

int anchor_register(int span, float *H, int w, int h, float mpx, float *quad, const AnchorFind &find, AnchorHit *A,
                    LensModel *out, int *moved) {
  float pos[MAX_NODES];
  const int nx = node_lattice(span, pos);
  float Ha[9];
  std::memcpy(Ha, H, sizeof Ha);
  LensModel mf;
  bool have = false;
  int nf = 0;
  if (moved) *moved = 0;
  for (int k = 0; k < ANCH_N; k++) A[k].ok = 0;
  for (int round = 0; round < ANCH_ROUNDS; round++) {
    const int now = find(Ha, have ? &mf : nullptr, A);
    if (now == nf || now < 4) break;
    nf = now;
    float Hn[9];
    if (anchor_homo(A, Hn)) std::memcpy(Ha, Hn, sizeof Ha);
    if (!(have = anchor_model_free(Ha, w, h, A, mf, mpx))) break;
    nf = 0;
    for (int k = 0; k < ANCH_N; k++) nf += A[k].ok;
  }
  if (!have || nf < ANCH_MIN) return 0;
  // the anchors and the corners about the corners' H; a corner further than ANCH_CORNER modules from where the
  // anchors' own map (Ha and mf) puts it is left out of the fit
  std::vector<double> qx, qy, dx, dy;
  LensModel mc;
  lens_axis(w, h, NAN, NAN, mc);
  for (int c = 0; c < 4; c++) {
    float x, y, u, v;
    corner_node(pos, nx, c, &x, &y);
    project_f(H, x, y, &u, &v);
    qx.push_back((u - mc.ax) / mc.R);
    qy.push_back((v - mc.ay) / mc.R);
    dx.push_back(0);
    dy.push_back(0);
  }
  for (int k = 0; k < ANCH_N; k++) {
    if (!A[k].ok) continue;
    float hu, hv;
    project_f(H, A[k].X, A[k].Y, &hu, &hv);
    qx.push_back((hu - mc.ax) / mc.R);
    qy.push_back((hv - mc.ay) / mc.R);
    dx.push_back(A[k].u - hu);
    dy.push_back(A[k].v - hv);
  }
  const int np = static_cast<int>(qx.size());
  std::vector<uint8_t> drop(np, 0);
  for (int c = 0; c < 4; c++) {
    float x, y, pu, pv, mu, mv;
    corner_node(pos, nx, c, &x, &y);
    project_f(Ha, x, y, &pu, &pv);
    lens_at(mf, pu, pv, &mu, &mv);
    drop[c] = std::hypot(pu + mu - quad[2 * c], pv + mv - quad[2 * c + 1]) > ANCH_CORNER * mpx;
  }
  const bool fit = lensm_fit(np, qx, qy, dx, dy, nullptr, drop, 0, mc);
  // a corner the anchors disagree with goes where the model over the anchors and the kept corners puts it, and H is
  // made again from the four
  int nm = 0;
  for (int c = 0; fit && c < 4; c++) {
    if (!drop[c]) continue;
    float x, y, hu, hv, mu, mv;
    corner_node(pos, nx, c, &x, &y);
    project_f(H, x, y, &hu, &hv);
    lens_at(mc, hu, hv, &mu, &mv);
    quad[2 * c] = hu + mu;
    quad[2 * c + 1] = hv + mv;
    nm++;
  }
  if (moved) *moved = nm;
  if (nm) {
    float src[8], Hc[9];
    for (int c = 0; c < 4; c++) corner_node(pos, nx, c, &src[2 * c], &src[2 * c + 1]);
    if (homography4(src, quad, Hc)) std::memcpy(H, Hc, sizeof Hc);
  }
  for (int k = 0; k < ANCH_N; k++)
    if (A[k].ok) {
      float hu, hv;
      project_f(H, A[k].X, A[k].Y, &hu, &hv);
      A[k].dx = A[k].u - hu;
      A[k].dy = A[k].v - hv;
    }
  LensModel m;
  if (!anchor_model(pos, nx, H, w, h, A, m)) return 0;
  if (out) *out = m;
  return 1;
}

void reg_model(int span, const float *H0, int w, int h, RegModel &m) {
  float pos[MAX_NODES];
  const int nx = node_lattice(span, pos);
  m.nx = nx;
  m.node.assign(pos, pos + nx);
  // the lens field: k1 |q|^2 q D per unit k1 at each node's place under H0, about the frame's centre
  const double ax = w / 2.0, ay = h / 2.0, D = std::hypot(w / 2.0, h / 2.0);
  m.lens.assign(static_cast<size_t>(nx) * nx * 2, 0.0);
  for (int j = 0; j < nx; j++)
    for (int i = 0; i < nx; i++) {
      const double mx = m.node[i], my = m.node[j];
      const double W = H0[6] * mx + H0[7] * my + H0[8];
      const double px = (H0[0] * mx + H0[1] * my + H0[2]) / W, py = (H0[3] * mx + H0[4] * my + H0[5]) / W;
      const double qx = (px - ax) / D, qy = (py - ay) / D, r2 = qx * qx + qy * qy;
      m.lens[(j * nx + i) * 2] = r2 * qx * D;
      m.lens[(j * nx + i) * 2 + 1] = r2 * qy * D;
    }
  // the anchors' mesh: knots at the lattice's ends (0) and the anchors' centres; interior nodes only
  constexpr int KN = REG_ANCH + 2;
  const double X0 = m.node[0], SX = m.node[nx - 1] - X0, ha = 0.5 * (span / 64);
  double t[KN];
  t[0] = 0;
  t[KN - 1] = 1;
  for (int i = 0; i < REG_ANCH; i++) t[i + 1] = (anchor_m0(span, i) + ha - X0) / SX;
  m.phi.assign(static_cast<size_t>(nx) * nx * REG_ANCH * REG_ANCH, 0.0);
  std::vector<int> bs(nx);
  std::vector<double> ws(nx * 4);
  for (int i = 0; i < nx; i++) herm4(t, KN, (m.node[i] - X0) / SX, &bs[i], &ws[i * 4]);
  for (int j = 1; j < nx - 1; j++)
    for (int i = 1; i < nx - 1; i++)
      for (int b = 0; b < 4; b++)
        for (int a = 0; a < 4; a++) {
          const int qx = bs[i] + a, qy = bs[j] + b;
          if (qx < 1 || qx > REG_ANCH || qy < 1 || qy > REG_ANCH) continue;  // the ends' knots hold 0
          m.phi[(static_cast<size_t>(j) * nx + i) * REG_ANCH * REG_ANCH + (qy - 1) * REG_ANCH + (qx - 1)] +=
              ws[j * 4 + b] * ws[i * 4 + a];
        }
}

void reg_nodes(const RegModel &m, const float *params, std::vector<float> &nodes) {
  const int nx = m.nx, NA = REG_ANCH * REG_ANCH;
  const double k1 = params[8];
  const float *anch = params + 9;
  nodes.assign(static_cast<size_t>(nx) * nx * 2, 0.f);
  for (int k = 0; k < nx * nx; k++) {
    double dx = k1 * m.lens[2 * k], dy = k1 * m.lens[2 * k + 1];
    const double *ph = &m.phi[static_cast<size_t>(k) * NA];
    for (int a = 0; a < NA; a++) {
      dx += ph[a] * anch[2 * a];
      dy += ph[a] * anch[2 * a + 1];
    }
    nodes[2 * k] = static_cast<float>(dx);
    nodes[2 * k + 1] = static_cast<float>(dy);
  }
}

void reg_chain(const RegModel &m, const double *node_grad, double *grad) {
  const int nx = m.nx, NA = REG_ANCH * REG_ANCH;
  double gk = 0;
  for (int a = 0; a < 2 * NA; a++) grad[9 + a] = 0;
  for (int k = 0; k < nx * nx; k++) {
    const double gx = node_grad[2 * k], gy = node_grad[2 * k + 1];
    gk += gx * m.lens[2 * k] + gy * m.lens[2 * k + 1];
    const double *ph = &m.phi[static_cast<size_t>(k) * NA];
    for (int a = 0; a < NA; a++) {
      grad[9 + 2 * a] += ph[a] * gx;
      grad[9 + 2 * a + 1] += ph[a] * gy;
    }
  }
  grad[8] = gk;
}

}  // namespace lz
