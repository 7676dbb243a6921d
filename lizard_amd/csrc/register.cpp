// The registration model on the host: the lens field at the lattice's nodes and the anchors' mesh weights, both in
// double, computed once per (ring, frame size, reference homography).
#include "register.hpp"

#include <cmath>

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
