// The registration model beside the held {H, nodes} form: a radial lens term and the 7 x 7 anchors' mesh give the
// node lattice's displacements, and its refinement on the GPU (the hand-written adjoint of the warp and the loss, and
// Gauss-Newton over the homography).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "anchor.hpp"
#include "lens.hpp"
#include "spectrum.hpp"

namespace lz {

constexpr int REG_ANCH = 7;                                 // anchors a side (SPEC 6.13)
constexpr int REG_PARAMS = 8 + 1 + 2 * REG_ANCH * REG_ANCH; // homography, k1, anchors' (dx, dy)

// Anchor square i's first module on either axis in ring span (squares span / 64 modules a side).
int anchor_m0(int span, int i);
constexpr int REG_GN = 9;                                   // Gauss-Newton parameters: h0..h7 and k1
constexpr int REG_NODE0 = 1 + REG_GN + REG_GN * (REG_GN + 1) / 2 + 5;  // a piece's node gradient (4 corners x 2)
constexpr int REG_MET0 = REG_NODE0 + 8;                     // a piece's node metric (4 corners x gx^2, gx gy, gy^2)
constexpr int REG_PART = REG_MET0 + 12;                     // floats of partial sums a piece of a cell
constexpr int REG_SPLIT = 1;                                // pieces a node cell's rows are split into

// The node lattice's displacements as a linear function of k1 and the anchors for one ring, frame size and reference
// homography H0: nodes = k1 lens + sum over anchors of (dx, dy) phi.
struct RegModel {
  int nx = 0;
  std::vector<double> node;  // nx lattice positions (modules)
  std::vector<double> lens;  // nx * nx * 2: |q|^2 q D at node (i, j)'s place under H0 (liblizard/src/acquire.c:1506-1512)
  std::vector<double> phi;   // nx * nx * 49: anchor a's weight at node (i, j) (acquire.c:2196-2210, :2319-2352)
};
void reg_model(int span, const float *H0, int w, int h, RegModel &m);
void reg_nodes(const RegModel &m, const float *params, std::vector<float> &nodes);
// grad_k1 and grad_anchors (98) from the gradient with respect to the nodes (nx * nx * 2).
void reg_chain(const RegModel &m, const double *node_grad, double *grad);

// The lens model of the border (liblizard/src/acquire.c:1470-1528): a node's displacement from H as the first-order
// change of the homography (x, y on both axes, the shift, the two perspective terms -(g . q) q) plus K (0..3) radial
// terms k |q|^(2k) q, q = (the node's place under H - the axis) / R, R the farthest frame corner's distance from the
// axis; the axis is the frame's centre unless given.
struct LensModel {
  int K = -1;            // radial terms; -1: no model (H alone)
  double t[LENS_P] = {};  // coefficients: 8 affine and perspective, then K radial
  double ax = 0, ay = 0, R = 1;
};
// The axis and R for a w x h frame (axis x, y; NaN: the centre).
void lens_axis(int w, int h, float axis_x, float axis_y, LensModel &m);
// The model's displacement at image point (u, v).
void lens_at(const LensModel &m, double u, double v, float *dx, float *dy);
// The interior of the node lattice from its border (acquire.c:1531-1605): the lens model fitted by least squares to
// the border nodes in use (the corners and the marked nodes use[] allows; all where use is null), K by AICc (at most 1
// where the axis lies inside the lattice's quad), then every interior node the model plus the Coons patch of what the
// model leaves on the border. nodes: nx * nx (dx, dy), the border read, the interior written. Returns K (-1: none fitted,
// H alone and the patch).
int border_fill(int span, const float *H, int w, int h, float axis_x, float axis_y, float *nodes, const uint8_t *use,
                LensModel *out);
// An anchor's reading: its centre (X, Y) in modules, its displacement from H (dx, dy) in pixels, found or not.
struct AnchorReading {
  float X, Y, dx, dy;
  int ok;
};
// The interior from the border and the anchors (acquire.c:2219-2358): the lens model refitted over the border nodes in
// use and the anchors found, robustly (border nodes further from it than the larger of tau_floor = mpx / 2 and three
// robust spreads are dropped and take the model's displacement; three rounds), the model at every node plus the Coons
// patch of what it leaves on the border, then a cubic Hermite mesh of what the anchors still see through that (each
// reading shrunk by the share of it that is signal; knots 0 at the lattice's ends, a missing anchor's knot the mean
// of its found neighbours). Returns the model's K.
int anchor_fill(int span, const float *H, int w, int h, float axis_x, float axis_y, float *nodes, const uint8_t *use,
                const AnchorReading *anchors, float mpx, LensModel *out);

// The homography through four point pairs (src modules -> dst pixels), h8 = 1 (acquire.c:643-654, its scalar form).
bool homography4(const float *src, const float *dst, float *H);

// The anchors' search through H and, when given, a lens model about it: every anchor of A not yet found searched
// (anchor.hpp); returns how many are found in all.
using AnchorFind = std::function<int(const float *H, const LensModel *lm, AnchorHit *A)>;
// The registration from the anchors outward (acquire.c:2128-2191): the anchors searched through H, then, round after
// round (at most 6, while more are found), those still missing through the homography the found ones say and the lens
// model they say about it, each round's readings held to the robust rule (acquire.c:2099-2117); none where fewer than
// 12 are found. Then each corner of quad (the lattice's corners in pixels, top left and clockwise) further than 3
// modules (mpx pixels each) from where the anchors' map puts it is moved to where the lens model over the anchors and
// the kept corners puts it, and H made again from the four corners (the reference searches such a corner again with
// its node template first and keeps a match that scores 0.35 or more); moved receives how many. Last, the anchors'
// displacements about H and the lens model of the corners and the anchors (acquire.c:2038-2057) in out. Returns 1
// with that model, 0 without.
int anchor_register(int span, float *H, int w, int h, float mpx, float *quad, const AnchorFind &find, AnchorHit *A,
                    LensModel *out, int *moved);

// Device state of a refinement: the homography, the intensity model w = alpha r + beta, and the change of k1 from
// the staged nodes' (dk1: the nodes are theirs plus dk1 times the unit lens field).
struct RegState {
  float H[9];
  float alpha, beta;
  float dk1;
  float pad[4];
};

// Per piece of a node cell (grid (nx - 1) x (nx - 1) REG_SPLIT, pieces (cj REG_SPLIT + piece) (nx - 1) + ci): sums of
// e^2, J^T e (9), J^T J (45, upper triangle) over h0..h7 and k1, the intensity statistics (sum w, sum r, sum w r,
// sum r r, count), e (gx, gy) spread to the cell's four corners (8) and each corner's share of the warp's metric
// (w^2 gx^2, w^2 gx gy, w^2 gy^2 with w the corner's bilinear weight; 12), e = w - alpha r - beta the residual of the
// warped capture w against the reference picture r. d_lens (nx * nx, device; null: k1's column 0 and dk1 unused) is the
// unit lens field at the nodes. With d_dl (n x n, device) e is instead d_dl / 2 at each sample: a loss's derivative
// with respect to the warped picture (launch_spectral_adjoint), so that the sums' 2 J^T e is that loss's gradient.
void launch_reg_accum(const FramePlan &p, const WarpFrame *d_frame, const RegState *d_state, const float *d_ref,
                      const float2 *d_lens, float *d_part, hipStream_t s, const float *d_dl = nullptr);
// Sums the cells' parts in FP64 and, when update, solves the damped normal equations over the first nparam (8: H; 9:
// H and k1) parameters in FP64 and moves H (and dk1); sets alpha and beta to the least-squares intensity fit of this
// pass; trace[step] = the pass's loss.
void launch_gn_step(const FramePlan &p, const float *d_part, double *d_sum, RegState *d_state, float damping, int update,
                    int nparam, float *d_trace, int step, hipStream_t s);

}  // namespace lz
