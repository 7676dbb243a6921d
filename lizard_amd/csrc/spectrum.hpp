// The finish path from picture samples to soft values on the GPU: anchor erase, detrend, 2D FFT, pilot alignment and
// the per-sub-channel soft demap (liblizard/src/lizard_phy.c:1566-1772). A FramePlan holds one format's tables (one
// version at one ring), built on the host once and kept on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <vector>

#include "common.hpp"
#include "format.hpp"

namespace lz {

constexpr int ANCH_K = 7;           // anchors per axis (SPEC 6.13)
constexpr int ANCH_MARGIN = 2;      // samples erased round a square
constexpr int BORDER = 15;          // modules between the symbol's edge and the picture (mark + guard)
constexpr int ALIGN_TURNS = 4096;   // steps of the pilots' turn table
constexpr int MAX_NODES = 40;       // node lattice points a side (layout.h OB_MAX_NODES)
constexpr int FIT_ROWS = 4;         // rows of the shading fit's lattice (every eighth sample) a fit workgroup sums
constexpr int FIT_BETA = 6 * 1536 / 8 / FIT_ROWS;    // the fit's partial sums a frame (6 per part, n <= 1536), then
constexpr int FIT_STRIDE = FIT_BETA + 8;              // the frame's shading beta (6) at FIT_BETA

// One frame's registration and capture for the warp: H maps module coordinates to capture pixels (h8 = 1), nodes
// the node lattice's displacements in pixels (ny rows of nx), img a tight uint8 luma capture of w x h.
// A capture and its registration as the warp reads them. w, h: the whole frame (the read's clamp); pixel (x, y) at
// img + y stride + x, where only the pixels the warp can reach need be stored (img then points before the stored
// rectangle by its offset).
struct WarpFrame {
  float H[9];
  const float2 *nodes;
  const uint8_t *img;
  int w, h;
  int stride;
};
// The sampler's pixel offsets are 32-bit: rows under 2^24 bytes apart and stride h under 2^31 bytes.
inline bool warp_frame_fits(int h, int stride) {
  return stride > 0 && h > 0 && stride < (1 << 24) && static_cast<long long>(stride) * h < (1LL << 31);
}

// Device-side view of a plan (by value into kernels).
struct PlanDev {
  int n, subch, E, blocks, npilots;
  int elo[ANCH_K], ehi[ANCH_K];
  double hole_inv[36];
  double sx2, sq2;
  const float *cx, *qx;                    // n each: the detrend basis c, q
  const float *x1r, *x1i, *x2r, *x2i;      // n each: their DFTs
  const float2 *tw;                        // n twiddles exp(-2 pi i k / n)
  const float2 *rot;                       // ALIGN_TURNS (cos, sin)
  const int32_t *coef;                     // subch * 320: u (low 16 bits, signed) and v (high 16)
  const int32_t *pilot;                    // npilots: coefficient index | frame slot parity << 31 ... (see spectrum.hip)
  const uint8_t *white;                    // frame slot whitening
  // the warp (picture sample x or y -> module coordinate and node cell; the same table on both axes)
  int nx, ny;
  float node_x[MAX_NODES], node_y[MAX_NODES];
  float node_inv[MAX_NODES];               // 1 / (node_x[i + 1] - node_x[i]), rounded once (IEEE division)
  const float *mcoord;                     // n: module coordinate of sample index t
  const int32_t *mcell;                    // n: node cell of sample index t (the same on both axes)
  int cell0[MAX_NODES];                    // first sample index of node cell i (cell0[nx - 1] = n)
  const int4 *blk;                         // per block: first slot, code's sent bits nt, slots nb, code index
  const int32_t *col_off, *col_ent;        // per column u < E: its coefficients (ci | source index << 19 | conj << 30)
};

struct FramePlan {
  int version = 0, span = 0, n = 0, subch = 0, E = 0;
  const Format *fmt = nullptr;
  PlanDev dev{};
  DevBuf tables, coef, pilot, tw, rot, white, mtab, blk, cols;
  std::vector<int> pilot_host;  // coefficient indices of the pilots, in block order
  std::vector<int> pilot_block;
};

// Builds (and uploads) the plan for LIZARD-16V in the ring of the given span (128, 256 or 512 modules).
std::unique_ptr<FramePlan> make_plan(int version, int span);

// Raw device pointers to one run's work areas and outputs; frame f of a run sits at f times each per-frame size.
struct SoftBufs {
  double *beta = nullptr;  // the shading fit's partial sums: FIT_STRIDE doubles a frame
  double *rowsum = nullptr, *proj = nullptr;
  float2 *X = nullptr, *C = nullptr;  // row-pass output [u][y]; the coefficients in coefficient order
  float *align = nullptr;
  int8_t *llr = nullptr;
  double *est = nullptr;
  float *rms = nullptr;
  float *coef_dbg = nullptr;  // null: no debug copy of the coefficients (then the soft floats, 2S floats each)
  float *blk_pilot = nullptr; // null: no pilot readings
  const float *otf = nullptr;  // null: no equalisation; else a blur sigma (picture samples) a frame, 0 for none
  int *zero_int = nullptr;     // an int the run's first kernel sets to 0 (a later launch's queue counter), or null
  SoftBufs at(const FramePlan &p, int f) const;  // frame f's slice
};

// Per-frame device work areas for one decode.
struct FrameBufs {
  DevBuf grid, X, C, rowsum, beta, proj, align, llr, est, rms, coef_dbg, pilot_r, otf;
  bool has_otf = false;
  void reserve(const FramePlan &p, int frames, bool debug);
  SoftBufs view(bool debug);
  // The known blur of each frame (sigma in picture samples, 0 to 8; 0 none) that the demap equalises; all zero or
  // null: none (the reference's demap). Copied on stream s.
  void set_otf(const float *sigma, int frames, hipStream_t s);
};

// Waves a workgroup of the row FFT, the column FFT and the demap (1, 2 or 4): launch shapes that leave every
// result's bits unchanged, so an autotuner may pick them by timing. And the radix plan of each FFT's 64-point stage
// across the lanes (R1 x 64 / R1, R1 = 4, 8 or 16): a different order of float operations, every plan within float
// rounding of the others and of FP64.
struct LaunchShape {
  int rows = 4, cols = 4, demap = 4;
  int rows_radix = 8, cols_radix = 8;
};

// Enqueues grid (row-major n x n float per frame, frames back to back) -> soft values, estimates, pilot shift.
void launch_soft(const FramePlan &p, const SoftBufs &b, const float *d_grid, int frames, hipStream_t s);
// The same with each frame's grid at its own address (d_grids: frames device pointers, on the device).
void launch_soft_grids(const FramePlan &p, const SoftBufs &b, const float *const *d_grids, int frames, hipStream_t s,
                       const LaunchShape &sh = LaunchShape{});
// The same from captures through their registrations (warp fused into the erase fit and the row transform).
void launch_soft_warp(const FramePlan &p, const SoftBufs &b, const WarpFrame *d_frames, int frames, hipStream_t s,
                      const LaunchShape &sh = LaunchShape{});
inline void launch_soft(const FramePlan &p, FrameBufs &b, const float *d_grid, int frames, bool debug, hipStream_t s) {
  launch_soft(p, b.view(debug), d_grid, frames, s);
}
inline void launch_soft_warp(const FramePlan &p, FrameBufs &b, const WarpFrame *d_frames, int frames, bool debug,
                             hipStream_t s) {
  launch_soft_warp(p, b.view(debug), d_frames, frames, s);
}
// The warp alone: n x n picture samples a frame, row-major.
void launch_warp(const FramePlan &p, const WarpFrame *d_frames, int frames, float *d_grid, hipStream_t s);
// The sender's pictures: frame slots (0/1, subch * 640 a frame) -> n x n float pictures in [0, 1], row-major;
// d_G a work area of E * n complex a frame.
void launch_modulate(const FramePlan &p, const uint8_t *d_bits, float2 *d_G, float *d_pic, int frames, hipStream_t s);
// The refinement's spectral loss and its adjoint through the FFT, one frame: the residual e = w - alpha r - beta
// (n x n, device) transformed (d_X: E n complex; d_C: the format's coefficients in coefficient order), per 256
// coefficients the sum of mask |C|^2 (d_part, FP64; the loss is 2 / n^2 times their sum), and the loss's derivative
// with respect to w (d_dl, n x n) through the inverse transform of the masked coefficients (d_G: E n complex).
// Returns the number of parts.
int launch_spectral_adjoint(const FramePlan &p, const float *d_w, const float *d_ref, float alpha, float beta,
                            const uint8_t *d_mask, float2 *d_X, float2 *d_C, double *d_part, float2 *d_G, float *d_dl,
                            hipStream_t s);
// The node lattice of a ring's layout (module coordinates; layout.c:110-118).
int node_lattice(int span, float *pos);

}  // namespace lz
