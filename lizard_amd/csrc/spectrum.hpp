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
constexpr int FFT_MAX_STAGES = 12;

// Device-side view of a plan (by value into kernels).
struct PlanDev {
  int n, subch, E, blocks, npilots, nstages;
  int radix[FFT_MAX_STAGES];
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
};

struct FramePlan {
  int version = 0, span = 0, n = 0, subch = 0, E = 0;
  const Format *fmt = nullptr;
  PlanDev dev{};
  DevBuf tables, coef, pilot, tw, rot, white;
  std::vector<int> radix;
  std::vector<int> pilot_host;  // coefficient indices of the pilots, in block order
  std::vector<int> pilot_block;
};

// Builds (and uploads) the plan for LIZARD-16V in the ring of the given span (128, 256 or 512 modules).
std::unique_ptr<FramePlan> make_plan(int version, int span);

// Per-frame device work areas for one decode.
struct FrameBufs {
  DevBuf grid, X, HP, rowsum, beta, proj, align, llr, est, rms, coef_dbg, pilot_r;
  void reserve(const FramePlan &p, int frames, bool debug);
};

// Enqueues grid (row-major n x n float per frame, frames back to back) -> soft values, estimates, pilot shift.
void launch_soft(const FramePlan &p, FrameBufs &b, const float *d_grid, int frames, bool debug, hipStream_t s);

}  // namespace lz
