// The 7 x 7 anchors' search on the GPU (liblizard/src/acquire.c:1876-2036 anchor_find, its scalar form): one workgroup
// an anchor, a window of the capture resampled about the anchor's place, the pattern's normalised correlation at every
// whole step within three modules from the window's integral images, the best step and a parabola each way.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lens.hpp"
#include "spectrum.hpp"

namespace lz {

constexpr int ANCH_N = ANCH_K * ANCH_K;              // anchors in all (ANCH_K a side, internal.h:186 OB_ANCH_K)
constexpr int ANCH_TS = 16;                          // the pattern's grid a side (OB_ANCH_TS)
constexpr int ANCH_REACH = 3;                        // modules searched either way (acquire.c:1716)
constexpr int ANCH_WMAX = 64;                        // the largest window a side (anchors of 2 modules or more)

// What one search reads: the capture (device, rows w bytes apart), H (module -> pixel, h8 = 1) and optionally a lens
// model about it (lens_on), the ring's anchor squares, the pattern's constants, and the anchors already found (skip
// bit k: anchor k keeps its record).
struct AnchorJob {
  const uint8_t *img;
  int w, h, span, lens_on;
  float H[9];
  float m0[ANCH_K];           // anchor square i's first module on either axis
  float tm, tn;               // the pattern's mean and norm after its mean is taken out
  int i25, j25, i6, j6;       // the pattern's inner squares along an axis (the grid cells under 0.25 and 0.6)
  unsigned long long skip;
  int K, pad;
  double t[LENS_P], ax, ay, R;  // the lens model: K radial terms, coefficients, axis, R
};

// An anchor's record (acquire.c:1697 anchor_t): its centre (X, Y) in modules, where it was found less where H puts
// it (dx, dy) pixels, the correlation's peak, where it was found (u, v) pixels, found or not.
struct AnchorHit {
  float X, Y, dx, dy, peak, u, v;
  int ok;
};

// A search's device (and pinned host) area: the job, then the 49 records.
constexpr size_t ANCH_JOB_BYTES = (sizeof(AnchorJob) + 255) / 256 * 256;
constexpr size_t ANCH_AREA = ANCH_JOB_BYTES + ANCH_N * sizeof(AnchorHit);

// The pattern's constants for a ring (float, the reference's order).
void anchor_pattern(int span, AnchorJob &j);
// frames jobs (device) -> frames x 49 records (device); anchors a job skips are left as they are.
void launch_anchor_find(const AnchorJob *d_jobs, int frames, AnchorHit *d_hits, hipStream_t s);

}  // namespace lz
