// The finder: a symbol's registration (ring, homography, radial lens, orientation) from a capture alone, the job of
// liblizard/src/acquire.c's find_frame + refine_node + the lens fit (:4631, :1242, :1470-1512), in a design of this
// library's own. The GPU does the pixel work:
//   k_find_vote   per pixel a Sobel gradient, thinned to its ridge (non-maximum suppression along the gradient), and
//                 a vote into a Hough accumulator over (gradient direction, distance from the frame's centre): the
//                 direction keeps its sign, so a dark-to-light edge and a light-to-dark one never share a bin
//   k_find_max    the accumulator's maximum
//   k_find_peaks  its local maxima above a fraction of that, with a 3 x 5 neighbourhood sum as their score
//   k_find_probe  per probe (a point and a normal) the nearest dark-to-light edge along the normal, to a fraction of a
//                 pixel (the derivative's peak, refined by a parabola)
// The host (find.cpp) turns peaks into the border's four sides (the solid line's outer and inner edges, a light
// module outside it and the light gap inside), fits the homography and the lens to the probed edge points, and settles
// ring and orientation by the timing track.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <memory>
#include <vector>

#include "common.hpp"

namespace lz {

constexpr int FIND_THETA = 720;       // direction bins over a full turn
constexpr int FIND_MAX_PEAKS = 4096;  // local maxima kept
constexpr int FIND_MAX_PROBES = 1 << 16;

struct FindPeak {
  float theta, rho, votes, score;  // gradient direction (radians), rho = n . (p - centre) in pixels
};
struct FindProbe {
  float x, y, nx, ny;  // the point and the unit normal, pointing from dark to light
  float reach;         // how far along the normal to look, either way (pixels)
  int frame;           // which frame of a batch (its image in the launch's table)
};
struct FindEdge {
  float x, y, w, off;  // the edge's point, its gradient (grey levels a pixel; 0 none), its offset along the normal
};

void launch_find_vote(const uint8_t *img, int w, int h, int stride, float thr, const float2 *trig, int nr, float *acc,
                      hipStream_t s);
void launch_find_peaks(const float *acc, int nr, float frac, unsigned *d_max, FindPeak *peaks, int *count,
                       hipStream_t s);
void launch_find_probe(const uint8_t *const *imgs, int w, int h, int stride, const FindProbe *probes, int n,
                       FindEdge *out, hipStream_t s);
void launch_find_zero(float *acc, size_t count, unsigned *d_max, int *count_out, hipStream_t s);

// What the finder settles on: H maps module coordinates (0 .. span + 30) to undistorted capture pixels (h8 = 1), k1
// the radial lens about the frame's centre (q over the farthest corner's distance); quad the border line's outer
// corners (module (1, 1), (L - 1, 1), (L - 1, L - 1), (1, L - 1)) in the capture, lens included.
struct FindResult {
  int found = 0, span = 0;
  double H[9] = {};
  double k1 = 0;
  float quad[8] = {};
  float track = 0, contrast = 0, rms = 0;  // the track's agreement, the border's contrast, the fit's rms (modules)
  int points = 0;                          // edge points the final fit used
  float ms[6] = {};  // host clock: upload, vote + peaks, sides, first probes, rings, refinement
};

struct FinderPool;  // host threads for a batch's per-frame geometry

// Pinned host memory (hipHostMalloc) that grows as asked, for the copies between the stages.
struct HostBuf {
  void *p = nullptr;
  size_t bytes = 0;
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (p) hip_check(hipHostFree(p), "hipHostFree");
    p = nullptr;
    bytes = 0;
    hip_check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc");
    bytes = n;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// A lane's buffers for the border's nodes (border.hpp): the ring's layout and order (kept per ring), the job and the
// nodes, the capture under the border's blur.
struct BorderBufs {
  int span = 0, levels = 0, count = 0;
  DevBuf kinds, sched, job, nodes, rim;
  HostBuf h;
};

// One host thread's probe buffers and stream, so that a frame's stages run without waiting on the other frames.
struct FinderLane {
  hipStream_t s = nullptr;
  bool own = false;  // the stream was made for this lane
  const uint8_t *d_img = nullptr;     // the device image of the frame the lane searches now
  DevBuf probes, edges, anch;         // anch: the anchors' search job and records
  HostBuf h_probes, h_edges, h_peaks, h_anch;
  BorderBufs border;
  FinderLane() = default;
  FinderLane(const FinderLane &) = delete;
  FinderLane &operator=(const FinderLane &) = delete;
  ~FinderLane() {
    if (own && s) (void)hipStreamDestroy(s);
  }
};

struct FinderBufs {
  DevBuf acc, trig, max, peaks, count, img, imgtab;
  FinderLane main;  // the caller's stream, for a single frame
  int trig_nr = 0;
  std::shared_ptr<FinderPool> pool;
  std::vector<std::unique_ptr<FinderLane>> lanes;  // one a pool thread
};

// The whole search on a capture already on the device (d_img, rows w bytes apart) with its host copy (img, rows
// stride bytes apart) for the host's reads.
void find_symbol(FinderBufs &b, const uint8_t *d_img, const uint8_t *img, int w, int h, int stride, hipStream_t s,
                 FindResult &r);
// The device images find_symbols needs room for with F frames: one a lane of the pool (made here if not yet), or 1.
int finder_slots(FinderBufs &b, int F);
// The same for F frames of one size at once (d_imgs: room for finder_slots(F) device images, rows w apart, each w h
// bytes apart in one buffer; imgs: the host frames, rows stride apart), each frame uploaded into its lane's slot and
// searched on a host thread of its own with a stream of its own. Each frame's result is the one find_symbol gives it. done (may be empty)
// is called with each frame's index and the lane that searched it on that lane's thread, as soon as its result is in
// place.
void find_symbols(FinderBufs &b, int F, uint8_t *d_imgs, const uint8_t *const *imgs, int w, int h, int stride,
                  hipStream_t s, FindResult *r, const std::function<void(int, FinderLane &)> &done);

}  // namespace lz
