// The border's nodes on the GPU (liblizard/src/acquire.c:1242-1468 refine_node, its scalar form, and the order
// acquire_core refines them in, :5046-5058): each marked node of the lattice's border found by its template of the
// border's own cells, from where its measured neighbours and the lens model put it; and the capture the templates read,
// its border band under a box a module wide (acquire.c:4922-4960 border_blur).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "lens.hpp"
#include "spectrum.hpp"

namespace lz {

// A node of the lattice (acquire.c internal.h:123 node_t): its displacement from H (pixels), its template's score,
// measured or not.
struct NodeState {
  float dx, dy, score;
  int done;
};

// One frame's refinement: the capture the templates read (device, rows w bytes apart), the layout's module map (Lw x Lw
// kinds, device), the marked nodes in their order and levels (device: order[level[l] .. level[l + 1]) are level l's,
// no two of them neighbours, each after every neighbour that precedes it in the reference's order), the node lattice
// (device, nx * nx, in and out), H and optionally a lens model about it.
struct BorderJob {
  const uint8_t *img, *kinds;
  const short *order, *level;
  NodeState *nodes;
  int w, h, span, nx, Lw, levels, lens_on, K;
  float H[9];
  float pos[MAX_NODES];
  double t[LENS_P], ax, ay, R;
};

// The marked nodes of a ring's lattice in the reference's order (ring by ring outward from the corners, rows within a
// ring) grouped into levels as above. Returns the levels.
int border_schedule(int span, std::vector<short> &order, std::vector<short> &level);
// The border band's four rectangles in pixels (x0, y0, x1, y1 each; empty ones x1 <= x0) for a box of radius r.
void border_band(int span, const float *H, int w, int h, int r, int rects[16]);
// out = img with every pixel inside one of the rectangles the rounded mean of its (2r + 1)^2 box (edges repeated).
void launch_border_blur(const uint8_t *img, uint8_t *out, int w, int h, int r, const int rects[16], hipStream_t s);
// Every frame's marked nodes refined, level after level.
void launch_border_nodes(const BorderJob *d_jobs, int frames, hipStream_t s);

}  // namespace lz
