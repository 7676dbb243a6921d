/* lizard_amd: LIZARD's receive hot path on an MI355X (gfx950), behind liblizard's decode API.
 *
 * Two layers of C ABI, plain pointers and sizes only:
 *   1. liblizard's own decode API (liblizard/include/lizard.h:125-155): the same symbols, struct layouts, pixel formats
 *      and error codes, so a caller of the reference links against this library unchanged.
 *   2. lizx_*: the reference's internal seams of the same path, which a GPU half plugs into (lizard_finish_ext,
 *      liblizard/src/lizard_phy.h:236-240; ldpc_decode_stall, liblizard/src/ldpc.h:47-49), and their batched forms.
 *
 * Errors: a function returns LIZ_OK (0) or a count, or a negative LIZ_E_*; one returning a handle returns NULL. The
 * calling thread's last failure is liz_last_error() / liz_last_error_code() (liblizard/api/api.cpp:16-36). No C++
 * exception crosses this boundary. A handle is used from one thread at a time; different handles from different
 * threads at once (liblizard/include/lizard.h:11-13). Each handle owns one HIP stream on its device.
 */
#ifndef LIZARD_AMD_H
#define LIZARD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIZ_API __attribute__((visibility("default")))

/* ---- liblizard's decode API (liblizard/include/lizard.h) -------------------------------------------------------- */

enum {
  LIZ_OK = 0,
  LIZ_E_ARG = -1,
  LIZ_E_NOMEM = -2,
  LIZ_E_IO = -3,
  LIZ_E_UNSUPPORTED = -4,
  LIZ_E_FORMAT = -5,
  LIZ_E_TOOBIG = -6,
  LIZ_E_STATE = -7,
  LIZ_E_INTERNAL = -8
}; /* lizard.h:50-60 */

enum { LIZ_BLOCK = 476, LIZ_ID_BYTES = 4, LIZ_PAYLOAD = 472, LIZ_MIN_BLOCKS = 4, LIZ_MAX_BLOCKS = 64 }; /* lizard.h:62-73 */

typedef enum { LIZ_GREY8 = 1, LIZ_RGBX8 = 2, LIZ_RGBA8 = 3, LIZ_BGRA8 = 4 } liz_pixfmt; /* lizard.h:74-79 */

/* lizard.h:132-148, field for field */
typedef struct {
  int found;
  int ring;
  int n;
  int word;
  int blocks;
  int fps;
  int held_used;
  int total;
  int verified;
  float quad[8];
  int pilot_blocks;
  float pilot_r[2];
  float pilot_sd[2];
} liz_decoded;

/* lizard.h:64-70 and :87-98, field for field: a format (blocks: its size 4..64, the version its word names; ring 0..2
   or LIZ_RING_DEFAULT; fps 1..255; codes 1 or 2 symbols side by side) and what it paints. */
enum { LIZ_RINGS = 3, LIZ_RING_DEFAULT = -1, LIZ_GAP_MODULES = 12 };
typedef struct {
  int blocks, ring, fps, codes;
} liz_format;
typedef struct {
  int n, span, pxm, side, width, height, gap, frame_blocks;
} liz_geometry;

/* Replace lizard.h:102-112 (api/format.cpp:28-87): the cells a side of a ring, a format's geometry, the display room a
   size needs, the blocks a size carries, the largest size that fits a room. Host only. */
LIZ_API int liz_ring_cells(int ring);
LIZ_API int liz_geometry_of(const liz_format *format, liz_geometry *out);
LIZ_API double liz_room_for(int blocks, int ring);
LIZ_API int liz_blocks_for(int blocks);
LIZ_API int liz_pick(double w, double h, int codes, int ring, int top_blocks);

/* Replace lizard.h:116-123 (api/codec.cpp:47-99): the per-frame encoder on the GPU. paint: blocks (frame_blocks x 476
   bytes), picture (its count mod 4 the pilots'), out width x height pixels of fmt, rows stride bytes apart, grey on
   white margins: the bit side, the picture's inverse 2D DFT, the Lanczos-3 resample of the periodic picture into the
   page with its guard interval, the anchors, the border and its word, the quiet zone (lizard_phy.c:970-1236, :672-716),
   each code's symbol in its place with the gap after it. */
typedef struct liz_encoder liz_encoder;
LIZ_API liz_encoder *liz_encoder_new(const liz_format *format);
LIZ_API void liz_encoder_free(liz_encoder *e);
LIZ_API int liz_encoder_geometry(const liz_encoder *e, liz_geometry *out);
LIZ_API int liz_encoder_paint(liz_encoder *e, const uint8_t *blocks, uint32_t picture, uint8_t *out, int stride,
                              liz_pixfmt fmt);

typedef struct liz_decoder liz_decoder;

/* Replaces lizard.h:17 liz_version. */
LIZ_API const char *liz_version(void);
/* Replace lizard.h:41-43 (api/api.cpp:33-34): the calling thread's last failure. */
LIZ_API const char *liz_last_error(void);
LIZ_API int liz_last_error_code(void);
/* Replaces lizard.h:127 (api/codec.cpp:101-114). nmax 384..1536, 0 = 1536. The handle owns a HIP stream and the
   device buffers for the largest picture; NULL and LIZ_E_UNSUPPORTED where no GPU is usable. */
LIZ_API liz_decoder *liz_decoder_new(int nmax);
/* Replaces lizard.h:128. */
LIZ_API void liz_decoder_free(liz_decoder *d);
/* Replaces lizard.h:130 (api/codec.cpp:118). */
LIZ_API int liz_decoder_max_blocks(const liz_decoder *d);
/* Replaces lizard.h:154 (api/codec.cpp:121-160). A frame decodes through the registration held by lizx_decoder_hold
   (the seam the reference's own split decode uses, liblizard/src/lizard_phy.h:193-199, lizard_held_reg
   lizard_phy.c:1899-1900), or, with none held, through the one lizx_find finds on it (lizard_acquire's place; not
   found: returns 0 with out->found = 0, out->ring = -1). The frame's own format word is read through that registration (acquire.c:4503, fmt.h:75) and the
   version taken as liblizard/src/any.c:48-51 takes it: the word's (out->word = 1, out->fps, *held set to it), else
   *held when it is 4..64 (out->held_used = 1), else the version the hold states (0: none); with no version, or a
   picture larger than nmax, the frame is not decoded (returns 0, out->n = 0). out->quad: the lattice's corner nodes
   in the image. */
LIZ_API int liz_decode(liz_decoder *d, const uint8_t *px, int w, int h, int stride, liz_pixfmt fmt, int *held,
                       uint8_t *verified, liz_decoded *out);

/* Replaces lizard.h:48 (api/api.cpp:37): 1, the codec's vector path (here the GPU's) compiled in. */
LIZ_API int liz_simd(void);
/* Replaces lizard.h:164 (api/codec.cpp:177-179, src/shake.c): the test stream's payload for an id, SHAKE256 (FIPS 202)
   of its four little-endian bytes, LIZ_PAYLOAD bytes. */
LIZ_API void liz_stream_fill(uint32_t id, uint8_t *payload);

/* Replaces lizard.h:161 (api/codec.cpp:162-175, core/rx/crops.h): where a camera frame's symbols are looked for:
   layout 1 the centre square, 2 two squares side by side along the long side; rect[k] = x, y, w, h; returns how many. */
LIZ_API int liz_layout_rects(int w, int h, int layout, int rect[2][4]);

/* ---- lizx: the path's seams ------------------------------------------------------------------------------------- */

/* Per-block outcome codes, as the reference's lizard_block_stats (liblizard/src/lizard_phy.h:188-190): iterations
   used (> 0), LIZX_ITS_CAP (-1, never converged), LIZX_ITS_DECLINED (-2, decline gate), LIZX_ITS_STALL (-3, stall
   rule), 0 not tried. */
enum { LIZX_ITS_CAP = -1, LIZX_ITS_DECLINED = -2, LIZX_ITS_STALL = -3 };

/* Format facts for a version (4..64; LIZARD-16V): info[0..5] = n, subch, blocks, tiers, vblocks, slots used. Host
   only, no device. Returns 0 or LIZ_E_FORMAT. Mirrors lizard_init's layout (liblizard/src/lizard_phy.c:485-578). */
LIZ_API int lizx_format_info(int version, int *info);

/* The encoder's bit side of a frame (liblizard/src/lizard_phy.c:1087-1150: CRC-32C, LDPC encode, the bit map and
   whitening, the pilots' tails, the padding), host: blocks * 476 bytes in, subch * 640 frame slots (0/1) out, cnt the
   painted pictures' count mod 4. Returns 0 or a negative LIZ_E_*. */
LIZ_API int lizx_frame_slots(int version, const uint8_t *blocks, int cnt, uint8_t *slots);
/* Block b of LIZARD-16V: its code's rate index, its first frame slot and its slot count (liblizard/src/lizard_phy.h
   block_tier / block_slot, lizard_phy.c:485-637); returns the block count. */
LIZ_API int lizx_format_blocks(int version, int *rate, int *slot0, int *nslots);
/* The coefficient order (research/SPEC.md 6.3) of LIZARD-16V: u and v of each of its subch * 320 coefficients. */
LIZ_API int lizx_format_coefs(int version, int16_t *u, int16_t *v);
/* A camera for synthetic captures: w x h pixels; H maps module coordinates (0 .. span + 30 on a side) to capture
   pixels; a radial lens k1 |q|^2 q D about the frame's centre, q the offset over the farthest corner's distance D
   (the reference's lens form, liblizard/src/acquire.c:1470-1512) and a second radial term k2 |q|^4 q D (the
   reference's second lens term, which a single k1 does not hold); Gaussian blur sigma (pixels, edges repeated); AWGN
   sigma (grey levels) from a counter-based generator seeded by seed. */
typedef struct {
  int w, h;
  double H[9];
  double k1;
  float blur, noise;
  uint64_t seed;
  double k2;
} lizx_camera;
/* The channel on the GPU: picture (n x n floats in [0, 1], device, LIZARD-16V shown in the ring of span modules;
   grey 128 outside it) -> capture (h x w uint8, device). No counterpart in the reference's decoder (a measurement
   input maker; the model of lizard_amd/sim.py's capture). */
LIZ_API int lizx_capture(liz_decoder *d, int version, int span, const float *d_picture, const lizx_camera *cam,
                         uint8_t *d_img);
/* The same camera on the whole symbol as a page shows it (liblizard/src/lizard_phy.c:1150-1236 and :700-716 paint):
   the border of lizx_border_kinds with the word for (version, fps), the 7 x 7 anchors, the picture taken as periodic
   (its 3-module guard interval the picture wrapped round its edge), 3 paint pixels a module, white round it. The
   camera's H maps module coordinates of the symbol (0 .. span + 30) to capture pixels. No counterpart in the
   reference's decoder (a measurement input maker). */
LIZ_API int lizx_capture_symbol(liz_decoder *d, int version, int span, int fps, const float *d_picture,
                                const lizx_camera *cam, uint8_t *d_img);
/* lizx_frame_slots on the GPU for frames x (blocks x 476 bytes) in device memory -> frames x subch x 640 slot bytes
   (0/1) in device memory; bit for bit lizx_frame_slots'. */
LIZ_API int lizx_frame_slots_device(liz_decoder *d, int version, int frames, const uint8_t *d_blocks, int cnt,
                                    uint8_t *d_slots);
/* The sender's pictures on the GPU: frames x (blocks x 476 bytes) (host, frame after frame) with pilot count cnt ->
   frames n x n float pictures in [0, 1], row-major, into device memory d_pictures (frame f at f n n floats). The
   bit side is lizx_frame_slots'; the picture is research/SPEC.md 6.5-6.7's (QPSK at each coefficient, the
   Hermitian spectrum's unscaled inverse DFT, clipped at 2 sqrt(640 subch)), the one lizard_encode resamples into its
   drive (liblizard/src/lizard_phy.c:1087-1200); the bit side (CRC, LDPC, bit map, whitening) on the GPU too.
   Returns when the pictures are written; *kernel_ms (optional) receives the GPU time after the blocks' upload (HIP
   events). */
LIZ_API int lizx_modulate(liz_decoder *d, int version, int frames, const uint8_t *blocks, int cnt, float *d_pictures,
                          float *kernel_ms);

/* lizard_finish_ext with ext_llr (liblizard/src/lizard_phy.c:1566, :1753-1871): the frame's soft values made
   elsewhere: llr = 16 * version * 640 int8 in frame slot order (whitened), est = one information estimate per
   sub-channel. Decline gate, bit map, layered min-sum LDPC (30 iterations, stall 9 / 0.95), bit pack and CRC-32C on
   the GPU. Host buffers in: blocks (blocks * 476 B, every block's decoded bytes), ok (1 where the CRC verified), its
   (per block, codes above), blk_est (per block). Returns the verified count. */
LIZ_API int lizx_finish_llr(liz_decoder *d, int version, const int8_t *llr, const float *est, uint8_t *blocks,
                            uint8_t *ok, int8_t *its, float *blk_est);

/* Optional outputs of the soft-value stages, for callers that check them (each pointer may be NULL). */
typedef struct {
  int8_t *llr;      /* 16 * version * 640 soft values, frame slot order (whitened), as lizard_phy.c's llrf */
  double *sub_est;  /* per sub-channel information estimate (axis_info, lizard_phy.c:75) */
  float *sub_rms;   /* per sub-channel rms axis value sqrt(m2 / 2) */
  float *coef;      /* 16 * version * 320 coefficients (re, im), coefficient order, after the detrend, before the
                       pilots' turn (lizard_debug's coefficients, lizard_phy.c:1719) */
  float align[2];   /* out: the pilots' grid shift in samples along u and v, or 0, 0 (lizard_align) */
  float *blk_pilot; /* per block: the pilot tail's mean of +-y / rms (lizard_phy.c:1795-1809), NaN where none */
  float *soft;      /* 16 * version * 640 floats: each slot's k y before the clamp to +-10 and the rounding of 8 k y
                       (lizard_phy.c:1764-1768), frame slot order; to tell rounding ties from real differences */
} lizx_soft;

/* lizard_finish with the caller's grid (liblizard/src/lizard_phy.c:1566 with grid, ext_llr NULL; lizard_phy.h:
   234-240): n x n picture samples in [0, 1], row-major (row y, column x; n = the format's picture size), of
   LIZARD-16V in the ring of the given span (128, 256 or 512 modules). On the GPU: anchor erase, detrend, 2D FFT,
   pilot alignment, soft demap, then the path of lizx_finish_llr. Outputs as lizx_finish_llr, and soft (optional). */
LIZ_API int lizx_finish_grid(liz_decoder *d, int version, int span, const float *grid, uint8_t *blocks, uint8_t *ok,
                             int8_t *its, float *blk_est, lizx_soft *soft);

/* The node lattice of a ring (span 128, 256 or 512 modules): its points' module coordinates, the same on both axes
   (liblizard/src/layout.c:110-118). Writes up to 40 floats to pos (may be NULL); returns the count a side. Host. */
LIZ_API int lizx_node_lattice(int span, float *pos);

/* A registration as the reference holds it (ob_reg_t, liblizard/src/internal.h:121-124): H maps module coordinates
   to capture pixels (9 floats, row-major, h[8] = 1: u = (h0 x + h1 y + h2) / (h6 x + h7 y + 1), v likewise with h3,
   h4, h5); nodes are the node lattice's displacements in pixels, nx * nx pairs (dx, dy), row by row. */
typedef struct {
  float H[9];
  const float *nodes;
} lizx_registration;

/* The format word (research/SPEC.md 5.3-5.4; liblizard/src/fmt.h, fmt.c, rs.c): the border band's RS(bytes, 3) word
   naming the symbol's version and display rate. Host.
   lizx_word_geometry: its word cells a side and codeword bytes in ring span (fmt.c ob_fmt_side_cells, ob_fmt_bytes).
   lizx_word_encode: the codeword (magic, version, fps, Reed-Solomon parity; ob_fmt_encode); returns its bytes.
   lizx_word_decode: soft values q[side * cells + i] in -0.5 .. 0.5, positive dark -> 1 with *version, *fps, or 0
   (hard decisions, then the least confident bytes as erasures; ob_fmt_decode).
   lizx_read_word: the same from a capture (GREY8 w x h, rows stride bytes apart) through a registration of ring span,
   the band read against the timing track (acquire.c:4503 ob_thin_read_fmt); q (optional) receives the 4 * cells soft
   values. 0 also where the border shows no contrast. */
LIZ_API int lizx_word_geometry(int span, int *cells, int *bytes);
/* The symbol's border as modules (liblizard/src/layout.c ob_layout_init + ob_layout_set_fmt): (span + 30)^2 kinds
   row-major, 0 picture and guard, 1 light, 2 dark, | 8 on the word's cells; version 0: no word (its cells light).
   Returns the count. Host only. */
LIZ_API int lizx_border_kinds(int span, int version, int fps, uint8_t *kinds);
LIZ_API int lizx_word_encode(int span, int version, int fps, uint8_t *cw);
LIZ_API int lizx_word_decode(int span, const float *q, int *version, int *fps);
LIZ_API int lizx_read_word(int span, const uint8_t *img, int w, int h, int stride, const lizx_registration *reg,
                           float *q, int *version, int *fps);

/* The warp alone (the reference's picture sampler, liblizard/src/acquire.c:5121-5196 ob_sample_grid): the n x n
   picture samples of LIZARD-16V in the ring of span, through reg, from a w x h GREY8 capture (rows stride bytes
   apart), into grid (row-major). The samples are the reference's scalar grid walk's, to the bit. */
LIZ_API int lizx_warp(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                      const lizx_registration *reg, float *grid);

/* lizard_finish on a capture through a registration (the reference's split decode with the registration held:
   lizard_phy.c:1566-1876 after lizard_acquire): warp, erase, detrend, FFT, pilot shift, soft values, decline gate,
   LDPC, CRC, all on the GPU. Outputs as lizx_finish_grid. */
LIZ_API int lizx_decode_reg(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                            const lizx_registration *reg, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est,
                            lizx_soft *soft);

/* The single-frame paths' equalisation by a known Gaussian blur of sigma picture samples (0 to 8; 0: none, the
   default and the reference's demap), as lizx_batch_otf, for the decoder's following decodes. */
LIZ_API int lizx_decoder_otf(liz_decoder *d, float sigma);
/* Whether the decoder replays its single-frame decodes (liz_decode with a registration held or found,
   lizx_decode_reg) as one HIP graph of the frame's kernels, captured once per format and work areas (on, 1, the
   default), or launches them one by one (0). Results are the same bits either way. No counterpart in the reference
   (liblizard's decoder runs on the CPU, api/codec.cpp:121-160). */
LIZ_API int lizx_decoder_graphs(liz_decoder *d, int on);
/* Waves a workgroup (1, 2 or 4; 4 the default) of the single-frame decode's row FFT, column FFT and demap: launch
   shapes only, the same bits at every one, so a caller may pick them by timing (a lone frame leaves most of the
   chip idle, where smaller workgroups spread it over more CUs). No counterpart in the reference (its decoder runs
   on the CPU). */
LIZ_API int lizx_decoder_shape(liz_decoder *d, int rows, int cols, int demap);
/* The hang bound (liblizard/core/wg/wg.h:48-55 HANG_MS, 4000 ms; LIZ_HANG_MS in the environment sets it at first
   use, as the reference's does): a handle's wait for its own GPU work that lasts longer takes the work as hung, fails
   with LIZ_E_STATE, and the handle refuses every later call (free it and make another, as the reference drops a lost
   device). 0: no bound. Process-wide. Returns 0. */
LIZ_API int lizx_hang_bound(double ms);
/* The finder's registration from the anchors outward. mode 0: H and the nodes of the finder's radial lens alone.
   Otherwise, after the border's fit, the 7 x 7 anchors are searched and registered (liblizard/src/acquire.c:2128-2191
   anchor_register, as lizx_anchor_register) and, where 12 or more are found, the lattice's interior is made from the
   border nodes and the anchors (acquire.c:2219-2358, as lizx_reg_anchor_fill): mode 1 (the default) the finder's own
   border nodes (its fit through thousands of edge points), all trusted; mode 2 the reference's chain
   (acquire.c:5042-5063): H through the lattice's corners, the border's marked nodes refined by their templates from
   the anchors' lens model (as lizx_border_refine), those scoring 0.35 or more trusted. Returns 0. */
LIZ_API int lizx_finder_anchors(liz_decoder *d, int mode);
/* The anchors' search on the GPU (liblizard/src/acquire.c:1876-2036 anchor_find, its scalar form, to the bit): every
   anchor of a ring of span modules searched on a capture (host, w x h grey, rows stride bytes apart, read at gamma 1)
   through H (module -> pixel, h8 = 1) and optionally a lens model about it (15 doubles as lizx_reg_border_fill gives
   it; null: H alone). anchors: 49 x 8 floats, row after row: the centre X, Y (modules), dx, dy (where it was found
   less where H puts it, pixels), the correlation's peak, where it was found u, v (pixels), 1 if found (else 0).
   Returns the anchors found. */
LIZ_API int lizx_anchor_find(liz_decoder *d, int span, const uint8_t *img, int w, int h, int stride, const float *H,
                             const double *model, float *anchors);
/* The registration from the anchors outward (liblizard/src/acquire.c:2128-2191 anchor_register): the anchors
   searched through H, then, round after round (at most 6, while more are found), those still missing through the
   homography and the lens model the found ones say (each round's readings held to the robust rule); none where fewer
   than 12 are found. A corner of quad (the lattice's corners in pixels, 8 floats, top left and clockwise; in and out)
   further than 3 modules (mpx pixels each) from where the anchors' map puts it is moved to where the lens model over
   the anchors and the kept corners puts it and H (in and out) made again from the four: the reference searches such
   a corner again with its node template first (acquire.c:1242), which this entry does not carry; moved (optional)
   receives how many. anchors as lizx_anchor_find's, about the final H; model (15 doubles) the corners' and anchors'
   lens model (acquire.c:2038-2057). Returns 1 where that model fits, else 0. */
LIZ_API int lizx_anchor_register(liz_decoder *d, int span, const uint8_t *img, int w, int h, int stride, float *H,
                                 float mpx, float *quad, float *anchors, double *model, int *moved);
/* The border's marked nodes refined on the GPU (liblizard/src/acquire.c:1242-1468 refine_node, its scalar form, to the
   bit, in acquire_core's order, :5046-5058): each non-corner border node of the lattice searched with its template of
   the border's own cells (the picture and the format word left out; 13 x 5 modules along a side) from the mean of its
   measured neighbours less the lens model (15 doubles; null: none) plus the model at the node, on a capture (host,
   rows stride bytes apart) whose border band reads under a box of radius mpx / 2 pixels where that is 1 or more
   (acquire.c:4922-4960 border_blur); the corners held at 0. nodes: nx * nx x 4 floats (dx, dy pixels, score, done).
   Returns nx. */
LIZ_API int lizx_border_refine(liz_decoder *d, int span, const uint8_t *img, int w, int h, int stride, const float *H,
                               float mpx, const double *model, float *nodes);
/* Holds a registration for liz_decode (the state lizard_acquire leaves for lizard_finish); reg NULL lets it go, and
   liz_decode then finds each frame's symbol itself. The nodes are copied.
   span 128/256/512. version 0: the frames' words (or *held) name the format; 4..64: the format for frames whose word
   does not read and with none held (frames painted without the border word; no counterpart in the reference). */
LIZ_API int lizx_decoder_hold(liz_decoder *d, int version, int span, const lizx_registration *reg);

/* The finder (no entry of its own in the reference, whose liz_decode runs lizard_acquire inside: acquire.c:4631
   find_frame, the corner nodes, the lens fit): from a capture alone (GREY8 w x h, rows stride bytes apart) the
   symbol's ring (*span 128/256/512) and its registration: H (9 floats, module -> pixel) and nodes (caller's room for
   lizx_node_lattice(span)^2 pairs, 40 x 40 at most: the radial lens's field at the nodes). quad (optional, 8 floats):
   the border line's outer corners in the capture, TL TR BR BL. info (optional, 11 floats): k1, the timing track's
   agreement (about 1), the border's contrast (0..1), the fit's rms (modules), edge points used, then the host clock's
   milliseconds: upload, vote + peaks, sides, first probes, rings, refinement. Returns 1 found, 0
   not. Its own design: a Hough vote of the gradients (GPU), the solid line's two edges probed to a fraction of a
   pixel (GPU), the homography and lens fitted to them, ring and orientation by the timing track (host). */
LIZ_API int lizx_find(liz_decoder *d, const uint8_t *img, int w, int h, int stride, int *span, float *H, float *nodes,
                      float *quad, float *info);
/* The finder on 1..256 frames of one size at once (GREY8, rows stride bytes apart): each frame uploaded and searched on
   one of the handle's host threads (up to 16, made on the first call), a stream each; frame f's result is the one
   lizx_find gives it. No counterpart in the reference, whose finder takes a frame a call. Out per frame: spans[f] (0
   not found), found[f], H + 9 f, nodes + 3200 f (room for 40 x 40 pairs each), and where versions is not NULL the
   version its format word names through that registration (lizx_read_word's), -1 where it has none or was not
   found. Returns the frames found. */
LIZ_API int lizx_find_batch(liz_decoder *d, int frames, const uint8_t *const *imgs, int w, int h, int stride,
                            int *spans, float *H, float *nodes, int *found, int *versions);

/* ---- replays: the .lizard file (research/SPEC.md 15) ---------------------------------------------------------- */

/* A replay read a frame at a time (liblizard/core/rx/lizfile.h Reader): NULL with liz_last_error where the file is not
   a LIZARD replay of version 1 (a ZIP of stored entries, its comment "LIZARD replay 1", meta.json with frames and
   sizes, an NNNN.zst entry a frame; every entry's CRC-32 checked). frame: the frame's w h luma, rows w apart (the
   median predictor's residuals, one zstd frame, undone; zstd is the system's libzstd). meta: 0 meta.json, 1
   stats.jsonl ("" none), 2 the layout ("1:1" or "2:1"). */
typedef struct lizx_replay lizx_replay;
LIZ_API lizx_replay *lizx_replay_open(const char *path);
LIZ_API void lizx_replay_close(lizx_replay *r);
LIZ_API int lizx_replay_frames(const lizx_replay *r);
LIZ_API int lizx_replay_frame_info(const lizx_replay *r, int i, int *w, int *h, double *ms);
LIZ_API int lizx_replay_frame(lizx_replay *r, int i, uint8_t *luma);
LIZ_API const char *lizx_replay_meta(const lizx_replay *r, int which);
/* A replay written as it goes (lizfile.h Writer): frames in order, then finish writes stats.jsonl (NULL or "" none)
   and the caller's meta.json; written through path.part, renamed over path at finish. finish returns the file's
   bytes. */
typedef struct lizx_replay_writer lizx_replay_writer;
LIZ_API lizx_replay_writer *lizx_replay_create(const char *path);
LIZ_API int lizx_replay_add(lizx_replay_writer *w, const uint8_t *luma, int width, int height, int stride);
LIZ_API int64_t lizx_replay_finish(lizx_replay_writer *w, const char *meta, const char *stats);
LIZ_API void lizx_replay_writer_free(lizx_replay_writer *w);

/* ---- the transfer's BLAKE3 tree (a verified file's chunks and root) --------------------------------------------- */

/* Every chunk's chaining value of a file in device memory (16-byte aligned): the file cut into chunks of
   2^chunk_log2 bytes (10..24, at most 16,383 chunks, the last shorter), each a whole subtree of BLAKE3's tree at its
   own offset; 32 bytes a chunk into d_cvs (device, 16-byte aligned). Replaces xfer_chunk_cv over every chunk
   (liblizard/src/xfer.h:175, xfer.c:217-221). kernel_ms (optional): the launches' time on the decoder's stream.
   Returns the chunk count. */
LIZ_API int lizx_b3_chunk_cvs(liz_decoder *d, const uint8_t *d_file, uint64_t length, int chunk_log2, uint8_t *d_cvs,
                              float *kernel_ms);
/* BLAKE3 (32 bytes, what b3sum prints) of device data (16-byte aligned) on the GPU. Replaces xfer_b3_hash
   (xfer.h:181) for data already on the device. */
LIZ_API int lizx_b3_hash_device(liz_decoder *d, const uint8_t *d_data, uint64_t length, uint8_t *out,
                                float *kernel_ms);
/* The file's root from its 2..16,383 chunks' chaining values (host). Replaces xfer_root (xfer.h:169). */
LIZ_API int lizx_b3_root(const uint8_t *cvs, uint32_t chunks, uint8_t *root);
/* BLAKE3 of host data on the host (xfer_b3_hash, xfer.h:181). */
LIZ_API int lizx_b3_hash(const uint8_t *data, uint64_t length, uint8_t *out);

/* The transfer's control blocks (host): the header and manifest block payloads of the transfer's wire format,
   version 3 (liblizard/src/xfer.h:40-92), held to the same rules on writing and reading (xfer.c:11-132). The struct
   has xfer_header_t's fields in its order (xfer.h:132-138), so a binding passes either; the returns are xfer.h's
   (0, or -1, -2 version, -3 hash, -4 codec for a header). */
typedef struct {
  uint8_t version, hash, chunk_log2, codec, seed, name_len, type_len;
  uint64_t length;
  uint32_t chunks, sent;
  uint8_t root[32];
  uint8_t name[255], type[158];
} lizx_xfer_header;
enum { LIZX_XFER_RESERVED = -1, LIZX_XFER_DATA = 0, LIZX_XFER_HEADER = 1, LIZX_XFER_MANIFEST = 2 };
/* ceil(length / 2^chunk_log2), 0 past 16,383 chunks or for chunk_log2 outside 10..24 (xfer_chunk_count). */
LIZ_API uint32_t lizx_xfer_chunk_count(uint64_t length, int chunk_log2);
/* Bytes of chunk index (the file's own), 0 past the last (xfer_chunk_len). */
LIZ_API uint64_t lizx_xfer_chunk_len(const lizx_xfer_header *h, uint32_t index);
/* A chunk's fountain source blocks for sent bytes, ceil(sent / 472) (xfer_blocks). */
LIZ_API uint32_t lizx_xfer_blocks(uint32_t sent);
/* 0 where sent and seed can be chunk index's (xfer_sent_ok). */
LIZ_API int lizx_xfer_sent_ok(const lizx_xfer_header *h, uint32_t index, uint32_t sent, int seed);
/* Manifest blocks of a file of chunks chunks, 0 for fewer than 2 (xfer_manifest_blocks). */
LIZ_API int lizx_xfer_manifest_blocks(uint32_t chunks);
/* A block id's kind: LIZX_XFER_DATA, _HEADER, _MANIFEST or _RESERVED (xfer_id_kind, xfer.h:123-126). */
LIZ_API int lizx_xfer_id_kind(uint32_t id);
/* A header for a file of length bytes in chunks of 2^chunk_log2 (xfer_header_init). */
LIZ_API int lizx_xfer_header_init(lizx_xfer_header *h, uint64_t length, int chunk_log2);
/* The 472-byte payload from h, -1 where h breaks a rule (xfer_header_write). */
LIZ_API int lizx_xfer_header_write(const lizx_xfer_header *h, uint8_t *payload);
/* h from a 472-byte payload (xfer_header_parse). */
LIZ_API int lizx_xfer_header_parse(const uint8_t *payload, lizx_xfer_header *h);
/* Manifest block m of a file of chunks chunks whose root is root (xfer_manifest_write). */
LIZ_API int lizx_xfer_manifest_write(const uint8_t *cvs, const uint32_t *sent, const uint8_t *seeds, uint32_t chunks,
                                     const uint8_t *root, uint32_t m, uint8_t *payload);
/* Manifest block m's entries into cvs, sent and seeds at 12 m; nothing written on -1 (xfer_manifest_parse). */
LIZ_API int lizx_xfer_manifest_parse(const uint8_t *payload, const lizx_xfer_header *h, uint32_t m, uint8_t *cvs,
                                     uint32_t *sent, uint8_t *seeds);
/* 0 where the 2..16,383 chaining values combine to root (xfer_manifest_check). */
LIZ_API int lizx_xfer_manifest_check(const uint8_t *cvs, uint32_t chunks, const uint8_t *root);

/* ---- lizx_ldpc: the LDPC decoder alone, batched ----------------------------------------------------------------- */

typedef struct lizx_ldpc lizx_ldpc;

/* Rate indices as liblizard/src/ldpc.h:31-33: 6 (7/8), 4 (3/4), 3 (2/3), 2 (1/2). dims (12 ints): n, k, m, z, mb,
   kb, np, nt, edges, slots_total, message bytes, posterior bytes. Host only. Mirrors ldpc_tables (ldpc.c:1177). */
LIZ_API int lizx_ldpc_dims(int rate, int *dims);
/* A decoder on a HIP device (-1: the current one), with its own stream. NULL on failure. */
LIZ_API lizx_ldpc *lizx_ldpc_new(int device);
LIZ_API void lizx_ldpc_free(lizx_ldpc *h);

/* Flags for the decode calls. */
enum {
  LIZX_LDPC_CRC = 1,     /* also pack bits 0..3807 MSB-first into 476 bytes and test CRC-32C against bits 3808..3839 */
  LIZX_LDPC_TRACE = 2,   /* also write every codeword's final posteriors and messages (trace buffers) */
  LIZX_LDPC_PLAIN = 4    /* one launch per codeword (the plain per-codeword path) instead of one persistent launch */
};

/* ldpc_decode_stall (liblizard/src/ldpc.h:49, ldpc.c:1080) over count codewords at once, host memory. Codeword i has
   rate index rate[i]; its n int8 LLRs (codeword order, > 0 means bit 0, the unsent bits as the caller knows them)
   start at llr + off[i]; its n hard decisions (0/1 a byte) go to bits + off[i]. its[i]: iterations, -1 or -3.
   max_iter >= 1; stall_it 0 disables the stall rule. With LIZX_LDPC_CRC, bytes[i * 476] and crc_ok[i] are written.
   With LIZX_LDPC_TRACE, post[off[i]] (int16, n per codeword, in the reference decoder's workspace order: data bits in
   code order, parity as [block row][position]) and msg[i * 26880] (int8, slots_total * 128: [layer slot][position])
   are written. Returns the count of codewords whose syndrome cleared, or a negative LIZ_E_*. */
LIZ_API int lizx_ldpc_decode(lizx_ldpc *h, int count, const uint8_t *rate, const int8_t *llr, const int64_t *off,
                             uint8_t *bits, int8_t *its, int max_iter, int stall_it, float stall_ratio, int flags,
                             uint8_t *bytes, uint8_t *crc_ok, int16_t *post, int8_t *msg);

/* The same on device memory already resident in HBM, enqueued on the handle's stream with no host synchronisation:
   every pointer is a device pointer, and the task list is made on the device. A rate index with no code is not
   decoded (its 0). lizx_ldpc_sync waits for the stream; lizx_ldpc_last_ms is then the time, by HIP events on the
   handle's stream, from the first decode enqueued since the previous sync to the end of the last. */
LIZ_API int lizx_ldpc_decode_device(lizx_ldpc *h, int count, const uint8_t *d_rate, const int8_t *d_llr,
                                    const int64_t *d_off, uint8_t *d_bits, int8_t *d_its, int max_iter, int stall_it,
                                    float stall_ratio, int flags, uint8_t *d_bytes, uint8_t *d_crc_ok);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc). lizx_ldpc_decode_device with the same arguments, results
   and stream, but its task list and queue counter live in a fixed work area the library holds on each GPU for the
   process's life (16384 codewords, 512 KB); it calls no hipMalloc, hipFree or allocating routine. count above 16384
   is refused (LIZ_E_ARG). Every call on one GPU shares that GPU's area: calls that overlap on one GPU (from several
   handles or threads, or enqueued on streams that run at once) race and must be serialised by the caller. */
LIZ_API int lizx_ldpc_decode_device_noalloc(lizx_ldpc *h, int count, const uint8_t *d_rate, const int8_t *d_llr,
                                            const int64_t *d_off, uint8_t *d_bits, int8_t *d_its, int max_iter,
                                            int stall_it, float stall_ratio, int flags, uint8_t *d_bytes,
                                            uint8_t *d_crc_ok);
LIZ_API int lizx_ldpc_sync(lizx_ldpc *h);
/* The autotuner of the persistent grid: times these device-resident codewords (reps runs a candidate) at every grid
   size from as many workgroups a CU as fit down to 2, keeps the fastest for this handle's later decodes (the bits do
   not depend on it); *per_cu (optional) receives it (0: as many as fit). */
LIZ_API int lizx_ldpc_tune(lizx_ldpc *h, int count, const uint8_t *d_rate, const int8_t *d_llr, const int64_t *d_off,
                           int8_t *d_its, int max_iter, int stall_it, float stall_ratio, int reps, int *per_cu);
LIZ_API float lizx_ldpc_last_ms(lizx_ldpc *h);
/* Sets the persistent grid's size for this handle's later decodes without timing anything: per_cu workgroups a CU
   (0: as many as fit; above that, as many as fit). The grid is min(count, CUs x per_cu) workgroups, each taking
   codewords off the queue until it is empty, so with per_cu = 1 or 2 every workgroup decodes several codewords of
   a large batch. The bits never depend on it. Returns the workgroups a CU now used. */
LIZ_API int lizx_ldpc_shape(lizx_ldpc *h, int per_cu);

/* ---- Registration refinement (no counterpart in the reference, whose registration is estimated on the CPU;
   liblizard/src/acquire.c:1470-1512 lens model, :2196-2352 anchors' mesh). A registration's parameters, params[107]:
   h0..h7 of the homography (module coordinates -> capture pixels, h8 = 1), k1 of a radial term k1 |q|^2 q D about the
   frame's centre (q the offset over the farthest corner's distance D), and the 7 x 7 anchors' (dx, dy) pixels, row
   after row. The node lattice's displacements are k1 times the lens field at each node's place under a reference
   homography H0 (9 floats) plus the anchors' cubic Hermite mesh (interior nodes; 0 at the lattice's ends). */
#define LIZX_REG_PARAMS 107
/* The node lattice's interior from its border as the reference builds it (liblizard/src/acquire.c:1531-1605
   border_fill_axis): the lens model of the border (the first-order change of H, an affine map, a shift and two
   perspective terms, beside K = 0 to 3 radial terms k |q|^(2k) q about the camera's axis, q over the farthest frame
   corner's distance; acquire.c:1470-1512) fitted by least squares to the corners and the border nodes use[] trusts
   (null: all), K chosen by AICc and at most 1 where the axis lies inside the lattice's quad; then every interior node is
   the model plus the Coons patch of what the model leaves on the border. nodes: nx * nx (dx, dy) pixels, host, the
   border read and the interior written; axis: 2 floats (null or NaN: the frame's centre). model (optional, 15
   doubles) receives K, the axis x, y, R and the 11 coefficients (zero past 8 + K). Returns K (-1: no model fitted). */
LIZ_API int lizx_reg_border_fill(int span, const float *H, int w, int h, const float *axis, float *nodes,
                                 const uint8_t *use, double *model);
/* The node lattice from its border and the 7 x 7 anchors as the reference builds it (acquire.c:2219-2358
   border_fill_anchors): the lens model refitted over the border nodes in use and the anchors found, robustly (border
   nodes further from it than the larger of mpx / 2 and three robust spreads are dropped and take the model's
   displacement, three rounds), the model plus the Coons patch of what it leaves on the border at every interior node,
   then the cubic Hermite mesh (acquire.c:2196-2210) of what the anchors still see, each reading shrunk by the share of
   it that is signal, the ends' knots 0. anchors: 49 x 5 floats, row after row: the centre's module X, Y, the
   displacement from H dx, dy in pixels, 1 if found (else 0). mpx: pixels a module. Otherwise as lizx_reg_border_fill;
   returns the model's K. */
LIZ_API int lizx_reg_anchor_fill(int span, const float *H, int w, int h, const float *axis, float *nodes,
                                 const uint8_t *use, const float *anchors, float mpx, double *model);
/* A lens model's displacement (model as lizx_reg_border_fill gives it) at every node's place under H (lensm_at,
   acquire.c:1615-1620); returns nx. */
LIZ_API int lizx_reg_lens_field(int span, const float *H, const double *model, float *nodes);
/* The nodes (ny * nx (dx, dy) pairs, host) of these parameters in a w x h frame; returns nx. */
LIZ_API int lizx_reg_nodes(int span, const float *H0, int w, int h, const float *params, float *nodes);
/* The refinement loss sum (w - alpha r - beta)^2 over the picture's samples, w the capture warped through params, r
   the reference picture (n x n floats, row-major, host: the blurred re-modulation of what is known), and its gradient
   with respect to all LIZX_REG_PARAMS parameters (the hand-written adjoint of the warp: per-sample sums on the GPU,
   the chain through the nodes on the host in FP64). */
LIZ_API int lizx_reg_loss_grad(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                               const float *H0, const float *params, const float *ref, float alpha, float beta,
                               double *loss, double *grad);
/* The loss on the coefficients and its hand-written adjoint through the FFT and the warp (an extension; no
   counterpart in the reference): with e = w - alpha r - beta (w the capture warped through H and the model's nodes,
   r the reference picture, n x n, as lizx_reg_loss_grad) and E its unscaled 2D DFT, loss = (2 / n^2) times the sum of
   |E(u, v)|^2 over the format's coefficients where mask (one byte a coefficient, coefficient order, subch * 320) is
   set: a decision-directed loss restricted to the coefficients whose values are known (the verified blocks' and the
   padding's), so that blocks not yet decoded add nothing to it. Over every coefficient it is lizx_reg_loss_grad's
   loss (Parseval). grad (107 doubles) as lizx_reg_loss_grad's: the inverse transform of the masked coefficients is
   the loss's derivative with respect to the warped picture, carried through the warp's adjoint to H, k1 and the
   anchors. */
LIZ_API int lizx_reg_loss_grad_spectral(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h,
                                        int stride, const float *H0, const float *params, const float *ref,
                                        const uint8_t *mask, float alpha, float beta, double *loss, double *grad);
/* Refinement of h0..h7 and k1 (params[0..8], in and out) and, with anchors nonzero, the 49 anchors' offsets
   (params[9..106]) on the spectral loss of lizx_reg_loss_grad_spectral, the intensity fit (alpha, beta) taken once
   from the first pass. Each step: the Gauss-Newton direction of the warp's J^T J over h0..h7 and k1 (FP64 on the
   host) on that loss's gradient, then, with anchors, the anchors' direction under the nodes' metric carried through
   the anchors' mesh (each node's 2 x 2 block of the warp's J^T J; a 98 x 98 FP64 solve on the host); each direction's
   step halved from the masked share's full step until the loss falls (at most 6 halvings; a step where neither falls
   ends the refinement). trace (optional, steps + 1 floats): the loss before the first step and after each. No
   counterpart in the reference. */
LIZ_API int lizx_reg_refine_spectral(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h,
                                     int stride, const float *H0, float *params, const float *ref,
                                     const uint8_t *mask, int steps, int anchors, float *trace);
/* steps Gauss-Newton steps over h0..h7 on the GPU (k1 and the anchors held; alpha and beta refitted each pass; the
   damped 8 x 8 normal equations solved in FP64 on the device); params' h0..h7 updated in place; trace (optional,
   steps + 1 floats) receives the loss before each step and after the last. */
LIZ_API int lizx_reg_refine(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                            const float *H0, float *params, const float *ref, int steps, float *trace);
/* The same steps over h0..h7 and k1 together (the anchors held): the 9 x 9 normal equations, k1's column the capture's
   gradient along the lens field's blend at each sample; params' h0..h7 and k1 updated in place. */
LIZ_API int lizx_reg_refine_k1(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                               const float *H0, float *params, const float *ref, int steps, float *trace);
/* The same Gauss-Newton steps over H (9 floats, h8 = 1, in and out) with a node lattice held as it is (host, nx * nx
   (dx, dy) pixels: a finder's or the anchors' registration), so a registration of any form is refined. */
LIZ_API int lizx_reg_refine_nodes(liz_decoder *d, int version, int span, const uint8_t *img, int w, int h, int stride,
                                  float *H, const float *nodes, const float *ref, int steps, float *trace);

/* ---- Batches: many frames of mixed formats through one launch sequence (liblizard/include/lizard.h:285-287
   liz_receiver_batch_cap, liblizard/core/rx/receiver.h:77-95 Receiver::push: the reference batches whole frames per
   receiver; here a batch is one HIP graph: per format the fused warp + erase + row FFT, the column FFT, the pilot
   shift and the soft values, then one persistent LDPC launch over every frame's blocks and their CRCs). */
typedef struct lizx_batch lizx_batch;
typedef struct {
  const uint8_t *img;  /* device: tight uint8 luma capture, w x h (LIZX_SRC_CAPTURE) */
  int w, h;
  float H[9];          /* module coordinates -> capture pixels, h8 = 1 */
  const float *nodes;  /* device: the node lattice's (dx, dy) pixels, ny rows of nx */
  const float *grid;   /* device: n x n picture samples, row-major (LIZX_SRC_GRID) */
} lizx_frame;
enum { LIZX_SRC_CAPTURE = 0, LIZX_SRC_GRID = 1 };
enum { LIZX_STAGE_SOFT = 0, LIZX_STAGE_BLOCKS = 1 };
/* LIZX_RUN_GRAPH: the batch as one captured HIP graph (captured on the first run, replayed after); LIZX_RUN_STREAM:
   the same launches without the graph; LIZX_RUN_FRAMES: frame by frame, each frame's own launches (one persistent
   LDPC launch over its blocks), waiting for each frame before the next, as a decoder called once a frame runs;
   LIZX_RUN_PLAIN: frame by frame as LIZX_RUN_FRAMES with the LDPC launched once a codeword. Every mode gives the
   same bits. */
enum { LIZX_RUN_GRAPH = 0, LIZX_RUN_STREAM = 1, LIZX_RUN_FRAMES = 2, LIZX_RUN_PLAIN = 3 };
LIZ_API lizx_batch *lizx_batch_new(int device);
LIZ_API void lizx_batch_free(lizx_batch *b);
/* frames frames, frame i of LIZARD-16 version[i] in the ring of span span[i] (128, 256 or 512 modules); source
   LIZX_SRC_*, stages LIZX_STAGE_*. Builds the plans and the device work areas once; a plan may be run many times. */
LIZ_API int lizx_batch_plan(lizx_batch *b, int frames, const int *version, const int *span, int source, int stages);
/* dims: frames, total blocks, total soft-value bytes, distinct formats. */
LIZ_API int lizx_batch_dims(const lizx_batch *b, int *dims);
/* Enqueues one run over frames[0 .. frames-1] (device-resident inputs) in a LIZX_RUN_* mode. Asynchronous except
   LIZX_RUN_FRAMES and LIZX_RUN_PLAIN, which wait per frame. */
LIZ_API int lizx_batch_run(lizx_batch *b, const lizx_frame *frames, int mode);
/* The autotuner: times each launch shape (waves a workgroup of the row FFT, column FFT and demap: 1, 2 or 4; the
   persistent LDPC grid's workgroups a CU: as many as fit, 2, 4, 6 or 8) on these frames, reps runs a candidate, by
   coordinate descent, and keeps the fastest. Shapes never change a result's bits. choice (optional, 4 ints) receives
   the shape kept. The next LIZX_RUN_GRAPH run captures a new graph. */
LIZ_API int lizx_batch_tune(lizx_batch *b, const lizx_frame *frames, int reps, int *choice);
/* Sets a launch shape (4 ints, as lizx_batch_tune reports). */
LIZ_API int lizx_batch_shape(lizx_batch *b, const int *choice);
/* The autotuner over launch shapes and radix plans: lizx_batch_tune's search, then the radix plan of the row and of
   the column FFT (the 64-point stage across a wave's lanes as R1 x 64/R1, R1 = 4, 8 or 16; 8 by default) by the same
   coordinate descent. A radix plan, unlike a shape, orders the FFT's float operations differently: every plan is
   within float rounding of the others and within 1e-5 relative L2 of FP64, so soft values may differ between plans
   at rounding ties. choice (optional, 6 ints) receives lizx_batch_tune's 4, then the row and column radix plans.
   No counterpart in the reference (its FFT plans are fixed per size, liblizard/src/fft.c:38-64). */
LIZ_API int lizx_batch_tune_plans(lizx_batch *b, const lizx_frame *frames, int reps, int *choice);
/* Sets a launch shape and the radix plans (6 ints, as lizx_batch_tune_plans reports). */
LIZ_API int lizx_batch_shape_plans(lizx_batch *b, const int *choice);
/* Equalisation by each frame's known Gaussian blur: sigma (frames floats, caller order; 0 to 8 picture samples, 0
   none; null: none) for the following runs. The demap then divides each coefficient (u, v), after the pilot turn, by
   the blur's OTF exp(-2 pi^2 sigma^2 (u^2 + v^2) / n^2) before the sub-channel statistics and soft values. No
   counterpart in the reference, whose demap does not equalise (its sub-channel scale absorbs the blur,
   research/SPEC.md 6.10); with no blur set the results are the reference's. */
LIZ_API int lizx_batch_otf(lizx_batch *b, const float *sigma);
/* Waits for the runs; *ms (optional) receives the GPU time from the first run's start since the last sync to the last
   run's end. */
LIZ_API int lizx_batch_sync(lizx_batch *b, float *ms);
/* The last run's blocks, frame after frame in block order: total blocks * 476 bytes (only verified ones written),
   verdicts, iterations (-2 declined, -1/-3 failed), block estimates. Returns the verified count. */
LIZ_API int lizx_batch_blocks(lizx_batch *b, uint8_t *blocks, uint8_t *ok, int8_t *its, float *blk_est);
/* Frame i's soft values (subch * 640), sub-channel estimates and rms, and pilot shift (each optional). */
LIZ_API int lizx_batch_soft(lizx_batch *b, int frame, int8_t *llr, double *sub_est, float *sub_rms, float *align);

/* The systematic IRA encoder (liblizard/src/ldpc.h:44, ldpc.c:117-126), host: count codewords of one rate index;
   data holds count * k bits (0/1 a byte), cw receives count * n. Returns 0 or LIZ_E_ARG. */
LIZ_API int lizx_ldpc_encode(int rate, int count, const uint8_t *data, uint8_t *cw);

/* CRC-32C (Castagnoli; research/SPEC.md 7.1, liblizard/src/layout.c:44-54 ob_crc32c), host. */
LIZ_API uint32_t lizx_crc32c(const uint8_t *p, size_t n);

/* A plain device copy of bytes (16-byte vectors, grid-stride), timed with HIP events: the practical HBM bandwidth the
   rooflines are stated against, and the known-size kernel the HBM counters are calibrated with. *ms receives the
   kernel's duration. */
LIZ_API int lizx_copy_probe(void *d_dst, const void *d_src, size_t bytes, float *ms);

/* Device memory helpers for callers that hold no HIP runtime of their own (tests, benches): allocate, free, copy
   in/out (synchronous). */
LIZ_API void *lizx_dev_alloc(size_t bytes);
LIZ_API void lizx_dev_free(void *p);
LIZ_API int lizx_dev_upload(void *dst, const void *src, size_t bytes);
LIZ_API int lizx_dev_download(void *dst, const void *src, size_t bytes);
/* Number of usable HIP devices (0 where none). */
LIZ_API int lizx_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
