/* ORACLE — test infrastructure only.
 *
 * A plain-C restatement of LIZARD's receive path arithmetic, used as the checker for the HIP product. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only to check or to time the CPU
 * baseline; the product library never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to the reference repository). Every
 * function here is pinned by at least one check that does not retype its formula (tests/test_oracle_*.py):
 *   or_crc32c            RFC 3720 check value 0xE3069283 + research/SPEC.md 7.4 block CRCs + compiled reference
 *   or_whiten            SPEC 7.3 first 128 bits + SPEC 7.4 per-block whitening rows
 *   or_perm              SPEC 7.3 stride table and perm(0..5) at 3/4
 *   or_tiers_for         SPEC 3.1 profile table (LIZARD-64 ... -1024)
 *   or_pos               SPEC 6.3 first 24 (u, v) + last-coefficient table
 *   or_code_*            adopted FNV-1a hashes of H (liblizard/test/ldpc_table.c:19)
 *   or_encode            SPEC 7.4 parity rows and whole-codeword SHA-256
 *   or_decode            compiled reference decoder (oracle/_ref, scalar form) on noisy words: bits, iterations,
 *                        final posteriors and messages; invariants (clean word -> 1 iteration)
 *   or_finish_llr        compiled reference lizard_finish_ext(ext_llr) on the same inputs
 */
#ifndef LIZARD_ORACLE_H
#define LIZARD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint32_t or_crc32c(const uint8_t *p, int n);
void or_whiten(uint8_t *out, int n);
int or_perm_stride(int nt);
void or_perm(int *perm, int nt);
int or_block_slots(int rate);
int or_sent_bits(int rate);
int or_tiers_for(int subch, int *rate, int *blocks);
int or_n_for(int subch);
/* pos[subch * 320] in the strip-block layout; returns vblocks (blocks of bw = min(n/2, 64) columns) or -1 */
int or_pos(int n, int subch, int *pos);

typedef struct or_code or_code;
or_code *or_code_new(int rate);
void or_code_free(or_code *c);
/* dims: n, k, m, z, mb, kb, np, nt, edges, slots_total */
void or_code_dims(const or_code *c, int *dims);
uint32_t or_code_hash(const or_code *c);
void or_encode(const or_code *c, const uint8_t *data, uint8_t *cw);
/* Returns iterations used, -1 (cap) or -3 (stall). L_out: n int16 in the reference decoder's workspace layout (data
   in code order, parity as [block row][position]); R_out: slots_total * 128 int8 ([layer base + slot][position]). */
int or_decode(const or_code *c, const int8_t *llr, uint8_t *out, int max_iter, int stall_it, float stall_ratio,
              int16_t *L_out, int8_t *R_out);

/* A frame format (n, subch): the layout facts the finish path reads. */
typedef struct or_format or_format;
or_format *or_format_new(int n, int subch);
void or_format_free(or_format *f);
/* info: n, subch, blocks, tiers, vblocks, used_slots */
void or_format_info(const or_format *f, int *info);
/* per block: tier rate index, first frame slot, slot count */
void or_format_blocks(const or_format *f, int *rate, int *slot0, int *nslots);
const int *or_format_pos(const or_format *f);

/* The finish path from soft values on (lizard_phy.c:1753-1871 with ext_llr): llr = subch * 640 int8 in frame slot
   order (whitened), est = subch floats. Outputs per block: ok, its (-2 declined, -1 cap, -3 stall, >0), est, and the
   476 decoded bytes. Returns the verified count. */
int or_finish_llr(const or_format *f, const int8_t *llr, const float *est, uint8_t *blocks, uint8_t *ok, int8_t *its,
                  float *blk_est);

/* The encoder's slot bits for a frame (lizard_phy.c:1087-1142 + map_out :121-153): blocks * 476 bytes in, frame
   slots (subch * 640, 0/1) out, c = pilot count mod 4. */
void or_frame_slots(const or_format *f, const uint8_t *blocks, int c, uint8_t *slots);

#ifdef __cplusplus
}
#endif
#endif
