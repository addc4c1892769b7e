/*
 * oracle.h - CPU restatement of GPUseed's seeding path (seed_gen.cu).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the
 * CPU baseline ("port") that bench.py times; the product (libgpuseed.so, the
 * seed_gen command line) never links, loads or calls it.
 *
 * Every function cites the seed_gen.cu lines whose behaviour it restates.
 *
 * Pinning: the tests check every function against brute force on small texts
 * (sorted suffixes, substring search), hand-worked known answers and the
 * reference's literal constants.  Two behaviours follow from reading the
 * reference alone, and are parity unpinned: that a code above 3 after the
 * 13-base window never ends a task (the test at seed_gen.cu:2246/2387 reads
 * the never-assigned outer `base` of 2125; that such a code leaves the
 * interval unchanged is pinned, from find_occ_gpu's 32-bit arithmetic at 333
 * and 2173-2174, in test_oracle_rank.py) and the stability of the device
 * segmented sorts (seed_gen.cu:1233, 3570), assumed stable.  The reference cannot run here, and its repository holds
 * no tests or golden vectors.
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_index or_index;

/* bwt_restore_bwt + bwt_restore_sa (seed_gen.cu:2685-2736).
 * Returns 0, -4 (missing file), -5 (primary/seq_len mismatch) or -2 (memory). */
int or_index_load(const char* prefix, or_index** out);
void or_index_free(or_index* idx);
uint32_t or_index_seq_len(const or_index* idx);
uint32_t or_index_primary(const or_index* idx);
uint32_t or_index_sa_intv(const or_index* idx);
void or_index_L2(const or_index* idx, uint32_t L2[5]);

/* Rank primitives, exposed for the known-answer tests.
 * or_occ: bwt_occ_gpu (seed_gen.cu:195-216); row in [0, seq_len] or 0xFFFFFFFF.
 * or_step: one backward-extension step through find_occ_gpu (329-368) as used
 *          at 2169-2174: returns (l', u') packed as l' | u' << 32.
 * or_lf: bwt_inv_psi_gpu (455-484) including its primary+1 behaviour.
 * or_locate: locate_seeds_gpu (668-684) for one SA row. */
uint32_t or_occ(const or_index* idx, uint32_t row, uint32_t c);
uint64_t or_step(const or_index* idx, uint32_t l, uint32_t u, uint32_t c);
uint32_t or_lf(const or_index* idx, uint32_t row);
uint32_t or_locate(const or_index* idx, uint32_t row);
/* The true suffix-array value of a row (seq_len for row 0, the empty
 * suffix): the sampled-suffix-array walk with the textbook LF, i.e. without
 * the primary+1 behaviour of or_lf.  Not part of the reference's output;
 * it checks the device's full suffix array, inverse and text. */
uint32_t or_sa_true(const or_index* idx, uint32_t row);

/* pack_4bit_fow / pack_4bit_rev code of one input byte (seed_gen.cu:2536-2578). */
uint32_t or_code_fwd(uint8_t ascii);
uint32_t or_code_rev(uint8_t ascii);

typedef struct {
  int k;            /* -k, minimum seed length (seed_gen.cu:2821, 2829)          */
  int smem;         /* -s (2832)                                                  */
  int prune;        /* 0: every task extended to its end (the definition);
                       1: distance-1 pruning (the baseline algorithm)            */
  uint32_t max_occ; /* 0: locate every row (the reference); else seeds whose
                       interval is larger get no positions                        */
  int threads;      /* 0: all online CPUs                                        */
  int use_table;    /* 1: precompute the 13-mer table like 2887-2895            */
  /* Work model of the device kernels (output unchanged, counters only):
   * smem_stop: in SMEM mode a strand's tasks stop once one reaches the read
   *            start (every later task would end there and be filtered out);
   * ext_len:   14..16: a task with a clean ext_len-base window first looks
   *            up the interval of that pattern (one lookup, counted in
   *            w_ext) and extends from there, or falls back to the 13-mer
   *            table when the pattern does not occur; 0 or 13: off. */
  int smem_stop;
  int ext_len;
  /* prune_sat: size records saturate at this value and a saturated record
   * never prunes, like the device's 8-bit (255) and 16-bit (65535) records;
   * 0: exact records. */
  uint32_t prune_sat;
  /* direct_min: work model of the device's direct extension.  A task whose
   * interval has kept a single row for direct_min backward steps, with no
   * size record of 1 at the next position, reads the reference text instead
   * of stepping: one suffix-array read (the text position), one 16-byte read
   * per 64-base text chunk compared (w_text), and one inverse-suffix-array
   * read for the row where the run ends, skipped when the task emits nothing
   * there or is pruned inside the run; an N ends the run (the row is read
   * back and backward steps go on).  Counted in w_direct; the steps it
   * replaces are not counted in w_blocks64/96.  The model holds for reads
   * within the device's 160-base LDS window.  0: off. */
  uint32_t direct_min;
  /* miss_skip: with ext_len <= k, a task whose clean ext_len-base window does
   * not occur ends at once instead of falling back to the 13-mer table.  Its
   * candidate could only be shorter than ext_len, hence than k, so it emits
   * nothing.  Its left end cannot change an in-lane SMEM filter decision:
   * left ends of finished tasks do not increase as the end decreases, so a
   * later candidate of length >= k lies left of both this task's left end
   * and the previous one's.  Its size records only prune (an early exit
   * that never changes the output).  Likewise, with k >= ext_len + 3, a task
   * whose clean (ext_len + 3)-base window is absent ends after the lookup
   * (the device table's presence bits).  Output unchanged; counters only. */
  int miss_skip;
  /* pair_probe: work model of the device's pair table (one byte per 18-base
   * core: which single bases follow it and which precede it in the text).
   * With k >= 19 and miss_skip, once a task's clean 19-base window was
   * found absent, the next task whose 20 bases read[e-20, e) are clean reads
   * the byte of its core read[e-19, e-1) first: one lookup (w_pair) says
   * whether its window read[e-19, e) occurs (a following base) and whether
   * the next task's window read[e-20, e-1) does (a preceding base).  A task
   * whose window is absent ends without its extended-table lookup; the next
   * task, answered by the same byte, needs no lookup at all when its window
   * is absent.  A present window is then looked up as usual (w_ext).
   * pair_probe 2 (the device's): a task whose 22 bases read[e-22, e) are
   * clean reads a second byte in the same operation, that of the core
   * read[e-21, e-3), which answers the windows of task ends e-2 and e-3
   * likewise.  Output unchanged; counters only.  0: off. */
  int pair_probe;
  /* tandem: work model of the device's tandem jumps (maximum period, at most
   * 8; 0: off).  Inside a stretch [R, ..) of the read with period p (code(z)
   * == code(z + p) <= 3 for every z >= R up to the stretch's end), the
   * pattern read[y, e) with R <= y < e is its last p codes (the unit)
   * repeated, so its interval depends only on the unit and on e - y: the
   * device reads it from a table of every unit of at most 6 bases and every
   * length in [20, 256] (or_tandem_row).  A task with no stretch open whose
   * 20 codes left of its end e have a period p <= tandem (the smallest one)
   * finds R by scanning left and opens a stretch if e - R >= tandem_span.  A
   * task of an open stretch with e - R >= 20 (else the stretch closes) takes
   * n = min(e - R, tandem_len) and the longest m <= n whose repeat occurs
   * (one table read, w_tandem): with m < 13 it ends as a 13-mer table miss
   * would (no candidate, left end unchanged); else its interval is the
   * repeat's of length m at y = e - m, in place of its table lookup and the
   * steps down to y, and it goes on from y as after a lookup, or ends there
   * (its interval emptying) when m < n.  Output unchanged (intervals are
   * exact; the skipped size records only prune); counters only.  Needs
   * prune. */
  uint32_t tandem;
  uint32_t tandem_span, tandem_len, tandem_pad;
  /* tandem_check: recompute every jumped-to interval from the text of the
   * read (table lookup and steps), and check that one more base empties it
   * where the jump ends the task; count disagreements in w_tandem_bad. */
  int tandem_check;
} or_opts;

typedef struct {
  uint32_t qbeg, qend; /* bit 31 set on both for the reverse strand (2402)   */
  uint32_t l, u;       /* SA interval rows                                     */
  uint32_t n_pos;      /* positions emitted for this seed                      */
  uint32_t pad;
} or_seed;

typedef struct {
  uint64_t n_reads;
  uint32_t* read_nseeds; /* seeds with at least one position, per read      */
  uint64_t* read_seed_off;
  uint32_t* read_npos;
  uint64_t* read_pos_off;
  uint64_t n_seeds;
  or_seed* seeds;
  uint64_t n_pos;
  uint32_t* pos;
  /* work counters (distance-1 pruning, whatever opts.prune says) */
  uint64_t w_tasks;       /* tasks with a valid 13-base window               */
  uint64_t w_steps;       /* backward-extension steps                        */
  uint64_t w_blocks64;    /* distinct 64-base occ blocks read by those steps */
  uint64_t w_blocks96;    /* distinct 96-base blocks (device layout)        */
  uint64_t w_lf;          /* LF steps of the SA walk                          */
  uint64_t w_sa;          /* SA samples read                                  */
  uint64_t w_steps_unpruned; /* steps with no pruning at all                 */
  uint64_t n_short;       /* reads shorter than k (no tasks)                  */
  uint64_t w_ext;         /* extended-table lookups (opts.ext_len)            */
  uint64_t w_direct;      /* direct-extension reads (opts.direct_min): 4-byte
                             suffix-array and inverse reads and, of them,     */
  uint64_t w_text;        /* 16-byte text reads                               */
  uint64_t w_pair;        /* pair-table bytes read (opts.pair_probe)          */
  uint64_t w_pair_ops;    /* operations that read them (one or two bytes)     */
  uint64_t w_tandem;      /* tandem-table entries read (opts.tandem)          */
  uint64_t w_tandem_bad;  /* jumped-to intervals that disagree (tandem_check) */
  uint64_t w_stretches;   /* tandem stretches opened                          */
} or_result;

/* Seeds reads[0..n): bases+offsets[i] holds lens[i] bytes of read i (raw line
 * bytes, '\r' included, seed_gen.cu:2955-2990).  Reads are numbered in input
 * order; the MEM look-ahead (624-667) stays inside the reference's batches of
 * >= 1,000,000 bases (2910-2991), which are recomputed from lens[]. */
int or_seed_reads(const or_index* idx, const char* bases, const uint64_t* offsets,
                  const uint32_t* lens, uint64_t n, const or_opts* opts, or_result* res);
void or_result_free(or_result* res);

/* The -o text output of one run (seed_gen.cu:3625-3648), reads numbered from
 * first_read_no (1-based). */
int or_format_text(const or_result* res, int smem, uint64_t first_read_no, FILE* out);

/* Entries [20, 256] of the tandem table's row of unit `key` (p bases for
 * key in [(4^p - 4)/3, (4^(p+1) - 4)/3), p = 1..6; code key - (4^p - 4)/3 =
 * sum of code_j * 4^j, the unit's last base as j = 0): 237 entries of four
 * words {l, u, m, 0}, for the string of length n = 20 + index that ends
 * with the unit repeated, m the longest length <= n that occurs and [l, u]
 * its interval (the empty string's [0, seq_len] for m = 0).  No counterpart
 * in the reference. */
int or_tandem_row(const or_index* X, uint32_t key, uint32_t* out4);

#ifdef __cplusplus
}
#endif

#endif
