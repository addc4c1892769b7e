/*
 * oracle.c - CPU restatement of GPUseed's seeding path (seed_gen.cu).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  The
 * product never links or calls this file.
 *
 * What it restates, with the reference lines it follows:
 *   index files ........... bwt_restore_bwt / bwt_restore_sa   2685-2736
 *   rank .................. pop_count_partial/full, bwt_occ_gpu 142-216
 *   backward step ......... find_occ_gpu 329-368, used at 2169-2174 / 2310-2315
 *   13-mer table .......... pre_calc_seed_intervals_gpu 2648-2675
 *   read encoding ......... pack_4bit_fow / pack_4bit_rev 2536-2578
 *   tasks and extension ... find_seed_intervals_gpu 2087-2413
 *   SMEM filter/order ..... filter_seed_intervals_gpu 991-1018, the segmented
 *                           sort at 1233 (stable, key = qbeg), count 865-880
 *   MEM filter/counts ..... filter_seed_intervals_gpu_mem 1103-1139,
 *                           count_seed_intervals_gpu_mem 927-989,
 *                           seeds_to_threads_mem 624-667, sort at 3569-3570
 *   SA walk ............... bwt_inv_psi_gpu 455-484, locate_seeds_gpu 668-684
 *   batching .............. main 2910-2991 (batches of >= 1,000,000 bases)
 *   text output ........... main 3625-3648
 *
 * Behaviour the reference leaves undefined is resolved as follows:
 *   - the outer `base` read at 2246/2387 is never assigned: a code > 3 after
 *     the 13-base window leaves the interval unchanged (N-transparent);
 *   - the warp-shuffle pruning (2192-2212, 2333-2351) is not modelled: it only
 *     removes candidates that the SMEM filter or the MEM counts remove anyway,
 *     so the output is computed from the full candidate lists (prune = 0) or
 *     with exact distance-1 pruning (prune = 1), which gives the same output;
 *   - a read with exactly one forward MEM entry gets that entry's rows (the
 *     reference leaves that output slot unwritten, 1110);
 *   - reads shorter than k produce no seeds (the reference writes task maps
 *     out of range for them, 2630).
 * k must be >= 14: for k <= 13 the reference emits window-only candidates
 * whose values come from the table's break points (2259 vs 2398).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define PRECALC 13 /* pre_calc_seed_len, seed_gen.cu:2822 */
#define BATCH_BASES 1000000u /* BASE_BATCH_SIZE, seed_gen.cu:2814 */

struct or_index {
  uint32_t primary, seq_len, sa_intv;
  uint32_t L2[5];
  uint64_t bwt_size;
  uint32_t* bwt;
  uint64_t n_sa;
  uint32_t* sa;
  uint64_t* table; /* l | u << 32 for every 13-mer, built on demand */
  pthread_mutex_t lock;
};

/* ---- index files (seed_gen.cu:2685-2736) -------------------------------- */

int or_index_load(const char* prefix, or_index** out) {
  *out = NULL;
  or_index* X = (or_index*)calloc(1, sizeof(or_index));
  if (!X) return -2;
  pthread_mutex_init(&X->lock, NULL);
  size_t pl = strlen(prefix);
  char* fn = (char*)malloc(pl + 8);
  memcpy(fn, prefix, pl);
  strcpy(fn + pl, ".bwt");
  FILE* fp = fopen(fn, "rb");
  if (!fp) {
    free(fn);
    free(X);
    return -4;
  }
  fseek(fp, 0, SEEK_END);
  long sz = ftell(fp);
  X->bwt_size = (uint64_t)(sz - 20) >> 2; /* (filesize - 5 words) / 4, 2726 */
  X->bwt = (uint32_t*)calloc(X->bwt_size + 8, 4); /* +8: uint4 reads stay inside */
  fseek(fp, 0, SEEK_SET);
  int rc = 0;
  if (fread(&X->primary, 4, 1, fp) != 1 || fread(X->L2 + 1, 4, 4, fp) != 4 ||
      fread(X->bwt, 4, X->bwt_size, fp) != X->bwt_size)
    rc = -4;
  fclose(fp);
  X->seq_len = X->L2[4];
  if (!rc) {
    strcpy(fn + pl, ".sa");
    fp = fopen(fn, "rb");
    if (!fp) rc = -4;
    uint32_t h[7];
    if (!rc && fread(h, 4, 7, fp) != 7) rc = -4;
    if (!rc && h[0] != X->primary) rc = -5;            /* 2697-2700 */
    if (!rc && h[6] != X->seq_len) rc = -5;            /* 2704-2707 */
    if (!rc && h[5] == 0) rc = -5;
    if (!rc) {
      X->sa_intv = h[5];
      X->n_sa = ((uint64_t)X->seq_len + X->sa_intv) / X->sa_intv; /* 2709 */
      X->sa = (uint32_t*)calloc(X->n_sa, 4);
      X->sa[0] = 0xFFFFFFFFu; /* 2711 */
      if (fread(X->sa + 1, 4, X->n_sa - 1, fp) != X->n_sa - 1) rc = -4;
    }
    if (fp) fclose(fp);
  }
  free(fn);
  if (rc) {
    or_index_free(X);
    return rc;
  }
  *out = X;
  return 0;
}

void or_index_free(or_index* X) {
  if (!X) return;
  free(X->bwt);
  free(X->sa);
  free(X->table);
  pthread_mutex_destroy(&X->lock);
  free(X);
}

uint32_t or_index_seq_len(const or_index* X) { return X->seq_len; }
uint32_t or_index_primary(const or_index* X) { return X->primary; }
uint32_t or_index_sa_intv(const or_index* X) { return X->sa_intv; }
void or_index_L2(const or_index* X, uint32_t L2[5]) { memcpy(L2, X->L2, sizeof X->L2); }

/* ---- rank (seed_gen.cu:142-216) ----------------------------------------- */

static inline uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

static inline uint32_t pop_full(uint32_t w, uint32_t c) { /* 152-158 */
  uint32_t odd = ((c & 2) ? w : ~w) >> 1;
  uint32_t even = (c & 1) ? w : ~w;
  return popc(odd & even & 0x55555555u);
}

static inline uint32_t pop_partial(uint32_t w, uint32_t c, uint32_t mask_bits) { /* 142-150 */
  w &= ~((1u << (mask_bits << 1)) - 1u);
  uint32_t odd = ((c & 2) ? w : ~w) >> 1;
  uint32_t even = (c & 1) ? w : ~w;
  uint32_t m = popc(odd & even & 0x55555555u);
  return c == 0 ? m - mask_bits : m;
}

/* count of c in the $-removed BWT [0..k], k a $-removed index */
static inline uint32_t occ_idx(const or_index* X, uint32_t k, uint32_t c) {
  const uint32_t* blk = X->bwt + (uint64_t)(k >> 6) * 8;
  const uint32_t* w = blk + 4;
  uint32_t kw = (k & 63) >> 4;
  uint32_t n = blk[c];
  if (kw > 0) n += pop_full(w[0], c);
  if (kw > 1) n += pop_full(w[1], c);
  if (kw > 2) n += pop_full(w[2], c);
  return n + pop_partial(w[kw], c, (~k) & 15);
}

uint32_t or_occ(const or_index* X, uint32_t k, uint32_t c) { /* 195-216 */
  if (k == X->seq_len) return X->L2[c + 1] - X->L2[c];
  if (k == 0xFFFFFFFFu) return 0;
  if (k >= X->primary) --k;
  return occ_idx(X, k, c);
}

/* One step of 2169-2174: the interval of c.P from the interval [l,u] of P.
 * find_occ_gpu returns (l-1-L2[c], u-L2[c]) for c > 3 (333), so the step
 * leaves the interval unchanged; the L2_gpu[5] read for c == 5 cancels. */
static inline void step(const or_index* X, uint32_t* l, uint32_t* u, uint32_t c) {
  if (c > 3) return;
  uint32_t o1 = or_occ(X, *l - 1, c), o2 = or_occ(X, *u, c);
  *l = X->L2[c] + o1 + 1;
  *u = X->L2[c] + o2;
}

uint64_t or_step(const or_index* X, uint32_t l, uint32_t u, uint32_t c) {
  step(X, &l, &u, c);
  return (uint64_t)l | ((uint64_t)u << 32);
}

/* bwt_inv_psi_gpu (455-484): the decrement comes before the primary test, so
 * rows primary and primary+1 both map to 0. */
uint32_t or_lf(const or_index* X, uint32_t k) {
  if (k > X->primary) --k;
  if (k == X->primary) return 0;
  const uint32_t* blk = X->bwt + (uint64_t)(k >> 6) * 8;
  uint32_t c = (blk[4 + ((k & 63) >> 4)] >> (((~k) & 15) << 1)) & 3;
  if (k == X->seq_len) return X->L2[c + 1] - X->L2[c];
  return X->L2[c] + occ_idx(X, k, c);
}

/* locate_seeds_gpu (668-684) for one row */
static uint32_t locate_count(const or_index* X, uint32_t row, uint64_t* lf_steps) {
  uint32_t itr = 0;
  /* rows past the last one only arise from the batch-wide MEM row skip
   * (624-667) and have no position; the reference would read past its arrays */
  if (row > X->seq_len) return 0xFFFFFFFFu;
  while (row % X->sa_intv) {
    ++itr;
    row = or_lf(X, row);
  }
  if (lf_steps) *lf_steps += itr;
  return X->sa[row / X->sa_intv] + itr;
}

uint32_t or_locate(const or_index* X, uint32_t row) { return locate_count(X, row, NULL); }

/* Textbook LF (row of the suffix one position to the left) and the sampled
 * suffix-array walk built on it.  The reference's walk differs only at row
 * primary+1 (or_lf above); this one is the true SA, which the device's direct
 * extension reads (not an output of the reference). */
static uint32_t lf_true(const or_index* X, uint32_t row) {
  uint32_t k = row - (row > X->primary ? 1u : 0u);
  const uint32_t* blk = X->bwt + (uint64_t)(k >> 6) * 8;
  uint32_t c = (blk[4 + ((k & 63) >> 4)] >> (((~k) & 15) << 1)) & 3;
  return X->L2[c] + occ_idx(X, k, c);
}

uint32_t or_sa_true(const or_index* X, uint32_t row) {
  if (row > X->seq_len) return 0xFFFFFFFFu;
  if (row == 0) return X->seq_len;
  uint32_t itr = 0;
  while (row % X->sa_intv) {
    if (row == X->primary) return itr; /* text position 0: the walk would wrap to $ */
    ++itr;
    row = lf_true(X, row);
  }
  return row == 0 ? X->seq_len : X->sa[row / X->sa_intv] + itr;
}

/* ---- 13-mer table (2648-2675) ------------------------------------------- */

static uint64_t kmer13(const or_index* X, uint32_t p) {
  uint32_t l = 0, u = X->seq_len;
  for (int i = 0; i < PRECALC; ++i) {
    step(X, &l, &u, (p >> (i << 1)) & 3);
    if (l > u) break;
  }
  return (uint64_t)l | ((uint64_t)u << 32);
}

typedef struct {
  void (*fn)(void* ctx, uint64_t i, int tid);
  void* ctx;
  uint64_t n;
  _Atomic uint64_t next;
  uint64_t chunk;
} pfor_t;

typedef struct {
  pfor_t* pf;
  int tid;
} pfor_arg;

static void* pfor_worker(void* a) {
  pfor_arg* pa = (pfor_arg*)a;
  pfor_t* pf = pa->pf;
  for (;;) {
    uint64_t b = atomic_fetch_add(&pf->next, pf->chunk);
    if (b >= pf->n) break;
    uint64_t e = b + pf->chunk < pf->n ? b + pf->chunk : pf->n;
    for (uint64_t i = b; i < e; ++i) pf->fn(pf->ctx, i, pa->tid);
  }
  return NULL;
}

static void parallel_for(uint64_t n, int threads, uint64_t chunk, void (*fn)(void*, uint64_t, int),
                         void* ctx) {
  pfor_t pf;
  pf.fn = fn;
  pf.ctx = ctx;
  pf.n = n;
  atomic_init(&pf.next, 0);
  pf.chunk = chunk ? chunk : 1;
  if (threads <= 1 || n <= pf.chunk) {
    for (uint64_t i = 0; i < n; ++i) fn(ctx, i, 0);
    return;
  }
  pthread_t th[256];
  pfor_arg args[256];
  if (threads > 256) threads = 256;
  for (int t = 0; t < threads; ++t) {
    args[t].pf = &pf;
    args[t].tid = t;
    pthread_create(&th[t], NULL, pfor_worker, &args[t]);
  }
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}

static void table_fill(void* ctx, uint64_t i, int tid) {
  (void)tid;
  or_index* X = (or_index*)ctx;
  const uint64_t per = 1u << 16;
  for (uint64_t p = i * per; p < (i + 1) * per; ++p) X->table[p] = kmer13(X, (uint32_t)p);
}

static int ensure_table(or_index* X, int threads) {
  pthread_mutex_lock(&X->lock);
  if (!X->table) {
    uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) << (2 * PRECALC));
    if (!t) {
      pthread_mutex_unlock(&X->lock);
      return -2;
    }
    X->table = t;
    parallel_for((1ull << (2 * PRECALC)) >> 16, threads, 1, table_fill, X);
  }
  pthread_mutex_unlock(&X->lock);
  return 0;
}

/* ---- encoding (2536-2578) ------------------------------------------------ */

uint32_t or_code_fwd(uint8_t a) { return (0x24031005u >> ((a & 7) << 2)) & 15; }
uint32_t or_code_rev(uint8_t a) { return (0x14002035u >> ((a & 7) << 2)) & 15; }

/* ---- per-strand candidates (find_seed_intervals_gpu, 2087-2413) ----------
 * Both strands run the same procedure on a "virtual" sequence v[0..L):
 *   forward: v[i] = fwd code of read[i]; task t has end e = L - t (2132-2264);
 *   reverse: v[i] = rev code of read[L-1-i]; the reference's reverse task t
 *            (start t, extending rightwards, 2272-2404) is the virtual task
 *            with end e = L - t, and a virtual candidate [b, e) is the reverse
 *            seed qbeg = L - e, qend = L - b.
 * Candidates come out in decreasing virtual end. */

typedef struct {
  uint32_t b, e, l, u;
} cand_t;

typedef struct {
  uint64_t tasks, steps, blocks64, blocks96, lf, sa, ext, direct, text, pair, pair_ops, tandem, tandem_bad, stretches;
} work_t;

#define TANDEM_MIN 20u  /* shortest stretch length a jump reads (the table's first length) */
#define TANDEM_LEN 256u /* the table's longest length */

/* the interval of the clean pattern v[b, e), e - b >= 13: its 13-mer table
 * entry, then one backward step per further base */
static void interval_of(const or_index* X, const uint8_t* v, uint32_t b, uint32_t e, uint32_t* l, uint32_t* u) {
  uint32_t p13 = 0;
  for (uint32_t q = 0; q < PRECALC; ++q) p13 |= (uint32_t)v[e - 1 - q] << (q << 1);
  const uint64_t t = X->table ? X->table[p13] : kmer13(X, p13);
  *l = (uint32_t)t;
  *u = (uint32_t)(t >> 32);
  for (uint32_t q = e - PRECALC; q-- > b && *l <= *u;) step(X, l, u, v[q]);
}

/* whether the clean 19-base window v[e-19, e) occurs in the text */
static int window19(const or_index* X, const uint8_t* v, uint32_t e) {
  uint32_t l, u;
  interval_of(X, v, e - 19, e, &l, &u);
  return l <= u;
}

/* the smallest period p <= pmax of the 20 codes v[e-20, e), all clean:
 * v[z] == v[z + p] for every z in [e-20, e-p); 0 if none */
static uint32_t tandem_period(const uint8_t* v, uint32_t e, uint32_t pmax) {
  for (uint32_t z = e - 20; z < e; ++z)
    if (v[z] > 3) return 0;
  for (uint32_t p = 1; p <= pmax; ++p) {
    uint32_t z = e - 20;
    while (z + p < e && v[z] == v[z + p]) ++z;
    if (z + p == e) return p;
  }
  return 0;
}

/* Backward search of unit (p codes, code j = (unit >> 2j) & 3 at distance j
 * from the end) repeated, one base per length from the empty string's
 * interval [0, seq_len]: the longest m <= n whose string occurs, its
 * interval in *l, *u. */
static uint32_t tandem_walk(const or_index* X, uint32_t unit, uint32_t p, uint32_t n, uint32_t* l, uint32_t* u) {
  uint32_t cl = 0, cu = X->seq_len, m = 0, j = 0;
  for (uint32_t len = 1; len <= n; ++len) {
    uint32_t nl = cl, nu = cu;
    step(X, &nl, &nu, (unit >> (2 * j)) & 3u);
    if (nl > nu) break;
    cl = nl;
    cu = nu;
    m = len;
    j = j + 1 == p ? 0 : j + 1;
  }
  *l = cl;
  *u = cu;
  return m;
}

int or_tandem_row(const or_index* X, uint32_t key, uint32_t* out4) {
  if (!X || !out4 || key >= 5460u) return -1;
  uint32_t p = 1;
  while (p < 6 && key >= ((1u << (2 * (p + 1))) - 4u) / 3u) ++p;
  const uint32_t unit = key - ((1u << (2 * p)) - 4u) / 3u;
  uint32_t l, u;
  const uint32_t m = tandem_walk(X, unit, p, TANDEM_LEN, &l, &u);
  for (uint32_t n = TANDEM_MIN; n <= TANDEM_LEN; ++n) {
    uint32_t* o = out4 + 4 * (n - TANDEM_MIN);
    if (n < m) { /* the repeat of length n: the same walk, stopped at n */
      tandem_walk(X, unit, p, n, &o[0], &o[1]);
      o[2] = n;
    } else {
      o[0] = l;
      o[1] = u;
      o[2] = m;
    }
    o[3] = 0;
  }
  return 0;
}

static inline uint32_t row_to_idx(const or_index* X, uint32_t row) {
  return row - (row >= X->primary ? 1u : 0u);
}

static void count_blocks(const or_index* X, uint32_t l, uint32_t u, work_t* w) {
  /* rows l-1 and u; -1 and seq_len are answered without memory in the
   * reference's layout (331-336); 96-base blocks have no such special case */
  uint32_t r1 = l - 1, r2 = u;
  int s1 = r1 == 0xFFFFFFFFu, s2 = r2 == 0xFFFFFFFFu;
  uint32_t i1 = s1 ? 0 : row_to_idx(X, r1), i2 = s2 ? 0 : row_to_idx(X, r2);
  int sp1 = s1 || r1 == X->seq_len, sp2 = s2 || r2 == X->seq_len;
  w->blocks64 += (!sp1) + (!sp2) - ((!sp1 && !sp2 && (i1 >> 6) == (i2 >> 6)) ? 1 : 0);
  w->blocks96 += (!s1) + (!s2) - ((!s1 && !s2 && i1 / 96 == i2 / 96) ? 1 : 0);
}

/* One strand's candidates (find_seed_intervals_gpu's forward procedure,
 * 2087-2264; the reverse strand runs it on reverse-complement codes).  The
 * output is the same whatever prune, smem_stop and ext_len say; they only
 * change which steps run and are counted (see or_opts). */
static uint32_t strand_candidates(const or_index* X, const uint8_t* v, uint32_t L, uint32_t k,
                                  int prune, int smem_stop, uint32_t ext_len, uint32_t sat,
                                  uint32_t direct_min, int miss_skip, int pair_probe, int smem,
                                  const or_opts* o, cand_t* out, uint32_t* A, work_t* w) {
  uint32_t n = 0;
  if (L < k) return 0;
  /* tandem-jump model (or_opts.tandem): the open stretch [tR, ..) of period
   * tp (0: none) */
  const uint32_t TP = (o->tandem && prune) ? (o->tandem > 8 ? 8 : o->tandem) : 0;
  const uint32_t tlen = o->tandem_len < TANDEM_MIN ? TANDEM_MIN : o->tandem_len;
  uint32_t tp = 0, tR = 0;
  /* pair-table model (or_opts.pair_probe), with the 16-base extended table:
   * probe is set while the last looked-up clean 19-base window was absent;
   * the last probe also answered the windows of the known_n task ends
   * known_e, known_e - 1, ...: bit i of known_m set iff that of known_e - i
   * occurs */
  const int pair = pair_probe && miss_skip && ext_len == 16 && k >= 19;
  int probe = 0;
  int64_t known_e = -1;
  uint32_t known_n = 0, known_m = 0;
#define KNOWN(t) (known_e - (int64_t)(t) >= 0 && known_e - (int64_t)(t) < (int64_t)known_n)
#define KNOWN_PRESENT(t) ((known_m >> (uint32_t)(known_e - (int64_t)(t))) & 1u)
  if (prune) memset(A, 0, sizeof(uint32_t) * L);
  uint32_t last_b = 0xFFFFFFFFu; /* left end of the last unpruned task (the device's in-lane SMEM filter) */
  for (uint32_t e = L; e >= k; --e) {
    uint32_t p = 0;
    int bad = 0;
    for (uint32_t j = 0; j < PRECALC; ++j) { /* 2132-2143 */
      uint32_t c = v[e - 1 - j];
      if (c > 3) {
        bad = 1;
        break;
      }
      p |= c << (j << 1);
    }
    if (bad) continue;
    uint32_t l = 0, u = 0, b = 0;
    int have = 0;
    /* a task whose window the last pair byte found absent is skipped before
     * anything else (the device's next_task) */
    const int known_absent = pair && KNOWN(e) && !KNOWN_PRESENT(e);
    /* tandem model: this task jumps, or not */
    int jump = 0, tend = 0;
    if (TP && !known_absent) {
      if (tp && !(e >= tR + TANDEM_MIN)) tp = 0; /* the stretch closes */
      if (!tp && e >= 20) {
        const uint32_t q = tandem_period(v, e, TP);
        if (q) {
          int32_t z = (int32_t)e - 21;
          while (z >= 0 && v[z] <= 3 && v[z] == v[z + q]) --z;
          if (e - (uint32_t)(z + 1) >= o->tandem_span) {
            tp = q;
            tR = (uint32_t)(z + 1);
            ++w->stretches;
          }
        }
      }
      if (tp) {
        const uint32_t n = e - tR < tlen ? e - tR : tlen;
        uint32_t unit = 0;
        for (uint32_t j = 0; j < tp; ++j) unit |= (uint32_t)v[e - 1 - j] << (2 * j);
        const uint32_t m = tandem_walk(X, unit, tp, n, &l, &u);
        ++w->tandem;
        if (m < PRECALC) continue; /* as a 13-mer table miss */
        b = e - m;
        jump = have = 1;
        tend = m < n;
        if (pair) probe = 0;
        if (o->tandem_check) {
          uint32_t cl, cu;
          interval_of(X, v, b, e, &cl, &cu);
          if (cl != l || cu != u) ++w->tandem_bad;
          if (tend) { /* one more base of the read must empty the interval */
            if (b == 0 || v[b - 1] > 3) {
              ++w->tandem_bad;
            } else {
              step(X, &cl, &cu, v[b - 1]);
              if (cl <= cu) ++w->tandem_bad;
            }
          }
        }
      }
    }
    if (jump) {
      /* the table's interval stands for the lookup and the steps above b */
    } else if (ext_len > PRECALC && e >= ext_len) {
      int clean = 1;
      for (uint32_t j = PRECALC; j < ext_len; ++j) clean &= v[e - 1 - j] <= 3;
      if (clean) { /* one lookup answers the first ext_len bases when they occur */
        uint64_t iv = X->table ? X->table[p] : kmer13(X, p);
        l = (uint32_t)iv;
        u = (uint32_t)(iv >> 32);
        for (uint32_t j = PRECALC; j < ext_len && l <= u; ++j) step(X, &l, &u, v[e - 1 - j]);
        int absent3 = 0; /* the clean (ext_len + 3)-base window does not occur */
        const int win3 = miss_skip && k >= ext_len + 3 && e >= ext_len + 3 && v[e - ext_len - 1] <= 3 &&
                         v[e - ext_len - 2] <= 3 && v[e - ext_len - 3] <= 3;
        if (l <= u && win3) {
          /* the table's presence bits */
          uint32_t l3 = l, u3 = u;
          step(X, &l3, &u3, v[e - ext_len - 1]);
          if (l3 <= u3) step(X, &l3, &u3, v[e - ext_len - 2]);
          if (l3 <= u3) step(X, &l3, &u3, v[e - ext_len - 3]);
          absent3 = l3 > u3;
        }
        if (pair && win3) { /* the 19-base window read[e-19, e) is clean */
          const int present = l <= u && !absent3;
          if (KNOWN(e)) { /* the last probe answered this window */
            if (present) ++w->ext;
          } else if (probe && e >= 20 && v[e - 20] <= 3) { /* this task's pair byte(s) */
            /* with 22 clean codes (pair_probe 2) a second byte, of the core
             * read[e-21, e-3), answers the windows of e-2 and e-3 */
            const int dual = pair_probe >= 2 && e >= 22 && v[e - 21] <= 3 && v[e - 22] <= 3;
            w->pair += dual ? 2 : 1;
            ++w->pair_ops;
            known_e = (int64_t)e - 1;
            known_n = dual ? 3 : 1;
            known_m = window19(X, v, e - 1) ? 1u : 0u;
            if (dual) known_m |= (window19(X, v, e - 2) ? 2u : 0u) | (window19(X, v, e - 3) ? 4u : 0u);
            if (present) ++w->ext;
          } else {
            ++w->ext;
          }
          probe = !present;
        } else {
          ++w->ext;
        }
        if (absent3) { /* absent: as below, with k >= ext_len + 3 */
          continue;
        }
        if (l <= u) {
          have = 1;
          b = e - ext_len;
        } else if (miss_skip && k >= ext_len) {
          /* the clean window does not occur, so no extension of the task
           * reaches ext_len bases, let alone k: it can emit nothing, and what
           * else it leaves (its left end for the in-lane SMEM filter, its
           * size records) cannot change any output (see or_opts.miss_skip) */
          continue;
        }
      }
    }
    if (!have) {
      ++w->tasks;
      uint64_t iv = X->table ? X->table[p] : kmer13(X, p);
      l = (uint32_t)iv;
      u = (uint32_t)(iv >> 32);
      if (l > u) continue; /* 2145 */
      b = e - PRECALC;
    }
    int pruned = 0, reached = !tend; /* a jump that ends the task: the next base empties it */
    if (prune) {
      uint32_t s = u - l + 1;
      if (sat && s >= sat) s = sat; /* a saturated record never prunes */
      if (A[b] == s && s != sat) pruned = 1;
      A[b] = s;
    }
    /* direct-extension model: in a run, the text position of read[b] and
     * the last 64-base text chunk read */
    int run = 0;
    uint32_t tq = 0, ones = 0;
    uint64_t chunk = ~0ull;
    for (int64_t x = (int64_t)b - 1; !pruned && !tend && x >= 0; --x) { /* 2154-2257 */
      uint32_t c = v[x], l2 = l, u2 = u;
      if (direct_min && prune && !run && c <= 3 && l == u && ones >= direct_min && A[x] != 1) {
        run = 1; /* suffix-array read: the text position of the single row */
        ++w->direct;
        tq = or_sa_true(X, l);
        chunk = ~0ull;
      }
      if (run && c > 3) { /* an N ends the run: the row of read[b] is read back */
        run = 0;
        ++w->direct;
      }
      if (c <= 3) {
        if (!run) {
          count_blocks(X, l, u, w);
        } else if (tq > 0 && (uint64_t)((tq - 1) >> 6) != chunk) {
          chunk = (tq - 1) >> 6; /* text chunk holding text[tq - 1] */
          ++w->direct;
          ++w->text;
        }
      }
      step(X, &l2, &u2, c);
      ++w->steps;
      if (l2 > u2) { /* 2246 */
        reached = 0;
        break;
      }
      if (c <= 3) ones = l2 == u2 ? ones + 1 : 0;
      l = l2;
      u = u2;
      b = (uint32_t)x; /* 2252-2255 */
      if (run) --tq;
      if (prune) {
        uint32_t s = u - l + 1;
        if (sat && s >= sat) s = sat;
        if (A[b] == s && s != sat) pruned = 1;
        A[b] = s;
      }
    }
    if (pruned) continue;
    const int emit = e - b >= k && (!smem || b < last_b);
    if (run && emit) ++w->direct; /* the row where the run ended */
    last_b = b;
    if (e - b >= k) out[n++] = (cand_t){b, e, l, u}; /* 2259-2264 */
    if (smem_stop && reached) break; /* later tasks end at 0 too and fail the SMEM filter */
  }
#undef KNOWN
#undef KNOWN_PRESENT
  return n;
}

/* ---- per-read state ------------------------------------------------------ */

typedef struct {
  cand_t* F; /* forward entries in slot order: increasing qend (2128) */
  cand_t* R; /* reverse entries in slot order: increasing qbeg (2268) */
  uint32_t nF, nR;
} lists_t;

typedef struct {
  or_seed* seeds;
  uint32_t* rows; /* later overwritten in place by positions */
  uint32_t nseeds, nrows;
} rout_t;

typedef struct {
  const or_index* X;
  const char* bases;
  const uint64_t* off;
  const uint32_t* len;
  const or_opts* o;
  uint64_t r0;     /* first read of the batch */
  lists_t* lists;  /* per read of the batch   */
  rout_t* outs;
  work_t* work;    /* per thread              */
  uint32_t** scratch; /* per thread, >= 2*maxL cand + L words */
  uint64_t* scratch_cap;
  int err;
} batch_ctx;

static void phase_candidates(void* vctx, uint64_t i, int tid) {
  batch_ctx* C = (batch_ctx*)vctx;
  uint64_t r = C->r0 + i;
  uint32_t L = C->len[r];
  lists_t* Ls = &C->lists[i];
  memset(Ls, 0, sizeof(*Ls));
  if (L < (uint32_t)C->o->k) return;
  const uint8_t* s = (const uint8_t*)C->bases + C->off[r];
  uint64_t need = (uint64_t)L * 2 + 8;
  if (C->scratch_cap[tid] < need) {
    free(C->scratch[tid]);
    C->scratch[tid] = (uint32_t*)malloc(need * 4);
    C->scratch_cap[tid] = need;
  }
  uint8_t* v = (uint8_t*)C->scratch[tid];
  uint32_t* A = C->scratch[tid] + (L + 3) / 4 + 1;
  uint32_t ntask = L - C->o->k + 1;
  cand_t* tmp = (cand_t*)malloc(sizeof(cand_t) * ntask * 2);
  Ls->F = tmp;
  Ls->R = tmp + ntask;
  for (uint32_t q = 0; q < L; ++q) v[q] = (uint8_t)or_code_fwd(s[q]);
  const int stop = C->o->smem && C->o->smem_stop;
  const uint32_t ext = (uint32_t)C->o->ext_len;
  const uint32_t sat = C->o->prune_sat;
  const uint32_t dmin = C->o->direct_min;
  uint32_t nf = strand_candidates(C->X, v, L, C->o->k, C->o->prune, stop, ext, sat, dmin, C->o->miss_skip,
                                  C->o->pair_probe, C->o->smem, C->o, tmp + ntask, A, &C->work[tid]);
  /* forward: decreasing end -> slot order (increasing end) */
  for (uint32_t q = 0; q < nf; ++q) Ls->F[q] = tmp[ntask + nf - 1 - q];
  Ls->nF = nf;
  for (uint32_t q = 0; q < L; ++q) v[q] = (uint8_t)or_code_rev(s[L - 1 - q]);
  uint32_t nr = strand_candidates(C->X, v, L, C->o->k, C->o->prune, stop, ext, sat, dmin, C->o->miss_skip,
                                  C->o->pair_probe, C->o->smem, C->o, Ls->R, A, &C->work[tid]);
  for (uint32_t q = 0; q < nr; ++q) { /* virtual [b,e) -> reverse (L-e, L-b) with strand bits */
    cand_t c = Ls->R[q];
    Ls->R[q] = (cand_t){(L - c.e) | 0x80000000u, (L - c.b) | 0x80000000u, c.l, c.u};
  }
  Ls->nR = nr;
}

static inline uint32_t isize(cand_t c) { return c.u - c.l + 1; }

/* the batch's compacted entry array (3400-3401): read r's F then R, read order */
typedef struct {
  const lists_t* lists;
  uint64_t nreads;
} batch_view;

/* entry j of read r relative view: returns pointer to the entry at global
 * offset `g` of the batch array; used by the MEM look-ahead (624-667) */
typedef struct {
  uint64_t* start; /* per read: first global index */
  uint64_t total;
} batch_index;

static const cand_t* entry_at(const batch_view* B, const batch_index* I, uint64_t g, uint64_t hint_read) {
  /* linear search from the hint read (look-ahead distances are tiny) */
  uint64_t r = hint_read;
  while (r + 1 < B->nreads && I->start[r + 1] <= g) ++r;
  while (r > 0 && I->start[r] > g) --r;
  uint64_t o = g - I->start[r];
  const lists_t* Ls = &B->lists[r];
  return o < Ls->nF ? &Ls->F[o] : &Ls->R[o - Ls->nF];
}

/* filtered MEM interval of an entry: (1,0) if killed (1110-1139) */
static cand_t mem_filtered(const lists_t* Ls, int is_rev, uint32_t o) {
  if (!is_rev) {
    cand_t c = Ls->F[o];
    if (o + 1 < Ls->nF && Ls->F[o + 1].b == c.b && isize(Ls->F[o + 1]) == isize(c))
      c.l = 1, c.u = 0;
    return c;
  }
  cand_t c = Ls->R[o];
  if (o > 0 && Ls->R[o - 1].e == c.e && isize(Ls->R[o - 1]) == isize(c)) c.l = 1, c.u = 0;
  return c;
}

static cand_t filtered_at(const batch_view* B, const batch_index* I, uint64_t g, uint64_t hint) {
  uint64_t r = hint;
  while (r + 1 < B->nreads && I->start[r + 1] <= g) ++r;
  while (r > 0 && I->start[r] > g) --r;
  uint64_t o = g - I->start[r];
  const lists_t* Ls = &B->lists[r];
  return o < Ls->nF ? mem_filtered(Ls, 0, (uint32_t)o) : mem_filtered(Ls, 1, (uint32_t)(o - Ls->nF));
}

typedef struct {
  batch_ctx* C;
  batch_view B;
  batch_index I;
} phase2_ctx;

static void emit_seed(rout_t* ro, uint32_t* cap_seeds, uint32_t* cap_rows, or_seed sd,
                      const uint32_t* rows, uint32_t nrows) {
  if (ro->nseeds == *cap_seeds) {
    *cap_seeds = *cap_seeds ? 2 * *cap_seeds : 8;
    ro->seeds = (or_seed*)realloc(ro->seeds, sizeof(or_seed) * *cap_seeds);
  }
  ro->seeds[ro->nseeds++] = sd;
  if (ro->nrows + nrows > *cap_rows) {
    while (ro->nrows + nrows > *cap_rows) *cap_rows = *cap_rows ? 2 * *cap_rows : 64;
    ro->rows = (uint32_t*)realloc(ro->rows, sizeof(uint32_t) * *cap_rows);
  }
  if (nrows) memcpy(ro->rows + ro->nrows, rows, sizeof(uint32_t) * nrows);
  ro->nrows += nrows;
}

static void phase_output(void* vctx, uint64_t i, int tid) {
  phase2_ctx* P = (phase2_ctx*)vctx;
  batch_ctx* C = P->C;
  const or_opts* o = C->o;
  const lists_t* Ls = &C->lists[i];
  rout_t* ro = &C->outs[i];
  memset(ro, 0, sizeof(*ro));
  uint32_t cap_s = 0, cap_r = 0;
  uint32_t* rowbuf = NULL;
  uint64_t rowcap = 0;
  /* stable merge by qbeg, forward first on ties (the sort at 1233 / 3570) */
  uint32_t a = 0, z = 0;
  while (a < Ls->nF || z < Ls->nR) {
    int take_f = z >= Ls->nR || (a < Ls->nF && Ls->F[a].b <= (Ls->R[z].b & 0x7FFFFFFFu));
    cand_t c;
    uint32_t nrows = 0;
    if (o->smem) {
      /* 991-1018: forward keeps the last entry of each qbeg, reverse the first of each qend */
      int killed;
      if (take_f) {
        c = Ls->F[a];
        killed = a + 1 < Ls->nF && Ls->F[a + 1].b == c.b;
        ++a;
      } else {
        c = Ls->R[z];
        killed = z > 0 && Ls->R[z - 1].e == c.e;
        ++z;
      }
      if (killed) continue;
      uint32_t sz = isize(c);
      or_seed sd = {c.b, c.e, c.l, c.u, 0, 0};
      if (o->max_occ == 0 || sz <= o->max_occ) nrows = sz;
      if (rowcap < nrows) {
        rowcap = nrows;
        rowbuf = (uint32_t*)realloc(rowbuf, sizeof(uint32_t) * rowcap);
      }
      for (uint32_t q = 0; q < nrows; ++q) rowbuf[q] = c.l + q; /* 555-573 */
      sd.n_pos = nrows;
      emit_seed(ro, &cap_s, &cap_r, sd, rowbuf, nrows);
    } else {
      /* MEM: counts 927-989 (in-read), rows 624-667 (batch-wide look-ahead) */
      uint64_t g0 = P->I.start[i];
      int is_rev = !take_f;
      uint32_t oi = take_f ? a : z;
      uint64_t g = g0 + (is_rev ? Ls->nF + oi : oi);
      cand_t f = mem_filtered(Ls, is_rev, oi);
      c = is_rev ? Ls->R[oi] : Ls->F[oi];
      if (take_f)
        ++a;
      else
        ++z;
      int32_t nint = (int32_t)(f.u - f.l + 1);
      int32_t cnt = nint;
      if (nint > 0) {
        if (!is_rev) {
          for (uint32_t p = 1; oi + p < Ls->nF && Ls->F[oi + p].b == c.b; ++p) {
            cand_t nf = mem_filtered(Ls, 0, oi + p);
            int32_t nn = (int32_t)(nf.u - nf.l + 1);
            if (nn > 0) {
              cnt = nint - nn;
              break;
            }
          }
        } else {
          for (uint32_t p = 1; p <= oi && Ls->R[oi - p].e == c.e; ++p) {
            cand_t nf = mem_filtered(Ls, 1, oi - p);
            int32_t nn = (int32_t)(nf.u - nf.l + 1);
            if (nn > 0) {
              cnt = nint - nn;
              break;
            }
          }
        }
      }
      if (cnt <= 0) continue;
      /* next non-empty entry with the same qbeg (forward, scanning on) or the
       * same qend (reverse, scanning back), anywhere in the batch (637, 645) */
      uint32_t nx = 1, ny = 0;
      if (!is_rev) {
        for (uint64_t gg = g + 1; gg < P->I.total; ++gg) {
          const cand_t* ne = entry_at(&P->B, &P->I, gg, i);
          if (ne->b != c.b) break;
          cand_t nf = filtered_at(&P->B, &P->I, gg, i);
          nx = nf.l;
          ny = nf.u;
          if (nf.u - nf.l + 1 > 0) break;
        }
      } else {
        for (uint64_t gg = g; gg-- > 0;) {
          const cand_t* ne = entry_at(&P->B, &P->I, gg, i);
          if (ne->e != c.e) break;
          cand_t nf = filtered_at(&P->B, &P->I, gg, i);
          nx = nf.l;
          ny = nf.u;
          if (nf.u - nf.l + 1 > 0) break;
        }
      }
      or_seed sd = {c.b, c.e, c.l, c.u, 0, 0};
      uint32_t want = (uint32_t)cnt;
      if (o->max_occ != 0 && isize(c) > o->max_occ) want = 0;
      if (rowcap < want) {
        rowcap = want;
        rowbuf = (uint32_t*)realloc(rowbuf, sizeof(uint32_t) * rowcap);
      }
      uint32_t got = 0;
      for (uint32_t q = 0; got < want; ++q) {
        uint32_t row = c.l + q;
        if (row < nx || row > ny) rowbuf[got++] = row;
      }
      nrows = got;
      sd.n_pos = nrows;
      emit_seed(ro, &cap_s, &cap_r, sd, rowbuf, nrows);
    }
  }
  free(rowbuf);
  /* SA walk, in place (668-684) */
  work_t* w = &C->work[tid];
  for (uint32_t q = 0; q < ro->nrows; ++q) {
    ro->rows[q] = locate_count(C->X, ro->rows[q], &w->lf);
    ++w->sa;
  }
}

static void batch_free_lists(lists_t* lists, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) free(lists[i].F);
}

int or_seed_reads(const or_index* Xc, const char* bases, const uint64_t* offsets, const uint32_t* lens,
                  uint64_t n, const or_opts* o, or_result* res) {
  memset(res, 0, sizeof(*res));
  if (!Xc || !o || (n && (!bases || !offsets || !lens))) return -1;
  if (o->k < PRECALC + 1) return -1;
  or_index* X = (or_index*)Xc;
  int threads = o->threads;
  if (threads <= 0) {
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    threads = c > 0 ? (int)c : 1;
  }
  if (threads > 256) threads = 256;
  if (o->use_table && ensure_table(X, threads)) return -2;

  res->n_reads = n;
  res->read_nseeds = (uint32_t*)calloc(n + 1, 4);
  res->read_npos = (uint32_t*)calloc(n + 1, 4);
  res->read_seed_off = (uint64_t*)calloc(n + 1, 8);
  res->read_pos_off = (uint64_t*)calloc(n + 1, 8);
  uint64_t cap_seeds = 1024, cap_pos = 4096;
  res->seeds = (or_seed*)malloc(sizeof(or_seed) * cap_seeds);
  res->pos = (uint32_t*)malloc(sizeof(uint32_t) * cap_pos);

  work_t* work = (work_t*)calloc((size_t)threads, sizeof(work_t));
  uint32_t** scratch = (uint32_t**)calloc((size_t)threads, sizeof(uint32_t*));
  uint64_t* scap = (uint64_t*)calloc((size_t)threads, sizeof(uint64_t));

  uint64_t r0 = 0;
  while (r0 < n) {
    /* the reference's batch: reads until the running base count reaches 1e6 */
    uint64_t r1 = r0, bases_in = 0;
    while (r1 < n && bases_in < BATCH_BASES) bases_in += lens[r1++];
    uint64_t nb = r1 - r0;
    batch_ctx C = {X, bases, offsets, lens, o, r0, NULL, NULL, work, scratch, scap, 0};
    C.lists = (lists_t*)calloc(nb, sizeof(lists_t));
    C.outs = (rout_t*)calloc(nb, sizeof(rout_t));
    parallel_for(nb, threads, 16, phase_candidates, &C);
    for (uint64_t i = 0; i < nb; ++i)
      if (lens[r0 + i] < (uint32_t)o->k) res->n_short++;
    phase2_ctx P;
    P.C = &C;
    P.B.lists = C.lists;
    P.B.nreads = nb;
    P.I.start = (uint64_t*)malloc(sizeof(uint64_t) * (nb + 1));
    uint64_t tot = 0;
    for (uint64_t i = 0; i < nb; ++i) {
      P.I.start[i] = tot;
      tot += C.lists[i].nF + C.lists[i].nR;
    }
    P.I.start[nb] = tot;
    P.I.total = tot;
    parallel_for(nb, threads, 16, phase_output, &P);
    for (uint64_t i = 0; i < nb; ++i) {
      rout_t* ro = &C.outs[i];
      uint64_t r = r0 + i;
      res->read_seed_off[r] = res->n_seeds;
      res->read_pos_off[r] = res->n_pos;
      res->read_nseeds[r] = ro->nseeds;
      res->read_npos[r] = ro->nrows;
      while (res->n_seeds + ro->nseeds > cap_seeds) cap_seeds *= 2;
      while (res->n_pos + ro->nrows > cap_pos) cap_pos *= 2;
      res->seeds = (or_seed*)realloc(res->seeds, sizeof(or_seed) * cap_seeds);
      res->pos = (uint32_t*)realloc(res->pos, sizeof(uint32_t) * cap_pos);
      if (ro->nseeds) memcpy(res->seeds + res->n_seeds, ro->seeds, sizeof(or_seed) * ro->nseeds);
      if (ro->nrows) memcpy(res->pos + res->n_pos, ro->rows, sizeof(uint32_t) * ro->nrows);
      res->n_seeds += ro->nseeds;
      res->n_pos += ro->nrows;
      free(ro->seeds);
      free(ro->rows);
    }
    free(P.I.start);
    batch_free_lists(C.lists, nb);
    free(C.lists);
    free(C.outs);
    r0 = r1;
  }
  res->read_seed_off[n] = res->n_seeds;
  res->read_pos_off[n] = res->n_pos;
  for (int t = 0; t < threads; ++t) {
    res->w_tasks += work[t].tasks;
    res->w_ext += work[t].ext;
    res->w_steps += work[t].steps;
    res->w_blocks64 += work[t].blocks64;
    res->w_blocks96 += work[t].blocks96;
    res->w_lf += work[t].lf;
    res->w_sa += work[t].sa;
    res->w_direct += work[t].direct;
    res->w_text += work[t].text;
    res->w_pair += work[t].pair;
    res->w_pair_ops += work[t].pair_ops;
    res->w_tandem += work[t].tandem;
    res->w_tandem_bad += work[t].tandem_bad;
    res->w_stretches += work[t].stretches;
    free(scratch[t]);
  }
  if (!o->prune) res->w_steps_unpruned = res->w_steps;
  free(work);
  free(scratch);
  free(scap);
  return 0;
}

void or_result_free(or_result* r) {
  if (!r) return;
  free(r->read_nseeds);
  free(r->read_seed_off);
  free(r->read_npos);
  free(r->read_pos_off);
  free(r->seeds);
  free(r->pos);
  memset(r, 0, sizeof(*r));
}

/* ---- text output (3625-3648) ------------------------------------------- */

int or_format_text(const or_result* res, int smem, uint64_t first_read_no, FILE* out) {
  static const char sign[2] = {'+', '-'};
  for (uint64_t r = 0; r < res->n_reads; ++r) {
    fprintf(out,
            "/*===================================%s seeds in read no. %d (Read[begin, End] -> "
            "starting position(s) on the reference)===================================*/\n",
            smem ? "SMEM" : "MEM", (int)(first_read_no + r));
    int32_t pb = -1, pe = -1;
    const or_seed* sd = res->seeds + res->read_seed_off[r];
    const uint32_t* pos = res->pos + res->read_pos_off[r];
    for (uint32_t s = 0; s < res->read_nseeds[r]; ++s) {
      for (uint32_t q = 0; q < sd[s].n_pos; ++q) {
        int32_t x = (int32_t)sd[s].qbeg, y = (int32_t)sd[s].qend;
        if (pb == x && pe == y)
          fprintf(out, ", %c%u", sign[(uint32_t)y >> 31], *pos);
        else
          fprintf(out, "\nRead[%d, %d] -> %c%u", (int32_t)(sd[s].qbeg & 0x7FFFFFFFu),
                  (int32_t)(sd[s].qend & 0x7FFFFFFFu), sign[(uint32_t)y >> 31], *pos);
        pb = x;
        pe = y;
        ++pos;
      }
    }
    fprintf(out, "\n");
  }
  return 0;
}
