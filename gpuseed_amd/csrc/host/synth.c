/*
 * Deterministic synthetic genomes and reads for the seeding workloads.
 *
 * The reference ships no data (its .gitignore excludes *.fa and *.fasta), so
 * every workload is generated here from a seed.  Output depends only on the
 * parameters and the seed, never on the thread count: the background is drawn
 * in fixed 1 Mbp blocks with one stream per block, repeat features are placed
 * from one sequential stream, and every read has its own stream.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "gs_host.h"

#define BLOCK (1u << 20)

static uint64_t stream_seed(uint64_t seed, uint64_t tag, uint64_t i) {
  uint64_t s = seed ^ (tag * 0xd1b54a32d192ed03ull) ^ (i * 0x9e3779b97f4a7c15ull);
  gs_splitmix64(&s);
  return s;
}

typedef struct {
  uint32_t tA, tC, tG; /* cumulative 16-bit thresholds */
} comp_t;

static comp_t make_comp(double gc) {
  comp_t c;
  double at = (1.0 - gc) * 0.5, cg = gc * 0.5;
  c.tA = (uint32_t)(at * 65536.0 + 0.5);
  c.tC = (uint32_t)((at + cg) * 65536.0 + 0.5);
  c.tG = (uint32_t)((at + 2 * cg) * 65536.0 + 0.5);
  return c;
}

static inline uint8_t draw_base(const comp_t* c, uint32_t r16) {
  return r16 < c->tA ? 0 : r16 < c->tC ? 1 : r16 < c->tG ? 2 : 3;
}

static void fill_iid(uint8_t* out, uint64_t len, const comp_t* c, uint64_t* s) {
  uint64_t i = 0;
  while (i < len) {
    uint64_t r = gs_splitmix64(s);
    for (int j = 0; j < 4 && i < len; ++j, ++i) out[i] = draw_base(c, (uint32_t)(r >> (16 * j)) & 0xFFFF);
  }
}

typedef struct {
  uint8_t* codes;
  uint64_t n, seed;
  comp_t comp;
  int tid, nthreads;
} bg_job;

static void* bg_worker(void* arg) {
  bg_job* j = (bg_job*)arg;
  uint64_t nblocks = (j->n + BLOCK - 1) / BLOCK;
  for (uint64_t b = (uint64_t)j->tid; b < nblocks; b += (uint64_t)j->nthreads) {
    uint64_t s = stream_seed(j->seed, 1, b);
    uint64_t beg = b * BLOCK, len = j->n - beg < BLOCK ? j->n - beg : BLOCK;
    fill_iid(j->codes + beg, len, &j->comp, &s);
  }
  return NULL;
}

static int pick_threads(int n_threads) {
  if (n_threads > 0) return n_threads > 256 ? 256 : n_threads;
  long c = sysconf(_SC_NPROCESSORS_ONLN);
  if (c < 1) c = 1;
  return c > 16 ? 16 : (int)c;
}

/* copy `len` bases of src into dst with per-base substitution probability div */
static void mutate_copy(uint8_t* dst, const uint8_t* src, uint64_t len, double div, uint64_t* s) {
  const uint64_t thr = (uint64_t)(div * 18446744073709551615.0);
  for (uint64_t i = 0; i < len; ++i) {
    uint8_t b = src[i];
    if (div > 0 && gs_splitmix64(s) < thr) b = (uint8_t)((b + 1 + gs_below(s, 3)) & 3);
    dst[i] = b;
  }
}

static uint32_t unif_u32(uint64_t* s, uint32_t lo, uint32_t hi) {
  return hi <= lo ? lo : lo + (uint32_t)gs_below(s, (uint64_t)(hi - lo) + 1);
}

int gs_synth_reference(const gs_synth_ref_params* p, uint8_t* codes, int n_threads) {
  if (!p || !codes || p->length == 0 || p->gc < 0 || p->gc > 1) return GS_EINVAL;
  const uint64_t n = p->length;
  comp_t comp = make_comp(p->gc);
  int nt = pick_threads(n_threads);
  if ((uint64_t)nt > (n + BLOCK - 1) / BLOCK) nt = (int)((n + BLOCK - 1) / BLOCK);
  pthread_t th[256];
  bg_job jobs[256];
  for (int t = 0; t < nt; ++t) {
    jobs[t] = (bg_job){codes, n, p->seed, comp, t, nt};
    if (pthread_create(&th[t], NULL, bg_worker, &jobs[t])) {
      for (int u = 0; u < t; ++u) pthread_join(th[u], NULL);
      return GS_ENOMEM;
    }
  }
  for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);

  uint64_t s = stream_seed(p->seed, 2, 0);
  /* interspersed repeat families */
  if (p->n_families > 0 && (p->copies_per_family > 0 || p->interspersed_fraction > 0)) {
    uint32_t fmax = p->family_len_max > p->family_len_min ? p->family_len_max : p->family_len_min;
    uint8_t* units = (uint8_t*)malloc((size_t)p->n_families * fmax);
    uint32_t* ulen = (uint32_t*)malloc(sizeof(uint32_t) * p->n_families);
    if (!units || !ulen) {
      free(units);
      free(ulen);
      return GS_ENOMEM;
    }
    for (uint32_t f = 0; f < p->n_families; ++f) {
      ulen[f] = unif_u32(&s, p->family_len_min, fmax);
      fill_iid(units + (size_t)f * fmax, ulen[f], &comp, &s);
    }
    uint64_t placed = 0, target = (uint64_t)(p->interspersed_fraction * (double)n);
    uint64_t copies = 0, want = (uint64_t)p->copies_per_family * p->n_families;
    for (;;) {
      if (p->copies_per_family > 0) {
        if (copies >= want) break;
      } else if (placed >= target) {
        break;
      }
      uint32_t f = p->copies_per_family > 0 ? (uint32_t)(copies / p->copies_per_family)
                                            : (uint32_t)gs_below(&s, p->n_families);
      uint64_t len = ulen[f];
      if (len > n) len = n;
      uint64_t pos = gs_below(&s, n - len + 1);
      double div = p->family_div_min + (p->family_div_max - p->family_div_min) * gs_unif(&s);
      mutate_copy(codes + pos, units + (size_t)f * fmax, len, div, &s);
      placed += len;
      ++copies;
    }
    free(units);
    free(ulen);
  }
  /* segmental duplications: copy a stretch of the current genome elsewhere */
  if (p->segdup_fraction > 0 && p->segdup_len_max > 0) {
    uint64_t placed = 0, target = (uint64_t)(p->segdup_fraction * (double)n);
    uint8_t* tmp = (uint8_t*)malloc(p->segdup_len_max > p->segdup_len_min ? p->segdup_len_max : p->segdup_len_min);
    if (!tmp) return GS_ENOMEM;
    while (placed < target) {
      uint64_t len = unif_u32(&s, p->segdup_len_min, p->segdup_len_max);
      if (len > n / 2) len = n / 2;
      if (len == 0) break;
      uint64_t src = gs_below(&s, n - len + 1), dst = gs_below(&s, n - len + 1);
      double div = p->segdup_div_min + (p->segdup_div_max - p->segdup_div_min) * gs_unif(&s);
      mutate_copy(tmp, codes + src, len, div, &s);
      memcpy(codes + dst, tmp, len);
      placed += len;
    }
    free(tmp);
  }
  /* exact tandem repeats */
  if (p->tandem_fraction > 0 && p->tandem_unit_max > 0) {
    uint64_t placed = 0, target = (uint64_t)(p->tandem_fraction * (double)n);
    uint8_t unit[256];
    while (placed < target) {
      uint32_t ul = unif_u32(&s, p->tandem_unit_min ? p->tandem_unit_min : 1, p->tandem_unit_max);
      if (ul > 256) ul = 256;
      uint32_t cp = unif_u32(&s, p->tandem_copies_min, p->tandem_copies_max);
      uint64_t len = (uint64_t)ul * cp;
      if (len == 0) break;
      if (len > n) len = n;
      fill_iid(unit, ul, &comp, &s);
      uint64_t pos = gs_below(&s, n - len + 1);
      for (uint64_t i = 0; i < len; ++i) codes[pos + i] = unit[i % ul];
      placed += len;
    }
  }
  return GS_OK;
}

/* ---- reads --------------------------------------------------------------- */

static const gs_read_class* pick_class(const gs_synth_read_params* p, uint64_t* s) {
  double tot = 0;
  for (uint32_t c = 0; c < p->n_classes; ++c) tot += p->classes[c].weight;
  double x = gs_unif(s) * tot;
  for (uint32_t c = 0; c + 1 < p->n_classes; ++c) {
    if (x < p->classes[c].weight) return &p->classes[c];
    x -= p->classes[c].weight;
  }
  return &p->classes[p->n_classes - 1];
}

static int check_read_params(const gs_synth_read_params* p) {
  if (!p || p->n_classes == 0 || p->n_classes > 4) return GS_EINVAL;
  for (uint32_t c = 0; c < p->n_classes; ++c)
    if (p->classes[c].len_min == 0 || p->classes[c].len_max < p->classes[c].len_min ||
        p->classes[c].weight < 0)
      return GS_EINVAL;
  return GS_OK;
}

int gs_synth_read_lengths(const gs_synth_read_params* p, uint32_t* lens) {
  int rc = check_read_params(p);
  if (rc) return rc;
  if (!lens) return GS_EINVAL;
  for (uint64_t i = 0; i < p->n_reads; ++i) {
    uint64_t s = stream_seed(p->seed, 3, i);
    const gs_read_class* c = pick_class(p, &s);
    lens[i] = unif_u32(&s, c->len_min, c->len_max);
  }
  return GS_OK;
}

typedef struct {
  const uint8_t* ref;
  uint64_t n;
  const gs_synth_read_params* p;
  const uint64_t* offsets;
  char* out;
  uint32_t* origin;
  uint8_t* strand;
  int tid, nthreads;
  int rc;
} rd_job;

static const char ACGT[4] = {'A', 'C', 'G', 'T'};

static void make_read(const rd_job* j, uint64_t i) {
  const gs_synth_read_params* p = j->p;
  uint64_t s = stream_seed(p->seed, 3, i);
  const gs_read_class* c = pick_class(p, &s);
  const uint32_t L = unif_u32(&s, c->len_min, c->len_max);
  char* o = j->out + j->offsets[i];
  /* room for deletions: at most 3 reference bases per output base */
  uint64_t span = (uint64_t)L + L / 4 + 16;
  if (span > j->n) span = j->n;
  uint64_t pos = gs_below(&s, j->n - span + 1);
  int rev = p->both_strands ? (int)(gs_splitmix64(&s) & 1) : 0;
  const uint64_t tsub = (uint64_t)(c->sub_rate * 18446744073709551615.0);
  const uint64_t tind = (uint64_t)(c->indel_rate * 18446744073709551615.0);
  uint64_t r = pos;
  uint32_t k = 0;
  while (k < L) {
    if (r >= j->n) r = pos; /* wrap inside the sampled span (only for tiny references) */
    uint64_t x = gs_splitmix64(&s);
    if (x < tind) {
      uint32_t il = 1 + (uint32_t)gs_below(&s, 3);
      if (gs_splitmix64(&s) & 1) {
        for (uint32_t q = 0; q < il && k < L; ++q) o[k++] = ACGT[gs_below(&s, 4)];
      } else {
        r += il;
      }
      continue;
    }
    uint8_t b = j->ref[r++];
    if (x - tind < tsub && x >= tind) b = (uint8_t)((b + 1 + gs_below(&s, 3)) & 3);
    o[k++] = ACGT[b];
  }
  if (rev) {
    for (uint32_t a = 0, z = L - 1; a < z; ++a, --z) {
      char t = o[a];
      o[a] = o[z];
      o[z] = t;
    }
    for (uint32_t a = 0; a < L; ++a)
      o[a] = o[a] == 'A' ? 'T' : o[a] == 'C' ? 'G' : o[a] == 'G' ? 'C' : 'A';
  }
  if (p->n_read_fraction > 0 && gs_unif(&s) < p->n_read_fraction) o[gs_below(&s, L)] = 'N';
  if (j->origin) j->origin[i] = (uint32_t)pos;
  if (j->strand) j->strand[i] = (uint8_t)rev;
}

static void* rd_worker(void* arg) {
  rd_job* j = (rd_job*)arg;
  const uint64_t chunk = 4096;
  for (uint64_t b = (uint64_t)j->tid * chunk; b < j->p->n_reads; b += (uint64_t)j->nthreads * chunk) {
    uint64_t e = b + chunk < j->p->n_reads ? b + chunk : j->p->n_reads;
    for (uint64_t i = b; i < e; ++i) make_read(j, i);
  }
  return NULL;
}

int gs_synth_reads(const uint8_t* ref, uint64_t n, const gs_synth_read_params* p, const uint64_t* offsets,
                   char* out, uint32_t* origin_pos, uint8_t* origin_strand, int n_threads) {
  int rc = check_read_params(p);
  if (rc) return rc;
  if (!ref || !offsets || !out || n < 32) return GS_EINVAL;
  int nt = pick_threads(n_threads);
  pthread_t th[256];
  rd_job jobs[256];
  for (int t = 0; t < nt; ++t) {
    jobs[t] = (rd_job){ref, n, p, offsets, out, origin_pos, origin_strand, t, nt, 0};
    if (pthread_create(&th[t], NULL, rd_worker, &jobs[t])) {
      for (int u = 0; u < t; ++u) pthread_join(th[u], NULL);
      return GS_ENOMEM;
    }
  }
  for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
  return GS_OK;
}

/* ---- presets for the benchmark configurations ----------------------------
 * 1: 1 Mbp, 20 x 300 bp families x 10 copies at 2 %; 10,000 x 100 bp reads,
 *    1 % substitutions, 0.1 % indels, 2 % of reads with one N.
 * 2: 100 Mbp, 5 % interspersed (30 families, 300-3000 bp, 5-15 %), 1 % tandem;
 *    1,000,000 x 150 bp, 0.5 % substitutions, 0.05 % indels.
 * 3: 3.1 Gbp, 45 % interspersed (50 families, 300-6000 bp, 5-20 %), 3 % tandem,
 *    5 % segmental duplications (1-50 kbp, 1-5 %); 10,000,000 x 150 bp reads.
 * 4: config-3 genome; 2,000,000 reads: 70 % 100-250 bp at 0.5 % error,
 *    25 % 1-5 kbp and 5 % 5-20 kbp at 8 % error of which 60 % indels.
 * 5: config-3 genome and config-4 read mix (request sizes are drawn by the
 *    serving harness). */
int gs_synth_preset(int config, gs_synth_ref_params* rp, gs_synth_read_params* qp) {
  if (config < 1 || config > 5 || !rp || !qp) return GS_EINVAL;
  memset(rp, 0, sizeof(*rp));
  memset(qp, 0, sizeof(*qp));
  rp->seed = 0;
  rp->gc = 0.41;
  qp->seed = 1 + (uint64_t)config;
  qp->both_strands = 1;
  qp->n_classes = 1;
  switch (config) {
    case 1:
      rp->length = 1000000;
      rp->n_families = 20;
      rp->family_len_min = rp->family_len_max = 300;
      rp->copies_per_family = 10;
      rp->family_div_min = rp->family_div_max = 0.02;
      qp->n_reads = 10000;
      qp->classes[0] = (gs_read_class){1.0, 100, 100, 0.01, 0.001};
      qp->n_read_fraction = 0.02;
      break;
    case 2:
      rp->length = 100000000;
      rp->n_families = 30;
      rp->family_len_min = 300;
      rp->family_len_max = 3000;
      rp->interspersed_fraction = 0.05;
      rp->family_div_min = 0.05;
      rp->family_div_max = 0.15;
      rp->tandem_fraction = 0.01;
      rp->tandem_unit_min = 1;
      rp->tandem_unit_max = 6;
      rp->tandem_copies_min = 20;
      rp->tandem_copies_max = 200;
      qp->n_reads = 1000000;
      qp->classes[0] = (gs_read_class){1.0, 150, 150, 0.005, 0.0005};
      break;
    default:
      rp->length = 3100000000ull;
      rp->n_families = 50;
      rp->family_len_min = 300;
      rp->family_len_max = 6000;
      rp->interspersed_fraction = 0.45;
      rp->family_div_min = 0.05;
      rp->family_div_max = 0.20;
      rp->tandem_fraction = 0.03;
      rp->tandem_unit_min = 1;
      rp->tandem_unit_max = 6;
      rp->tandem_copies_min = 20;
      rp->tandem_copies_max = 200;
      rp->segdup_fraction = 0.05;
      rp->segdup_len_min = 1000;
      rp->segdup_len_max = 50000;
      rp->segdup_div_min = 0.01;
      rp->segdup_div_max = 0.05;
      if (config == 3) {
        qp->n_reads = 10000000;
        qp->classes[0] = (gs_read_class){1.0, 150, 150, 0.005, 0.0005};
      } else {
        qp->n_reads = 2000000;
        qp->n_classes = 3;
        qp->classes[0] = (gs_read_class){0.70, 100, 250, 0.0045, 0.0005};
        qp->classes[1] = (gs_read_class){0.25, 1000, 5000, 0.032, 0.048};
        qp->classes[2] = (gs_read_class){0.05, 5000, 20000, 0.032, 0.048};
      }
      break;
  }
  return GS_OK;
}
