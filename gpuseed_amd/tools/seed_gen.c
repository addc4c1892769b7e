/*
 * seed_gen - command-line drop-in for GPUseed's seed_gen (seed_gen.cu:
 * 2817-3790), running on an MI355X through libgpuseed's C ABI.
 *
 *   seed_gen [-k INT] [-s] [-o] [long options] <index-prefix> <reads.fa>
 *
 * Same flags as the reference (getopt "k:so", seed_gen.cu:2825-2839):
 * -k minimum seed length (atoi, default 20), -s SMEM mode (default MEM),
 * -o print the seeds to stdout.  Same inputs (<prefix>.bwt, <prefix>.sa;
 * reads one per line, '>' lines skipped, seed_gen.cu:2955-2990) and the same
 * stdout text (3625-3648).  stderr carries progress logs only.
 *
 * Long options (no counterpart in the reference):
 *   --kernel N         0 plain per-read kernel, 1 batched persistent kernel
 *                      with 8-bit size records, 2 with 16-bit ones, 3 (default)
 *                      with 4-bit ones (more waves per CU)
 *   --max-occ N        seeds with more than N occurrences get no positions
 *                      (default 0: every occurrence, as the reference)
 *   --device-bases N   bases uploaded per device batch (default 256,000,000);
 *                      device batches hold whole reference batches, so MEM
 *                      output does not depend on this value
 *   --device N         HIP device ordinal (default 0)
 *   --reference-capacity
 *                      stop like the reference when a 1,000,000-base batch has
 *                      more positions than its fixed output buffer holds
 *                      (seed_gen.cu:3553-3558); off by default, the library
 *                      has no such limit
 *   --kmer-len N       13 (default): the reference's 13-mer table only;
 *                      14..16: also an N-mer interval table with presence
 *                      bits (16 * 4^N bytes of HBM) that answers a task's
 *                      first N bases with one lookup; the output is the same
 *   --direct           also the full suffix array, its inverse and the text
 *                      (8 bytes per base plus 2 bits): the batched kernels
 *                      then compare single-row matches with the text 64
 *                      bases per load; the output is the same
 *   --pairs            also the pair table (4^18 bytes: which bases follow and
 *                      precede every 18-base pattern): with --kmer-len 16
 *                      and -k >= 19, one lookup answers two adjacent 19-base
 *                      windows of a read strand that does not match; the
 *                      output is the same
 *   --autotune         pick the batched kernel's launch parameters by timing
 *                      them on the first device batch (same output)
 *
 * Exit status 0 on success, 1 on any error (the reference's exit(1) and
 * exit(EXIT_FAILURE) cases: missing or inconsistent index files, output
 * overflow under --reference-capacity).  Differences: -k below 14 is refused
 * (the 13-base table lookup makes shorter seeds a different procedure,
 * seed_gen.cu:2132-2145, 2259); a missing reads file is an error instead of a
 * crash; the empty trailing batch the reference runs when the input ends
 * exactly at a batch boundary is not run.
 *
 * Three threads: a reader cuts the input into device batches, the main thread
 * seeds them, a writer formats the text; each hand-off holds at most two
 * batches.
 */
#define _GNU_SOURCE
#include <getopt.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include "gpuseed.h"

#define BASE_BATCH 1000000u /* the reference's batch size in bases (seed_gen.cu:2814) */

static double now(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

static void* xrealloc(void* p, size_t n) {
  void* q = realloc(p, n ? n : 1);
  if (!q) {
    fprintf(stderr, "seed_gen: out of host memory\n");
    exit(1);
  }
  return q;
}

/* ---- a device batch: whole reference batches of reads ------------------- */
typedef struct chunk {
  char* bases;
  size_t nbytes, bcap;
  uint64_t* offs;
  uint32_t* lens;
  uint64_t n, ncap;
  uint64_t first_read; /* 0-based number of the first read in the input       */
  uint64_t* rb_end;    /* reads [rb_end[i-1], rb_end[i]) form reference batch i */
  uint64_t* rb_bases;  /* bases of reference batch i                           */
  size_t nrb, rbcap;
  gs_result res;
  int rc;
  struct chunk* next;
} chunk;

static void chunk_free(chunk* c) {
  if (!c) return;
  free(c->bases);
  free(c->offs);
  free(c->lens);
  free(c->rb_end);
  free(c->rb_bases);
  gs_result_free(&c->res);
  free(c);
}

/* ---- bounded FIFO between two threads ----------------------------------- */
typedef struct {
  pthread_mutex_t mu;
  pthread_cond_t cv;
  chunk *head, *tail;
  int count, cap, closed;
} fifo;

static void fifo_init(fifo* q, int cap) {
  pthread_mutex_init(&q->mu, NULL);
  pthread_cond_init(&q->cv, NULL);
  q->head = q->tail = NULL;
  q->count = 0;
  q->cap = cap;
  q->closed = 0;
}

static void fifo_put(fifo* q, chunk* c) {
  pthread_mutex_lock(&q->mu);
  while (q->count >= q->cap) pthread_cond_wait(&q->cv, &q->mu);
  c->next = NULL;
  if (q->tail)
    q->tail->next = c;
  else
    q->head = c;
  q->tail = c;
  q->count++;
  pthread_cond_broadcast(&q->cv);
  pthread_mutex_unlock(&q->mu);
}

static void fifo_close(fifo* q) {
  pthread_mutex_lock(&q->mu);
  q->closed = 1;
  pthread_cond_broadcast(&q->cv);
  pthread_mutex_unlock(&q->mu);
}

/* NULL once the queue is closed and drained */
static chunk* fifo_get(fifo* q) {
  pthread_mutex_lock(&q->mu);
  while (!q->head && !q->closed) pthread_cond_wait(&q->cv, &q->mu);
  chunk* c = q->head;
  if (c) {
    q->head = c->next;
    if (!q->head) q->tail = NULL;
    q->count--;
    pthread_cond_broadcast(&q->cv);
  }
  pthread_mutex_unlock(&q->mu);
  return c;
}

/* ---- reader ---------------------------------------------------------------- */
typedef struct {
  FILE* fp;
  uint64_t device_bases;
  fifo* out;
  uint64_t n_reads, n_bases;
  double t_read;
} reader_args;

static void* reader_main(void* a) {
  reader_args* R = (reader_args*)a;
  char* line = NULL;
  size_t lcap = 0;
  uint64_t read_no = 0;
  int eof = 0;
  while (!eof) {
    chunk* c = (chunk*)calloc(1, sizeof(chunk));
    if (!c) {
      fprintf(stderr, "seed_gen: out of host memory\n");
      exit(1);
    }
    c->first_read = read_no;
    double t0 = now();
    /* whole reference batches until the device batch is full (2955-2991) */
    while (!eof && c->nbytes < R->device_bases) {
      uint64_t batch_bases = 0;
      while (batch_bases < BASE_BATCH) {
        ssize_t nb = getline(&line, &lcap, R->fp);
        if (nb < 0) {
          eof = 1;
          break;
        }
        if (line[0] == '>') continue;
        /* one read per line, its last byte dropped (the newline, 2967-2970) */
        const size_t L = (size_t)nb - 1;
        if (c->nbytes + L > c->bcap) {
          c->bcap = (c->nbytes + L) * 2 + (1u << 20);
          c->bases = (char*)xrealloc(c->bases, c->bcap);
        }
        if (c->n == c->ncap) {
          c->ncap = c->ncap ? c->ncap * 2 : 65536;
          c->offs = (uint64_t*)xrealloc(c->offs, c->ncap * 8);
          c->lens = (uint32_t*)xrealloc(c->lens, c->ncap * 4);
        }
        memcpy(c->bases + c->nbytes, line, L);
        c->offs[c->n] = c->nbytes;
        c->lens[c->n] = (uint32_t)L;
        c->n++;
        c->nbytes += L;
        batch_bases += L;
      }
      if (c->n == (c->nrb ? c->rb_end[c->nrb - 1] : 0)) break; /* nothing read */
      if (c->nrb == c->rbcap) {
        c->rbcap = c->rbcap ? c->rbcap * 2 : 64;
        c->rb_end = (uint64_t*)xrealloc(c->rb_end, c->rbcap * 8);
        c->rb_bases = (uint64_t*)xrealloc(c->rb_bases, c->rbcap * 8);
      }
      c->rb_end[c->nrb] = c->n;
      c->rb_bases[c->nrb] = batch_bases;
      c->nrb++;
    }
    R->t_read += now() - t0;
    if (!c->n) {
      chunk_free(c);
      break;
    }
    read_no += c->n;
    R->n_reads += c->n;
    R->n_bases += c->nbytes;
    fifo_put(R->out, c);
  }
  free(line);
  fifo_close(R->out);
  return NULL;
}

/* ---- writer ---------------------------------------------------------------- */
typedef struct {
  fifo* in;
  int print, smem, k, reference_capacity;
  double t_write;
  uint64_t n_pos;
  volatile int failed; /* set once: later batches are dropped unprinted */
} writer_args;

static void* writer_main(void* a) {
  writer_args* W = (writer_args*)a;
  chunk* c;
  while ((c = fifo_get(W->in))) {
    if (W->failed) {
      chunk_free(c);
      continue;
    }
    double t0 = now();
    gs_result view = c->res;
    int overflow = 0;
    uint64_t over_count = 0;
    int64_t over_cap = 0;
    if (W->reference_capacity) {
      /* the reference's output buffer: 3 * 2 * 2 * (bases rounded up to 8 -
         (k - 1) * reads) positions per batch, compared as unsigned (3555) */
      uint64_t r0 = 0;
      for (size_t b = 0; b < c->nrb; ++b) {
        const uint64_t r1 = c->rb_end[b];
        const uint64_t bs8 = (c->rb_bases[b] + 7) / 8 * 8;
        const int32_t cap = (int32_t)(12 * ((int64_t)bs8 - (int64_t)(W->k - 1) * (int64_t)(r1 - r0)));
        const uint64_t np = c->res.read_pos_off[r1] - c->res.read_pos_off[r0];
        if ((uint32_t)np > (uint32_t)cap) {
          overflow = 1;
          over_count = np;
          over_cap = cap;
          view.n_reads = r0;
          break;
        }
        r0 = r1;
      }
    }
    if (W->print && view.n_reads) {
      fflush(stdout);
      int rc = gs_format_text(&view, W->smem, c->first_read + 1, STDOUT_FILENO);
      if (rc) {
        fprintf(stderr, "seed_gen: writing the output failed (%d)\n", rc);
        W->failed = 1;
      }
    }
    W->n_pos += c->res.n_pos;
    W->t_write += now() - t0;
    if (overflow) {
      fprintf(stderr, "n_seeds_sum_fow_rev (%llu) is more than allocated size(%d)\n",
              (unsigned long long)over_count, (int)over_cap);
      W->failed = 1;
    }
    chunk_free(c);
  }
  return NULL;
}

/* ---- index files: the reference's checks and messages (2685-2736) --------- */
/* The reference's checks of the index headers (seed_gen.cu:2692-2723), with
 * its messages and exit statuses.  Headers in upstream bwa 0.7's 64-bit
 * layout (whose 64-bit fields agree) are left to gs_index_open, which checks
 * them completely. */
static void check_index_files(const char* prefix) {
  size_t n = strlen(prefix) + 8;
  char* fn = (char*)malloc(n);
  uint32_t bwt_head[10], sa_head[14];
  memset(bwt_head, 0, sizeof bwt_head);
  memset(sa_head, 0, sizeof sa_head);
  snprintf(fn, n, "%s.bwt", prefix);
  FILE* fp = fopen(fn, "rb");
  if (!fp) {
    fprintf(stderr, "Unable to open .bwt file.");
    exit(1);
  }
  if (fread(bwt_head, 4, 10, fp) < 5) {
    fprintf(stderr, "seed_gen: %s is truncated\n", fn);
    exit(1);
  }
  fclose(fp);
  snprintf(fn, n, "%s.sa", prefix);
  fp = fopen(fn, "rb");
  if (!fp) {
    fprintf(stderr, "Unable to open .sa file.");
    exit(1);
  }
  if (fread(sa_head, 4, 14, fp) < 7) {
    fprintf(stderr, "seed_gen: %s is truncated\n", fn);
    exit(1);
  }
  fclose(fp);
  free(fn);
  uint64_t b64[5], s64[7];
  memcpy(b64, bwt_head, sizeof b64);
  memcpy(s64, sa_head, sizeof s64);
  if (b64[4] && s64[0] == b64[0] && s64[6] == b64[4] && s64[5] && b64[0] <= b64[4] && b64[1] <= b64[2] &&
      b64[2] <= b64[3] && b64[3] <= b64[4])
    return; /* bwa 0.7 layout */
  if (sa_head[0] != bwt_head[0]) {
    fprintf(stderr, "SA-BWT inconsistency: primary is not the same.\n");
    exit(EXIT_FAILURE);
  }
  if (sa_head[6] != bwt_head[4]) {
    fprintf(stderr, "SA-BWT inconsistency: seq_len is not the same.\n");
    exit(EXIT_FAILURE);
  }
}

static const char* errname(int rc) {
  switch (rc) {
    case GS_EINVAL: return "invalid argument";
    case GS_ENOMEM: return "out of memory";
    case GS_EHIP: return "HIP runtime error";
    case GS_EIO: return "index file missing or unreadable";
    case GS_EFORMAT: return "index files inconsistent";
    case GS_ENODEV: return "no gfx950 device";
    case GS_ERANGE: return "input beyond a limit";
    default: return "error";
  }
}

int main(int argc, char** argv) {
  double t_start = now();
  gs_opts o;
  memset(&o, 0, sizeof o);
  o.k = 20;
  o.kernel = 3;
  int print = 0, device = 0, reference_capacity = 0, kmer_len = 13, autotune = 0, direct = 0, pairs = 0, c, li;
  uint64_t device_bases = 256000000ull;
  static const struct option longopts[] = {{"kernel", required_argument, 0, 1},
                                           {"max-occ", required_argument, 0, 2},
                                           {"device-bases", required_argument, 0, 3},
                                           {"device", required_argument, 0, 4},
                                           {"reference-capacity", no_argument, 0, 5},
                                           {"kmer-len", required_argument, 0, 6},
                                           {"autotune", no_argument, 0, 7},
                                           {"direct", no_argument, 0, 8},
                                           {"pairs", no_argument, 0, 9},
                                           {0, 0, 0, 0}};
  while ((c = getopt_long(argc, argv, "k:so", longopts, &li)) >= 0) {
    switch (c) {
      case 'k': o.k = atoi(optarg); break;
      case 's': o.smem = 1; break;
      case 'o': print = 1; break;
      case 1: o.kernel = atoi(optarg); break;
      case 2: o.max_occ = (uint32_t)strtoul(optarg, NULL, 10); break;
      case 3: device_bases = strtoull(optarg, NULL, 10); break;
      case 4: device = atoi(optarg); break;
      case 5: reference_capacity = 1; break;
      case 6: kmer_len = atoi(optarg); break;
      case 7: autotune = 1; break;
      case 8: direct = 1; break;
      case 9: pairs = 1; break;
      default: break; /* like the reference: getopt's message, then go on */
    }
  }
  if (optind + 2 > argc) {
    fprintf(stderr, "Usage: seed_gen [-k INT] [-s] [-o] [--kernel N] [--max-occ N] [--device-bases N] "
                    "[--device N] [--reference-capacity] [--kmer-len 13..16] [--direct] [--pairs] [--autotune] "
                    "<index-prefix> <reads.fa>\n");
    return 1;
  }
  if (o.k < 14) {
    fprintf(stderr, "seed_gen: -k must be at least 14 (seeds start from a 13-base table lookup)\n");
    return 1;
  }
  if (o.kernel < 0 || o.kernel > 3) {
    fprintf(stderr, "seed_gen: --kernel must be 0, 1, 2 or 3\n");
    return 1;
  }
  if (kmer_len < 13 || kmer_len > 16) {
    fprintf(stderr, "seed_gen: --kmer-len must be 13, 14, 15 or 16\n");
    return 1;
  }
  if (device_bases < 1) device_bases = 1;
  const char* prefix = argv[optind];
  fprintf(stderr, "Loading index from file and copying it to GPU...\n");
  check_index_files(prefix);
  FILE* fp = fopen(argv[optind + 1], "r");
  if (!fp) {
    fprintf(stderr, "seed_gen: cannot open %s\n", argv[optind + 1]);
    return 1;
  }
  setvbuf(fp, NULL, _IOFBF, 1 << 22);
  gs_index* idx = NULL;
  double t0 = now();
  int rc = gs_index_open(prefix, device, &idx);
  if (rc) {
    fprintf(stderr, "seed_gen: loading the index failed: %s (%d)\n", errname(rc), rc);
    return 1;
  }
  gs_index_info info;
  gs_index_get_info(idx, &info);
  fprintf(stderr, "primary=%u seq_len=%u sa_intv=%u L2=[%u,%u,%u,%u,%u]\n", info.primary, info.seq_len,
          info.sa_intv, info.L2[0], info.L2[1], info.L2[2], info.L2[3], info.L2[4]);
  fprintf(stderr, "Index on the GPU in %.3f seconds (read %.3f, upload %.3f, repack %.3f, 13-mer table %.3f), %.2f GB\n",
          now() - t0, info.t_read, info.t_upload, info.t_repack, info.t_table, info.device_bytes / 1e9);
  if (kmer_len > 13) {
    rc = gs_index_set_kmer_len(idx, kmer_len);
    if (rc) {
      fprintf(stderr, "seed_gen: building the %d-mer table failed: %s (%d)\n", kmer_len, errname(rc), rc);
      return 1;
    }
    gs_index_get_info(idx, &info);
    fprintf(stderr, "%d-mer interval table in %.3f seconds, %.2f GB on the GPU\n", kmer_len, info.t_ext,
            info.device_bytes / 1e9);
  }
  if (direct) {
    rc = gs_index_set_direct(idx, 1);
    if (rc) {
      fprintf(stderr, "seed_gen: building the direct-extension structures failed: %s (%d)\n", errname(rc), rc);
      return 1;
    }
    gs_index_get_info(idx, &info);
    fprintf(stderr, "direct-extension structures in %.3f seconds, %.2f GB on the GPU\n", info.t_direct,
            info.device_bytes / 1e9);
  }
  if (pairs) {
    rc = gs_index_set_pairs(idx, 1);
    if (rc) {
      fprintf(stderr, "seed_gen: building the pair table failed: %s (%d)\n", errname(rc), rc);
      return 1;
    }
    gs_index_get_info(idx, &info);
    fprintf(stderr, "pair table in %.3f seconds, %.2f GB on the GPU\n", info.t_pairs, info.device_bytes / 1e9);
  }

  const double t_ready = now();  /* index and its structures on the GPU */
  fifo q_in, q_out;
  fifo_init(&q_in, 2);
  fifo_init(&q_out, 2);
  reader_args R = {fp, device_bases, &q_in, 0, 0, 0.0};
  writer_args W = {&q_out, print, o.smem, o.k, reference_capacity, 0.0, 0, 0};
  pthread_t tr, tw;
  if (pthread_create(&tr, NULL, reader_main, &R) || pthread_create(&tw, NULL, writer_main, &W)) {
    fprintf(stderr, "seed_gen: cannot start threads\n");
    return 1;
  }
  double t_gpu = 0.0, t_kernels = 0.0;
  uint64_t n_batches = 0;
  chunk* ck;
  int seed_rc = 0;
  while ((ck = fifo_get(&q_in))) {
    if (W.failed || seed_rc) { /* drain the reader after a failure */
      chunk_free(ck);
      continue;
    }
    double t1 = now();
    if (autotune && o.kernel != 0) { /* launch parameters from the first device batch */
      gs_tune_result tr;
      rc = gs_autotune(idx, ck->bases, ck->offs, ck->lens, ck->n, &o, 0, &tr);
      if (rc) {
        fprintf(stderr, "seed_gen: autotuning failed: %s (%d)\n", errname(rc), rc);
        seed_rc = rc;
        chunk_free(ck);
        continue;
      }
      o = tr.opts;
      fprintf(stderr, "autotune: kernel %d, wg %u, seg_tasks %u, walk_rows %u: %.3f s against %.3f s (%d tried)\n",
              o.kernel, (unsigned)o.wg, (unsigned)o.seg_tasks, (unsigned)o.walk_rows, tr.seconds,
              tr.default_seconds, tr.n_tried);
      autotune = 0;
      t1 = now();
    }
    rc = gs_seed(idx, ck->bases, ck->offs, ck->lens, ck->n, &o, &ck->res);
    if (rc) {
      fprintf(stderr, "seed_gen: seeding failed: %s (%d)\n", errname(rc), rc);
      seed_rc = rc;
      chunk_free(ck);
      continue;
    }
    t_gpu += now() - t1;
    t_kernels += ck->res.t_seed + ck->res.t_post + ck->res.t_locate;
    ++n_batches;
    fprintf(stderr, "Batch of %llu reads (%llu bases): %llu seeds, %llu positions, %.3f s on the GPU\n",
            (unsigned long long)ck->n, (unsigned long long)ck->nbytes, (unsigned long long)ck->res.n_seeds,
            (unsigned long long)ck->res.n_pos, now() - t1);
    fifo_put(&q_out, ck);
  }
  fifo_close(&q_out);
  pthread_join(tr, NULL);
  pthread_join(tw, NULL);
  fclose(fp);
  const double t_end = now(), t_total = t_end - t_start, t_seeding = t_end - t_ready;
  gs_index_close(idx);
  if (seed_rc || W.failed) return EXIT_FAILURE;
  fprintf(stderr,
          "%s seeds: %llu reads, %llu bases, %llu device batches, %llu positions\n"
          "time: total %.3f s, input %.3f s, device %.3f s (kernels %.3f s), output %.3f s; %.1f reads/s\n"
          "index ready after %.3f s; seeding %.3f s (%.1f reads/s)\n",
          o.smem ? "SMEM" : "MEM", (unsigned long long)R.n_reads, (unsigned long long)R.n_bases,
          (unsigned long long)n_batches, (unsigned long long)W.n_pos, t_total, R.t_read, t_gpu, t_kernels,
          W.t_write, t_total > 0 ? R.n_reads / t_total : 0.0, t_ready - t_start, t_seeding,
          t_seeding > 0 ? R.n_reads / t_seeding : 0.0);
  return 0;
}
