/*
 * oracle_seed - the oracle behind a seed_gen-like command line.
 * TEST INFRASTRUCTURE ONLY (parity checker / CPU baseline).
 *
 *   oracle_seed [-k INT] [-s] [-o] [-t THREADS] [-p 0|1] [-m MAX_OCC] [-T] <prefix> <reads.fa>
 *
 * Reads are taken the way seed_gen.cu:2955-2990 takes them: lines starting
 * with '>' are skipped and every other line is one read of its bytes minus the
 * last one.  -T builds the 13-mer table first (like 2887-2895).
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include "oracle.h"

static double now(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

int main(int argc, char** argv) {
  or_opts o = {20, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int print = 0, c;
  while ((c = getopt(argc, argv, "k:sot:p:m:T")) >= 0) {
    switch (c) {
      case 'k': o.k = atoi(optarg); break;
      case 's': o.smem = 1; break;
      case 'o': print = 1; break;
      case 't': o.threads = atoi(optarg); break;
      case 'p': o.prune = atoi(optarg); break;
      case 'm': o.max_occ = (uint32_t)strtoul(optarg, NULL, 10); break;
      case 'T': o.use_table = 1; break;
      default: break;
    }
  }
  if (optind + 2 > argc) {
    fprintf(stderr, "usage: oracle_seed [-k INT] [-s] [-o] [-t N] [-p 0|1] [-m MAX] [-T] prefix reads.fa\n");
    return 1;
  }
  or_index* X = NULL;
  int rc = or_index_load(argv[optind], &X);
  if (rc) {
    fprintf(stderr, "index load failed (%d)\n", rc);
    return 1;
  }
  FILE* fp = fopen(argv[optind + 1], "r");
  if (!fp) {
    fprintf(stderr, "cannot open %s\n", argv[optind + 1]);
    return 1;
  }
  size_t cap = 1 << 20, used = 0, ncap = 1024, n = 0;
  char* buf = (char*)malloc(cap);
  uint64_t* off = (uint64_t*)malloc(ncap * 8);
  uint32_t* len = (uint32_t*)malloc(ncap * 4);
  char* line = NULL;
  size_t lcap = 0;
  ssize_t nb;
  while ((nb = getline(&line, &lcap, fp)) >= 0) {
    if (line[0] == '>') continue;
    size_t L = (size_t)(nb - 1);
    while (used + L > cap) cap *= 2;
    buf = (char*)realloc(buf, cap);
    if (n == ncap) {
      ncap *= 2;
      off = (uint64_t*)realloc(off, ncap * 8);
      len = (uint32_t*)realloc(len, ncap * 4);
    }
    memcpy(buf + used, line, L);
    off[n] = used;
    len[n] = (uint32_t)L;
    used += L;
    ++n;
  }
  free(line);
  fclose(fp);
  double t0 = now();
  or_result res;
  if (o.use_table) {
    /* build the table outside the timed region (the reference times it apart, 2888-2895) */
    or_result tmp;
    or_seed_reads(X, buf, off, len, 0, &o, &tmp);
    or_result_free(&tmp);
    t0 = now();
  }
  rc = or_seed_reads(X, buf, off, len, n, &o, &res);
  double t1 = now();
  if (rc) {
    fprintf(stderr, "seeding failed (%d)\n", rc);
    return 1;
  }
  fprintf(stderr,
          "oracle: %zu reads, %.3f s, %.1f reads/s, seeds=%lu pos=%lu tasks=%lu steps=%lu "
          "blocks64=%lu blocks96=%lu lf=%lu sa=%lu short=%lu\n",
          n, t1 - t0, n / (t1 - t0), (unsigned long)res.n_seeds, (unsigned long)res.n_pos,
          (unsigned long)res.w_tasks, (unsigned long)res.w_steps, (unsigned long)res.w_blocks64,
          (unsigned long)res.w_blocks96, (unsigned long)res.w_lf, (unsigned long)res.w_sa,
          (unsigned long)res.n_short);
  if (print) or_format_text(&res, o.smem, 1, stdout);
  or_result_free(&res);
  or_index_free(X);
  free(buf);
  free(off);
  free(len);
  return 0;
}
