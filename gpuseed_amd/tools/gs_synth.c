/*
 * gs_synth - writes the synthetic reference and reads of a benchmark
 * configuration as FASTA.
 *
 *   gs_synth ref   -c CONFIG [-L LENGTH] [-S SEED] -o ref.fa
 *   gs_synth reads -c CONFIG -r ref.fa [-n NREADS] [-S SEED] -o reads.fa
 *
 * The reference is one record ">synthetic" with 80-column lines; reads are
 * single-line records ">r<i>" (the form seed_gen reads, seed_gen.cu:2955-2990).
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "gpuseed.h"

static uint8_t* read_fasta_codes(const char* fn, uint64_t* n_out) {
  FILE* fp = fopen(fn, "r");
  if (!fp) return NULL;
  size_t cap = 1 << 20, n = 0;
  uint8_t* c = (uint8_t*)malloc(cap);
  char* line = NULL;
  size_t lcap = 0;
  ssize_t nb;
  while ((nb = getline(&line, &lcap, fp)) >= 0) {
    if (line[0] == '>') continue;
    for (ssize_t i = 0; i < nb; ++i) {
      int v = -1;
      switch (line[i] | 0x20) {
        case 'a': v = 0; break;
        case 'c': v = 1; break;
        case 'g': v = 2; break;
        case 't': v = 3; break;
        default: break;
      }
      if (v < 0) continue;
      if (n == cap) {
        cap *= 2;
        c = (uint8_t*)realloc(c, cap);
      }
      c[n++] = (uint8_t)v;
    }
  }
  free(line);
  fclose(fp);
  *n_out = n;
  return c;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: gs_synth ref|reads -c CONFIG [-L LEN] [-n N] [-S SEED] [-r ref.fa] -o out.fa\n");
    return 1;
  }
  const char* what = argv[1];
  int config = 1, c;
  uint64_t len = 0, nreads = 0, seed = 0;
  int seed_set = 0;
  const char *out = NULL, *ref = NULL;
  optind = 2;
  while ((c = getopt(argc, argv, "c:L:n:S:o:r:")) >= 0) {
    switch (c) {
      case 'c': config = atoi(optarg); break;
      case 'L': len = strtoull(optarg, NULL, 10); break;
      case 'n': nreads = strtoull(optarg, NULL, 10); break;
      case 'S': seed = strtoull(optarg, NULL, 10); seed_set = 1; break;
      case 'o': out = optarg; break;
      case 'r': ref = optarg; break;
      default: return 1;
    }
  }
  gs_synth_ref_params rp;
  gs_synth_read_params qp;
  if (gs_synth_preset(config, &rp, &qp) != GS_OK || !out) {
    fprintf(stderr, "bad config or missing -o\n");
    return 1;
  }
  FILE* fo = fopen(out, "w");
  if (!fo) {
    perror(out);
    return 1;
  }
  static const char ACGT[4] = {'A', 'C', 'G', 'T'};
  if (!strcmp(what, "ref")) {
    if (len) rp.length = len;
    if (seed_set) rp.seed = seed;
    uint8_t* codes = (uint8_t*)malloc(rp.length);
    if (!codes || gs_synth_reference(&rp, codes, 0) != GS_OK) {
      fprintf(stderr, "generation failed\n");
      return 1;
    }
    fprintf(fo, ">synthetic config=%d length=%llu seed=%llu\n", config,
            (unsigned long long)rp.length, (unsigned long long)rp.seed);
    char lb[81];
    for (uint64_t i = 0; i < rp.length; i += 80) {
      uint64_t m = rp.length - i < 80 ? rp.length - i : 80;
      for (uint64_t j = 0; j < m; ++j) lb[j] = ACGT[codes[i + j]];
      lb[m] = '\n';
      fwrite(lb, 1, m + 1, fo);
    }
    free(codes);
  } else if (!strcmp(what, "reads")) {
    if (!ref) {
      fprintf(stderr, "reads needs -r ref.fa\n");
      return 1;
    }
    uint64_t n = 0;
    uint8_t* codes = read_fasta_codes(ref, &n);
    if (!codes || n < 32) {
      fprintf(stderr, "cannot read reference %s\n", ref);
      return 1;
    }
    if (nreads) qp.n_reads = nreads;
    if (seed_set) qp.seed = seed;
    uint32_t* lens = (uint32_t*)malloc(sizeof(uint32_t) * qp.n_reads);
    uint64_t* offs = (uint64_t*)malloc(sizeof(uint64_t) * (qp.n_reads + 1));
    gs_synth_read_lengths(&qp, lens);
    uint64_t tot = 0;
    for (uint64_t i = 0; i < qp.n_reads; ++i) {
      offs[i] = tot;
      tot += lens[i];
    }
    char* buf = (char*)malloc(tot ? tot : 1);
    if (gs_synth_reads(codes, n, &qp, offs, buf, NULL, NULL, 0) != GS_OK) {
      fprintf(stderr, "read generation failed\n");
      return 1;
    }
    for (uint64_t i = 0; i < qp.n_reads; ++i) {
      fprintf(fo, ">r%llu\n", (unsigned long long)(i + 1));
      fwrite(buf + offs[i], 1, lens[i], fo);
      fputc('\n', fo);
    }
    free(buf);
    free(lens);
    free(offs);
    free(codes);
  } else {
    fprintf(stderr, "unknown mode %s\n", what);
    return 1;
  }
  return fclose(fo) ? 1 : 0;
}
