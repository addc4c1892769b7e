/*
 * gs_build_index - builds <prefix>.bwt and <prefix>.sa from a FASTA file.
 *
 *   gs_build_index [-r RATIO] [-p PREFIX] ref.fa
 *
 * Same command line as build_index.sh (build_index.sh:24-50: -r is the
 * suffix-array sampling interval, default 7; -p defaults to the FASTA name),
 * same output files (build_index.sh:63-75), without the bwa fork: all records
 * are concatenated, A/C/G/T (either case) map to 0..3 and any other letter is
 * replaced by a pseudo-random base drawn from a fixed-seed stream.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "gpuseed.h"

int main(int argc, char** argv) {
  int ratio = 7, c;
  const char* prefix = NULL;
  while ((c = getopt(argc, argv, "r:p:")) >= 0) {
    switch (c) {
      case 'r': ratio = atoi(optarg); break;
      case 'p': prefix = optarg; break;
      default: break;
    }
  }
  if (optind >= argc) {
    fprintf(stderr, "Must specify the FASTA file\nExiting\n");
    return 1;
  }
  const char* fa = argv[optind];
  if (!prefix) prefix = fa;
  if (ratio <= 0) {
    fprintf(stderr, "invalid ratio %d\n", ratio);
    return 1;
  }
  FILE* fp = fopen(fa, "r");
  if (!fp) {
    perror(fa);
    return 1;
  }
  size_t cap = 1 << 20, n = 0;
  uint8_t* codes = (uint8_t*)malloc(cap);
  char* line = NULL;
  size_t lcap = 0;
  ssize_t nb;
  uint64_t rng = 11;
  while ((nb = getline(&line, &lcap, fp)) >= 0) {
    if (line[0] == '>') continue;
    for (ssize_t i = 0; i < nb; ++i) {
      char ch = line[i];
      if (ch == '\n' || ch == '\r' || ch == ' ' || ch == '\t') continue;
      int v;
      switch (ch | 0x20) {
        case 'a': v = 0; break;
        case 'c': v = 1; break;
        case 'g': v = 2; break;
        case 't': v = 3; break;
        default:
          rng = rng * 6364136223846793005ull + 1442695040888963407ull;
          v = (int)(rng >> 62);
          break;
      }
      if (n == cap) {
        cap *= 2;
        codes = (uint8_t*)realloc(codes, cap);
      }
      codes[n++] = (uint8_t)v;
    }
  }
  free(line);
  fclose(fp);
  fprintf(stderr, "Building the suffix position array with compression ratio = %d\n", ratio);
  int rc = gs_build_index_host(codes, n, (uint32_t)ratio, prefix);
  free(codes);
  if (rc != GS_OK) {
    fprintf(stderr, "index construction failed (%d)\n", rc);
    return 1;
  }
  fprintf(stderr, "Wrote %s.bwt and %s.sa (%zu bases)\n", prefix, prefix, n);
  return 0;
}
