/*
 * gs_build_index - builds <prefix>.bwt and <prefix>.sa from a FASTA file.
 *
 *   gs_build_index [-r RATIO] [-p PREFIX] ref.fa
 *   gs_build_index [-r RATIO] -p PREFIX -b BWA_PREFIX
 *   gs_build_index -p PREFIX -c BWA07_PREFIX
 *
 * Same command line as build_index.sh (build_index.sh:24-50: -r is the
 * suffix-array sampling interval, default 7; -p defaults to the FASTA name),
 * same output files (build_index.sh:63-75), without the bwa fork: all records
 * are concatenated, A/C/G/T (either case) map to 0..3 and any other letter is
 * replaced by a pseudo-random base drawn from a fixed-seed stream.
 * -b takes the reference from bwa's BWA_PREFIX.pac/.ann/.amb instead (the
 * concatenated sequences with bwa's own random bases in runs of N) and lists
 * its sequences; -c converts the .bwt/.sa of upstream bwa 0.7 at BWA07_PREFIX
 * to the layout seed_gen reads.
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
  const char* bwa_prefix = NULL;
  const char* convert = NULL;
  while ((c = getopt(argc, argv, "r:p:b:c:")) >= 0) {
    switch (c) {
      case 'r': ratio = atoi(optarg); break;
      case 'p': prefix = optarg; break;
      case 'b': bwa_prefix = optarg; break;
      case 'c': convert = optarg; break;
      default: break;
    }
  }
  if (convert || bwa_prefix) {
    if (!prefix || ratio <= 0) {
      fprintf(stderr, "-b and -c need -p PREFIX (and a positive -r)\n");
      return 1;
    }
    int rc;
    if (convert) {
      rc = gs_index_convert(convert, prefix);
      if (rc == GS_OK) fprintf(stderr, "Wrote %s.bwt and %s.sa from %s\n", prefix, prefix, convert);
    } else {
      gs_bwa_ref* R = NULL;
      rc = gs_bwa_ref_load(bwa_prefix, &R);
      if (rc == GS_OK) {
        for (int i = 0; i < gs_bwa_ref_n_contigs(R); ++i) {
          const char* name = NULL;
          uint64_t off = 0, len = 0;
          gs_bwa_ref_contig(R, i, &name, &off, &len);
          fprintf(stderr, "%s\toffset %llu\tlength %llu\n", name, (unsigned long long)off, (unsigned long long)len);
        }
        fprintf(stderr, "Building the suffix position array with compression ratio = %d\n", ratio);
        rc = gs_build_index_host(gs_bwa_ref_codes(R), gs_bwa_ref_length(R), (uint32_t)ratio, prefix);
        if (rc == GS_OK)
          fprintf(stderr, "Wrote %s.bwt and %s.sa (%llu bases, %d sequences, %d ambiguous runs)\n", prefix, prefix,
                  (unsigned long long)gs_bwa_ref_length(R), gs_bwa_ref_n_contigs(R), gs_bwa_ref_n_holes(R));
        gs_bwa_ref_free(R);
      }
    }
    if (rc != GS_OK) {
      fprintf(stderr, "failed (%d)\n", rc);
      return 1;
    }
    return 0;
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
