/*
 * gs_roofline - measures the random-access roofline of the current GPU.
 *
 *   gs_roofline [-b GiB] [-i iters] [-g grid] [-r reps] [-l line] [-m mode] [-f inflight]
 *   gs_roofline -c GiB [-r reps]     (streaming copy of GiB, reads GiB and writes GiB)
 *
 * Modes: 0 independent whole-line reads, 1 dependent chains, 2 one 16-byte
 * chunk of a random line, 3 cooperative (LINE/16 lanes per line), 4 random
 * dwords.  Without -l/-m/-f it sweeps line sizes 16..128 B over all modes,
 * printing one JSON object per configuration on stdout.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>

#include "gpuseed.h"

static int run(double gib, int line, int mode, int inflight, int grid, int iters, int reps) {
  gs_roofline_result r;
  int rc = gs_roofline_random((uint64_t)(gib * 1073741824.0), line, mode, inflight, grid, iters,
                              reps, &r);
  if (rc != GS_OK) {
    fprintf(stderr, "gs_roofline_random failed: %d\n", rc);
    return rc;
  }
  printf("{\"buffer_gib\": %.2f, \"line_bytes\": %d, \"mode\": \"%s\", \"inflight\": %d, "
         "\"grid\": %d, \"iters\": %d, \"seconds\": %.6f, \"glines_per_s\": %.3f, "
         "\"gbytes_per_s\": %.1f}\n",
         gib, line, mode == 0 ? "independent" : mode == 1 ? "dependent" : mode == 2 ? "subline16" : mode == 3 ? "coop" : "dword", inflight, grid, iters, r.seconds,
         r.lines_per_s * 1e-9, r.bytes_per_s * 1e-9);
  fflush(stdout);
  return 0;
}

int main(int argc, char** argv) {
  double gib = 8.0;
  int iters = 256, grid = 8192, reps = 3, line = 0, mode = -1, inflight = 0, c;
  double copy_gib = 0;
  while ((c = getopt(argc, argv, "b:i:g:r:l:m:f:c:")) >= 0) {
    switch (c) {
      case 'b': gib = atof(optarg); break;
      case 'i': iters = atoi(optarg); break;
      case 'g': grid = atoi(optarg); break;
      case 'r': reps = atoi(optarg); break;
      case 'l': line = atoi(optarg); break;
      case 'm': mode = atoi(optarg); break;
      case 'f': inflight = atoi(optarg); break;
      case 'c': copy_gib = atof(optarg); break;
      default: return 2;
    }
  }
  if (copy_gib > 0) { /* streaming copy of known size, for hardware byte-counter calibration */
    double sec = 0;
    uint64_t bytes = (uint64_t)(copy_gib * 1073741824.0);
    if (gs_stream_copy_bytes(bytes, reps, &sec) != GS_OK) return 1;
    printf("{\"copy_bytes\": %llu, \"seconds\": %.6f, \"gbytes_per_s\": %.1f}\n",
           (unsigned long long)bytes, sec, 2.0 * bytes / sec * 1e-9);
    return 0;
  }
  static const int lines[] = {16, 32, 64, 128};
  static const int infl[] = {1, 2, 4};
  for (int li = 0; li < 4; ++li) {
    if (line && lines[li] != line) continue;
    for (int m = 0; m < 5; ++m) {
      if (mode >= 0 && m != mode) continue;
      for (int fi = 0; fi < 3; ++fi) {
        if (m != 0 && fi > 0) continue;
        if (m == 4 && lines[li] != 16) continue;
        if (inflight && infl[fi] != inflight) continue;
        if (run(gib, lines[li], m, infl[fi], grid, iters, reps)) return 1;
      }
    }
  }
  return 0;
}
