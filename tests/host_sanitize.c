/*
 * Drives the host C code of libgpuseed (csrc/host/*.c: synthetic generators,
 * SA-IS and the index writer, the index readers, the bwa 0.7 readers and
 * converter, the text formatter) in a build with AddressSanitizer and
 * UBSan.  No GPU code is linked.  Built and run by test_host_sanitize.py:
 *   host_sanitize <tmp dir> <bwa 0.7 index prefix> <bwa 0.7 reference prefix>
 */
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "gpuseed.h"
#include "host/gs_host.h"

#define CHECK(x)                                                  \
  do {                                                            \
    if (!(x)) {                                                   \
      fprintf(stderr, "check failed: %s (line %d)\n", #x, __LINE__); \
      return 1;                                                   \
    }                                                             \
  } while (0)

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint32_t next32(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (uint32_t)rng;
}

int main(int argc, char** argv) {
  CHECK(argc == 4);
  char prefix[4096], conv[4096];
  snprintf(prefix, sizeof prefix, "%s/synth", argv[1]);
  snprintf(conv, sizeof conv, "%s/converted", argv[1]);

  /* every preset; a small reference and reads of the uneven-length mix */
  gs_synth_ref_params rp;
  gs_synth_read_params qp;
  for (int c = 1; c <= 5; ++c) CHECK(gs_synth_preset(c, &rp, &qp) == GS_OK);
  CHECK(gs_synth_preset(4, &rp, &qp) == GS_OK);
  rp.length = 300000;
  uint8_t* codes = malloc(rp.length);
  CHECK(codes && gs_synth_reference(&rp, codes, 3) == GS_OK);
  qp.n_reads = 300;
  uint32_t* lens = malloc(4 * qp.n_reads);
  uint64_t* offs = malloc(8 * (qp.n_reads + 1));
  CHECK(lens && offs && gs_synth_read_lengths(&qp, lens) == GS_OK);
  offs[0] = 0;
  for (uint64_t i = 0; i < qp.n_reads; ++i) offs[i + 1] = offs[i] + lens[i];
  char* reads = malloc(offs[qp.n_reads]);
  uint32_t* opos = malloc(4 * qp.n_reads);
  uint8_t* ostr = malloc(qp.n_reads);
  CHECK(reads && opos && ostr && gs_synth_reads(codes, rp.length, &qp, offs, reads, opos, ostr, 3) == GS_OK);

  /* index writer (SA-IS) and reader */
  CHECK(gs_build_index_host(codes, rp.length, 7, prefix) == GS_OK);
  gs_host_index idx;
  CHECK(gs_host_index_load(prefix, &idx) == GS_OK);
  CHECK(idx.seq_len == rp.length && idx.sa_intv == 7);
  gs_host_index_free(&idx);
  CHECK(!gs_host_is_bwa07(prefix));

  /* the text formatter on a result with random seeds, both modes */
  gs_result res;
  memset(&res, 0, sizeof res);
  res.n_reads = 200;
  res.read_nseeds = calloc(res.n_reads, 4);
  res.read_npos = calloc(res.n_reads, 4);
  res.read_seed_off = calloc(res.n_reads + 1, 8);
  res.read_pos_off = calloc(res.n_reads + 1, 8);
  CHECK(res.read_nseeds && res.read_npos && res.read_seed_off && res.read_pos_off);
  for (uint64_t r = 0; r < res.n_reads; ++r) {
    res.read_nseeds[r] = next32() % 6;
    res.read_seed_off[r + 1] = res.read_seed_off[r] + res.read_nseeds[r];
  }
  res.n_seeds = res.read_seed_off[res.n_reads];
  res.seeds = calloc(res.n_seeds + 1, sizeof(gs_seed_rec));
  CHECK(res.seeds);
  for (uint64_t r = 0; r < res.n_reads; ++r) {
    uint32_t np = 0;
    for (uint64_t s = res.read_seed_off[r]; s < res.read_seed_off[r + 1]; ++s) {
      gs_seed_rec* q = &res.seeds[s];
      const uint32_t rev = (next32() & 1) ? 0x80000000u : 0u;
      q->qbeg = (next32() % 100) | rev;
      q->qend = ((q->qbeg & 0x7FFFFFFFu) + 19 + next32() % 30) | rev;
      q->l = next32() % 1000;
      q->u = q->l + next32() % 4;
      q->n_pos = q->u - q->l + 1;
      np += q->n_pos;
    }
    res.read_npos[r] = np;
    res.read_pos_off[r + 1] = res.read_pos_off[r] + np;
  }
  res.n_pos = res.read_pos_off[res.n_reads];
  res.pos = calloc(res.n_pos + 1, 4);
  CHECK(res.pos);
  for (uint64_t i = 0; i < res.n_pos; ++i) res.pos[i] = next32();
  const int fd = open("/dev/null", O_WRONLY);
  CHECK(fd >= 0);
  CHECK(gs_format_text(&res, 1, 1, fd) == GS_OK);
  CHECK(gs_format_text(&res, 0, 1001, fd) == GS_OK);
  close(fd);
  free(res.read_nseeds);
  free(res.read_npos);
  free(res.read_seed_off);
  free(res.read_pos_off);
  free(res.seeds);
  free(res.pos);

  /* bwa 0.7 index: detection, loading, conversion to the fork layout */
  CHECK(gs_host_is_bwa07(argv[2]));
  CHECK(gs_host_index_load_bwa07(argv[2], &idx) == GS_OK);
  const uint64_t n07 = idx.seq_len;
  gs_host_index_free(&idx);
  CHECK(gs_index_convert(argv[2], conv) == GS_OK);
  CHECK(gs_host_index_load(conv, &idx) == GS_OK);
  CHECK(idx.seq_len == n07);
  gs_host_index_free(&idx);

  /* bwa 0.7 reference files */
  gs_bwa_ref* ref = NULL;
  CHECK(gs_bwa_ref_load(argv[3], &ref) == GS_OK);
  const uint64_t rl = gs_bwa_ref_length(ref);
  CHECK(gs_bwa_ref_codes(ref) != NULL && gs_bwa_ref_n_contigs(ref) > 0);
  for (uint64_t p = 0; p < rl; p += 1 + rl / 97) {
    int contig = -1;
    uint64_t off = 0;
    CHECK(gs_bwa_ref_locate(ref, p, &contig, &off) == GS_OK && contig >= 0);
  }
  gs_bwa_ref_free(ref);

  free(codes);
  free(lens);
  free(offs);
  free(reads);
  free(opos);
  free(ostr);
  printf("host code clean\n");
  return 0;
}
