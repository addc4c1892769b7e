/*
 * Reference-format FM-index files (.bwt / .sa) as consumed by seed_gen.
 *
 * Layout (seed_gen.cu:29,52-58 for the block macros; 2685-2736 for the readers;
 * build_index.sh:54-75 for the two-pass bwa build that produces them):
 *   .bwt : u32 primary, u32 L2[1..4], then the $-removed BWT in blocks of
 *          OCC_INTERVAL = 64 bases: 4 u32 cumulative counts (A,C,G,T before the
 *          block) followed by 4 u32 words of 16 bases each, first base in the
 *          two most significant bits; a final counts-only block closes the body.
 *   .sa  : u32 primary, u32 L2[1..4], u32 sa_intv, u32 seq_len, then
 *          SA[i*sa_intv] for i = 1 .. n_sa-1, n_sa = (seq_len + sa_intv) / sa_intv.
 * All integers are the reference's 32-bit bwtint_t (seed_gen.cu:22).
 */
#include <errno.h>
#include <stdlib.h>
#include <string.h>

#include "gs_host.h"

uint64_t gs_bwt_words(uint64_t n) { return (n + 15) / 16 + 4 * ((n + 63) / 64 + 1); }

static int write_all(FILE* fp, const void* p, size_t sz, size_t cnt) {
  return fwrite(p, sz, cnt, fp) == cnt ? 0 : GS_EIO;
}

/* Build the reference-layout body from the text and its suffix array. */
static void bwt_body_from_sa(const uint8_t* codes, uint32_t n, const uint32_t* SA, uint32_t primary,
                             uint32_t* body) {
  uint32_t c[4] = {0, 0, 0, 0};
  uint64_t k = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t row = i < primary ? i : i + 1; /* $-removed index -> SA row */
    uint32_t ch = codes[SA[row] - 1];
    if ((i & 63) == 0) {
      memcpy(body + k, c, sizeof c);
      k += 4;
    }
    if ((i & 15) == 0) ++k;
    body[k - 1] |= ch << ((15 - (i & 15)) << 1);
    ++c[ch];
  }
  memcpy(body + k, c, sizeof c);
}

int gs_build_index_host(const uint8_t* codes, uint64_t n64, uint32_t sa_intv, const char* prefix) {
  if (!codes || !prefix || n64 == 0 || n64 >= 0xFFFFFFFEull || sa_intv == 0) return GS_EINVAL;
  const uint32_t n = (uint32_t)n64;
  uint32_t* SA = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  if (!SA) return GS_ENOMEM;
  int rc = gs_sais_u32(codes, n, SA);
  if (rc) {
    free(SA);
    return rc;
  }
  uint32_t L2[5] = {0, 0, 0, 0, 0};
  for (uint32_t i = 0; i < n; ++i) L2[1 + codes[i]]++;
  for (int c = 2; c <= 4; ++c) L2[c] += L2[c - 1];
  uint32_t primary = 0;
  for (uint32_t i = 0; i <= n; ++i)
    if (SA[i] == 0) {
      primary = i;
      break;
    }
  const uint64_t words = gs_bwt_words(n);
  uint32_t* body = (uint32_t*)calloc(words, sizeof(uint32_t));
  if (!body) {
    free(SA);
    return GS_ENOMEM;
  }
  bwt_body_from_sa(codes, n, SA, primary, body);

  size_t plen = strlen(prefix);
  char* fn = (char*)malloc(plen + 8);
  if (!fn) {
    free(SA);
    free(body);
    return GS_ENOMEM;
  }
  rc = 0;
  memcpy(fn, prefix, plen);
  strcpy(fn + plen, ".bwt");
  FILE* fp = fopen(fn, "wb");
  if (!fp) rc = GS_EIO;
  if (!rc) rc = write_all(fp, &primary, 4, 1);
  if (!rc) rc = write_all(fp, L2 + 1, 4, 4);
  if (!rc) rc = write_all(fp, body, 4, words);
  if (fp && fclose(fp) && !rc) rc = GS_EIO;
  if (!rc) {
    strcpy(fn + plen, ".sa");
    fp = fopen(fn, "wb");
    if (!fp) rc = GS_EIO;
    const uint32_t n_sa = (uint32_t)(((uint64_t)n + sa_intv) / sa_intv);
    if (!rc) rc = write_all(fp, &primary, 4, 1);
    if (!rc) rc = write_all(fp, L2 + 1, 4, 4);
    if (!rc) rc = write_all(fp, &sa_intv, 4, 1);
    if (!rc) rc = write_all(fp, &n, 4, 1);
    for (uint32_t i = 1; !rc && i < n_sa; ++i) rc = write_all(fp, &SA[(uint64_t)i * sa_intv], 4, 1);
    if (fp && fclose(fp) && !rc) rc = GS_EIO;
  }
  free(fn);
  free(SA);
  free(body);
  return rc;
}

void gs_host_index_free(gs_host_index* idx) {
  if (!idx) return;
  free(idx->bwt);
  free(idx->sa);
  memset(idx, 0, sizeof(*idx));
}

/* Mirrors bwt_restore_bwt / bwt_restore_sa (seed_gen.cu:2685-2736), with the
 * exit() calls turned into return codes and the missing size checks added.
 * Files in upstream bwa 0.7's 64-bit layout are recognised by their header
 * and size and converted (bwa07.c). */
int gs_host_index_load(const char* prefix, gs_host_index* idx) {
  if (!prefix || !idx) return GS_EINVAL;
  if (gs_host_is_bwa07(prefix)) return gs_host_index_load_bwa07(prefix, idx);
  memset(idx, 0, sizeof(*idx));
  size_t plen = strlen(prefix);
  char* fn = (char*)malloc(plen + 8);
  if (!fn) return GS_ENOMEM;
  memcpy(fn, prefix, plen);
  strcpy(fn + plen, ".bwt");
  FILE* fp = fopen(fn, "rb");
  if (!fp) {
    free(fn);
    return GS_EIO;
  }
  int rc = 0;
  long fsz = -1;
  if (fseek(fp, 0, SEEK_END) == 0) fsz = ftell(fp);
  if (fsz < 20) rc = GS_EFORMAT;
  if (!rc && fseek(fp, 0, SEEK_SET)) rc = GS_EIO;
  if (!rc) {
    idx->bwt_words = (uint64_t)(fsz - 20) >> 2;
    idx->bwt = (uint32_t*)malloc(sizeof(uint32_t) * (idx->bwt_words ? idx->bwt_words : 1));
    if (!idx->bwt) rc = GS_ENOMEM;
  }
  if (!rc && (fread(&idx->primary, 4, 1, fp) != 1 || fread(idx->L2 + 1, 4, 4, fp) != 4 ||
              fread(idx->bwt, 4, idx->bwt_words, fp) != idx->bwt_words))
    rc = GS_EIO;
  fclose(fp);
  idx->L2[0] = 0;
  idx->seq_len = idx->L2[4];
  if (!rc && (idx->seq_len == 0 || idx->bwt_words < gs_bwt_words(idx->seq_len) ||
              idx->primary > idx->seq_len))
    rc = GS_EFORMAT;
  if (!rc) {
    strcpy(fn + plen, ".sa");
    fp = fopen(fn, "rb");
    if (!fp) rc = GS_EIO;
    uint32_t hdr[7];
    if (!rc && fread(hdr, 4, 7, fp) != 7) rc = GS_EIO;
    if (!rc && hdr[0] != idx->primary) rc = GS_EFORMAT;
    if (!rc && hdr[6] != idx->seq_len) rc = GS_EFORMAT;
    if (!rc && hdr[5] == 0) rc = GS_EFORMAT;
    if (!rc) {
      idx->sa_intv = hdr[5];
      idx->n_sa = ((uint64_t)idx->seq_len + idx->sa_intv) / idx->sa_intv;
      idx->sa = (uint32_t*)malloc(sizeof(uint32_t) * idx->n_sa);
      if (!idx->sa) rc = GS_ENOMEM;
    }
    if (!rc) {
      idx->sa[0] = 0xFFFFFFFFu;
      if (fread(idx->sa + 1, 4, idx->n_sa - 1, fp) != idx->n_sa - 1) rc = GS_EIO;
    }
    if (fp) fclose(fp);
  }
  free(fn);
  if (rc) gs_host_index_free(idx);
  return rc;
}
