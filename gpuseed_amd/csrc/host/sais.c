/*
 * Suffix-array construction by induced sorting (SA-IS), written from the
 * published algorithm: G. Nong, S. Zhang, W. H. Chan, "Two Efficient
 * Algorithms for Linear Time Suffix Array Construction", IEEE Trans. Comput.
 * 60(10), 2011 (conference version DCC 2009).
 *
 * This replaces the suffix-array step of the bwa fork that build_index.sh
 * drives (build_index.sh:63, `bwa index -b sa`), which is not vendored.  The
 * index bytes are fully determined by the text, so any correct suffix sorter
 * reproduces them.
 *
 * Indices are 32-bit (the reference's bwtint_t is uint32_t, seed_gen.cu:22),
 * so the text with its sentinel must be shorter than 2^32 - 1.
 */
#include <stdlib.h>
#include <string.h>

#include "gs_host.h"

#define EMPTY 0xFFFFFFFFu

typedef struct {
  const void* s;
  int wide; /* 0: uint8 symbols, 1: uint32 symbols */
} sym_t;

static inline uint32_t sym(sym_t T, uint32_t i) {
  return T.wide ? ((const uint32_t*)T.s)[i] : ((const uint8_t*)T.s)[i];
}

/* type bit: 1 = S-type */
static inline int is_s(const uint8_t* t, uint32_t i) { return (t[i >> 3] >> (i & 7)) & 1; }
static inline void set_s(uint8_t* t, uint32_t i) { t[i >> 3] |= (uint8_t)(1u << (i & 7)); }
static inline int is_lms(const uint8_t* t, uint32_t i) { return i > 0 && is_s(t, i) && !is_s(t, i - 1); }

static void bucket_bounds(sym_t T, uint32_t n, uint32_t K, uint32_t* bkt, int ends) {
  memset(bkt, 0, sizeof(uint32_t) * K);
  for (uint32_t i = 0; i < n; ++i) bkt[sym(T, i)]++;
  uint32_t sum = 0;
  for (uint32_t c = 0; c < K; ++c) {
    sum += bkt[c];
    bkt[c] = ends ? sum : sum - bkt[c];
  }
}

static void induce_l(sym_t T, uint32_t* SA, uint32_t n, uint32_t K, const uint8_t* t, uint32_t* bkt) {
  bucket_bounds(T, n, K, bkt, 0);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t p = SA[i];
    if (p != EMPTY && p > 0 && !is_s(t, p - 1)) SA[bkt[sym(T, p - 1)]++] = p - 1;
  }
}

static void induce_s(sym_t T, uint32_t* SA, uint32_t n, uint32_t K, const uint8_t* t, uint32_t* bkt) {
  bucket_bounds(T, n, K, bkt, 1);
  for (uint32_t i = n; i-- > 0;) {
    uint32_t p = SA[i];
    if (p != EMPTY && p > 0 && is_s(t, p - 1)) SA[--bkt[sym(T, p - 1)]] = p - 1;
  }
}

/* T[0..n) with T[n-1] the unique smallest symbol (0); alphabet [0,K). */
static int sais_rec(sym_t T, uint32_t* SA, uint32_t n, uint32_t K) {
  uint8_t* t = (uint8_t*)calloc(((size_t)n + 7) / 8, 1);
  uint32_t* bkt = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)K);
  if (!t || !bkt) {
    free(t);
    free(bkt);
    return GS_ENOMEM;
  }
  /* classify: S if T[i] < T[i+1], or equal and T[i+1] is S; the sentinel is S */
  set_s(t, n - 1);
  for (uint32_t i = n - 1; i-- > 0;) {
    uint32_t a = sym(T, i), b = sym(T, i + 1);
    if (a < b || (a == b && is_s(t, i + 1))) set_s(t, i);
  }

  /* stage 1: sort LMS substrings by one round of induced sorting */
  for (uint32_t i = 0; i < n; ++i) SA[i] = EMPTY;
  bucket_bounds(T, n, K, bkt, 1);
  for (uint32_t i = 1; i < n; ++i)
    if (is_lms(t, i)) SA[--bkt[sym(T, i)]] = i;
  induce_l(T, SA, n, K, t, bkt);
  induce_s(T, SA, n, K, t, bkt);

  /* compact the sorted LMS positions into SA[0..n1) */
  uint32_t n1 = 0;
  for (uint32_t i = 0; i < n; ++i)
    if (is_lms(t, SA[i])) SA[n1++] = SA[i];

  /* name LMS substrings; equal substrings get equal names */
  for (uint32_t i = n1; i < n; ++i) SA[i] = EMPTY;
  uint32_t name = 0, prev = EMPTY;
  for (uint32_t i = 0; i < n1; ++i) {
    uint32_t pos = SA[i];
    int diff = 0;
    for (uint32_t d = 0;; ++d) {
      if (prev == EMPTY || sym(T, pos + d) != sym(T, prev + d) || is_s(t, pos + d) != is_s(t, prev + d)) {
        diff = 1;
        break;
      }
      if (d > 0 && (is_lms(t, pos + d) || is_lms(t, prev + d))) break;
    }
    if (diff) {
      ++name;
      prev = pos;
    }
    SA[n1 + (pos >> 1)] = name - 1;
  }
  /* gather names in text order into the tail: reduced string S1 = SA[n-n1..n) */
  for (uint32_t i = n, j = n; i-- > n1;)
    if (SA[i] != EMPTY) SA[--j] = SA[i];

  uint32_t* S1 = SA + n - n1;
  uint32_t* SA1 = SA;
  int rc = 0;
  if (name < n1) {
    sym_t T1 = {S1, 1};
    rc = sais_rec(T1, SA1, n1, name);
    if (rc) {
      free(t);
      free(bkt);
      return rc;
    }
  } else {
    for (uint32_t i = 0; i < n1; ++i) SA1[S1[i]] = i;
  }

  /* stage 3: induce the full SA from the sorted LMS suffixes */
  for (uint32_t i = 1, j = 0; i < n; ++i)
    if (is_lms(t, i)) S1[j++] = i; /* S1 now maps reduced index -> text position */
  for (uint32_t i = 0; i < n1; ++i) SA1[i] = S1[SA1[i]];
  for (uint32_t i = n1; i < n; ++i) SA[i] = EMPTY;
  bucket_bounds(T, n, K, bkt, 1);
  for (uint32_t i = n1; i-- > 0;) {
    uint32_t p = SA[i];
    SA[i] = EMPTY;
    SA[--bkt[sym(T, p)]] = p;
  }
  induce_l(T, SA, n, K, t, bkt);
  induce_s(T, SA, n, K, t, bkt);
  free(t);
  free(bkt);
  return 0;
}

int gs_sais_u32(const uint8_t* codes, uint32_t n, uint32_t* sa) {
  if (!codes || !sa || n == 0 || n >= 0xFFFFFFFEu) return GS_EINVAL;
  /* shift symbols by one so that 0 is free for the sentinel */
  uint8_t* s = (uint8_t*)malloc((size_t)n + 1);
  if (!s) return GS_ENOMEM;
  for (uint32_t i = 0; i < n; ++i) {
    if (codes[i] > 3) {
      free(s);
      return GS_EINVAL;
    }
    s[i] = (uint8_t)(codes[i] + 1);
  }
  s[n] = 0;
  sym_t T = {s, 0};
  int rc = sais_rec(T, sa, n + 1, 5);
  free(s);
  return rc;
}
