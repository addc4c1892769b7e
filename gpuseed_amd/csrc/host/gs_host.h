/* Internal declarations shared by the host-side C sources of libgpuseed. */
#ifndef GS_HOST_H
#define GS_HOST_H

#include <stdint.h>
#include <stdio.h>

#include "gpuseed.h"

#ifdef __cplusplus
extern "C" {
#endif

/* splitmix64 (Steele, Lea, Flood 2014): tiny, fast, good enough for synthetic data. */
static inline uint64_t gs_splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
/* uniform double in [0,1) */
static inline double gs_unif(uint64_t* s) { return (gs_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
/* uniform integer in [0, n) */
static inline uint64_t gs_below(uint64_t* s, uint64_t n) {
  return (uint64_t)(((unsigned __int128)gs_splitmix64(s) * n) >> 64);
}

/* Suffix array of codes[0..n) (alphabet 0..3) followed by a unique smallest
 * sentinel: sa[0..n] with sa[0] == n.  Returns 0 or a negative GS_ code. */
int gs_sais_u32(const uint8_t* codes, uint32_t n, uint32_t* sa);

/* Number of u32 words of a reference-format .bwt body for a text of n bases. */
uint64_t gs_bwt_words(uint64_t n);

/* Host copy of an index in the reference's on-disk layout. */
typedef struct {
  uint32_t primary;
  uint32_t L2[5];
  uint32_t seq_len;
  uint64_t bwt_words; /* size of bwt[] in u32                              */
  uint32_t* bwt;      /* 64-base blocks: 4 counts + 4 words of 16 bases    */
  uint32_t sa_intv;
  uint64_t n_sa;
  uint32_t* sa; /* sa[0] = 0xFFFFFFFF, sa[i] = SA[i*sa_intv]            */
} gs_host_index;

int gs_host_index_load(const char* prefix, gs_host_index* idx);
void gs_host_index_free(gs_host_index* idx);

/* Upstream bwa 0.7 files (64-bit integers, 128-base blocks): detection by
 * header and size, and loading into the fork layout above. */
int gs_host_is_bwa07(const char* prefix);
int gs_host_index_load_bwa07(const char* prefix, gs_host_index* idx);

#ifdef __cplusplus
}
#endif

#endif
