/*
 * Index and reference files of upstream bwa 0.7 (the layout `bwa index`
 * writes), read into the layouts this library uses.
 *
 * seed_gen reads the 32-bit files of GPUseed's bwa fork (seed_gen.cu:
 * 2685-2736: u32 header, 64-base occurrence blocks of 4 u32 counts and 4 words
 * of bases).  Upstream bwa 0.7 writes 64-bit integers and 128-base blocks:
 *   .bwt : u64 primary, u64 L2[1..4], then per block of 128 positions of the
 *          $-removed BWT 4 u64 counts (A,C,G,T before the block, as 8 u32
 *          words) and 8 u32 words of 16 bases (first base in the two most
 *          significant bits); a final counts-only block closes the body, so
 *          the body has (n+15)/16 + 8*((n+127)/128 + 1) words.
 *   .sa  : u64 primary, u64 L2[1..4], u64 sa_intv, u64 seq_len, then
 *          SA[i*sa_intv] as u64 for i = 1 .. n_sa-1, n_sa = (seq_len +
 *          sa_intv) / sa_intv.
 * Its text is usually the reference followed by its reverse complement
 * (seq_len = 2 * l_pac); any text of fewer than 2^32 - 2 bases loads.  The
 * occurrence structure is rebuilt in the fork's 64-base layout, so the rest of
 * the library (and seed_gen's procedure) applies unchanged.
 *
 *   .pac : the forward reference, 4 bases per byte, base l in bits
 *          ((~l & 3) << 1) of byte l >> 2; then, if l_pac % 4 == 0, one zero
 *          byte; then one byte holding l_pac % 4.  Runs of N were replaced by
 *          random bases when bwa wrote it (the .amb file lists them).
 *   .ann : "l_pac n_seqs seed", then per sequence "gi name[ anno]" and
 *          "offset len n_ambs".
 *   .amb : "l_pac n_seqs n_holes", then per hole "offset len char".
 */
#include <ctype.h>
#include <stdlib.h>
#include <string.h>

#include "gs_host.h"

static uint64_t bwa07_words(uint64_t n) { return (n + 15) / 16 + 8 * ((n + 127) / 128 + 1); }

static char* path_of(const char* prefix, const char* ext) {
  size_t a = strlen(prefix), b = strlen(ext);
  char* fn = (char*)malloc(a + b + 1);
  if (!fn) return NULL;
  memcpy(fn, prefix, a);
  memcpy(fn + a, ext, b + 1);
  return fn;
}

static long file_size(FILE* fp) {
  if (fseek(fp, 0, SEEK_END)) return -1;
  long s = ftell(fp);
  if (fseek(fp, 0, SEEK_SET)) return -1;
  return s;
}

int gs_host_is_bwa07(const char* prefix) {
  char* fn = path_of(prefix, ".bwt");
  if (!fn) return 0;
  FILE* fp = fopen(fn, "rb");
  free(fn);
  if (!fp) return 0;
  uint64_t h[5];
  const long fsz = file_size(fp);
  const int ok = fsz >= 40 && fread(h, 8, 5, fp) == 5;
  fclose(fp);
  if (!ok) return 0;
  const uint64_t n = h[4];
  if (n == 0 || n >= (1ull << 40) || h[0] > n || h[1] > h[2] || h[2] > h[3] || h[3] > h[4]) return 0;
  return (uint64_t)fsz == 40 + 4 * bwa07_words(n);
}

/* bwa 0.7's .bwt/.sa as a fork-layout host index (same text, same rows). */
int gs_host_index_load_bwa07(const char* prefix, gs_host_index* idx) {
  if (!prefix || !idx) return GS_EINVAL;
  memset(idx, 0, sizeof(*idx));
  char* fn = path_of(prefix, ".bwt");
  if (!fn) return GS_ENOMEM;
  FILE* fp = fopen(fn, "rb");
  free(fn);
  if (!fp) return GS_EIO;
  int rc = 0;
  uint64_t h[5];
  const long fsz = file_size(fp);
  if (fsz < 40 || fread(h, 8, 5, fp) != 5) rc = GS_EIO;
  const uint64_t n = rc ? 0 : h[4];
  if (!rc && (n == 0 || h[0] > n || h[1] > h[2] || h[2] > h[3] || h[3] > h[4] ||
              (uint64_t)fsz != 40 + 4 * bwa07_words(n)))
    rc = GS_EFORMAT;
  if (!rc && n >= 0xFFFFFFFEull) rc = GS_ERANGE; /* e.g. both strands of a human genome: 6.2 G */
  uint32_t* src = NULL;
  const uint64_t src_words = rc ? 0 : bwa07_words(n);
  if (!rc && !(src = (uint32_t*)malloc(4 * src_words))) rc = GS_ENOMEM;
  if (!rc && fread(src, 4, src_words, fp) != src_words) rc = GS_EIO;
  fclose(fp);
  if (!rc) {
    idx->primary = (uint32_t)h[0];
    idx->L2[0] = 0;
    for (int c = 1; c <= 4; ++c) idx->L2[c] = (uint32_t)h[c];
    idx->seq_len = (uint32_t)n;
    idx->bwt_words = gs_bwt_words(n);
    idx->bwt = (uint32_t*)calloc(idx->bwt_words, 4);
    if (!idx->bwt) rc = GS_ENOMEM;
  }
  if (!rc) {
    /* walk the bases, checking every stored 64-bit count, and write the
     * 64-base blocks: 4 u32 counts, then 4 words of bases */
    uint64_t cnt[4] = {0, 0, 0, 0};
    uint64_t k = 0;
    for (uint64_t i = 0; i < n && !rc; i += 16) {
      const uint64_t blk = i >> 7;
      const uint32_t* sb = src + blk * 16;
      if ((i & 127) == 0) {
        for (int c = 0; c < 4; ++c) {
          const uint64_t stored = (uint64_t)sb[2 * c] | ((uint64_t)sb[2 * c + 1] << 32);
          if (stored != cnt[c]) rc = GS_EFORMAT;
        }
      }
      if ((i & 63) == 0) {
        for (int c = 0; c < 4; ++c) idx->bwt[k + c] = (uint32_t)cnt[c];
        k += 4;
      }
      const uint32_t w = sb[8 + ((i & 127) >> 4)];
      idx->bwt[k++] = w;
      const uint64_t m = n - i < 16 ? n - i : 16;
      for (uint64_t j = 0; j < m; ++j) ++cnt[(w >> ((15 - j) << 1)) & 3];
      if (m < 16 && (w & ((1u << ((16 - m) << 1)) - 1u))) rc = GS_EFORMAT; /* bases past the end */
    }
    if (!rc) {
      /* the closing counts follow the last (possibly partial) block's bases */
      const uint64_t rem = n & 127;
      const uint32_t* last = src + 16 * (n >> 7) + (rem ? 8 + (rem + 15) / 16 : 0);
      for (int c = 0; c < 4; ++c) {
        const uint64_t stored = (uint64_t)last[2 * c] | ((uint64_t)last[2 * c + 1] << 32);
        if (stored != cnt[c]) rc = GS_EFORMAT;
        idx->bwt[k + c] = (uint32_t)cnt[c];
      }
      for (int c = 0; c < 4; ++c)
        if (idx->L2[c + 1] - idx->L2[c] != cnt[c]) rc = GS_EFORMAT;
    }
  }
  free(src);
  if (!rc) {
    fn = path_of(prefix, ".sa");
    if (!fn) rc = GS_ENOMEM;
    fp = fn ? fopen(fn, "rb") : NULL;
    free(fn);
    if (!rc && !fp) rc = GS_EIO;
    uint64_t s[7];
    if (!rc && fread(s, 8, 7, fp) != 7) rc = GS_EIO;
    if (!rc && (s[0] != h[0] || s[6] != n || s[5] == 0 || s[5] > 0xFFFFFFFFull)) rc = GS_EFORMAT;
    for (int c = 1; !rc && c <= 4; ++c)
      if (s[c] != h[c]) rc = GS_EFORMAT;
    if (!rc) {
      idx->sa_intv = (uint32_t)s[5];
      idx->n_sa = (n + idx->sa_intv) / idx->sa_intv;
      const long want = 56 + 8 * (long)(idx->n_sa - 1);
      if (file_size(fp) != want || fseek(fp, 56, SEEK_SET)) rc = GS_EFORMAT;
    }
    if (!rc && !(idx->sa = (uint32_t*)malloc(4 * idx->n_sa))) rc = GS_ENOMEM;
    if (!rc) {
      idx->sa[0] = 0xFFFFFFFFu;
      uint64_t buf[4096];
      for (uint64_t i = 1; i < idx->n_sa && !rc;) {
        const uint64_t m = idx->n_sa - i < 4096 ? idx->n_sa - i : 4096;
        if (fread(buf, 8, m, fp) != m) rc = GS_EIO;
        for (uint64_t j = 0; j < m && !rc; ++j) {
          if (buf[j] >= n) rc = GS_EFORMAT;
          idx->sa[i + j] = (uint32_t)buf[j];
        }
        i += m;
      }
    }
    if (fp) fclose(fp);
  }
  if (rc) gs_host_index_free(idx);
  return rc;
}

static int write_index(const gs_host_index* X, const char* prefix) {
  int rc = 0;
  char* fn = path_of(prefix, ".bwt");
  if (!fn) return GS_ENOMEM;
  FILE* fp = fopen(fn, "wb");
  free(fn);
  if (!fp) return GS_EIO;
  if (fwrite(&X->primary, 4, 1, fp) != 1 || fwrite(X->L2 + 1, 4, 4, fp) != 4 ||
      fwrite(X->bwt, 4, X->bwt_words, fp) != X->bwt_words)
    rc = GS_EIO;
  if (fclose(fp) && !rc) rc = GS_EIO;
  if (rc) return rc;
  fn = path_of(prefix, ".sa");
  if (!fn) return GS_ENOMEM;
  fp = fopen(fn, "wb");
  free(fn);
  if (!fp) return GS_EIO;
  if (fwrite(&X->primary, 4, 1, fp) != 1 || fwrite(X->L2 + 1, 4, 4, fp) != 4 || fwrite(&X->sa_intv, 4, 1, fp) != 1 ||
      fwrite(&X->seq_len, 4, 1, fp) != 1 || fwrite(X->sa + 1, 4, X->n_sa - 1, fp) != X->n_sa - 1)
    rc = GS_EIO;
  if (fclose(fp) && !rc) rc = GS_EIO;
  return rc;
}

int gs_index_convert(const char* src_prefix, const char* dst_prefix) {
  if (!src_prefix || !dst_prefix) return GS_EINVAL;
  gs_host_index X;
  int rc = gs_host_index_load(src_prefix, &X);
  if (rc) return rc;
  rc = write_index(&X, dst_prefix);
  gs_host_index_free(&X);
  return rc;
}

/* ---- .pac / .ann / .amb -------------------------------------------------- */

typedef struct {
  char* name;
  uint64_t offset, len;
} contig_t;

typedef struct {
  uint64_t offset, len;
  char c;
} hole_t;

struct gs_bwa_ref {
  uint64_t l_pac;
  uint8_t* codes;
  int n_contigs, n_holes;
  contig_t* contigs;
  hole_t* holes;
};

void gs_bwa_ref_free(gs_bwa_ref* R) {
  if (!R) return;
  for (int i = 0; i < R->n_contigs; ++i) free(R->contigs[i].name);
  free(R->contigs);
  free(R->holes);
  free(R->codes);
  free(R);
}

/* the next line of fp into buf (without the newline); 0 at end of file or
 * for a line longer than the buffer */
static int read_line(FILE* fp, char* buf, size_t cap) {
  if (!fgets(buf, (int)cap, fp)) return 0;
  size_t n = strlen(buf);
  if (n == cap - 1 && buf[n - 1] != '\n') return 0;
  while (n && (buf[n - 1] == '\n' || buf[n - 1] == '\r')) buf[--n] = 0;
  return 1;
}

static int load_ann(const char* prefix, gs_bwa_ref* R) {
  char* fn = path_of(prefix, ".ann");
  if (!fn) return GS_ENOMEM;
  FILE* fp = fopen(fn, "r");
  free(fn);
  if (!fp) return GS_EIO;
  int rc = 0;
  char* line = (char*)malloc(1 << 16);
  long long l_pac = 0;
  int n_seqs = 0;
  unsigned seed = 0;
  if (!line) rc = GS_ENOMEM;
  if (!rc && (!read_line(fp, line, 1 << 16) || sscanf(line, "%lld %d %u", &l_pac, &n_seqs, &seed) != 3 ||
              l_pac <= 0 || n_seqs <= 0 || n_seqs > (1 << 24)))
    rc = GS_EFORMAT;
  if (!rc && !(R->contigs = (contig_t*)calloc((size_t)n_seqs, sizeof(contig_t)))) rc = GS_ENOMEM;
  if (!rc) R->l_pac = (uint64_t)l_pac;
  uint64_t next = 0;
  for (int i = 0; !rc && i < n_seqs; ++i) {
    int gi = 0, pos = 0;
    if (!read_line(fp, line, 1 << 16) || sscanf(line, "%d %n", &gi, &pos) < 1) {
      rc = GS_EFORMAT;
      break;
    }
    const char* nm = line + pos;
    size_t len = 0;
    while (nm[len] && !isspace((unsigned char)nm[len])) ++len;
    if (!len || !(R->contigs[i].name = (char*)malloc(len + 1))) {
      rc = len ? GS_ENOMEM : GS_EFORMAT;
      break;
    }
    memcpy(R->contigs[i].name, nm, len);
    R->contigs[i].name[len] = 0;
    R->n_contigs = i + 1;
    long long off = 0;
    long long ln = 0;
    int n_ambs = 0;
    if (!read_line(fp, line, 1 << 16) || sscanf(line, "%lld %lld %d", &off, &ln, &n_ambs) != 3 || off < 0 ||
        ln < 0 || (uint64_t)off != next || (uint64_t)(off + ln) > R->l_pac) {
      rc = GS_EFORMAT;
      break;
    }
    R->contigs[i].offset = (uint64_t)off;
    R->contigs[i].len = (uint64_t)ln;
    next = (uint64_t)(off + ln);
  }
  if (!rc && next != R->l_pac) rc = GS_EFORMAT;
  free(line);
  fclose(fp);
  return rc;
}

static int load_amb(const char* prefix, gs_bwa_ref* R) {
  char* fn = path_of(prefix, ".amb");
  if (!fn) return GS_ENOMEM;
  FILE* fp = fopen(fn, "r");
  free(fn);
  if (!fp) return GS_EIO;
  int rc = 0;
  char line[256];
  long long l_pac = 0;
  int n_seqs = 0, n_holes = 0;
  if (!read_line(fp, line, sizeof line) || sscanf(line, "%lld %d %d", &l_pac, &n_seqs, &n_holes) != 3 ||
      (uint64_t)l_pac != R->l_pac || n_seqs != R->n_contigs || n_holes < 0)
    rc = GS_EFORMAT;
  if (!rc && n_holes && !(R->holes = (hole_t*)calloc((size_t)n_holes, sizeof(hole_t)))) rc = GS_ENOMEM;
  for (int i = 0; !rc && i < n_holes; ++i) {
    long long off = 0, ln = 0;
    char c = 0;
    if (!read_line(fp, line, sizeof line) || sscanf(line, "%lld %lld %c", &off, &ln, &c) != 3 || off < 0 || ln <= 0 ||
        (uint64_t)(off + ln) > R->l_pac) {
      rc = GS_EFORMAT;
      break;
    }
    R->holes[i].offset = (uint64_t)off;
    R->holes[i].len = (uint64_t)ln;
    R->holes[i].c = c;
    R->n_holes = i + 1;
  }
  fclose(fp);
  return rc;
}

static int load_pac(const char* prefix, gs_bwa_ref* R) {
  char* fn = path_of(prefix, ".pac");
  if (!fn) return GS_ENOMEM;
  FILE* fp = fopen(fn, "rb");
  free(fn);
  if (!fp) return GS_EIO;
  int rc = 0;
  const uint64_t nb = (R->l_pac + 3) / 4;
  uint8_t* pac = NULL;
  if ((uint64_t)file_size(fp) != R->l_pac / 4 + 2) rc = GS_EFORMAT;
  if (!rc && (!(pac = (uint8_t*)malloc(nb + 2)) || !(R->codes = (uint8_t*)malloc(R->l_pac)))) rc = GS_ENOMEM;
  if (!rc && fread(pac, 1, R->l_pac / 4 + 2, fp) != R->l_pac / 4 + 2) rc = GS_EIO;
  if (!rc && pac[R->l_pac / 4 + 1] != (uint8_t)(R->l_pac % 4)) rc = GS_EFORMAT;
  for (uint64_t l = 0; !rc && l < R->l_pac; ++l) R->codes[l] = (pac[l >> 2] >> ((~l & 3) << 1)) & 3;
  free(pac);
  fclose(fp);
  return rc;
}

int gs_bwa_ref_load(const char* prefix, gs_bwa_ref** out) {
  if (!prefix || !out) return GS_EINVAL;
  *out = NULL;
  gs_bwa_ref* R = (gs_bwa_ref*)calloc(1, sizeof(gs_bwa_ref));
  if (!R) return GS_ENOMEM;
  int rc = load_ann(prefix, R);
  if (!rc) rc = load_amb(prefix, R);
  if (!rc) rc = load_pac(prefix, R);
  if (rc) {
    gs_bwa_ref_free(R);
    return rc;
  }
  *out = R;
  return GS_OK;
}

uint64_t gs_bwa_ref_length(const gs_bwa_ref* R) { return R ? R->l_pac : 0; }
const uint8_t* gs_bwa_ref_codes(const gs_bwa_ref* R) { return R ? R->codes : NULL; }
int gs_bwa_ref_n_contigs(const gs_bwa_ref* R) { return R ? R->n_contigs : 0; }
int gs_bwa_ref_n_holes(const gs_bwa_ref* R) { return R ? R->n_holes : 0; }

int gs_bwa_ref_contig(const gs_bwa_ref* R, int i, const char** name, uint64_t* offset, uint64_t* len) {
  if (!R || i < 0 || i >= R->n_contigs) return GS_EINVAL;
  if (name) *name = R->contigs[i].name;
  if (offset) *offset = R->contigs[i].offset;
  if (len) *len = R->contigs[i].len;
  return GS_OK;
}

int gs_bwa_ref_hole(const gs_bwa_ref* R, int i, uint64_t* offset, uint64_t* len, char* c) {
  if (!R || i < 0 || i >= R->n_holes) return GS_EINVAL;
  if (offset) *offset = R->holes[i].offset;
  if (len) *len = R->holes[i].len;
  if (c) *c = R->holes[i].c;
  return GS_OK;
}

int gs_bwa_ref_locate(const gs_bwa_ref* R, uint64_t pos, int* contig, uint64_t* offset) {
  if (!R || !contig || !offset) return GS_EINVAL;
  if (pos >= R->l_pac) return GS_ERANGE;
  int lo = 0, hi = R->n_contigs - 1;
  while (lo < hi) { /* last contig whose offset <= pos */
    const int mid = (lo + hi + 1) / 2;
    if (R->contigs[mid].offset <= pos)
      lo = mid;
    else
      hi = mid - 1;
  }
  *contig = lo;
  *offset = pos - R->contigs[lo].offset;
  return GS_OK;
}
