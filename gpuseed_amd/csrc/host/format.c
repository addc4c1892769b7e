/*
 * The -o text writer (seed_gen.cu:3625-3648), byte-identical to the
 * reference's fprintf loop and fast enough not to dominate a run:
 * hand-rolled decimal conversion, one buffer per slice of reads, slices
 * formatted on several threads and written in read order.
 *
 * Per read: the header line, then for each position of each seed either
 * ", <sign><pos>" when (qbeg, qend) (raw, with the strand bit) equals the
 * previous position's pair, or "\nRead[<qbeg>, <qend>] -> <sign><pos>" with
 * the strand bit cleared; then "\n".  The read number is printed with %d of
 * an int, as the reference does (3630-3631).
 */
#define _GNU_SOURCE
#include <errno.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "gpuseed.h"

static const char kHead[2][2][128] = {
    {"/*===================================MEM seeds in read no. ",
     " (Read[begin, End] -> starting position(s) on the reference)===================================*/\n"},
    {"/*===================================SMEM seeds in read no. ",
     " (Read[begin, End] -> starting position(s) on the reference)===================================*/\n"}};

static char* put_u32(char* p, uint32_t v) {
  char t[10];
  int n = 0;
  do {
    t[n++] = (char)('0' + v % 10u);
    v /= 10u;
  } while (v);
  while (n) *p++ = t[--n];
  return p;
}

static char* put_i32(char* p, int32_t v) {
  if (v < 0) {
    *p++ = '-';
    return put_u32(p, 0u - (uint32_t)v);
  }
  return put_u32(p, (uint32_t)v);
}

/* upper bound of the text of reads [r0, r1) */
static size_t bound(const gs_result* res, uint64_t r0, uint64_t r1) {
  const size_t nseeds = res->read_seed_off[r1] - res->read_seed_off[r0];
  const size_t npos = res->read_pos_off[r1] - res->read_pos_off[r0];
  return (r1 - r0) * 200u + nseeds * 40u + npos * 13u + 16u;
}

typedef struct {
  const gs_result* res;
  int smem;
  uint64_t first_read_no, r0, r1;
  char* buf;
  size_t len;
  int err;
} slice;

static void format_slice(slice* s) {
  const gs_result* res = s->res;
  size_t cap = bound(res, s->r0, s->r1);
  char* p = (char*)malloc(cap);
  s->buf = p;
  if (!p) {
    s->err = GS_ENOMEM;
    return;
  }
  const char* h0 = kHead[s->smem ? 1 : 0][0];
  const char* h1 = kHead[s->smem ? 1 : 0][1];
  const size_t l0 = strlen(h0), l1 = strlen(h1);
  static const char sign[2] = {'+', '-'};
  for (uint64_t r = s->r0; r < s->r1; ++r) {
    memcpy(p, h0, l0);
    p += l0;
    p = put_i32(p, (int32_t)(uint32_t)(s->first_read_no + r));
    memcpy(p, h1, l1);
    p += l1;
    const gs_seed_rec* sd = res->seeds + res->read_seed_off[r];
    const uint32_t* pos = res->pos + res->read_pos_off[r];
    const uint32_t ns = res->read_nseeds[r];
    int32_t pb = -1, pe = -1; /* raw pair of the last printed position */
    for (uint32_t j = 0; j < ns; ++j) {
      const uint32_t n = sd[j].n_pos;
      if (!n) continue;
      const char sg = sign[sd[j].qend >> 31];
      const int32_t x = (int32_t)sd[j].qbeg, y = (int32_t)sd[j].qend;
      const int cont = pb == x && pe == y;
      pb = x;
      pe = y;
      if (cont) {
        *p++ = ',';
        *p++ = ' ';
      } else {
        memcpy(p, "\nRead[", 6);
        p += 6;
        p = put_i32(p, (int32_t)(sd[j].qbeg & 0x7FFFFFFFu));
        *p++ = ',';
        *p++ = ' ';
        p = put_i32(p, (int32_t)(sd[j].qend & 0x7FFFFFFFu));
        memcpy(p, "] -> ", 5);
        p += 5;
      }
      *p++ = sg;
      p = put_u32(p, *pos++);
      for (uint32_t q = 1; q < n; ++q) {
        *p++ = ',';
        *p++ = ' ';
        *p++ = sg;
        p = put_u32(p, *pos++);
      }
    }
    *p++ = '\n';
  }
  s->len = (size_t)(p - s->buf);
}

static void* slice_main(void* a) {
  format_slice((slice*)a);
  return NULL;
}

static int write_all(int fd, const char* p, size_t n) {
  while (n) {
    ssize_t w = write(fd, p, n);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) return GS_EIO;
    p += w;
    n -= (size_t)w;
  }
  return GS_OK;
}

static int n_threads(void) {
  const char* e = getenv("GS_FORMAT_THREADS");
  long t = e ? strtol(e, NULL, 10) : sysconf(_SC_NPROCESSORS_ONLN);
  if (t < 1) t = 1;
  if (t > 16) t = 16;
  return (int)t;
}

GS_API int gs_format_text(const gs_result* res, int smem, uint64_t first_read_no, int fd) {
  if (!res || fd < 0) return GS_EINVAL;
  if (!res->n_reads) return GS_OK;
  enum { kMaxSlices = 16 };
  const size_t target = (size_t)8 << 20; /* text bytes per slice */
  const int T = n_threads();
  slice sl[kMaxSlices];
  pthread_t th[kMaxSlices];
  uint64_t r = 0;
  int rc = GS_OK;
  while (r < res->n_reads && rc == GS_OK) {
    /* cut up to T slices of about `target` bytes each */
    int ns = 0;
    while (ns < T && r < res->n_reads) {
      uint64_t lo = r + 1, hi = res->n_reads;
      if (bound(res, r, hi) > target) {
        while (lo < hi) { /* smallest end whose bound exceeds the target */
          uint64_t mid = lo + (hi - lo) / 2;
          if (bound(res, r, mid) > target)
            hi = mid;
          else
            lo = mid + 1;
        }
      } else {
        lo = hi;
      }
      slice s = {res, smem, first_read_no, r, lo, NULL, 0, GS_OK};
      sl[ns++] = s;
      r = lo;
    }
    int started[kMaxSlices] = {0};
    for (int i = 1; i < ns; ++i) started[i] = pthread_create(&th[i], NULL, slice_main, &sl[i]) == 0;
    format_slice(&sl[0]);
    for (int i = 1; i < ns; ++i) {
      if (started[i])
        pthread_join(th[i], NULL);
      else
        format_slice(&sl[i]);
    }
    for (int i = 0; i < ns; ++i) {
      if (rc == GS_OK) rc = sl[i].err;
      if (rc == GS_OK) rc = write_all(fd, sl[i].buf, sl[i].len);
      free(sl[i].buf);
    }
  }
  return rc;
}
