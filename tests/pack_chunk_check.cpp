// CPU check of gs::pack_chunk (aligned word loads) against the byte-by-byte
// definition of the 4-bit codes (pack_4bit_fow / pack_4bit_rev,
// seed_gen.cu:2536-2578), with the
// read placed against inaccessible pages on either side, so that a load
// outside the read's own aligned words would fault.  Built and run by
// test_pack_chunk.py.
#include <sys/mman.h>
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "gs_device.h"

static uint32_t cf(uint32_t b) { return (0x24031005u >> ((b & 7u) << 2)) & 15u; }
static uint32_t cr(uint32_t b) { return (0x14002035u >> ((b & 7u) << 2)) & 15u; }
static uint4 old_chunk(const uint8_t* s, uint32_t L, uint32_t strand, uint32_t c) {
  uint32_t w[4] = {0, 0, 0, 0};
  for (uint32_t j = 0; j < 32; ++j) {
    const uint32_t i = c * 32 + j;
    if (i >= L) break;
    const uint32_t v = strand ? cr(s[L - 1 - i]) : cf(s[i]);
    w[j >> 3] |= v << ((j & 7) << 2);
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}
int main() {
  const size_t pg = sysconf(_SC_PAGESIZE);
  // [guard][2 pages data][guard]
  uint8_t* m = (uint8_t*)mmap(nullptr, 4 * pg, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
  mprotect(m, pg, PROT_NONE);
  mprotect(m + 3 * pg, pg, PROT_NONE);
  uint8_t* lo = m + pg; uint8_t* hi = m + 3 * pg;
  srand(7);
  const char al[] = "ACGTNacgtnRYKMSWBDHV\r@X";
  long n = 0, bad = 0;
  for (int trial = 0; trial < 40; ++trial) {
    for (size_t i = 0; i < 2 * pg; ++i) lo[i] = (trial & 1) ? (uint8_t)rand() : (uint8_t)al[rand() % (sizeof(al) - 1)];
    for (uint32_t L = 1; L <= 300; ++L) {
      for (int place = 0; place < 8; ++place) {
        // 0-3: read start d bytes after the lower guard; 4-7: read ending e bytes before the upper guard
        const uint8_t* s = place < 4 ? lo + place : hi - L - (place - 4);
        const uint32_t nch = (L + 31) / 32;
        for (uint32_t strand = 0; strand < 2; ++strand)
          for (uint32_t c = 0; c < nch + 1; ++c) {
            const uint4 a = gs::pack_chunk(s, L, strand, c), b = old_chunk(s, L, strand, c);
            ++n;
            if (a.x != b.x || a.y != b.y || a.z != b.z || a.w != b.w) {
              if (bad++ < 5) printf("diff L=%u place=%d strand=%u c=%u\n", L, place, strand, c);
            }
          }
      }
    }
  }
  printf("%ld chunks compared, %ld differ\n", n, bad);
  return bad != 0;
}
