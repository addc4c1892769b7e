// Device-side index layout and rank primitives shared by the HIP kernels.
//
// Occurrence layout ("occ96"): the $-removed BWT is cut into blocks of 96
// positions.  For every block b and symbol c there is one 16-byte entry
//     { count of c before the block, 96-bit bitmap of c inside the block }
// stored at occ[4*b + c].  A rank query occ(c, i) is then one 16-byte load
// plus three popcounts.  The choice follows the measured random-access
// behaviour of MI355X: about 48 G distinct random requests/s whatever their
// size up to a 128-byte line, while one lane issuing several loads to one line
// pays per load; a single 16-byte request per query is the cheapest shape.
// The reference instead reads a 32-byte block (4 counts + 64 bases) with a
// uint4 load and a count load (seed_gen.cu:195-216, 329-368).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gs {

constexpr uint32_t kBlockBases = 96;
constexpr int kPrecalc = 13;  // pre_calc_seed_len (seed_gen.cu:2822)
constexpr uint32_t kMaxQuirk = 64;  // rows of the primary+1 walk kept for locating from the full suffix array
constexpr int kPairCore = 18;       // bases of a pair-table core: its bytes answer 19-base windows

// floor(i / 96) for any 32-bit i: (i >> 5) / 3 via the exact multiply-high
// reciprocal of 3.
__device__ __forceinline__ uint32_t div96(uint32_t i) { return __umulhi(i >> 5, 0xAAAAAAABu) >> 1; }

// n low bits set, n in [0, 32]
__device__ __forceinline__ uint32_t low_mask(uint32_t n) { return n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u); }

// count stored in the entry plus the set bits of positions [0, bit] of the block
__device__ __forceinline__ uint32_t rank_in_entry(uint4 e, uint32_t bit) {
  const uint32_t r = bit + 1;  // 1..96 positions counted
  const uint32_t n1 = r > 32 ? r - 32 : 0, n2 = r > 64 ? r - 64 : 0;
  return e.x + __popc(e.y & low_mask(r)) + __popc(e.z & low_mask(n1)) + __popc(e.w & low_mask(n2));
}

struct DevIndex {
  const uint4* occ;      // [n_blocks * 4]
  const uint2* kmer;     // 4^13 (l, u) intervals
  // optional 4^ext_len entries (ext_len 14..16), else null: {l, u} of the
  // pattern and a 64-bit presence mask (bit c1 + 4*c2 + 16*c3 set iff the
  // pattern extended to the left by c1, then c2, then c3 occurs)
  const uint4* kext;
  const uint32_t* sa;    // sampled suffix array, sa[0] = 0xFFFFFFFF
  // optional direct-extension structures, else null: the full suffix array
  // (sa_full[row] = text position, sa_full[0] = seq_len) and its inverse
  // (isa[pos] = row), seq_len + 1 words each, and the text at 2 bits per base
  // (base i at bits 2*(i & 15) of word i >> 4), read 64 bases per 16 bytes
  const uint32_t* sa_full;
  const uint32_t* isa;
  const uint4* text;
  // with sa_full: the rows whose reference walk passes row primary+1, each
  // with the position that walk reports (at most kMaxQuirk; none if the
  // list did not fit, and then positions come from the walk)
  const uint2* quirk;
  // optional pair table (4^18 bytes), else null: for every 18-base core, bit
  // c (0..3) set iff the core followed by base c occurs in the text, bit
  // 4 + c iff base c followed by the core occurs; core index as the interval
  // tables' (sum of code_j * 4^j, the core's last base as j = 0)
  const uint8_t* pair;
  // tandem table (kTandemKeys * kTandemRows 16-byte entries), see gs_internal.h
  const uint4* tandem;
  uint32_t n_quirk, locate_direct;
  uint32_t primary, seq_len, sa_intv, ext_len;  // ext_len 0: no extended table
  uint64_t sa_div;       // ceil(2^64 / sa_intv) (0 when sa_intv == 1)
  uint32_t L2[5];
};

__device__ __forceinline__ uint32_t row_to_idx(uint32_t row, uint32_t primary) {
  return row - (row >= primary ? 1u : 0u);
}

// One backward-extension step: the interval of c.P from the interval [l, u]
// of P (seed_gen.cu:2169-2174 with find_occ_gpu, 329-368).  c must be 0..3.
// Rows l-1 and u are answered from one entry when they share a block.
__device__ __forceinline__ void extend(const DevIndex& X, uint32_t c, uint32_t& l, uint32_t& u) {
  const uint32_t r1 = l - 1u;
  const uint32_t i2 = row_to_idx(u, X.primary);
  const uint32_t b2 = div96(i2);
  const bool have1 = r1 != 0xFFFFFFFFu;
  const uint32_t i1 = have1 ? row_to_idx(r1, X.primary) : 0u;
  const uint32_t b1 = div96(i1);
  const uint4 e2 = X.occ[(size_t)b2 * 4 + c];
  uint4 e1 = e2;
  if (have1 && b1 != b2) e1 = X.occ[(size_t)b1 * 4 + c];
  const uint32_t o2 = rank_in_entry(e2, i2 - b2 * kBlockBases);
  const uint32_t o1 = have1 ? rank_in_entry(e1, i1 - b1 * kBlockBases) : 0u;
  const uint32_t base = X.L2[c];
  l = base + o1 + 1u;
  u = base + o2;
}

// LF of the suffix-array walk with the reference's ordering of the primary
// test (bwt_inv_psi_gpu, seed_gen.cu:455-484): rows primary and primary+1
// both map to row 0.  The symbol at the row is found from the four bitmaps
// of its block (one 64-byte span).
__device__ __forceinline__ uint32_t lf_walk(const DevIndex& X, uint32_t k) {
  if (k > X.primary) --k;
  if (k == X.primary) return 0;
  const uint32_t b = div96(k), bit = k - b * kBlockBases;
  const uint4* p = X.occ + (size_t)b * 4;
  const uint32_t w = bit >> 5, sh = bit & 31;
  uint4 e0 = p[0], e1 = p[1], e2 = p[2], e3 = p[3];
  auto has = [&](const uint4& e) -> uint32_t {
    const uint32_t word = w == 0 ? e.y : (w == 1 ? e.z : e.w);
    return (word >> sh) & 1u;
  };
  uint4 e = e3;
  uint32_t c = 3;
  if (has(e0)) {
    e = e0;
    c = 0;
  } else if (has(e1)) {
    e = e1;
    c = 1;
  } else if (has(e2)) {
    e = e2;
    c = 2;
  }
  return X.L2[c] + rank_in_entry(e, bit);
}

__device__ __forceinline__ uint32_t div_sa(const DevIndex& X, uint32_t row) {
  return X.sa_div ? (uint32_t)__umul64hi((uint64_t)row, X.sa_div) : row;
}

// read byte -> code, seed_gen.cu:2540 (forward) and 2562 (reverse complement)
constexpr uint32_t kLutFwd = 0x24031005u, kLutRev = 0x14002035u;
__device__ __forceinline__ uint32_t code_fwd(uint32_t byte) { return (kLutFwd >> ((byte & 7u) << 2)) & 15u; }
__device__ __forceinline__ uint32_t code_rev(uint32_t byte) { return (kLutRev >> ((byte & 7u) << 2)) & 15u; }

// Chunk c (codes 32c .. 32c+31) of a read's virtual strand in 4-bit codes,
// two per byte, even position in the low nibble; codes at or past the read's
// length L are 0.  Forward code i is the code of byte i, reverse code i the
// reverse-complement code of byte L-1-i (pack_4bit_fow / pack_4bit_rev,
// seed_gen.cu:2536-2578).  The chunk's codes come from at most 32
// consecutive bytes of the read, loaded as the 4-byte-aligned words that hold
// them, all at once: a word is read only if it holds a byte of the read, so
// no load leaves [s rounded down to 4 bytes, s + L + 3).
__host__ __device__ __forceinline__ uint4 pack_chunk(const uint8_t* s, uint32_t L, uint32_t strand, uint32_t c) {
  // the window of 32 bytes, in read positions: forward [32c, 32c + 32),
  // reverse [L - 32c - 32, L - 32c) (its start may be negative)
  const int32_t lo = strand ? (int32_t)L - 32 * (int32_t)c - 32 : 32 * (int32_t)c;
  const uint32_t d = (uint32_t)((uintptr_t)s & 3u);  // s's offset in its aligned word
  const uint32_t* const base = reinterpret_cast<const uint32_t*>(s - d);
  const int32_t m0 = (lo + (int32_t)d) >> 2;  // the window's first word (floor)
  const uint32_t off = (uint32_t)(lo + (int32_t)d - 4 * m0);
  uint32_t w[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int32_t m = m0 + k, first = 4 * m - (int32_t)d;  // read position of the word's first byte
    w[k] = (first + 4 > 0 && first < (int32_t)L) ? base[m] : 0u;
  }
  // the window's bytes in order (reverse: last byte first)
  uint32_t x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = (uint32_t)((((uint64_t)w[k + 1] << 32) | w[k]) >> (8 * off));
  if (strand) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t t = x[k];
      x[k] = __builtin_bswap32(x[7 - k]);
      x[7 - k] = __builtin_bswap32(t);
    }
  }
  const uint32_t lut = strand ? kLutRev : kLutFwd;
  const uint32_t nvalid = L > 32 * c ? L - 32 * c : 0u;  // codes of the chunk inside the read
  uint32_t out[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t byte = x[2 * q + (j >> 2)] >> (8 * (j & 3));
      v |= ((lut >> ((byte & 7u) << 2)) & 15u) << (4 * j);
    }
    const uint32_t keep = nvalid >= 8u * (q + 1) ? 32u : (nvalid > 8u * q ? 4 * (nvalid - 8u * q) : 0u);
    out[q] = keep == 32u ? v : (v & ((1u << keep) - 1u));
  }
  return make_uint4(out[0], out[1], out[2], out[3]);
}

}  // namespace gs
