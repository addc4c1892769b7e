// Device-resident FM-index: load, validate, repack into occ96, 13-mer table.
//
// Replaces the reference's index path: bwt_restore_bwt / bwt_restore_sa
// (seed_gen.cu:2685-2736), the cudaMemcpy upload (2864-2876) and
// pre_calc_seed_intervals_gpu (2648-2675, launched at 2887-2895).
#include <string.h>
#include <sys/time.h>

#include <chrono>

#include "gs_internal.h"
#include "host/gs_host.h"

using gs::DevIndex;

static double wall() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

uint64_t gs_reciprocal64(uint32_t d) {
  if (d <= 1) return 0;
  const unsigned __int128 one = (unsigned __int128)1 << 64;
  unsigned __int128 q = one / d;
  if (q * d != one) q += 1;
  return (uint64_t)q;
}

DevIndex gs_index::dev() const {
  DevIndex d;
  d.occ = occ;
  d.kmer = kmer;
  d.kext = kext;
  d.sa = sa;
  d.sa_full = sa_full;
  d.isa = isa;
  d.text = reinterpret_cast<const uint4*>(text);
  d.quirk = quirk;
  d.pair = pair;
  d.tandem = tandem;
  d.n_quirk = n_quirk;
  d.locate_direct = sa_full && n_quirk <= gs::kMaxQuirk ? 1u : 0u;
  d.primary = primary;
  d.seq_len = seq_len;
  d.sa_intv = sa_intv;
  d.ext_len = kext ? ext_len : 0;
  d.sa_div = gs_reciprocal64(sa_intv);
  for (int c = 0; c < 5; ++c) d.L2[c] = L2[c];
  return d;
}

namespace {

// symbol at $-removed position p of a reference-layout body (seed_gen.cu:52-58)
__device__ __forceinline__ uint32_t ref_word(const uint32_t* bwt, uint64_t p) {
  return bwt[(p >> 6) * 8 + 4 + ((p & 63) >> 4)];
}

// occurrences of c among the 16 bases of one reference word (pop_count_full, 152-158)
__device__ __forceinline__ uint32_t count_in_word(uint32_t w, uint32_t c) {
  uint32_t odd = ((c & 2) ? w : ~w) >> 1;
  uint32_t even = (c & 1) ? w : ~w;
  return __popc(odd & even & 0x55555555u);
}

// Every 64-base block's counts must equal the previous block's counts plus its
// bases, and the closing counts (the last four words of the body) must equal
// the last block's counts plus its bases; any mismatch sets *err.  Files
// written by build_index.sh satisfy this.
__global__ void check_ref_counts(const uint32_t* __restrict__ bwt, uint32_t n, uint64_t n64,
                                 uint64_t tail, uint32_t* __restrict__ err) {
  const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j > n64) return;
  if (j == 0) {
    if (bwt[0] | bwt[1] | bwt[2] | bwt[3]) atomicOr(err, 1u);
    return;
  }
  const uint32_t* prev = bwt + (j - 1) * 8;
  const uint32_t* cur = j == n64 ? bwt + tail : bwt + j * 8;
  const uint64_t first = (j - 1) * 64;
  const uint32_t valid = n - first < 64 ? (uint32_t)(n - first) : 64u;
  uint32_t s[4] = {prev[0], prev[1], prev[2], prev[3]};
  for (uint32_t q = 0; q < valid; ++q) {
    const uint32_t w = prev[4 + (q >> 4)];
    s[(w >> ((15 - (q & 15)) << 1)) & 3u]++;
  }
  for (uint32_t c = 0; c < 4; ++c)
    if (s[c] != cur[c]) atomicOr(err, 1u);
}

__global__ void repack_occ96(const uint32_t* __restrict__ bwt, uint32_t n, uint64_t n_blocks,
                             uint4* __restrict__ occ) {
  const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (b >= n_blocks) return;
  const uint64_t p0 = b * gs::kBlockBases;
  const uint64_t j = p0 >> 6;
  uint32_t cnt[4];
  for (int c = 0; c < 4; ++c) cnt[c] = bwt[j * 8 + c];
  for (uint64_t p = j << 6; p < p0; p += 16) {
    const uint32_t w = ref_word(bwt, p);
    for (uint32_t c = 0; c < 4; ++c) cnt[c] += count_in_word(w, c);
  }
  uint32_t bits[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (uint32_t w6 = 0; w6 < 6; ++w6) {
    const uint64_t p = p0 + 16 * w6;
    if (p >= n) break;
    const uint32_t word = ref_word(bwt, p);
    const uint32_t m = n - p < 16 ? (uint32_t)(n - p) : 16u;
    for (uint32_t q = 0; q < m; ++q) {
      const uint32_t c = (word >> ((15 - q) << 1)) & 3u;
      const uint32_t pos = 16 * w6 + q;
      bits[c][pos >> 5] |= 1u << (pos & 31);
    }
  }
  for (uint32_t c = 0; c < 4; ++c) occ[b * 4 + c] = make_uint4(cnt[c], bits[c][0], bits[c][1], bits[c][2]);
}

// One level of the 13-mer table: entries of level i+1 from level i.  Entry p
// after i characters depends on the low 2i bits of p; character i is applied
// next; an empty interval is carried unchanged (the break at 2666-2668).
__global__ void kmer_level(const DevIndex X, const uint2* __restrict__ prev, uint2* __restrict__ next,
                           uint32_t level, uint64_t first, uint64_t end) {
  const uint64_t t = first + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (t >= end) return;
  const uint64_t q = t & ((1ull << (2 * level)) - 1);
  const uint32_t c = (uint32_t)(t >> (2 * level));
  uint32_t l = 0, u = X.seq_len;
  if (level > 0) {
    const uint2 iv = prev[q];
    l = iv.x;
    u = iv.y;
  }
  if (l <= u) gs::extend(X, c, l, u);
  next[t] = make_uint2(l, u);
}

// Presence bits of one extended-table entry: which of the 64 three-base
// left extensions c3.c2.c1.P of pattern P occur (bit c1 + 4*c2 + 16*c3),
// found by a depth-first search that stops at empty intervals.
__global__ void kmer_presence(const DevIndex X, const uint2* __restrict__ level, uint4* __restrict__ out,
                              uint64_t first, uint64_t end) {
  const uint64_t t = first + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (t >= end) return;
  const uint2 iv = level[t];
  uint64_t mask = 0;
  if (iv.x <= iv.y) {
    for (uint32_t c1 = 0; c1 < 4; ++c1) {
      uint32_t l1 = iv.x, u1 = iv.y;
      gs::extend(X, c1, l1, u1);
      if (l1 > u1) continue;
      for (uint32_t c2 = 0; c2 < 4; ++c2) {
        uint32_t l2 = l1, u2 = u1;
        gs::extend(X, c2, l2, u2);
        if (l2 > u2) continue;
        for (uint32_t c3 = 0; c3 < 4; ++c3) {
          uint32_t l3 = l2, u3 = u2;
          gs::extend(X, c3, l3, u3);
          if (l3 <= u3) mask |= 1ull << (c1 + 4 * c2 + 16 * c3);
        }
      }
    }
  }
  out[t] = make_uint4(iv.x, iv.y, (uint32_t)mask, (uint32_t)(mask >> 32));
}

// The full suffix array, its inverse and the text, from the occurrence
// layout and the sampled suffix array: thread i starts at sampled row
// i * sa_intv, whose text position the sampled array holds, and walks LF
// (the textbook one: the row of the suffix one position to the left) until
// it reaches position 0 or another sampled row, recording each row's
// position and the base it precedes.  Every row is visited by exactly one
// thread; positions fall at every step, so each walk ends.  The walk checks
// the index against itself: a walk that lands on a sampled row must arrive
// with the position the sampled array holds there, the walk that reaches
// position 0 must be at row primary, and the walks together must visit
// n + 1 rows.  err[0] collects bit 1 for a wrong landing, bit 2 for a row or
// position out of range; err[1..2] count the rows visited (64 bits).
__global__ void direct_chains(const DevIndex X, uint64_t first, uint64_t end, uint32_t* __restrict__ sa_full,
                              uint32_t* __restrict__ isa, uint8_t* __restrict__ text8,
                              uint32_t* __restrict__ err) {
  const uint64_t i = first + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= end) return;
  const uint32_t n = X.seq_len;
  uint32_t row = (uint32_t)i * X.sa_intv;
  uint32_t pos = i ? X.sa[i] : n;
  uint32_t visited = 0, bad = 0;
  while (true) {
    if (row > n || pos > n) {
      bad |= 2u;
      break;
    }
    if (sa_full) sa_full[row] = pos;  // both null: only the text is wanted
    if (isa) isa[pos] = row;
    ++visited;
    if (pos == 0) {  // the suffix at position 0, preceded by $: row primary
      if (row != X.primary) bad |= 1u;
      break;
    }
    const uint32_t kk = row - (row > X.primary ? 1u : 0u);
    const uint32_t b = gs::div96(kk), bit = kk - b * gs::kBlockBases;
    const uint4* p = X.occ + (size_t)b * 4;
    const uint32_t w = bit >> 5, sh = bit & 31;
    uint32_t c = 3;
    uint4 e = p[3];
    for (uint32_t s = 0; s < 3; ++s) {
      const uint4 es = p[s];
      const uint32_t word = w == 0 ? es.y : (w == 1 ? es.z : es.w);
      if ((word >> sh) & 1u) {
        c = s;
        e = es;
        break;
      }
    }
    text8[pos - 1] = (uint8_t)c;
    row = X.L2[c] + gs::rank_in_entry(e, bit);
    --pos;
    const uint32_t j = gs::div_sa(X, row);
    if (row - j * X.sa_intv == 0) {  // another thread's start: it must hold this position
      if (row > n || X.sa[j] != pos || row == 0) bad |= 1u;
      break;
    }
  }
  if (bad) atomicOr(err, bad);
  atomicAdd(reinterpret_cast<unsigned long long*>(err + 2), (unsigned long long)visited);
}

// 16 bases per word, base i at bits 2*(i & 15); words past the text are 0
__global__ void pack_text(const uint8_t* __restrict__ text8, uint64_t n, uint32_t* __restrict__ text,
                          uint64_t n_words) {
  const uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  uint32_t v = 0;
  for (uint32_t q = 0; q < 16; ++q) {
    const uint64_t pos = 16 * w + q;
    if (pos < n) v |= (uint32_t)(text8[pos] & 3u) << (2 * q);
  }
  text[w] = v;
}

// Pair table: the 18-base core at text position p (its index: the sum of
// text[p + 17 - j] * 4^j) gets bit text[p + 18] (a base that follows it) and
// bit 4 + text[p - 1] (a base that precedes it).  One thread per core; the
// text at 2 bits per base, 16 per word.
__global__ void pair_bits(const uint32_t* __restrict__ text, uint64_t n, uint64_t first, uint64_t end,
                          uint32_t* __restrict__ table) {
  const uint64_t p = first + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (p >= end || p + gs::kPairCore > n) return;
  auto base = [&](uint64_t i) -> uint32_t { return (text[i >> 4] >> (2 * (i & 15))) & 3u; };
  uint64_t idx = 0;
  for (int j = 0; j < gs::kPairCore; ++j) idx |= (uint64_t)base(p + gs::kPairCore - 1 - j) << (2 * j);
  uint32_t bits = 0;
  if (p + gs::kPairCore < n) bits |= 1u << base(p + gs::kPairCore);
  if (p > 0) bits |= 16u << base(p - 1);
  if (bits) atomicOr(table + (idx >> 2), bits << (8 * (idx & 3)));
}

// One row of the tandem table per thread: key -> unit (p bases, its last
// base the lowest code), then backward search of the unit repeated, one base
// per length, from the interval of the empty string (rows 0..seq_len).  Once
// a length is absent every longer one is too, so m stops growing.
__global__ void tandem_rows(const DevIndex X, uint4* __restrict__ out) {
  const uint32_t key = blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= gs::kTandemKeys) return;
  uint32_t p = 1;
  while (p < gs::kTandemMax && key >= gs::tandem_key_base(p + 1)) ++p;
  const uint32_t unit = key - gs::tandem_key_base(p);
  uint32_t l = 0, u = X.seq_len, m = 0;
  bool alive = true;
  uint32_t j = 0;  // code j of the unit comes next (base n - 1 - j of a string of length n)
  for (uint32_t n = 1; n <= gs::kTandemLen; ++n) {
    if (alive) {
      uint32_t nl = l, nu = u;
      gs::extend(X, (unit >> (2 * j)) & 3u, nl, nu);
      if (nl > nu) {
        alive = false;
      } else {
        l = nl;
        u = nu;
        m = n;
      }
    }
    if (n >= gs::kTandemMin) out[(size_t)key * gs::kTandemRows + (n - gs::kTandemMin)] = make_uint4(l, u, m, 0);
    j = j + 1 == p ? 0 : j + 1;
  }
}

inline unsigned grid_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// One table level in launches of at most launch_chunk() work-items (level
// 16 has 2^32 entries).
void launch_level(const DevIndex& D, const uint2* prev, uint2* next, uint32_t level, uint64_t cnt,
                  hipStream_t st) {
  const uint64_t step = gs::launch_chunk();
  for (uint64_t first = 0; first < cnt; first += step) {
    const uint64_t n = cnt - first < step ? cnt - first : step;
    hipLaunchKernelGGL(kmer_level, dim3(grid_for(n, 256)), dim3(256), 0, st, D, prev, next, level, first,
                       first + n);
  }
}

std::atomic<uint64_t> g_launch_chunk{1ull << 30};

}  // namespace

uint64_t gs::launch_chunk() { return g_launch_chunk.load(std::memory_order_relaxed); }

// Test support: smaller launches make every chunked loop run past its first
// chunk on a small index (0 restores 2^30).  No counterpart in the reference.
extern "C" GS_API int gs_set_launch_chunk(uint64_t items) {
  if (items > (1ull << 30)) return GS_EINVAL;
  g_launch_chunk.store(items ? items : (1ull << 30), std::memory_order_relaxed);
  return GS_OK;
}

void gs::index_acquire(gs_index* X) { X->refs.fetch_add(1, std::memory_order_relaxed); }

int gs_build_kmer_table(gs_index* X) {
  const uint64_t full = 1ull << (2 * gs::kPrecalc);
  uint2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&X->kmer, full * sizeof(uint2)) != hipSuccess) return GS_ENOMEM;
  X->device_bytes += full * sizeof(uint2);
  if (hipMalloc(&a, (full / 4) * sizeof(uint2)) != hipSuccess) return GS_ENOMEM;
  if (hipMalloc(&b, (full / 4) * sizeof(uint2)) != hipSuccess) {
    (void)hipFree(a);
    return GS_ENOMEM;
  }
  const DevIndex D = X->dev();
  const uint2* prev = nullptr;
  for (uint32_t lev = 0; lev < (uint32_t)gs::kPrecalc; ++lev) {
    const uint64_t cnt = 1ull << (2 * (lev + 1));
    uint2* next = lev + 1 == (uint32_t)gs::kPrecalc ? X->kmer : ((lev & 1) ? b : a);
    launch_level(D, prev, next, lev, cnt, X->stream);
    prev = next;
  }
  hipError_t e = hipStreamSynchronize(X->stream);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return GS_EHIP;
  // the tandem table (about 21 MB), from the occ entries alone
  const uint64_t tbytes = (uint64_t)gs::kTandemKeys * gs::kTandemRows * sizeof(uint4);
  if (hipMalloc(&X->tandem, tbytes) != hipSuccess) {
    X->tandem = nullptr;
    return GS_ENOMEM;
  }
  X->device_bytes += tbytes;
  hipLaunchKernelGGL(tandem_rows, dim3(grid_for(gs::kTandemKeys, 64)), dim3(64), 0, X->stream, X->dev(), X->tandem);
  return hipStreamSynchronize(X->stream) == hipSuccess ? GS_OK : GS_EHIP;
}

// The extended table: the intervals of all 4^len patterns (len 14..16),
// entry p = sum of code_j * 4^j in backward-search order, empty patterns
// stored as l > u, each with the presence bits of its 64 three-base left
// extensions.  The intervals are built from the 13-mer table one level at a
// time with the same kernel, then kmer_presence fills the 16-byte entries.
// Seeding answers the first len bases of a task with one lookup, ends the
// task at once when that pattern, or (with k >= len + 3) its clean
// (len + 3)-base extension, is absent, and goes back to the 13-mer table only
// when the pattern is absent with k < len or its window holds a code above 3.
extern "C" GS_API int gs_index_set_kmer_len(gs_index* X, int len) {
  if (!X || len < gs::kPrecalc || len > 16) return GS_EINVAL;
  if (X->servers.load() > 0) return GS_EINVAL;  // their workers may be reading the table
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  if (X->kext) {
    GS_HIP(hipDeviceSynchronize());  // batches run on streams of their own
    (void)hipFree(X->kext);
    X->device_bytes -= (1ull << (2 * X->ext_len)) * sizeof(uint4);
    X->kext = nullptr;
    X->ext_len = 0;
  }
  if (len == gs::kPrecalc) return GS_OK;
  const double t0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  uint2* levels[3] = {nullptr, nullptr, nullptr};  // levels 14..len
  int rc = GS_OK;
  const DevIndex D = X->dev();
  const uint2* prev = X->kmer;
  for (int lev = gs::kPrecalc; lev < len && !rc; ++lev) {
    const uint64_t cnt = 1ull << (2 * (lev + 1));
    uint2*& next = levels[lev + 1 - 14];
    if (hipMalloc(&next, cnt * sizeof(uint2)) != hipSuccess) {
      rc = GS_ENOMEM;
      break;
    }
    launch_level(D, prev, next, (uint32_t)lev, cnt, X->stream);
    if (hipGetLastError() != hipSuccess) rc = GS_EHIP;
    prev = next;
  }
  if (!rc && hipStreamSynchronize(X->stream) != hipSuccess) rc = GS_EHIP;
  for (int lev = 14; lev < len; ++lev) {  // only the last level is read from here on
    (void)hipFree(levels[lev - 14]);
    levels[lev - 14] = nullptr;
  }
  uint4* table = nullptr;
  const uint64_t cnt = 1ull << (2 * len);
  if (!rc && hipMalloc(&table, cnt * sizeof(uint4)) != hipSuccess) rc = GS_ENOMEM;
  if (!rc) {
    const uint64_t step = gs::launch_chunk();
    for (uint64_t first = 0; first < cnt; first += step) {
      const uint64_t n = cnt - first < step ? cnt - first : step;
      hipLaunchKernelGGL(kmer_presence, dim3(grid_for(n, 256)), dim3(256), 0, X->stream, D,
                         levels[len - 14], table, first, first + n);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(X->stream) != hipSuccess) rc = GS_EHIP;
  }
  (void)hipFree(levels[len - 14]);
  if (!rc) {
    X->kext = table;
    X->ext_len = (uint32_t)len;
    X->device_bytes += cnt * sizeof(uint4);
  } else {
    (void)hipFree(table);
  }
  const double t1 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  X->t_ext = t1 - t0;
  return rc;
}

static void free_direct(gs_index* X) {
  const uint64_t padded = ((uint64_t)X->seq_len + 1 + 3) & ~3ull;  // as allocated
  if (X->sa_full) X->device_bytes -= 2 * padded * sizeof(uint32_t) + X->text_words * sizeof(uint32_t);
  (void)hipFree(X->sa_full);
  (void)hipFree(X->isa);
  (void)hipFree(X->text);
  (void)hipFree(X->quirk);
  X->sa_full = X->isa = X->text = nullptr;
  X->quirk = nullptr;
  X->n_quirk = 0;
  X->text_words = 0;
}

// The reference's walk (bwt_inv_psi_gpu, seed_gen.cu:455-484) sends row
// primary+1 to row 0 instead of to its true LF row, so a row whose walk
// reaches primary+1 before any sampled row reports, instead of its text
// position, the number of steps it took to get there (sa[0] = 0xFFFFFFFF
// plus the steps, seed_gen.cu:668-684).  Those are the rows of text positions
// P, P+1, ... (P = the position of row primary+1) up to the first sampled
// row; the list lets locate_rows read every other position from the full
// suffix array in one load.
static int find_quirk_rows(gs_index* X) {
  X->n_quirk = 0;
  const uint32_t q = X->primary + 1;
  if (q > X->seq_len || q % X->sa_intv == 0) return GS_OK;  // no such row, or sampled: no detour
  uint32_t P = 0;
  GS_HIP(hipMemcpy(&P, X->sa_full + q, 4, hipMemcpyDeviceToHost));
  const uint64_t avail = (uint64_t)X->seq_len + 1 - P;
  const uint32_t m = (uint32_t)(avail < gs::kMaxQuirk + 1 ? avail : gs::kMaxQuirk + 1);
  std::vector<uint32_t> rows(m);
  GS_HIP(hipMemcpy(rows.data(), X->isa + P, (size_t)m * 4, hipMemcpyDeviceToHost));
  std::vector<uint2> list;
  for (uint32_t j = 0; j < m && rows[j] % X->sa_intv != 0; ++j) list.push_back(make_uint2(rows[j], j));
  if (list.size() > gs::kMaxQuirk) {
    X->n_quirk = gs::kMaxQuirk + 1;  // does not fit: locate walks
    return GS_OK;
  }
  X->n_quirk = (uint32_t)list.size();
  if (list.empty()) return GS_OK;
  GS_HIP(hipMalloc(&X->quirk, list.size() * sizeof(uint2)));
  GS_HIP(hipMemcpy(X->quirk, list.data(), list.size() * sizeof(uint2), hipMemcpyHostToDevice));
  return GS_OK;
}

// Direct extension: once a task's interval has a single row, the seeding
// kernels can read that row's text position from the full suffix array,
// compare the read with the text 64 bases per 16-byte load instead of one
// backward step per base, and read the row where the match ends from the
// inverse suffix array.  The structures are derived here from the index
// itself (direct_chains), so they describe exactly its text.
extern "C" GS_API int gs_index_set_direct(gs_index* X, int on) {
  if (!X) return GS_EINVAL;
  if (X->servers.load() > 0) return GS_EINVAL;  // their workers may be reading the structures
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  if (X->sa_full) {
    GS_HIP(hipDeviceSynchronize());  // batches run on streams of their own
    free_direct(X);
  }
  if (!on) return GS_OK;
  const double t0 = wall();
  const uint64_t n = X->seq_len, rows = n + 1;
  const uint64_t words = 4 * ((n + 63) / 64 + 1);  // whole 64-base chunks plus one
  uint8_t* text8 = nullptr;
  uint32_t* d_err = nullptr;  // [0] error bits, [2..3] rows visited
  int rc = GS_OK;
  // whole 16-byte chunks: the persistent kernel reads the 16 bytes that hold a word
  const uint64_t padded = (rows + 3) & ~3ull;
  if (hipMalloc(&X->sa_full, padded * 4) != hipSuccess || hipMalloc(&X->isa, padded * 4) != hipSuccess ||
      hipMalloc(&X->text, words * 4) != hipSuccess || hipMalloc(&text8, n + 1) != hipSuccess ||
      hipMalloc(&d_err, 16) != hipSuccess)
    rc = GS_ENOMEM;
  if (!rc) {
    X->text_words = words;
    X->device_bytes += 2 * padded * 4 + words * 4;
    const DevIndex D = X->dev();
    uint32_t h_err[4] = {0, 0, 0, 0};
    if (hipMemsetAsync(d_err, 0, 16, X->stream) != hipSuccess) rc = GS_EHIP;
    const uint64_t step = gs::launch_chunk();
    for (uint64_t first = 0; !rc && first < X->n_sa; first += step) {
      const uint64_t m = X->n_sa - first < step ? X->n_sa - first : step;
      hipLaunchKernelGGL(direct_chains, dim3(grid_for(m, 256)), dim3(256), 0, X->stream, D, first, first + m,
                         X->sa_full, X->isa, text8, d_err);
    }
    if (!rc)
      hipLaunchKernelGGL(pack_text, dim3(grid_for(words, 256)), dim3(256), 0, X->stream, text8, n, X->text, words);
    if (!rc && (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(h_err, d_err, 16, hipMemcpyDeviceToHost, X->stream) != hipSuccess ||
                hipStreamSynchronize(X->stream) != hipSuccess))
      rc = GS_EHIP;
    // every row once, every landing on the position the sampled array holds
    const uint64_t visited = (uint64_t)h_err[2] | ((uint64_t)h_err[3] << 32);
    if (!rc && (h_err[0] != 0 || visited != rows)) rc = GS_EFORMAT;
  }
  (void)hipFree(d_err);
  (void)hipFree(text8);
  if (!rc) rc = find_quirk_rows(X);
  if (rc) {
    if (!X->text_words) {  // not counted yet
      (void)hipFree(X->sa_full);
      (void)hipFree(X->isa);
      (void)hipFree(X->text);
      X->sa_full = X->isa = X->text = nullptr;
    } else {
      free_direct(X);
    }
  }
  X->t_direct = wall() - t0;
  return rc;
}

// Pair table: one byte per 18-base core saying which bases follow and which
// precede it in the text, so that one lookup answers whether two adjacent
// 19-base windows of a read occur.  Built from the text of the
// direct-extension structures when the index has them, else from a text
// derived here by the same walk of the index.
extern "C" GS_API int gs_index_set_pairs(gs_index* X, int on) {
  if (!X) return GS_EINVAL;
  if (X->servers.load() > 0) return GS_EINVAL;  // their workers may be reading the table
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  const uint64_t bytes = 1ull << (2 * gs::kPairCore);
  if (X->pair) {
    GS_HIP(hipDeviceSynchronize());  // batches run on streams of their own
    (void)hipFree(X->pair);
    X->pair = nullptr;
    X->device_bytes -= bytes;
  }
  if (!on) return GS_OK;
  const double t0 = wall();
  const uint64_t n = X->seq_len, rows = n + 1;
  const uint64_t words = 4 * ((n + 63) / 64 + 1);
  const uint32_t* text = X->text;
  uint32_t* own_text = nullptr;
  uint8_t* text8 = nullptr;
  uint32_t* d_err = nullptr;
  int rc = GS_OK;
  hipStream_t st = X->stream;
  const uint64_t step = gs::launch_chunk();
  if (!text) {
    if (hipMalloc(&own_text, words * 4) != hipSuccess || hipMalloc(&text8, n + 1) != hipSuccess ||
        hipMalloc(&d_err, 16) != hipSuccess)
      rc = GS_ENOMEM;
    uint32_t h_err[4] = {0, 0, 0, 0};
    if (!rc && hipMemsetAsync(d_err, 0, 16, st) != hipSuccess) rc = GS_EHIP;
    const DevIndex D = X->dev();
    for (uint64_t first = 0; !rc && first < X->n_sa; first += step) {
      const uint64_t m = X->n_sa - first < step ? X->n_sa - first : step;
      hipLaunchKernelGGL(direct_chains, dim3(grid_for(m, 256)), dim3(256), 0, st, D, first, first + m,
                         (uint32_t*)nullptr, (uint32_t*)nullptr, text8, d_err);
    }
    if (!rc)
      hipLaunchKernelGGL(pack_text, dim3(grid_for(words, 256)), dim3(256), 0, st, text8, n, own_text, words);
    if (!rc && (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(h_err, d_err, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess))
      rc = GS_EHIP;
    const uint64_t visited = (uint64_t)h_err[2] | ((uint64_t)h_err[3] << 32);
    if (!rc && (h_err[0] != 0 || visited != rows)) rc = GS_EFORMAT;
    (void)hipFree(text8);
    (void)hipFree(d_err);
    text = own_text;
  }
  if (!rc && hipMalloc(&X->pair, bytes) != hipSuccess) {
    X->pair = nullptr;
    rc = GS_ENOMEM;
  }
  if (!rc && hipMemsetAsync(X->pair, 0, bytes, st) != hipSuccess) rc = GS_EHIP;
  const uint64_t cores = n >= (uint64_t)gs::kPairCore ? n - gs::kPairCore + 1 : 0;
  for (uint64_t first = 0; !rc && first < cores; first += step) {
    const uint64_t m = cores - first < step ? cores - first : step;
    hipLaunchKernelGGL(pair_bits, dim3(grid_for(m, 256)), dim3(256), 0, st, text, n, first, first + m,
                       reinterpret_cast<uint32_t*>(X->pair));
  }
  if (!rc && (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)) rc = GS_EHIP;
  (void)hipFree(own_text);
  if (rc) {
    (void)hipFree(X->pair);
    X->pair = nullptr;
  } else {
    X->device_bytes += bytes;
  }
  X->t_pairs = wall() - t0;
  return rc;
}

extern "C" GS_API int gs_index_read_pairs(const gs_index* X, uint64_t first, uint64_t count, uint8_t* out) {
  const uint64_t entries = 1ull << (2 * gs::kPairCore);
  if (!X || !out || !X->pair || first > entries || count > entries - first) return GS_EINVAL;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  GS_HIP(hipMemcpy(out, X->pair + first, count, hipMemcpyDeviceToHost));
  return GS_OK;
}

extern "C" GS_API int gs_index_read_direct(const gs_index* X, uint32_t first, uint32_t count, uint32_t* sa_out,
                                           uint32_t* isa_out, uint8_t* text_out) {
  if (!X || !X->sa_full || (uint64_t)first + count > (uint64_t)X->seq_len + 1) return GS_EINVAL;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  if (sa_out) GS_HIP(hipMemcpy(sa_out, X->sa_full + first, (size_t)count * 4, hipMemcpyDeviceToHost));
  if (isa_out) GS_HIP(hipMemcpy(isa_out, X->isa + first, (size_t)count * 4, hipMemcpyDeviceToHost));
  if (text_out && count) {
    const uint64_t w0 = first >> 4, w1 = ((uint64_t)first + count + 15) >> 4;
    std::vector<uint32_t> w(w1 - w0);
    GS_HIP(hipMemcpy(w.data(), X->text + w0, w.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t q = 0; q < count; ++q) {
      const uint64_t pos = (uint64_t)first + q;
      text_out[q] = (uint8_t)((w[(pos >> 4) - w0] >> (2 * (pos & 15))) & 3u);
    }
  }
  return GS_OK;
}

static void free_index(gs_index* X) {
  if (!X) return;
  if (X->work) gs::batch_free(X->work);
  X->work = nullptr;
  (void)hipFree(X->occ);
  (void)hipFree(X->sa);
  (void)hipFree(X->kmer);
  (void)hipFree(X->kext);
  (void)hipFree(X->pair);
  (void)hipFree(X->tandem);
  free_direct(X);
  if (X->stream) (void)hipStreamDestroy(X->stream);
  delete X;
}

static int pick_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GS_ENODEV;
  if (device < 0 || device >= n) return GS_ENODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return GS_ENODEV;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GS_ENODEV;
  if (hipSetDevice(device) != hipSuccess) return GS_ENODEV;
  return GS_OK;
}

// Upload a host index in the reference layout and build every device structure.
static int index_from_host(const gs_host_index* H, int device, gs_index** out) {
  int rc = pick_device(device);
  if (rc) return rc;
  gs_index* X = new gs_index();
  X->device = device;
  X->primary = H->primary;
  X->seq_len = H->seq_len;
  X->sa_intv = H->sa_intv;
  for (int c = 0; c < 5; ++c) X->L2[c] = H->L2[c];
  X->n_sa = H->n_sa;
  X->n_blocks = ((uint64_t)H->seq_len + gs::kBlockBases - 1) / gs::kBlockBases;
  uint32_t* d_bwt = nullptr;
  uint32_t* d_err = nullptr;
  uint32_t err = 0;
  const double t0 = wall();
  if (hipStreamCreateWithFlags(&X->stream, hipStreamNonBlocking) != hipSuccess) {
    free_index(X);
    return GS_EHIP;
  }
  // the occ entries cover every block plus one spare so that no load can leave the array
  if (hipMalloc(&X->occ, (X->n_blocks + 1) * 4 * sizeof(uint4)) != hipSuccess ||
      hipMalloc(&X->sa, X->n_sa * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&d_bwt, H->bwt_words * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&d_err, sizeof(uint32_t)) != hipSuccess) {
    (void)hipFree(d_bwt);
    (void)hipFree(d_err);
    free_index(X);
    return GS_ENOMEM;
  }
  X->device_bytes = (X->n_blocks + 1) * 4 * sizeof(uint4) + X->n_sa * sizeof(uint32_t);
  rc = GS_OK;
  if (hipMemcpyAsync(d_bwt, H->bwt, H->bwt_words * 4, hipMemcpyHostToDevice, X->stream) != hipSuccess ||
      hipMemcpyAsync(X->sa, H->sa, X->n_sa * 4, hipMemcpyHostToDevice, X->stream) != hipSuccess ||
      hipMemsetAsync(d_err, 0, 4, X->stream) != hipSuccess ||
      hipMemsetAsync(X->occ + X->n_blocks * 4, 0, 4 * sizeof(uint4), X->stream) != hipSuccess ||
      hipStreamSynchronize(X->stream) != hipSuccess)
    rc = GS_EHIP;
  const double t1 = wall();
  if (!rc) {
    const uint64_t n64 = ((uint64_t)H->seq_len + 63) / 64;
    hipLaunchKernelGGL(check_ref_counts, dim3(grid_for(n64 + 1, 256)), dim3(256), 0, X->stream, d_bwt,
                       H->seq_len, n64, gs_bwt_words(H->seq_len) - 4, d_err);
    hipLaunchKernelGGL(repack_occ96, dim3(grid_for(X->n_blocks, 256)), dim3(256), 0, X->stream, d_bwt,
                       H->seq_len, X->n_blocks, X->occ);
    if (hipMemcpyAsync(&err, d_err, 4, hipMemcpyDeviceToHost, X->stream) != hipSuccess ||
        hipStreamSynchronize(X->stream) != hipSuccess)
      rc = GS_EHIP;
  }
  (void)hipFree(d_bwt);
  (void)hipFree(d_err);
  const double t2 = wall();
  if (!rc && err) rc = GS_EFORMAT;
  if (!rc) rc = gs_build_kmer_table(X);
  const double t3 = wall();
  if (rc) {
    free_index(X);
    return rc;
  }
  X->t_upload = t1 - t0;
  X->t_repack = t2 - t1;
  X->t_table = t3 - t2;
  (void)gs::noalloc_setup(device);  // a failed lookup makes gs_seed_noalloc return GS_EHIP
  *out = X;
  return GS_OK;
}

extern "C" GS_API int gs_index_open(const char* prefix, int device, gs_index** out) {
  if (!prefix || !out) return GS_EINVAL;
  *out = nullptr;
  const double t0 = wall();
  gs_host_index H;
  int rc = gs_host_index_load(prefix, &H);
  if (rc) return rc;
  // the reference's counts must agree with L2 (the totals close the body)
  const uint64_t tail = H.bwt_words >= 4 ? gs_bwt_words(H.seq_len) - 4 : 0;
  for (int c = 0; c < 4 && !rc; ++c)
    if (H.bwt[tail + c] != H.L2[c + 1] - H.L2[c]) rc = GS_EFORMAT;
  for (int c = 0; c < 4 && !rc; ++c)
    if (H.L2[c + 1] < H.L2[c]) rc = GS_EFORMAT;
  const double t1 = wall();
  if (!rc) rc = index_from_host(&H, device, out);
  if (!rc) (*out)->t_read = t1 - t0;
  gs_host_index_free(&H);
  return rc;
}

void gs::index_release(gs_index* X) {
  if (X->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) {
    (void)hipSetDevice(X->device);
    free_index(X);
  }
}

// Drops the caller's reference: the index is freed now, or with the last
// batch or server made on it.  gs_seed's own batch goes at once.
extern "C" GS_API int gs_index_close(gs_index* X) {
  if (!X) return GS_EINVAL;
  if (X->work) {
    (void)hipSetDevice(X->device);
    gs::batch_free(X->work);
    X->work = nullptr;
  }
  gs::index_release(X);
  return GS_OK;
}

extern "C" GS_API int gs_index_get_info(const gs_index* X, gs_index_info* info) {
  if (!X || !info) return GS_EINVAL;
  memset(info, 0, sizeof(*info));
  info->seq_len = X->seq_len;
  info->primary = X->primary;
  info->sa_intv = X->sa_intv;
  for (int c = 0; c < 5; ++c) info->L2[c] = X->L2[c];
  info->device_bytes = X->device_bytes;
  info->n_blocks = X->n_blocks;
  info->t_read = X->t_read;
  info->t_upload = X->t_upload;
  info->t_repack = X->t_repack;
  info->t_table = X->t_table;
  info->t_sa = X->t_sa;
  info->t_write = X->t_write;
  info->sa_rounds = X->sa_rounds;
  info->kmer_len = X->kext ? (int)X->ext_len : gs::kPrecalc;
  info->t_ext = X->t_ext;
  info->direct = X->sa_full ? 1 : 0;
  info->t_direct = X->t_direct;
  info->pairs = X->pair ? 1 : 0;
  info->t_pairs = X->t_pairs;
  return GS_OK;
}

// Test access to the device structures: copies occ entries [first, first+count)
// of block-major order and table entries to host buffers.
extern "C" GS_API int gs_index_read_occ(const gs_index* X, uint64_t first, uint64_t count, uint32_t* out4) {
  if (!X || !out4) return GS_EINVAL;
  const uint64_t entries = ((uint64_t)X->n_blocks + 1) * 4;
  if (first > entries || count > entries - first) return GS_EINVAL;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  GS_HIP(hipMemcpy(out4, X->occ + first, count * sizeof(uint4), hipMemcpyDeviceToHost));
  return GS_OK;
}

extern "C" GS_API int gs_index_read_tandem(const gs_index* X, uint64_t first, uint64_t count, uint32_t* out4) {
  if (!X || !out4 || !X->tandem) return GS_EINVAL;
  const uint64_t entries = (uint64_t)gs::kTandemKeys * gs::kTandemRows;
  if (first > entries || count > entries - first) return GS_EINVAL;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  GS_HIP(hipMemcpy(out4, X->tandem + first, count * sizeof(uint4), hipMemcpyDeviceToHost));
  return GS_OK;
}

extern "C" GS_API int gs_index_read_kext(const gs_index* X, uint64_t first, uint64_t count, uint32_t* out4) {
  if (!X || !out4 || !X->kext) return GS_EINVAL;
  const uint64_t entries = 1ull << (2 * X->ext_len);
  if (first > entries || count > entries - first) return GS_EINVAL;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  GS_HIP(hipMemcpy(out4, X->kext + first, count * sizeof(uint4), hipMemcpyDeviceToHost));
  return GS_OK;
}

extern "C" GS_API int gs_index_read_kmer(const gs_index* X, uint64_t first, uint64_t count, uint32_t* out2) {
  const uint64_t entries = 1ull << (2 * gs::kPrecalc);
  if (!X || !out2 || first > entries || count > entries - first) return GS_EINVAL;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  GS_HIP(hipMemcpy(out2, X->kmer + first, count * sizeof(uint2), hipMemcpyDeviceToHost));
  return GS_OK;
}
