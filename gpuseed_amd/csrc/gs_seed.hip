// Seeding kernels: backward extension, SMEM/MEM finalisation, SA walk.
//
// Reference path replaced (seed_gen.cu): find_seed_intervals_gpu (2087-2413),
// DeviceSelect/DeviceScan/DeviceReduce and sum_arrays (3383-3430),
// filter/count kernels and the device-side segmented sort (865-1251),
// seeds_to_threads(_mem) and locate_seeds_gpu (555-724).
//
// Design (one read = two extension streams, forward and reverse):
//  * A lane owns one stream.  Streams are written in a common "virtual" form:
//    the forward stream reads the forward codes left-to-right, the reverse
//    stream reads the reverse-complement codes of the reversed read, and both
//    run the reference's forward-task procedure (2132-2264) on task ends
//    e = L, L-1, ..., k.  A virtual candidate [b, e) of the reverse stream is
//    the reference's reverse seed (L-e, L-b) (2272-2404).
//  * Tasks of a stream run one after another in the same lane, so pruning is
//    exact distance-1 pruning: task e stops at left end x as soon as its
//    interval size equals the size the previous tasks recorded for x (then
//    every occurrence of read[x,e) continues with read[e], so the task's seed
//    has the same left end as a longer one).  The reference approximates this
//    with 30-lane shuffles (2192-2212); its output does not depend on which
//    candidates are pruned, and neither does this one.
//  * Every loop iteration issues at most one memory operation per lane: a
//    13-mer table read or one occ entry (two when l-1 and u fall in different
//    96-base blocks).  The loop stays converged on the loads; all pointer
//    chasing between loads is LDS-only work on the staged read bytes and the
//    per-position size record.
//  * SMEM filtering happens in the lane: tasks finish in decreasing end order
//    and left ends never increase, so a task is a survivor exactly when its
//    left end is smaller than the previous finished task's (991-1018).  For
//    the same reason an SMEM stream ends when a task reaches the read start:
//    every later task would end there too and be filtered out.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "gs_internal.h"

using gs::DevIndex;

namespace {

struct SeedParams {
  DevIndex X;
  const uint8_t* bases;     // batch bytes, padded by 16
  const uint64_t* offs;     // per read
  const uint32_t* lens;     // per read
  const uint64_t* task_off; // per read: sum of tasks of earlier reads
  uint32_t n_reads, k, smem, pad;
  uint4* cand;              // per read: 2*task_off .. : forward region then reverse region
  uint32_t* nF;
  uint32_t* nR;
  void* scratch;            // size records of reads longer than the LDS slots
};

template <typename AT>
struct SatMax {
  static constexpr uint32_t value = (uint32_t)(AT)(~(AT)0);
};

constexpr int kSeedBlock = 128;

template <typename AT, int LMAX>
struct Slot {
  static constexpr int bytes = ((LMAX + 2 * LMAX * (int)sizeof(AT)) + 15) & ~15;
};

template <typename AT, int LMAX>
__global__ void __launch_bounds__(kSeedBlock) seed_streams(SeedParams P) {
  constexpr int PAIR = Slot<AT, LMAX>::bytes;
  constexpr uint32_t SAT = SatMax<AT>::value;
  __shared__ __align__(16) uint8_t lds[(kSeedBlock / 2) * PAIR];
  const uint32_t gid = blockIdx.x * kSeedBlock + threadIdx.x;
  const uint32_t r = gid >> 1;
  const uint32_t strand = gid & 1;
  uint8_t* slot = lds + (threadIdx.x >> 1) * PAIR;
  const bool live = r < P.n_reads;
  const uint32_t L = live ? P.lens[r] : 0;
  const uint64_t off = live ? P.offs[r] : 0;
  const bool in_lds = L <= (uint32_t)LMAX;

  // stage the read bytes: the two lanes of a pair take alternate aligned 16-byte chunks
  if (live && in_lds && L) {
    const uint64_t c0 = off >> 4, c1 = (off + L - 1) >> 4;
    for (uint64_t ch = c0 + strand; ch <= c1; ch += 2) {
      const uint4 v = reinterpret_cast<const uint4*>(P.bases)[ch];
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      for (int q = 0; q < 16; ++q) {
        const uint64_t pos = ch * 16 + q;
        if (pos >= off && pos < off + L) slot[pos - off] = (uint8_t)(w[q >> 2] >> ((q & 3) * 8));
      }
    }
  }
  __syncthreads();

  const uint8_t* src = in_lds ? slot : P.bases + off;
  AT* A = in_lds ? reinterpret_cast<AT*>(slot + LMAX) + strand * LMAX
                 : reinterpret_cast<AT*>(P.scratch) + 2 * off + strand * L;
  auto V = [&](int32_t i) -> uint32_t {
    return strand == 0 ? gs::code_fwd(src[i]) : gs::code_rev(src[L - 1 - i]);
  };

  const uint32_t k = P.k;
  const uint32_t ntask = L >= k ? L - k + 1 : 0;
  bool done = !live || L < k;
  for (uint32_t q = 0; !done && q < L; ++q) A[q] = 0;

  int32_t e = (int32_t)L;
  uint32_t p = 0, nbad = 0;
  if (!done)
    for (int j = 0; j < gs::kPrecalc; ++j) {
      const uint32_t c = V((int32_t)L - 1 - j);
      nbad += c > 3;
      p |= (c & 3u) << (2 * j);
    }
  int phase = 0;  // 0: task e waits for its table lookup; 1: task e is extending
  bool ext_miss = false;  // the task's extended-table pattern is absent: use the 13-mer path
  int32_t x = 0;
  uint32_t l = 0, u = 0, b = 0, s = 0, last_b = 0xFFFFFFFFu, nout = 0;
  uint4* outF = P.cand + (live ? 2 * P.task_off[r] : 0);

  auto next_task = [&]() -> bool {
    ext_miss = false;
    if ((uint32_t)(e - 1) < k) return false;
    const uint32_t cn = V(e - 14), co = V(e - 1);
    nbad = nbad + (cn > 3 ? 1u : 0u) - (co > 3 ? 1u : 0u);
    p = (p >> 2) | ((cn & 3u) << 24);
    --e;
    return true;
  };
  // size record at x: prune if the previous tasks recorded the same size there
  auto prune_at = [&](int32_t pos, uint32_t size) -> bool {
    const uint32_t sv = size >= SAT ? SAT : size;
    if (size < SAT && (uint32_t)A[pos] == sv) return true;
    A[pos] = (AT)sv;
    return false;
  };
  auto finish = [&]() {
    if ((uint32_t)e - b >= k && (!P.smem || b < last_b)) {
      if (strand == 0) {
        outF[ntask - 1 - nout] = make_uint4(b, (uint32_t)e, l, u);
      } else {
        outF[ntask + nout] = make_uint4((L - (uint32_t)e) | 0x80000000u, (L - b) | 0x80000000u, l, u);
      }
      ++nout;
    }
    last_b = b;
  };

  while (true) {
    int op = 0;
    uint32_t c = 0;
    while (!done) {
      if (phase == 0) {
        if (nbad == 0) {
          op = 1;
          break;
        }
        if (!next_task()) done = true;
        continue;
      }
      if (x < 0) {
        finish();
        // SMEM: once a task reaches the read start, every later task does too
        // (its pattern is a substring of this one's), so none can pass the
        // filter b < last_b = 0: the stream is done
        phase = 0;
        if (P.smem || !next_task()) done = true;
        continue;
      }
      c = V(x);
      if (c <= 3) {
        op = 2;
        break;
      }
      b = (uint32_t)x;  // code > 3 after the window: interval kept, position consumed
      if (prune_at(x, s)) {
        phase = 0;
        if (!next_task()) done = true;
        continue;
      }
      --x;
    }
    if (__ballot(op != 0) == 0) break;

    // one memory operation per lane
    const uint4* a1 = nullptr;
    const uint4* a2 = nullptr;
    uint32_t i1 = 0, i2 = 0, b1 = 0, b2 = 0;
    bool have1 = false;
    uint32_t tlen = gs::kPrecalc, tidx = p;  // table lookup: pattern length and entry
    uint32_t pbit = 64;  // presence bit of the clean 3 bases left of an extended window (64: none)
    if (op == 1) {
      a1 = reinterpret_cast<const uint4*>(P.X.kmer) + (p >> 1);  // two 8-byte entries per load
      if (P.X.ext_len && !ext_miss && e >= (int32_t)P.X.ext_len) {
        uint32_t q = p, ok = 1;
        for (uint32_t j = gs::kPrecalc; j < P.X.ext_len; ++j) {
          const uint32_t cj = V(e - 1 - (int32_t)j);
          ok &= cj <= 3 ? 1u : 0u;
          q |= (cj & 3u) << (2 * j);
        }
        if (ok) {
          tlen = P.X.ext_len;
          tidx = q;
          a1 = P.X.kext + q;
          if (k >= tlen + 3 && e >= (int32_t)tlen + 3) {
            const uint32_t c1 = V(e - (int32_t)tlen - 1), c2 = V(e - (int32_t)tlen - 2),
                           c3 = V(e - (int32_t)tlen - 3);
            if ((c1 | c2 | c3) <= 3) pbit = c1 + 4 * c2 + 16 * c3;
          }
        }
      }
    } else if (op == 2) {
      const uint32_t r1 = l - 1u;
      have1 = r1 != 0xFFFFFFFFu;
      i2 = gs::row_to_idx(u, P.X.primary);
      b2 = gs::div96(i2);
      i1 = have1 ? gs::row_to_idx(r1, P.X.primary) : 0u;
      b1 = gs::div96(i1);
      a1 = P.X.occ + (size_t)b2 * 4 + c;
      if (have1 && b1 != b2) a2 = P.X.occ + (size_t)b1 * 4 + c;
    }
    uint4 d1 = make_uint4(0, 0, 0, 0), d2;
    if (a1) d1 = *a1;
    d2 = d1;
    if (a2) d2 = *a2;

    if (op == 1) {
      const bool ext = tlen > gs::kPrecalc;
      const uint32_t il = ext ? d1.x : ((tidx & 1) ? d1.z : d1.x), iu = ext ? d1.y : ((tidx & 1) ? d1.w : d1.y);
      const bool absent = pbit < 64 && !(((pbit < 32 ? d1.z : d1.w) >> (pbit & 31)) & 1u);
      if (il > iu) {
        // absent clean tlen-base window: with k >= tlen the task can emit
        // nothing and ends (as in seed_persistent); else the 13-mer path decides
        if (ext && P.k < tlen)
          ext_miss = true;
        else if (!next_task())
          done = true;
      } else if (absent) {  // absent clean (tlen + 3)-base window, k >= tlen + 3: ends likewise
        if (!next_task()) done = true;
      } else {
        l = il;
        u = iu;
        s = u - l + 1;
        b = (uint32_t)e - tlen;
        if (prune_at((int32_t)b, s)) {
          if (!next_task()) done = true;
        } else {
          x = (int32_t)b - 1;
          phase = 1;
        }
      }
    } else if (op == 2) {
      const uint32_t o2 = gs::rank_in_entry(d1, i2 - b2 * gs::kBlockBases);
      const uint32_t o1 = have1 ? gs::rank_in_entry(d2, i1 - b1 * gs::kBlockBases) : 0u;
      const uint32_t base = P.X.L2[c];
      const uint32_t l2 = base + o1 + 1u, u2 = base + o2;
      if (l2 > u2) {
        finish();
        phase = 0;
        if (!next_task()) done = true;
      } else {
        l = l2;
        u = u2;
        s = u - l + 1;
        b = (uint32_t)x;
        if (prune_at(x, s)) {
          phase = 0;
          if (!next_task()) done = true;
        } else {
          --x;
        }
      }
    }
  }
  if (live) (strand == 0 ? P.nF : P.nR)[r] = nout;
}

// ---- finalisation ---------------------------------------------------------

struct PostParams {
  const uint4* cand;
  const uint32_t* nF;
  const uint32_t* nR;
  const uint64_t* task_off;
  const uint32_t* lens;
  const uint32_t* batch;  // reference batch of each read
  uint32_t n_reads, k, smem, max_occ;
  uint32_t* nseeds;       // count pass
  uint32_t* npos;
  const uint64_t* seed_off;  // write pass
  const uint64_t* pos_off;
  gs_seed_rec* seeds;
  uint32_t* rows;
  uint32_t long_min;          // reads longer than this go to finalize_long
  uint32_t n_long;
  const uint32_t* long_reads;
};

struct ReadLists {
  const uint4* F;  // increasing qend
  const uint4* R;  // increasing qbeg
  uint32_t nF, nR;
};

__device__ __forceinline__ ReadLists lists_of(const PostParams& P, uint32_t r) {
  ReadLists Ls;
  const uint32_t L = P.lens[r];
  const uint32_t ntask = L >= P.k ? L - P.k + 1 : 0;
  const uint4* base = P.cand + 2 * P.task_off[r];
  Ls.nF = P.nF[r];
  Ls.nR = P.nR[r];
  Ls.F = base + ntask - Ls.nF;
  Ls.R = base + ntask;
  return Ls;
}

__device__ __forceinline__ uint32_t isize(uint4 c) { return c.w - c.z + 1u; }

// MEM filter (filter_seed_intervals_gpu_mem, 1103-1139): an entry is dropped
// when its neighbour towards the longer seed has the same qbeg/qend and size.
__device__ __forceinline__ bool mem_killed(const ReadLists& Ls, int rev, uint32_t i) {
  if (!rev) {
    const uint4 a = Ls.F[i];
    if (i + 1 >= Ls.nF) return false;
    const uint4 n = Ls.F[i + 1];
    return n.x == a.x && isize(n) == isize(a);
  }
  if (i == 0) return false;
  const uint4 a = Ls.R[i], pv = Ls.R[i - 1];
  return pv.y == a.y && isize(pv) == isize(a);
}

// rows an entry contributes (count_seed_intervals_gpu_mem, 927-989; in-read)
__device__ uint32_t mem_count(const ReadLists& Ls, int rev, uint32_t i) {
  if (mem_killed(Ls, rev, i)) return 0;
  const uint4 a = rev ? Ls.R[i] : Ls.F[i];
  const uint32_t sz = isize(a);
  if (!rev) {
    for (uint32_t j = i + 1; j < Ls.nF && Ls.F[j].x == a.x; ++j)
      if (!mem_killed(Ls, 0, j)) return sz - isize(Ls.F[j]);
  } else {
    for (uint32_t j = i; j-- > 0 && Ls.R[j].y == a.y;)
      if (!mem_killed(Ls, 1, j)) return sz - isize(Ls.R[j]);
  }
  return sz;
}

// interval whose rows an entry skips (seeds_to_threads_mem, 624-667): the next
// non-empty entry with the same qbeg (forward, scanning on) or qend (reverse,
// scanning back) anywhere in the same reference batch; (1, 0) if none.
__device__ uint2 mem_skip(const PostParams& P, uint32_t r, int rev, uint32_t i) {
  const ReadLists L0 = lists_of(P, r);
  const uint4 a = rev ? L0.R[i] : L0.F[i];
  uint2 skip = make_uint2(1, 0);
  const uint32_t bid = P.batch[r];
  if (!rev) {
    uint32_t rr = r;
    int part = 0;  // 0: F, 1: R
    uint32_t j = i + 1;
    while (rr < P.n_reads && P.batch[rr] == bid) {
      const ReadLists Ls = lists_of(P, rr);
      const uint32_t n = part == 0 ? Ls.nF : Ls.nR;
      if (j < n) {
        const uint4 en = part == 0 ? Ls.F[j] : Ls.R[j];
        if (en.x != a.x) break;
        if (!mem_killed(Ls, part, j)) {
          skip = make_uint2(en.z, en.w);
          break;
        }
        ++j;
        continue;
      }
      j = 0;
      if (part == 0) {
        part = 1;
      } else {
        part = 0;
        ++rr;
      }
    }
  } else {
    int64_t rr = r;
    int part = 1;
    int64_t j = (int64_t)i - 1;
    while (rr >= 0 && P.batch[rr] == bid) {
      const ReadLists Ls = lists_of(P, (uint32_t)rr);
      if (j >= 0) {
        const uint4 en = part == 0 ? Ls.F[j] : Ls.R[j];
        if (en.y != a.y) break;
        if (!mem_killed(Ls, part, (uint32_t)j)) {
          skip = make_uint2(en.z, en.w);
          break;
        }
        --j;
        continue;
      }
      if (part == 1) {
        part = 0;
        j = (int64_t)Ls.nF - 1;
      } else {
        --rr;
        if (rr < 0 || P.batch[rr] != bid) break;
        part = 1;
        j = (int64_t)lists_of(P, (uint32_t)rr).nR - 1;
      }
    }
  }
  return skip;
}

// One lane per read.  Count pass (write == false): seeds and positions per
// read.  Write pass: seed records and suffix-array rows at the scanned offsets,
// merged by qbeg with forward entries first on ties (the stable segmented sort
// at 1233 and 3570).
// One entry of the merged lists: its seed record and rows (write pass), or
// only its counts; the seeds and positions it adds.
template <bool WRITE>
__device__ __forceinline__ void finalize_entry(const PostParams& P, uint32_t r, const ReadLists& Ls, int rev,
                                               uint32_t i, uint64_t so, uint64_t po, uint32_t& ns,
                                               uint32_t& np) {
  const uint4 en = rev ? Ls.R[i] : Ls.F[i];
  const uint32_t sz = isize(en);
  uint32_t cnt = P.smem ? sz : mem_count(Ls, rev, i);
  if (cnt == 0) return;
  if (P.max_occ && sz > P.max_occ) cnt = 0;
  if (WRITE) {
    gs_seed_rec rec;
    rec.qbeg = en.x;
    rec.qend = en.y;
    rec.l = en.z;
    rec.u = en.w;
    rec.n_pos = cnt;
    rec.pad = 0;
    P.seeds[so + ns] = rec;
    if (cnt) {
      uint2 skip = make_uint2(1, 0);
      if (!P.smem) skip = mem_skip(P, r, rev, i);
      uint32_t row = en.z, got = 0;
      while (got < cnt) {
        if (row < skip.x || row > skip.y) P.rows[po + np + got++] = row;
        ++row;
      }
    }
  }
  ++ns;
  np += cnt;
}

// the next entry of the merge: forward entries first on ties of qbeg
__device__ __forceinline__ bool take_forward(const ReadLists& Ls, uint32_t a, uint32_t z) {
  return z >= Ls.nR || (a < Ls.nF && Ls.F[a].x <= (Ls.R[z].x & 0x7FFFFFFFu));
}

template <bool WRITE>
__global__ void __launch_bounds__(256) finalize_reads(PostParams P) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= P.n_reads) return;
  if (P.lens[r] > P.long_min && P.lens[r] >= P.k) return;  // finalize_long's (the reads of long_reads)
  const ReadLists Ls = lists_of(P, r);
  uint32_t ns = 0, np = 0;
  uint64_t so = WRITE ? P.seed_off[r] : 0, po = WRITE ? P.pos_off[r] : 0;
  uint32_t a = 0, z = 0;
  while (a < Ls.nF || z < Ls.nR) {
    const bool takeF = take_forward(Ls, a, z);
    finalize_entry<WRITE>(P, r, Ls, takeF ? 0 : 1, takeF ? a++ : z++, so, po, ns, np);
  }
  if (!WRITE) {
    P.nseeds[r] = ns;
    P.npos[r] = np;
  }
}

// Reads longer than long_min bases (thousands of entries each): one wave per
// read.  Lane l takes positions [l*T/64, (l+1)*T/64) of the merged order of
// the T = nF + nR entries; its starting split between the two lists is found
// by binary search on the merge path (the number a of forward entries among
// the first d merged is the smallest a at which forward entry a does not come
// before reverse entry d-a-1).  The count pass sums the lanes' seeds and
// positions; the write pass counts first, scans the lanes' counts for their
// output offsets, then merges its part again and writes.
template <bool WRITE>
__global__ void __launch_bounds__(256) finalize_long(PostParams P) {
  const uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (w >= P.n_long) return;
  const uint32_t r = P.long_reads[w];
  const ReadLists Ls = lists_of(P, r);
  const uint32_t T = Ls.nF + Ls.nR;
  const uint32_t d0 = (uint32_t)((uint64_t)T * lane / 64), d1 = (uint32_t)((uint64_t)T * (lane + 1) / 64);
  uint32_t lo = d0 > Ls.nR ? d0 - Ls.nR : 0, hi = d0 < Ls.nF ? d0 : Ls.nF;
  while (lo < hi) {
    const uint32_t a = (lo + hi) >> 1, b = d0 - a;
    if (b > 0 && Ls.F[a].x <= (Ls.R[b - 1].x & 0x7FFFFFFFu))
      lo = a + 1;  // forward entry a is merged before reverse entry b-1
    else
      hi = a;
  }
  const uint32_t a0 = lo, z0 = d0 - lo;
  uint32_t ns = 0, np = 0;
  {
    uint32_t a = a0, z = z0;
    for (uint32_t d = d0; d < d1; ++d) {
      const bool takeF = take_forward(Ls, a, z);
      finalize_entry<false>(P, r, Ls, takeF ? 0 : 1, takeF ? a++ : z++, 0, 0, ns, np);
    }
  }
  // inclusive scans over the lanes (64-bit positions: a read's rows can
  // exceed 2^32 only through pathological counts, so sum in 64 bits)
  uint64_t sn = ns, sp = np;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t tn = __shfl_up(sn, o, 64), tp = __shfl_up(sp, o, 64);
    if ((int)lane >= o) {
      sn += tn;
      sp += tp;
    }
  }
  if (!WRITE) {
    if (lane == 63) {
      P.nseeds[r] = (uint32_t)sn;
      P.npos[r] = (uint32_t)sp;
    }
    return;
  }
  const uint64_t so = P.seed_off[r] + (sn - ns), po = P.pos_off[r] + (sp - np);
  uint32_t ws = 0, wp = 0;
  uint32_t a = a0, z = z0;
  for (uint32_t d = d0; d < d1; ++d) {
    const bool takeF = take_forward(Ls, a, z);
    finalize_entry<true>(P, r, Ls, takeF ? 0 : 1, takeF ? a++ : z++, so, po, ws, wp);
  }
}

// ---- exclusive scan of per-read counts into 64-bit offsets ---------------
// Three launches: per-workgroup totals, a single-workgroup scan of those
// totals, then per-workgroup local scans offset by the scanned totals.  Waves
// scan with shuffles; the workgroup combines its four wave totals in LDS.

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;  // per thread
constexpr uint32_t kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
  const int lane = threadIdx.x & 63;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// exclusive scan of v across the workgroup; returns the workgroup total in *tot
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* tot) {
  __shared__ uint64_t wsum[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int w = 0; w < kScanBlock / 64; ++w) {
    if (w < wid) before += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *tot = all;
  return before + inc - v;
}

__global__ void __launch_bounds__(kScanBlock) scan_tile_totals(const uint32_t* __restrict__ in, uint64_t n,
                                                               uint64_t* __restrict__ tiles) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint64_t s = 0;
  for (int q = 0; q < kScanItems; ++q)
    if (base + q < n) s += in[base + q];
  uint64_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanBlock) scan_tiles(uint64_t* __restrict__ tiles, uint64_t ntiles,
                                                         uint64_t* __restrict__ total) {
  uint64_t carry = 0;
  for (uint64_t base = 0; base < ntiles; base += kScanBlock) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t v = i < ntiles ? tiles[i] : 0;
    uint64_t tot;
    const uint64_t ex = block_excl_scan(v, &tot);
    if (i < ntiles) tiles[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kScanBlock) scan_apply(const uint32_t* __restrict__ in, uint64_t n,
                                                         const uint64_t* __restrict__ tiles,
                                                         uint64_t* __restrict__ out) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t s = 0;
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = base + q < n ? in[base + q] : 0u;
    s += v[q];
  }
  uint64_t tot;
  uint64_t run = tiles[blockIdx.x] + block_excl_scan(s, &tot);
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (base + kScanItems >= n && base < n) out[n] = run;  // the thread holding the last item
}

// ---- suffix-array walk (locate_seeds_gpu, 668-684) -------------------------
// R rows per thread, walked together so that their dependent lookups overlap;
// the rows of a thread are R grid-strides apart, so every load of a round is
// coalesced across the wave.
template <int R>
__global__ void __launch_bounds__(256) locate_rows(const DevIndex X, uint32_t* __restrict__ rows,
                                                   uint64_t n) {
  const uint64_t T = (uint64_t)gridDim.x * 256;
  const uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (X.locate_direct) {
    // one load from the full suffix array per row; the rows whose reference
    // walk detours through row primary+1 report what that walk reports
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint64_t i = t + j * T;
      if (i >= n) continue;
      const uint32_t r = rows[i];
      uint32_t pos = 0xFFFFFFFFu;  // rows past the last one (MEM row skip) and row 0 ($) have none
      if (r != 0 && r <= X.seq_len) {
        pos = X.sa_full[r];
        for (uint32_t q = 0; q < X.n_quirk; ++q)
          if (X.quirk[q].x == r) pos = X.quirk[q].y;
      }
      rows[i] = pos;
    }
    return;
  }
  uint32_t row[R], itr[R];
  uint32_t live = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint64_t i = t + j * T;
    row[j] = 0;
    itr[j] = 0;
    if (i >= n) continue;
    row[j] = rows[i];
    // Rows past the last one come only from the reference's batch-wide MEM
    // row skip (seed_gen.cu:637/645) and have no position.
    if (row[j] > X.seq_len)
      rows[i] = 0xFFFFFFFFu;
    else
      live |= 1u << j;
  }
  while (live) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (!(live & (1u << j))) continue;
      const uint32_t q = gs::div_sa(X, row[j]);
      if (row[j] == q * X.sa_intv) {
        rows[t + j * T] = X.sa[q] + itr[j];
        live &= ~(1u << j);
      } else if (itr[j] > X.seq_len) {  // a walk on a consistent index ends within seq_len + 1 steps
        rows[t + j * T] = 0xFFFFFFFFu;
        live &= ~(1u << j);
      } else {
        ++itr[j];
        row[j] = gs::lf_walk(X, row[j]);
      }
    }
  }
}

inline unsigned grid_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void zero_bytes(uint8_t* __restrict__ p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---- work items of the persistent kernel, written at load time ---------------
// Whole-read streams (reads held whole in the LDS window) come first, two
// items per read (forward, reverse) in read order; then every read cut into
// segments, segment-major: items (read | seg << 33) and (read | 1 << 32 |
// seg << 33) for seg = 0..nseg-1, with the read in long_reads and its first
// item in item0.  A whole-read stream holds up to ~150 task ends and a
// segment at most seg_tasks, so the last items a persistent kernel hands
// out are the small ones.

__device__ __forceinline__ bool read_is_long(uint32_t L, uint32_t k, uint32_t nseg) {
  return nseg > 1 || (L > gs::kPersistLmax && L >= k);
}

// per read: items as a whole-read stream, items as segments, 1 if segmented
__global__ void item_counts(const uint32_t* __restrict__ lens, uint32_t n, uint32_t k, uint32_t seg_tasks,
                            uint32_t* __restrict__ c_whole, uint32_t* __restrict__ c_seg,
                            uint32_t* __restrict__ c_long) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t L = lens[i];
  const uint32_t nseg = gs::n_segments(L, k, seg_tasks);
  const bool lng = read_is_long(L, k, nseg);
  c_whole[i] = lng ? 0u : 2u;
  c_seg[i] = lng ? 2u * nseg : 0u;
  c_long[i] = lng ? 1u : 0u;
}

// o_whole, o_seg, o_long: exclusive scans of the counts (n + 1 entries)
__global__ void item_fill(const uint32_t* __restrict__ lens, uint32_t n, uint32_t k, uint32_t seg_tasks,
                          const uint64_t* __restrict__ o_whole, const uint64_t* __restrict__ o_seg,
                          const uint64_t* __restrict__ o_long, uint64_t* __restrict__ items,
                          uint32_t* __restrict__ long_reads, uint64_t* __restrict__ item0) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t L = lens[i];
  const uint32_t nseg = gs::n_segments(L, k, seg_tasks);
  if (!read_is_long(L, k, nseg)) {
    const uint64_t p = o_whole[i];
    items[p] = i;
    items[p + 1] = i | (1ull << 32);
    return;
  }
  const uint64_t j = o_long[i], first = o_whole[n] + o_seg[i];
  long_reads[j] = i;
  item0[j] = first;
  for (uint64_t g = 0; g < nseg; ++g) {
    items[first + 2 * g] = i | (g << 33);
    items[first + 2 * g + 1] = i | (1ull << 32) | (g << 33);
  }
}

// ---- device buffers of one gs_seed call -------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  bool own = true;  // false: a view into memory this buffer must not free or replace
  void view(void* q, size_t bytes) {
    p = q;
    cap = bytes;
    own = false;
  }
  int ensure(size_t bytes) {
    if (bytes <= cap) return GS_OK;
    if (!own) return GS_ENOMEM;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = bytes + (bytes / 4 < (256u << 20) ? bytes / 4 : (256u << 20)) + 256;  // headroom for reuse
    if (hipMalloc(&p, want) != hipSuccess) return GS_ENOMEM;
    cap = want;
    return GS_OK;
  }
  template <typename T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
  ~DevBuf() {
    if (p && own) (void)hipFree(p);
  }
};

int scan_counts(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* tiles, uint64_t* total,
                hipStream_t st) {
  const uint64_t ntiles = (n + kScanTile - 1) / kScanTile;
  if (ntiles == 0) return GS_OK;
  hipLaunchKernelGGL(scan_tile_totals, dim3((unsigned)ntiles), dim3(kScanBlock), 0, st, in, n, tiles);
  hipLaunchKernelGGL(scan_tiles, dim3(1), dim3(kScanBlock), 0, st, tiles, ntiles, total);
  hipLaunchKernelGGL(scan_apply, dim3((unsigned)ntiles), dim3(kScanBlock), 0, st, in, n, tiles, out);
  return GS_OK;
}

template <typename AT, int LMAX>
void launch_seed(const SeedParams& P, hipStream_t st) {
  const uint64_t lanes = 2ull * P.n_reads;
  hipLaunchKernelGGL((seed_streams<AT, LMAX>), dim3(grid_for(lanes, kSeedBlock)), dim3(kSeedBlock), 0,
                     st, P);
}

}  // namespace

// ---- device-resident batches ------------------------------------------------

struct gs_batch {
  gs_index* X = nullptr;
  bool holds = false;  // keeps a reference to X (see batch_new)
  hipStream_t stream = nullptr;  // a batch's own stream: batches of one index may run concurrently
  gs_opts opts;
  uint64_t n_reads = 0, span = 0, tasks = 0;
  uint32_t maxlen = 0;
  DevBuf bases, offs, lens, toff, batch, cand, nF, nR, ns, np, soff, poff, tiles, tot, seeds, rows, scratch;
  DevBuf codes, code_off, queue, counters;  // persistent kernel
  DevBuf items, seg_meta, long_reads, item0, item_tmp;  // item_tmp: counts and scans of the item plan
  DevBuf seg_link, dyn_items, dyn_meta, dyn_link, dyn_ready, dyn_ctl, wmeta, wlink;  // work splitting
  uint64_t code_chunks = 0, n_items = 0;
  uint32_t n_long = 0, seg_tasks = 64, dyn_cap = 0;
  uint64_t n_seeds = 0, n_pos = 0;
  bool ran = false;
  hipEvent_t ev[7] = {};
  hipEvent_t kev[2] = {};  // around the extension kernel's own launch (t_kernel)
  double t_h2d = 0;
  uint64_t* h_tot = nullptr;  // pinned: totals of the count stage
  // captured pipeline (gs_batch_capture) and its pinned staging
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_fresh = false;
  uint64_t graph_seeds = 0, graph_pos = 0;
  char* hs_bases = nullptr;
  uint64_t* hs_offs = nullptr;
  uint32_t* hs_lens = nullptr;
  uint32_t *hr_ns = nullptr, *hr_np = nullptr, *hr_rows = nullptr;
  uint64_t *hr_soff = nullptr, *hr_poff = nullptr;
  gs_seed_rec* hr_seeds = nullptr;
  // pinned result staging of batch_download (grow-only; the server reuses it
  // round after round)
  struct Pinned {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
      if (bytes <= cap) return GS_OK;
      if (p) (void)hipHostFree(p);
      p = nullptr;
      cap = 0;
      const size_t want = bytes + bytes / 4 + 4096;
      if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        return GS_ENOMEM;
      }
      cap = want;
      return GS_OK;
    }
    void release() {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      cap = 0;
    }
  };
  Pinned dl_ns, dl_np, dl_soff, dl_poff, dl_seeds, dl_pos;
  Pinned st_a, st_b;            // staged_d2h's two chunks
  hipEvent_t cev[2] = {};       // their copies' completion
};

static void free_graph(gs_batch* B);
static void free_host_results(gs_batch* B);

void* gs::result_alloc(size_t bytes) { return malloc(bytes); }

constexpr size_t kStageChunk = 64u << 20;  // pinned chunks of staged uploads and downloads

static void destroy_batch(gs_batch* B) {
  if (!B) return;
  free_graph(B);
  free_host_results(B);
  for (auto* q : {&B->dl_ns, &B->dl_np, &B->dl_soff, &B->dl_poff, &B->dl_seeds, &B->dl_pos, &B->st_a, &B->st_b})
    q->release();
  for (auto& e : B->cev)
    if (e) (void)hipEventDestroy(e);
  if (B->h_tot) (void)hipHostFree(B->h_tot);
  for (auto& e : B->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : B->kev)
    if (e) (void)hipEventDestroy(e);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  gs_index* X = B->X;
  const bool holds = B->holds;
  delete B;
  if (holds) gs::index_release(X);
}

// ---- host plan of a batch -----------------------------------------------------
// Everything gs_batch_create derives on the CPU: the byte pieces to upload, the
// offsets rebased onto them, task offsets, reference-batch ids, code chunks and
// the persistent kernel's work items.  A plan can hold several groups of reads
// (one per request of a server round); batch ids never span two groups.

int gs::plan_check(const gs_opts* o) {
  if (!o) return GS_EINVAL;
  if (o->k < gs::kPrecalc + 1 || o->k > 1 << 20) return GS_EINVAL;
  if (o->seg_tasks > gs::kSegTasksMax) return GS_EINVAL;
  if (o->wg != 0 && o->wg != 64 && o->wg != 128 && o->wg != 256) return GS_EINVAL;
  if (o->walk_rows != 0 && o->walk_rows != 1 && o->walk_rows != 2 && o->walk_rows != 4) return GS_EINVAL;
  if (o->kernel < 0 || o->kernel > 3) return GS_EINVAL;
  if (o->split_tasks > 64 && o->split_tasks != 0xFFFF) return GS_EINVAL;
  if (o->split_work == 0xFFFF) return GS_EINVAL;
  if (o->direct_min > 64 && o->direct_min != 0xFFFF) return GS_EINVAL;
  if (o->tandem > gs::kTandemMax && o->tandem != 0xFFFF) return GS_EINVAL;
  return GS_OK;
}

void gs::plan_reset(BatchPlan& P, const gs_opts& o) {
  P = BatchPlan();
  P.opts = o;
  P.seg_tasks = o.seg_tasks ? o.seg_tasks : 64u;
}

// room for `add` more elements, growing geometrically so that many small
// additions (a server round's chunks) stay linear
template <typename V>
static void grow(V& v, size_t add) {
  const size_t need = v.size() + add;
  if (need > v.capacity()) v.reserve(std::max(need, 2 * v.capacity()));
}

int gs::plan_add(BatchPlan& P, const char* bases, const uint64_t* offsets, const uint32_t* lens, uint64_t r0,
                 uint64_t r1) {
  if (r1 <= r0) return GS_OK;
  const size_t nr = (size_t)(r1 - r0);
  grow(P.offs, nr);
  grow(P.lens, nr);
  grow(P.task_off, nr + 1);
  grow(P.rbatch, nr);
  grow(P.code_off, nr + 1);
  const uint32_t k = (uint32_t)P.opts.k;
  const uint32_t batch_bases = P.opts.batch_bases ? P.opts.batch_bases : 1000000u;
  uint64_t lo = UINT64_MAX, hi = 0;
  for (uint64_t i = r0; i < r1; ++i) {
    if (offsets[i] < lo) lo = offsets[i];
    if (offsets[i] + lens[i] > hi) hi = offsets[i] + lens[i];
  }
  if (hi < lo) lo = hi = 0;
  if (P.n_reads + (r1 - r0) > 0xFFFFFFFFull / 2) return GS_ERANGE;
  const uint64_t base = P.span;
  if (hi > lo) P.pieces.push_back({bases + lo, hi - lo});
  P.span += hi - lo;
  uint64_t acc = 0;
  uint32_t bid = P.next_batch_id;
  for (uint64_t i = r0; i < r1; ++i) {
    const uint32_t L = lens[i];
    P.offs.push_back(offsets[i] - lo + base);
    P.lens.push_back(L);
    P.task_off.push_back(P.tasks);
    P.tasks += L >= k ? L - k + 1 : 0;
    P.rbatch.push_back(bid);
    acc += L;
    if (acc >= batch_bases) {  // a batch closes after the read that reaches the limit (2955-2991)
      ++bid;
      acc = 0;
    }
    if (L > P.maxlen) P.maxlen = L;
    ++P.n_reads;
    P.code_off.push_back(P.code_chunks);
    P.code_chunks += 2 * ((L + 31) / 32);
    const uint32_t nseg = gs::n_segments(L, k, P.seg_tasks);
    if (nseg > 1 || (L > gs::kPersistLmax && L >= k)) ++P.n_long;
    P.n_items += 2ull * nseg;
  }
  P.next_batch_id = bid + 1;  // the next group starts a batch of its own
  if (P.n_items > 0xF0000000ull) return GS_ERANGE;
  return GS_OK;
}

// Closes the per-read offset arrays; the work items (whole-read streams
// first, then the segments of long reads, so that the last items a
// persistent kernel hands out are small) are written on the device by
// batch_load.
void gs::plan_finish(BatchPlan& P) {
  P.task_off.push_back(P.tasks);
  P.code_off.push_back(P.code_chunks);
}

uint64_t gs::plan_device_bytes(const BatchPlan& P) {
  const uint64_t nr1 = P.n_reads + 1;
  return (P.span + 32) + 16 * (P.code_chunks + 1) + 16 * (2 * P.tasks + 1) + 8 * P.n_items +
         (P.n_long ? 16 * P.n_items : 0) + 12 * P.n_long + nr1 * (8 * 5 + 4 * 7 + 36) +
         (P.maxlen > 256 ? 4 * (P.span + 16) : 0);
}

static int batch_events(gs_batch* B) {
  if (!B->stream && hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) return GS_EHIP;
  for (auto& e : B->ev)
    if (!e && hipEventCreate(&e) != hipSuccess) return GS_EHIP;
  for (auto& e : B->kev)
    if (!e && hipEventCreate(&e) != hipSuccess) return GS_EHIP;
  if (!B->h_tot && hipHostMalloc((void**)&B->h_tot, 16, hipHostMallocDefault) != hipSuccess) return GS_ENOMEM;
  return GS_OK;
}

// Device buffers of a batch of `span` bases, n reads, `tasks` task slots per
// strand, `code_chunks` 16-byte code chunks, n_items work items (n_long of
// them long reads), reads of at most maxlen bases and an overflow queue of
// dyn_cap pieces; buffers only grow.
static int ensure_buffers(gs_batch* B, uint64_t span, uint64_t n, uint64_t tasks, uint64_t code_chunks,
                          uint64_t n_items, uint64_t n_long, uint32_t maxlen, uint32_t dyn_cap) {
  const uint64_t nr1 = n + 1;
  if (B->codes.ensure(16 * (code_chunks + 1)) || B->code_off.ensure(8 * nr1) ||
      B->items.ensure(8 * (n_items + 1)) || B->long_reads.ensure(4 * (n_long + 1)) ||
      B->item_tmp.ensure(36 * nr1 + 8) ||
      B->item0.ensure(8 * (n_long + 1)) || B->seg_meta.ensure(n_long ? 16 * n_items : 16) ||
      B->bases.ensure(span + 32) || B->offs.ensure(8 * nr1) || B->lens.ensure(4 * nr1) ||
      B->toff.ensure(8 * nr1) || B->batch.ensure(4 * nr1) || B->cand.ensure(16 * (2 * tasks + 1)) ||
      B->nF.ensure(4 * nr1) || B->nR.ensure(4 * nr1) || B->ns.ensure(4 * nr1) || B->np.ensure(4 * nr1) ||
      B->soff.ensure(8 * nr1) || B->poff.ensure(8 * nr1) || B->tiles.ensure(8 * (nr1 / kScanTile + 2)) ||
      B->tot.ensure(16) || (maxlen > 256 && B->scratch.ensure(4 * (span + 16))) ||
      B->queue.ensure(16) || B->counters.ensure(gs::kCounterBytes + gs::kColdBytes) ||
      B->seg_link.ensure(n_long ? 4 * n_items : 16) ||
      (dyn_cap &&
       (B->dyn_items.ensure(16ull * dyn_cap) || B->dyn_meta.ensure(16ull * dyn_cap) ||
        B->dyn_link.ensure(4ull * dyn_cap) || B->dyn_ready.ensure(4ull * dyn_cap) ||
        B->dyn_ctl.ensure(4ull * gs::kDynCtlWords) || B->wmeta.ensure(32 * nr1) || B->wlink.ensure(8 * nr1))))
    return GS_ENOMEM;
  return GS_OK;
}

int gs::batch_reserve(gs_batch* B, uint64_t bases, uint64_t reads, uint32_t seg_tasks) {
  if (!B) return GS_EINVAL;
  if (hipSetDevice(B->X->device) != hipSuccess) return GS_EHIP;
  if (!seg_tasks) seg_tasks = 64;
  // every read and long-read segment is an item of each strand; a read has
  // at most as many tasks as bases and code chunks as 32-base spans (+1)
  const uint64_t items = 2 * (reads + bases / seg_tasks + 1);
  const uint32_t dyn = (uint32_t)std::min<uint64_t>(16u << 20, std::max<uint64_t>(65536, items));
  int rc = ensure_buffers(B, bases, reads, bases, 2 * (bases / 32 + reads), items, reads, 0xFFFFFFFFu, dyn);
  if (!rc && bases > (4u << 20)) {  // the pinned chunks of staged uploads
    const size_t ch = std::min<uint64_t>(bases, kStageChunk);
    if (B->st_a.ensure(ch) || B->st_b.ensure(ch)) rc = GS_ENOMEM;
  }
  return rc;
}

// Copies len bytes to dst by calling part(a, z) for disjoint ranges
// [a, z) of [0, len) on up to 8 threads (the calling thread takes the first
// part, and every part no thread could be started for).
template <typename F>
static void parallel_parts(size_t len, F&& part) {
  constexpr size_t kThreads = 8, kMinPart = 1u << 20;
  const size_t nt = std::max<size_t>(1, std::min(kThreads, len / kMinPart));
  const size_t each = (len + nt - 1) / nt;
  std::vector<std::thread> th;
  size_t done = std::min(len, each);  // bytes covered by this thread or started helpers
  try {
    for (size_t t = 1; t < nt; ++t) {
      const size_t a = std::min(len, t * each), z = std::min(len, a + each);
      th.emplace_back([=, &part] { part(a, z); });
      done = z;
    }
  } catch (...) {  // no thread to spare: this one copies the rest
  }
  part(0, std::min(len, each));
  if (done < len) part(done, len);
  for (auto& t : th) t.join();
}

// The batch's read bytes (the plan's pieces, back to back) to dst through two
// pinned chunks: the bytes of chunk c+1 are gathered on several threads while
// chunk c's DMA runs.  Pageable sources otherwise go through the runtime's
// own staging, one thread at a fraction of the link's rate.
static int staged_upload(gs_batch* B, char* dst, const std::vector<std::pair<const char*, uint64_t>>& pieces,
                         uint64_t total) {
  constexpr size_t kChunk = kStageChunk;
  hipStream_t st = B->stream;
  if (total <= (4u << 20)) {  // small: the pieces directly
    uint64_t at = 0;
    for (const auto& pc : pieces) {
      if (hipMemcpyAsync(dst + at, pc.first, pc.second, hipMemcpyHostToDevice, st) != hipSuccess) return GS_EHIP;
      at += pc.second;
    }
    return GS_OK;
  }
  const size_t ch = std::min<uint64_t>(total, kChunk);
  if (B->st_a.ensure(ch) || B->st_b.ensure(ch)) return GS_ENOMEM;
  for (auto& e : B->cev)
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return GS_EHIP;
  std::vector<uint64_t> start(pieces.size() + 1, 0);  // offset of each piece in the batch
  for (size_t i = 0; i < pieces.size(); ++i) start[i + 1] = start[i] + pieces[i].second;
  char* stage[2] = {static_cast<char*>(B->st_a.p), static_cast<char*>(B->st_b.p)};
  const size_t nch = (total + ch - 1) / ch;
  for (size_t c = 0; c < nch; ++c) {
    // the buffer of chunk c held chunk c-2, whose DMA must be done
    if (c >= 2 && hipEventSynchronize(B->cev[c & 1]) != hipSuccess) return GS_EHIP;
    const uint64_t off = c * ch, len = std::min<uint64_t>(ch, total - off);
    char* sp = stage[c & 1];
    parallel_parts(len, [&](size_t a, size_t z) {
      uint64_t g = off + a;  // batch offset of the next byte to copy
      size_t i = std::upper_bound(start.begin(), start.end(), g) - start.begin() - 1;
      while (g < off + z) {
        const uint64_t in = g - start[i], n = std::min(pieces[i].second - in, off + z - g);
        memcpy(sp + (g - off), pieces[i].first + in, n);
        g += n;
        ++i;
      }
    });
    if (hipMemcpyAsync(dst + off, sp, len, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipEventRecord(B->cev[c & 1], st) != hipSuccess)
      return GS_EHIP;
  }
  return GS_OK;
}

// Uploads a finished plan into B, reusing B's device buffers where they are
// large enough.  B's results and any captured graph are dropped.
int gs::batch_load(gs_batch* B, const BatchPlan& P) {
  gs_index* X = B->X;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  int rc = batch_events(B);
  if (rc) return rc;
  free_graph(B);
  B->opts = P.opts;
  B->n_reads = P.n_reads;
  B->span = P.span;
  B->tasks = P.tasks;
  B->maxlen = P.maxlen;
  B->code_chunks = P.code_chunks;
  B->n_items = P.n_items;
  B->n_long = (uint32_t)P.n_long;
  B->seg_tasks = P.seg_tasks;
  B->n_seeds = B->n_pos = 0;
  B->ran = false;
  const uint64_t n = P.n_reads, nr1 = n + 1;
  // Work splitting pays where the tail of a launch is a large share of it:
  // batches of at most kSplitItemsPerLane work items per lane of the
  // persistent grid (in larger batches the tail is amortised and keeping
  // the overflow queue live costs more than it saves).  An explicit
  // split_tasks turns it on whatever the batch size.
  bool split_on = P.opts.split_tasks != 0xFFFF && P.opts.kernel != 0;
  if (split_on && P.opts.split_tasks == 0) {
    unsigned grid = 0;
    const int wg = P.opts.wg ? P.opts.wg : 128;
    if (gs::persistent_grid(P.opts.kernel >= 1 ? P.opts.kernel - 1 : 0, wg, &grid) != GS_OK || !grid ||
        B->n_items > (uint64_t)gs::kSplitItemsPerLane * grid * (uint64_t)wg)
      split_on = false;
  }
  B->dyn_cap = split_on ? (uint32_t)std::min<uint64_t>(16u << 20, std::max<uint64_t>(65536, B->n_items)) : 0u;
  if (ensure_buffers(B, B->span, n, B->tasks, B->code_chunks, B->n_items, B->n_long, B->maxlen, B->dyn_cap))
    return GS_ENOMEM;
  hipStream_t st = B->stream;
  if (hipEventRecord(B->ev[0], st) != hipSuccess) return GS_EHIP;
  if ((rc = staged_upload(B, B->bases.as<char>(), P.pieces, P.span))) return rc;
  hipLaunchKernelGGL(zero_bytes, dim3(1), dim3(32), 0, st, B->bases.as<uint8_t>() + B->span, 32u);
  if (n &&
      (hipMemcpyAsync(B->offs.p, P.offs.data(), 8 * n, hipMemcpyHostToDevice, st) != hipSuccess ||
       hipMemcpyAsync(B->lens.p, P.lens.data(), 4 * n, hipMemcpyHostToDevice, st) != hipSuccess ||
       hipMemcpyAsync(B->toff.p, P.task_off.data(), 8 * nr1, hipMemcpyHostToDevice, st) != hipSuccess ||
       hipMemcpyAsync(B->batch.p, P.rbatch.data(), 4 * n, hipMemcpyHostToDevice, st) != hipSuccess ||
       hipMemcpyAsync(B->code_off.p, P.code_off.data(), 8 * nr1, hipMemcpyHostToDevice, st) != hipSuccess))
    return GS_EHIP;
  if (n) {  // the work items, from the lengths just uploaded
    uint32_t* cw = B->item_tmp.as<uint32_t>();
    uint32_t *cs = cw + nr1, *cl = cw + 2 * nr1;
    uint64_t* ow = reinterpret_cast<uint64_t*>(B->item_tmp.as<uint8_t>() + ((12 * nr1 + 7) & ~7ull));
    uint64_t *os = ow + nr1, *ol = ow + 2 * nr1;
    const uint32_t k = (uint32_t)P.opts.k;
    hipLaunchKernelGGL(item_counts, dim3(grid_for(n, 256)), dim3(256), 0, st, B->lens.as<uint32_t>(), (uint32_t)n,
                       k, B->seg_tasks, cw, cs, cl);
    scan_counts(cw, n, ow, B->tiles.as<uint64_t>(), B->tot.as<uint64_t>(), st);
    scan_counts(cs, n, os, B->tiles.as<uint64_t>(), B->tot.as<uint64_t>(), st);
    scan_counts(cl, n, ol, B->tiles.as<uint64_t>(), B->tot.as<uint64_t>(), st);
    hipLaunchKernelGGL(item_fill, dim3(grid_for(n, 256)), dim3(256), 0, st, B->lens.as<uint32_t>(), (uint32_t)n, k,
                       B->seg_tasks, ow, os, ol, B->items.as<uint64_t>(), B->long_reads.as<uint32_t>(),
                       B->item0.as<uint64_t>());
    if (hipGetLastError() != hipSuccess) return GS_EHIP;
  }
  if (hipEventRecord(B->ev[1], st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return GS_EHIP;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, B->ev[0], B->ev[1]);
  B->t_h2d = ms * 1e-3;
  return GS_OK;
}

gs_batch* gs::batch_new(gs_index* X, bool hold) {
  gs_batch* B = new gs_batch();
  B->X = X;
  B->holds = hold;
  if (hold) gs::index_acquire(X);
  return B;
}

void gs::batch_free(gs_batch* B) { destroy_batch(B); }

extern "C" GS_API int gs_batch_create(gs_index* X, const char* bases, const uint64_t* offsets,
                                      const uint32_t* lens, uint64_t n_reads, const gs_opts* opts,
                                      gs_batch** out) {
  if (!X || !opts || !out || (n_reads && (!bases || !offsets || !lens))) return GS_EINVAL;
  *out = nullptr;
  int rc = gs::plan_check(opts);
  if (rc) return rc;
  if (n_reads > 0xFFFFFFFFull / 2) return GS_ERANGE;
  gs::BatchPlan P;
  gs::plan_reset(P, *opts);
  if ((rc = gs::plan_add(P, bases, offsets, lens, 0, n_reads))) return rc;
  gs::plan_finish(P);
  gs_batch* B = gs::batch_new(X);
  if ((rc = gs::batch_load(B, P))) {
    destroy_batch(B);
    return rc;
  }
  *out = B;
  return GS_OK;
}

// ---- the device pipeline of a batch -----------------------------------------
// Three stages on the index's stream: extension (pack, persistent kernel,
// segment join; or the plain kernel), counts (per-read counts, scans, totals
// to pinned host memory), records (seed records, SA rows, SA walk).  An eager
// run waits for the totals between counts and records to size the outputs; a
// captured graph replays all three with the sizes of the run before it.

static void fill_params(gs_batch* B, SeedParams& S, PostParams& Q) {
  gs_index* X = B->X;
  S.X = X->dev();
  S.bases = B->bases.as<uint8_t>();
  S.offs = B->offs.as<uint64_t>();
  S.lens = B->lens.as<uint32_t>();
  S.task_off = B->toff.as<uint64_t>();
  S.n_reads = (uint32_t)B->n_reads;
  S.k = (uint32_t)B->opts.k;
  S.smem = B->opts.smem ? 1 : 0;
  S.pad = 0;
  S.cand = B->cand.as<uint4>();
  S.nF = B->nF.as<uint32_t>();
  S.nR = B->nR.as<uint32_t>();
  S.scratch = B->scratch.p;
  Q.cand = S.cand;
  Q.nF = S.nF;
  Q.nR = S.nR;
  Q.task_off = S.task_off;
  Q.lens = S.lens;
  Q.batch = B->batch.as<uint32_t>();
  Q.n_reads = S.n_reads;
  Q.k = S.k;
  Q.smem = S.smem;
  Q.max_occ = B->opts.max_occ;
  Q.nseeds = B->ns.as<uint32_t>();
  Q.npos = B->np.as<uint32_t>();
  Q.seed_off = B->soff.as<uint64_t>();
  Q.pos_off = B->poff.as<uint64_t>();
  Q.seeds = B->seeds.as<gs_seed_rec>();
  Q.rows = B->rows.as<uint32_t>();
  // reads longer than the persistent kernel's window are finalized a wave each
  Q.long_min = B->n_long ? gs::kPersistLmax : 0xFFFFFFFFu;
  Q.n_long = B->n_long;
  Q.long_reads = B->long_reads.as<uint32_t>();
}

// kev (null under graph capture): two events recorded on st right before and
// after the extension kernel itself, without the code packing, the queue's
// set-up and the join of split streams around it
static int enqueue_extension(gs_batch* B, const SeedParams& S, int kernel, unsigned pgrid, hipStream_t st,
                             hipEvent_t* kev = nullptr) {
  if (!B->n_reads) return GS_OK;
  if (kernel == 0) {
    if (kev && hipEventRecord(kev[0], st) != hipSuccess) return GS_EHIP;
    if (B->maxlen <= 160)
      launch_seed<uint16_t, 160>(S, st);
    else
      launch_seed<uint16_t, 256>(S, st);
    if (hipGetLastError() != hipSuccess) return GS_EHIP;
    return kev && hipEventRecord(kev[1], st) != hipSuccess ? GS_EHIP : GS_OK;
  }
  gs::PersistParams PP;
  PP.X = S.X;
  PP.codes = B->codes.as<uint4>();
  PP.code_off = B->code_off.as<uint64_t>();
  PP.lens = S.lens;
  PP.task_off = S.task_off;
  PP.items = B->items.as<uint64_t>();
  PP.n_items = B->n_items;
  PP.n_reads = S.n_reads;
  PP.k = S.k;
  PP.smem = S.smem;
  PP.seg_tasks = B->seg_tasks;
  PP.cand = S.cand;
  PP.nF = S.nF;
  PP.nR = S.nR;
  PP.seg_meta = B->seg_meta.as<uint4>();
  PP.seg_link = B->seg_link.as<uint32_t>();
  PP.queue = B->queue.as<uint32_t>();
  PP.counters = B->counters.as<uint64_t>();
  PP.direct_min = !S.X.isa || B->opts.direct_min == 0xFFFF ? 0u : (B->opts.direct_min ? B->opts.direct_min : 4u);
  PP.dyn_items = B->dyn_items.as<uint4>();
  PP.dyn_meta = B->dyn_meta.as<uint4>();
  PP.dyn_link = B->dyn_link.as<uint32_t>();
  PP.dyn_ready = B->dyn_ready.as<uint32_t>();
  PP.dyn_ctl = B->dyn_ctl.as<uint32_t>();
  PP.wmeta = B->wmeta.as<uint4>();
  PP.wlink = B->wlink.as<uint32_t>();
  PP.dyn_cap = B->dyn_cap;
  PP.split_tasks = B->opts.split_tasks ? B->opts.split_tasks : 8u;
  PP.split_work = 64u * (B->opts.split_work ? B->opts.split_work : 32u);
  PP.epoch = 1;  // dyn_ready is cleared by launch_persistent before every launch
  // tandem jumps read the index's tandem table
  const bool tandem = B->opts.tandem != 0xFFFF && B->X->tandem;
  PP.tandem = tandem ? (B->opts.tandem ? std::min<uint32_t>(B->opts.tandem, gs::kTandemMax) : gs::kTandemMax) : 0u;
  gs::JoinParams J;
  J.long_reads = B->long_reads.as<uint32_t>();
  J.item0 = B->item0.as<uint64_t>();
  J.lens = S.lens;
  J.task_off = S.task_off;
  J.seg_meta = PP.seg_meta;
  J.cand = S.cand;
  J.nF = S.nF;
  J.nR = S.nR;
  J.n_long = B->n_long;
  J.k = S.k;
  J.smem = S.smem;
  J.seg_tasks = B->seg_tasks;
  J.seg_link = PP.seg_link;
  J.dyn_items = PP.dyn_items;
  J.dyn_meta = PP.dyn_meta;
  J.dyn_link = PP.dyn_link;
  J.wmeta = PP.wmeta;
  J.wlink = PP.wlink;
  J.n_reads = S.n_reads;
  J.dyn_cap = PP.dyn_cap;
  J.split_tasks = PP.split_tasks;
  J.pad = 0;
  int rc = gs::launch_pack_codes(S.bases, S.offs, S.lens, PP.code_off, S.n_reads, J.long_reads, B->n_long,
                                 B->codes.as<uint4>(), st);
  if (!rc) rc = gs::launch_persistent(PP, kernel - 1, B->opts.wg, pgrid, st, kev ? kev[0] : nullptr,
                                      kev ? kev[1] : nullptr);
  if (!rc) rc = gs::launch_join(J, st);
  return rc;
}

static int enqueue_counts(gs_batch* B, const PostParams& Q, hipStream_t st) {
  const uint64_t n = B->n_reads;
  if (n) {
    hipLaunchKernelGGL(finalize_reads<false>, dim3(grid_for(n, 256)), dim3(256), 0, st, Q);
    if (Q.n_long) hipLaunchKernelGGL(finalize_long<false>, dim3((Q.n_long + 3) / 4), dim3(256), 0, st, Q);
    scan_counts(Q.nseeds, n, B->soff.as<uint64_t>(), B->tiles.as<uint64_t>(), B->tot.as<uint64_t>(), st);
    scan_counts(Q.npos, n, B->poff.as<uint64_t>(), B->tiles.as<uint64_t>(), B->tot.as<uint64_t>() + 1, st);
    GS_HIP(hipMemcpyAsync(B->h_tot, B->tot.p, 16, hipMemcpyDeviceToHost, st));
  } else {  // no reads: the offset arrays still hold their closing 0
    hipLaunchKernelGGL(zero_bytes, dim3(1), dim3(32), 0, st, B->soff.as<uint8_t>(), 8u);
    hipLaunchKernelGGL(zero_bytes, dim3(1), dim3(32), 0, st, B->poff.as<uint8_t>(), 8u);
  }
  return hipGetLastError() == hipSuccess ? GS_OK : GS_EHIP;
}

static int enqueue_write(gs_batch* B, const PostParams& Q, hipStream_t st) {
  if (B->n_reads) hipLaunchKernelGGL(finalize_reads<true>, dim3(grid_for(B->n_reads, 256)), dim3(256), 0, st, Q);
  if (Q.n_long) hipLaunchKernelGGL(finalize_long<true>, dim3((Q.n_long + 3) / 4), dim3(256), 0, st, Q);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_EHIP;
}

// locate_rows over rows[0, n), R rows per thread, in launches of at most
// launch_chunk() rows (a dispatch's work-item count is a 32-bit quantity).
static void launch_locate(const DevIndex& X, uint32_t* rows, uint64_t n, unsigned R, hipStream_t st) {
  const uint64_t step = gs::launch_chunk();
  for (uint64_t first = 0; first < n; first += step) {
    const uint64_t m = n - first < step ? n - first : step;
    const unsigned grid = grid_for((m + R - 1) / R, 256);
    if (R == 4)
      hipLaunchKernelGGL(locate_rows<4>, dim3(grid), dim3(256), 0, st, X, rows + first, m);
    else if (R == 2)
      hipLaunchKernelGGL(locate_rows<2>, dim3(grid), dim3(256), 0, st, X, rows + first, m);
    else
      hipLaunchKernelGGL(locate_rows<1>, dim3(grid), dim3(256), 0, st, X, rows + first, m);
  }
}

static int enqueue_walk(gs_batch* B, const SeedParams& S, const PostParams& Q, hipStream_t st) {
  if (B->n_pos) launch_locate(S.X, Q.rows, B->n_pos, B->opts.walk_rows ? B->opts.walk_rows : 1u, st);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_EHIP;
}

static int persistent_grid_for(gs_batch* B, int kernel, unsigned* grid) {
  *grid = 0;
  if (kernel == 0 || !B->n_reads) return GS_OK;
  return gs::persistent_grid(kernel - 1, B->opts.wg, grid);
}

// Runs the whole device pipeline on an uploaded batch; results stay in HBM.
extern "C" GS_API int gs_batch_run(gs_batch* B, int kernel_override, gs_batch_stats* stats) {
  if (!B) return GS_EINVAL;
  gs_index* X = B->X;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  hipStream_t st = B->stream;
  const int kernel = kernel_override >= 0 ? kernel_override : B->opts.kernel;
  if (kernel < 0 || kernel > 3) return GS_EINVAL;
  unsigned pgrid = 0;
  int rc = persistent_grid_for(B, kernel, &pgrid);
  if (rc) return rc;
  SeedParams S;
  PostParams Q;
  fill_params(B, S, Q);
  B->h_tot[0] = B->h_tot[1] = 0;
  GS_HIP(hipEventRecord(B->ev[0], st));
  rc = enqueue_extension(B, S, kernel, pgrid, st, B->kev);
  if (rc) return rc;
  GS_HIP(hipEventRecord(B->ev[1], st));
  rc = enqueue_counts(B, Q, st);
  if (rc) return rc;
  GS_HIP(hipEventRecord(B->ev[2], st));
  GS_HIP(hipStreamSynchronize(st));
  B->n_seeds = B->h_tot[0];
  B->n_pos = B->h_tot[1];
  const void *seeds0 = B->seeds.p, *rows0 = B->rows.p;
  if (B->seeds.ensure(sizeof(gs_seed_rec) * (B->n_seeds + 1)) || B->rows.ensure(4 * (B->n_pos + 1)))
    return GS_ENOMEM;
  if (B->graph && (B->seeds.p != seeds0 || B->rows.p != rows0)) free_graph(B);  // it wrote the old buffers
  fill_params(B, S, Q);
  GS_HIP(hipEventRecord(B->ev[3], st));
  if ((rc = enqueue_write(B, Q, st))) return rc;
  GS_HIP(hipEventRecord(B->ev[4], st));
  if ((rc = enqueue_walk(B, S, Q, st))) return rc;
  GS_HIP(hipEventRecord(B->ev[5], st));
  GS_HIP(hipStreamSynchronize(st));
  B->ran = true;
  B->graph_fresh = false;
  if (stats) {
    float ms[5] = {0, 0, 0, 0, 0};
    (void)hipEventElapsedTime(&ms[0], B->ev[0], B->ev[1]);
    (void)hipEventElapsedTime(&ms[1], B->ev[1], B->ev[2]);
    (void)hipEventElapsedTime(&ms[2], B->ev[3], B->ev[4]);
    (void)hipEventElapsedTime(&ms[3], B->ev[4], B->ev[5]);
    (void)hipEventElapsedTime(&ms[4], B->ev[0], B->ev[5]);
    memset(stats, 0, sizeof(*stats));
    stats->t_h2d = B->t_h2d;
    stats->t_seed = ms[0] * 1e-3;
    stats->t_count = ms[1] * 1e-3;
    stats->t_write = ms[2] * 1e-3;
    stats->t_locate = ms[3] * 1e-3;
    stats->t_total = ms[4] * 1e-3;
    float kms = 0;
    if (B->n_reads && hipEventElapsedTime(&kms, B->kev[0], B->kev[1]) == hipSuccess) stats->t_kernel = kms * 1e-3;
    stats->n_reads = B->n_reads;
    stats->n_seeds = B->n_seeds;
    stats->n_pos = B->n_pos;
    if (kernel >= 1) {
      uint64_t c[4] = {0, 0, 0, 0};
      if (hipMemcpy(c, B->counters.p, 32, hipMemcpyDeviceToHost) == hipSuccess) {
        stats->n_table_reads = c[0];
        stats->n_occ_reads = c[1];
        stats->n_direct_reads = c[2];
        stats->n_tandem_reads = c[3];
      }
    }
  }
  return GS_OK;
}

static void free_graph(gs_batch* B) {
  if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->gexec = nullptr;
  B->graph = nullptr;
  B->graph_fresh = false;
}

static void free_host_results(gs_batch* B) {
  for (void* p : {(void*)B->hr_ns, (void*)B->hr_np, (void*)B->hr_soff, (void*)B->hr_poff, (void*)B->hr_seeds,
                  (void*)B->hr_rows, (void*)B->hs_bases, (void*)B->hs_offs, (void*)B->hs_lens})
    if (p) (void)hipHostFree(p);
  B->hr_ns = B->hr_np = nullptr;
  B->hr_soff = B->hr_poff = nullptr;
  B->hr_seeds = nullptr;
  B->hr_rows = nullptr;
  B->hs_bases = nullptr;
  B->hs_offs = nullptr;
  B->hs_lens = nullptr;
}

// Captures upload -> extension -> counts -> records -> download as one graph.
extern "C" GS_API int gs_batch_capture(gs_batch* B, int kernel) {
  if (!B || !B->ran) return GS_EINVAL;
  const int kern = kernel >= 0 ? kernel : B->opts.kernel;
  if (kern < 0 || kern > 3) return GS_EINVAL;
  gs_index* X = B->X;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  hipStream_t st = B->stream;
  free_graph(B);
  free_host_results(B);
  const uint64_t n = B->n_reads, nr1 = n + 1;
  // pinned staging of the inputs and of the results, sized by the last run
  if (hipHostMalloc((void**)&B->hs_bases, B->span + 1, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&B->hs_offs, 8 * nr1, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&B->hs_lens, 4 * nr1, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&B->hr_ns, 4 * nr1, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&B->hr_np, 4 * nr1, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&B->hr_soff, 8 * nr1, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&B->hr_poff, 8 * nr1, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&B->hr_seeds, sizeof(gs_seed_rec) * (B->n_seeds + 1), hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&B->hr_rows, 4 * (B->n_pos + 1), hipHostMallocDefault) != hipSuccess) {
    free_host_results(B);
    return GS_ENOMEM;
  }
  // the graph's upload nodes re-send the batch's resident inputs from these
  if ((B->span && hipMemcpyAsync(B->hs_bases, B->bases.p, B->span, hipMemcpyDeviceToHost, st) != hipSuccess) ||
      (n && (hipMemcpyAsync(B->hs_offs, B->offs.p, 8 * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipMemcpyAsync(B->hs_lens, B->lens.p, 4 * n, hipMemcpyDeviceToHost, st) != hipSuccess)) ||
      hipStreamSynchronize(st) != hipSuccess) {
    free_host_results(B);
    return GS_EHIP;
  }
  unsigned pgrid = 0;
  int rc = persistent_grid_for(B, kern, &pgrid);
  if (rc) return rc;
  SeedParams S;
  PostParams Q;
  fill_params(B, S, Q);
  GS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
  do {
    if ((B->span && hipMemcpyAsync(B->bases.p, B->hs_bases, B->span, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (n && (hipMemcpyAsync(B->offs.p, B->hs_offs, 8 * n, hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(B->lens.p, B->hs_lens, 4 * n, hipMemcpyHostToDevice, st) != hipSuccess))) {
      rc = GS_EHIP;
      break;
    }
    if ((rc = enqueue_extension(B, S, kern, pgrid, st))) break;
    if ((rc = enqueue_counts(B, Q, st))) break;
    if ((rc = enqueue_write(B, Q, st))) break;
    if ((rc = enqueue_walk(B, S, Q, st))) break;
    if ((n && (hipMemcpyAsync(B->hr_ns, B->ns.p, 4 * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
               hipMemcpyAsync(B->hr_np, B->np.p, 4 * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
               hipMemcpyAsync(B->hr_soff, B->soff.p, 8 * nr1, hipMemcpyDeviceToHost, st) != hipSuccess ||
               hipMemcpyAsync(B->hr_poff, B->poff.p, 8 * nr1, hipMemcpyDeviceToHost, st) != hipSuccess)) ||
        (B->n_seeds && hipMemcpyAsync(B->hr_seeds, B->seeds.p, sizeof(gs_seed_rec) * B->n_seeds,
                                      hipMemcpyDeviceToHost, st) != hipSuccess) ||
        (B->n_pos && hipMemcpyAsync(B->hr_rows, B->rows.p, 4 * B->n_pos, hipMemcpyDeviceToHost, st) != hipSuccess))
      rc = GS_EHIP;
  } while (0);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(st, &g);
  if (rc || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : GS_EHIP;
  }
  B->graph = g;
  if (hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
    free_graph(B);
    return GS_EHIP;
  }
  B->graph_seeds = B->n_seeds;
  B->graph_pos = B->n_pos;
  return GS_OK;
}

extern "C" GS_API int gs_batch_replay(gs_batch* B, gs_batch_stats* stats) {
  if (!B || !B->gexec) return GS_EINVAL;
  gs_index* X = B->X;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  hipStream_t st = B->stream;
  B->h_tot[0] = B->h_tot[1] = 0;
  GS_HIP(hipEventRecord(B->ev[0], st));
  GS_HIP(hipGraphLaunch(B->gexec, st));
  GS_HIP(hipEventRecord(B->ev[5], st));
  GS_HIP(hipStreamSynchronize(st));
  // the totals the graph copied back must be the sizes it was captured with
  if (B->n_reads && (B->h_tot[0] != B->graph_seeds || B->h_tot[1] != B->graph_pos)) return GS_ERANGE;
  B->n_seeds = B->graph_seeds;
  B->n_pos = B->graph_pos;
  B->graph_fresh = true;
  if (stats) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, B->ev[0], B->ev[5]);
    memset(stats, 0, sizeof(*stats));
    stats->t_total = ms * 1e-3;
    stats->n_reads = B->n_reads;
    stats->n_seeds = B->n_seeds;
    stats->n_pos = B->n_pos;
  }
  return GS_OK;
}

// Device-to-host copy of `bytes` into pageable memory through two pinned
// chunks: the DMA of chunk c+1 runs while chunk c is copied out, on several
// threads when it is large.  Pageable destinations otherwise go through the
// runtime's own staging at a fraction of the link's rate.
static int staged_d2h(gs_batch* B, void* dst, const void* src, size_t bytes) {
  constexpr size_t kChunk = kStageChunk;
  hipStream_t st = B->stream;
  if (bytes <= (4u << 20))  // small: one plain copy
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipStreamSynchronize(st) == hipSuccess
               ? GS_OK
               : GS_EHIP;
  const size_t ch = std::min(bytes, kChunk);
  if (B->st_a.ensure(ch) || B->st_b.ensure(ch)) return GS_ENOMEM;
  for (auto& e : B->cev)
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return GS_EHIP;
  char* stage[2] = {static_cast<char*>(B->st_a.p), static_cast<char*>(B->st_b.p)};
  const size_t nch = (bytes + ch - 1) / ch;
  auto issue = [&](size_t c) -> bool {
    const size_t off = c * ch, len = std::min(ch, bytes - off);
    return hipMemcpyAsync(stage[c & 1], static_cast<const char*>(src) + off, len, hipMemcpyDeviceToHost, st) ==
               hipSuccess &&
           hipEventRecord(B->cev[c & 1], st) == hipSuccess;
  };
  if (!issue(0)) return GS_EHIP;
  for (size_t c = 0; c < nch; ++c) {
    // the buffer of chunk c+1 held chunk c-1, already copied out
    if (c + 1 < nch && !issue(c + 1)) return GS_EHIP;
    if (hipEventSynchronize(B->cev[c & 1]) != hipSuccess) return GS_EHIP;
    const size_t off = c * ch, len = std::min(ch, bytes - off);
    char* d = static_cast<char*>(dst) + off;
    const char* sp = stage[c & 1];
    constexpr size_t kThreads = 8;
    const size_t part = (len + kThreads - 1) / kThreads;
    std::vector<std::thread> th;
    size_t done = part;  // bytes the helper threads do not cover
    try {
      for (size_t t = 1; t < kThreads; ++t) {
        const size_t a = std::min(len, t * part), z = std::min(len, a + part);
        th.emplace_back([=] { memcpy(d + a, sp + a, z - a); });
        done = std::min(len, (t + 1) * part);
      }
    } catch (...) {  // no thread to spare: this one copies the rest
    }
    memcpy(d, sp, std::min(len, part));
    if (done < len) memcpy(d + done, sp + done, len - done);
    for (auto& t : th) t.join();
  }
  return GS_OK;
}

extern "C" GS_API int gs_batch_fetch(gs_batch* B, gs_result* out) {
  if (!B || !out || !B->ran) return GS_EINVAL;
  memset(out, 0, sizeof(*out));
  if (hipSetDevice(B->X->device) != hipSuccess) return GS_EHIP;
  hipStream_t st = B->stream;
  const uint64_t n = B->n_reads, nr1 = n + 1;
  out->n_reads = n;
  out->n_seeds = B->n_seeds;
  out->n_pos = B->n_pos;
  out->read_nseeds = (uint32_t*)gs::result_alloc(4 * nr1);
  out->read_npos = (uint32_t*)gs::result_alloc(4 * nr1);
  out->read_seed_off = (uint64_t*)gs::result_alloc(8 * nr1);
  out->read_pos_off = (uint64_t*)gs::result_alloc(8 * nr1);
  out->seeds = (gs_seed_rec*)gs::result_alloc(sizeof(gs_seed_rec) * (B->n_seeds + 1));
  out->pos = (uint32_t*)gs::result_alloc(4 * (B->n_pos + 1));
  if (!out->read_nseeds || !out->read_npos || !out->read_seed_off || !out->read_pos_off || !out->seeds ||
      !out->pos) {
    gs_result_free(out);
    return GS_ENOMEM;
  }
  if (B->graph_fresh) {  // the last replay already downloaded everything
    memcpy(out->read_nseeds, B->hr_ns, 4 * n);
    memcpy(out->read_npos, B->hr_np, 4 * n);
    if (n) {
      memcpy(out->read_seed_off, B->hr_soff, 8 * nr1);
      memcpy(out->read_pos_off, B->hr_poff, 8 * nr1);
    } else {
      out->read_seed_off[0] = out->read_pos_off[0] = 0;
    }
    memcpy(out->seeds, B->hr_seeds, sizeof(gs_seed_rec) * B->n_seeds);
    memcpy(out->pos, B->hr_rows, 4 * B->n_pos);
    return GS_OK;
  }
  hipEvent_t e0 = B->ev[5], e1 = B->ev[6];
  int rc = GS_OK;
  if (hipEventRecord(e0, st) != hipSuccess) rc = GS_EHIP;
  // per-read arrays, then seeds and positions, each through the pinned chunks
  if (!rc && n) rc = staged_d2h(B, out->read_nseeds, B->ns.p, 4 * n);
  if (!rc && n) rc = staged_d2h(B, out->read_npos, B->np.p, 4 * n);
  if (!rc) rc = staged_d2h(B, out->read_seed_off, B->soff.p, 8 * nr1);
  if (!rc) rc = staged_d2h(B, out->read_pos_off, B->poff.p, 8 * nr1);
  if (!rc && B->n_seeds) rc = staged_d2h(B, out->seeds, B->seeds.p, sizeof(gs_seed_rec) * B->n_seeds);
  if (!rc && B->n_pos) rc = staged_d2h(B, out->pos, B->rows.p, 4 * B->n_pos);
  if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)) rc = GS_EHIP;
  if (rc) {
    gs_result_free(out);
    return rc;
  }
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  out->t_d2h = ms * 1e-3;
  return GS_OK;
}

int gs::batch_download(gs_batch* B, ResultView* v) {
  if (!B || !v || !B->ran) return GS_EINVAL;
  if (hipSetDevice(B->X->device) != hipSuccess) return GS_EHIP;
  const uint64_t n = B->n_reads, nr1 = n + 1;
  if (B->dl_ns.ensure(4 * nr1) || B->dl_np.ensure(4 * nr1) || B->dl_soff.ensure(8 * nr1) ||
      B->dl_poff.ensure(8 * nr1) || B->dl_seeds.ensure(sizeof(gs_seed_rec) * (B->n_seeds + 1)) ||
      B->dl_pos.ensure(4 * (B->n_pos + 1)))
    return GS_ENOMEM;
  hipStream_t st = B->stream;
  if ((n && (hipMemcpyAsync(B->dl_ns.p, B->ns.p, 4 * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipMemcpyAsync(B->dl_np.p, B->np.p, 4 * n, hipMemcpyDeviceToHost, st) != hipSuccess)) ||
      hipMemcpyAsync(B->dl_soff.p, B->soff.p, 8 * nr1, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(B->dl_poff.p, B->poff.p, 8 * nr1, hipMemcpyDeviceToHost, st) != hipSuccess ||
      (B->n_seeds && hipMemcpyAsync(B->dl_seeds.p, B->seeds.p, sizeof(gs_seed_rec) * B->n_seeds,
                                    hipMemcpyDeviceToHost, st) != hipSuccess) ||
      (B->n_pos && hipMemcpyAsync(B->dl_pos.p, B->rows.p, 4 * B->n_pos, hipMemcpyDeviceToHost, st) != hipSuccess) ||
      hipStreamSynchronize(st) != hipSuccess)
    return GS_EHIP;
  v->n_reads = n;
  v->n_seeds = B->n_seeds;
  v->n_pos = B->n_pos;
  v->nseeds = static_cast<const uint32_t*>(B->dl_ns.p);
  v->npos = static_cast<const uint32_t*>(B->dl_np.p);
  v->soff = static_cast<const uint64_t*>(B->dl_soff.p);
  v->poff = static_cast<const uint64_t*>(B->dl_poff.p);
  v->seeds = static_cast<const gs_seed_rec*>(B->dl_seeds.p);
  v->pos = static_cast<const uint32_t*>(B->dl_pos.p);
  return GS_OK;
}

int gs_batch_set_tuning(gs_batch* B, int kernel, int wg, int walk_rows) {
  if (!B || kernel < 0 || kernel > 3 || (wg != 64 && wg != 128 && wg != 256) ||
      (walk_rows != 1 && walk_rows != 2 && walk_rows != 4))
    return GS_EINVAL;
  B->opts.kernel = kernel;
  B->opts.wg = (uint16_t)wg;
  B->opts.walk_rows = (uint16_t)walk_rows;
  return GS_OK;
}

// Text positions of suffix-array rows as the reference's walk reports them
// (locate_seeds_gpu, seed_gen.cu:668-684), with the same kernel as the
// pipeline (from the full suffix array when the index has it).
extern "C" GS_API int gs_index_locate(gs_index* X, const uint32_t* rows, uint64_t n, uint32_t* out) {
  if (!X || (n && (!rows || !out))) return GS_EINVAL;
  if (!n) return GS_OK;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  uint32_t* d = nullptr;
  if (hipMalloc(&d, n * 4) != hipSuccess) return GS_ENOMEM;
  int rc = GS_OK;
  if (hipMemcpyAsync(d, rows, n * 4, hipMemcpyHostToDevice, X->stream) != hipSuccess) rc = GS_EHIP;
  if (!rc) {
    launch_locate(X->dev(), d, n, 1, X->stream);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, d, n * 4, hipMemcpyDeviceToHost, X->stream) != hipSuccess ||
        hipStreamSynchronize(X->stream) != hipSuccess)
      rc = GS_EHIP;
  }
  (void)hipFree(d);
  return rc;
}

int gs_batch_set_direct_min(gs_batch* B, uint16_t direct_min) {
  if (!B || (direct_min > 64 && direct_min != 0xFFFF)) return GS_EINVAL;
  B->opts.direct_min = direct_min;
  return GS_OK;
}

extern "C" GS_API int gs_batch_destroy(gs_batch* B) {
  if (!B) return GS_EINVAL;
  (void)hipSetDevice(B->X->device);
  destroy_batch(B);
  return GS_OK;
}

// One batch from host memory to host memory: create, run, fetch.
extern "C" GS_API int gs_seed(gs_index* X, const char* bases, const uint64_t* offsets,
                              const uint32_t* lens, uint64_t n_reads, const gs_opts* opts,
                              gs_result* out) {
  if (!out) return GS_EINVAL;
  memset(out, 0, sizeof(*out));
  if (!X || !opts || (n_reads && (!bases || !offsets || !lens))) return GS_EINVAL;
  int rc = gs::plan_check(opts);
  if (rc) return rc;
  if (n_reads > 0xFFFFFFFFull / 2) return GS_ERANGE;
  gs::BatchPlan P;
  gs::plan_reset(P, *opts);
  if ((rc = gs::plan_add(P, bases, offsets, lens, 0, n_reads))) return rc;
  gs::plan_finish(P);
  // the index keeps this batch for the next call: its device buffers are
  // reused (no allocation, which would synchronise the device, once they fit)
  if (!X->work) X->work = gs::batch_new(X, false);
  gs_batch* B = X->work;
  if ((rc = gs::batch_load(B, P))) return rc;
  gs_batch_stats s;
  rc = gs_batch_run(B, -1, &s);
  if (!rc) rc = gs_batch_fetch(B, out);
  if (!rc) {
    out->t_h2d = s.t_h2d;
    out->t_seed = s.t_seed;
    out->t_post = s.t_count + s.t_write;
    out->t_locate = s.t_locate;
  }
  return rc;
}

extern "C" GS_API void gs_result_free(gs_result* r) {
  if (!r) return;
  free(r->read_nseeds);
  free(r->read_npos);
  free(r->read_seed_off);
  free(r->read_pos_off);
  free(r->seeds);
  free(r->pos);
  memset(r, 0, sizeof(*r));
}

// ---- gs_seed_noalloc: the same pipeline without any allocation --------------
// Every device buffer of the call is a slice of `scratch`, a file-scope array
// that each GPU holds from the library's load to the end of the process.  The
// host side keeps no vectors either: the task map, code-chunk offsets, work
// items and reference-batch ids that gs_batch_create plans on the CPU are
// computed on the device from the uploaded offsets and lengths.  Calls on
// one GPU share its array, so calls that overlap on one GPU race.

namespace {

constexpr uint64_t kNoAllocBytes = 256ull << 20;
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocDevices];  // device address of `scratch` on each device (null: not looked up)
int g_scratch_rc[kNoAllocDevices];     // the lookup's result once it has run

// Per read: the offset rebased onto the uploaded span, the task and code
// chunk counts that the scans turn into offsets, and the two work items of a
// whole-read stream (items may be null: plain kernel).
__global__ void noalloc_prepare(uint64_t* __restrict__ offs, const uint32_t* __restrict__ lens, uint32_t n,
                                uint64_t lo, uint32_t k, uint32_t* __restrict__ ntask,
                                uint32_t* __restrict__ ncode, uint64_t* __restrict__ items) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t L = lens[i];
  offs[i] -= lo;
  ntask[i] = L >= k ? L - k + 1 : 0u;
  ncode[i] = 2u * ((L + 31u) / 32u);
  if (items) {
    items[2ull * i] = i;
    items[2ull * i + 1] = i | (1ull << 32);
  }
}

// Reference-batch id of every read: a batch closes after the read that
// brings it to batch_bases (seed_gen.cu:2955-2991), a running count with a
// reset.  One wave: the lanes load 64 lengths at a time and every lane
// follows the whole count, keeping the id of its own read.
__global__ void __launch_bounds__(64) noalloc_batch_ids(const uint32_t* __restrict__ lens, uint32_t n,
                                                        uint32_t batch_bases, uint32_t* __restrict__ batch) {
  const uint32_t lane = threadIdx.x;
  uint32_t id = 0;
  uint64_t acc = 0;
  for (uint32_t base = 0; base < n; base += 64) {
    const uint32_t i = base + lane;
    const uint32_t L = i < n ? lens[i] : 0u;
    const uint32_t m = n - base < 64u ? n - base : 64u;
    uint32_t mine = 0;
    for (uint32_t j = 0; j < m; ++j) {
      const uint32_t Lj = (uint32_t)__shfl((int)L, (int)j, 64);
      if (lane == j) mine = id;
      acc += Lj;
      if (acc >= batch_bases) {
        ++id;
        acc = 0;
      }
    }
    if (i < n) batch[i] = mine;
  }
}

inline uint64_t align256(uint64_t v) { return (v + 255) & ~255ull; }

}  // namespace

// Looks up this device's copy of `scratch` (the current device must be
// `device`); called where an index is set up on a device.
int gs::noalloc_setup(int device) {
  if (device < 0 || device >= kNoAllocDevices) return GS_EINVAL;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return GS_OK;
  if (g_scratch_rc[device]) return g_scratch_rc[device];
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    g_scratch_rc[device] = GS_EHIP;
    return GS_EHIP;
  }
  g_scratch_ptr[device] = p;
  return GS_OK;
}

extern "C" GS_API int gs_seed_noalloc(gs_index* X, const char* bases, const uint64_t* offsets,
                                      const uint32_t* lens, uint64_t n_reads, const gs_opts* opts,
                                      gs_result_buffers* out) {
  if (!X || !opts || !out || (n_reads && (!bases || !offsets || !lens))) return GS_EINVAL;
  if (!out->read_seed_off || !out->read_pos_off || (n_reads && (!out->read_nseeds || !out->read_npos)))
    return GS_EINVAL;
  out->n_seeds = out->n_pos = 0;
  int rc = gs::plan_check(opts);
  if (rc) return rc;
  if (n_reads > 0xFFFFFFFFull / 2) return GS_ERANGE;
  const uint32_t n = (uint32_t)n_reads, k = (uint32_t)opts->k;
  // sizes from the lengths alone (one pass, nothing allocated)
  uint64_t lo = UINT64_MAX, hi = 0, tasks = 0, chunks = 0;
  uint32_t maxlen = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t L = lens[i];
    if (offsets[i] < lo) lo = offsets[i];
    if (offsets[i] + L > hi) hi = offsets[i] + L;
    tasks += L >= k ? L - k + 1 : 0;
    chunks += 2 * ((L + 31) / 32);
    if (L > maxlen) maxlen = L;
  }
  if (hi < lo) lo = hi = 0;
  const uint64_t span = hi - lo, nr1 = (uint64_t)n + 1;
  // the persistent kernel takes reads held whole in its LDS window; a call
  // with longer reads runs the plain kernel (the same output)
  const int kernel = opts->kernel != 0 && maxlen <= gs::kPersistLmax ? opts->kernel : 0;
  const bool long_scratch = kernel == 0 && maxlen > 256;

  // slices of the scratch array, in order; outputs (seed records, rows) come
  // after the counts are known
  uint64_t at = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = at;
    at = align256(at + bytes);
    return o;
  };
  const uint64_t o_bases = take(span + 32), o_offs = take(8 * nr1), o_lens = take(4 * nr1),
                 o_toff = take(8 * nr1), o_batch = take(4 * nr1), o_code_off = take(8 * nr1),
                 o_items = kernel ? take(8 * (2 * (uint64_t)n + 1)) : 0,
                 o_codes = kernel ? take(16 * (chunks + 1)) : 0, o_cand = take(16 * (2 * tasks + 1)),
                 o_nF = take(4 * nr1), o_nR = take(4 * nr1), o_ns = take(4 * nr1), o_np = take(4 * nr1),
                 o_soff = take(8 * nr1), o_poff = take(8 * nr1), o_tiles = take(8 * (nr1 / kScanTile + 2)),
                 o_tot = take(16), o_queue = take(16), o_counters = take(gs::kCounterBytes + gs::kColdBytes),
                 o_long = long_scratch ? take(4 * (span + 16)) : 0;
  if (at > kNoAllocBytes) return GS_ERANGE;
  const uint64_t fixed = at;

  if (X->device < 0 || X->device >= kNoAllocDevices) return GS_EINVAL;
  if (hipSetDevice(X->device) != hipSuccess) return GS_EHIP;
  uint8_t* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = static_cast<uint8_t*>(g_scratch_ptr[X->device]);
  }
  if (!base) return GS_EHIP;  // the lookup at set-up failed
  hipStream_t st = X->stream;

  // a batch whose buffers are views into the scratch array; it owns nothing
  gs_batch B;
  B.X = X;
  B.stream = st;
  B.opts = *opts;
  B.opts.kernel = kernel;
  B.n_reads = n;
  B.span = span;
  B.tasks = tasks;
  B.maxlen = maxlen;
  B.code_chunks = chunks;
  B.n_items = kernel ? 2ull * n : 0;
  B.n_long = 0;
  B.seg_tasks = opts->seg_tasks ? opts->seg_tasks : 64u;
  B.dyn_cap = 0;
  uint64_t h_tot[2] = {0, 0};
  B.h_tot = h_tot;
  B.bases.view(base + o_bases, span + 32);
  B.offs.view(base + o_offs, 8 * nr1);
  B.lens.view(base + o_lens, 4 * nr1);
  B.toff.view(base + o_toff, 8 * nr1);
  B.batch.view(base + o_batch, 4 * nr1);
  B.code_off.view(base + o_code_off, 8 * nr1);
  if (kernel) {
    B.items.view(base + o_items, 8 * (2 * (uint64_t)n + 1));
    B.codes.view(base + o_codes, 16 * (chunks + 1));
  }
  B.cand.view(base + o_cand, 16 * (2 * tasks + 1));
  B.nF.view(base + o_nF, 4 * nr1);
  B.nR.view(base + o_nR, 4 * nr1);
  B.ns.view(base + o_ns, 4 * nr1);
  B.np.view(base + o_np, 4 * nr1);
  B.soff.view(base + o_soff, 8 * nr1);
  B.poff.view(base + o_poff, 8 * nr1);
  B.tiles.view(base + o_tiles, 8 * (nr1 / kScanTile + 2));
  B.tot.view(base + o_tot, 16);
  B.queue.view(base + o_queue, 16);
  B.counters.view(base + o_counters, gs::kCounterBytes + gs::kColdBytes);
  if (long_scratch) B.scratch.view(base + o_long, 4 * (span + 16));

  // inputs, then the plan on the device
  if ((span && hipMemcpyAsync(B.bases.p, bases + lo, span, hipMemcpyHostToDevice, st) != hipSuccess) ||
      (n && (hipMemcpyAsync(B.offs.p, offsets, 8ull * n, hipMemcpyHostToDevice, st) != hipSuccess ||
             hipMemcpyAsync(B.lens.p, lens, 4ull * n, hipMemcpyHostToDevice, st) != hipSuccess)))
    return GS_EHIP;
  hipLaunchKernelGGL(zero_bytes, dim3(1), dim3(32), 0, st, B.bases.as<uint8_t>() + span, 32u);
  if (n) {
    hipLaunchKernelGGL(noalloc_prepare, dim3(grid_for(n, 256)), dim3(256), 0, st, B.offs.as<uint64_t>(),
                       B.lens.as<uint32_t>(), n, lo, k, B.ns.as<uint32_t>(), B.np.as<uint32_t>(),
                       kernel ? B.items.as<uint64_t>() : (uint64_t*)nullptr);
    scan_counts(B.ns.as<uint32_t>(), n, B.toff.as<uint64_t>(), B.tiles.as<uint64_t>(), B.tot.as<uint64_t>(), st);
    scan_counts(B.np.as<uint32_t>(), n, B.code_off.as<uint64_t>(), B.tiles.as<uint64_t>(),
                B.tot.as<uint64_t>() + 1, st);
    hipLaunchKernelGGL(noalloc_batch_ids, dim3(1), dim3(64), 0, st, B.lens.as<uint32_t>(), n,
                       opts->batch_bases ? opts->batch_bases : 1000000u, B.batch.as<uint32_t>());
  }
  if (hipGetLastError() != hipSuccess) return GS_EHIP;

  // extension and counts, as gs_batch_run
  unsigned pgrid = 0;
  if ((rc = persistent_grid_for(&B, kernel, &pgrid))) return rc;
  SeedParams S;
  PostParams Q;
  fill_params(&B, S, Q);
  if ((rc = enqueue_extension(&B, S, kernel, pgrid, st))) return rc;
  if ((rc = enqueue_counts(&B, Q, st))) return rc;
  GS_HIP(hipStreamSynchronize(st));
  const uint64_t n_seeds = n ? h_tot[0] : 0, n_pos = n ? h_tot[1] : 0;
  out->n_seeds = n_seeds;
  out->n_pos = n_pos;
  const uint64_t o_seeds = fixed, o_rows = align256(o_seeds + sizeof(gs_seed_rec) * (n_seeds + 1));
  if (o_rows + 4 * (n_pos + 1) > kNoAllocBytes) return GS_ERANGE;
  if (n_seeds > out->seeds_cap || n_pos > out->pos_cap || (n_seeds && !out->seeds) || (n_pos && !out->pos))
    return GS_ERANGE;  // n_seeds and n_pos say what the call needs
  B.n_seeds = n_seeds;
  B.n_pos = n_pos;
  B.seeds.view(base + o_seeds, sizeof(gs_seed_rec) * (n_seeds + 1));
  B.rows.view(base + o_rows, 4 * (n_pos + 1));
  fill_params(&B, S, Q);
  if ((rc = enqueue_write(&B, Q, st))) return rc;
  if ((rc = enqueue_walk(&B, S, Q, st))) return rc;
  if ((n && (hipMemcpyAsync(out->read_nseeds, B.ns.p, 4ull * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipMemcpyAsync(out->read_npos, B.np.p, 4ull * n, hipMemcpyDeviceToHost, st) != hipSuccess)) ||
      hipMemcpyAsync(out->read_seed_off, B.soff.p, 8 * nr1, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(out->read_pos_off, B.poff.p, 8 * nr1, hipMemcpyDeviceToHost, st) != hipSuccess ||
      (n_seeds && hipMemcpyAsync(out->seeds, B.seeds.p, sizeof(gs_seed_rec) * n_seeds, hipMemcpyDeviceToHost,
                                 st) != hipSuccess) ||
      (n_pos && hipMemcpyAsync(out->pos, B.rows.p, 4 * n_pos, hipMemcpyDeviceToHost, st) != hipSuccess))
    return GS_EHIP;
  GS_HIP(hipStreamSynchronize(st));
  return GS_OK;
}
