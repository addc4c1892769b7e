// Persistent extension kernel (the batched form of the seeding path).
//
// Same stream procedure and output as seed_streams in gs_seed.hip (the plain
// one-stream-per-lane kernel), restructured for throughput on gfx950:
//  * a fixed grid that fills the chip; every lane pulls (read, strand)
//    streams from a device work queue (one atomic per wave refill, leader
//    lane + ballot) and starts the next stream as soon as its current one
//    ends, so no lane idles while its wave finishes a long stream;
//  * reads are pre-packed to 4-bit codes of both virtual strands by
//    pack_codes, so a refill is a few 16-byte loads and LDS stores;
//  * a lane's LDS slot holds its stream's codes (4 bits per base) and its
//    size record (one AT per position), indexed directly in LDS;
//  * the loop body is one memory operation (an interval-table entry, from
//    the extended table when the index has one, or one or two occ entries)
//    followed by a short processing step; the rare
//    transitions (windows with a code > 3, N-transparent positions, reaching
//    the read start) are resolved inside the processing step.
#include "gs_internal.h"

namespace {

__device__ __forceinline__ uint32_t pick4(uint32_t c, uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3) {
  return c == 0 ? a0 : (c == 1 ? a1 : (c == 2 ? a2 : a3));
}

template <typename AT, int LMAX, int BS>
__global__ void __launch_bounds__(BS) seed_persistent(gs::PersistParams P) {
  static_assert(LMAX % 32 == 0, "LMAX must be a multiple of 32");
  constexpr int CODE_BYTES = LMAX / 2;
  constexpr int A_BYTES = LMAX * (int)sizeof(AT);
  constexpr int SLOT = CODE_BYTES + A_BYTES;
  static_assert(SLOT % 16 == 0, "slot must be 16-byte aligned");
  constexpr uint32_t SAT = (uint32_t)(AT)(~(AT)0);
  __shared__ __align__(16) uint8_t lds[BS * SLOT];
  const uint32_t sbase = threadIdx.x * SLOT;
  AT* const A = reinterpret_cast<AT*>(lds + sbase + CODE_BYTES);
  const int lane = threadIdx.x & 63;
  const uint32_t k = P.k;
  const uint32_t L2_0 = P.X.L2[0], L2_1 = P.X.L2[1], L2_2 = P.X.L2[2], L2_3 = P.X.L2[3];
  const uint32_t primary = P.X.primary;
  const uint64_t n_items = P.n_items;

  bool have = false, finished = false;
  uint32_t r = 0, strand = 0, L = 0, ntask = 0, nout = 0, last_b = 0, p = 0, nbad = 0;
  uint32_t l = 0, u = 0, b = 0, s = 0, cx = 0;
  int32_t e = 0, x = 0, e_min = 0, w0 = 0;
  uint32_t t0 = 0, first_b = 0, flags = 0;
  uint64_t item = 0;
  int phase = 0;
  bool ext_miss = false;  // the task's extended-table pattern is absent: use the 13-mer path
  uint4* out = nullptr;
  const uint4* gcodes = nullptr;  // the stream's codes in device memory (segments only)
  uint64_t n_table = 0, n_occ = 0;

  // code of position i of the stream: the LDS window holds [w0, w0 + LMAX);
  // a segment's extension may run left of it and then reads device memory
  auto code = [&](int32_t i) -> uint32_t {
    const int32_t j = i - w0;
    if (j >= 0) return (lds[sbase + (j >> 1)] >> ((j & 1) << 2)) & 15u;
    const uint32_t w = reinterpret_cast<const uint32_t*>(gcodes + (i >> 5))[(i >> 3) & 3];
    return (w >> ((i & 7) << 2)) & 15u;
  };

  auto end_stream = [&]() {
    if (gcodes)
      P.seg_meta[item] = make_uint4(nout, first_b, last_b, flags);
    else
      (strand ? P.nR : P.nF)[r] = nout;
    have = false;
  };
  auto advance = [&]() -> bool {
    if (e - 1 < e_min) return false;
    const uint32_t cn = code(e - 14), co = code(e - 1);
    nbad = nbad + (cn > 3 ? 1u : 0u) - (co > 3 ? 1u : 0u);
    p = (p >> 2) | ((cn & 3u) << 24);
    --e;
    return true;
  };
  // next task with a clean 13-base window, or end of the stream
  auto next_task = [&]() {
    phase = 0;
    ext_miss = false;
    do {
      if (!advance()) {
        end_stream();
        return;
      }
    } while (nbad != 0);
  };
  // size record at pos (positions left of the window keep none: pruning is
  // only an early exit, so a missing record never changes the output)
  auto prune_at = [&](int32_t pos, uint32_t size) -> bool {
    const int32_t j = pos - w0;
    if (j < 0) return false;
    const uint32_t sv = size >= SAT ? SAT : size;
    if (size < SAT && (uint32_t)A[j] == sv) return true;
    A[j] = (AT)sv;
    return false;
  };
  auto finish_task = [&]() {
    const bool emit = (uint32_t)e - b >= k && (!P.smem || b < last_b);
    if (emit) {
      if (strand == 0)
        out[ntask - 1 - t0 - nout] = make_uint4(b, (uint32_t)e, l, u);
      else
        out[ntask + t0 + nout] = make_uint4((L - (uint32_t)e) | 0x80000000u, (L - b) | 0x80000000u, l, u);
      ++nout;
    }
    if (!(flags & 2u)) {  // the segment's first finished task
      first_b = b;
      flags = 2u | (emit ? 1u : 0u);
    }
    last_b = b;
  };
  // phase 1: settle x on the next position with a base code
  auto prepare_x = [&]() {
    while (true) {
      if (x < 0) {
        finish_task();
        if (P.smem)
          end_stream();  // later tasks of the item end at the read start too and are filtered out
        else
          next_task();
        return;
      }
      const uint32_t c = code(x);
      if (c <= 3) {
        cx = c;
        return;
      }
      b = (uint32_t)x;  // code > 3 after the window: interval kept, position consumed
      if (prune_at(x, s)) {
        next_task();
        return;
      }
      --x;
    }
  };
  auto start_stream = [&](uint64_t it) {
    item = it;
    const uint64_t v = P.items[it];
    r = (uint32_t)v;
    strand = (uint32_t)(v >> 32) & 1u;
    const uint32_t seg = (uint32_t)(v >> 33);
    L = P.lens[r];
    if (L < k) {
      (strand ? P.nR : P.nF)[r] = 0;
      return;
    }
    ntask = L - k + 1;
    const uint32_t nch = (L + 31) >> 5;
    const uint4* src = P.codes + P.code_off[r] + strand * nch;
    uint32_t top, tcount;
    if (L <= (uint32_t)LMAX) {  // the whole stream in the window
      gcodes = nullptr;
      t0 = 0;
      tcount = ntask;
      top = nch << 5;
      w0 = 0;
    } else {                    // one segment of task ends
      gcodes = src;
      t0 = seg * P.seg_tasks;
      tcount = min(P.seg_tasks, ntask - t0);
      top = ((L - t0) + 31u) & ~31u;
      w0 = top > (uint32_t)LMAX ? (int32_t)(top - LMAX) : 0;
    }
    const uint32_t c0 = (uint32_t)w0 >> 5, c1 = top >> 5;
    for (uint32_t c = c0; c < c1; ++c) *reinterpret_cast<uint4*>(lds + sbase + 16 * (c - c0)) = src[c];
    const uint32_t za = ((top - (uint32_t)w0) * (uint32_t)sizeof(AT) + 15) >> 4;
    for (uint32_t c = 0; c < za; ++c)
      *reinterpret_cast<uint4*>(lds + sbase + CODE_BYTES + 16 * c) = make_uint4(0, 0, 0, 0);
    out = P.cand + 2 * P.task_off[r];
    nout = 0;
    last_b = 0xFFFFFFFFu;
    first_b = 0xFFFFFFFFu;
    flags = 0;
    e = (int32_t)(L - t0);
    e_min = e - (int32_t)tcount + 1;
    p = 0;
    nbad = 0;
    for (int j = 0; j < gs::kPrecalc; ++j) {
      const uint32_t c = code(e - 1 - j);
      nbad += c > 3 ? 1u : 0u;
      p |= (c & 3u) << (2 * j);
    }
    have = true;
    phase = 0;
    if (nbad != 0) next_task();
  };

  while (true) {
    // lanes without a stream take the next items of the queue
    const bool need = !have && !finished;
    const uint64_t needmask = __ballot(need);
    if (needmask) {
      const int leader = __ffsll((unsigned long long)needmask) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(P.queue, (uint32_t)__popcll(needmask));
      base = __shfl(base, leader);
      if (need) {
        const uint64_t it = (uint64_t)base + __popcll(needmask & ((1ull << lane) - 1ull));
        if (it >= n_items)
          finished = true;
        else
          start_stream(it);
      }
    }
    const int op = have ? (phase == 0 ? 1 : 2) : 0;
    if (__ballot(op != 0) == 0) {
      if (__ballot(!finished) == 0) break;
      continue;
    }
    // one memory operation per lane
    const uint4* a1 = nullptr;
    const uint4* a2 = nullptr;
    uint32_t i1 = 0, i2 = 0, b1 = 0, b2 = 0;
    bool have1 = false;
    uint32_t tlen = gs::kPrecalc, tidx = p;  // table lookup: pattern length and entry
    if (op == 1) {
      const uint2* tbl = P.X.kmer;
      if (P.X.ext_len && !ext_miss && e >= (int32_t)P.X.ext_len) {
        uint32_t q = p, ok = 1;
        for (uint32_t j = gs::kPrecalc; j < P.X.ext_len; ++j) {
          const uint32_t c = code(e - 1 - (int32_t)j);
          ok &= c <= 3 ? 1u : 0u;
          q |= (c & 3u) << (2 * j);
        }
        if (ok) {
          tlen = P.X.ext_len;
          tidx = q;
          tbl = P.X.kext;
        }
      }
      a1 = reinterpret_cast<const uint4*>(tbl) + (tidx >> 1);
      ++n_table;
    } else if (op == 2) {
      const uint32_t r1 = l - 1u;
      have1 = r1 != 0xFFFFFFFFu;
      i2 = u - (u >= primary ? 1u : 0u);
      b2 = gs::div96(i2);
      i1 = have1 ? r1 - (r1 >= primary ? 1u : 0u) : 0u;
      b1 = gs::div96(i1);
      a1 = P.X.occ + (size_t)b2 * 4 + cx;
      if (have1 && b1 != b2) a2 = P.X.occ + (size_t)b1 * 4 + cx;
      n_occ += a2 ? 2 : 1;
    }
    uint4 d1 = make_uint4(0, 0, 0, 0);
    if (a1) d1 = *a1;
    uint4 d2 = d1;
    if (a2) d2 = *a2;

    if (op == 1) {
      const uint32_t il = (tidx & 1u) ? d1.z : d1.x, iu = (tidx & 1u) ? d1.w : d1.y;
      if (il > iu) {
        if (tlen > gs::kPrecalc)
          ext_miss = true;  // absent longer pattern: the 13-mer lookup and steps decide
        else
          next_task();
      } else {
        l = il;
        u = iu;
        s = u - l + 1u;
        b = (uint32_t)e - tlen;
        if (prune_at((int32_t)b, s)) {
          next_task();
        } else {
          x = (int32_t)b - 1;
          phase = 1;
          prepare_x();
        }
      }
    } else if (op == 2) {
      const uint32_t o2 = gs::rank_in_entry(d1, i2 - b2 * gs::kBlockBases);
      const uint32_t o1 = have1 ? gs::rank_in_entry(d2, i1 - b1 * gs::kBlockBases) : 0u;
      const uint32_t base = pick4(cx, L2_0, L2_1, L2_2, L2_3);
      const uint32_t l2 = base + o1 + 1u, u2 = base + o2;
      if (l2 > u2) {
        finish_task();
        next_task();
      } else {
        l = l2;
        u = u2;
        s = u - l + 1u;
        b = (uint32_t)x;
        if (prune_at(x, s)) {
          next_task();
        } else {
          --x;
          prepare_x();
        }
      }
    }
  }
  if (P.counters) {
    // wave totals, one atomic per wave
    for (int d = 32; d > 0; d >>= 1) {
      n_table += __shfl_down(n_table, d, 64);
      n_occ += __shfl_down(n_occ, d, 64);
    }
    if (lane == 0) {
      atomicAdd((unsigned long long*)&P.counters[0], (unsigned long long)n_table);
      atomicAdd((unsigned long long*)&P.counters[1], (unsigned long long)n_occ);
    }
  }
}

// 4-bit codes of both virtual strands: forward v[i] = code of byte i, reverse
// v[i] = reverse-complement code of byte L-1-i (pack_4bit_fow / pack_4bit_rev,
// seed_gen.cu:2536-2578); two codes per byte, even position in the low
// nibble, 32 codes per 16-byte chunk.
__device__ __forceinline__ uint4 pack_chunk(const uint8_t* s, uint32_t L, uint32_t strand, uint32_t c) {
  uint32_t w[4] = {0, 0, 0, 0};
  for (uint32_t j = 0; j < 32; ++j) {
    const uint32_t i = c * 32 + j;
    if (i >= L) break;
    const uint32_t v = strand ? gs::code_rev(s[L - 1 - i]) : gs::code_fwd(s[i]);
    w[j >> 3] |= v << ((j & 7) << 2);
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// reads of length <= kPersistLmax: one thread per (read, strand, chunk)
__global__ void __launch_bounds__(256) pack_codes(const uint8_t* __restrict__ bases, const uint64_t* __restrict__ offs,
                                                  const uint32_t* __restrict__ lens,
                                                  const uint64_t* __restrict__ code_off, uint32_t n_reads,
                                                  uint4* __restrict__ codes) {
  constexpr uint32_t cmax = gs::kPersistLmax >> 5;
  const uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  const uint64_t r = t / (2 * cmax);
  if (r >= n_reads) return;
  const uint32_t rest = (uint32_t)(t - r * 2 * cmax);
  const uint32_t strand = rest / cmax, c = rest - strand * cmax;
  const uint32_t L = lens[r];
  if (L > gs::kPersistLmax) return;
  const uint32_t nch = (L + 31) >> 5;
  if (c >= nch) return;
  codes[code_off[r] + strand * nch + c] = pack_chunk(bases + offs[r], L, strand, c);
}

// longer reads: one workgroup per read
__global__ void __launch_bounds__(256) pack_codes_long(const uint8_t* __restrict__ bases,
                                                       const uint64_t* __restrict__ offs,
                                                       const uint32_t* __restrict__ lens,
                                                       const uint64_t* __restrict__ code_off,
                                                       const uint32_t* __restrict__ long_reads,
                                                       uint4* __restrict__ codes) {
  const uint32_t r = long_reads[blockIdx.x];
  const uint32_t L = lens[r];
  const uint32_t nch = (L + 31) >> 5;
  const uint8_t* s = bases + offs[r];
  for (uint32_t q = threadIdx.x; q < 2 * nch; q += 256) {
    const uint32_t strand = q >= nch ? 1u : 0u;
    codes[code_off[r] + q] = pack_chunk(s, L, strand, q - strand * nch);
  }
}

// Joins the segments of a long read's stream (one thread per stream): moves
// each segment's candidates next to the previous segment's and settles the
// SMEM filter at segment boundaries.  A segment filters its own tasks in the
// lane, but its first finished task has no predecessor there: it survives iff
// its left end is smaller than the last finished task's of earlier segments
// (finish_task of the persistent kernel, 991-1018).  Candidates only move
// toward the read's end of the region, in an order that reads every source
// before it can be overwritten.
__global__ void __launch_bounds__(256) join_segments(gs::JoinParams J) {
  const uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t >= 2ull * J.n_long) return;
  const uint32_t r = J.long_reads[t >> 1], strand = (uint32_t)(t & 1);
  const uint32_t L = J.lens[r];
  const uint32_t ntask = L - J.k + 1;
  const uint32_t S = J.seg_tasks;
  const uint32_t nseg = (ntask + S - 1) / S;
  uint4* out = J.cand + 2 * J.task_off[r];
  const uint4* meta = J.seg_meta + J.item0[t >> 1] + strand;
  uint32_t K = 0, LB = 0xFFFFFFFFu;
  for (uint32_t g = 0; g < nseg; ++g) {
    const uint4 m = meta[2 * (uint64_t)g];
    const uint32_t t0 = g * S;
    const uint32_t drop = (J.smem && (m.w & 1u) && !(m.y < LB)) ? 1u : 0u;
    for (uint32_t i = drop; i < m.x; ++i) {
      if (strand == 0)
        out[ntask - 1 - K - (i - drop)] = out[ntask - 1 - t0 - i];
      else
        out[ntask + K + (i - drop)] = out[ntask + t0 + i];
    }
    K += m.x - drop;
    if (m.w & 2u) LB = m.z;
  }
  (strand ? J.nR : J.nF)[r] = K;
}

__global__ void zero_words(uint32_t* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename AT, int LMAX, int BS>
int grid_t(unsigned* grid) {
  int dev = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GS_EHIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return GS_EHIP;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, seed_persistent<AT, LMAX, BS>, BS, 0) != hipSuccess ||
      per_cu < 1)
    return GS_EHIP;
  *grid = (unsigned)(per_cu * prop.multiProcessorCount);
  return GS_OK;
}

template <typename AT, int LMAX, int BS>
int launch_t(const gs::PersistParams& P, unsigned grid, hipStream_t st) {
  hipLaunchKernelGGL((seed_persistent<AT, LMAX, BS>), dim3(grid), dim3(BS), 0, st, P);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_EHIP;
}

}  // namespace

namespace gs {

int launch_pack_codes(const uint8_t* bases, const uint64_t* offs, const uint32_t* lens, const uint64_t* code_off,
                      uint32_t n_reads, const uint32_t* long_reads, uint32_t n_long, uint4* codes, hipStream_t st) {
  const uint64_t threads = (uint64_t)n_reads * 2 * (kPersistLmax >> 5);
  if (threads)
    hipLaunchKernelGGL(pack_codes, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, bases, offs, lens,
                       code_off, n_reads, codes);
  if (n_long)
    hipLaunchKernelGGL(pack_codes_long, dim3(n_long), dim3(256), 0, st, bases, offs, lens, code_off, long_reads,
                       codes);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_EHIP;
}

int launch_join(const JoinParams& J, hipStream_t st) {
  if (!J.n_long) return GS_OK;
  const uint64_t threads = 2ull * J.n_long;
  hipLaunchKernelGGL(join_segments, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, J);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_EHIP;
}

// Workgroups that fill every CU at the kernel's occupancy (LDS slots bound it).
int persistent_grid(int variant, int wg, unsigned* grid) {
  constexpr int L = kPersistLmax;
  switch (variant * 1000 + (wg ? wg : 128)) {
    case 64: return grid_t<uint8_t, L, 64>(grid);
    case 128: return grid_t<uint8_t, L, 128>(grid);
    case 256: return grid_t<uint8_t, L, 256>(grid);
    case 1064: return grid_t<uint16_t, L, 64>(grid);
    case 1128: return grid_t<uint16_t, L, 128>(grid);
    case 1256: return grid_t<uint16_t, L, 256>(grid);
    default: return GS_EINVAL;
  }
}

int launch_persistent(const PersistParams& P, int variant, int wg, unsigned grid, hipStream_t st) {
  if (!grid) return GS_EINVAL;
  hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, st, P.queue, 1u);
  if (P.counters) hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, st, (uint32_t*)P.counters, 4u);
  constexpr int L = kPersistLmax;
  switch (variant * 1000 + (wg ? wg : 128)) {
    case 64: return launch_t<uint8_t, L, 64>(P, grid, st);
    case 128: return launch_t<uint8_t, L, 128>(P, grid, st);
    case 256: return launch_t<uint8_t, L, 256>(P, grid, st);
    case 1064: return launch_t<uint16_t, L, 64>(P, grid, st);
    case 1128: return launch_t<uint16_t, L, 128>(P, grid, st);
    case 1256: return launch_t<uint16_t, L, 256>(P, grid, st);
    default: return GS_EINVAL;
  }
}

}  // namespace gs
