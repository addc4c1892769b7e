// Persistent extension kernel (the batched form of the seeding path).
//
// Same stream procedure and output as seed_streams in gs_seed.hip (the plain
// one-stream-per-lane kernel), restructured for throughput on gfx950:
//  * a fixed grid that fills the chip; every lane pulls work items from a
//    device queue (one atomic per wave refill, leader lane + ballot) and
//    starts the next one as soon as its current one ends;
//  * reads are pre-packed to 4-bit codes of both virtual strands by
//    pack_codes, so a refill is a few 16-byte loads and LDS stores;
//  * a lane's LDS slot holds its item's codes (4 bits per base) and its size
//    record (4, 8 or 16 bits per position), indexed directly in LDS;
//  * the loop body is one memory operation (an interval-table entry, from
//    the extended table when the index has one, or one or two occ entries,
//    or a direct-extension read: a suffix-array word, a 64-base text chunk,
//    an inverse-suffix-array word) followed by a short processing step; the
//    rare transitions (windows with a code > 3, N-transparent positions,
//    reaching the read start) are resolved inside the processing step;
//  * tasks that cannot reach k bases (absent clean 16- or 19-base window,
//    from the extended table's interval and presence bits) end after their
//    lookup; each task end keeps its 19 codes in a rolling register window;
//  * text chunks are compared by the whole wave, one loading lane at a time;
//  * work splitting: the cost of a stream is heavy-tailed (in young repeat
//    families interval sizes shrink at every step, so distance-1 pruning
//    never stops a task and each task walks far).  Once the main queue is
//    exhausted and fewer items run than half the lanes, an item that has used
//    split_work memory operations since it started or last split, and still
//    has at least 2 * split_tasks task ends to do (for reads held whole in
//    the LDS window at most kSplitWorkWhole and kSplitTasksWhole), hands half
//    of them to an
//    overflow queue, which idle lanes take (a piece starts without size
//    records, so splitting while every lane is busy would only add work);
//    the join kernels put the pieces back in task order and settle the
//    SMEM filter at their boundaries, as for the segments of long reads.
//    A split keeps the upper half of the remaining task ends in the lane
//    (with its size records) and pushes the lower half as one piece, which
//    may split again; pieces are chained in task order.  Lanes never wait
//    for one another: a piece is written by the lane that reserved it in the
//    same iteration, a lane that claimed a piece whose entry is not visible
//    yet checks again next iteration, and a lane leaves once the main queue
//    is exhausted, no piece is left to claim and no item is running anywhere
//    (so no piece can appear any more).
#include <algorithm>
#include <type_traits>

#include "gs_internal.h"

namespace {

// tag type of 4-bit size records (seed_persistent's record width is taken
// from its first template argument: Nib, uint8_t or uint16_t)
struct Nib {};

__device__ __forceinline__ uint32_t pick4(uint32_t c, uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3) {
  return c == 0 ? a0 : (c == 1 ? a1 : (c == 2 ? a2 : a3));
}

constexpr uint32_t kNone = 0xFFFFFFFFu;
// workgroups (of four waves) of the helper kernels that take reads or
// streams in turn: 16 per CU on 256 CUs
constexpr uint32_t kHelperGrid = 4096;

// 8 4-bit codes (code i in bits 4i..4i+3) -> their low two bits packed (code
// i in bits 2i, 2i+1) and a bit per code above 3 (bit i)
__device__ __forceinline__ uint32_t nib_codes(uint32_t x) {
  uint32_t c = x & 0x33333333u;
  c = (c | (c >> 2)) & 0x0F0F0F0Fu;
  c = (c | (c >> 4)) & 0x00FF00FFu;
  return (c | (c >> 8)) & 0x0000FFFFu;
}
__device__ __forceinline__ uint32_t nib_bad(uint32_t x) {
  uint32_t b = ((x >> 2) | (x >> 3)) & 0x11111111u;
  b = (b | (b >> 3)) & 0x03030303u;
  b = (b | (b >> 6)) & 0x000F000Fu;
  return (b | (b >> 12)) & 0x000000FFu;
}
// 32 2-bit fields in reverse order
__device__ __forceinline__ uint64_t rev2(uint64_t x) {
  const uint64_t r = __builtin_bitreverse64(x);
  return ((r >> 1) & 0x5555555555555555ull) | ((r & 0x5555555555555555ull) << 1);
}
constexpr uint32_t kBad13 = (1u << gs::kPrecalc) - 1u;  // the 13-mer window's codes in a 19-code window

// A 16-byte chunk of 32 4-bit codes as the persistent kernel's LDS slot
// holds it: {codes at 2 bits in reverse order (low, high word), a bit per
// code above 3 in reverse order, 0}.
__device__ __forceinline__ uint4 slot_chunk(uint4 v) {
  const uint64_t f = (uint64_t)nib_codes(v.x) | ((uint64_t)nib_codes(v.y) << 16) |
                     ((uint64_t)nib_codes(v.z) << 32) | ((uint64_t)nib_codes(v.w) << 48);
  const uint32_t fb = nib_bad(v.x) | (nib_bad(v.y) << 8) | (nib_bad(v.z) << 16) | (nib_bad(v.w) << 24);
  const uint64_t rv = rev2(f);
  return make_uint4((uint32_t)rv, (uint32_t)(rv >> 32), __builtin_bitreverse32(fb), 0u);
}

__device__ __forceinline__ uint32_t load_acquire(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// The persistent kernel's arguments: what its loop reads on every
// iteration, and the device copy of the rest (gs::PersistCold, read where it
// is used).
struct PersistArgs {
  gs::DevIndex X;
  const uint4* codes;
  const gs::PersistCold* cold;
  uint32_t* queue;
  uint64_t n_items;
  uint32_t k, smem, direct_min, dyn_cap, split_work, split_tasks, split_below, refill_group, tandem;
};

// Waves per SIMD the register allocation must allow: 4 for the 4-bit
// records (their LDS slot lets 16 waves share a CU, and this memory-bound
// loop wants them all; the few values that do not fit, such as the
// candidate pointer, are reloaded where a task emits), no bound for the
// wider records, whose LDS slots allow fewer waves anyway.
template <typename AT>
constexpr int kMinWaves = std::is_same<AT, Nib>::value ? 4 : 1;

template <typename AT, int LMAX, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(kMinWaves<AT>)))
seed_persistent(PersistArgs P) {
  static_assert(LMAX % 32 == 0, "LMAX must be a multiple of 32");
  static_assert(LMAX == (int)gs::kPersistLmax, "pack_codes writes reads of up to kPersistLmax in the slot's form");
  // size records of RB bits per position: 4 (Nib, two per byte, low nibble
  // first), 8 or 16; a record saturates at 2^RB - 1 and then never prunes
  constexpr int RB = std::is_same<AT, Nib>::value ? 4 : 8 * (int)sizeof(AT);
  // A lane's LDS slot: the size records (RB bits per position of the
  // window), then the window's codes at 2 bits in reverse order (position j
  // at index r = w0 + LMAX - 1 - j, 16 per word, two words of padding), then
  // a bit per position set for a code above 3 (same index, one padding word
  // with every bit set: positions left of the read start).  In this order
  // the 22 codes left of any task end are one run of bits, read with three
  // word loads.
  constexpr int A_BYTES = LMAX * RB / 8;
  constexpr int CW = LMAX / 16 + 2;  // code words
  constexpr int BW = LMAX / 32 + 1;  // words of the code-above-3 bits
  constexpr int SLOT = (A_BYTES + 4 * (CW + BW) + 15) & ~15;
  static_assert(A_BYTES % 16 == 0, "records must keep 16-byte alignment");
  constexpr uint32_t SAT = (1u << RB) - 1u;
  __shared__ __align__(16) uint8_t lds[BS * SLOT];
  const uint32_t sbase = threadIdx.x * SLOT;
  // record j of the slot starting at byte `slot`
  auto rec_get = [&](uint32_t slot, int32_t j) __attribute__((always_inline)) -> uint32_t {
    uint8_t* a = lds + slot;
    if constexpr (RB == 4)
      return (a[j >> 1] >> ((j & 1) << 2)) & 15u;
    else if constexpr (RB == 8)
      return a[j];
    else
      return reinterpret_cast<uint16_t*>(a)[j];
  };
  auto rec_set = [&](uint32_t slot, int32_t j, uint32_t v) __attribute__((always_inline)) {
    uint8_t* a = lds + slot;
    if constexpr (RB == 4) {
      const uint32_t sh = (j & 1) << 2;
      a[j >> 1] = (uint8_t)((a[j >> 1] & ~(15u << sh)) | (v << sh));
    } else if constexpr (RB == 8) {
      a[j] = (uint8_t)v;
    } else {
      reinterpret_cast<uint16_t*>(a)[j] = (uint16_t)v;
    }
  };
  const int lane = threadIdx.x & 63;
  const uint32_t k = P.k;
  const uint32_t L2_0 = P.X.L2[0], L2_1 = P.X.L2[1], L2_2 = P.X.L2[2], L2_3 = P.X.L2[3];
  const uint32_t primary = P.X.primary;
  const uint64_t n_items = P.n_items;
  // three control words, 256 bytes apart (separate cache lines and channels):
  // pieces reserved by splitting lanes, the next piece to claim, items being
  // processed
  const gs::PersistCold* const C = P.cold;
  auto reserved = [&]() __attribute__((always_inline)) { return C->dyn_ctl; };
  auto next_piece = [&]() __attribute__((always_inline)) { return C->dyn_ctl + 64; };
  // items running: one counter per group of waves (gs::kRunSlots of them,
  // 128 bytes apart), so that no single address takes every wave's atomics;
  // a wave adds to its own and reads the total with one load per lane
  auto run_slot = [&]() __attribute__((always_inline)) -> uint32_t* {
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * BS + threadIdx.x) >> 6);
    return C->dyn_ctl + gs::kRunBase + gs::kRunStride * (wave % gs::kRunSlots);
  };
  auto running_total = [&]() __attribute__((always_inline)) -> uint32_t {
    uint32_t v = lane < (int)gs::kRunSlots ? __hip_atomic_load(C->dyn_ctl + gs::kRunBase + gs::kRunStride * lane,
                                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                           : 0u;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
  };

  bool have = false, finished = false, main_done = false;
  int kind = 0;  // 0: a whole read's stream, 1: a segment of a long read, 2: a split-off piece
  uint32_t r = 0, strand = 0, L = 0, nout = 0, last_b = 0;
  // the 22 codes left of the task end e: code(e-1-j) at bits 2j of wcode for
  // j = 0..21, and bit j of wbad set if that code is above 3 or its position
  // is left of the read start; the 13-mer pattern is the low 26 bits, an
  // extended window and its 3 presence bases the bits above, code(e-20)
  // (the base left of the 19-base window) bits 38-39, and code(e-21),
  // code(e-22) (for a second pair byte) the top four
  uint64_t wcode = 0;
  uint32_t wbad = 0;
  // pair table: probe is set while the last clean 19-base window looked up
  // was absent; the last probe also answered the windows of the task ends
  // known_e, known_e - 1, ... (their number in bits 4-5 of known_m; bit i
  // set iff the window of known_e - i occurs)
  const bool pair_on = P.X.pair && P.X.ext_len == 16 && k >= 19;
  bool probe = false;
  int32_t known_e = -1;
  uint32_t known_m = 0;
  uint32_t l = 0, u = 0, b = 0, cx = 0;
  int32_t e = 0, x = 0, e_min = 0, w0 = 0;
  uint32_t t0 = 0, first_b = 0, flags = 0, work = 0;
  uint32_t pending = kNone;                                // claimed piece not visible yet
  int32_t drun = 0;                                         // items started minus ended this iteration
  uint32_t poll_wait = 0;                                   // wave-uniform: iterations to the next poll
  uint32_t fold_wait = 0;                                   // wave-uniform: iterations since the count was folded
  bool in_tail = false;                                     // wave-uniform: items split (work splitting on)
  uint32_t backoff = 1;                                     // wave-uniform: sleeps while all lanes wait
  uint32_t link = kNone;                                   // first piece split off this item
  uint32_t item = 0;  // index of the item (or piece) in its list
  int phase = 0;
  bool need_next = false;  // the task ended: next_task at the end of the iteration
  bool ext_miss = false;  // the task's extended-table pattern is absent: use the 13-mer path
  // direct extension: single-row steps of the task so far, the text position
  // of read[b] in a run, and why a run reads the inverse suffix array (1: the
  // next step empties, 2: the read start, 3: an N at x)
  uint32_t ones = 0, tq = 0;
  int rout = 0;
  uint4* out = nullptr;
  uint32_t gcodes = 0;  // the stream's first 16-byte code chunk (a batch has fewer than 2^32)
  uint32_t n_table = 0, n_occ = 0, n_direct = 0, n_tandem = 0;  // this lane's memory operations
  // Tandem jumps.  In a stretch [tR, ..) of period tp (code(z) == code(z + tp)
  // <= 3 for z >= tR up to the stretch's end), read[y, e) for tR <= y < e
  // inside it is its last tp codes repeated, so its interval is the tandem
  // table's for that unit and length e - y.  A task end e of the stretch
  // with e - tR >= kTandemMin reads the entry of length n = min(e - tR,
  // kTandemLen) (phase 6, the entry index in tq) in place of its table lookup
  // and of the steps down to e - n, and extends from there as after a lookup;
  // if the entry's longest occurring length m is below n, the backward
  // extension ends at e - m (below 13: as a 13-mer table miss).  Skipped
  // size records only prune.  tp 0: no stretch.
  uint32_t tp = 0;
  int32_t tR = 0;
  const bool tandem_on = P.tandem != 0;

  // code of position i of the stream: the LDS window holds [w0, w0 + LMAX);
  // an extension may run left of it and then reads the 4-bit codes in
  // device memory (only in segments of long reads: a read held whole has
  // w0 = 0, and its chunks are stored in the slot's form)
  auto code = [&](int32_t i) __attribute__((always_inline)) -> uint32_t {
    const int32_t j = i - w0;
    if (j >= 0) {
      const uint32_t rr = (uint32_t)(LMAX - 1 - j);
      const uint32_t* cw = reinterpret_cast<const uint32_t*>(lds + sbase + A_BYTES);
      if ((cw[CW + (rr >> 5)] >> (rr & 31u)) & 1u) return 4u;
      return (cw[rr >> 4] >> ((rr & 15u) << 1)) & 3u;
    }
    const uint32_t w = reinterpret_cast<const uint32_t*>(P.codes + gcodes + (uint32_t)(i >> 5))[(i >> 3) & 3];
    return (w >> ((i & 7) << 2)) & 15u;
  };

  // the 22 codes left of task end ee into wcode / wbad (code(ee-1-j) at bits
  // 2j, bit j set for a code above 3 or a position left of the read start):
  // one run of bits of the slot, unless the window reaches left of a
  // segment's LDS window (then position by position, from device memory)
  // the same 22 codes and bits from the slot alone, for ee - 22 >= w0 or
  // w0 == 0 (positions left of the read start read the zero code words and
  // the all-ones bit word after the window)
  auto window_bits = [&](int32_t ee, uint64_t& w, uint32_t& bad) __attribute__((always_inline)) {
    const uint32_t r1 = (uint32_t)(w0 + LMAX - ee);  // index of position ee - 1
    const uint32_t* cw = reinterpret_cast<const uint32_t*>(lds + sbase + A_BYTES);
    const uint32_t wi = r1 >> 4, o = (r1 & 15u) << 1;
    const uint64_t lo = (uint64_t)cw[wi] | ((uint64_t)cw[wi + 1] << 32);
    const uint64_t hi = cw[wi + 2];
    w = (o ? (lo >> o) | (hi << (64 - o)) : lo) & ((1ull << 44) - 1ull);
    const uint32_t bi = CW + (r1 >> 5), bo = r1 & 31u;
    const uint64_t bb = (uint64_t)cw[bi] | ((uint64_t)cw[bi + 1] << 32);
    bad = (uint32_t)(bb >> bo) & 0x3FFFFFu;
  };
  auto load_window = [&](int32_t ee) __attribute__((always_inline)) {
    if (w0 > 0 && ee - 22 < w0) {
      uint64_t w = 0;
      uint32_t bad = 0;
#pragma unroll 1
      for (int j = 0; j < 22; ++j) {
        const int32_t pos = ee - 1 - j;
        const uint32_t c = pos >= 0 ? code(pos) : 15u;
        bad |= (c > 3 ? 1u : 0u) << j;
        w |= (uint64_t)(c & 3u) << (2 * j);
      }
      wcode = w;
      wbad = bad;
      return;
    }
    window_bits(ee, wcode, wbad);
  };
  auto end_stream = [&]() __attribute__((always_inline)) {
    const uint4 meta = make_uint4(nout, first_b, last_b, flags);
    if (kind == 2) {
      C->dyn_meta[item] = meta;
      C->dyn_link[item] = link;
    } else if (kind == 1) {
      C->seg_meta[item] = meta;
      if (P.dyn_cap) C->seg_link[item] = link;  // allocated only with splitting on
    } else if (link != kNone) {  // a whole read's stream that was split: joined later
      C->wmeta[2 * r + strand] = meta;
      C->wlink[2 * r + strand] = link;
      (strand ? C->nR : C->nF)[r] = gs::kSplitStream;
    } else {
      (strand ? C->nR : C->nF)[r] = nout;
    }
    have = false;
    --drun;  // folded into `running` once per wave and iteration
  };
  // Hands the lower half of the task ends still to do (e-1 .. e_min) to the
  // overflow queue as one piece and keeps the upper half, whose size records
  // this lane holds; the piece is linked right after this item's own tasks
  // (ahead of pieces split off earlier, which cover later task ends).  False
  // if the queue is full (then the item goes on whole).
  auto split = [&]() __attribute__((always_inline)) -> bool {
    const uint32_t n_rem = (uint32_t)(e - e_min);
    const uint32_t n_push = n_rem / 2;
    if (__hip_atomic_load(reserved(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= P.dyn_cap) return false;
    const uint32_t id = atomicAdd(reserved(), 1u);
    if (id >= P.dyn_cap) return false;  // reserved past the end: nothing to write, nobody claims it
    const uint32_t t_first = L - (uint32_t)(e_min + (int32_t)n_push - 1);  // task index of end e_min + n_push - 1
    C->dyn_items[id] = make_uint4(r, strand, t_first, n_push);
    C->dyn_link[id] = link;  // initial value: the piece's own successor
    __threadfence();
    store_release(C->dyn_ready + id, C->epoch);
    link = id;
    e_min += (int32_t)n_push;
    work = 0;
    return true;
  };
  // At a task's start: a jump (phase 6, the table entry in tq) while the
  // open stretch reaches kTandemMin codes left of e; with no stretch open,
  // opens one when the 20 codes left of e have a period and the repeat runs
  // on for kTandemSpan codes or more.
  auto t_start = [&]() __attribute__((always_inline)) {
    if (tp && e - tR < (int32_t)gs::kTandemMin) tp = 0;  // the stretch ends
    if (!tp && e >= 20 && !(wbad & 0xFFFFFu)) {
      // the smallest period up to kTandemMax (every shift a constant, one
      // select per period), kept if it is at most P.tandem (a smaller
      // period would be the smallest)
      // A period p also makes every multiple of p below 20 a period, and
      // each p <= 6 divides 4, 5 or 6: codes without those three periods
      // (nearly all) skip the rest.
      static_assert(gs::kTandemMax == 6, "the prefilter covers periods 1..6");
      const uint64_t w40 = wcode & ((1ull << 40) - 1ull);
      auto period = [&](uint32_t p) __attribute__((always_inline)) -> bool {
        return !((w40 ^ (w40 >> (2 * p))) & ((1ull << (40 - 2 * p)) - 1ull));
      };
      uint32_t q = 0;
      if (period(4) | period(5) | period(6)) {
#pragma unroll
        for (uint32_t p = gs::kTandemMax; p >= 1; --p)
          if (period(p)) q = p;
        if (q > P.tandem) q = 0;
      }
      if (q) {
        // R: z falls while code(z) == code(z + q) <= 3, 22 positions per
        // step from the slot (the first position that differs, holds a code
        // above 3 or lies left of the read start stops it), one at a time
        // near a segment window's left edge
        int32_t z = e - 21;
        while (z >= 0) {
          if (w0 > 0 && z - 21 < w0) {
            const uint32_t c = code(z);
            if (c > 3 || c != code(z + (int32_t)q)) break;
            --z;
            continue;
          }
          uint64_t wa, wb;
          uint32_t ba, bq;
          window_bits(z + 1, wa, ba);                // code(z - j) at bits 2j
          window_bits(z + 1 + (int32_t)q, wb, bq);   // code(z + q - j)
          const uint64_t d = wa ^ wb;
          const uint64_t g = (d | (d >> 1)) & 0x5555555555555555ull;  // bit 2j: position z - j differs
          const uint32_t jd = g ? (uint32_t)__builtin_ctzll(g) >> 1 : 32u;
          const uint32_t bm = ba | bq;
          const uint32_t jb = bm ? (uint32_t)__builtin_ctz(bm) : 32u;
          const uint32_t j0 = jd < jb ? jd : jb;
          if (j0 < 22u) {
            z -= (int32_t)j0;
            break;
          }
          z -= 22;
        }
        if (e - (z + 1) >= (int32_t)gs::kTandemSpan) {
          tp = q;
          tR = z + 1;
        }
      }
    }
    if (tp) {
      const uint32_t n = min((uint32_t)(e - tR), gs::kTandemLen);
      const uint32_t key = gs::tandem_key_base(tp) + ((uint32_t)wcode & ((1u << (2 * tp)) - 1u));
      tq = key * gs::kTandemRows + (n - gs::kTandemMin);
      phase = 6;
    }
  };
  // next task with a clean 13-base window, or end of the item
  auto next_task = [&]() __attribute__((always_inline)) {
    phase = 0;
    ext_miss = false;
    ones = 0;
    // split only in the tail (in_tail: the main queue is exhausted and fewer
    // items run than half the lanes), so pieces find idle lanes instead of
    // adding work; reads held whole in the window (and pieces of them) split
    // sooner: their pieces reload only the read's few code chunks
    const bool whole = L <= (uint32_t)LMAX;
    const uint32_t sw = whole ? min(P.split_work, gs::kSplitWorkWhole) : P.split_work;
    const uint32_t stk = whole ? min(P.split_tasks, gs::kSplitTasksWhole) : P.split_tasks;
    if (in_tail && work > sw && e - e_min >= 2 * (int32_t)stk) split();
    // the next task end: past those whose windows the last probe found
    // absent (their windows are clean), then past 13-base windows holding a
    // code above 3, jumping below the highest such code at once
    int32_t ne = e - 1;
    for (uint32_t ki = (uint32_t)(known_e - ne); ki < (known_m >> 4) && !((known_m >> ki) & 1u); ++ki) {
      --ne;
      probe = true;
    }
    while (true) {
      if (ne < e_min) {
        end_stream();
        return;
      }
      load_window(ne);
      const uint32_t m13 = wbad & kBad13;
      if (!m13) break;
      ne -= 32 - (int32_t)__clz(m13);
    }
    e = ne;
    b = (uint32_t)e;  // the task's left end until its lookup
    if (tandem_on) t_start();
  };
  // size record at pos (positions left of the window keep none: pruning is
  // only an early exit, so a missing record never changes the output)
  auto prune_at = [&](int32_t pos, uint32_t size) __attribute__((always_inline)) -> bool {
    const int32_t j = pos - w0;
    if (j < 0) return false;
    const uint32_t sv = size >= SAT ? SAT : size;
    if (size < SAT && rec_get(sbase, j) == sv) return true;
    rec_set(sbase, j, sv);
    return false;
  };
  auto finish_task = [&]() __attribute__((always_inline)) {
    const bool emit = (uint32_t)e - b >= k && (!P.smem || b < last_b);
    if (emit) {
      if (strand == 0)
        out[L - k - t0 - nout] = make_uint4(b, (uint32_t)e, l, u);  // task slots: L - k + 1 per strand
      else
        out[L - k + 1 + t0 + nout] = make_uint4((L - (uint32_t)e) | 0x80000000u, (L - b) | 0x80000000u, l, u);
      ++nout;
    }
    if (!(flags & 2u)) {  // the item's first finished task
      first_b = b;
      flags = 2u | (emit ? 1u : 0u);
    }
    last_b = b;
  };
  // phase 1: settle x on the next position with a base code.  Returns 0 if
  // the task goes on, 1 if it was pruned, 3 if it reached the read start.
  auto prepare_x = [&]() __attribute__((always_inline)) -> int {
    while (true) {
      if (x < 0) return 3;
      const uint32_t c = code(x);
      if (c <= 3) {
        cx = c;
        return 0;
      }
      b = (uint32_t)x;  // code > 3 after the window: interval kept, position consumed
      if (prune_at(x, u - l + 1u)) return 1;
      --x;
    }
  };
  // Direct extension.  Once the interval has kept a single row for
  // direct_min backward steps and no size record of 1 waits at x (which would
  // prune the task at once), the task reads the row's text position from the
  // full suffix array (phase 3) and compares the read with the text 64 bases
  // per 16-byte load (phase 4): while the bases agree, the pattern occurs
  // exactly once, one position further left, so each step's interval is one
  // row and its size record is 1.  A mismatch or text position 0 means the
  // next step empties.  The row where the match ends comes from the inverse
  // suffix array (phase 5) when it is needed: for a candidate that is
  // emitted, or to go on with backward steps past an N.
  auto try_direct = [&]() __attribute__((always_inline)) {
    if (P.direct_min && l == u && ones >= P.direct_min) {
      const int32_t j = x - w0;  // runs stay inside the LDS window
      if (j >= 0 && rec_get(sbase, j) != 1u) phase = 3;
    }
  };
  // how a run ends: oc 1 the next step empties, 2 the read start, 3 an N at x;
  // returns the task result (0 while the inverse read is pending)
  auto run_end = [&](int oc) __attribute__((always_inline)) -> int {
    if (oc == 3 || ((uint32_t)e - b >= k && (!P.smem || b < last_b))) {
      rout = oc;
      phase = 5;
      return 0;
    }
    return oc == 1 ? 2 : 3;  // nothing is emitted, so the row is not needed
  };
  // Loads task ends e_hi .. e_hi - tcount + 1 of stream (r, strand) into the
  // lane: the LDS code window [w0, top) covers every task's 13-base start.
  auto start_item = [&](uint32_t tstart, uint32_t tcount, uint32_t link0) __attribute__((always_inline)) {
    const uint32_t nch = (L + 31) >> 5;
    gcodes = (uint32_t)C->code_off[r] + strand * nch;
    const uint4* src = P.codes + gcodes;
    uint32_t top;
    t0 = tstart;
    if (L <= (uint32_t)LMAX) {  // the whole stream fits the window
      top = nch << 5;
      w0 = 0;
    } else {
      top = ((L - t0) + 31u) & ~31u;
      w0 = top > (uint32_t)LMAX ? (int32_t)(top - LMAX) : 0;
    }
    // each 32-code chunk of the window, reversed: chunk c holds positions
    // 32c .. 32c + 31, which get the slot indices 32q + 31 .. 32q with
    // q = LMAX/32 - 1 - (c - c0)
    const uint32_t c0 = (uint32_t)w0 >> 5, c1 = top >> 5;
    uint32_t* cw = reinterpret_cast<uint32_t*>(lds + sbase + A_BYTES);
    const bool whole = L <= (uint32_t)LMAX;  // pack_codes wrote its chunks in the slot's form
    for (uint32_t c = c0; c < c1; ++c) {
      uint4 v = src[c];
      if (!whole) v = slot_chunk(v);
      const uint32_t q = (uint32_t)(LMAX / 32 - 1) - (c - c0);
      cw[2 * q] = v.x;
      cw[2 * q + 1] = v.y;
      cw[CW + q] = v.z;
    }
    cw[CW - 2] = 0;
    cw[CW - 1] = 0;
    cw[CW + BW - 1] = 0xFFFFFFFFu;  // left of the read start (a segment's reads left of w0 go to device memory)
    const uint32_t za = ((top - (uint32_t)w0) * (uint32_t)RB / 8 + 15) >> 4;
    for (uint32_t c = 0; c < za; ++c) *reinterpret_cast<uint4*>(lds + sbase + 16 * c) = make_uint4(0, 0, 0, 0);
    out = C->cand + 2 * C->task_off[r];
    nout = 0;
    last_b = 0xFFFFFFFFu;
    first_b = 0xFFFFFFFFu;
    flags = 0;
    work = 0;
    link = link0;  // set before the item can end (a piece whose windows all hold an N ends here)
    e = (int32_t)(L - t0);
    e_min = e - (int32_t)tcount + 1;
    load_window(e);
    have = true;
    ++drun;
    phase = 0;
    ext_miss = false;
    ones = 0;
    probe = false;
    known_e = -1;
    known_m = 0;
    b = (uint32_t)e;
    tp = 0;
    need_next = (wbad & kBad13) != 0;  // no work done yet, so this never splits
    if (tandem_on && !need_next) t_start();
  };
  auto start_main = [&](uint64_t it) __attribute__((always_inline)) {
    item = (uint32_t)it;
    const uint64_t v = C->items[it];
    r = (uint32_t)v;
    strand = (uint32_t)(v >> 32) & 1u;
    const uint32_t seg = (uint32_t)(v >> 33);
    L = C->lens[r];
    if (L < k) {
      (strand ? C->nR : C->nF)[r] = 0;
      return;
    }
    // one call site of start_item for both kinds (it is long)
    uint32_t ts = 0, tc = L - k + 1;
    kind = 0;
    if (L > (uint32_t)LMAX) {
      kind = 1;
      ts = seg * C->seg_tasks;
      tc = min(C->seg_tasks, L - k + 1 - ts);
    }
    start_item(ts, tc, kNone);
  };
  auto start_piece = [&](uint32_t id) __attribute__((always_inline)) {
    const uint4 d = C->dyn_items[id];
    const uint32_t lk = C->dyn_link[id];
    item = id;
    kind = 2;
    r = d.x;
    strand = d.y;
    L = C->lens[r];
    start_item(d.z, d.w, lk);  // pieces split off this one go in front of its successor
  };
  // items started minus items ended since the wave's last fold, one atomic
  // per wave every 32 iterations, and at once when the wave has no item
  // (its count is then final); a split pushes its pieces (and fences) before
  // its item ends.  A wave's folded total never falls below zero (its items'
  // starts are folded no later than their ends), and a count that lags only
  // delays the split trigger or keeps waiting lanes polling a little longer.
  auto fold_running = [&]() __attribute__((always_inline)) {
    if ((++fold_wait & 31u) && __ballot(have) != 0) return;
    // the split trigger's state for the wave: the main queue is exhausted (a
    // lane of the wave found it so) and fewer items run than split_below
    const uint32_t total = running_total();
    in_tail = __ballot(main_done) != 0 && total < P.split_below;
    if (__ballot(drun != 0) == 0) return;
    int d = drun;
    drun = 0;
    for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
    if (lane == 0 && d) atomicAdd(run_slot(), (uint32_t)d);
  };
  // overflow queue, polled by one lane per wave: claims pieces for the
  // wave's idle lanes with one CAS, or finds that none can come any more.
  // Relaxed loads only (an early exit is always safe: a lane that splits
  // its item takes its own pieces if nobody else does).
  auto poll_pieces = [&]() __attribute__((always_inline)) {
    const bool idle = !have && !finished && main_done && pending == kNone;
    const uint64_t idlemask = __ballot(idle);
    if (!idlemask) return;
    if (poll_wait) {
      --poll_wait;
      return;
    }
    const int leader = __ffsll((unsigned long long)idlemask) - 1;
    uint32_t got = 0, first = 0, none = 0, empty = 0;
    if (lane == leader) {
      const uint32_t nx = __hip_atomic_load(next_piece(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t rs = __hip_atomic_load(reserved(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t lim = rs < P.dyn_cap ? rs : P.dyn_cap;
      if (nx < lim) {
        const uint32_t want = min((uint32_t)__popcll(idlemask), lim - nx);
        if (atomicCAS(next_piece(), nx, nx + want) == nx) {
          got = want;
          first = nx;
        }
      } else {
        empty = 1;
      }
    }
    got = __shfl(got, leader);
    first = __shfl(first, leader);
    // nothing to claim: a poller wave (one in kPollerEvery) reads the running
    // count, which decides whether any piece can still come; the others
    // leave at once, so that the waiting waves' loads of the few control
    // lines stay rare (every wave polling them slowed the items still
    // running 2.4 times)
    if (__shfl(empty, leader)) {
      const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * BS + threadIdx.x) >> 6);
      if (wave % gs::kPollerEvery != 0 || running_total() == 0) none = 1;
    }
    if (idle) {
      const uint32_t rank = (uint32_t)__popcll(idlemask & ((1ull << lane) - 1ull));
      if (rank < got)
        pending = first + rank;
      else if (none)
        finished = true;
    }
    poll_wait = got ? 0u : 256u;  // waves with busy lanes poll rarely
  };
  // a claimed piece starts once its entry is visible
  auto take_pending = [&]() __attribute__((always_inline)) {
    if (pending == kNone || have) return;
    if (__hip_atomic_load(C->dyn_ready + pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != C->epoch) return;
    __threadfence();
    const uint32_t id = pending;
    pending = kNone;
    start_piece(id);
  };

  while (true) {
    // lanes without an item take the next one of the main queue
    const bool need = !have && !finished && !main_done;
    const uint64_t needmask = __ballot(need);
    // refills go in groups: the whole wave runs the refill path (code copy,
    // record clear, first pattern) whenever any lane takes it, so idle lanes
    // wait until refill_group of them are idle or the wave has no work
    if (needmask && ((uint32_t)__popcll(needmask) >= P.refill_group || __ballot(have) == 0)) {
      const int leader = __ffsll((unsigned long long)needmask) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(P.queue, (uint32_t)__popcll(needmask));
      base = __shfl(base, leader);
      if (need) {
        const uint64_t it = (uint64_t)base + __popcll(needmask & ((1ull << lane) - 1ull));
        if (it >= n_items) {
          main_done = true;
          if (!P.dyn_cap) finished = true;
        } else {
          start_main(it);
        }
      }
    }
    if (P.dyn_cap) {
      poll_pieces();
      take_pending();
    }
    const int op = have && !need_next ? (phase == 0 ? 1 : (phase == 1 ? 2 : phase)) : 0;
    if (__ballot(op != 0 || need_next) == 0) {
      if (P.dyn_cap) fold_running();
      if (__ballot(!finished) == 0) break;
      // the whole wave waits for pieces: back off (about 3 us per sleep, up to 16)
      for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(127);
      backoff = backoff < 16 ? 2 * backoff : 16;
      poll_wait = 0;
      continue;
    }
    backoff = 1;
    // One memory operation per lane: a 16-byte load (and for op 2 perhaps a
    // second one).  The 1-, 4- and 8-byte items (pair byte, suffix-array and
    // inverse words, recorded intervals) come from the aligned 16 bytes that
    // hold them, `sub` saying which part: one load instruction and address
    // for every kind, the same single random access each.
    const uint4* a1 = nullptr;
    const uint4* a2 = nullptr;
    uint32_t sub = 0, sub2 = 0;
    bool pair_read = false;  // op 1 reads the pair byte instead of a table entry
    bool dual = false;       // and a second one
    uint32_t i1 = 0, i2 = 0, b1 = 0, b2 = 0;
    bool have1 = false;
    uint32_t tlen = gs::kPrecalc, tidx = (uint32_t)wcode & 0x3FFFFFFu;  // table lookup: pattern length and entry
    uint32_t pbit = 64;  // presence bit of the clean 3 bases left of an extended window (64: none)
    if (op == 1) {
      a1 = reinterpret_cast<const uint4*>(P.X.kmer) + (tidx >> 1);  // two 8-byte entries per load
      const uint32_t el = P.X.ext_len;
      // the extended window: codes 13 .. el-1 clean and inside the read
      if (el && !ext_miss && !(wbad & (((1u << el) - 1u) & ~kBad13))) {
        tlen = el;
        tidx = (uint32_t)(wcode & ((1ull << (2 * el)) - 1ull));
        a1 = P.X.kext + tidx;
        // presence bit c1 + 4*c2 + 16*c3 of the next three bases, if clean
        if (k >= tlen + 3 && !(wbad & (7u << el))) pbit = (uint32_t)(wcode >> (2 * el)) & 63u;
      }
      // after an absent window, a task whose 20 codes are clean and whose
      // window no probe answered reads the pair byte of its 18-base core
      // read[e-19, e-1) instead, and with read[e-22, e-20) clean too that of
      // read[e-21, e-3) in the same operation
      if (pair_on && probe && (uint32_t)(known_e - e) >= (known_m >> 4) && !(wbad & 0xFFFFFu)) {
        const uint64_t pi = (wcode >> 2) & ((1ull << 36) - 1ull);
        a1 = reinterpret_cast<const uint4*>(P.X.pair) + (pi >> 4);
        sub = (uint32_t)pi & 15u;
        pair_read = true;
        if (!(wbad & (3u << 20))) {
          const uint64_t pj = (wcode >> 6) & ((1ull << 36) - 1ull);
          a2 = reinterpret_cast<const uint4*>(P.X.pair) + (pj >> 4);
          sub2 = (uint32_t)pj & 15u;
          dual = true;
        }
      }
      n_table += dual ? 2 : 1;
      work += dual ? 2 : 1;
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
      work += a2 ? 2 : 1;
    } else if (op == 3) {  // the suffix-array word of row l (arrays padded to 16 bytes)
      a1 = reinterpret_cast<const uint4*>(P.X.sa_full) + (l >> 2);
      sub = l & 3u;
    } else if (op == 4) {
      a1 = P.X.text + ((tq - 1u) >> 6);  // tq >= 1 whenever phase 4 is entered
    } else if (op == 5) {  // the inverse word of text position tq
      a1 = reinterpret_cast<const uint4*>(P.X.isa) + (tq >> 2);
      sub = tq & 3u;
    } else if (op == 6) {  // tandem-table entry tq
      a1 = P.X.tandem + tq;
      ++n_tandem;
      ++work;
    }
    if (op >= 3 && op <= 5) {
      ++n_direct;
      ++work;
    }
    uint4 d1 = make_uint4(0, 0, 0, 0);
    if (a1) d1 = *a1;
    uint4 d2 = d1;
    if (a2) d2 = *a2;
    // ops 3 and 5: word `sub`; the pair byte: byte `sub`
    const uint32_t word = pick4(pair_read ? sub >> 2 : sub, d1.x, d1.y, d1.z, d1.w);
    const uint32_t v4 = pair_read ? (word >> ((sub & 3u) << 3)) & 255u : word;

    // Text comparison of the lanes in phase 4, one such lane at a time with
    // the whole wave: lane i checks position x - i of that lane's read against
    // text position tq - 1 - i (codes and size records from the lane's LDS
    // slot, the text from the chunk it loaded), up to the end of the chunk or
    // of the LDS window.  The first mismatch, N or size record of 1 ends the
    // comparison; lanes before it write the records of 1.  A chunk costs a
    // few wave instructions instead of a loop over 64 positions in one lane.
    uint32_t run_g = 0;  // positions matched (and recorded) by this lane's comparison
    int run_kind = 0;    // 0: all matched; 1: a mismatch, 3: an N at position x - run_g; 4: pruned there
    {
      uint64_t runners = __ballot(op == 4);
      while (runners) {
        const int R = __ffsll((unsigned long long)runners) - 1;
        runners &= runners - 1;
        const uint32_t r_tq = (uint32_t)__builtin_amdgcn_readlane((int)tq, R);
        const int32_t r_x = __builtin_amdgcn_readlane(x, R);
        const int32_t r_w0 = __builtin_amdgcn_readlane(w0, R);
        const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)d1.x, R);
        const uint32_t t1 = (uint32_t)__builtin_amdgcn_readlane((int)d1.y, R);
        const uint32_t t2 = (uint32_t)__builtin_amdgcn_readlane((int)d1.z, R);
        const uint32_t t3 = (uint32_t)__builtin_amdgcn_readlane((int)d1.w, R);
        const uint32_t tp = r_tq - 1u;  // r_tq >= 1 in phase 4
        const int32_t jx = r_x - r_w0;  // >= 0: runs start inside the window and stop at its edge
        int32_t m = (int32_t)(tp & 63u) + 1;  // positions of the chunk at or below tp
        m = m < jx + 1 ? m : jx + 1;
        const uint32_t rs = ((threadIdx.x & ~63u) + (uint32_t)R) * SLOT;  // the runner's LDS slot
        const bool valid = lane < m;
        const int32_t j = jx - lane;
        uint32_t cc = 0, rec = 0, tb = 0;
        if (valid) {
          const uint32_t rr = (uint32_t)(LMAX - 1 - j);
          const uint32_t* rcw = reinterpret_cast<const uint32_t*>(lds + rs + A_BYTES);
          cc = ((rcw[CW + (rr >> 5)] >> (rr & 31u)) & 1u) ? 4u : (rcw[rr >> 4] >> ((rr & 15u) << 1)) & 3u;
          rec = rec_get(rs, j);
          const uint32_t t = tp - (uint32_t)lane;
          tb = (pick4((t >> 4) & 3u, t0, t1, t2, t3) >> ((t & 15u) << 1)) & 3u;
        }
        const uint64_t bad = __ballot(valid && cc != tb);  // a code above 3 never equals a base
        const uint64_t nmask = __ballot(valid && cc > 3);
        const uint64_t prn = __ballot(valid && cc == tb && rec == 1u);
        const uint32_t fb = bad ? (uint32_t)__ffsll((unsigned long long)bad) - 1u : 64u;
        const uint32_t fp = prn ? (uint32_t)__ffsll((unsigned long long)prn) - 1u : 64u;
        uint32_t g = fb < fp ? fb : fp;  // the first event, 64 if none
        g = g < (uint32_t)m ? g : (uint32_t)m;
        if constexpr (RB == 4) {
          // two positions per byte: the lane of a byte's higher position (or
          // the first lane, at an even top position) writes the whole byte,
          // keeping the nibble of a position outside the matched range
          if ((uint32_t)lane < g && ((j & 1) || lane == 0)) {
            uint8_t* a = lds + rs + (j >> 1);
            if ((j & 1) && (uint32_t)lane + 1 < g)
              *a = 0x11;  // positions j and j - 1 both matched
            else
              *a = (uint8_t)((*a & ~(15u << ((j & 1) << 2))) | (1u << ((j & 1) << 2)));
          }
        } else {
          if ((uint32_t)lane < g) rec_set(rs, j, 1u);
        }
        int kind = 0;
        if (fp < fb)
          kind = 4;
        else if (fb < 64u)
          kind = ((nmask >> fb) & 1u) ? 3 : 1;
        if (lane == R) {
          run_g = g;
          run_kind = kind;
        }
      }
    }

    // the new interval: a table entry (op 1) or one backward step (op 2);
    // both then share one copy of the task logic, so a wave whose lanes are
    // in different phases runs it once
    uint32_t nl = 0, nu = 0, nb = 0;
    bool absent = false;  // the clean (tlen + 3)-base window does not occur
    if (op == 1) {
      if (tlen > gs::kPrecalc) {
        nl = d1.x;
        nu = d1.y;
        absent = pbit < 64 && !(((pbit < 32 ? d1.z : d1.w) >> (pbit & 31)) & 1u);
      } else {
        nl = (tidx & 1u) ? d1.z : d1.x;
        nu = (tidx & 1u) ? d1.w : d1.y;
      }
      nb = (uint32_t)e - tlen;
    } else if (op == 2) {
      const uint32_t o2 = gs::rank_in_entry(d1, i2 - b2 * gs::kBlockBases);
      const uint32_t o1 = have1 ? gs::rank_in_entry(d2, i1 - b1 * gs::kBlockBases) : 0u;
      const uint32_t base = pick4(cx, L2_0, L2_1, L2_2, L2_3);
      nl = base + o1 + 1u;
      nu = base + o2;
      nb = (uint32_t)x;
    } else if (op == 6) {  // the entry's interval of length m = d1.z
      nl = d1.x;
      nu = d1.y;
      nb = (uint32_t)e - d1.z;
    }
    // How the task goes on: 0 continues, 1 pruned or a table miss (no
    // candidate, 2145), 2 its interval emptied, 3 it reached the read start;
    // `prep`: settle x on the next base (phase 1) and see whether a direct
    // run starts.  Each tail below has one copy for every lane.
    int res = 0;
    bool prep = false;
    if (pair_read) {
      // the pair byte: bit code(e-1) says whether this task's window
      // read[e-19, e) occurs, bit 4 + code(e-20) whether the next task's
      // window read[e-20, e-1) does; an absent window ends the task as an
      // absent presence bit would, a present one is looked up next
      // a second byte's bit code(e-3) and bit 4 + code(e-22) answer the
      // windows of e-2 and e-3
      known_e = e - 1;
      uint32_t m = (v4 >> (4u + ((uint32_t)(wcode >> 38) & 3u))) & 1u;
      if (dual) {
        const uint32_t vb = (pick4(sub2 >> 2, d2.x, d2.y, d2.z, d2.w) >> ((sub2 & 3u) << 3)) & 255u;
        m |= ((vb >> ((uint32_t)(wcode >> 4) & 3u)) & 1u) << 1;
        m |= ((vb >> (4u + ((uint32_t)(wcode >> 42) & 3u))) & 1u) << 2;
      }
      known_m = m | ((dual ? 3u : 1u) << 4);
      if (!((v4 >> ((uint32_t)wcode & 3u)) & 1u))
        res = 1;
      else
        probe = false;
    } else if (op >= 3 && op <= 5) {
      if (op == 3) {  // the text position of read[b]
        tq = v4;
        if (tq == 0)
          res = run_end(1);  // the pattern starts the text: no step can extend it
        else
          phase = 4;
      } else if (op == 4) {  // the wave compared this lane's chunk above
        int oc = 0;  // 0: chunk used up, load the next one
        if (run_kind == 4) {
          oc = 4;
          b = (uint32_t)(x - (int32_t)run_g);  // pruned at the position after the matched ones
        } else {
          if (run_g) {
            b = (uint32_t)(x - (int32_t)run_g + 1);  // one row, run_g positions further left
            tq -= run_g;
            x = (int32_t)b - 1;
          }
          if (run_kind)
            oc = run_kind;  // 1: the step at x empties; 3: x holds an N
          else if (x < 0)
            oc = 2;
          else if (x < w0)
            oc = 3;  // left of the window: backward steps go on from the row read back
          else if (code(x) > 3)
            oc = 3;  // checked before the next chunk's load, which an N does not need
          else if (tq == 0)
            oc = 1;  // the pattern now starts the text: the next step empties
        }
        if (oc == 4)
          res = 1;
        else if (oc != 0)
          res = run_end(oc);
      } else {  // op 5: the row of read[b .. e)
        l = u = v4;
        if (rout == 3)
          prep = true;  // consumes the N as a backward step would
        else
          res = rout == 1 ? 2 : 3;
      }
    } else if (op != 0) {
      // a tandem entry whose repeat occurs for fewer than 13 bases ends the
      // task as a 13-mer table miss would
      if (nl > nu || (op == 6 && d1.z < (uint32_t)gs::kPrecalc)) {
        // An absent clean tlen-base window: the task's candidate could only be
        // shorter than tlen.  With k >= tlen it can emit nothing, and what
        // else it would leave cannot change the output (left ends of finished
        // tasks do not grow as the end falls, so a later candidate of length
        // >= k lies left of this task's left end and of the previous one's;
        // size records only prune), so it ends here.  Otherwise the 13-mer
        // lookup and steps decide.
        if (op == 1 && tlen > gs::kPrecalc && k < tlen)
          ext_miss = true;
        else
          res = op == 2 ? 2 : 1;
        if (pair_on && op == 1 && pbit < 64) probe = true;  // an absent clean 19-base window
      } else if (absent) {
        res = 1;  // the same holds for an absent clean (tlen + 3)-base window with k >= tlen + 3
        if (pair_on) probe = true;
      } else {
        // a present clean 19-base window (a jump's window occurs too)
        if (pair_on && ((op == 1 && pbit < 64) || op == 6)) probe = false;
        if (op == 2) ones = nl == nu ? ones + 1u : 0u;
        l = nl;
        u = nu;
        b = nb;
        if (prune_at((int32_t)b, u - l + 1u)) {
          res = 1;
        } else if (op == 6 && d1.z < min((uint32_t)(e - tR), gs::kTandemLen)) {
          res = 2;  // the repeat one base longer does not occur: the extension ends here
        } else {
          x = (int32_t)b - 1;
          prep = true;
        }
      }
    }
    if (prep) {
      phase = 1;
      res = prepare_x();
      if (res == 0) try_direct();
    }
    if (res >= 2) finish_task();
    if (res == 3 && P.smem)
      end_stream();  // later tasks of the item end at the read start too and are filtered out
    else if (res != 0)
      need_next = true;
    if (need_next) {  // the one place a task ends into the next one
      need_next = false;
      next_task();
    }
    if (P.dyn_cap) fold_running();
  }
  if (C->counters) {
    // wave totals, one atomic per wave
    uint64_t c0 = n_table, c1 = n_occ, c2 = n_direct, c3 = n_tandem;
    for (int d = 32; d > 0; d >>= 1) {
      c0 += __shfl_down(c0, d, 64);
      c1 += __shfl_down(c1, d, 64);
      c2 += __shfl_down(c2, d, 64);
      c3 += __shfl_down(c3, d, 64);
    }
    if (lane == 0) {
      atomicAdd((unsigned long long*)&C->counters[0], (unsigned long long)c0);
      atomicAdd((unsigned long long*)&C->counters[1], (unsigned long long)c1);
      atomicAdd((unsigned long long*)&C->counters[2], (unsigned long long)c2);
      atomicAdd((unsigned long long*)&C->counters[3], (unsigned long long)c3);
    }
  }
}

// One piece of a stream (its task ends from t0 on, `m` = {candidates,
// first finished b, last finished b, flags}) moved next to the stream's
// earlier pieces; K candidates so far, LB = left end of the last finished
// task so far.  A piece filters its own tasks in the lane, but its first
// finished task had no predecessor there: it survives iff its left end is
// smaller than LB (finish_task, 991-1018).  Candidate i moves from t0 + i to
// K + i - drop (counted from the end the strand's region is packed against),
// never toward a higher source index, so a candidate's destination is a
// source that was read before it: moving `width` consecutive candidates at
// once (all reads of a group before its writes) keeps that order.  width is
// 1 for one thread, 64 for a wave whose lanes hold the same K and LB.
__device__ __forceinline__ void join_piece(uint4* out, uint32_t ntask, uint32_t strand, uint32_t t0, uint4 m,
                                           uint32_t smem, uint32_t& K, uint32_t& LB, uint32_t lane = 0,
                                           uint32_t width = 1) {
  const uint32_t drop = (smem && (m.w & 1u) && !(m.y < LB)) ? 1u : 0u;
  for (uint32_t i0 = drop; i0 < m.x; i0 += width) {
    const uint32_t i = i0 + lane;
    if (i < m.x) {
      if (strand == 0)
        out[ntask - 1 - K - (i - drop)] = out[ntask - 1 - t0 - i];
      else
        out[ntask + K + (i - drop)] = out[ntask + t0 + i];
    }
  }
  K += m.x - drop;
  if (m.w & 2u) LB = m.z;
}

// the pieces linked after an item, in task order (at most dyn_cap of them)
__device__ __forceinline__ void join_chain(const gs::JoinParams& J, uint4* out, uint32_t ntask, uint32_t strand,
                                           uint32_t link, uint32_t& K, uint32_t& LB, uint32_t lane = 0,
                                           uint32_t width = 1) {
  for (uint32_t guard = 0; link != kNone && guard < J.dyn_cap; ++guard) {
    join_piece(out, ntask, strand, J.dyn_items[link].z, J.dyn_meta[link], J.smem, K, LB, lane, width);
    link = J.dyn_link[link];
  }
}

// Streams of long reads, one wave per stream at a time (a 20 kbp read has
// about 20,000 candidates per strand to move): segments in order, each
// followed by the chain of pieces split off it; every lane follows the same
// metadata and the lanes move 64 candidates at a time.  A fixed grid of
// waves takes the streams in turn (a workgroup per four streams left the
// launch bound by workgroup dispatch), and the metadata of 64 segments is
// loaded at once, one segment per lane, then read from the lanes in order.
__global__ void __launch_bounds__(256) join_segments(gs::JoinParams J) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n_streams = 2ull * J.n_long, n_waves = gridDim.x * 4ull;
  for (uint64_t t = blockIdx.x * 4ull + (threadIdx.x >> 6); t < n_streams; t += n_waves) {
    const uint32_t r = J.long_reads[t >> 1], strand = (uint32_t)(t & 1);
    const uint32_t L = J.lens[r];
    const uint32_t ntask = L - J.k + 1;
    const uint32_t S = J.seg_tasks;
    const uint32_t nseg = (ntask + S - 1) / S;
    uint4* out = J.cand + 2 * J.task_off[r];
    const uint64_t i0 = J.item0[t >> 1] + strand;
    uint32_t K = 0, LB = 0xFFFFFFFFu;
    for (uint32_t g0 = 0; g0 < nseg; g0 += 64) {
      uint4 mine = make_uint4(0, 0, 0, 0);
      uint32_t link = kNone;
      if (g0 + lane < nseg) {
        const uint64_t it = i0 + 2 * (uint64_t)(g0 + lane);
        mine = J.seg_meta[it];
        if (J.dyn_cap) link = J.seg_link[it];
      }
      const uint32_t ng = min(64u, nseg - g0);
      for (uint32_t j = 0; j < ng; ++j) {
        const uint4 m = make_uint4(__shfl(mine.x, j), __shfl(mine.y, j), __shfl(mine.z, j), __shfl(mine.w, j));
        const uint32_t lk = __shfl(link, j);
        join_piece(out, ntask, strand, (g0 + j) * S, m, J.smem, K, LB, lane, 64);
        if (J.dyn_cap) join_chain(J, out, ntask, strand, lk, K, LB, lane, 64);
      }
    }
    if (lane == 0) (strand ? J.nR : J.nF)[r] = K;
  }
}

// Whole-read streams that were split (one thread per stream of the batch;
// the others were final already).
__global__ void __launch_bounds__(256) join_whole(gs::JoinParams J) {
  const uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t >= 2ull * J.n_reads) return;
  const uint32_t r = (uint32_t)(t >> 1), strand = (uint32_t)(t & 1);
  uint32_t* const n_out = strand ? J.nR : J.nF;
  if (n_out[r] != gs::kSplitStream) return;
  const uint32_t ntask = J.lens[r] - J.k + 1;
  uint4* out = J.cand + 2 * J.task_off[r];
  uint32_t K = 0, LB = 0xFFFFFFFFu;
  join_piece(out, ntask, strand, 0, J.wmeta[t], J.smem, K, LB);
  join_chain(J, out, ntask, strand, J.wlink[t], K, LB);
  n_out[r] = K;
}

using gs::pack_chunk;

// reads of length <= kPersistLmax: one thread per (read, strand, chunk),
// each chunk stored in the form the persistent kernel's LDS slot holds
// (slot_chunk), which a refill then copies without converting
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
  codes[code_off[r] + strand * nch + c] = slot_chunk(pack_chunk(bases + offs[r], L, strand, c));
}

// longer reads: a fixed grid of waves taking the reads in turn, a wave's
// lanes one chunk each (a workgroup per read left the launch bound by
// workgroup dispatch: 1.4 M workgroups of about a hundred chunks each)
__global__ void __launch_bounds__(256) pack_codes_long(const uint8_t* __restrict__ bases,
                                                       const uint64_t* __restrict__ offs,
                                                       const uint32_t* __restrict__ lens,
                                                       const uint64_t* __restrict__ code_off,
                                                       const uint32_t* __restrict__ long_reads, uint32_t n_long,
                                                       uint4* __restrict__ codes) {
  const uint32_t lane = threadIdx.x & 63, n_waves = gridDim.x * 4;
  for (uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6); w < n_long; w += n_waves) {
    const uint32_t r = long_reads[w];
    const uint32_t L = lens[r];
    const uint32_t nch = (L + 31) >> 5;
    const uint8_t* s = bases + offs[r];
    uint4* dst = codes + code_off[r];
    for (uint32_t q = lane; q < 2 * nch; q += 64) {
      const uint32_t strand = q >= nch ? 1u : 0u;
      dst[q] = pack_chunk(s, L, strand, q - strand * nch);
    }
  }
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

// one thread: the launch's cold fields into their device slot
__global__ void write_cold(gs::PersistCold c, gs::PersistCold* dst) { *dst = c; }

template <typename AT, int LMAX, int BS>
int launch_t(const gs::PersistParams& P, unsigned grid, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  gs::PersistCold c;
  c.code_off = P.code_off;
  c.lens = P.lens;
  c.task_off = P.task_off;
  c.items = P.items;
  c.cand = P.cand;
  c.nF = P.nF;
  c.nR = P.nR;
  c.seg_meta = P.seg_meta;
  c.seg_link = P.seg_link;
  c.counters = P.counters;
  c.dyn_items = P.dyn_items;
  c.dyn_meta = P.dyn_meta;
  c.dyn_link = P.dyn_link;
  c.dyn_ready = P.dyn_ready;
  c.dyn_ctl = P.dyn_ctl;
  c.wmeta = P.wmeta;
  c.wlink = P.wlink;
  c.seg_tasks = P.seg_tasks;
  c.epoch = P.epoch;
  gs::PersistCold* dst = reinterpret_cast<gs::PersistCold*>(reinterpret_cast<uint8_t*>(P.counters) + gs::kCounterBytes);
  hipLaunchKernelGGL(write_cold, dim3(1), dim3(1), 0, st, c, dst);
  PersistArgs A;
  A.X = P.X;
  A.codes = P.codes;
  A.cold = dst;
  A.queue = P.queue;
  A.n_items = P.n_items;
  A.k = P.k;
  A.smem = P.smem;
  A.direct_min = P.direct_min;
  A.dyn_cap = P.dyn_cap;
  A.split_work = P.split_work;
  A.split_tasks = P.split_tasks;
  A.split_below = grid * (uint32_t)BS / 2;  // half the lanes of the launch
  A.refill_group = gs::kRefillGroup;
  A.tandem = P.tandem;
  if (e0 && hipEventRecord(e0, st) != hipSuccess) return GS_EHIP;
  hipLaunchKernelGGL((seed_persistent<AT, LMAX, BS>), dim3(grid), dim3(BS), 0, st, A);
  if (hipGetLastError() != hipSuccess) return GS_EHIP;
  return e1 && hipEventRecord(e1, st) != hipSuccess ? GS_EHIP : GS_OK;
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
    hipLaunchKernelGGL(pack_codes_long, dim3(std::min<uint32_t>((n_long + 3) / 4, kHelperGrid)), dim3(256), 0, st,
                       bases, offs, lens, code_off, long_reads, n_long, codes);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_EHIP;
}

int launch_join(const JoinParams& J, hipStream_t st) {
  if (J.n_long) {
    const uint64_t waves = 2ull * J.n_long;  // one stream per wave at a time, four waves per workgroup
    hipLaunchKernelGGL(join_segments, dim3((unsigned)std::min<uint64_t>((waves + 3) / 4, kHelperGrid)), dim3(256), 0,
                       st, J);
  }
  if (J.dyn_cap && J.n_reads) {
    const uint64_t threads = 2ull * J.n_reads;
    hipLaunchKernelGGL(join_whole, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, J);
  }
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
    case 2064: return grid_t<Nib, L, 64>(grid);
    case 2128: return grid_t<Nib, L, 128>(grid);
    case 2256: return grid_t<Nib, L, 256>(grid);
    default: return GS_EINVAL;
  }
}

int launch_persistent(const PersistParams& P, int variant, int wg, unsigned grid, hipStream_t st, hipEvent_t e0,
                      hipEvent_t e1) {
  if (!grid) return GS_EINVAL;
  hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, st, P.queue, 1u);
  if (P.dyn_cap) {  // an empty overflow queue: no piece reserved, claimed or written
    hipLaunchKernelGGL(zero_words, dim3((gs::kDynCtlWords + 63) / 64), dim3(64), 0, st, P.dyn_ctl, gs::kDynCtlWords);
    hipLaunchKernelGGL(zero_words, dim3((P.dyn_cap + 63) / 64), dim3(64), 0, st, P.dyn_ready, P.dyn_cap);
  }
  if (!P.counters) return GS_EINVAL;  // its buffer also holds the launch's PersistCold
  hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, st, (uint32_t*)P.counters, 8u);
  constexpr int L = kPersistLmax;
  switch (variant * 1000 + (wg ? wg : 128)) {
    case 64: return launch_t<uint8_t, L, 64>(P, grid, st, e0, e1);
    case 128: return launch_t<uint8_t, L, 128>(P, grid, st, e0, e1);
    case 256: return launch_t<uint8_t, L, 256>(P, grid, st, e0, e1);
    case 1064: return launch_t<uint16_t, L, 64>(P, grid, st, e0, e1);
    case 1128: return launch_t<uint16_t, L, 128>(P, grid, st, e0, e1);
    case 1256: return launch_t<uint16_t, L, 256>(P, grid, st, e0, e1);
    case 2064: return launch_t<Nib, L, 64>(P, grid, st, e0, e1);
    case 2128: return launch_t<Nib, L, 128>(P, grid, st, e0, e1);
    case 2256: return launch_t<Nib, L, 256>(P, grid, st, e0, e1);
    default: return GS_EINVAL;
  }
}

}  // namespace gs
