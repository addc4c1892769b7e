// Host-side internals shared by the HIP translation units of libgpuseed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <utility>
#include <vector>

#include "gpuseed.h"
#include "gs_device.h"

#define GS_HIP(call)                        \
  do {                                      \
    hipError_t e_ = (call);                 \
    if (e_ != hipSuccess) return GS_EHIP;   \
  } while (0)

struct gs_index {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t primary = 0, seq_len = 0, sa_intv = 0;
  uint32_t L2[5] = {0, 0, 0, 0, 0};
  uint64_t n_blocks = 0;  // 96-base blocks
  uint64_t n_sa = 0;
  uint4* occ = nullptr;
  uint32_t* sa = nullptr;
  uint2* kmer = nullptr;
  uint4* kext = nullptr;  // extended interval table with presence bits (gs_index_set_kmer_len)
  uint32_t ext_len = 0;
  // direct-extension structures (gs_index_set_direct): full suffix array and
  // its inverse (seq_len + 1 words each), the text at 2 bits per base
  uint32_t* sa_full = nullptr;
  uint32_t* isa = nullptr;
  uint32_t* text = nullptr;
  uint64_t text_words = 0;
  // rows whose reference walk passes row primary+1 and their reported
  // positions (locate from the full suffix array); n_quirk > kMaxQuirk: walk
  uint2* quirk = nullptr;
  uint32_t n_quirk = 0;
  uint8_t* pair = nullptr;  // pair table (gs_index_set_pairs), 4^18 bytes
  uint4* tandem = nullptr;  // tandem table, built with the 13-mer table
  double t_pairs = 0;
  uint64_t device_bytes = 0;
  double t_read = 0, t_upload = 0, t_repack = 0, t_table = 0, t_sa = 0, t_write = 0, t_ext = 0, t_direct = 0;
  int sa_rounds = 0;
  gs_batch* work = nullptr;  // gs_seed's batch, kept for the next call (its buffers only grow)
  // Lifetime: the caller's handle holds one reference, every batch and
  // server made on the index one more; gs_index_close drops the caller's
  // and the device structures are freed with the last one, so a batch or
  // server that outlives the caller's handle still reads valid memory.
  std::atomic<int> refs{1};
  std::atomic<int> servers{0};  // live servers: their workers launch at any time
  gs::DevIndex dev() const;
};

namespace gs {
void index_acquire(gs_index* X);
void index_release(gs_index* X);  // frees the index with its last reference
// Work-items per launch of the chunked kernels (table levels, presence
// bits, direct-structure chains, suffix-array walk): a dispatch's work-item
// count is a 32-bit quantity, so launches over more items than this are
// split.  2^30 unless a test sets it smaller (gs_set_launch_chunk).
uint64_t launch_chunk();
// allocation of result arrays handed to callers (freed with free())
void* result_alloc(size_t bytes);
// Looks up the device address of gs_seed_noalloc's scratch array on `device`
// (which must be current); called once per device where an index is set up.
int noalloc_setup(int device);
}  // namespace gs

namespace gs {

// Parameters of the persistent extension kernel (gs_persist.hip).
//
// Work items: one per (read, strand) for reads of at most kPersistLmax bases;
// reads longer than that are cut into segments of seg_tasks consecutive task
// ends, one item per (read, segment, strand).  items[i] = read | (segment << 33)
// | (strand << 32); the items of a read are adjacent, segment-major.
// The fields of a persistent launch that the kernel's loop uses only at
// item starts and ends, in work splitting and for its counters.  The launch
// writes them to device memory (the 256 bytes after the counters) and the
// kernel reads them with scalar loads where it uses them, so that its
// arguments fit the scalar registers instead of spilling into vector ones.
struct PersistCold {
  const uint64_t* code_off;
  const uint32_t* lens;
  const uint64_t* task_off;
  const uint64_t* items;
  uint4* cand;
  uint32_t* nF;
  uint32_t* nR;
  uint4* seg_meta;
  uint32_t* seg_link;
  uint64_t* counters;
  uint4* dyn_items;
  uint4* dyn_meta;
  uint32_t* dyn_link;
  uint32_t* dyn_ready;
  uint32_t* dyn_ctl;
  uint4* wmeta;
  uint32_t* wlink;
  uint32_t seg_tasks, epoch;
};
constexpr uint32_t kCounterBytes = 32;   // four 64-bit counters
constexpr uint32_t kColdBytes = 256;     // PersistCold, after the counters
static_assert(sizeof(PersistCold) <= kColdBytes, "PersistCold must fit its slot");

struct PersistParams {
  DevIndex X;
  const uint4* codes;        // 4-bit virtual codes, per read: forward chunks then reverse chunks
  const uint64_t* code_off;  // per read, in 16-byte chunks
  const uint32_t* lens;
  const uint64_t* task_off;
  const uint64_t* items;
  uint64_t n_items;
  uint32_t n_reads, k, smem, seg_tasks;
  uint4* cand;               // same candidate layout as the plain kernel
  uint32_t* nF;
  uint32_t* nR;
  uint4* seg_meta;           // per item of a segmented read: {emitted, first b, last b, flags}
  uint32_t* seg_link;        // per item of a segmented read: first piece split off it, or ~0
  uint32_t* queue;           // work-queue counter
  uint64_t* counters;        // optional: [0] table reads, [1] occ entry reads, [2] direct-extension reads,
                             // [3] tandem-table reads
  uint32_t direct_min;       // size-1 steps before a task reads the text (0: no direct extension)
  // work splitting (dyn_cap 0: off)
  uint4* dyn_items;          // pieces {read, strand, first task, tasks}
  uint4* dyn_meta;           // per piece, as seg_meta
  uint32_t* dyn_link;        // per piece: the next piece in task order, or ~0
  uint32_t* dyn_ready;       // per piece: == epoch once its entry is written (cleared per launch)
  uint32_t* dyn_ctl;         // [0] pieces reserved, [64] next piece to claim, [kRunBase + kRunStride * i]
                             // items running, counter i of kRunSlots (kDynCtlWords words)
  uint4* wmeta;              // per whole-read stream (2 * read + strand) that was split
  uint32_t* wlink;
  uint32_t dyn_cap, split_work, split_tasks, epoch;
  uint32_t split_below;           // split only while fewer items than this run (set at launch)
  uint32_t refill_group;          // idle lanes of a wave that trigger a refill (set at launch)
  uint32_t tandem;                // tandem jumps: largest period (0: off; needs the index's tandem table)
};

// Joins pieces (segments of long reads, split-off pieces) into the per-read
// candidate layout.
struct JoinParams {
  const uint32_t* long_reads;   // reads cut into segments
  const uint64_t* item0;        // per long read: index of its first item
  const uint32_t* lens;
  const uint64_t* task_off;
  const uint4* seg_meta;
  const uint32_t* seg_link;
  const uint4* dyn_items;
  const uint4* dyn_meta;
  const uint32_t* dyn_link;
  const uint4* wmeta;
  const uint32_t* wlink;
  uint4* cand;
  uint32_t* nF;
  uint32_t* nR;
  uint32_t n_long, k, smem, seg_tasks;
  uint32_t n_reads, dyn_cap, split_tasks, pad;
};

constexpr uint32_t kSplitStream = 0xFFFFFFFFu;  // nF/nR of a whole-read stream left for join_whole
// Work splitting by default only in batches of at most this many items per
// lane of the persistent grid: on the 3.1 Gbp index, 80,000 150 bp reads
// (0.6 items per lane) 7.5 ms with splitting, 11 ms without; 160,000 (1.2 per
// lane) 18.8 and 10.7 ms; 10,000 mixed-length reads (1.6 per lane) 7.5 and
// 6.2 ms.
constexpr uint32_t kSplitItemsPerLane = 1;
// Split thresholds for reads held whole in the LDS window (memory operations,
// task ends); the options' values apply when smaller.  A launch of 10,000
// 150 bp reads on the 3.1 Gbp index: 8.8 ms with the defaults of long reads
// (2048, 8), 3.7 ms with (128, 4), 3.6 ms with (64, 2); 40,000 reads: 8.8,
// 4.9 and 9.7 ms.  Long reads' segments keep the options' values: with
// (128, 4) for every item a mixed-length launch of 10,000 reads took 12.9
// against 9.3 ms.
constexpr uint32_t kSplitWorkWhole = 128;
constexpr uint32_t kSplitTasksWhole = 4;
// The running-items count of work splitting, spread over kRunSlots counters
// 128 bytes apart: waves add to counter (wave mod kRunSlots), readers sum
// them.  One counter, folded once per wave and iteration, serialised every
// wave's atomics on one address: 10,000 mixed-length reads took 9.1 ms per
// launch with splitting on against 6.3 ms with it off.
constexpr uint32_t kRunBase = 128, kRunStride = 32, kRunSlots = 16;
constexpr uint32_t kDynCtlWords = kRunBase + kRunStride * kRunSlots;
// Waves that wait for split-off pieces once they have nothing to do: one in
// kPollerEvery (about one per CU); the others leave at their first poll
// that finds no piece.  On the 3.1 Gbp index, 80,000 150 bp reads: 8.0 ->
// 7.0 ms per launch; 2,000 mixed-length reads: 4.5 -> 3.8 ms.  (Forced on
// in a launch of 2,000,000 150 bp reads, 15 items per lane, splitting after
// 128 operations makes it twice as long, 51 against 23 ms: the pieces start
// without size records; the default keeps it off at that size.)
constexpr uint32_t kPollerEvery = 16;
constexpr uint32_t kRefillGroup = 12;           // idle lanes of a wave that trigger a refill
                                                // (2.5M 150 bp reads, 3.1 Gbp index, an early
                                                // form of the kernel: 1 lane 72.5 ms, 4: 58.0 ms,
                                                // 8: 58.2 ms, 16: 59.6 ms; 10M reads with the
                                                // current one, alternating runs on one box: 8:
                                                // 68.7-69.0 ms, 12: 67.6-68.1 ms median)

constexpr uint32_t kPersistLmax = 160;  // LDS code window of a persistent lane, in bases
// Tandem jumps of the persistent kernel: periods 1..kTandemMax; a stretch
// [R, E) opens at a task end E whose 20 codes to the left have a period and
// spans at least kTandemSpan positions; a task end e of the stretch with
// e - R >= kTandemMin reads the tandem table's row of length
// min(e - R, kTandemLen).  With the 20 codes of the rolling window a period
// is at most 19; the configurations' tandem repeats have periods 1-6.
constexpr uint32_t kTandemMax = 6;
constexpr uint32_t kTandemSpan = 24;
constexpr uint32_t kTandemMin = 20;
constexpr uint32_t kTandemLen = 256;
// The tandem table: for every unit of 1..kTandemMax bases (key base(p) +
// sum of code_j * 4^j, the unit's last base as j = 0, base(p) = (4^p - 4) / 3)
// and every length n in [kTandemMin, kTandemLen], the 16-byte entry
// {l, u, m, 0} of the string of length n that ends with the unit repeated
// (the unit's last base at its end): m = the longest length <= n whose
// string occurs, [l, u] the interval of that length.
constexpr uint32_t kTandemKeys = (4096u * 4u - 4u) / 3u;  // 5460
constexpr uint32_t kTandemRows = kTandemLen - kTandemMin + 1;
__host__ __device__ constexpr uint32_t tandem_key_base(uint32_t p) { return ((1u << (2 * p)) - 4u) / 3u; }
constexpr uint32_t kSegTasksMax = 112;  // task ends per segment: the window keeps every 13-base start

// segments of a read of length L (1 for reads held whole in the LDS window)
__host__ __device__ inline uint32_t n_segments(uint32_t L, uint32_t k, uint32_t seg_tasks) {
  if (L <= kPersistLmax || L < k) return 1;
  const uint32_t ntask = L - k + 1;
  return (ntask + seg_tasks - 1) / seg_tasks;
}

int launch_pack_codes(const uint8_t* bases, const uint64_t* offs, const uint32_t* lens, const uint64_t* code_off,
                      uint32_t n_reads, const uint32_t* long_reads, uint32_t n_long, uint4* codes, hipStream_t st);
// variant 0: 8-bit size records, 1: 16-bit size records; wg 64, 128 or 256
// (0 = 128); grid from persistent_grid (called outside any graph capture)
int persistent_grid(int variant, int wg, unsigned* grid);
// e0, e1 (optional): recorded on st right before and after seed_persistent
int launch_persistent(const PersistParams& P, int variant, int wg, unsigned grid, hipStream_t st,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
int launch_join(const JoinParams& J, hipStream_t st);

}  // namespace gs

namespace gs {

// The CPU side of a batch (see gs_batch_create): built on any thread, then
// uploaded by batch_load.  Per read only; the work items of the persistent
// kernel (n_items of them, n_long reads cut into segments) are written on the
// device when the batch is loaded.
struct BatchPlan {
  gs_opts opts{};
  uint32_t seg_tasks = 64, maxlen = 0, next_batch_id = 0;
  uint64_t n_reads = 0, span = 0, tasks = 0, code_chunks = 0, n_items = 0, n_long = 0;
  std::vector<std::pair<const char*, uint64_t>> pieces;  // bytes uploaded back to back
  std::vector<uint64_t> offs, task_off, code_off;
  std::vector<uint32_t> lens, rbatch;
};

int plan_check(const gs_opts* o);
void plan_reset(BatchPlan& P, const gs_opts& o);
// appends reads [r0, r1) of one source as a group of their own
int plan_add(BatchPlan& P, const char* bases, const uint64_t* offsets, const uint32_t* lens, uint64_t r0,
             uint64_t r1);
void plan_finish(BatchPlan& P);
uint64_t plan_device_bytes(const BatchPlan& P);  // device buffers the plan needs, outputs excluded
// hold: the batch keeps a reference to X (every batch but gs_seed's own,
// which the index owns and frees)
gs_batch* batch_new(gs_index* X, bool hold = true);
int batch_load(gs_batch* B, const BatchPlan& P);
// Grows B's device buffers for any plan of at most `bases` bases and `reads`
// reads, so that loading such plans allocates (and synchronises) nothing.
int batch_reserve(gs_batch* B, uint64_t bases, uint64_t reads, uint32_t seg_tasks);
void batch_free(gs_batch* B);

// Results of the batch's last run, downloaded into pinned staging that the
// batch keeps (valid until the next download or the batch's end).
struct ResultView {
  uint64_t n_reads, n_seeds, n_pos;
  const uint32_t *nseeds, *npos;
  const uint64_t *soff, *poff;
  const gs_seed_rec* seeds;
  const uint32_t* pos;
};
int batch_download(gs_batch* B, ResultView* v);

}  // namespace gs

// Tuning fields of an existing batch (the autotuner's knobs that need no new
// work items): kernel variant, workgroup size, suffix-array walk rows.
int gs_batch_set_tuning(gs_batch* B, int kernel, int wg, int walk_rows);
int gs_batch_set_direct_min(gs_batch* B, uint16_t direct_min);  // as gs_opts.direct_min

// ceil(2^64 / d) for d >= 2 (exact quotient for every 32-bit numerator), 0 for d == 1
uint64_t gs_reciprocal64(uint32_t d);

// Builds the 4^13 interval table on the index's stream (pre_calc_seed_intervals_gpu,
// seed_gen.cu:2648-2675), level by level.
int gs_build_kmer_table(gs_index* X);
