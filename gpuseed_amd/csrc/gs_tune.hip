// Autotuner: picks the batched pipeline's launch parameters on a sample of
// the caller's reads by timing the whole device pipeline (HIP events).
//
// Knobs, searched one after another (each step keeps the best so far):
//  1. seg_tasks (task ends per work item of reads longer than the LDS window;
//     only if the sample has such reads): 32, 64, 96, 112;
//  2. kernel variant (4-, 8- or 16-bit size records, which sets the LDS slot and
//     so the waves per CU) x workgroup size 64 / 128 / 256;
//  3. suffix-array walk rows per thread: 1, 2, 4;
//  4. with the index's direct-extension structures: single-row steps before a
//     task reads the text, 4 (the default), 2, 8, or never.
// Every candidate must produce the same number of seeds and positions as the
// first run; the output of the pipeline depends on none of the knobs.
#include <string.h>

#include "gs_internal.h"

namespace {

struct Timer {
  gs_batch* B = nullptr;
  int reps = 3;
  uint64_t seeds = 0, pos = 0;
  bool have_ref = false;
  int tried = 0;
  // best pipeline time of `reps` runs, or a negative code
  int time(double* best) {
    *best = 1e30;
    for (int r = 0; r < reps; ++r) {
      gs_batch_stats s;
      int rc = gs_batch_run(B, -1, &s);
      if (rc) return rc;
      if (!have_ref) {
        seeds = s.n_seeds;
        pos = s.n_pos;
        have_ref = true;
      } else if (s.n_seeds != seeds || s.n_pos != pos) {
        return GS_EHIP;  // variants disagree: must not happen
      }
      if (s.t_total < *best) *best = s.t_total;
    }
    ++tried;
    return GS_OK;
  }
};

}  // namespace

extern "C" GS_API int gs_autotune(gs_index* idx, const char* bases, const uint64_t* offsets, const uint32_t* lens,
                                  uint64_t n_reads, const gs_opts* base, int reps, gs_tune_result* out) {
  if (!idx || !base || !out || !n_reads || !bases || !offsets || !lens) return GS_EINVAL;
  memset(out, 0, sizeof(*out));
  gs_opts cur = *base;
  if (cur.kernel == 0) cur.kernel = 3;
  if (!cur.wg) cur.wg = 128;
  if (!cur.walk_rows) cur.walk_rows = 1;
  if (!cur.seg_tasks) cur.seg_tasks = 64;
  bool has_long = false;
  for (uint64_t i = 0; i < n_reads && !has_long; ++i) has_long = lens[i] > gs::kPersistLmax;
  Timer T;
  T.reps = reps > 0 ? reps : 3;
  double best = 0;
  int rc = GS_OK;

  // 1. work-item size of long reads: one batch per value
  static const uint32_t segs[] = {32, 64, 96, 112};
  gs_batch* B = nullptr;
  if (has_long) {
    uint32_t best_seg = cur.seg_tasks;
    double best_t = 1e30;
    for (uint32_t sg : segs) {
      gs_opts o = cur;
      o.seg_tasks = sg;
      gs_batch* C = nullptr;
      if ((rc = gs_batch_create(idx, bases, offsets, lens, n_reads, &o, &C))) return rc;
      T.B = C;
      double t = 0;
      rc = gs_batch_run(C, -1, nullptr);  // warm-up
      if (!rc) rc = T.time(&t);
      gs_batch_destroy(C);
      if (rc) return rc;
      if (sg == cur.seg_tasks) out->default_seconds = t;
      if (t < best_t) {
        best_t = t;
        best_seg = sg;
      }
    }
    cur.seg_tasks = best_seg;
  }
  if ((rc = gs_batch_create(idx, bases, offsets, lens, n_reads, &cur, &B))) return rc;
  T.B = B;
  do {
    if ((rc = gs_batch_run(B, -1, nullptr))) break;  // warm-up
    if ((rc = T.time(&best))) break;
    if (!has_long) out->default_seconds = best;
    // 2. size-record width x workgroup size
    static const int wgs[] = {64, 128, 256};
    int bk = cur.kernel, bw = cur.wg;
    for (int kv = 1; kv <= 3 && !rc; ++kv)
      for (int wg : wgs) {
        if (kv == cur.kernel && wg == cur.wg) continue;
        if ((rc = gs_batch_set_tuning(B, kv, wg, cur.walk_rows))) break;
        double t = 0;
        if ((rc = T.time(&t))) break;
        if (t < best) {
          best = t;
          bk = kv;
          bw = wg;
        }
      }
    if (rc) break;
    cur.kernel = bk;
    cur.wg = (uint16_t)bw;
    // 3. rows per thread of the suffix-array walk
    static const int rows[] = {1, 2, 4};
    int br = cur.walk_rows;
    for (int wr : rows) {
      if (wr == cur.walk_rows) continue;
      if ((rc = gs_batch_set_tuning(B, cur.kernel, cur.wg, wr))) break;
      double t = 0;
      if ((rc = T.time(&t))) break;
      if (t < best) {
        best = t;
        br = wr;
      }
    }
    if (rc) break;
    cur.walk_rows = (uint16_t)br;
    // 4. direct extension
    if (idx->sa_full) {
      static const uint16_t dmins[] = {4, 2, 8, 0xFFFF};
      uint16_t bd = cur.direct_min ? cur.direct_min : 4;
      for (uint16_t dm : dmins) {
        if (dm == bd) continue;
        if ((rc = gs_batch_set_direct_min(B, dm))) break;
        double t = 0;
        if ((rc = T.time(&t))) break;
        if (t < best) {
          best = t;
          bd = dm;
        }
      }
      if (rc) break;
      cur.direct_min = bd;
    }
  } while (0);
  gs_batch_destroy(B);
  if (rc) return rc;
  out->opts = cur;
  out->seconds = best;
  out->n_tried = T.tried;
  return GS_OK;
}
