// Seeding server: a resident index, requests submitted from any thread, and
// continuous batching of their reads into device rounds.
//
// A request's reads are cut at its reference batches (a batch closes after
// the read that brings it to batch_bases, seed_gen.cu:2955-2991), so a chunk
// of a request seeds exactly as the request would alone: SMEM seeds depend on
// the read only, MEM row skips never leave a batch (plan_add starts a new
// batch id for every chunk).  Threads:
//  * a planner builds the next round on the CPU while the device works:
//    every request with unscheduled reads first gets a fair share of the
//    round (round_bases / active requests, at least one reference batch),
//    then the round is filled in arrival order up to the round size and to
//    the device budget divided among the rounds that can hold memory at
//    once.  The round size is round_bases, growing with the backlog up to
//    round_growth x round_bases (a launch's tail is a fixed cost);
//  * `workers` device threads (default 3), each with its own batch and HIP
//    stream, take rounds in turn: upload, the whole pipeline (one
//    persistent-kernel launch over the reads of every request in the round),
//    download into the batch's pinned staging, then hand each chunk's share
//    to its request.  Several rounds on the GPU at once let the waves of the
//    next round fill the compute units that the tail of another leaves idle,
//    and hide each worker's host work (upload, download, hand-off).
// Requests arriving while rounds run join the next one; a large request is
// spread over several rounds, interleaved with the requests behind it, and
// its chunks are put back in read order when it is collected.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gs_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id = 0;
  const char* bases = nullptr;
  const uint64_t* offs = nullptr;
  const uint32_t* lens = nullptr;
  uint64_t n = 0, n_bases = 0;
  std::vector<uint64_t> cuts;  // read index after each reference batch; last = n
  size_t next_cut = 0;         // first cut not yet scheduled
  uint64_t next_read = 0;      // first read not yet scheduled
  uint64_t sched_bases = 0;    // bases of the reads before next_read
  uint64_t done_reads = 0;
  // results of one chunk, written by the worker that runs it, in the layout
  // of a gs_result (offsets relative to the chunk): a request of one chunk
  // hands these arrays to its caller as they are
  struct Piece {
    gs_result r{};
    ~Piece() { gs_result_free(&r); }
  };
  // by chunk ordinal, i.e. in read order; each piece is its own allocation so
  // that a worker filling one never races with the planner adding another
  std::vector<std::shared_ptr<Piece>> pieces;
  uint64_t n_seeds = 0, n_pos = 0;
  double t_submit = 0, t_start = -1, t_done = -1;
  uint32_t rounds = 0;
  uint32_t inflight = 0;   // chunks planned and not yet processed
  int rc = GS_OK;
  bool done = false;       // results final (or failed)
  bool collected = false;  // gs_server_wait has returned: free once inflight is 0
};

struct Chunk {
  Request* q;
  uint64_t r0, r1;  // reads of the request
  uint64_t at;      // first read in the round
  std::shared_ptr<Request::Piece> piece;
};

struct Round {
  gs::BatchPlan plan;
  std::vector<Chunk> chunks;
  uint64_t device_bytes = 0;
  int rc = GS_OK;  // the plan could not be built: the round's requests fail with it
};

}  // namespace

struct gs_server {
  gs_index* X = nullptr;
  gs_server_opts o{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_plan, cv_dev, cv_done;
  std::deque<Request*> active;          // requests with unscheduled reads, in arrival order
  std::map<uint64_t, Request*> reqs;    // every request not yet collected by gs_server_wait
  std::unique_ptr<Round> ready;         // planned round waiting for the device
  uint64_t next_id = 1;
  bool stop = false;
  int fatal = GS_OK;
  std::thread planner;
  std::vector<std::thread> devices;
  std::vector<gs_batch*> batches;  // one per device thread
  int running = 0;                 // rounds on the device
  bool building = false;           // the planner builds a round outside the lock
  uint64_t rounds_taken = 0;       // ordinal of the next round a device thread takes
  uint64_t fail_round = UINT64_MAX;  // test support: this round fails with fail_rc before its upload
  int fail_rc = GS_OK;
  gs_server_stats st{};
  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Reads of q from its first unscheduled cut up to whole reference batches
// holding at least `want` bases (one batch at least); returns the end read.
uint64_t take(Request* q, uint64_t want, uint64_t* bases_out) {
  uint64_t b = 0, r = q->next_read;
  size_t c = q->next_cut;
  while (c < q->cuts.size()) {
    const uint64_t end = q->cuts[c];
    for (uint64_t i = r; i < end; ++i) b += q->lens[i];
    r = end;
    ++c;
    if (b >= want) break;
  }
  *bases_out = b;
  return r;
}

void advance(Request* q, uint64_t end) {
  while (q->next_cut < q->cuts.size() && q->cuts[q->next_cut] <= end) ++q->next_cut;
  q->next_read = end;
}

// Chooses the next round's chunks from the active requests (the server lock
// is held); build_round then makes its plan without the lock.
std::unique_ptr<Round> plan_round(gs_server* S) {
  auto R = std::make_unique<Round>();
  uint64_t n_reads = 0;
  // rounds grow with the backlog, up to round_growth x round_bases: a launch
  // has a fixed cost (its tail), so larger rounds raise throughput under
  // load, at the price of longer waits for the requests that arrive behind
  // them; a short queue keeps rounds small
  uint64_t backlog = 0;
  for (Request* q : S->active) backlog += q->n_bases - q->sched_bases;
  const uint64_t grown = backlog / (uint64_t)S->o.workers;
  const uint64_t cap_max = S->o.round_bases * (uint64_t)S->o.round_growth;
  const uint64_t cap = grown < S->o.round_bases ? S->o.round_bases : (grown > cap_max ? cap_max : grown);
  const uint64_t mem_cap = S->o.budget_bytes / (uint64_t)(S->o.workers + 1);
  uint64_t bases = 0, est = 0;
  // device bytes of a chunk, an upper bound of what plan_device_bytes adds:
  // bytes, codes, two candidate slots per task, work items and per-read arrays
  auto chunk_bytes = [](uint64_t b, uint64_t reads) { return b * (1 + 1 + 32 + 24 + 4) + reads * 160; };
  auto add = [&](Request* q, uint64_t end) -> bool {
    if (end <= q->next_read) return true;
    uint64_t b = 0;
    for (uint64_t i = q->next_read; i < end; ++i) b += q->lens[i];
    const uint64_t e = chunk_bytes(b, end - q->next_read);
    if (!R->chunks.empty() && est + e > mem_cap) return false;  // over the budget: a later round
    if (n_reads + (end - q->next_read) > 0x7FFFFFFFull) return false;
    const uint64_t at = n_reads;
    n_reads += end - q->next_read;
    bases += b;
    est += e;
    q->sched_bases += b;
    auto piece = std::make_shared<Request::Piece>();
    q->pieces.push_back(piece);
    R->chunks.push_back({q, q->next_read, end, at, piece});
    ++q->inflight;
    advance(q, end);
    return true;
  };
  if (S->o.exclusive) {  // one whole request per round
    Request* q = S->active.front();
    add(q, q->n);
  } else {
    const uint64_t n_act = S->active.size();
    const uint64_t share = n_act ? (cap + n_act - 1) / n_act : cap;
    for (Request* q : S->active) {  // fair share first
      if (bases >= cap) break;
      uint64_t b = 0;
      const uint64_t end = take(q, share < cap - bases ? share : cap - bases, &b);
      if (!add(q, end)) break;
    }
    for (Request* q : S->active) {  // then arrival order
      while (bases < cap && q->next_read < q->n) {
        uint64_t b = 0;
        const uint64_t end = take(q, cap - bases, &b);
        if (!add(q, end)) break;
      }
      if (bases >= cap) break;
    }
  }
  if (R->chunks.empty()) {  // the first chunk of a round is always taken unless it cannot be planned
    Request* q = S->active.front();
    q->next_read = q->n;
    q->next_cut = q->cuts.size();
    q->rc = GS_ERANGE;
  }
  for (auto it = S->active.begin(); it != S->active.end();)
    it = (*it)->next_read >= (*it)->n ? S->active.erase(it) : it + 1;
  return R;
}

// The round's plan, from its chunks in order (no lock needed: the requests'
// read arrays do not change, and each request stays alive while it has
// chunks in flight).
void build_round(gs_server* S, Round& R) {
  gs::plan_reset(R.plan, S->o.opts);
  for (const Chunk& c : R.chunks) {
    const int rc = gs::plan_add(R.plan, c.q->bases, c.q->offs, c.q->lens, c.r0, c.r1);
    if (rc) {
      R.rc = rc;
      return;
    }
  }
  gs::plan_finish(R.plan);
  R.device_bytes = gs::plan_device_bytes(R.plan);
}

void finish_request(gs_server* S, Request* q, int rc) {
  q->rc = rc;
  q->done = true;
  q->t_done = S->now();
  S->cv_done.notify_all();
}

void planner_main(gs_server* S) {
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    S->cv_plan.wait(lk, [&] { return S->stop || (!S->ready && !S->active.empty()); });
    if (S->active.empty() && S->stop) break;
    if (S->ready || S->active.empty()) continue;
    Request* front = S->active.front();
    const double tp0 = S->now();
    std::unique_ptr<Round> R = plan_round(S);
    if (R->chunks.empty()) {  // nothing could be planned for the front request
      if (!front->done) finish_request(S, front, GS_ERANGE);
      continue;
    }
    S->building = true;  // workers wait for this round even if no request is active
    lk.unlock();
    build_round(S, *R);
    lk.lock();
    S->st.t_plan += S->now() - tp0;
    S->building = false;
    S->ready = std::move(R);
    S->cv_dev.notify_all();
  }
  S->cv_dev.notify_all();
}

void device_main(gs_server* S, gs_batch* B) {
  std::unique_lock<std::mutex> lk(S->mu);
  while (true) {
    const double tw = S->now();
    S->cv_dev.wait(lk, [&] { return S->ready || (S->stop && S->active.empty() && !S->building); });
    if (!S->ready) break;
    S->st.t_idle += S->now() - tw;
    std::unique_ptr<Round> R = std::move(S->ready);
    ++S->running;
    const int injected = S->rounds_taken++ == S->fail_round ? S->fail_rc : GS_OK;
    S->cv_plan.notify_all();
    const double t_round = S->now();
    std::vector<char> skip(R->chunks.size());  // requests that already failed keep nothing more
    for (size_t i = 0; i < R->chunks.size(); ++i) {
      Request* q = R->chunks[i].q;
      if (q->t_start < 0) q->t_start = t_round;
      ++q->rounds;
      skip[i] = q->done;
    }
    lk.unlock();
    // device work without the lock: the planner and the other workers go on
    int rc = R->rc ? R->rc : (injected ? injected : gs::batch_load(B, R->plan));
    const double t1 = S->now();
    gs_batch_stats bs;
    memset(&bs, 0, sizeof bs);
    if (!rc) rc = gs_batch_run(B, -1, &bs);
    const double t2 = S->now();
    gs::ResultView res;  // pinned staging of this worker's batch
    memset(&res, 0, sizeof res);
    if (!rc) rc = gs::batch_download(B, &res);
    const double t3 = S->now();
    if (!rc) {
      // each chunk's results into its piece: the pieces' arrays are allocated
      // here, then filled by up to 8 threads that take fixed-size parts of
      // them (per-read arrays, seed records, positions) from a shared list,
      // so a round of a few large chunks copies, and takes its page faults,
      // as fast as one of many small ones
      bool alloc_failed = false;
      for (size_t i = 0; i < R->chunks.size(); ++i) {
        const Chunk& c = R->chunks[i];
        if (skip[i]) continue;
        gs_result& o = c.piece->r;
        gs_result_free(&o);
        const uint64_t a = c.at, m = c.r1 - c.r0, z = a + m;
        o.n_reads = m;
        o.n_seeds = res.soff[z] - res.soff[a];
        o.n_pos = res.poff[z] - res.poff[a];
        o.read_nseeds = (uint32_t*)gs::result_alloc(4 * (m + 1));
        o.read_npos = (uint32_t*)gs::result_alloc(4 * (m + 1));
        o.read_seed_off = (uint64_t*)gs::result_alloc(8 * (m + 1));
        o.read_pos_off = (uint64_t*)gs::result_alloc(8 * (m + 1));
        o.seeds = (gs_seed_rec*)gs::result_alloc(sizeof(gs_seed_rec) * (o.n_seeds + 1));
        o.pos = (uint32_t*)gs::result_alloc(4 * (o.n_pos + 1));
        if (!o.read_nseeds || !o.read_npos || !o.read_seed_off || !o.read_pos_off || !o.seeds || !o.pos) {
          gs_result_free(&o);
          skip[i] = 1;
          alloc_failed = true;
        }
      }
      // parts: kind 0 reads [a, z) of the per-read arrays (the offsets
      // rebased to the chunk, the last one included with the chunk's last
      // part), kind 1 seed records, kind 2 positions
      struct Part {
        uint32_t chunk, kind;
        uint64_t a, z;
      };
      constexpr uint64_t kReadsPer = 1u << 17, kSeedsPer = 1u << 17, kPosPer = 1u << 20;
      std::vector<Part> parts;
      uint64_t bytes = 0;
      for (size_t i = 0; i < R->chunks.size(); ++i) {
        if (skip[i]) continue;
        const gs_result& o = R->chunks[i].piece->r;
        for (uint64_t a = 0; a < o.n_reads || a == 0; a += kReadsPer)
          parts.push_back({(uint32_t)i, 0u, a, std::min<uint64_t>(o.n_reads, a + kReadsPer)});
        for (uint64_t a = 0; a < o.n_seeds; a += kSeedsPer)
          parts.push_back({(uint32_t)i, 1u, a, std::min<uint64_t>(o.n_seeds, a + kSeedsPer)});
        for (uint64_t a = 0; a < o.n_pos; a += kPosPer)
          parts.push_back({(uint32_t)i, 2u, a, std::min<uint64_t>(o.n_pos, a + kPosPer)});
        bytes += 24 * o.n_reads + sizeof(gs_seed_rec) * o.n_seeds + 4 * o.n_pos;
      }
      std::atomic<size_t> next{0};
      auto copier = [&]() {
        for (size_t k = next++; k < parts.size(); k = next++) {
          const Part& pt = parts[k];
          const Chunk& c = R->chunks[pt.chunk];
          gs_result& o = c.piece->r;
          const uint64_t at = c.at, s0 = res.soff[at], p0 = res.poff[at];
          if (pt.kind == 0) {
            memcpy(o.read_nseeds + pt.a, res.nseeds + at + pt.a, 4 * (pt.z - pt.a));
            memcpy(o.read_npos + pt.a, res.npos + at + pt.a, 4 * (pt.z - pt.a));
            const uint64_t last = pt.z == o.n_reads ? pt.z : pt.z - 1;  // offsets have n_reads + 1 entries
            for (uint64_t j = pt.a; j <= last; ++j) {
              o.read_seed_off[j] = res.soff[at + j] - s0;
              o.read_pos_off[j] = res.poff[at + j] - p0;
            }
          } else if (pt.kind == 1) {
            memcpy(o.seeds + pt.a, res.seeds + s0 + pt.a, sizeof(gs_seed_rec) * (pt.z - pt.a));
          } else {
            memcpy(o.pos + pt.a, res.pos + p0 + pt.a, 4 * (pt.z - pt.a));
          }
        }
      };
      const size_t nt = std::min<size_t>({parts.size(), (size_t)8, (size_t)(bytes / (4u << 20)) + 1});
      std::vector<std::thread> helpers;
      try {
        for (size_t t = 1; t < nt; ++t) helpers.emplace_back(copier);
      } catch (...) {  // fewer helpers: the parts they would take stay in the list for this thread
      }
      copier();
      for (auto& h : helpers) h.join();
      if (alloc_failed) rc = GS_ENOMEM;
    }
    const double t_end = S->now();
    lk.lock();
    --S->running;
    S->st.n_rounds++;
    S->st.busy_seconds += t_end - t_round;
    S->st.device_seconds += bs.t_total;
    S->st.t_seed += bs.t_seed;
    S->st.t_post += bs.t_count + bs.t_write;
    S->st.t_locate += bs.t_locate;
    S->st.t_load += t1 - t_round;
    S->st.t_run += t2 - t1;
    S->st.t_fetch += t3 - t2;
    S->st.t_split += t_end - t3;
    S->st.n_reads += R->plan.n_reads;
    S->st.n_bases += R->plan.span;
    if (R->device_bytes > S->st.peak_round_bytes) S->st.peak_round_bytes = R->device_bytes;
    for (const Chunk& c : R->chunks) {
      Request* q = c.q;
      --q->inflight;
      if (!q->done) {
        if (rc) {
          // nothing more is planned for a failed request: it leaves the
          // active list now, so the planner never walks it once
          // gs_server_wait has freed it
          q->next_read = q->n;
          q->next_cut = q->cuts.size();
          auto it = std::find(S->active.begin(), S->active.end(), q);
          if (it != S->active.end()) S->active.erase(it);
          finish_request(S, q, rc);
        } else {
          q->done_reads += c.r1 - c.r0;
          q->n_seeds += c.piece->r.n_seeds;
          q->n_pos += c.piece->r.n_pos;
          if (q->done_reads == q->n) finish_request(S, q, GS_OK);
        }
      }
      if (q->collected && q->inflight == 0) delete q;
    }
    if (rc) S->fatal = rc;
  }
}

}  // namespace

extern "C" GS_API int gs_server_create(gs_index* X, const gs_server_opts* o, gs_server** out) {
  if (!X || !o || !out) return GS_EINVAL;
  *out = nullptr;
  int rc = gs::plan_check(&o->opts);
  if (rc) return rc;
  gs_server* S = new gs_server();
  S->X = X;
  gs::index_acquire(X);
  X->servers.fetch_add(1);
  S->o = *o;
  if (!S->o.budget_bytes) S->o.budget_bytes = 64ull << 30;
  if (!S->o.round_bases) S->o.round_bases = 16000000ull;
  if (S->o.opts.kernel == 0 && !S->o.exclusive) S->o.opts.kernel = 3;
  // rounds hold few work items per lane, so shorter segments of long reads
  // balance better (peak with 64: 1.31 M reads/s, 48: 1.42 M, 32: 1.48 M)
  if (!S->o.opts.seg_tasks) S->o.opts.seg_tasks = 32;
  if (S->o.workers <= 0) S->o.workers = S->o.exclusive ? 1 : 3;
  if (S->o.workers > 8) S->o.workers = 8;
  if (S->o.round_growth <= 0) S->o.round_growth = S->o.exclusive ? 1 : 4;
  if (S->o.round_growth > 64) S->o.round_growth = 64;
  S->t0 = Clock::now();
  for (int w = 0; w < S->o.workers; ++w) S->batches.push_back(gs::batch_new(X));
  // device buffers for the largest round up front (100 bases per read on
  // average at most), so that rounds allocate, and synchronise the device,
  // only when they exceed it
  const uint64_t max_round = S->o.round_bases * (uint64_t)S->o.round_growth;
  for (gs_batch* B : S->batches) {
    rc = gs::batch_reserve(B, max_round, max_round / 100, S->o.opts.seg_tasks);
    if (rc) break;
  }
  if (rc) {
    for (gs_batch* B : S->batches) gs::batch_free(B);
    X->servers.fetch_sub(1);
    gs::index_release(X);
    delete S;
    return rc;
  }
  S->planner = std::thread(planner_main, S);
  for (int w = 0; w < S->o.workers; ++w) S->devices.emplace_back(device_main, S, S->batches[w]);
  *out = S;
  return GS_OK;
}

extern "C" GS_API int gs_server_submit(gs_server* S, const char* bases, const uint64_t* offsets,
                                       const uint32_t* lens, uint64_t n_reads, uint64_t* ticket) {
  if (!S || !ticket || (n_reads && (!bases || !offsets || !lens))) return GS_EINVAL;
  if (n_reads > 0xFFFFFFFFull / 2) return GS_ERANGE;
  Request* q = new Request();
  q->bases = bases;
  q->offs = offsets;
  q->lens = lens;
  q->n = n_reads;
  const uint32_t bb = S->o.opts.batch_bases ? S->o.opts.batch_bases : 1000000u;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n_reads; ++i) {
    acc += lens[i];
    q->n_bases += lens[i];
    if (acc >= bb) {
      q->cuts.push_back(i + 1);
      acc = 0;
    }
  }
  if (q->cuts.empty() || q->cuts.back() != n_reads) q->cuts.push_back(n_reads);
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) {
    delete q;
    return GS_EINVAL;
  }
  q->id = S->next_id++;
  q->t_submit = S->now();
  S->reqs[q->id] = q;
  S->st.n_requests++;
  *ticket = q->id;
  if (n_reads == 0) {
    q->t_start = q->t_submit;
    finish_request(S, q, GS_OK);
    return GS_OK;
  }
  S->active.push_back(q);
  S->cv_plan.notify_all();
  return GS_OK;
}

// Frees a collected request now, or leaves it to the device thread when a
// failed request still has chunks in a round (the server lock is held).
static void release_request(Request* q) {
  if (q->inflight == 0)
    delete q;
  else
    q->collected = true;
}

extern "C" GS_API int gs_server_wait(gs_server* S, uint64_t ticket, gs_result* out, gs_request_info* info) {
  if (!S || !out) return GS_EINVAL;
  memset(out, 0, sizeof(*out));
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return GS_EINVAL;
  Request* q = it->second;
  S->cv_done.wait(lk, [&] { return q->done; });
  S->reqs.erase(it);
  if (info) {
    memset(info, 0, sizeof(*info));
    info->t_submit = q->t_submit;
    info->t_start = q->t_start;
    info->t_done = q->t_done;
    info->n_reads = q->n;
    info->n_seeds = q->n_seeds;
    info->n_pos = q->n_pos;
    info->n_rounds = q->rounds;
  }
  if (q->rc) {
    const int rc = q->rc;
    release_request(q);
    return rc;
  }
  // a successful request has no chunk left in any round: its pieces are final
  lk.unlock();
  const uint64_t n = q->n, nr1 = n + 1;
  int rc = GS_OK;
  if (q->pieces.size() == 1) {  // one chunk: its arrays are the result
    *out = q->pieces[0]->r;
    memset(&q->pieces[0]->r, 0, sizeof(gs_result));
  } else {
    out->n_reads = n;
    out->n_seeds = q->n_seeds;
    out->n_pos = q->n_pos;
    out->read_nseeds = (uint32_t*)gs::result_alloc(4 * nr1);
    out->read_npos = (uint32_t*)gs::result_alloc(4 * nr1);
    out->read_seed_off = (uint64_t*)gs::result_alloc(8 * nr1);
    out->read_pos_off = (uint64_t*)gs::result_alloc(8 * nr1);
    out->seeds = (gs_seed_rec*)gs::result_alloc(sizeof(gs_seed_rec) * (out->n_seeds + 1));
    out->pos = (uint32_t*)gs::result_alloc(4 * (out->n_pos + 1));
    if (!out->read_nseeds || !out->read_npos || !out->read_seed_off || !out->read_pos_off || !out->seeds ||
        !out->pos) {
      gs_result_free(out);
      rc = GS_ENOMEM;
    } else {
      uint64_t so = 0, po = 0, i = 0;
      for (const auto& pp : q->pieces) {  // chunks in read order
        const gs_result& pc = pp->r;
        memcpy(out->read_nseeds + i, pc.read_nseeds, 4 * pc.n_reads);
        memcpy(out->read_npos + i, pc.read_npos, 4 * pc.n_reads);
        for (uint64_t j = 0; j < pc.n_reads; ++j) {
          out->read_seed_off[i + j] = so + pc.read_seed_off[j];
          out->read_pos_off[i + j] = po + pc.read_pos_off[j];
        }
        memcpy(out->seeds + so, pc.seeds, sizeof(gs_seed_rec) * pc.n_seeds);
        memcpy(out->pos + po, pc.pos, 4 * pc.n_pos);
        i += pc.n_reads;
        so += pc.n_seeds;
        po += pc.n_pos;
      }
      out->read_seed_off[n] = so;
      out->read_pos_off[n] = po;
      if (i != n || so != out->n_seeds || po != out->n_pos) {
        gs_result_free(out);
        rc = GS_EHIP;  // pieces do not add up: must not happen
      }
    }
  }
  lk.lock();
  release_request(q);
  return rc;
}

extern "C" GS_API int gs_server_get_stats(gs_server* S, gs_server_stats* st) {
  if (!S || !st) return GS_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  *st = S->st;
  st->elapsed = S->now();
  return S->fatal;
}

extern "C" GS_API double gs_server_clock(gs_server* S) { return S ? S->now() : 0.0; }

extern "C" GS_API int gs_server_destroy(gs_server* S) {
  if (!S) return GS_EINVAL;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv_plan.notify_all();
    S->cv_dev.notify_all();
  }
  S->planner.join();
  for (auto& t : S->devices) t.join();
  for (auto& kv : S->reqs) delete kv.second;  // submitted but never collected
  for (gs_batch* B : S->batches) gs::batch_free(B);
  const int rc = S->fatal;
  gs_index* X = S->X;
  delete S;
  X->servers.fetch_sub(1);
  gs::index_release(X);
  return rc;
}

// Test support: the round_no-th round a device thread takes (from 0) fails
// with rc before its upload, as a failed upload or launch would.  No
// counterpart in the reference.
extern "C" GS_API int gs_server_inject_failure(gs_server* S, uint64_t round_no, int rc) {
  if (!S || rc >= 0) return GS_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  S->fail_round = round_no;
  S->fail_rc = rc;
  return GS_OK;
}
